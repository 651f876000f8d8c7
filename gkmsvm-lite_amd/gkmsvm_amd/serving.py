"""A continuous-batching request loop over one resident model.

The model's SV tables, coefficients and self-kernels stay in HBM.  Clients
submit scoring, variant (score_variants, kernel method) and GkmExplain
requests from any thread and get a Future.  A launcher thread admits pending
requests into GPU steps while earlier steps run: each step packs every
admitted scoring/variant row into one batch (rows of any length, encoded in
one pass and sent in one pinned copy) and launches the index or pairwise
kernel (or the l-mer table lookup) once for all of them on its own HIP
stream; the step's explain requests follow as one batch per mode.  A
harvester thread waits on each step's completion event and hands out its
results.  Up to ``streams`` steps are in flight, so a request that arrives
while others are on the GPU is admitted as soon as a stream frees up and its
work overlaps theirs.  Admission is by memory: a step takes requests while the bytes it
would hold (histograms, planes, outputs) stay within ``hbm_budget``; a
request larger than the whole budget is rejected with MemoryError.

Each row's result depends only on that row (the per-row reduction runs in a
fixed order), so a request scores identically alone and in any mix.  For
non-RBF models the server builds the exact l-mer tables at start (``tables``):
scoring and variant rows then cost one table lookup per window and explain
rows l lookups per position, independent of the number of SVs.
"""

from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass

import numpy as np
import torch

from . import engine as E
from . import tables as T
from .codec import encode_batch
from .inputs import short_rows  # noqa: F401  (documented NaN semantics)


class Request:
    """One client request: kind "score" | "variant" | "explain"; seqs [str] or,
    for variants, [(ref, alt)]; its Future, timestamps, and the rows, longest
    sequence and HBM bytes it takes (computed once)."""

    __slots__ = ("kind", "seqs", "mode", "future", "t_submit", "t_admit", "t_done",
                 "n_rows", "lmax", "nbytes")

    def __init__(self, kind: str, seqs: list, mode: str = "hypothetical"):
        self.kind = kind
        self.seqs = seqs
        self.mode = mode
        self.future = Future()
        self.t_submit = self.t_admit = self.t_done = 0.0
        if kind == "variant":
            self.n_rows = 2 * len(seqs)
            self.lmax = max((max(len(r), len(a)) for r, a in seqs), default=0)
        else:
            self.n_rows = len(seqs)
            self.lmax = max((len(q) for q in seqs), default=0)
        self.nbytes = 0

    def rows(self) -> int:
        return self.n_rows

    def max_len(self) -> int:
        return self.lmax


@dataclass
class _Step:
    reqs: list
    stream: torch.cuda.Stream
    done: torch.cuda.Event
    score_host: torch.Tensor | None
    spans: list               # (request, start row, n rows) for scoring/variant rows
    explain_out: list         # (request, host tensor [rows, 4, L], first row, rows, length)
    keep: list                # device tensors kept alive until the step completes


class GkmServer:
    """Serve one model; see the module docstring."""

    def __init__(self, model, *, hbm_budget: int = 8 << 30, max_rows: int = 4096,
                 streams: int = 2, max_explain: int = 64, tables: bool = True):
        model.rocm()
        self.model = model
        self.sv = model._svs()
        self.kp = model._params()
        # A server amortises the exact l-mer tables over its whole stream: with
        # ``tables`` it builds them up front wherever the model allows them
        # (non-RBF; use_lmer_table / use_explain_table not False), so scoring
        # and variant rows cost one lookup per window and explain rows l per
        # position, whatever the number of SVs.
        self.table = None
        self.explain_tables = False
        if tables and model.use_lmer_table is not False:
            self.table = model._build_lmer_table()
        if (tables and model.use_explain_table is not False and self.table is not None
                and T.explain_table_bytes(self.kp.l) <= T.EXPLAIN_TABLE_BYTES):
            for pert in (False, True):
                self._build_explain(pert)
            self.explain_tables = True
        self.hbm_budget = int(hbm_budget)
        self.max_rows = int(max_rows)
        self.max_explain = int(max_explain)
        self._streams = [torch.cuda.Stream(device=self.kp.device) for _ in range(streams)]
        self._free = list(self._streams)
        self._pending: deque = deque()
        self._inflight: list = []
        self._lock = threading.Condition()
        self._stop = False
        self._thread = self._harvester = None
        self._launcher_done = False
        self.completed: list = []

    def _build_explain(self, pert: bool):
        from .explain import explain_tables

        key = ("explain", pert)
        if key not in self.sv.tables:
            ta, tb, _ = explain_tables(self.kp.base, pert)
            if "V" not in self.sv.tables:
                self.sv.tables["V"] = T.lmer_weights(self.sv)
            self.sv.tables[key] = T.explain_table(self.sv, ta, tb, self.sv.tables["V"])

    # -- client side ----------------------------------------------------------------

    def submit(self, kind: str, seqs, mode: str = "hypothetical") -> Future:
        if kind not in ("score", "variant", "explain"):
            raise ValueError(f"unknown request kind {kind!r}")
        req = Request(kind, list(seqs), mode)
        req.t_submit = time.perf_counter()
        need = req.nbytes = self._bytes(req)
        if need > self.hbm_budget:
            req.future.set_exception(MemoryError(
                f"request needs {need} bytes of HBM, over the server's budget {self.hbm_budget}"))
            return req.future
        with self._lock:
            self._pending.append(req)
            self._lock.notify_all()
        return req.future

    def start(self):
        if self._thread is None:
            self._stop = False
            self._launcher_done = False
            self._thread = threading.Thread(target=self._loop, name="gkm-server", daemon=True)
            self._harvester = threading.Thread(target=self._harvest_loop, name="gkm-results",
                                               daemon=True)
            self._thread.start()
            self._harvester.start()
        return self

    def stop(self):
        """Serve what is pending, then stop both threads."""
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        if self._thread is not None:
            self._thread.join()
            self._harvester.join()
            self._thread = self._harvester = None

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- admission ----------------------------------------------------------------------

    def _bytes(self, req: Request) -> int:
        S, nm = self.sv.S, self.kp.nm
        L = req.max_len()
        planes = 2 * (L // 32 + 4) * 8 + L + 8 * L  # planes, codes, packed windows
        if req.kind == "explain":
            per_sv = 0 if self.explain_tables else S * (8 + nm * 8)
            return len(req.seqs) * (per_sv + 4 * L * 8 * 2 + planes)
        per_sv = 0 if self.table is not None else S * nm * 8
        return req.rows() * (per_sv + planes + 16)

    def _admit(self):
        """Pop the requests of the next step (at least one if any is pending)."""
        batch, used, rows, n_explain = [], 0, 0, 0
        while self._pending:
            req = self._pending[0]
            need = req.nbytes
            is_ex = req.kind == "explain"
            if batch and (used + need > self.hbm_budget
                          or (not is_ex and rows + req.rows() > self.max_rows)
                          or (is_ex and n_explain >= self.max_explain)):
                break
            self._pending.popleft()
            batch.append(req)
            used += need
            if is_ex:
                n_explain += 1
            else:
                rows += req.rows()
        return batch

    # -- the worker ---------------------------------------------------------------------

    def _loop(self):
        """The launcher: whenever requests are pending and a stream is free, it
        admits the next step and launches it (results are the harvester's)."""
        torch.cuda.set_device(self.kp.device)
        while True:
            with self._lock:
                while not (self._pending and self._free):
                    if self._stop and not self._pending:
                        self._launcher_done = True
                        self._lock.notify_all()
                        return
                    self._lock.wait()
                batch = self._admit()
                stream = self._free.pop()
            try:
                st = self._launch(batch, stream)
            except Exception as exc:  # surface launch errors on the requests
                with self._lock:
                    self._free.append(stream)
                    self._lock.notify_all()
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(exc)
                continue
            with self._lock:
                self._inflight.append(st)
                self._lock.notify_all()

    def _harvest_loop(self):
        """The harvester: waits on each step's completion event (oldest first,
        blocking in the runtime rather than polling), hands out its results and
        frees its stream."""
        torch.cuda.set_device(self.kp.device)
        while True:
            with self._lock:
                while not self._inflight:
                    if self._launcher_done:
                        return
                    self._lock.wait()
                st = self._inflight[0]
            st.done.synchronize()
            self._deliver(st)
            with self._lock:
                self._inflight.pop(0)
                self._free.append(st.stream)
                self._lock.notify_all()

    def _deliver(self, st):
        now = time.perf_counter()
        # one copy of the step's scores (the pinned buffer is reused); each
        # request gets its own slice of it
        scores = st.score_host.numpy().copy() if st.score_host is not None else None
        for req, r0, n in st.spans:
            vals = scores[r0:r0 + n]
            if req.kind == "variant":
                half = n // 2
                res = vals[half:] - vals[:half]  # alt - ref
            else:
                res = vals
            req.t_done = now
            req.future.set_result(res)
            self.completed.append(req)
        for req, host, r0, n, L in st.explain_out:
            if req.future.done():  # rejected at launch
                continue
            req.t_done = now
            req.future.set_result(host[r0:r0 + n, :, :L].numpy().copy())
            self.completed.append(req)

    def _launch(self, batch, stream) -> _Step:
        kp, sv, model = self.kp, self.sv, self.model
        t = time.perf_counter()
        spans, strings, keep = [], [], []
        for req in batch:
            req.t_admit = t
            if req.kind == "score":
                rows = req.seqs
            elif req.kind == "variant":
                rows = [r for r, _ in req.seqs] + [a for _, a in req.seqs]
            else:
                continue
            spans.append((req, len(strings), len(rows)))
            strings.extend(rows)
        score_host = None
        explains = [r for r in batch if r.kind == "explain"]
        with torch.cuda.stream(stream):
            if strings:
                # every scoring row of the step encoded in one pass; a request
                # with an empty string or a letter other than ACGTN fails alone
                codes, lens, bad = encode_batch(strings, kp.l)
                if bad.any():
                    for req, r0, n in spans:
                        if bad[r0:r0 + n].any():
                            req.future.set_exception(ValueError(
                                "Invalid sequence (empty, or a letter other than ACGTN)"))
                    spans = [sp for sp in spans if not sp[0].future.done()]
                xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
                out = torch.empty(len(strings), dtype=torch.float64, device=kp.device)
                E.score_rows(sv, xs, model.bias, out, table=self.table)
                out = torch.where(xs.len < kp.l, torch.full_like(out, float("nan")), out)
                score_host = torch.empty(out.shape, dtype=torch.float64, pin_memory=True)
                score_host.copy_(out, non_blocking=True)
                keep += [xs, out]
            ex_out = []
            by_mode: dict = {}
            for req in explains:
                by_mode.setdefault(req.mode, []).append(req)
            for mode, reqs in by_mode.items():
                # every explain request of this mode in one call (rows of any
                # length; each row's attributions depend only on that row)
                from .explain import _MODE_ALIASES, _explain_rows_table, gkmexplain

                seqs = [q for req in reqs for q in req.seqs]
                resolved = _MODE_ALIASES.get(mode)
                if self.explain_tables and resolved in ("hypothetical", "perturbation"):
                    # straight to the explain table: the step's rows encoded in
                    # one pass, as gkmexplain does through the same table
                    codes, lens, bad = encode_batch(seqs, kp.l)
                    r0 = 0
                    for req in reqs:
                        if bad[r0:r0 + len(req.seqs)].any():
                            req.future.set_exception(ValueError(
                                "Invalid sequence (empty, or a letter other than ACGTN)"))
                        r0 += len(req.seqs)
                    pert = resolved == "perturbation"
                    res = torch.empty((len(seqs), 4, codes.shape[1]), dtype=torch.float64,
                                      device=kp.device)
                    _explain_rows_table(model, codes, lens, sv.tables[("explain", pert)], pert,
                                        res)
                    short = torch.from_numpy(lens < kp.l).to(kp.device, non_blocking=True)
                    res = torch.where(short[:, None, None], torch.full_like(res, float("nan")), res)
                else:
                    res = gkmexplain(model, seqs, mode=mode, batch_size=len(seqs))
                host = torch.empty(res.shape, dtype=torch.float64, pin_memory=True)
                host.copy_(res, non_blocking=True)
                keep.append(res)
                r0 = 0
                for req in reqs:
                    ex_out.append((req, host, r0, len(req.seqs), req.max_len()))
                    r0 += len(req.seqs)
            ev = torch.cuda.Event()
            ev.record(stream)
        return _Step(batch, stream, ev, score_host, spans, ex_out, keep)


def latency_stats(reqs) -> dict:
    """p50 / p99 / mean request latency (submit to result) in milliseconds."""
    lat = np.array([(r.t_done - r.t_submit) * 1e3 for r in reqs])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None, "n": 0}
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": int(lat.size)}
