"""A continuous-batching request loop over one resident model.

The model's SV tables, coefficients and self-kernels stay in HBM.  Clients
submit scoring, variant (score_variants, kernel method) and GkmExplain
requests from any thread and get a handle (result(), done(), as a Future's).
A launcher thread admits pending
requests into GPU steps while earlier steps run: each step packs every
admitted scoring/variant row into one batch (rows of any length; their bytes
go over in one pinned copy and are decoded on the device, where rows with a
letter other than ACGTN are flagged to fail alone) and launches the index or pairwise
kernel (or the l-mer table lookup) once for all of them; the step's explain
requests follow as one batch per mode.  A harvester thread waits on each
step's completion event and hands out its results.  Up to ``streams`` steps
(2 on the per-SV kernels, 1 with the tables by default) are in flight, each
with its own pinned buffers, so a request that arrives
while others are on the GPU is admitted as soon as a slot frees up: the host
decodes, stages and launches its step while the one before runs.  The steps
share one HIP stream and so run on the device in admission order: the GPU
goes from one step to the next without waiting on the host, and a step's
short kernels after its index launch (explain rows, the reduction, the copies
back) never queue behind the next step's persistent launch, which holds every
CU.  (Steps on streams of their own ran side by side and finished together;
ordering them with hipStreamWaitEvent blocked the launching thread until the
event had completed.)  While a step is in flight the launcher lets arrivals
join the next step for up to ``coalesce_s`` past the oldest one's (that step
would wait behind the one in flight anyway), so a burst that arrives over a
fraction of a millisecond becomes one step rather than several.  Admission
is by memory: a step takes requests while the bytes it would hold
(histograms, planes, outputs) stay within ``hbm_budget``; a request larger
than the whole budget is rejected with MemoryError.

Each row's result depends only on that row (the per-row reduction runs in a
fixed order), so a request scores identically alone and in any mix.  For
non-RBF models the server builds the exact l-mer tables at start (``tables``):
scoring and variant rows then cost one table lookup per window and explain
rows l lookups per position, independent of the number of SVs.
"""

from __future__ import annotations

import gc
import threading
import time
from collections import deque
from itertools import chain
from concurrent.futures import TimeoutError as FutureTimeout
from dataclasses import dataclass

import numpy as np
import torch

from . import engine as E
from . import tables as T
from .codec import encode_strings
from . import _lib

_BAD_SEQ = "Invalid sequence (empty, or a letter other than ACGTN)"

# gc.freeze() nesting across running servers (the heap alive when the first
# one starts stays out of the collector's scans until the last one stops)
_GC_LOCK = threading.Lock()
_GC_FROZEN = 0


def _gc_freeze():
    global _GC_FROZEN
    with _GC_LOCK:
        if _GC_FROZEN == 0:
            gc.collect()
            gc.freeze()
        _GC_FROZEN += 1


def _gc_unfreeze():
    global _GC_FROZEN
    with _GC_LOCK:
        _GC_FROZEN -= 1
        if _GC_FROZEN == 0:
            gc.unfreeze()


class Request:
    """One client request -- kind "score" | "variant" | "explain"; seqs [str]
    or, for variants, [(ref, alt)]; timestamps; the rows, longest sequence and
    HBM bytes it takes (computed once) -- and the handle submit() returns:
    result(timeout), exception(timeout), done() and add_done_callback(fn) as
    on a concurrent.futures.Future.  A request owns no lock: every waiter
    sleeps on its server's delivery condition, which the server notifies once
    per step after setting all of the step's results."""

    __slots__ = ("kind", "seqs", "mode", "t_submit", "t_admit", "t_done", "n_rows", "lmax",
                 "nbytes", "_cond", "_done", "_value", "_error", "_callbacks")

    def __init__(self, kind: str, seqs: list, mode: str = "hypothetical", cond=None):
        self.kind = kind
        self.seqs = seqs
        self.mode = mode
        self._cond = threading.Condition() if cond is None else cond
        self._done = False
        self._value = self._error = self._callbacks = None
        self.t_submit = self.t_admit = self.t_done = 0.0
        if kind == "variant":
            self.n_rows = 2 * len(seqs)
            self.lmax = max(map(len, chain.from_iterable(seqs)), default=0)
        else:
            self.n_rows = len(seqs)
            self.lmax = max(map(len, seqs), default=0)
        self.nbytes = 0

    def rows(self) -> int:
        return self.n_rows

    def max_len(self) -> int:
        return self.lmax

    # -- the result handle ---------------------------------------------------------

    def done(self) -> bool:
        return self._done

    def _wait(self, timeout):
        if not self._done:
            with self._cond:
                if not self._cond.wait_for(lambda: self._done, timeout):
                    raise FutureTimeout(f"no result within {timeout} s")

    def result(self, timeout: float | None = None):
        """The request's result (scores [n] float64, variant effects [n], or
        attributions [n, 4, L]); raises the request's error."""
        self._wait(timeout)
        if self._error is not None:
            raise self._error
        return self._value

    def exception(self, timeout: float | None = None):
        self._wait(timeout)
        return self._error

    def add_done_callback(self, fn) -> None:
        """fn(request) once the result is set (at once when it already is)."""
        with self._cond:
            if not self._done:
                if self._callbacks is None:
                    self._callbacks = []
                self._callbacks.append(fn)
                return
        fn(self)


class _Slot:
    """A step in flight: the server's stream and this slot's pinned buffers."""

    __slots__ = ("stream", "pool")

    def __init__(self, stream):
        self.stream = stream
        self.pool = {}


@dataclass
class _Step:
    reqs: list
    slot: _Slot
    done: torch.cuda.Event
    score_host: torch.Tensor | None
    score_short: np.ndarray | None  # scoring rows shorter than l (scored NaN)
    bad_host: torch.Tensor | None   # [rows] int32: a letter other than ACGTN (device flags)
    empty: np.ndarray | None        # [rows] empty strings
    spans: list               # (request, rows, alt rows): scoring rows (a slice, None) or a
                              # variant's reference and alternate rows (slices or lists)
    explain_out: list         # (request, host [rows, 4, L], first row, rows, length,
                              #  short rows, first row in the step's rows or None)
    keep: list                # device tensors kept alive until the step completes


class GkmServer:
    """Serve one model; see the module docstring."""

    def __init__(self, model, *, hbm_budget: int = 8 << 30, max_rows: int = 4096,
                 streams: int | None = None, max_explain: int = 64, tables: bool = True,
                 history: int = 1 << 16, stage_len: int = 2048, freeze_gc: bool = True,
                 coalesce_s: float | None = None):
        """hbm_budget: device bytes a step may hold; max_rows / max_explain:
        scoring+variant rows and explain requests per step; streams: steps in
        flight (slots with their own pinned buffers, on one HIP stream; by
        default 2 on the per-SV kernels, 1 with the tables, whose steps are
        host work -- a second one only adds thread contention);
        tables: build the exact l-mer tables at start where the model
        allows them; history: completed requests kept for latency_stats;
        stage_len: pinned host buffers are sized at start for steps of rows up
        to this many bases (longer rows grow them); freeze_gc: while the
        server runs, the objects alive when it starts are moved out of the
        cyclic collector's generations (gc.freeze, process-wide), so a full
        collection scans only what was allocated since -- without it, a
        collection over the interpreter's whole heap (tens of ms with torch
        loaded) lands inside some step; coalesce_s: while a step is in flight,
        how long past the oldest pending request's arrival the launcher lets
        later arrivals join its step (the step would queue behind the one in
        flight on the device anyway); default 0.5 ms on the per-SV kernels,
        0 with the tables, whose steps the host bounds."""
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
        self.coalesce_s = (5e-4 if self.table is None else 0.0) if coalesce_s is None \
            else float(coalesce_s)
        self.max_explain = int(max_explain)
        stream = torch.cuda.Stream(device=self.kp.device)
        if streams is None:
            streams = 1 if self.table is not None else 2
        self._slots = [_Slot(stream) for _ in range(max(1, int(streams)))]
        self._free = list(self._slots)
        # pinned staging and result buffers, one set per slot: a slot's step
        # is delivered (its results copied out) before the slot takes a new
        # step; sized here so no step allocates pinned memory
        self._presize(int(stage_len))
        self._prepare(int(stage_len))
        # the tables and caches above were written on the caller's current
        # stream, which the server's stream does not wait on: finish them now
        torch.cuda.current_stream(self.kp.device).synchronize()
        self.freeze_gc = bool(freeze_gc)
        self._froze = False
        self._pending: deque = deque()
        # rows (scoring / variant), bytes and explain requests pending
        self._pending_rows = self._pending_bytes = self._pending_explain = 0
        self._inflight: list = []
        self._lock = threading.Condition()
        # what every request's waiters sleep on (notified once per step)
        self._delivered = threading.Condition(threading.Lock())
        self._stop = False
        self._thread = self._harvester = None
        self._launcher_done = False
        # (kind, t_submit, t_admit, t_done) of the last `history` completed
        # requests, for latency_stats: tuples of atoms, which the cyclic
        # collector stops tracking, rather than the requests themselves
        self.completed: deque = deque(maxlen=int(history))

    def _presize(self, stage_len: int):
        """Allocate each slot's pinned buffers for a full step of rows up to
        stage_len bases: staged string bytes, scores, explain outputs, flags."""
        rows = self.max_rows + self.max_explain
        for st in self._slots:
            self._host_buffer(st, "stage", (rows * stage_len + 12 * rows + 8,), torch.uint8)
            self._host_buffer(st, "out", (self.max_rows + self.max_explain * 4 * stage_len,))
            # variants scored through their references add two cut rows a pair
            self._host_buffer(st, "bad", (rows + self.max_rows,), torch.int32)
            if self.table is None:
                self._host_buffer(st, "score", (2 * self.max_rows,))
            for i in range(2):
                self._host_buffer(st, f"explain{i}", (self.max_explain * 4 * stage_len,))

    def _prepare(self, stage_len: int):
        """Build the device state that steps would otherwise build lazily on
        the server's stream, where another stream could read it half-written: the
        window-weight table, the explain weight tables, the SVs' l-mer
        weights V (explain without RBF), the SVs' l-mer index and the explain
        neighbourhoods."""
        from .explain import _explain_ab, _explain_masks, _lmer_applies, explain_tables

        kp, sv = self.kp, self.sv
        kp.wtab(stage_len)
        if self.table is not None:
            return
        for pert in (False, True):
            ta, tb, cmax = explain_tables(kp.base, pert)
            _explain_ab(sv, ta, tb, pert)
        if self.model.use_lmer_index is False:
            return
        for pert in (False, True):
            ta, tb, cmax = explain_tables(kp.base, pert)
            if _lmer_applies(self.model, sv, kp, ta, tb, cmax):
                if "V" not in sv.tables:
                    sv.tables["V"] = T.lmer_weights(sv)
                _explain_masks(sv, ta, tb, cmax, pert)
        if sv.index() is None:
            return
        for pert in (False, True):
            ta, tb, cmax = explain_tables(kp.base, pert)
            if cmax >= 0:
                _explain_masks(sv, ta, tb, cmax, pert)

    def _build_explain(self, pert: bool):
        from .explain import explain_tables

        key = ("explain", pert)
        if key not in self.sv.tables:
            ta, tb, _ = explain_tables(self.kp.base, pert)
            if "V" not in self.sv.tables:
                self.sv.tables["V"] = T.lmer_weights(self.sv)
            self.sv.tables[key] = T.explain_table(self.sv, ta, tb, self.sv.tables["V"])

    # -- client side ----------------------------------------------------------------

    def _request(self, kind, seqs, mode, t):
        """(request, None) to queue, or (request, its result item) when it is
        settled at submission: empty (its empty result) or larger than the
        HBM budget (MemoryError)."""
        if kind not in ("score", "variant", "explain"):
            raise ValueError(f"unknown request kind {kind!r}")
        req = Request(kind, list(seqs), mode, self._delivered)
        req.t_submit = t
        if req.n_rows == 0:  # nothing to compute: the empty result at once
            req.t_admit = t
            return req, (req, np.empty((0, 4, 0) if kind == "explain" else 0), None)
        need = req.nbytes = self._bytes(req)
        if need > self.hbm_budget:
            return req, (req, None, MemoryError(
                f"request needs {need} bytes of HBM, over the server's budget {self.hbm_budget}"))
        return req, None

    def submit(self, kind: str, seqs, mode: str = "hypothetical") -> Request:
        """Queue a request; returns its handle (see Request)."""
        req, settled = self._request(kind, seqs, mode, time.perf_counter())
        if settled is not None:
            self._finish([settled])
            return req
        with self._lock:
            self._pending.append(req)
            self._count_pending(req, 1)
            self._lock.notify_all()
        return req

    def _count_pending(self, req, sign):
        self._pending_bytes += sign * req.nbytes
        if req.kind == "explain":
            self._pending_explain += sign
        else:
            self._pending_rows += sign * req.n_rows

    def submit_many(self, items) -> list:
        """Queue requests given as (kind, seqs) or (kind, seqs, mode) together,
        under one lock and one wake-up of the launcher; returns their handles
        in order, as submit would one at a time."""
        t = time.perf_counter()
        reqs, queued, settled = [], [], []
        for item in items:
            req, done = self._request(item[0], item[1],
                                      item[2] if len(item) > 2 else "hypothetical", t)
            reqs.append(req)
            if done is None:
                queued.append(req)
            else:
                settled.append(done)
        if settled:
            self._finish(settled)
        if queued:
            with self._lock:
                self._pending.extend(queued)
                for req in queued:
                    self._count_pending(req, 1)
                self._lock.notify_all()
        return reqs

    def _finish(self, items):
        """Set results [(request, value, error)] (requests already done are
        skipped), record the completed ones, wake every waiter once, then run
        the requests' done-callbacks."""
        now = time.perf_counter()
        called = []
        with self._delivered:
            for req, value, error in items:
                if req._done:
                    continue
                req._value, req._error, req.t_done = value, error, now
                req._done = True
                if error is None and req.n_rows:
                    self.completed.append((req.kind, req.t_submit, req.t_admit, now))
                if req._callbacks:
                    called.append(req)
            self._delivered.notify_all()
        for req in called:
            for fn in req._callbacks:
                fn(req)

    def start(self):
        if self._thread is None:
            if self.freeze_gc and not self._froze:
                _gc_freeze()
                self._froze = True
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
        if self._froze:
            _gc_unfreeze()
            self._froze = False

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
        if (self._pending_bytes <= self.hbm_budget and self._pending_rows <= self.max_rows
                and self._pending_explain <= self.max_explain):  # all of them
            batch = list(self._pending)
            self._pending.clear()
            self._pending_rows = self._pending_bytes = self._pending_explain = 0
            return batch
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
            self._count_pending(req, -1)
            batch.append(req)
            used += need
            if is_ex:
                n_explain += 1
            else:
                rows += req.rows()
        return batch

    def _hold(self) -> float:
        """Seconds to hold the pending requests back (0: admit now).  While a
        step is in flight the next one only queues behind it on the device,
        so arrivals get until coalesce_s past the oldest pending request's to
        join it; no hold when the GPU is idle, when stopping, or once the
        pending rows fill a step."""
        if not self._inflight or self._stop or self.coalesce_s <= 0:
            return 0.0
        left = self._pending[0].t_submit + self.coalesce_s - time.perf_counter()
        if left <= 0 or self._pending_rows >= self.max_rows:
            return 0.0
        return left

    # -- the worker ---------------------------------------------------------------------

    def _loop(self):
        """The launcher: whenever requests are pending and a slot is free, it
        admits the next step (after the hold of _hold) and launches it
        (results are the harvester's)."""
        torch.cuda.set_device(self.kp.device)
        while True:
            with self._lock:
                while True:
                    while not (self._pending and self._free):
                        if self._stop and not self._pending:
                            self._launcher_done = True
                            self._lock.notify_all()
                            return
                        self._lock.wait()
                    hold = self._hold()
                    if hold <= 0:
                        break
                    self._lock.wait(hold)
                batch = self._admit()
                slot = self._free.pop()
            try:
                st = self._launch(batch, slot)
            except Exception as exc:  # surface launch errors on the requests
                with self._lock:
                    self._free.append(slot)
                    self._lock.notify_all()
                self._finish([(r, None, exc) for r in batch])
                continue
            with self._lock:
                self._inflight.append(st)
                self._lock.notify_all()

    def _harvest_loop(self):
        """The harvester: waits on each step's completion event (oldest first,
        blocking in the runtime rather than polling), hands out its results and
        frees its slot."""
        torch.cuda.set_device(self.kp.device)
        while True:
            with self._lock:
                while not self._inflight:
                    if self._launcher_done:
                        return
                    self._lock.wait()
                st = self._inflight[0]
            try:
                st.done.synchronize()
                self._deliver(st)
            except Exception as exc:  # a device error or a delivery bug fails the
                self._finish([(r, None, exc) for r in st.reqs])  # step's requests, not the server
            with self._lock:
                self._inflight.pop(0)
                self._free.append(st.slot)
                self._lock.notify_all()

    def _host_buffer(self, slot, key: str, shape, dtype=torch.float64) -> torch.Tensor:
        """A view of `shape` on the slot's pinned buffer `key` (grown when
        too small)."""
        pool = slot.pool
        n = int(np.prod(shape))
        buf = pool.get(key)
        if buf is None or buf.numel() < n or buf.dtype != dtype:
            old = buf.numel() if buf is not None and buf.dtype == dtype else 0
            buf = torch.empty(max(n, 2 * old), dtype=dtype, pin_memory=True)
            pool[key] = buf
        return buf[:n].view(shape)

    def _staging(self, slot, strings) -> torch.Tensor:
        """The slot's pinned staging buffer, large enough for the strings'
        bytes, their int64 starts and int32 lengths."""
        n = len(strings)
        need = -(-sum(map(len, strings)) // 8) * 8 + 12 * n
        return self._host_buffer(slot, "stage", (need,), torch.uint8)

    def _deliver(self, st):
        # rows with a letter other than ACGTN (flagged on the device) or empty
        bad = None
        if st.bad_host is not None:
            bad = st.bad_host.numpy() != 0
            bad |= st.empty
            if not bad.any():
                bad = None
        # one copy of the step's scores out of the slot's pinned buffer; each
        # request gets its own slice of it; rows shorter than l score NaN
        scores = None
        if st.score_host is not None:
            scores = st.score_host.numpy().copy()
            if st.score_short is not None:
                scores[st.score_short] = np.nan
        items = []
        for req, rows, alt_rows in st.spans:
            if bad is not None and (bad[rows].any()
                                    or (alt_rows is not None and bad[alt_rows].any())):
                items.append((req, None, ValueError(_BAD_SEQ)))  # fails alone
            elif alt_rows is None:
                items.append((req, scores[rows], None))
            else:
                items.append((req, scores[alt_rows] - scores[rows], None))  # alt - ref
        for req, host, r0, n, L, short, g0 in st.explain_out:
            if req.done():  # rejected at launch
                continue
            if g0 is not None and bad is not None and bad[g0:g0 + n].any():
                items.append((req, None, ValueError(_BAD_SEQ)))
                continue
            res = host[r0:r0 + n, :, :L].numpy().copy()
            if short is not None and short[r0:r0 + n].any():
                res[short[r0:r0 + n]] = np.nan
            items.append((req, res, None))
        self._finish(items)

    def _rows(self, batch, t, pairs: bool = False):
        """The step's scoring/variant rows: (spans, strings, number of allele
        pairs in score_rows' pairs layout), and its explain requests.  A
        variant's reference and alternate are two rows; with `pairs`, the
        pairs E.variant_cut takes go last instead, in four blocks --
        references, each alternate cut to its windows over the variant, the
        references cut there, the alternates."""
        spans, strings, explains, cuts = [], [], [], []
        l = self.kp.l
        for req in batch:
            req.t_admit = t
            if req.kind == "score":
                r0 = len(strings)
                spans.append((req, slice(r0, r0 + len(req.seqs)), None))
                strings.extend(req.seqs)
            elif req.kind == "variant":
                r0, m = len(strings), len(req.seqs)
                if not pairs:
                    for r, a in req.seqs:
                        strings.append(r)
                        strings.append(a)
                    spans.append((req, slice(r0, r0 + 2 * m, 2), slice(r0 + 1, r0 + 2 * m, 2)))
                    continue
                ref_rows, alt_rows = [], []
                for r, a in req.seqs:
                    c = E.variant_cut(r, a, l)
                    if c is None:
                        ref_rows.append(len(strings))
                        alt_rows.append(len(strings) + 1)
                        strings.append(r)
                        strings.append(a)
                    else:  # rows set once the blocks are laid out
                        cuts.append((ref_rows, alt_rows, len(ref_rows), r, a, c))
                        ref_rows.append(-1)
                        alt_rows.append(-1)
                spans.append((req, ref_rows, alt_rows))
            else:
                explains.append(req)
        P = len(cuts)
        if P:
            b = len(strings)
            strings.extend([r for _, _, _, r, _, _ in cuts])
            strings.extend([a[c0:c1] for _, _, _, _, a, (c0, c1) in cuts])
            strings.extend([r[c0:c1] for _, _, _, r, _, (c0, c1) in cuts])
            strings.extend([a for _, _, _, _, a, _ in cuts])
            for k, (ref_rows, alt_rows, i, _, _, _) in enumerate(cuts):
                ref_rows[i] = b + k
                alt_rows[i] = b + 3 * P + k
        return spans, strings, explains, P

    def _explain_calls(self, reqs, slot, keep, ex_out):
        """Explain requests through gkmexplain, one call per mode (rows of any
        length; each row's attributions depend only on that row); a request
        with an empty string or a letter other than ACGTN fails alone."""
        from .explain import gkmexplain

        by_mode: dict = {}
        for req in reqs:
            by_mode.setdefault(req.mode, []).append(req)
        for i, (mode, group) in enumerate(by_mode.items()):
            _, _, bad = encode_strings([q for req in group for q in req.seqs])
            good, r0 = [], 0
            for req in group:
                if bad[r0:r0 + len(req.seqs)].any():
                    self._finish([(req, None, ValueError(_BAD_SEQ))])
                else:
                    good.append(req)
                r0 += len(req.seqs)
            if not good:
                continue
            seqs = [q for req in good for q in req.seqs]
            res = gkmexplain(self.model, seqs, mode=mode, batch_size=len(seqs))
            hv = self._host_buffer(slot, f"explain{i}", tuple(res.shape))
            hv.copy_(res, non_blocking=True)
            keep.append(res)
            rr = 0
            for req in good:
                ex_out.append((req, hv, rr, len(req.seqs), req.max_len(), None, None))
                rr += len(req.seqs)

    def _launch_tables(self, batch, slot) -> _Step:
        """A step through the exact l-mer tables: the scoring/variant rows and
        the explain rows the explain tables serve go to the device in one
        pinned copy of their bytes, are decoded and packed there (rows with
        other letters flagged, failing alone when delivered) and get their
        self-kernels in one launch; then one table-score launch over the
        scoring rows and one table-explain launch per mode write one device
        buffer, which comes back in one pinned copy.  Explain requests of
        other modes go through gkmexplain."""
        from .explain import _MODE_ALIASES

        kp, sv, model = self.kp, self.sv, self.model
        stream = slot.stream
        spans, strings, explains, _ = self._rows(batch, time.perf_counter())
        ns = len(strings)
        groups = {"hypothetical": [], "perturbation": []} if self.explain_tables else {}
        others = []
        for req in explains:
            g = groups.get(_MODE_ALIASES.get(req.mode))
            (g if g is not None else others).append(req)
        ex_groups = []  # (perturbation, requests, first row, rows, output width)
        for mode, reqs in groups.items():
            r0 = len(strings)
            for req in reqs:
                strings.extend(req.seqs)
            if len(strings) > r0:
                width = max(kp.l, max(req.max_len() for req in reqs))
                ex_groups.append((mode == "perturbation", reqs, r0, len(strings) - r0, width))
        score_host = score_short = bad_host = empty = None
        ex_out, keep = [], []
        with torch.cuda.stream(stream):
            if strings:
                xs, dbad = E.DeviceSeqs.from_strings(strings, kp.device, rc=kp.rc, stream=stream,
                                                     min_width=kp.l,
                                                     staging=self._staging(slot, strings))
                lens = xs.lengths
                empty = lens == 0
                short = lens < kp.l  # NaN rows, set when delivered
                dx = E.self_kraw(kp, xs, stream) if kp.normalize else None
                fwd, _ = xs.windows(kp.l, stream)
                wt, _ = kp.wtab(xs.lmax)
                W = fwd.shape[1]
                total = ns + sum(n * 4 * w for _, _, _, n, w in ex_groups)
                dev = torch.empty(total, dtype=torch.float64, device=kp.device)
                sh = E.stream_handle(stream)
                if ns:
                    E._timed("k_table_score",
                             lambda: float(np.clip(lens[:ns] - kp.l + 1, 0, None).sum()) * 12.0,
                             stream, lambda: _lib.call(
                                 "gkm_table_score", E._p(fwd), E._p(xs.len), ns, W, kp.l,
                                 E._p(wt), E._p(self.table), E._p(dx), int(kp.normalize),
                                 float(model.bias), E._p(dev), sh))
                off = ns
                for pert, reqs, r0, n, w in ex_groups:
                    E._timed("k_table_explain",
                             lambda: (float(np.clip(lens[r0:r0 + n] - kp.l + 1, 0, None).sum())
                                      * (4 + 32 * kp.l) + n * 32.0 * w),
                             stream, lambda: _lib.call(
                                 "gkm_table_explain", E._p(fwd[r0:]), E._p(xs.len[r0:]), n, W, w,
                                 kp.l, E._p(wt), E._p(sv.tables[("explain", pert)]),
                                 E._p(dx[r0:]) if dx is not None else None, int(kp.normalize),
                                 int(pert), E._p(dev[off:]), sh))
                    off += n * 4 * w
                host = self._host_buffer(slot, "out", (total,))
                host.copy_(dev, non_blocking=True)
                bad_host = self._host_buffer(slot, "bad", (len(strings),), torch.int32)
                bad_host.copy_(dbad, non_blocking=True)
                keep += [xs, dev, dx, dbad]
                if ns:
                    score_host = host[:ns]
                    score_short = short[:ns] if short[:ns].any() else None
                off = ns
                for _, reqs, r0, n, w in ex_groups:
                    hv = host[off:off + n * 4 * w].view(n, 4, w)
                    sg = short[r0:r0 + n] if short[r0:r0 + n].any() else None
                    rr = 0
                    for req in reqs:
                        ex_out.append((req, hv, rr, len(req.seqs), req.max_len(), sg, r0 + rr))
                        rr += len(req.seqs)
                    off += n * 4 * w
            self._explain_calls(others, slot, keep, ex_out)
            ev = torch.cuda.Event()
            ev.record(stream)
        return _Step(batch, slot, ev, score_host, score_short, bad_host, empty, spans, ex_out,
                     keep)

    def _launch(self, batch, slot) -> _Step:
        """A step through the per-SV kernels (or, with tables, _launch_tables):
        every scoring/variant row of the step decoded on the device from one
        pinned copy of their bytes and scored in one batch."""
        if self.table is not None:
            return self._launch_tables(batch, slot)
        kp, sv, model = self.kp, self.sv, self.model
        stream = slot.stream
        spans, strings, explains, pairs = self._rows(batch, time.perf_counter(),
                                                     pairs=E.pairs_apply(sv))
        score_host = score_short = bad_host = empty = None
        ex_out, keep = [], []
        with torch.cuda.stream(stream):
            if strings:
                xs, dbad = E.DeviceSeqs.from_strings(strings, kp.device, rc=kp.rc, stream=stream,
                                                     min_width=kp.l,
                                                     staging=self._staging(slot, strings))
                empty = xs.lengths == 0
                out = torch.empty(len(strings), dtype=torch.float64, device=kp.device)
                E.score_rows(sv, xs, model.bias, out, stream=stream, pairs=pairs)
                short = xs.lengths < kp.l
                score_short = short if short.any() else None
                score_host = self._host_buffer(slot, "score", out.shape)
                score_host.copy_(out, non_blocking=True)
                bad_host = self._host_buffer(slot, "bad", (len(strings),), torch.int32)
                bad_host.copy_(dbad, non_blocking=True)
                keep += [xs, out, dbad]
            self._explain_calls(explains, slot, keep, ex_out)
            ev = torch.cuda.Event()
            ev.record(stream)
        return _Step(batch, slot, ev, score_host, score_short, bad_host, empty, spans, ex_out,
                     keep)


def latency_stats(records) -> dict:
    """p50 / p99 / mean request latency (submit to result) in milliseconds of
    GkmServer.completed records (kind, t_submit, t_admit, t_done) or of
    Request objects."""
    lat = np.array([(r[3] - r[1]) * 1e3 if isinstance(r, tuple) else (r.t_done - r.t_submit) * 1e3
                    for r in records])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None, "n": 0}
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": int(lat.size)}
