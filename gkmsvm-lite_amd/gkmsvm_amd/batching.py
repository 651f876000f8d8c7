"""Batched scoring of many queries of uneven lengths.

Queries are sorted by length and cut into length buckets (each a multiple of
32 bases); a bucket's rows go to the GPU in steps of ``rows`` rows padded to the
bucket's length -- rows inside a step keep their own lengths, so padding costs
memory only.  For each bucket shape the autotuner times the candidate
configurations once (the l-mer index with a few block targets, the
brute-force kernel with a few launch shapes) on the bucket's first step and
keeps the fastest; every configuration computes the same integers, so the
choice never changes a result.  Models with an exact l-mer table (non-RBF;
built under "auto" when the call's per-SV cost exceeds the build) score every
bucket through the table instead: one lookup per window.  With ``graphs`` each bucket shape is captured
once as a HIP graph (packing, self-kernels, histograms, combine) over static
input buffers and replayed for every later step of that shape.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from . import engine as E
from .inputs import nan_rows, read_batch


class _Plan:
    __slots__ = ("use_index", "shape", "target_blocks", "ms", "table")

    def __init__(self, use_index, shape=None, target_blocks=None, ms=float("inf"), table=None):
        self.use_index = use_index
        self.shape = shape
        self.target_blocks = target_blocks
        self.ms = ms
        self.table = table

    def __repr__(self):
        if self.table is not None:
            return f"table({self.ms:.2f} ms)"
        if self.use_index:
            return f"index(target_blocks={self.target_blocks}, {self.ms:.2f} ms)"
        return f"brute(shape={self.shape}, {self.ms:.2f} ms)"


def _score_step(sv, kp, codes_t, len_t, lengths, out, plan: _Plan):
    """One step on device inputs (codes_t [R, Lb] uint8, len_t [R] int32)."""
    xs = E.DeviceSeqs.__new__(E.DeviceSeqs)
    xs.n, xs.lmax = codes_t.shape
    xs.nw = xs.lmax // 32 + 4
    xs.lengths = lengths
    xs.device = kp.device
    xs.codes = codes_t
    xs.len = len_t
    xs.planes = torch.empty((xs.n, xs.nw, 2), dtype=torch.int32, device=kp.device)
    xs.planes_rc = torch.empty_like(xs.planes) if kp.rc else None
    xs._windows = {}
    _lib.call("gkm_pack_planes", E._p(xs.codes), E._p(xs.len), xs.n, xs.lmax, xs.nw,
              E._p(xs.planes), E._p(xs.planes_rc), E.stream_handle())
    if plan.table is not None:
        return E.table_rows(kp, xs, plan.table, 0.0, out)
    dx = E.self_kraw(kp, xs) if kp.normalize else None
    if plan.use_index and sv.wfix is not None:  # exact level sums from the index itself
        return E.index_score(sv, xs, 0, xs.n, dx, 0.0, out, target_blocks=plan.target_blocks)
    hist = torch.empty((xs.n, sv.S, kp.nm), dtype=torch.int64, device=kp.device)
    if plan.use_index:
        E.index_hist(kp, sv.index(), xs, 0, xs.n, sv.S, out=hist,
                     target_blocks=plan.target_blocks)
    else:
        E.cross_hist(kp, xs, 0, xs.n, sv.seqs, sv.groups, out=hist, shape=plan.shape)
    if sv.wfix is not None:
        _lib.call("gkm_combine_fixed", E._p(hist), xs.n, sv.S, kp.nm, E._p(kp.tab),
                  E._p(dx) if dx is not None else None, E._p(sv.wfix), sv.fix_shift,
                  int(kp.normalize), 0.0, E._p(out), E.stream_handle())
    else:
        _lib.call("gkm_combine_score", E._p(hist), xs.n, sv.S, kp.nm, E._p(kp.tab),
                  E._p(dx) if dx is not None else None, E._p(sv.ds), E._p(sv.coef),
                  int(kp.normalize), int(kp.rbf), kp.gamma, 0.0, E._p(out), E.stream_handle())
    return out


class BatchScorer:
    """Score many sequences of uneven lengths against one resident model."""

    def __init__(self, model, *, rows: int = 256, bucket: int = 128, graphs: bool = True,
                 autotune: bool = True, persistent: bool = True):
        """rows, bucket: rows per step and length granularity of the bucketed
        path; graphs: capture each bucket shape as a HIP graph; autotune: time
        the candidate plans (chunk sizes of the persistent launch) once;
        persistent: kernels without RBF whose batch goes through the l-mer
        index score every row in one persistent launch (gkm_index_score_queue)
        instead of bucket by bucket."""
        model.rocm()
        self.model = model
        self.sv = model._svs()
        self.kp = model._params()
        self.rows = int(rows)
        self.bucket = max(32, int(bucket) // 32 * 32)
        self.graphs = graphs
        self.autotune = autotune
        self.plans: dict = {}
        self._graphs: dict = {}
        self.persistent = persistent
        self.queue_chunk: int | None = None  # autotuned windows per persistent work item
        self.queue_ms: dict = {}
        # the bucket histogram must fit the hist budget of one step
        self.rows = max(1, min(self.rows, E.rows_per_chunk(self.sv.S, self.kp.nm)))

    # -- autotuning ----------------------------------------------------------------

    def _candidates(self, Lb):
        table = self.model._lmer_table_for()
        if table is not None:  # one lookup per window: no per-SV plan comes close
            return [_Plan(False, table=table)]
        cands = []
        if self.kp.l <= E.INDEX_MAX_L and self.kp.dmax >= 0:
            nt = self.sv.index().n_tiles
            cands += [_Plan(True, target_blocks=E.index_blocks([Lb] * self.rows, self.kp.l, nt, c))
                      for c in (16, 32, 64)]
        norient = 2 if self.kp.rc else 1
        cands += [_Plan(False, shape=E.launch_shape(self.rows, self.sv.S, norient))]
        cands += [_Plan(False, shape=s) for s in ((1, 1), (2, 1), (1, 2), (4, 1))]
        return cands

    def _plan_for(self, Lb, codes_t, len_t, lengths, out):
        key = (Lb, codes_t.shape[0])
        if key in self.plans:
            return self.plans[key]
        if not self.autotune:
            plan = _Plan(self.sv.use_index_for(lengths, self.model.use_lmer_index),
                         target_blocks=None, table=self.model._lmer_table_for())
            self.plans[key] = plan
            return plan
        best = None
        for cand in self._candidates(Lb):
            if self.model.use_lmer_index is False and cand.use_index:
                continue
            if (self.model.use_lmer_index is True and not cand.use_index
                    and cand.table is None):
                continue
            _score_step(self.sv, self.kp, codes_t, len_t, lengths, out, cand)  # warm-up
            torch.cuda.synchronize()
            t = time.perf_counter()
            _score_step(self.sv, self.kp, codes_t, len_t, lengths, out, cand)
            torch.cuda.synchronize()
            cand.ms = (time.perf_counter() - t) * 1e3
            if best is None or cand.ms < best.ms:
                best = cand
        self.plans[key] = best
        return best

    # -- steps ----------------------------------------------------------------------------

    def _run_step(self, Lb, codes: np.ndarray, lengths: np.ndarray) -> torch.Tensor:
        dev = self.kp.device
        R = self.rows
        n = codes.shape[0]
        pad = np.zeros((R, Lb), dtype=np.uint8)
        pad[:n, : codes.shape[1]] = codes
        lens = np.zeros(R, dtype=np.int64)
        lens[:n] = lengths
        key = (Lb, R)
        g = self._graphs.get(key)
        if g is not None:
            g["codes"].copy_(torch.from_numpy(pad))
            g["len"].copy_(torch.from_numpy(lens.astype(np.int32)))
            g["graph"].replay()
            return g["out"][:n]
        codes_t = torch.from_numpy(pad).to(dev)
        len_t = torch.from_numpy(lens.astype(np.int32)).to(dev)
        out = torch.empty(R, dtype=torch.float64, device=dev)
        plan = self._plan_for(Lb, codes_t, len_t, lens, out)
        if not self.graphs:
            return _score_step(self.sv, self.kp, codes_t, len_t, lens, out, plan)[:n]
        # capture this shape: lengths live on the device, so any rows up to Lb
        # bases replay correctly (host-side launch sizing uses the bucket's
        # longest row, which bounds every later step's)
        full = np.full(R, Lb, dtype=np.int64)
        if self.sv.index() is not None and plan.use_index:
            self.sv.index()
        _score_step(self.sv, self.kp, codes_t, len_t, full, out, plan)  # warm-up (LDS limits)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            _score_step(self.sv, self.kp, codes_t, len_t, full, out, plan)
        self._graphs[key] = {"graph": graph, "codes": codes_t, "len": len_t, "out": out}
        graph.replay()
        return out[:n]

    def __call__(self, x, *, lengths=None) -> torch.Tensor:
        """[B, 1] float64 decision values (NaN for rows shorter than l)."""
        batch = read_batch(x, lengths)
        l = self.kp.l
        short = batch.short(l)
        codes, lens = batch.codes, batch.lengths
        # under "auto" the whole call decides whether the l-mer table pays
        self.model._plan_lmer_table(float(np.clip(lens - l + 1, 0, None).sum()))
        if self.model._lmer_table_for() is not None:
            # plans and graphs made before the table existed are stale
            self.plans = {k: v for k, v in self.plans.items() if v.table is not None}
            self._graphs = {k: g for k, g in self._graphs.items() if k in self.plans}
        if self._use_queue(lens):
            return nan_rows(self._score_queue(codes, lens)[:, None], short)
        order = np.argsort(lens, kind="stable")
        out = torch.empty(len(lens), dtype=torch.float64, device=self.kp.device)
        i = 0
        while i < len(order):
            # bucket of the next row's length; take up to `rows` rows that fit it
            Lb = max(l, -(-int(lens[order[i]]) // self.bucket) * self.bucket)
            j = i
            while j < len(order) and j - i < self.rows and lens[order[j]] <= Lb:
                j += 1
            idx = order[i:j]
            res = self._run_step(Lb, codes[idx][:, :Lb], lens[idx])
            out[torch.from_numpy(idx).to(out.device)] = res + self.model.bias
            i = j
        return nan_rows(out[:, None], short)

    # -- the persistent launch -------------------------------------------------------

    QUEUE_ROWS = 1 << 17   # rows per persistent launch (device memory of one launch)
    QUEUE_CHUNKS = (16, 24, 32, 64)
    QUEUE_TUNE_ROWS = 1024

    def _use_queue(self, lens) -> bool:
        if not self.persistent or self.sv.wfix is None or self.model._lmer_table_for() is not None:
            return False
        if self.kp.l > E.INDEX_MAX_L or self.kp.dmax < 0:
            return False
        return self.sv.use_index_for(lens, self.model.use_lmer_index)

    def _queue_rows(self, codes, lens, chunk) -> torch.Tensor:
        kp = self.kp
        xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
        dx = E.self_kraw(kp, xs) if kp.normalize else None
        out = torch.empty(xs.n, dtype=torch.float64, device=kp.device)
        return E.index_score_queue(self.sv, xs, dx, self.model.bias, out, chunk=chunk)

    def _tune_queue(self, codes, lens):
        """Windows per work item: each candidate timed once on a sample of the
        rows spread over their lengths (every candidate gives the same bits)."""
        if not self.autotune:
            self.queue_chunk = E.QUEUE_CHUNK
            return
        order = np.argsort(lens, kind="stable")
        pick = order[np.linspace(0, len(order) - 1, min(len(order), self.QUEUE_TUNE_ROWS))
                     .astype(np.int64)]
        c, ln = codes[pick], lens[pick]
        best = None
        for chunk in self.QUEUE_CHUNKS:
            self._queue_rows(c, ln, chunk)  # warm-up
            torch.cuda.synchronize()
            t = time.perf_counter()
            self._queue_rows(c, ln, chunk)
            torch.cuda.synchronize()
            self.queue_ms[chunk] = (time.perf_counter() - t) * 1e3
            if best is None or self.queue_ms[chunk] < self.queue_ms[best]:
                best = chunk
        self.queue_chunk = best

    def _score_queue(self, codes, lens) -> torch.Tensor:
        """[B] decision values of every row through persistent launches of up
        to QUEUE_ROWS rows each."""
        if self.queue_chunk is None:
            self._tune_queue(codes, lens)
        parts = []
        for s in range(0, len(lens), self.QUEUE_ROWS):
            e = min(len(lens), s + self.QUEUE_ROWS)
            w = max(self.kp.l, int(lens[s:e].max()))
            parts.append(self._queue_rows(np.ascontiguousarray(codes[s:e, :w]), lens[s:e],
                                          self.queue_chunk))
        return parts[0] if len(parts) == 1 else torch.cat(parts)
