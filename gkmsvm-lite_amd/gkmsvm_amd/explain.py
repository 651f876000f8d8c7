"""GkmExplain attributions and the mutation impact score on the GPU
(gkmsvm-lite explain/core.py:34-248, explain/weights.py:9-57,
explain/impact.py:51-194)."""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import engine as E
from .codec import codes_to_one_hot
from .inputs import mask_positions, nan_rows, read_batch, read_pair, stacked

_MODE_ALIASES = {
    0: "importance",
    1: "hypothetical",
    2: "perturbation",
    "importance": "importance",
    "hypothetical": "hypothetical",
    "perturbation": "perturbation",
}


def explain_tables(base, perturbation: bool):
    """(table_a, table_b, cmax): per-mismatch-count attribution weights, l + 2
    entries each, and the largest count with a non-zero weight.

    Importance (explain/weights.py:19-34): alpha[m] = h[m] / (l - m) for a
    position where the bases agree, kappa[m] = h[m-1] / (l - m + 1) where they
    differ, m <= d.  Perturbation (explain/weights.py:37-57): the kernel change
    of one substitution, h[m+1] - h[m] (match -> mismatch) and h[m-1] - h[m]
    (mismatch -> match)."""
    l = base.l
    h = base._mismatch_table
    d = l if base.d is None else base.d
    a = np.zeros(l + 2)
    b = np.zeros(l + 2)
    if perturbation:
        for m in range(l + 1):
            a[m] = (h[m + 1] if m + 1 <= l else 0.0) - h[m]
            if m >= 1:
                b[m] = h[m - 1] - h[m]
    else:
        for m in range(min(d, l) + 1):
            if l - m > 0:
                a[m] = h[m] / (l - m)
            if m >= 1:
                b[m] = h[m - 1] / (l - m + 1)
    nz = [m for m in range(l + 1) if a[m] != 0 or b[m] != 0]
    return a, b, (max(nz) if nz else -1)


# Window-pair comparisons per second of the per-SV explain kernels (MI355X):
# the per-SV cost that an explain table has to beat
SWEEP_COMPARISONS_PER_S = 5e12

# gkm_index_explain rates (MI355X, l = 11): (window, mask) probes and index
# entries per second; with the sweep's rate, the cost model that picks the
# per-SV explain kernel
INDEX_EXPLAIN_PROBES_PER_S = 3.5e10   # (window, mask, SV tile) records read (l-mer records: per probe)
INDEX_EXPLAIN_ENTRIES_PER_S = 2.0e11

# gkm_lmer_explain (the SVs' l-mer weights V): (window, mask) probes per second
LMER_EXPLAIN_PROBES_PER_S = 1.5e11
LMER_PROBE_BYTES = 13.0  # its algorithmic bytes: the mask, its level and V[v]
# The l-mer path's limits (kernels/lmer_index.py:48-58, explain/lmer.py:21-29):
# V is taken up to l = 12 and, past LMER_MAX_NEIGHBORS neighbours a window,
# only where the SVs hold LMER_SV_WINDOWS_PER_NEIGHBOR windows per neighbour
LMER_MAX_L = 12
LMER_MAX_NEIGHBORS = 8192
LMER_SV_WINDOWS_PER_NEIGHBOR = 6
# sorted windows per block of gkm_lmer_explain_sorted: one per wave (a serving
# step's 10 x 1 kb rows: 0.67 ms in position order, 0.50 ms sorted in blocks
# of 4 windows, 0.56 of 8, 0.76 of 16 -- fewer blocks, less parallelism)
LMER_SORTED_CHUNK = 4

# Explain kernel: "auto" -- for a kernel without RBF, the SVs' l-mer weights
# where the l-mer path applies (explain/lmer.py:31-150: V, summed over the SVs
# once and cached), else the l-mer index when the cost model favours it and
# the model allows it, else the sweep --, "lmer" (gkm_lmer_explain: every
# query window visits its neighbourhood in V; no RBF), "index"
# (gkm_index_explain: the same neighbourhood in the SVs' l-mer index, per-SV
# coefficients) or "sweep" (gkm_explain: lanes = 64 diagonals of one query x
# SV pair).  All compute the same attributions.
EXPLAIN_KERNEL = "auto"


def _sv_chunk(B: int, S: int) -> int:
    """SVs per block of the sweep kernel: enough blocks to fill the chip."""
    spb = max(1, (B * S) // 8192)
    return max(spb, -(-S // 65535))


def gkmexplain(model, x, *, mode="importance", lengths=None, batch_size: int = 50,
               verbose: bool = False) -> torch.Tensor:
    """[B, 4, L] float64 GkmExplain attributions on the model's GPU: importance
    (hypothetical x one-hot; sums to score - bias), hypothetical, or
    perturbation (lsgkm mode 4).  0 past each sequence's length, NaN rows for
    sequences shorter than l."""
    resolved = _MODE_ALIASES.get(mode)
    if resolved is None:
        raise ValueError(
            "mode must be 'importance', 'hypothetical', or 'perturbation' "
            f"(or 0/1/2), got {mode!r}"
        )
    dev = model._require_device()
    batch = read_batch(x, lengths)
    base = model.kernel.base
    short = batch.short(base.l, "attributions")
    codes, lens = batch.codes, batch.lengths
    B, L = codes.shape
    A = model._explain_table_for(B, lens, resolved == "perturbation") if L >= base.l else None
    # the table kernel writes every entry (0 past each row's length)
    out = (torch.empty if A is not None else torch.zeros)((B, 4, L), dtype=torch.float64,
                                                          device=dev)
    if L >= base.l:
        bs = max(1, int(batch_size)) if batch_size else B
        if A is not None:
            bs = max(bs, 4096)  # a table row costs l lookups per position
        bs = min(bs, 65535)
        starts = range(0, B, bs)
        if verbose:
            from tqdm import tqdm

            starts = tqdm(starts, desc="GkmExplain")
        for s in starts:
            e = min(B, s + bs)
            if A is not None:
                _explain_rows_table(model, codes[s:e], lens[s:e], A,
                                    resolved == "perturbation", out[s:e])
            else:
                _explain_rows(model, codes[s:e], lens[s:e], resolved, out[s:e])
    if resolved == "importance":
        out = out * torch.from_numpy(codes_to_one_hot(codes)).to(dev)
    if A is not None:
        return nan_rows(out, short)
    return mask_positions(out, batch, short)


def _explain_rows(model, codes, lens, resolved, out):
    kp = model._params()
    sv = model._svs()
    pert = resolved == "perturbation"
    ta, tb, cmax = explain_tables(kp.base, pert)
    if cmax < 0:
        return
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    B = xs.n
    st = E.stream_handle()
    dx = E.self_kraw(kp, xs) if kp.normalize else torch.ones(B, dtype=torch.float64,
                                                            device=kp.device)
    ta_d, tb_d = _explain_ab(sv, ta, tb, pert)
    kernel = _explain_kernel(model, sv, kp, xs, ta, tb, cmax)
    if kernel == "lmer":
        if "V" not in sv.tables:
            from . import tables as T

            sv.tables["V"] = T.lmer_weights(sv)
        masks, levels = _explain_masks(sv, ta, tb, cmax, pert)
        fwd, _ = xs.windows(kp.l)
        Wx = fwd.shape[1]
        wq, _ = kp.wtab(xs.lmax)
        # every row's windows sorted by l-mer (PAD, 0xFFFFFFFF, last as
        # INT32_MAX): a block's windows share high digits, and so do the V
        # entries of their neighbourhoods
        flat = fwd.reshape(-1)
        keys, order = torch.sort(torch.where(flat < 0, torch.full_like(flat, 2**31 - 1), flat))
        xloc = order.to(torch.int32)
        # the kernel adds into out itself where it is the zero-filled [B, 4, lmax]
        # block gkmexplain passes
        direct = out.is_contiguous() and tuple(out.shape) == (B, 4, xs.lmax)
        buf = out if direct else torch.zeros((B, 4, xs.lmax), dtype=torch.float64,
                                             device=kp.device)
        probes = float(np.clip(xs.lengths - kp.l + 1, 0, None).sum()) * len(masks)
        E._timed("k_lmer_explain_sorted", lambda: probes * LMER_PROBE_BYTES, None,
                 lambda: _lib.call(
                     "gkm_lmer_explain_sorted", int(pert), E._p(keys), E._p(xloc), flat.numel(),
                     E._p(xs.len), B, Wx, xs.lmax, E._p(sv.tables["V"]),
                     E._p(dx) if kp.normalize else None, int(kp.normalize), E._p(masks),
                     E._p(levels), len(masks), kp.l, E._p(wq), LMER_SORTED_CHUNK, E._p(ta_d),
                     E._p(tb_d), E._p(buf), st))
        if not direct:
            out += buf
        return
    buf = torch.zeros((B, 4, xs.lmax), dtype=torch.float64, device=kp.device)
    hist = None
    if kp.rbf:
        hist = E.sv_hist(sv, xs, 0, B, use_index=model.use_lmer_index)
    ec = torch.empty((B, sv.S), dtype=torch.float64, device=kp.device)
    _lib.call("gkm_explain_coef", E._p(hist), B, sv.S, kp.nm, E._p(kp.tab), E._p(dx),
              E._p(sv.ds), E._p(sv.coef), int(kp.normalize), int(kp.rbf), kp.gamma, E._p(ec), st)
    nm = cmax + 1
    wt, wn = kp.wtab(max(xs.lmax, sv.seqs.lmax))
    norient = 2 if kp.rc else 1
    units = lambda: E.comparisons(kp, xs.lengths, sv.seqs.lengths, norient)  # noqa: E731
    if kernel == "index":
        idx = sv.index()
        masks, levels = _explain_masks(sv, ta, tb, cmax, pert)
        fwd, _ = xs.windows(kp.l)
        Wx = fwd.shape[1]
        wq, _ = kp.wtab(xs.lmax)
        chunk = max(4, min(64, -(-Wx * B // 1024)))
        probes = float(np.clip(xs.lengths - kp.l + 1, 0, None).sum()) * len(masks)
        E._timed("k_index_explain",
                 lambda: probes * (E.INDEX_PROBE_BYTES + idx.per_lmer * E.INDEX_HIT_BYTES),
                 None, lambda: _lib.call(
                     "gkm_index_explain", int(pert), E._p(fwd), E._p(xs.len), B, Wx, xs.lmax,
                     E._p(idx.offsets), E._p(idx.ids), E._p(idx.vals), E._p(_explain_records(idx)),
                     E._p(masks),
                     E._p(levels), len(masks), kp.l, E._p(wq), sv.S, idx.tile, idx.n_tiles,
                     chunk, E._p(ta_d), E._p(tb_d), E._p(ec), E._p(buf), st))
    else:
        E._timed("k_sweep", units, None, lambda: _lib.call(
            "gkm_explain", int(pert), E._p(xs.planes), E._p(xs.len), xs.lmax, B, xs.nw,
            E._p(sv.seqs.planes), E._p(sv.seqs.planes_rc) if kp.rc else None,
            E._p(sv.seqs.len), sv.S, sv.seqs.nw, kp.l, nm, cmax, E._p(wt), wn,
            _sv_chunk(B, sv.S), E._p(ta_d), E._p(tb_d), E._p(ec), E._p(buf), st))
    out += buf


def _explain_ab(sv, ta, tb, pert):
    """The explain tables ta, tb of this mode on the device (cached: a copy
    from pageable host memory would wait for the work queued on the stream)."""
    key = ("explain_ab", bool(pert))
    if key not in sv.tables:
        sv.tables[key] = (torch.from_numpy(np.ascontiguousarray(ta)).to(sv.kp.device),
                          torch.from_numpy(np.ascontiguousarray(tb)).to(sv.kp.device))
    return sv.tables[key]


def _explain_masks(sv, ta, tb, cmax, pert):
    """(masks, levels) on the device: the XOR neighbourhood of the levels
    0..cmax with a non-zero explain weight (cached per mode)."""
    key = ("explain_masks", bool(pert))
    if key not in sv.tables:
        weight = np.abs(np.asarray(ta)) + np.abs(np.asarray(tb))
        masks, levels = E.neighbor_masks(sv.kp.l, weight, cmax)
        sv.tables[key] = (torch.from_numpy(masks.view(np.int32)).to(sv.kp.device),
                          torch.from_numpy(levels).to(sv.kp.device))
    return sv.tables[key]


def _explain_records(idx):
    """The records gkm_index_explain reads for this index: the l-mer records
    (unweighted, at most LREC_MAX_TILES tiles), else the per-key records of a
    weighted index or of more tiles, else none (the bounds)."""
    if idx.lmer_records is not None:
        return idx.lmer_records
    if idx.vals is not None or idx.n_tiles > E.LREC_MAX_TILES:
        return idx.records
    return None


def _path_seconds(sv, kp, x_lengths, ta, tb, cmax):
    """(index, sweep) estimated GPU seconds of explaining rows of these
    lengths through the per-SV kernels (index None when it does not apply)."""
    norient = 2 if kp.rc else 1
    pair_s = E.comparisons(kp, x_lengths, sv.seqs.lengths, norient) / SWEEP_COMPARISONS_PER_S
    if kp.l > E.INDEX_MAX_L or cmax < 0:
        return None, pair_s
    from math import comb

    weight = np.abs(np.asarray(ta)) + np.abs(np.asarray(tb))
    n_masks = sum(comb(kp.l, m) * 3**m for m in range(cmax + 1) if weight[m] != 0)
    wx = float(np.clip(np.asarray(x_lengths) - kp.l + 1, 0, None).sum())
    per_lmer = norient * float(np.clip(sv.seqs.lengths - kp.l + 1, 0, None).sum()) / 4**kp.l
    probes = wx * n_masks
    n_tiles = -(-sv.S // E.index_tile(kp.nm, not kp.weighted, sv.S))
    if not kp.weighted and n_tiles <= E.LREC_MAX_TILES:
        n_tiles = 1  # one l-mer record per probe covers every tile
    index_s = (probes * n_tiles / INDEX_EXPLAIN_PROBES_PER_S
               + probes * per_lmer / INDEX_EXPLAIN_ENTRIES_PER_S)
    return index_s, pair_s


def _lmer_applies(model, sv, kp, ta, tb, cmax) -> bool:
    """Whether the l-mer weights path takes these rows (explain/lmer.py:61-79):
    no RBF, the l-mer index allowed, l <= LMER_MAX_L, and a neighbourhood of at
    most LMER_MAX_NEIGHBORS l-mers or SVs with enough windows per neighbour."""
    if kp.rbf or model.use_lmer_index is False or kp.l > LMER_MAX_L or cmax < 0:
        return False
    n = _n_masks(kp, ta, tb, cmax)
    if n <= LMER_MAX_NEIGHBORS:
        return True
    norient = 2 if kp.rc else 1
    sv_windows = norient * float(np.clip(sv.seqs.lengths - kp.l + 1, 0, None).sum())
    return n * LMER_SV_WINDOWS_PER_NEIGHBOR <= sv_windows


def _n_masks(kp, ta, tb, cmax) -> int:
    from math import comb

    weight = np.abs(np.asarray(ta)) + np.abs(np.asarray(tb))
    return sum(comb(kp.l, m) * 3**m for m in range(cmax + 1) if weight[m] != 0)


def per_sv_seconds(model, x_lengths, perturbation: bool) -> float:
    """Estimated GPU seconds of the cheapest explain kernel without the
    explain table for rows of these lengths (what that table has to beat)."""
    kp = model._params()
    sv = model._svs()
    ta, tb, cmax = explain_tables(kp.base, perturbation)
    if _lmer_applies(model, sv, kp, ta, tb, cmax):
        wx = float(np.clip(np.asarray(x_lengths) - kp.l + 1, 0, None).sum())
        return wx * _n_masks(kp, ta, tb, cmax) / LMER_EXPLAIN_PROBES_PER_S
    index_s, pair_s = _path_seconds(sv, kp, x_lengths, ta, tb, cmax)
    if index_s is None or model.use_lmer_index is False:
        return pair_s
    return index_s if model.use_lmer_index is True else min(index_s, pair_s)


def _explain_kernel(model, sv, kp, xs, ta, tb, cmax) -> str:
    """The per-SV explain kernel for these rows (EXPLAIN_KERNEL, and under
    "auto" the cheaper of the index and the sweep by the cost model;
    model.use_lmer_index = False keeps the sweep, True the index)."""
    if EXPLAIN_KERNEL != "auto":
        return EXPLAIN_KERNEL
    if _lmer_applies(model, sv, kp, ta, tb, cmax):
        return "lmer"
    use = model.use_lmer_index
    if use is False or kp.l > E.INDEX_MAX_L or cmax < 0:
        return "sweep"
    if use is True:
        return "index"
    index_s, pair_s = _path_seconds(sv, kp, xs.lengths, ta, tb, cmax)
    return "index" if index_s < pair_s else "sweep"


def _explain_rows_table(model, codes, lens, A, pert, out):
    """GkmExplain of rows from the explain table (explain/table.py:172-196)."""
    kp = model._params()
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    dx = E.self_kraw(kp, xs) if kp.normalize else None
    fwd, _ = xs.windows(kp.l)
    wt, _ = kp.wtab(xs.lmax)
    W = np.clip(xs.lengths - kp.l + 1, 0, None)
    E._timed("k_table_explain", lambda: float(W.sum()) * (4 + 32 * kp.l) + xs.n * 32.0 * xs.lmax,
             None, lambda: _lib.call(
                 "gkm_table_explain", E._p(fwd), E._p(xs.len), xs.n, fwd.shape[1], xs.lmax,
                 kp.l, E._p(wt), E._p(A), E._p(dx), int(kp.normalize), int(pert), E._p(out),
                 E.stream_handle()))


def mutation_impact(model, ref, alt, *, ref_lengths=None, alt_lengths=None,
                    batch_size: int | None = None, verbose: bool = False) -> torch.Tensor:
    """[B, 1] GkmExplain mutation impact (lsgkm gkmexplain -m 5): sum_j coef_j
    dK_j with K and dK normalised by the reference allele's norm; RBF kernels
    scale dK_j by (exp(g(K_j - 1)) - exp(-g)) / K_j (0 for K_j <= 0)."""
    dev = model._require_device()
    kp = model._params()
    sv = model._svs()
    rb, ab = read_pair(ref, alt, ref_lengths, alt_lengths)
    B = rb.n
    # a variant whose either allele has no window scores NaN
    short = stacked(rb, ab).short(kp.l, "mutation impacts")
    if short is not None:
        short = short.reshape(2, -1).any(axis=0)
    ref, alt, rl, al = rb.codes, ab.codes, rb.lengths, ab.lengths
    out = torch.zeros(B, dtype=torch.float64, device=dev)
    bs = B if batch_size is None else max(1, int(batch_size))
    st = E.stream_handle()
    # the per-SV paths visit every window of both alleles (explain/impact.py:110-120)
    W = np.clip(np.maximum(rl, al) - kp.l + 1, 0, None)
    model._plan_lmer_table(float(2 * W.sum()))
    table = model._lmer_table_for()
    for s in range(0, B, bs):
        e = min(B, s + bs)
        n = e - s
        codes = np.concatenate([ref[s:e], alt[s:e]])
        lens = np.concatenate([rl[s:e], al[s:e]])
        xs = E.DeviceSeqs(codes, lens, dev, rc=kp.rc)
        if kp.normalize:
            rs = E.DeviceSeqs(ref[s:e], rl[s:e], dev, rc=kp.rc)
            dref = E.self_kraw(kp, rs)
        else:
            dref = torch.ones(n, dtype=torch.float64, device=dev)
        if table is not None:
            # linear in the SVs: the table's alt minus ref sums over the
            # reference allele's norm (explain/impact.py:139-148)
            raw = torch.empty(2 * n, dtype=torch.float64, device=dev)
            E.table_rows(kp, xs, table, 0.0, raw, normalize=False)
            dk = raw[n:] - raw[:n]
            if kp.normalize:
                dk = dk / torch.sqrt(torch.clamp(dref, min=1e-10))
            out[s:e] = dk
            continue
        hist = E.sv_hist(sv, xs, 0, 2 * n, use_index=model.use_lmer_index)
        _lib.call("gkm_combine_impact", E._p(hist), n, sv.S, kp.nm, E._p(kp.tab), E._p(dref),
                  E._p(sv.ds), E._p(sv.coef), int(kp.normalize), int(kp.rbf), kp.gamma,
                  E._p(out[s:e]), st)
    return nan_rows(out[:, None], short)
