"""Exact l-mer tables of non-RBF models, built and held in HBM.

For -t 0/1/2/4 (no RBF) the decision value is linear in the query's windows
apart from its own normalisation (gkmsvm-lite lmer_table.py:1-27):

    f(x) = sum_i wt_i w(x_i) / sqrt(K_raw(x, x)) + b,   w(u) = sum_v V[v] h[m(u, v)],

V[v] = sum_s alpha_s / sqrt(K_raw(s, s)) x (window-weighted occurrences of the
l-mer v in SV s and its reverse complement).  Scoring then costs one lookup
per window plus the query's self-kernel, whatever the number of SVs.  The
GkmExplain table (explain/table.py:14-33) holds every window's [l, 4]
attribution, so an explanation is l lookups per position.

Both tables are XOR convolutions of V over the 4^l l-mers, built here with
the radix-4 Walsh-Hadamard transform on the GPU (gkm_wht4): the transform of
the pair weight depends only on the number of non-zero digits, so its
spectrum is a Krawtchouk sum evaluated exactly on the host.  V comes from a
CSR of the SV windows sorted by l-mer (entries in SV order), so every table
is deterministic.  Results equal the per-SV kernel path up to floating-point
summation order.
"""

from __future__ import annotations

from math import comb, fsum

import numpy as np
import torch

from . import _lib
from . import engine as E

# Largest l with a score table (lmer_table.py:31: 4^12 float64 = 128 MiB)
MAX_TABLE_L = 12
# Largest l with an explain table (explain/table.py:34-35: [4^12, 4, 12] f64 = 6.4 GB)
EXPLAIN_TABLE_MAX_L = 12
# Largest explain table "auto" builds: [4^12, 12, 4] float64 fits in 8 GiB, a
# small share of an MI355X's 288 GB of HBM (the reference stops at 2 GiB on the CPU)
EXPLAIN_TABLE_BYTES = 8 << 30

# GPU build costs for the "auto" decisions (MI355X): sorting the SV windows by
# l-mer, per window; the transforms and fills, per table entry
TABLE_SORT_S_PER_WINDOW = 3e-10
TABLE_S_PER_LMER = 2.5e-10
EXPLAIN_TABLE_S_PER_ENTRY = 1.2e-10


def applies(kp: E.KernelParams) -> bool:
    """Whether the model has a score table: no RBF and l <= MAX_TABLE_L."""
    return not kp.rbf and kp.l <= MAX_TABLE_L


def explain_applies(kp: E.KernelParams) -> bool:
    return not kp.rbf and kp.l <= EXPLAIN_TABLE_MAX_L


def krawtchouk(m: int, w: int, n: int) -> int:
    """K_m(w; n) for the alphabet of 4 letters: sum over z with m non-zero
    digits (of n) of the Walsh character of y with w non-zero digits."""
    return sum((-1) ** j * 3 ** (m - j) * comb(w, j) * comb(n - w, m - j) for j in range(m + 1))


def spectrum(values, n: int) -> np.ndarray:
    """lam[w] = sum_m values[m] K_m(w; n), w = 0..n: the Walsh-Hadamard
    transform of g(z) = values[m(z)] over n digits (exact integer Krawtchouk
    values, one rounding per product, an exactly rounded sum)."""
    vals = [float(values[m]) if m < len(values) else 0.0 for m in range(n + 1)]
    return np.array([fsum(vals[m] * krawtchouk(m, w, n) for m in range(n + 1))
                     for w in range(n + 1)], dtype=np.float64)


def sv_coefficients(sv: E.SVModel) -> torch.Tensor:
    """c_s = alpha_s / sqrt(max(K_raw(s, s), 1e-10)) (alpha_s without
    normalisation), in the SV model's order (svm.py:485-506)."""
    if sv.kp.normalize:
        return sv.coef / torch.sqrt(torch.clamp(sv.ds, min=1e-10))
    return sv.coef.clone()


def lmer_weights(sv: E.SVModel, stream=None) -> torch.Tensor:
    """V [4^l] float64 (lmer_table.py:35-57, explain/common.py:87-96)."""
    kp = sv.kp
    l = kp.l
    nl = 1 << (2 * l)
    seqs = sv.seqs
    dev = kp.device
    fwd, rcw = seqs.windows(l, stream)
    W = fwd.shape[1]
    blocks = [fwd] + ([rcw] if kp.rc else [])
    keys = torch.cat(blocks, dim=1).to(torch.int64) & 0xFFFFFFFF  # [S, nb*W], SV-major
    nw = keys.shape[1]
    flat = keys.reshape(-1)
    valid = flat < nl
    svid = torch.arange(sv.S, device=dev, dtype=torch.int32).repeat_interleave(nw)
    vals = None
    if kp.weighted:
        # a window's weight by its distance from its SV's centre; an RC window
        # carries the weight of the forward window it came from
        wt, _ = kp.wtab(seqs.lmax)
        ws = torch.from_numpy(np.clip(seqs.lengths - l + 1, 0, None)).to(dev)[:, None]
        j = torch.arange(W, device=dev)[None, :]
        d = [(ws // 2 - j).abs()] + ([(ws // 2 - (ws - 1 - j)).abs()] if kp.rc else [])
        dist = torch.cat(d, dim=1).reshape(-1).clamp(0, wt.numel() - 1)
        vals = wt[dist][valid].contiguous()
    flat = flat[valid]
    svid = svid[valid]
    order = torch.argsort(flat, stable=True)
    svid = svid[order].contiguous()
    if vals is not None:
        vals = vals[order].contiguous()
    counts = torch.bincount(flat, minlength=nl)
    if flat.numel() >= 2**31:
        raise MemoryError("over 2^31 SV windows")
    offsets = torch.zeros(nl + 1, dtype=torch.int64, device=dev)
    offsets[1:] = torch.cumsum(counts, 0)
    offsets = offsets.to(torch.int32)
    del keys, flat, order, counts
    c = sv_coefficients(sv)
    V = torch.empty(nl, dtype=torch.float64, device=dev)
    _lib.call("gkm_lmer_weights", E._p(offsets), E._p(svid), E._p(vals), E._p(c), l, E._p(V),
              E.stream_handle(stream))
    return V


def score_table(sv: E.SVModel, V: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """w [4^l] float64: w(u) = sum_v V[v] h[m(u, v)] (build_lmer_table,
    lmer_table.py:60-98)."""
    kp = sv.kp
    l = kp.l
    st = E.stream_handle(stream)
    f = (lmer_weights(sv, stream) if V is None else V).clone()
    lam = torch.from_numpy(spectrum(kp.h[: l + 1], l)).to(kp.device)
    _lib.call("gkm_wht4", E._p(f), l, 1, st)
    _lib.call("gkm_table_spectrum", E._p(f), l, E._p(lam), 4.0 ** -l, st)
    _lib.call("gkm_wht4", E._p(f), l, 1, st)
    return f


def explain_table(sv: E.SVModel, table_a, table_b, V: torch.Tensor | None = None,
                  stream=None) -> torch.Tensor:
    """A [4^l, l, 4] float64: A[u, k, d] = sum_v V[v] [(u ^ v)_k == d]
    t_d(m(u, v)), t_0 = table_a (the bases agree), t_1..3 = table_b
    (_build_explain_table, explain/table.py:91-113)."""
    kp = sv.kp
    l = kp.l
    st = E.stream_handle(stream)
    fv = (lmer_weights(sv, stream) if V is None else V).clone()
    _lib.call("gkm_wht4", E._p(fv), l, 1, st)
    ta = np.asarray(table_a, dtype=np.float64)
    tb = np.asarray(table_b, dtype=np.float64)
    # lam_d over the l - 1 other digits: t_d at m' + [d != 0] mismatches
    lam_a = torch.from_numpy(spectrum(ta[:l], l - 1)).to(kp.device)
    lam_b = torch.from_numpy(spectrum(tb[1 : l + 1], l - 1)).to(kp.device)
    A = torch.empty((1 << (2 * l), l, 4), dtype=torch.float64, device=kp.device)
    _lib.call("gkm_explain_fill", E._p(fv), l, E._p(lam_a), E._p(lam_b), 4.0 ** -l, E._p(A), st)
    del fv
    _lib.call("gkm_wht4", E._p(A), l, 4 * l, st)
    return A


def build_seconds(sv: E.SVModel) -> float:
    """Estimated GPU time to build the score table."""
    nb = 2 if sv.kp.rc else 1
    windows = float(np.clip(sv.seqs.lengths - sv.kp.l + 1, 0, None).sum()) * nb
    return windows * TABLE_SORT_S_PER_WINDOW + 4.0 ** sv.kp.l * TABLE_S_PER_LMER


def explain_build_seconds(sv: E.SVModel) -> float:
    return build_seconds(sv) + 4.0 ** sv.kp.l * 4 * sv.kp.l * EXPLAIN_TABLE_S_PER_ENTRY


def explain_table_bytes(l: int) -> int:
    return (1 << (2 * l)) * 4 * l * 8
