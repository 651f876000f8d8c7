"""CPU oracle for the gapped k-mer hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, the smoke check and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker (or as the timed
CPU baseline); the product never calls it.

It restates gkmsvm-lite's algorithm (citations are ``path:line`` relative to
the reference's ``src/gkmsvm/``): the exact integer parts (packing, mismatch
histograms, self-kernels, the per-SV explain loop, the l-mer index) run in
plain C (``gkm_oracle.c``, built by ``oracle/Makefile``); the small
floating-point parts (h[m] tables, window weights, normalisation, RBF, the
alpha-weighted reduction, ISM by full recomputation) are numpy here.

Parity pins (tests/test_oracle_*.py): LS-GKM gkmpredict scores for -t 0..6
(fixed, N-bearing and ragged sets), gkmexplain modes 0/1/4/5, the ENCODE
10-SV gkmpredict scores, the hand values of tests/test_kernels/test_direct.py
(l=3 k=2, [330,120,36,8,1] and the d cut), a numpy one-hot mismatch count on
random inputs, brute force == l-mer index, completeness of importance, and
ISM = f(mutant) - f(x) by construction.
"""

from __future__ import annotations

import ctypes
import gzip
import os
import subprocess
from dataclasses import dataclass, field
from math import comb

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libgkm_oracle.so")
PAD = np.uint32(0xFFFFFFFF)

_lib = None


def build() -> str:
    """Compile gkm_oracle.c with gcc (oracle/Makefile); returns the .so path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def use_native_build() -> str:
    """Compile gkm_oracle.c for this host's CPU (-O3 -march=native, OpenMP)
    into a temporary directory and route every oracle call through it; for
    timing the CPU baseline on the machine it runs on.  Returns the .so path."""
    global _lib
    import tempfile

    out = os.path.join(tempfile.mkdtemp(prefix="gkm_oracle_native_"), "libgkm_oracle.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fPIC", "-fopenmp", "-std=c11", "-shared",
                    "-o", out, os.path.join(_HERE, "gkm_oracle.c"), "-lm"], check=True)
    _lib = _configure(ctypes.CDLL(out))
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = _configure(ctypes.CDLL(_SO))
    return _lib


def _configure(L):
    """Declare the C entry points' argument types on a loaded oracle library."""
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    I = ctypes.c_int
    D = ctypes.c_double
    L.orc_pack.argtypes = [P, P, I64, I64, I, P]
    L.orc_rc_windows.argtypes = [P, I64, I64, I, P]
    L.orc_pair_hist.argtypes = [P, P, I64, I64, P, P, P, P, I64, I64, I, P]
    L.orc_self_hist.argtypes = [P, P, P, P, I64, I64, I, P]
    L.orc_explain.argtypes = [P, I64, I64, P, I64, I64, P, P, P, I, I, I64, I,
                              P, P, P]
    L.orc_neighbor_count.argtypes = [I, I]
    L.orc_neighbor_count.restype = I64
    L.orc_neighbor_masks.argtypes = [I, I, P, P, P]
    L.orc_neighbor_masks.restype = I64
    L.orc_index_build.argtypes = [P, P, P, P, I64, I64, I, P, P, P]
    L.orc_index_build.restype = I64
    L.orc_index_reduce.argtypes = [P, P, I64, I64, P, P, P, P, P, I64, I, P, P,
                                   P, I64, I, D, P]
    L.orc_threads.restype = I
    return L


def _p(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# Codes, tables and weights
# ---------------------------------------------------------------------------

_CODE = np.zeros(256, dtype=np.uint8)  # every non-ACGT letter reads as A
for _i, _b in enumerate("ACGT"):
    _CODE[ord(_b)] = _i
    _CODE[ord(_b.lower())] = _i


def encode(seq: str) -> np.ndarray:
    """[L] uint8 codes; N and other letters as A (codec.py:79-108)."""
    return _CODE[np.frombuffer(seq.encode("ascii"), dtype=np.uint8)]


def batch(seqs) -> tuple[np.ndarray, np.ndarray]:
    """([n, Lmax] codes zero-padded, [n] lengths) from strings or code arrays."""
    arrs = [encode(s) if isinstance(s, str) else np.asarray(s, np.uint8) for s in seqs]
    lens = np.array([len(a) for a in arrs], dtype=np.int64)
    out = np.zeros((len(arrs), int(lens.max())), dtype=np.uint8)
    for i, a in enumerate(arrs):
        out[i, : len(a)] = a
    return out, lens


def est_table(l: int, k: int, truncate: bool) -> np.ndarray:
    """Estimated l-mer pair weights by mismatch count, kernels/esttrunc.py:40-110
    (LS-GKM calc_gkm_kernel_lmerest_wt)."""
    b = 4
    prev = [[1.0] * (k + 1) for _ in range(k + 1)]
    cur = [[1.0] * (k + 1) for _ in range(k + 1)]
    for iL in range(1, l + 1):
        for iK in range(1, k + 1):
            cur[iK][0] = prev[iK][0] + (b - 1) * prev[iK - 1][0]
            for jM in range(1, iK + 1):
                cur[iK][jM] = (cur[iK - 1][jM - 1] * (iK - iL)) / iK
        cur, prev = prev, cur
    nnorm = comb(l, k) * (b**l)
    wm = [prev[k][i] / nnorm for i in range(k + 1)]
    kern = []
    for m in range(l + 1):
        v = 0.0
        for i in range(min(m, k) + 1):
            if l - m >= k - i:
                v += wm[i] * comb(l - m, k - i) * comb(m, i)
        kern.append(v)
    if truncate:
        tr, live = [], True
        for v in kern:
            if v < 1e-50:
                live = False
            tr.append(v if live else 0.0)
        kern = tr
    res = [0.0] * (l + 1)
    for m in range(l + 1):
        w = 0.0
        for m1 in range(l + 1):
            if kern[m1] == 0.0:
                continue
            for m2 in range(l + 1):
                if kern[m2] == 0.0:
                    continue
                for t in range(min(m, l) + 1):
                    r = m1 + m2 - 2 * t - l + m
                    diff = m1 - t
                    if r < 0 or diff > l - m or r > diff:
                        continue
                    cc = (comb(m, t) * comb(l - m, diff) * comb(diff, r)
                          * ((b - 1) ** t) * ((b - 2) ** r))
                    w += cc * kern[m1] * kern[m2]
        res[l - m] = w
    return np.array(res, dtype=np.float64)


KTYPES = {0: "gkm_cnt", 1: "gkm_estfull", 2: "gkm_esttrunc", 3: "gkmrbf",
          4: "wgkm", 5: "wgkmrbf", 6: "gkm_cnt_rbf"}


@dataclass
class Kern:
    """Kernel parameters as svm.py:95-106 builds them from a params dict."""

    ktype: str
    l: int
    k: int
    d: int | None
    rc: bool = True
    gamma: float | None = None  # RBF kernels only
    M: int | None = None  # weighted kernels only
    H: float | None = None
    table: np.ndarray = field(default=None)

    @classmethod
    def make(cls, ktype, p: dict):
        if isinstance(ktype, int) or str(ktype).isdigit():
            ktype = KTYPES[int(ktype)]
        l, k, rc = p["L"], p["k"], p.get("include_rc", True)
        if ktype == "gkm_cnt":  # kernels/direct.py:222-240, no d by default
            d = p.get("d", None)
        else:
            d = p.get("d", 3)
        gamma = p.get("gamma", 1.0) if ktype in ("gkmrbf", "wgkmrbf", "gkm_cnt_rbf") else None
        M = p.get("M", 50) if ktype in ("wgkm", "wgkmrbf") else None
        H = float(p.get("H", 50.0)) if ktype in ("wgkm", "wgkmrbf") else None
        if ktype in ("gkm_cnt", "gkm_cnt_rbf"):
            table = np.array([comb(l - m, k) if l - m >= k and (d is None or m <= d)
                              else 0 for m in range(l + 1)], dtype=np.float64)
        else:
            table = est_table(l, k, ktype != "gkm_estfull")
            if d < l:
                table[d + 1:] = 0.0
        return cls(ktype, l, k, d, rc, gamma, M, H, table)

    @property
    def dmax(self) -> int:
        nz = np.nonzero(self.table)[0]
        return int(nz.max()) if len(nz) else -1

    @property
    def weighted(self) -> bool:
        return self.M is not None


def window_weights(W: int, M: int, H: float) -> np.ndarray:
    """kernels/weighted.py:38-44 (_window_weights), the same numpy expression."""
    center = W // 2
    dist = np.abs(center - np.arange(W))
    wt = np.floor(M * np.exp(-1 * np.log(2) * dist / H) + 1)
    return np.minimum(wt, M).astype(np.uint8)


def row_weights(kern: Kern, lengths, W: int, rc: bool = False):
    """kernels/packed.py:105-122: [n, W] uint8 (0 on PAD), reversed for RC rows."""
    if not kern.weighted:
        return None
    out = np.zeros((len(lengths), W), dtype=np.uint8)
    for r, n in enumerate(lengths):
        nw = int(n) - kern.l + 1
        if nw > 0:
            out[r, :nw] = window_weights(nw, kern.M, kern.H)
    return np.ascontiguousarray(out[:, ::-1]) if rc else out


# ---------------------------------------------------------------------------
# Packing and histograms (C)
# ---------------------------------------------------------------------------


def pack(codes, lengths, l):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    lengths = np.ascontiguousarray(lengths, dtype=np.int64)
    n, Lmax = codes.shape
    W = Lmax - l + 1
    out = np.empty((n, W), dtype=np.uint32)
    lib().orc_pack(_p(codes), _p(lengths), n, Lmax, l, _p(out))
    return out


def rc_windows(fwd, l):
    fwd = np.ascontiguousarray(fwd, dtype=np.uint32)
    out = np.empty_like(fwd)
    lib().orc_rc_windows(_p(fwd), fwd.shape[0], fwd.shape[1], l, _p(out))
    return out


@dataclass
class Packed:
    fwd: np.ndarray
    rc: np.ndarray | None
    w: np.ndarray | None
    w_rc: np.ndarray | None
    lengths: np.ndarray


def pack_both(kern: Kern, codes, lengths) -> Packed:
    """kernels/packed.py:171-187: forward windows, RC windows from them."""
    fwd = pack(codes, lengths, kern.l)
    rc = rc_windows(fwd, kern.l) if kern.rc else None
    W = fwd.shape[1]
    w = row_weights(kern, lengths, W)
    w_rc = row_weights(kern, lengths, W, rc=True) if kern.rc else None
    return Packed(fwd, rc, w, w_rc, np.asarray(lengths, dtype=np.int64))


def pair_hist(kern: Kern, px: Packed, py: Packed) -> np.ndarray:
    """[B, S, dmax+1] uint64 histograms of x's forward windows against y's
    forward (and RC) windows."""
    dm = kern.dmax
    B, Wx = px.fwd.shape
    S, Wy = py.fwd.shape
    out = np.empty((B, S, dm + 1), dtype=np.uint64)
    c = np.ascontiguousarray
    lib().orc_pair_hist(
        _p(c(px.fwd)), _p(None if px.w is None else c(px.w)), B, Wx,
        _p(c(py.fwd)), _p(None if py.w is None else c(py.w)),
        _p(None if py.rc is None else c(py.rc)),
        _p(None if py.w_rc is None else c(py.w_rc)), S, Wy, dm, _p(out))
    return out


def self_hist(kern: Kern, px: Packed) -> np.ndarray:
    dm = kern.dmax
    B, W = px.fwd.shape
    out = np.empty((B, dm + 1), dtype=np.uint64)
    c = np.ascontiguousarray
    lib().orc_self_hist(_p(c(px.fwd)), _p(None if px.w is None else c(px.w)),
                        _p(None if px.rc is None else c(px.rc)),
                        _p(None if px.w_rc is None else c(px.w_rc)), B, W, dm, _p(out))
    return out


def combine(kern: Kern, hist: np.ndarray) -> np.ndarray:
    """K_raw = sum_m table[m] N[m] (float64)."""
    h = kern.table[: hist.shape[-1]]
    return (hist.astype(np.float64) * h).sum(axis=-1)


# ---------------------------------------------------------------------------
# Model-level oracles
# ---------------------------------------------------------------------------


@dataclass
class Model:
    kern: Kern
    sv_codes: np.ndarray
    sv_lengths: np.ndarray
    coef: np.ndarray
    bias: float

    def __post_init__(self):
        self.sv = pack_both(self.kern, self.sv_codes, self.sv_lengths)
        self.sv_diag = combine(self.kern, self_hist(self.kern, self.sv))


def _finish(kern, k):
    return np.exp(kern.gamma * (k - 1.0)) if kern.gamma is not None else k


def raw(model: Model, codes, lengths):
    """K_raw [B, S] and K_raw(x, x) [B]."""
    px = pack_both(model.kern, codes, lengths)
    return (combine(model.kern, pair_hist(model.kern, px, model.sv)),
            combine(model.kern, self_hist(model.kern, px)), px)


def score(model: Model, codes, lengths) -> np.ndarray:
    """[B] decision values, svm.py:43-61 (_normalized_dot) + bias; NaN for rows
    shorter than l (inputs.py:112-134)."""
    kr, dx, _ = raw(model, codes, lengths)
    k = kr / np.maximum(np.sqrt(dx[:, None] * model.sv_diag[None, :]), 1e-10)
    out = (model.coef[None, :] * _finish(model.kern, k)).sum(axis=1) + model.bias
    return np.where(np.asarray(lengths) < model.kern.l, np.nan, out)


def gram(kern: Kern, codes, lengths, codes_y=None, lengths_y=None) -> np.ndarray:
    """Normalised, finished Gram, gram.py:21-106 semantics."""
    px = pack_both(kern, codes, lengths)
    py = px if codes_y is None else pack_both(kern, codes_y, lengths_y)
    kr = combine(kern, pair_hist(kern, px, py))
    dx = combine(kern, self_hist(kern, px))
    dy = dx if codes_y is None else combine(kern, self_hist(kern, py))
    k = kr / np.maximum(np.sqrt(dx[:, None] * dy[None, :]), 1e-10)
    return _finish(kern, k)


def explain_tables(kern: Kern, perturbation: bool):
    """explain/weights.py:9-57: (table_a, table_b, min_matches)."""
    l, h = kern.l, kern.table
    d = l if kern.d is None else kern.d
    a = np.zeros(l + 2)
    b = np.zeros(l + 2)
    if perturbation:
        for m in range(l + 1):
            now = h[m]
            more = h[m + 1] if m + 1 < len(h) else 0.0
            a[m] = more - now
            if m >= 1:
                b[m] = h[m - 1] - now
    else:
        for m in range(min(d, l) + 1):
            if l - m > 0:
                a[m] = h[m] / (l - m)
        for m in range(1, min(d, l) + 1):
            b[m] = h[m - 1] / (l - m + 1)
    nz = [m for m in range(l + 1) if a[m] != 0 or b[m] != 0]
    return a, b, l - max(nz, default=-1)


def rbf_scale(k_norm, gamma):
    """explain/common.py:27-32 (_rbf_scale)."""
    pos = k_norm > 0
    safe = np.where(pos, k_norm, 1.0)
    return np.where(pos, (np.exp(gamma * (safe - 1)) - np.exp(-gamma)) / safe, 0.0)


def explain(model: Model, codes, lengths, mode: str) -> np.ndarray:
    """GkmExplain [B, 4, L] via the per-SV loop, explain/core.py:116-248 (mode
    importance = hypothetical x one-hot, explain/core.py:119-124), masked as
    inputs.py:137-145."""
    kern = model.kern
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    lengths = np.asarray(lengths, dtype=np.int64)
    B, L = codes.shape
    if mode == "importance":
        hyp = explain(model, codes, lengths, "hypothetical")
        onehot = np.zeros((B, 4, L))
        onehot[np.arange(B)[:, None], codes, np.arange(L)[None, :]] = 1.0
        return hyp * onehot
    pert = mode == "perturbation"
    ta, tb, mm = explain_tables(kern, pert)
    kr, dx, px = raw(model, codes, lengths)
    ec = model.coef[None, :] / np.maximum(np.sqrt(dx[:, None] * model.sv_diag[None, :]), 1e-10)
    if kern.gamma is not None:
        ec = ec * rbf_scale(kr / np.maximum(np.sqrt(dx[:, None] * model.sv_diag[None, :]), 1e-10),
                            kern.gamma)
    sv = model.sv
    py = sv.fwd if sv.rc is None else np.concatenate([sv.fwd, sv.rc], axis=1)
    wx = wy = None
    if kern.weighted:
        wx = px.w.astype(np.float64)
        wy = (sv.w if sv.rc is None else np.concatenate([sv.w, sv.w_rc], axis=1)).astype(np.float64)
        wx, wy = np.ascontiguousarray(wx), np.ascontiguousarray(wy)
    out = np.zeros((B, 4, L))
    py = np.ascontiguousarray(py)
    ec = np.ascontiguousarray(ec)
    lib().orc_explain(_p(np.ascontiguousarray(px.fwd)), B, px.fwd.shape[1], _p(py), py.shape[0],
                      py.shape[1], _p(ta), _p(tb), _p(ec), kern.l, mm, L, int(pert),
                      _p(wx), _p(wy), _p(out))
    return _per_position(out, lengths, kern.l)


def _per_position(out, lengths, l):
    inside = np.arange(out.shape[2])[None, None, :] < np.asarray(lengths)[:, None, None]
    out = np.where(inside, out, 0.0)
    return np.where((np.asarray(lengths) < l)[:, None, None], np.nan, out)


def ism(model: Model, codes, lengths) -> np.ndarray:
    """[B, 4, L] f(mutant) - f(x) for every single-base substitution, by scoring
    every mutant from scratch (the definition ism.py:21-64 computes by window
    deltas; tests/test_ism.py:33-106 checks the same identity)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    lengths = np.asarray(lengths, dtype=np.int64)
    B, L = codes.shape
    ref = score(model, codes, lengths)
    mut = np.repeat(codes[:, None, None, :], 4, axis=1).repeat(L, axis=2)  # [B,4,L,L]
    idx = np.arange(L)
    for v in range(4):
        mut[:, v, idx, idx] = v
    mlen = np.repeat(lengths, 4 * L)
    f = score(model, mut.reshape(B * 4 * L, L), mlen).reshape(B, 4, L)
    return _per_position(f - ref[:, None, None], lengths, model.kern.l)


def impact(model: Model, ref_codes, ref_len, alt_codes, alt_len) -> np.ndarray:
    """[n] gkmexplain -m 5 mutation impact, explain/impact.py:22-48 (_impact_cpu)."""
    kern = model.kern
    kr, dr, _ = raw(model, ref_codes, ref_len)
    ka, _, _ = raw(model, alt_codes, alt_len)
    z = np.maximum(np.sqrt(dr[:, None] * model.sv_diag[None, :]), 1e-10)
    k = kr / z
    dk = (ka - kr) / z
    if kern.gamma is not None:
        scale = np.where(k > 0, (np.exp(kern.gamma * (k - 1.0)) - np.exp(-kern.gamma))
                         / np.where(k > 0, k, 1.0), 0.0)
        dk = dk * scale
    out = (model.coef[None, :] * dk).sum(axis=1)
    short = (np.asarray(ref_len) < kern.l) | (np.asarray(alt_len) < kern.l)
    return np.where(short, np.nan, out)


# ---------------------------------------------------------------------------
# Exact l-mer tables (lmer_table.py, explain/table.py) and DeltaSVM
# ---------------------------------------------------------------------------


def lmer_weights(model: Model) -> np.ndarray:
    """V [4^l] (lmer_table.py:35-57, sv_lmer_weights): per l-mer, the sum over
    SV windows (forward and RC, each with its window weight) of
    coef / sqrt(max(K_raw(sv, sv), 1e-10)) (svm.py:497-502)."""
    kern = model.kern
    n = 4 ** kern.l
    c = model.coef / np.sqrt(np.clip(model.sv_diag, 1e-10, None))
    sv = model.sv
    V = np.zeros(n)
    for packed, w in ((sv.fwd, sv.w), (sv.rc, sv.w_rc)):
        if packed is None:
            continue
        wt = c[:, None] if w is None else c[:, None] * w.astype(np.float64)
        keys = packed.ravel().astype(np.int64)
        wt = np.broadcast_to(wt, packed.shape).ravel()
        keep = keys < n
        V += np.bincount(keys[keep], weights=wt[keep], minlength=n)
    return V


def mask_plan(kern: Kern):
    """kernels/masked.py:44-74 (mask_plan): [(j, a_j, positions)] with
    h(m) = sum_t a_t C(l - m, l - dmax + t), or None when it does not apply."""
    from itertools import combinations

    l, dmax = kern.l, kern.dmax
    min_matches = l - dmax
    if dmax < 0 or min_matches < 1:
        return None
    js = list(range(l - dmax, l + 1))
    A = np.array([[comb(l - m, j) for j in js] for m in range(dmax + 1)], dtype=np.float64)
    a = np.linalg.solve(A, kern.table[: dmax + 1])
    tol = 1e-12 * np.abs(a).max()
    return [(j, float(c), list(combinations(range(l), j))) for j, c in zip(js, a)
            if abs(c) > tol]


def lmer_table(model: Model) -> np.ndarray:
    """w [4^l], w(u) = sum_sv c_sv K_raw(u, sv), by the masked marginals of
    lmer_table.py:60-98 (build_lmer_table; without the RC pairing shortcut)."""
    kern = model.kern
    l = kern.l
    plan = mask_plan(kern)
    if plan is None:
        raise ValueError("no mask plan for this kernel")
    V = lmer_weights(model).reshape((4,) * l)  # axis a holds position l - 1 - a
    w = np.zeros_like(V)
    for _, coef, positions in plan:
        for S in positions:
            keep = {l - 1 - int(p) for p in S}
            drop = tuple(a for a in range(l) if a not in keep)
            w += coef * (V.sum(axis=drop, keepdims=True) if drop else V)
    return w.ravel()


def digit_mismatches(l: int) -> np.ndarray:
    """[4^l, 4^l] m(u, v) for every pair of packed l-mers (small l only)."""
    z = np.arange(4 ** l, dtype=np.int64)
    x = z[:, None] ^ z[None, :]
    m = np.zeros(x.shape, dtype=np.int64)
    for k in range(l):
        m += ((x >> (2 * k)) & 3) != 0
    return m


def lmer_table_brute(model: Model) -> np.ndarray:
    """w(u) = sum_v V[v] h[m(u, v)] over all pairs of l-mers (l <= 6)."""
    l = model.kern.l
    h = np.concatenate([model.kern.table, np.zeros(2)])
    return h[digit_mismatches(l)] @ lmer_weights(model)


def explain_table_brute(model: Model, perturbation: bool) -> np.ndarray:
    """A [4^l, l, 4]: A[u, k, d] = sum_v V[v] [(u ^ v)_k == d] t_d(m(u, v)),
    t_0 = table_a, t_1..3 = table_b (the definition explain/table.py:14-26
    states; l <= 5)."""
    l = model.kern.l
    ta, tb, _ = explain_tables(model.kern, perturbation)
    V = lmer_weights(model)
    m = digit_mismatches(l)
    z = np.arange(4 ** l, dtype=np.int64)
    x = z[:, None] ^ z[None, :]
    A = np.zeros((4 ** l, l, 4))
    for k in range(l):
        xk = (x >> (2 * k)) & 3
        for d in range(4):
            t = ta if d == 0 else tb
            A[:, k, d] = np.where(xk == d, t[m], 0.0) @ V
    return A


def score_with_table(model: Model, w: np.ndarray, codes, lengths) -> np.ndarray:
    """[B] f(x) = sum_i wt_i w(x_i) / sqrt(max(K_raw(x, x), 1e-10)) + b
    (svm.py:559-582, _table_raw / _score_lmer_table); NaN for short rows."""
    kern = model.kern
    px = pack_both(kern, codes, lengths)
    dx = combine(kern, self_hist(kern, px))
    keys = px.fwd.astype(np.int64)
    pad = keys >= 4 ** kern.l
    vals = np.where(pad, 0.0, w[np.where(pad, 0, keys)])
    if px.w is not None:
        vals = vals * px.w
    raw_ = vals.sum(axis=1)
    out = raw_ / np.sqrt(np.clip(dx, 1e-10, None)) + model.bias
    return np.where(np.asarray(lengths) < kern.l, np.nan, out)


def deltasvm_weights(model: Model, w: np.ndarray, query_norm: bool) -> np.ndarray:
    """svm.py:667-705: the table in DeltaSVM order (first base most
    significant), times the one-window weight, divided by sqrt(K_raw(u, u))
    with query_norm."""
    kern = model.kern
    l = kern.l
    out = w.reshape((4,) * l).transpose(tuple(range(l - 1, -1, -1))).ravel().copy()
    wt0 = 1.0 if not kern.weighted else float(window_weights(1, kern.M, kern.H)[0])
    out *= wt0
    if query_norm:
        idx = np.arange(4 ** l)
        digits = (idx[:, None] // 4 ** np.arange(l - 1, -1, -1)) % 4
        agree = digits == 3 - digits[:, ::-1]
        h = kern.table
        selfk = np.full(len(idx), h[0])
        if kern.rc:
            selfk = selfk + h[(~agree).sum(axis=1)]
        out = out / np.sqrt(np.clip(selfk * wt0 ** 2, 1e-10, None))
    return out


def deltasvm_score(weights, l: int, k: int, codes, lengths, include_rc=True,
                   bias=0.0) -> np.ndarray:
    """[B] DeltaSVM scores, deltasvm.py:22-44 (_score_strand_numba) per strand
    at each row's own length, the RC strand 3 - x reversed (deltasvm.py:107-147)."""
    from itertools import combinations

    combos = list(combinations(range(l), k))
    out = np.zeros(len(lengths))
    for b, n in enumerate(np.asarray(lengths)):
        x = np.asarray(codes[b, :n], dtype=np.int64)
        strands = [x] + ([3 - x[::-1]] if include_rc else [])
        tot = 0.0
        for s in strands:
            acc = 0.0
            for i in range(len(s) - l + 1):
                for c in combos:
                    idx = 0
                    for j in c:
                        idx = idx * 4 + int(s[i + j])
                    acc += weights[idx]
            tot += acc
        out[b] = tot + bias
    return out


# ---------------------------------------------------------------------------
# The reference's CPU scoring algorithm (l-mer index + fused reduce)
# ---------------------------------------------------------------------------


class LmerIndex:
    """kernels/lmer_index.py:161-200 over the model's SVs."""

    def __init__(self, model: Model):
        kern = model.kern
        self.model = model
        l = kern.l
        dm = kern.dmax
        n = lib().orc_neighbor_count(l, dm)
        self.masks = np.empty(n, dtype=np.uint32)
        self.levels = np.empty(n, dtype=np.int32)
        n = lib().orc_neighbor_masks(l, dm, _p(np.ascontiguousarray(kern.table)),
                                     _p(self.masks), _p(self.levels))
        self.masks, self.levels = self.masks[:n], self.levels[:n]
        self.weights = np.ascontiguousarray(kern.table[self.levels])
        sv = model.sv
        S, W = sv.fwd.shape
        self.offsets = np.empty((1 << (2 * l)) + 1, dtype=np.int64)
        cap = S * W * (2 if sv.rc is not None else 1)
        self.ids = np.empty(cap, dtype=np.int32)
        self.vals = np.empty(cap, dtype=np.uint8) if kern.weighted else None
        c = np.ascontiguousarray
        ne = lib().orc_index_build(_p(c(sv.fwd)), _p(None if sv.w is None else c(sv.w)),
                                   _p(None if sv.rc is None else c(sv.rc)),
                                   _p(None if sv.w_rc is None else c(sv.w_rc)), S, W, l,
                                   _p(self.offsets), _p(self.ids), _p(self.vals))
        self.n_entries = int(ne)
        self.inv_ds = 1.0 / np.sqrt(np.clip(model.sv_diag, 1e-20, None))

    def score(self, codes, lengths) -> np.ndarray:
        """[B] decision values, kernels/lmer_index_cpu.py:71-120 + bias."""
        m = self.model
        kern = m.kern
        px = pack_both(kern, codes, lengths)
        dx = combine(kern, self_hist(kern, px))
        inv_dx = np.ascontiguousarray(1.0 / np.sqrt(np.clip(dx, 1e-20, None)))
        B, W = px.fwd.shape
        out = np.empty(B)
        rbf = kern.gamma is not None
        c = np.ascontiguousarray
        lib().orc_index_reduce(_p(c(px.fwd)), _p(None if px.w is None else c(px.w)), B, W,
                               _p(self.offsets), _p(self.ids), _p(self.vals),
                               _p(self.masks), _p(self.weights), len(self.masks), kern.l,
                               _p(inv_dx), _p(c(self.inv_ds)), _p(c(m.coef)), len(m.coef),
                               int(rbf), float(kern.gamma or 0.0), _p(out))
        out = out + m.bias
        return np.where(np.asarray(lengths) < kern.l, np.nan, out)


def threads() -> int:
    return int(lib().orc_threads())


# ---------------------------------------------------------------------------
# LS-GKM files (the fixtures' formats)
# ---------------------------------------------------------------------------


def _open(path):
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path)


def read_fasta(path):
    recs, name, seq = [], None, []
    with _open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith(">"):
                if name is not None:
                    recs.append((name, "".join(seq)))
                name, seq = line[1:].strip(), []
            else:
                seq.append(line)
    if name is not None:
        recs.append((name, "".join(seq)))
    return recs


def load_lsgkm(path) -> Model:
    """LS-GKM model text: header, 'SV', then '<coef> <sequence>' lines; bias =
    -rho (importers/lsgkm.py:37-97)."""
    head, coefs, seqs, in_sv = {}, [], [], False
    with _open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("SV"):
                in_sv = True
                continue
            if in_sv:
                c, s = line.split()[:2]
                coefs.append(float(c))
                seqs.append(s)
            else:
                kv = line.split(None, 1)
                if len(kv) == 2:
                    head[kv[0]] = kv[1]
    kt = head["kernel_type"]
    kt = int(kt) if kt.isdigit() else kt
    p = {"L": int(head["L"]), "k": int(head["k"]),
         "include_rc": int(head.get("norc", 0)) == 0}
    for key, cast in (("d", int), ("gamma", float), ("M", int), ("H", float)):
        if key in head:
            p[key] = cast(head[key])
    codes, lens = batch(seqs)
    return Model(Kern.make(kt, p), codes, lens, np.array(coefs),
                 -float(head["rho"].split()[0]))
