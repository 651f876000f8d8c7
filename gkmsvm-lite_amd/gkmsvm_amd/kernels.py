"""Gapped k-mer kernel definitions: the pair weight h[m] by mismatch count for
LS-GKM's -t 0..6, the window weights of -t 4/-t 5 and the RBF finish.

Class names, parameters and aliases follow gkmsvm-lite (kernels/direct.py:209-253,
kernels/esttrunc.py:10-110, kernels/weighted.py:389-502, kernels/rbf.py:10-57,
svm.py:64-152).  The tables are host-side float64 and travel to the GPU as
kernel arguments; the window weights travel as a table by distance from the
sequence centre.
"""

from __future__ import annotations

from math import comb

import numpy as np

MAX_L = 15  # a window's 2-bit code fits in 30 bits; LS-GKM allows l <= 12


class GkmKernel:
    """Base: window length l, informative positions k, RC and normalisation."""

    def __init__(self, l: int, k: int, *, normalize: bool = True, include_rc: bool = True):
        if k > l:
            raise ValueError(f"k ({k}) must be <= l ({l})")
        if l < 1:
            raise ValueError(f"l must be >= 1, got {l}")
        if k < 1:
            raise ValueError(f"k must be >= 1, got {k}")
        self.l = l
        self.k = k
        self.normalize = normalize
        self.include_rc = include_rc

    @property
    def base(self) -> "GkmKernel":
        return self

    gamma = None  # RBF kernels set it

    def finish(self, k_norm):
        return k_norm

    def window_weights(self, W: int):
        return None

    def weight_table(self, max_windows: int):
        """uint8 [max_windows // 2 + 1]: window weight by distance from the
        centre, or None when every window counts once."""
        return None


class DirectGkmKernel(GkmKernel):
    """-t 0 (gkm_cnt): h[m] = C(l - m, k) for l - m >= k; pairs with more than
    d mismatches count 0 (d None: every pair)."""

    def __init__(self, l: int, k: int, *, d: int | None = None, normalize: bool = True,
                 include_rc: bool = True):
        if l > MAX_L:
            raise ValueError(f"l must be <= {MAX_L}, got {l}")
        super().__init__(l, k, normalize=normalize, include_rc=include_rc)
        self.d = d
        self._mismatch_table = np.array(
            [comb(l - m, k) if l - m >= k and (d is None or m <= d) else 0
             for m in range(l + 1)],
            dtype=np.float64,
        )
        self._min_matches = self._get_min_matches()

    def _get_min_matches(self) -> int:
        nz = np.nonzero(self._mismatch_table)[0]
        return self.l - int(nz.max()) if len(nz) else self.l + 1

    @property
    def dmax(self) -> int:
        """Most mismatches with a non-zero pair weight (-1: none)."""
        return self.l - self._min_matches

    def h_padded(self) -> np.ndarray:
        """h[0..l] plus a trailing zero (l + 2 entries)."""
        out = np.zeros(self.l + 2)
        out[: self.l + 1] = self._mismatch_table
        return out


def estimated_lmer_table(l: int, k: int, truncate: bool) -> np.ndarray:
    """Pair weights of LS-GKM's estimated l-mer kernel by mismatch count
    (calc_gkm_kernel_lmerest_wt in LS-GKM's libsvm_gkm.c; gkmsvm-lite
    kernels/esttrunc.py:40-110).

    wm[i]: the weight of a gapped k-mer pair differing at i of its k positions,
    from a two-row recursion over window positions; kern[m] sums those over the
    gapped k-mers two windows with m mismatches share; the result h[m]
    convolves kern with itself over the base-change patterns between windows.
    """
    b = 4
    rows = [[1.0] * (k + 1) for _ in range(k + 1)]
    nxt = [[1.0] * (k + 1) for _ in range(k + 1)]
    for pos in range(1, l + 1):
        for kk in range(1, k + 1):
            nxt[kk][0] = rows[kk][0] + (b - 1) * rows[kk - 1][0]
            for j in range(1, kk + 1):
                nxt[kk][j] = (nxt[kk - 1][j - 1] * (kk - pos)) / kk
        rows, nxt = nxt, rows
    total = comb(l, k) * b**l
    wm = [rows[k][i] / total for i in range(k + 1)]
    kern = []
    for m in range(l + 1):
        acc = 0.0
        for i in range(min(m, k) + 1):
            if l - m >= k - i:
                acc += wm[i] * comb(l - m, k - i) * comb(m, i)
        kern.append(acc)
    if truncate:
        cut = next((i for i, v in enumerate(kern) if v < 1e-50), len(kern))
        kern = [v if i < cut else 0.0 for i, v in enumerate(kern)]
    h = np.zeros(l + 1)
    for m in range(l + 1):
        acc = 0.0
        for m1, k1 in enumerate(kern):
            if k1 == 0.0:
                continue
            for m2, k2 in enumerate(kern):
                if k2 == 0.0:
                    continue
                for t in range(m + 1):
                    r = m1 + m2 - 2 * t - l + m
                    diff = m1 - t
                    if r < 0 or diff > l - m or r > diff:
                        continue
                    acc += (comb(m, t) * comb(l - m, diff) * comb(diff, r) * (b - 1) ** t
                            * (b - 2) ** r) * k1 * k2
        h[l - m] = acc
    return h


class EstTruncGkmKernel(DirectGkmKernel):
    """-t 2 (gkm_esttrunc, truncate=True) and -t 1 (gkm_estfull, truncate=False):
    the estimated l-mer kernel, pairs past d mismatches zeroed."""

    def __init__(self, l: int, k: int, *, d: int = 3, normalize: bool = True,
                 include_rc: bool = True, truncate: bool = True):
        super().__init__(l, k, normalize=normalize, include_rc=include_rc)
        self.d = d
        table = estimated_lmer_table(l, k, truncate)
        if d < l:
            table[d + 1:] = 0.0
        self._mismatch_table = table
        self._min_matches = self._get_min_matches()


def window_weights(W: int, M: int, H: float) -> np.ndarray:
    """[W] uint8: window i of a W-window sequence weighs
    min(M, floor(M 2^(-|W//2 - i| / H) + 1)) (kernels/weighted.py:38-44)."""
    return distance_weights(W // 2 + 1, M, H)[np.abs(W // 2 - np.arange(W))]


def distance_weights(n: int, M: int, H: float) -> np.ndarray:
    """[n] uint8 window weight at distance 0..n-1 from the centre, evaluated with
    the same float64 expression as the reference, element by element."""
    dist = np.arange(n)
    wt = np.floor(M * np.exp(-1 * np.log(2) * dist / H) + 1)
    return np.minimum(wt, M).astype(np.uint8)


class CenterWeightedGkmKernel(EstTruncGkmKernel):
    """-t 4 (wgkm): the -t 2 kernel with integer window weights, M at the
    centre, halving every H windows, at least 1."""

    def __init__(self, l: int, k: int, *, d: int = 3, M: int = 50, H: float = 50.0,
                 normalize: bool = True, include_rc: bool = True):
        if not 1 <= M <= 255:
            raise ValueError(f"M must be in 1..255 (LS-GKM stores weights as uint8), got {M}")
        if H <= 0:
            raise ValueError(f"H must be positive, got {H}")
        super().__init__(l, k, d=d, normalize=normalize, include_rc=include_rc)
        self.M = int(M)
        self.H = float(H)

    def window_weights(self, W: int) -> np.ndarray:
        return window_weights(W, self.M, self.H)

    def weight_table(self, max_windows: int):
        return distance_weights(max(1, max_windows) // 2 + 1, self.M, self.H)


class RbfKernel(GkmKernel):
    """RBF on a normalised base kernel: K = exp(gamma (K_norm - 1))."""

    def __init__(self, base: GkmKernel, gamma: float):
        super().__init__(base.l, base.k, normalize=True, include_rc=base.include_rc)
        self._base = base
        self.gamma = float(gamma)

    @property
    def base(self) -> GkmKernel:
        return self._base

    def finish(self, k_norm):
        import torch

        if isinstance(k_norm, torch.Tensor):
            return torch.exp(self.gamma * (k_norm - 1))
        return np.exp(self.gamma * (k_norm - 1))


class RbfGkmKernel(RbfKernel):
    """-t 3 (gkmrbf): RBF on the -t 2 kernel."""

    def __init__(self, l: int, k: int, *, d: int = 3, gamma: float = 1.0,
                 normalize: bool = True, include_rc: bool = True):
        super().__init__(EstTruncGkmKernel(l, k, d=d, include_rc=include_rc), gamma)
        self.d = d


class CenterWeightedRbfGkmKernel(RbfKernel):
    """-t 5 (wgkmrbf): RBF on the -t 4 kernel."""

    def __init__(self, l: int, k: int, *, d: int = 3, M: int = 50, H: float = 50.0,
                 gamma: float = 1.0, normalize: bool = True, include_rc: bool = True):
        super().__init__(
            CenterWeightedGkmKernel(l, k, d=d, M=M, H=H, include_rc=include_rc), gamma
        )
        self.d = d
        self.M = M
        self.H = H


def _args(p, d=3, **extra):
    return dict(l=p["L"], k=p["k"], d=p.get("d", d), normalize=True,
                include_rc=p.get("include_rc", True), **extra)


def _gamma(p):
    return {"gamma": p.get("gamma", 1.0)}


def _weights(p):
    return {"M": p.get("M", 50), "H": p.get("H", 50.0)}


KERNEL_BUILDERS = {
    "gkm_cnt": lambda p: DirectGkmKernel(**_args(p, d=None)),
    "gkm_esttrunc": lambda p: EstTruncGkmKernel(**_args(p, truncate=True)),
    "gkm_estfull": lambda p: EstTruncGkmKernel(**_args(p, truncate=False)),
    "gkmrbf": lambda p: RbfGkmKernel(**_args(p, **_gamma(p))),
    "wgkm": lambda p: CenterWeightedGkmKernel(**_args(p, **_weights(p))),
    "wgkmrbf": lambda p: CenterWeightedRbfGkmKernel(**_args(p, **_weights(p), **_gamma(p))),
    "gkm_cnt_rbf": lambda p: RbfKernel(DirectGkmKernel(**_args(p)), **_gamma(p)),
}

KERNEL_ALIASES: dict = {
    "direct": "gkm_cnt",
    "estimated": "gkm_esttrunc",
    "estimated_full": "gkm_estfull",
    "rbf": "gkmrbf",
    "weighted": "wgkm",
    "weighted_rbf": "wgkmrbf",
    "direct_rbf": "gkm_cnt_rbf",
    0: "gkm_cnt",
    1: "gkm_estfull",
    2: "gkm_esttrunc",
    3: "gkmrbf",
    4: "wgkm",
    5: "wgkmrbf",
    6: "gkm_cnt_rbf",
}


def resolve_kernel_type(kernel_type) -> str:
    """Internal name of a kernel type given as a name, an alias or an LS-GKM
    ``-t`` integer (svm.py:126-152)."""
    if isinstance(kernel_type, str) and kernel_type.strip().isdigit():
        kernel_type = int(kernel_type)
    if isinstance(kernel_type, (int, np.integer)):
        name = KERNEL_ALIASES.get(int(kernel_type))
        if name is None:
            raise ValueError(
                f"Unknown kernel type integer {kernel_type}. "
                "Valid: 0-6 (LS-GKM -t flag; 6 from the kundajelab fork)"
            )
        return name
    if kernel_type in KERNEL_BUILDERS:
        return kernel_type
    name = KERNEL_ALIASES.get(kernel_type)
    if name is not None:
        return name
    raise NotImplementedError(
        f"Unknown kernel type {kernel_type!r}. Internal names: {', '.join(KERNEL_BUILDERS)}"
    )


def build_kernel(kernel_type, params: dict) -> GkmKernel:
    return KERNEL_BUILDERS[resolve_kernel_type(kernel_type)](params)
