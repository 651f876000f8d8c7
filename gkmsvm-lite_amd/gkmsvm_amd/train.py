"""Training: train_gkmsvm (C-SVC) and train_gkmsvr (epsilon-SVR) with the
kernel on the GPU (gkmsvm-lite train.py:26-450, solver.py:13-392).

Two solvers, as the reference's:

* "libsvm": the N x N Gram is computed on the GPU (upper triangle, mirrored;
  compute_gram) and solved by scikit-learn's LIBSVM with a precomputed kernel.
* "smo": the column-cached SMO of gkm_smo.c (second-order working-set
  selection, LRU column cache, no N x N matrix): on a cache miss the solver
  asks for a batch of kernel columns K(i, :), which KernelColumns computes on
  the GPU -- the batch's sequences as rows against every training sequence
  (through the l-mer index or the pairwise kernel), normalised and finished
  by the FP64 combine -- so training is no longer bounded by host RAM for the
  Gram.

"auto" picks "libsvm" while the Gram fits max_gram_gb (default: a quarter of
the host's available memory), else "smo".
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _lib
from . import engine as E
from .gram import _gram_index, compute_gram
from .inputs import read_batch
from .kernels import KERNEL_BUILDERS, resolve_kernel_type
from .svm import GkmSVM

_SMO_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libgkm_smo.so")
_smo = None
_COLUMNS_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
# Columns computed per cache miss (the missing one plus the likeliest next
# ones), so the GPU works on batches
SMO_PREFETCH = 256


def smo_lib():
    """The SMO solver library (libgkm_smo.so; raises if not built)."""
    global _smo
    if _smo is None:
        if not os.path.exists(_SMO_PATH):
            raise ImportError(f"{_SMO_PATH} is not built: run gkmsvm-lite_amd/gkmsvm_amd/build_ext.py")
        lib = ctypes.CDLL(_SMO_PATH)
        P = ctypes.c_void_p
        lib.gkm_smo_solve.restype = ctypes.c_int
        lib.gkm_smo_solve.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_double,
                                      ctypes.c_int64, ctypes.c_int, ctypes.c_int, _COLUMNS_FN,
                                      P, P, P, P]
        _smo = lib
    return _smo


def smo_solve_columns(columns, N: int, y, p, C, *, tol: float = 1e-3,
                      max_iter: int = 10_000_000, cache_size: int | None = None,
                      prefetch: int = SMO_PREFETCH):
    """Solve min 1/2 a'Qa + p'a, y'a = 0, 0 <= a <= C (Q_ts = y_t y_s K(t mod N,
    s mod N)) with gkm_smo.c; columns(idx) returns K(idx, :) as a [len(idx),
    N] float64 array.  Returns (a [n], rho, stats)."""
    lib = smo_lib()
    y = np.ascontiguousarray(y, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    C = np.ascontiguousarray(np.broadcast_to(np.asarray(C, dtype=np.float64), y.shape))
    n = y.shape[0]
    errors: list[BaseException] = []

    def cb(count, idx_ptr, n_cols, out_ptr, _ctx):
        try:
            idx = np.ctypeslib.as_array(idx_ptr, shape=(count,)).copy()
            cols = np.ascontiguousarray(columns(idx), dtype=np.float64)
            if cols.shape != (count, n_cols):
                raise ValueError(f"columns returned {cols.shape}, expected {(count, n_cols)}")
            ctypes.memmove(out_ptr, cols.ctypes.data, cols.nbytes)
        except BaseException as exc:  # raised again once the solver returns
            errors.append(exc)
            nan = np.full((count, n_cols), np.nan)
            ctypes.memmove(out_ptr, nan.ctypes.data, nan.nbytes)

    fn = _COLUMNS_FN(cb)
    alpha = np.zeros(n, dtype=np.float64)
    rho = np.zeros(1, dtype=np.float64)
    stats = np.zeros(4, dtype=np.float64)
    cache = N if cache_size is None else max(2, min(N, int(cache_size)))
    rc = lib.gkm_smo_solve(n, N, y.ctypes.data, p.ctypes.data, C.ctypes.data, float(tol),
                           int(max_iter), cache, int(prefetch), fn, None, alpha.ctypes.data,
                           rho.ctypes.data, stats.ctypes.data)
    if errors:
        raise RuntimeError("kernel column computation failed during SMO") from errors[0]
    if rc < 0:
        raise MemoryError("SMO: out of host memory for the column cache")
    return alpha, float(rho[0]), {"iterations": int(stats[0]), "misses": int(stats[1]),
                                  "columns": int(stats[2]), "gap": float(stats[3]),
                                  "converged": rc == 0}


class KernelColumns:
    """Normalised, finished kernel columns K(i, :) over a training set held on
    the GPU (solver.py:13-88, KernelColumnCache): the training sequences as
    columns (sorted by length, with their self-kernels and, when it is the
    cheaper path, an l-mer index), a batch of them as rows per call."""

    def __init__(self, kernel, codes: np.ndarray, lengths: np.ndarray, device=None):
        dev = E.device_of(device)
        self.kp = E.KernelParams(kernel, dev)
        self.codes = np.ascontiguousarray(codes)
        self.lengths = np.asarray(lengths, dtype=np.int64)
        self.N = len(self.lengths)
        self.order = np.argsort(self.lengths, kind="stable")
        self.ys = E.DeviceSeqs(self.codes[self.order], self.lengths[self.order], dev,
                               rc=self.kp.rc)
        self.groups = torch.from_numpy(E.group_bounds(self.lengths[self.order])).to(dev)
        self.dy = E.self_kraw(self.kp, self.ys) if self.kp.normalize else None
        # a column costs one row: the index pays once a few batches go through it
        self.index = _gram_index(self.kp, np.repeat(self.lengths.mean(), self.N), self.ys)

    def __call__(self, idx) -> np.ndarray:
        idx = np.asarray(idx, dtype=np.int64)
        kp = self.kp
        xs = E.DeviceSeqs(self.codes[idx], self.lengths[idx], kp.device, rc=kp.rc)
        B = len(idx)
        if self.index is not None:
            hist = E.index_hist(kp, self.index, xs, 0, B, self.N)
        else:
            hist = E.cross_hist(kp, xs, 0, B, self.ys, self.groups)
        dx = E.self_kraw(kp, xs) if kp.normalize else None
        ones = torch.ones(max(B, self.N), dtype=torch.float64, device=kp.device)
        tile = torch.empty((B, self.N), dtype=torch.float64, device=kp.device)
        _lib.call("gkm_combine_gram", E._p(hist), B, self.N, kp.nm, E._p(kp.tab),
                  E._p(dx if dx is not None else ones), E._p(self.dy if self.dy is not None
                                                             else ones),
                  int(kp.normalize), int(kp.rbf), kp.gamma, E._p(tile), E.stream_handle())
        out = np.empty((B, self.N), dtype=np.float64)
        out[:, self.order] = tile.cpu().numpy()
        return out


def _codes(seqs):
    batch = read_batch(seqs if not isinstance(seqs, np.ndarray) or seqs.ndim != 1
                       else list(seqs))
    return batch.codes, batch.lengths


def _concat(parts):
    L = max(c.shape[1] for c, _ in parts)
    X = np.zeros((sum(c.shape[0] for c, _ in parts), L), dtype=np.uint8)
    r = 0
    for c, _ in parts:
        X[r:r + c.shape[0], : c.shape[1]] = c
        r += c.shape[0]
    return X, np.concatenate([n for _, n in parts])


def _build_kernel(kernel_type, l, k, d, gamma, M, H, include_rc):
    kernel_type = resolve_kernel_type(kernel_type)
    params = {"L": l, "k": k, "d": d, "include_rc": include_rc}
    if kernel_type in ("gkmrbf", "wgkmrbf", "gkm_cnt_rbf"):
        params["gamma"] = gamma
    if kernel_type in ("wgkm", "wgkmrbf"):
        params["M"] = 50 if M is None else M
        params["H"] = 50.0 if H is None else H
    return KERNEL_BUILDERS[kernel_type](params), kernel_type, params


def _available_bytes() -> int:
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    return int(line.split()[1]) * 1024
    except OSError:
        pass
    return 16 << 30


def _pick_solver(solver: str, N: int, max_gram_gb: float | None) -> str:
    if solver not in ("auto", "libsvm", "smo"):
        raise ValueError(f"Unknown solver {solver!r}. Use 'auto', 'libsvm' or 'smo' "
                         "('nystrom' is not part of this engine).")
    if solver != "auto":
        return solver
    limit = (max_gram_gb * 1e9) if max_gram_gb is not None else _available_bytes() / 4
    return "libsvm" if 8.0 * N * N <= limit else "smo"


def _model(X, lengths, coefs, bias, kernel_type, params, sv_chunk_size, device):
    keep = np.flatnonzero(coefs != 0)
    sv_len = lengths[keep]
    sv = X[keep][:, : int(sv_len.max())]
    return GkmSVM(sv, coefs[keep].astype(np.float64), float(bias), kernel_type, params,
                  sv_lengths=sv_len if (sv_len != sv.shape[1]).any() else None,
                  sv_chunk_size=sv_chunk_size, device=device)


def train_gkmsvm(pos_seqs, neg_seqs, *, kernel_type="estimated", l: int = 11, k: int = 7,
                 d: int = 3, C: float = 1.0, pos_weight: float = 1.0, gamma: float = 1.0,
                 M: int | None = None, H: float | None = None, include_rc: bool = True,
                 solver: str = "auto", cache_size: int | None = None, tol: float = 1e-3,
                 max_iter: int = 10_000_000, max_gram_gb: float | None = None,
                 sv_chunk_size: int | None = None, gram_chunk_size: int = 1000,
                 device: str = "rocm", verbose: bool = False, **_unused) -> GkmSVM:
    """Train a gapped k-mer SVM classifier on positive and negative sequences
    (train.py:26-128).  Positives get C * pos_weight (LS-GKM -w).  solver:
    "auto", "libsvm" (GPU Gram + scikit-learn's LIBSVM) or "smo" (column-
    cached SMO, columns on the GPU).  Returns a GkmSVM on the ROCm device."""
    pc, pl = _codes(pos_seqs)
    nc, nl = _codes(neg_seqs)
    X, lengths = _concat([(pc, pl), (nc, nl)])
    if (lengths < l).any():
        raise ValueError(f"training sequences must be at least l = {l} long")
    y = np.r_[np.ones(pc.shape[0]), -np.ones(nc.shape[0])]
    kernel, kernel_type, params = _build_kernel(kernel_type, l, k, d, gamma, M, H, include_rc)
    N = X.shape[0]
    solver = _pick_solver(solver, N, max_gram_gb)
    if solver == "libsvm":
        L = X.shape[1]
        gram = compute_gram(kernel, X, lengths=lengths if (lengths != L).any() else None,
                            chunk_size=gram_chunk_size, verbose=verbose, device=device)
        from sklearn.svm import SVC

        clf = SVC(C=C, kernel="precomputed", tol=tol,
                  class_weight=None if pos_weight == 1.0 else {1: pos_weight})
        clf.fit(gram, y)
        coefs = np.zeros(N)
        coefs[clf.support_] = clf.dual_coef_[0]
        bias = float(clf.intercept_[0])
    else:
        cols = KernelColumns(kernel, X, lengths, device)
        Cv = np.where(y > 0, C * pos_weight, C)
        alpha, rho, stats = smo_solve_columns(cols, N, y, -np.ones(N), Cv, tol=tol,
                                              max_iter=max_iter, cache_size=cache_size)
        if verbose:
            print(f"SMO: {stats}")
        coefs = y * alpha
        bias = -rho
    return _model(X, lengths, coefs, bias, kernel_type, params, sv_chunk_size, device)


def train_gkmsvr(seqs, labels, *, kernel_type="estimated", l: int = 11, k: int = 7, d: int = 3,
                 C: float = 1.0, epsilon: float = 0.1, gamma: float = 1.0, M: int | None = None,
                 H: float | None = None, include_rc: bool = True, solver: str = "auto",
                 cache_size: int | None = None, tol: float = 1e-3, max_iter: int = 10_000_000,
                 max_gram_gb: float | None = None, sv_chunk_size: int | None = None,
                 gram_chunk_size: int = 1000, device: str = "rocm", verbose: bool = False,
                 **_unused) -> GkmSVM:
    """Train a gapped k-mer epsilon-SVR on sequences and real-valued labels
    (train.py:131-234; LIBSVM's formulation, as lsgkm-svr's gkmtrain -y 3).
    solver as train_gkmsvm; SMO works on the 2N dual variables with one
    column per sequence."""
    X, lengths = _codes(seqs)
    z = np.asarray(labels, dtype=np.float64).reshape(-1)
    N = X.shape[0]
    if z.shape[0] != N:
        raise ValueError(f"{N} sequences but {z.shape[0]} labels")
    if (lengths < l).any():
        raise ValueError(f"training sequences must be at least l = {l} long")
    kernel, kernel_type, params = _build_kernel(kernel_type, l, k, d, gamma, M, H, include_rc)
    solver = _pick_solver(solver, N, max_gram_gb)
    if solver == "libsvm":
        L = X.shape[1]
        gram = compute_gram(kernel, X, lengths=lengths if (lengths != L).any() else None,
                            chunk_size=gram_chunk_size, verbose=verbose, device=device)
        from sklearn.svm import SVR

        reg = SVR(C=C, epsilon=epsilon, kernel="precomputed", tol=tol)
        reg.fit(gram, z)
        coefs = np.zeros(N)
        coefs[reg.support_] = reg.dual_coef_[0]
        bias = float(reg.intercept_[0])
    else:
        cols = KernelColumns(kernel, X, lengths, device)
        y = np.r_[np.ones(N), -np.ones(N)]
        p = np.r_[epsilon - z, epsilon + z]
        alpha, rho, stats = smo_solve_columns(cols, N, y, p, C, tol=tol, max_iter=max_iter,
                                              cache_size=cache_size)
        if verbose:
            print(f"SMO: {stats}")
        coefs = alpha[:N] - alpha[N:]
        bias = -rho
    return _model(X, lengths, coefs, bias, kernel_type, params, sv_chunk_size, device)
