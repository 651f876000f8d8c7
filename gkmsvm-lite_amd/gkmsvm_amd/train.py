"""train_gkmsvm: Gram on the GPU, then sklearn's LIBSVM with a precomputed
kernel (gkmsvm-lite train.py:26-128, 307-450; only the libsvm solver)."""

from __future__ import annotations

import numpy as np

from .gram import compute_gram
from .inputs import as_batch, full_lengths
from .kernels import KERNEL_BUILDERS, resolve_kernel_type
from .svm import GkmSVM


def _codes(seqs):
    codes, lengths = as_batch(seqs if not isinstance(seqs, np.ndarray) or seqs.ndim != 1
                              else list(seqs))
    return codes, full_lengths(codes, lengths)


def _build_kernel(kernel_type, l, k, d, gamma, M, H, include_rc):
    kernel_type = resolve_kernel_type(kernel_type)
    params = {"L": l, "k": k, "d": d, "include_rc": include_rc}
    if kernel_type in ("gkmrbf", "wgkmrbf", "gkm_cnt_rbf"):
        params["gamma"] = gamma
    if kernel_type in ("wgkm", "wgkmrbf"):
        params["M"] = 50 if M is None else M
        params["H"] = 50.0 if H is None else H
    return KERNEL_BUILDERS[kernel_type](params), kernel_type, params


def train_gkmsvm(pos_seqs, neg_seqs, *, kernel_type="estimated", l: int = 11, k: int = 7,
                 d: int = 3, C: float = 1.0, pos_weight: float = 1.0, gamma: float = 1.0,
                 M: int | None = None, H: float | None = None, include_rc: bool = True,
                 solver: str = "libsvm", sv_chunk_size: int | None = None,
                 gram_chunk_size: int = 1000, device: str = "rocm", verbose: bool = False,
                 **_unused) -> GkmSVM:
    """Train a gapped k-mer SVM classifier on positive and negative sequences.

    The N x N Gram is computed on the GPU (upper triangle, mirrored) and solved
    by sklearn's SVC(kernel="precomputed"); positives get C * pos_weight.
    Returns a GkmSVM on the ROCm device.
    """
    if solver not in ("auto", "libsvm"):
        raise NotImplementedError(f"solver {solver!r}: only the precomputed-Gram libsvm solver")
    pc, pl = _codes(pos_seqs)
    nc, nl = _codes(neg_seqs)
    L = max(pc.shape[1], nc.shape[1])
    X = np.zeros((pc.shape[0] + nc.shape[0], L), dtype=np.uint8)
    X[: pc.shape[0], : pc.shape[1]] = pc
    X[pc.shape[0]:, : nc.shape[1]] = nc
    lengths = np.concatenate([pl, nl])
    if (lengths < l).any():
        raise ValueError(f"training sequences must be at least l = {l} long")
    y = np.array([1] * pc.shape[0] + [-1] * nc.shape[0])
    kernel, kernel_type, params = _build_kernel(kernel_type, l, k, d, gamma, M, H, include_rc)
    gram = compute_gram(kernel, X, lengths=lengths if (lengths != L).any() else None,
                        chunk_size=gram_chunk_size, verbose=verbose, device=device)
    from sklearn.svm import SVC

    clf = SVC(C=C, kernel="precomputed",
              class_weight=None if pos_weight == 1.0 else {1: pos_weight})
    clf.fit(gram, y)
    sv_idx = clf.support_
    sv_len = lengths[sv_idx]
    sv = X[sv_idx][:, : int(sv_len.max())]
    return GkmSVM(sv, np.asarray(clf.dual_coef_[0], dtype=np.float64), float(clf.intercept_[0]),
                  kernel_type, params, sv_lengths=sv_len if (sv_len != sv.shape[1]).any() else None,
                  sv_chunk_size=sv_chunk_size, device=device)
