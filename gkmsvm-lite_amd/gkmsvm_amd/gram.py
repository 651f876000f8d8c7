"""Gram matrices on the GPU (gkmsvm-lite gram.py:21-164)."""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import engine as E
from .inputs import read_batch


GRAM_DEVICE_BYTES = 32 << 30  # a Gram up to this size is assembled in HBM


def _gram_index(kp, x_lengths, ys, force: bool = False):
    """An l-mer index of the columns when it is the cheaper path (the cost
    model of SVModel.use_index_for, the columns as SVs) or forced, else None
    (also when the index does not apply)."""
    if kp.l > E.INDEX_MAX_L or kp.dmax < 0:
        return None
    if force:
        return E.DeviceIndex(kp, ys)
    from math import comb

    norient = 2 if kp.rc else 1
    wy = float(np.clip(ys.lengths - kp.l + 1, 0, None).sum())
    wx = float(np.clip(np.asarray(x_lengths) - kp.l + 1, 0, None).sum())
    n_masks = sum(comb(kp.l, m) * 3**m for m in range(kp.dmax + 1) if kp.h[m] != 0)
    n_tiles = -(-ys.n // E.index_tile(kp.nm, not kp.weighted, ys.n))
    probes = wx * n_masks
    rate = E.INDEX_PROBES_PER_S_WEIGHTED if kp.weighted else E.INDEX_PROBES_PER_S
    index_s = (probes * n_tiles / rate
               + probes * norient * wy / 4**kp.l / E.INDEX_ENTRIES_PER_S)
    # the pairwise kernel skips the lower triangle; the index probes every row
    brute_s = 0.5 * wx * norient * wy / E.BRUTE_COMPARISONS_PER_S
    return E.DeviceIndex(kp, ys) if index_s < brute_s else None


def compute_gram(kernel, X, Y=None, *, chunk_size: int = 1000, verbose: bool = False,
                 out: np.ndarray | None = None, lengths=None, y_lengths=None,
                 device=None, use_index="auto") -> np.ndarray:
    """[N, M] (or [N, N]) normalised, finished kernel matrix in host memory.

    Without Y, only the upper triangle (with the diagonal) is computed and
    mirrored, so the matrix is exactly symmetric.  Rows go to the GPU in blocks
    whose [rows, N] histograms stay under engine.MAX_HIST_BYTES; sequences are
    sorted by length on the device (64-column groups of one length) and the
    result is put back in input order.  The histograms come from the pairwise
    kernel or an l-mer index of the columns (use_index "auto": the cheaper by
    the cost model; True/False force it) -- the same integers either way, as
    the reference builds the Gram through its index (gram.py:82-106).
    chunk_size is accepted for API compatibility; the block size follows from
    memory.
    """
    dev = E.device_of(device)
    kp = E.KernelParams(kernel, dev)
    xb = read_batch(X, lengths)
    xc, xl = xb.codes, xb.lengths
    symmetric = Y is None
    if symmetric:
        yc, yl = xc, xl
    else:
        yb = read_batch(Y, y_lengths)
        yc, yl = yb.codes, yb.lengths
    N, M = xc.shape[0], yc.shape[0]
    gram = np.zeros((N, M), dtype=np.float64) if out is None else out
    st = E.stream_handle()
    ox = np.argsort(xl, kind="stable")
    xs = E.DeviceSeqs(xc[ox], xl[ox], dev, rc=kp.rc)
    if symmetric:
        oy, ys = ox, xs
    else:
        oy = np.argsort(yl, kind="stable")
        ys = E.DeviceSeqs(yc[oy], yl[oy], dev, rc=kp.rc)
    groups = torch.from_numpy(E.group_bounds(yl[oy])).to(dev)
    index = _gram_index(kp, xl, ys, use_index is True) if use_index is not False else None
    if use_index is True and index is None:
        raise ValueError("the l-mer index needs l <= 12 and a mismatch cutoff")
    ones_x = torch.ones(N, dtype=torch.float64, device=dev)
    dx = E.self_kraw(kp, xs) if kp.normalize else ones_x
    dy = dx if symmetric else (E.self_kraw(kp, ys) if kp.normalize
                               else torch.ones(M, dtype=torch.float64, device=dev))
    step = E.rows_per_chunk(M, kp.nm)
    # the whole sorted Gram in HBM when it fits (mirrored and put back in input
    # order on the device, then one copy to the host); else row blocks to host
    on_device = N * M * 8 <= GRAM_DEVICE_BYTES
    if on_device:
        g_dev = torch.empty((N, M), dtype=torch.float64, device=dev)
    else:
        sorted_g = np.empty((N, M), dtype=np.float64)
    tile = None
    rows = range(0, N, step)
    if verbose:
        from tqdm import tqdm

        rows = tqdm(rows, desc="Gram matrix")
    for r0 in rows:
        r1 = min(N, r0 + step)
        if index is not None:
            hist = E.index_hist(kp, index, xs, r0, r1, M, col_min=0 if symmetric else -1)
        else:
            hist = E.cross_hist(kp, xs, r0, r1, ys, groups, col_min=0 if symmetric else -1)
        if on_device:
            tile = g_dev[r0:r1]
        elif tile is None or tile.shape[0] != r1 - r0:
            tile = torch.empty((r1 - r0, M), dtype=torch.float64, device=dev)
        _lib.call("gkm_combine_gram", E._p(hist), r1 - r0, M, kp.nm, E._p(kp.tab),
                  E._p(dx[r0:r1]), E._p(dy), int(kp.normalize), int(kp.rbf), kp.gamma,
                  E._p(tile), st)
        if not on_device:
            sorted_g[r0:r1] = tile.cpu().numpy()
        del hist
    if on_device:
        if symmetric:  # upper triangle (with the diagonal) mirrored: exactly symmetric
            g_dev = torch.triu(g_dev)
            g_dev += torch.triu(g_dev, 1).T
        if not (np.array_equal(ox, np.arange(N)) and np.array_equal(oy, np.arange(M))):
            inv_x = torch.from_numpy(np.argsort(ox)).to(dev)
            inv_y = torch.from_numpy(np.argsort(oy)).to(dev)
            g_dev = g_dev.index_select(0, inv_x).index_select(1, inv_y)
        if gram.flags.c_contiguous and gram.dtype == np.float64:
            torch.from_numpy(gram).copy_(g_dev)
        else:
            gram[...] = g_dev.cpu().numpy()
        return gram
    if symmetric:
        upper = np.triu(sorted_g)
        sorted_g = upper + np.triu(upper, 1).T
    gram[np.ix_(ox, oy)] = sorted_g
    return gram
