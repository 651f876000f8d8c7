"""Input batches, allele pairs and the NaN / zero masking of rows and
positions past a sequence's length, as gkmsvm-lite's inputs.py:16-145 defines
them (host side of every call)."""

from __future__ import annotations

import warnings

import numpy as np

from .codec import encode_codes, to_codes

# LS-GKM reads at most this many bases of a sequence
LSGKM_MAX_LENGTH = 2047


def uniform_or_lengths(lengths, shape):
    """lengths as host int64 [B], or None when every row is shape[-1] long."""
    if lengths is None:
        return None
    if hasattr(lengths, "detach"):
        lengths = lengths.detach().cpu().numpy()
    lengths = np.asarray(lengths, dtype=np.int64).reshape(-1)
    B, L = shape[0], shape[-1]
    if lengths.shape[0] != B or (lengths < 0).any() or (lengths > L).any():
        raise ValueError(f"lengths must be {B} values in 0..{L}")
    return None if (lengths == L).all() else lengths


def as_batch(x, lengths=None):
    """(codes [B, L_max] uint8 on the host, lengths [B] int64 or None) from a
    [B, 4, L] one-hot array, [B, L] codes, or a list of [4, L_i] arrays, [L_i]
    codes or DNA strings (N allowed)."""
    if isinstance(x, (list, tuple)):
        if lengths is not None:
            raise ValueError("lengths go with a padded array, not a list")
        if not len(x):
            raise ValueError("no sequences")
        if isinstance(x[0], str):
            rows = [encode_codes(s, allow_n=True) for s in x]
        else:
            rows = [to_codes(np.asarray(a)[None])[0] for a in x]
        lens = np.array([len(r) for r in rows], dtype=np.int64)
        out = np.zeros((len(rows), int(lens.max())), dtype=np.uint8)
        for i, r in enumerate(rows):
            out[i, : len(r)] = r
        x, lengths = out, lens
    else:
        x = to_codes(x)
    lengths = uniform_or_lengths(lengths, x.shape)
    longest = int(x.shape[-1] if lengths is None else lengths.max())
    if longest > LSGKM_MAX_LENGTH:
        warnings.warn(
            f"sequences up to {longest} bp: LS-GKM reads only the first "
            f"{LSGKM_MAX_LENGTH} bases of longer ones",
            stacklevel=3,
        )
    return np.ascontiguousarray(x), lengths


def full_lengths(codes, lengths):
    """[B] int64 lengths (every row full length when lengths is None)."""
    if lengths is None:
        return np.full(codes.shape[0], codes.shape[-1], dtype=np.int64)
    return np.asarray(lengths, dtype=np.int64)


def allele_arrays(ref, alt, ref_lengths=None, alt_lengths=None):
    """(ref, alt, ref_lengths, alt_lengths) as batches of one length L, the
    shorter padded on the host."""
    ref, ref_lengths = as_batch(ref, ref_lengths)
    alt, alt_lengths = as_batch(alt, alt_lengths)
    if ref.shape[0] != alt.shape[0]:
        raise ValueError(f"ref and alt hold {ref.shape[0]} and {alt.shape[0]} sequences")
    L = max(ref.shape[-1], alt.shape[-1])

    def widen(x, n):
        if x.shape[-1] == L:
            return x, n
        n = np.full(x.shape[0], x.shape[-1], dtype=np.int64) if n is None else n
        return np.pad(x, ((0, 0), (0, L - x.shape[-1]))), n

    ref, ref_lengths = widen(ref, ref_lengths)
    alt, alt_lengths = widen(alt, alt_lengths)
    return ref, alt, ref_lengths, alt_lengths


def joint_lengths(L, *parts):
    """Lengths of rows from several batches of one length L (each part a
    (rows, lengths or None) pair), or None when every row is L long."""
    if all(n is None for _, n in parts):
        return None
    return np.concatenate(
        [np.full(rows, L, dtype=np.int64) if n is None else n for rows, n in parts]
    )


def rows_of(lengths, start, n):
    return None if lengths is None else lengths[start : start + n]


def short_rows(lengths, l, what="scores"):
    """[B] bool of rows shorter than l (no windows), or None; warns."""
    if lengths is None:
        return None
    short = lengths < l
    if not short.any():
        return None
    warnings.warn(
        f"{int(short.sum())} sequences are shorter than l = {l}: their {what} are NaN",
        stacklevel=3,
    )
    return short


def nan_rows(out, short):
    """out [B, ...] (torch tensor) with rows in short set to NaN."""
    if short is None:
        return out
    import torch

    mask = torch.as_tensor(short, device=out.device).reshape(-1, *[1] * (out.dim() - 1))
    return torch.where(mask, torch.full_like(out, float("nan")), out)


def per_position(out, lengths, short):
    """[B, 4, L] results with positions past each row's length 0 and short rows
    NaN (torch tensor)."""
    if lengths is None:
        return out
    import torch

    L = out.shape[2]
    inside = torch.arange(L, device=out.device)[None, None, :] < torch.as_tensor(
        lengths, device=out.device
    )[:, None, None]
    out = torch.where(inside, out, torch.zeros((), dtype=out.dtype, device=out.device))
    return nan_rows(out, short)
