"""DNA strings, one-hot arrays and base codes (A=0, C=1, G=2, T=3).

N (an all-zero one-hot column) and, with ``allow_other``, any other letter are
stored as A, which is how the kernels read them on the forward strand (as
LS-GKM does); their reverse complement then holds T.  Mirrors
codec.py:40-137 of gkmsvm-lite (one_hot_encode, encode_codes, to_codes).
"""

from __future__ import annotations

import numpy as np

BASES = "ACGT"

# byte -> code; 4 marks N/n, 5 any other letter
_LUT = np.full(256, 5, dtype=np.uint8)
for _i, _b in enumerate(BASES):
    _LUT[ord(_b)] = _i
    _LUT[ord(_b.lower())] = _i
_LUT[ord("N")] = 4
_LUT[ord("n")] = 4


def _raw_codes(sequence: str, allow_n: bool, allow_other: bool) -> np.ndarray:
    if len(sequence) == 0:
        raise ValueError("Sequence must be non-empty")
    raw = _LUT[np.frombuffer(sequence.encode("ascii"), dtype=np.uint8)]
    if (raw == 4).any() and not (allow_n or allow_other):
        raise ValueError("Invalid base in sequence (pass allow_n=True for N)")
    if (raw == 5).any() and not allow_other:
        raise ValueError(
            "Invalid base in sequence (pass allow_other=True to read non-ACGTN letters as N)"
        )
    return raw


def encode_codes(sequence: str, *, allow_n: bool = False, allow_other: bool = False) -> np.ndarray:
    """[L] uint8 base codes of a DNA string; N (and other letters with
    allow_other) stored as A."""
    raw = _raw_codes(sequence, allow_n, allow_other)
    return np.where(raw > 3, 0, raw).astype(np.uint8)


def encode_strings(sequences, min_width: int = 1):
    """Codes of many DNA strings in one pass: ([n, max(Lmax, min_width)] uint8
    padded with A, lengths [n] int64, bad [n] bool).  N is stored as A, as by
    encode_codes(allow_n=True); a string with any other letter, or empty, is
    flagged in bad (its row is left as decoded) for the caller to reject."""
    n = len(sequences)
    lens = np.fromiter((len(q) for q in sequences), dtype=np.int64, count=n)
    raw = _LUT[np.frombuffer("".join(sequences).encode("ascii", "replace"), dtype=np.uint8)]
    starts = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=starts[1:])
    other = raw == 5
    bad = lens == 0
    if other.any():
        cs = np.zeros(raw.size + 1, dtype=np.int64)
        np.cumsum(other, out=cs[1:])
        bad |= cs[starts[1:]] > cs[starts[:-1]]
    raw[raw > 3] = 0
    out = np.zeros((n, max(int(lens.max()) if n else 0, min_width)), dtype=np.uint8)
    for i in range(n):
        out[i, : lens[i]] = raw[starts[i]:starts[i + 1]]
    return out, lens, bad


def one_hot_encode(sequence: str, *, dtype=np.float32, allow_n: bool = False,
                   allow_other: bool = False) -> np.ndarray:
    """[4, L] one-hot (channels A, C, G, T); N is an all-zero column."""
    raw = _raw_codes(sequence, allow_n, allow_other)
    out = np.zeros((4, len(raw)), dtype=dtype)
    keep = raw < 4
    out[raw[keep], np.nonzero(keep)[0]] = 1
    return out


def to_codes(x) -> np.ndarray:
    """[B, L] uint8 codes of a [B, 4, L] one-hot batch (an all-zero column, N,
    becomes A), or [B, L] integer codes as they are.  Accepts NumPy arrays or
    torch tensors (moved to the host)."""
    if hasattr(x, "detach"):  # torch tensor
        x = x.detach().cpu().numpy()
    x = np.asarray(x)
    if x.ndim == 2:
        if x.dtype == np.uint8:
            return x
        if not (np.issubdtype(x.dtype, np.integer) or x.dtype == np.bool_):
            raise ValueError(f"[B, L] base codes must be integers, got {x.dtype}")
        return x.astype(np.uint8)
    if x.ndim != 3 or x.shape[1] != 4:
        raise ValueError(
            f"expected a [B, 4, L] one-hot array or [B, L] base codes, got {list(x.shape)}"
        )
    return np.argmax(x, axis=1).astype(np.uint8)


def codes_to_one_hot(codes, dtype=np.float64) -> np.ndarray:
    """[B, 4, L] one-hot of [B, L] codes (an N stored as A is one-hot A)."""
    codes = np.asarray(codes)
    B, L = codes.shape
    out = np.zeros((B, 4, L), dtype=dtype)
    out[np.arange(B)[:, None], codes.astype(np.int64), np.arange(L)[None, :]] = 1
    return out


def decode_codes(codes) -> str:
    return "".join(BASES[int(c)] for c in np.asarray(codes).ravel())


# -- one-hot utilities of the public API (codec.py:40-220 of gkmsvm-lite) -------------


def _is_torch(x) -> bool:
    return hasattr(x, "detach") and hasattr(x, "flip")


def one_hot_decode(arr) -> str:
    """The DNA string of a [4, L] one-hot array (an all-zero column, N, reads
    as A).  ValueError for another shape."""
    a = arr.detach().cpu().numpy() if _is_torch(arr) else np.asarray(arr)
    if a.ndim != 2 or a.shape[0] != 4:
        raise ValueError(f"one_hot_decode takes a [4, L] array, got shape {list(a.shape)}")
    return _LETTERS[np.argmax(a, axis=0)].tobytes().decode()


_LETTERS = np.frombuffer(BASES.encode(), dtype=np.uint8)


def reverse_complement(arr):
    """Reverse complement of a one-hot [4, L] or [B, 4, L] array (NumPy or
    torch): channels A<->T, C<->G swap (the channel order A C G T reversed)
    and positions run backwards.  ValueError for another shape."""
    if arr.ndim not in (2, 3) or arr.shape[-2] != 4:
        raise ValueError(f"reverse_complement takes [4, L] or [B, 4, L], got shape "
                         f"{list(arr.shape)}")
    if _is_torch(arr):
        return arr.flip((-2, -1)).contiguous()
    return np.ascontiguousarray(np.asarray(arr)[..., ::-1, ::-1])


def validate(arr) -> None:
    """ValueError unless arr [..., 4, L] is one-hot DNA: at every position
    exactly one non-zero channel and channels summing to 1."""
    a = arr.detach().cpu().numpy() if _is_torch(arr) else np.asarray(arr)
    if a.ndim < 2 or a.shape[-2] != 4:
        raise ValueError(f"a one-hot array has 4 channels on its second-to-last axis, got shape "
                         f"{list(a.shape)}")
    if not np.allclose(a.sum(axis=-2), 1.0):
        raise ValueError("not one-hot: the channels of some position do not sum to 1")
    if not ((a != 0).sum(axis=-2) == 1).all():
        raise ValueError("not one-hot: some position has more than one non-zero channel")


def encode_batch(sequences, *, dtype=np.float32, pad_value: float = 0.0,
                 allow_n: bool = False, allow_other: bool = False):
    """(one-hot [B, 4, L_max] of DNA strings, padded with pad_value past each
    string; mask [B, L_max] bool, True inside each string).  N is an all-zero
    column when allow_n (and any other letter too when allow_other); otherwise
    they raise ValueError, as does an empty list or an empty string."""
    if not len(sequences):
        raise ValueError("encode_batch needs at least one sequence")
    raw = [_raw_codes(s, allow_n, allow_other) for s in sequences]
    width = max(len(r) for r in raw)
    out = np.full((len(raw), 4, width), pad_value, dtype=dtype)
    mask = np.zeros((len(raw), width), dtype=bool)
    for i, r in enumerate(raw):
        n = len(r)
        mask[i, :n] = True
        col = np.zeros((4, n), dtype=dtype)
        keep = r < 4
        col[r[keep], np.flatnonzero(keep)] = 1
        out[i, :, :n] = col
    return out, mask
