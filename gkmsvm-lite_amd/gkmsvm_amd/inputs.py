"""Host-side sequence batches.

Every public call turns its input into a ``SeqBatch``: base codes
``[B, W]`` uint8 (A = 0, C = 1, G = 2, T = 3; rows padded with A past their
length) and the length of every row, always explicit.  ``padded`` records
whether the caller's rows differ in length (the reference's calls then carry
a lengths array; inputs.py:27-61); results past a row's length are 0 and
rows with no l-mer window are NaN, with a warning
(inputs.py:112-145 is the behaviour kept).

Accepted inputs: a [B, 4, L] one-hot array, [B, L] integer codes (with
``lengths`` when rows are padded), or a list of DNA strings (N read as A),
[L_i] code arrays or [4, L_i] one-hot arrays.  NumPy arrays and torch
tensors are both read on the host.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np

from .codec import encode_strings, to_codes

# LS-GKM keeps a sequence in a 2048-byte buffer (one byte for the
# terminator), so gkmtrain/gkmpredict see at most this many bases
LSGKM_MAX_LENGTH = 2047


def _host_int64(a) -> np.ndarray:
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.int64).reshape(-1)


def check_lengths(lengths, n: int, width: int) -> np.ndarray | None:
    """Validated row lengths (int64 [n]) for rows `width` wide, or None when
    every row is full width (so is a None argument).  ValueError unless
    there are n lengths, each in 0..width."""
    if lengths is None:
        return None
    lens = _host_int64(lengths)
    if lens.size != n or lens.min(initial=0) < 0 or lens.max(initial=0) > width:
        raise ValueError(f"expected {n} row lengths between 0 and {width}, got {lens.size} "
                         f"values{'' if lens.size == 0 else f' in {lens.min()}..{lens.max()}'}")
    return None if bool((lens == width).all()) else lens


def row_lengths(codes: np.ndarray, lengths) -> np.ndarray:
    """int64 [n] lengths of padded rows (the full width where lengths is None)."""
    if lengths is None:
        return np.full(codes.shape[0], codes.shape[-1], dtype=np.int64)
    return _host_int64(lengths)


@dataclass(frozen=True)
class SeqBatch:
    """Base codes [B, W] uint8 and every row's length; padded: the rows came
    with lengths (some row shorter than W)."""

    codes: np.ndarray
    lengths: np.ndarray
    padded: bool

    @property
    def n(self) -> int:
        return self.codes.shape[0]

    @property
    def width(self) -> int:
        return self.codes.shape[1]

    def given_lengths(self) -> np.ndarray | None:
        """The lengths as a padded batch states them (None: all full width)."""
        return self.lengths if self.padded else None

    def n_windows(self, l: int) -> np.ndarray:
        """l-mer windows per row."""
        return np.clip(self.lengths - l + 1, 0, None)

    def short(self, l: int, what: str = "scores") -> np.ndarray | None:
        """bool [B] of rows with fewer than l bases (no window), or None when
        there are none; warns when there are."""
        if not self.padded:
            return None
        few = self.lengths < l
        if not few.any():
            return None
        warnings.warn(f"{int(few.sum())} of {self.n} sequences have fewer than l = {l} bases, "
                      f"so they have no l-mer windows: their {what} are NaN", stacklevel=3)
        return few

    def widened(self, width: int) -> SeqBatch:
        """The same rows padded with A to `width` columns."""
        if width == self.width:
            return self
        codes = np.zeros((self.n, width), dtype=np.uint8)
        codes[:, : self.width] = self.codes
        return SeqBatch(codes, self.lengths, True)

    def rows(self, start: int, stop: int) -> SeqBatch:
        return SeqBatch(self.codes[start:stop], self.lengths[start:stop], self.padded)


def _pack_rows(rows: list) -> tuple[np.ndarray, np.ndarray]:
    lens = np.fromiter((r.shape[-1] for r in rows), dtype=np.int64, count=len(rows))
    codes = np.zeros((len(rows), int(lens.max())), dtype=np.uint8)
    for i, r in enumerate(rows):
        codes[i, : r.shape[-1]] = r
    return codes, lens


def read_batch(x, lengths=None) -> SeqBatch:
    """A SeqBatch of any accepted input (module docstring).  ValueError for an
    empty list, lengths given with a list, lengths out of range, or a string
    that is empty or holds a letter other than ACGTN; warns about sequences
    longer than LS-GKM reads."""
    if isinstance(x, (list, tuple)):
        if lengths is not None:
            raise ValueError("a list of sequences carries its own lengths: pass lengths only "
                             "with a padded [B, L] or [B, 4, L] array")
        if len(x) == 0:
            raise ValueError("the input holds no sequences")
        if isinstance(x[0], str):
            codes, lens, bad = encode_strings(list(x))
            if bad.any():
                i = int(np.argmax(bad))
                raise ValueError(f"sequence {i} is empty or holds a letter other than "
                                 f"A, C, G, T, N: {x[i][:40]!r}")
        else:
            codes, lens = _pack_rows([to_codes(np.asarray(a)[None])[0] for a in x])
        given = None if bool((lens == codes.shape[1]).all()) else lens
    else:
        codes = to_codes(x)
        given = check_lengths(lengths, codes.shape[0], codes.shape[-1])
        lens = row_lengths(codes, given)
    if lens.size and int(lens.max()) > LSGKM_MAX_LENGTH:
        warnings.warn(f"the longest sequence has {int(lens.max())} bases; LS-GKM itself would "
                      f"read only its first {LSGKM_MAX_LENGTH}", stacklevel=3)
    return SeqBatch(np.ascontiguousarray(codes), lens, given is not None)


def read_pair(ref, alt, ref_lengths=None, alt_lengths=None) -> tuple[SeqBatch, SeqBatch]:
    """Reference and alternative alleles as two batches of one width (the
    narrower padded); ValueError when they hold different numbers of
    sequences."""
    r = read_batch(ref, ref_lengths)
    a = read_batch(alt, alt_lengths)
    if r.n != a.n:
        raise ValueError(f"{r.n} reference and {a.n} alternative sequences: variants come "
                         "in pairs")
    width = max(r.width, a.width)
    return r.widened(width), a.widened(width)


def stacked(*batches: SeqBatch) -> SeqBatch:
    """Batches of one width as one, rows in argument order."""
    return SeqBatch(np.concatenate([b.codes for b in batches]),
                    np.concatenate([b.lengths for b in batches]),
                    any(b.padded for b in batches))


def nan_rows(out, rows):
    """The torch tensor out [B, ...] with the rows flagged in `rows` (bool
    [B] or None) set to NaN."""
    if rows is None:
        return out
    import torch

    flag = torch.as_tensor(np.asarray(rows, dtype=bool), device=out.device)
    return out.masked_fill(flag.view(-1, *([1] * (out.dim() - 1))), float("nan"))


def mask_positions(out, batch: SeqBatch, rows=None):
    """Per-position results out [B, 4, W] (torch) with 0 past each row's
    length and NaN in the flagged rows."""
    if not batch.padded:
        return nan_rows(out, rows)
    import torch

    pos = torch.arange(out.shape[2], device=out.device)
    past = pos[None, :] >= torch.as_tensor(batch.lengths, device=out.device)[:, None]
    return nan_rows(out.masked_fill(past[:, None, :], 0.0), rows)
