"""DeltaSVM: a linear scoring model over gapped k-mer weights, on the GPU
(gkmsvm-lite deltasvm.py:1-239).

A DeltaSVM holds one weight per k-mer (4^k, first base most significant) and
scores a sequence by summing, over its windows of length l and the C(l, k)
position combinations of each window, the weight of the gapped k-mer read
there, on both strands with ``include_rc`` (the reverse strand reads N as T,
as the gkm-SVM kernels do).  ``GkmSVM.to_deltasvm()`` builds one with k = l
from the model's exact l-mer table.
"""

from __future__ import annotations

from itertools import combinations

import numpy as np
import torch

from . import _lib
from . import engine as E
from .inputs import read_batch

_BASE_MAP = {"A": 0, "C": 1, "G": 2, "T": 3}


def kmer_to_index(kmer: str) -> int:
    """Index of a k-mer string, first base most significant (deltasvm.py:47-51)."""
    idx = 0
    for base in kmer:
        idx = idx * 4 + _BASE_MAP[base]
    return idx


def index_to_kmer(idx: int, k: int) -> str:
    bases = "ACGT"
    chars = []
    for _ in range(k):
        chars.append(bases[idx % 4])
        idx //= 4
    return "".join(reversed(chars))


class DeltaSVM:
    """Linear gapped k-mer scoring model: score = sum over windows and
    C(l, k) combos of weight[k-mer] (+ the reverse strand) + bias."""

    def __init__(self, weights, l: int, k: int, *, include_rc: bool = True, bias: float = 0.0,
                 device: str = "rocm"):
        weights = np.asarray(weights.detach().cpu() if hasattr(weights, "detach") else weights)
        expected = 4**k
        if weights.shape != (expected,):
            raise ValueError(f"weights must be [{expected}], got {list(weights.shape)}")
        if k > l:
            raise ValueError(f"k ({k}) must be <= l ({l})")
        self.weights = weights.astype(np.float64)
        self.bias = float(bias)
        self.l = int(l)
        self.k = int(k)
        self.include_rc = bool(include_rc)
        self._combos = np.array(list(combinations(range(l), k)), dtype=np.int32)
        self._device = None
        self._dev_weights = None
        self._dev_combos = None
        if E.placement(device) is not None:
            self.rocm(device)

    @property
    def num_kmers(self) -> int:
        return int((self.weights != 0).sum())

    def rocm(self, device=None):
        """Hold the weights on a ROCm GPU; returns self."""
        dev = E.device_of(device)
        if self._device != dev:
            self._device = dev
            self._dev_weights = torch.from_numpy(self.weights).to(dev)
            self._dev_combos = torch.from_numpy(self._combos).to(dev)
        return self

    cuda = rocm

    def cpu(self):
        self._device = None
        self._dev_weights = None
        self._dev_combos = None
        return self

    @property
    def device(self) -> str:
        return "cpu" if self._device is None else "rocm"

    def __call__(self, x, *, lengths=None, verbose: bool = False) -> torch.Tensor:
        """[B, 1] float64 scores; each row scored at its own length."""
        if self._device is None:
            raise RuntimeError(
                "this model is on the CPU: DeltaSVM scoring runs on ROCm GPUs only; "
                "call model.rocm()"
            )
        batch = read_batch(x, lengths)
        codes, lens = batch.codes, batch.lengths
        B, L = codes.shape
        out = torch.empty(B, dtype=torch.float64, device=self._device)
        c = torch.from_numpy(np.ascontiguousarray(codes)).to(self._device)
        ln = torch.from_numpy(lens.astype(np.int32)).to(self._device)
        _lib.call("gkm_deltasvm_score", E._p(c), E._p(ln), B, L, self.l, self.k,
                  E._p(self._dev_combos), len(self._combos), E._p(self._dev_weights),
                  int(self.include_rc), self.bias, E._p(out), E.stream_handle())
        return out[:, None]

    def score_variants(self, ref, alt, *, verbose: bool = False) -> torch.Tensor:
        """[B, 1] score(alt) - score(ref)."""
        return self(alt, verbose=verbose) - self(ref, verbose=verbose)
