"""GkmSVM: the drop-in model class (gkmsvm-lite svm.py:155-662).

Scoring, variant scoring and the device switch keep the reference's names,
arguments and shapes.  Computation runs on the ROCm GPU only (``model.rocm()``,
the default device); results come back as float64 torch tensors on that GPU,
as the reference returns device arrays for a GPU model.
"""

from __future__ import annotations

import numpy as np
import torch

from . import engine as E
from .codec import to_codes
from .inputs import (
    allele_arrays,
    as_batch,
    full_lengths,
    joint_lengths,
    nan_rows,
    rows_of,
    short_rows,
    uniform_or_lengths,
)
from .kernels import KERNEL_ALIASES, KERNEL_BUILDERS, build_kernel, resolve_kernel_type  # noqa: F401


def to_cpu(x):
    """A result as a NumPy array on the host."""
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


class GkmSVM:
    """Gapped k-mer SVM: score = sum_i coef_i K(x, sv_i) + bias."""

    def __init__(self, support_sequences, coefficients, bias: float, kernel_type,
                 kernel_params: dict, *, sv_lengths=None, sv_chunk_size: int | None = None,
                 device: str = "rocm"):
        sv = np.asarray(support_sequences.detach().cpu() if hasattr(support_sequences, "detach")
                        else support_sequences)
        if sv.ndim == 3 and sv.shape[1] != 4:
            raise ValueError(
                f"support_sequences must be [S, 4, L] one-hot or [S, L] base codes, "
                f"got {list(sv.shape)}"
            )
        sv = to_codes(sv)
        coef = np.asarray(coefficients.detach().cpu() if hasattr(coefficients, "detach")
                          else coefficients, dtype=np.float64)
        if coef.ndim != 1:
            raise ValueError(f"coefficients must be [S], got {list(coef.shape)}")
        if sv.shape[0] != coef.shape[0]:
            raise ValueError(
                f"Number of support sequences ({sv.shape[0]}) must match number of "
                f"coefficients ({coef.shape[0]})"
            )
        self.support_sequences = np.ascontiguousarray(sv)
        self.sv_lengths = uniform_or_lengths(sv_lengths, sv.shape)
        self.coefficients = coef
        self.bias = float(bias)
        self.kernel_type = resolve_kernel_type(kernel_type)
        self._kernel_params = dict(kernel_params)
        self.kernel = build_kernel(self.kernel_type, kernel_params)
        self.sv_chunk_size = sv_chunk_size
        # Score through the l-mer index of the SVs: "auto" picks the cheaper of
        # the index and the brute-force kernel per call; True/False force it
        self.use_lmer_index = "auto"
        self._device = None
        self._kp = None
        self._sv = None
        if device is not None and str(device) != "cpu":
            self.rocm(device)

    # -- device switch -------------------------------------------------------

    def rocm(self, device=None):
        """Hold the model on a ROCm GPU (SV tables resident in HBM); returns self."""
        dev = E.device_of(device)
        if self._device != dev:
            self._device = dev
            self._kp = None
            self._sv = None
        return self

    # the reference's name for the same switch
    cuda = rocm

    def cpu(self):
        """Drop the device tables; scoring then needs ``rocm()`` again."""
        self._device = None
        self._kp = None
        self._sv = None
        return self

    @property
    def device(self) -> str:
        return "cpu" if self._device is None else "rocm"

    def clear_caches(self) -> None:
        """Drop the resident SV tables (rebuilt on the next call)."""
        self._sv = None

    def _require_device(self):
        if self._device is None:
            raise RuntimeError(
                "this model is on the CPU: the gkm engine computes on ROCm GPUs only; "
                "call model.rocm()"
            )
        return self._device

    def _params(self) -> E.KernelParams:
        if self._kp is None:
            self._kp = E.KernelParams(self.kernel, self._require_device())
        return self._kp

    def _svs(self) -> E.SVModel:
        """The resident SV tables (built on first use)."""
        if self._sv is None:
            kp = self._params()
            self._sv = E.SVModel(kp, self.support_sequences,
                                 full_lengths(self.support_sequences, self.sv_lengths),
                                 self.coefficients)
        return self._sv

    @property
    def num_support_vectors(self) -> int:
        return self.support_sequences.shape[0]

    @property
    def kernel_params(self) -> dict:
        return dict(self._kernel_params)

    # -- scoring ------------------------------------------------------------------

    def _score_codes(self, codes: np.ndarray, lengths: np.ndarray) -> torch.Tensor:
        """[B] decision values of host codes (rows shorter than l included)."""
        sv = self._svs()
        xs = E.DeviceSeqs(codes, lengths, self._device, rc=self._params().rc)
        out = torch.empty(codes.shape[0], dtype=torch.float64, device=self._device)
        return E.score_rows(sv, xs, self.bias, out, use_index=self.use_lmer_index)

    def __call__(self, x, *, lengths=None, batch_size: int | None = None,
                 verbose: bool = False) -> torch.Tensor:
        """[B, 1] float64 decision values (NaN for sequences shorter than l).

        x: [B, 4, L] one-hot, [B, L] codes, or a list of [4, L_i] arrays, [L_i]
        codes or DNA strings; lengths: [B] lengths of a padded x's rows.
        """
        self._require_device()
        codes, lengths = as_batch(x, lengths)
        l = self.kernel.base.l
        short = short_rows(lengths, l)
        lens = full_lengths(codes, lengths)
        B = codes.shape[0]
        if short is not None and short.all():
            return nan_rows(torch.zeros((B, 1), dtype=torch.float64, device=self._device), short)
        if codes.shape[1] < l:
            raise ValueError(f"Sequence length {codes.shape[1]} is shorter than window length {l}")
        bs = B if batch_size is None else max(1, int(batch_size))
        parts = []
        starts = range(0, B, bs)
        if verbose:
            from tqdm import tqdm

            starts = tqdm(starts, desc="Scoring")
        for s in starts:
            parts.append(self._score_codes(codes[s:s + bs], lens[s:s + bs]))
        out = (parts[0] if len(parts) == 1 else torch.cat(parts))[:, None]
        return nan_rows(out, short)

    def score_variants(self, ref, alt, *, method: str = "kernel", ref_lengths=None,
                       alt_lengths=None, batch_size: int | None = None,
                       verbose: bool = False) -> torch.Tensor:
        """[B, 1] variant effect scores: score(alt) - score(ref) ("kernel") or
        the GkmExplain mutation impact ("gkmexplain", lsgkm gkmexplain -m 5)."""
        if method == "kernel":
            ref, alt, ref_lengths, alt_lengths = allele_arrays(ref, alt, ref_lengths, alt_lengths)
            n = ref.shape[0]
            scores = self(
                np.concatenate([alt, ref]),
                lengths=joint_lengths(ref.shape[-1], (n, alt_lengths), (n, ref_lengths)),
                batch_size=batch_size,
                verbose=verbose,
            )
            return scores[:n] - scores[n:]
        if method == "gkmexplain":
            from .explain import mutation_impact

            return mutation_impact(self, ref, alt, ref_lengths=ref_lengths,
                                   alt_lengths=alt_lengths, batch_size=batch_size,
                                   verbose=verbose)
        raise ValueError(f"method must be 'kernel' or 'gkmexplain', got {method!r}")

    def _sv_diag(self) -> torch.Tensor:
        """K_raw(sv, sv) [S] of the base kernel, in SV order."""
        return self._svs().diag_original_order()

    def _rows(self, lengths, start, n):
        return rows_of(lengths, start, n)
