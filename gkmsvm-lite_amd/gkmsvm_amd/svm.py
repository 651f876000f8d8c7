"""GkmSVM: the drop-in model class (gkmsvm-lite svm.py:155-662).

Scoring, variant scoring and the device switch keep the reference's names,
arguments and shapes.  Computation runs on the ROCm GPU only (``model.rocm()``,
the default device); results come back as float64 torch tensors on that GPU,
as the reference returns device arrays for a GPU model.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import engine as E
from . import tables as T
from .codec import to_codes
from .inputs import check_lengths, nan_rows, read_batch, read_pair, row_lengths, stacked
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
        self.sv_lengths = check_lengths(sv_lengths, sv.shape[0], sv.shape[-1])
        self.coefficients = coef
        self.bias = float(bias)
        self.kernel_type = resolve_kernel_type(kernel_type)
        self._kernel_params = dict(kernel_params)
        self.kernel = build_kernel(self.kernel_type, kernel_params)
        self.sv_chunk_size = sv_chunk_size
        # Score through the l-mer index of the SVs: "auto" picks the cheaper of
        # the index and the brute-force kernel per call; True/False force it
        self.use_lmer_index = "auto"
        # Exact l-mer weight table for non-RBF kernels (tables.py): "auto"
        # builds it when a call queries enough windows to pay for building it
        # (_plan_lmer_table) and uses it once built; True/False force it on or
        # off (svm.py:216-220)
        self.use_lmer_table: bool | str = "auto"
        # GkmExplain attributions of every l-mer window for non-RBF kernels,
        # one table per mode (hypothetical/importance, perturbation): "auto"
        # builds one when a call explains enough sequences and it fits
        # EXPLAIN_TABLE_BYTES; True/False force it on or off (svm.py:224-228)
        self.use_explain_table: bool | str = "auto"
        self._device = None
        self._kp = None
        self._sv = None
        if E.placement(device) is not None:
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
        """Drop the resident SV tables, l-mer index, l-mer table and explain
        tables (rebuilt on the next call; svm.py:241-251)."""
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
                                 row_lengths(self.support_sequences, self.sv_lengths),
                                 self.coefficients)
        return self._sv

    @property
    def num_support_vectors(self) -> int:
        return self.support_sequences.shape[0]

    @property
    def kernel_params(self) -> dict:
        return dict(self._kernel_params)

    # -- exact l-mer tables ---------------------------------------------------------

    def _build_lmer_table(self) -> torch.Tensor | None:
        """The score table w(u) = sum_sv c_sv K_raw(u, sv) on the GPU (cached),
        or None when the kernel has none (RBF, l > MAX_TABLE_L); svm.py:485-506."""
        sv = self._svs()
        if not T.applies(sv.kp):
            return None
        if "score" not in sv.tables:
            if "V" not in sv.tables:
                sv.tables["V"] = T.lmer_weights(sv)
            sv.tables["score"] = T.score_table(sv, sv.tables["V"])
        return sv.tables["score"]

    def _plan_lmer_table(self, n_windows: float) -> None:
        """Under "auto", build the l-mer table if querying n_windows windows
        through the per-SV paths would take longer than building it
        (svm.py:508-536, with MI355X costs)."""
        if self.use_lmer_table != "auto":
            return
        sv = self._svs()
        if "score" in sv.tables or not T.applies(sv.kp):
            return
        if sv.query_seconds(n_windows) >= T.build_seconds(sv):
            self._build_lmer_table()

    def _lmer_table_for(self) -> torch.Tensor | None:
        """The l-mer table if it applies and is on: built now when forced,
        used under "auto" only once built (svm.py:538-557)."""
        use = self.use_lmer_table
        if use is False:
            return None
        sv = self._svs()
        if use == "auto":
            return sv.tables.get("score")
        return self._build_lmer_table()

    def _explain_table_for(self, n_rows: int, x_lengths, perturbation: bool):
        """The explain table of this mode, built if use_explain_table calls for
        it (explain/table.py:150-169), or None."""
        use = self.use_explain_table
        sv = self._svs()
        kp = sv.kp
        if use is False or not T.explain_applies(kp):
            return None
        key = ("explain", bool(perturbation))
        if key not in sv.tables:
            if use == "auto":
                if T.explain_table_bytes(kp.l) > T.EXPLAIN_TABLE_BYTES:
                    return None
                from .explain import per_sv_seconds

                if per_sv_seconds(self, x_lengths, perturbation) < T.explain_build_seconds(sv):
                    return None
            from .explain import explain_tables

            ta, tb, _ = explain_tables(kp.base, perturbation)
            if "V" not in sv.tables:
                sv.tables["V"] = T.lmer_weights(sv)
            sv.tables[key] = T.explain_table(sv, ta, tb, sv.tables["V"])
        return sv.tables[key]

    def to_deltasvm(self, *, device: str = "rocm", query_norm: bool = False,
                    verbose: bool = False):
        """A DeltaSVM with one weight per l-mer: the model's kernel sum on that
        l-mer alone (svm.py:707-753).  Kernels with an l-mer table read it off
        the table; RBF kernels and l > MAX_TABLE_L score every l-mer through
        the kernel.  Without query_norm the weights drop the l-mer's own
        normalisation 1 / sqrt(K(u, u)); RBF kernels always keep it."""
        from .deltasvm import DeltaSVM

        kp = self._params()
        l = kp.l
        if l > self._MAX_DELTASVM_L:
            raise ValueError(
                f"to_deltasvm() requires l <= {self._MAX_DELTASVM_L} "
                f"(4^{l} = {4**l:,} weights would use {4**l * 4 / 1e9:.1f} GB). Got l={l}."
            )
        table = self._build_lmer_table()
        if table is not None:
            wt = kp.base.weight_table(1)
            wt0 = 1.0 if wt is None else float(wt[0])
            out = torch.empty(1 << (2 * l), dtype=torch.float64, device=kp.device)
            h = torch.tensor(kp.h[: l + 1], dtype=torch.float64, device=kp.device)
            _lib.call("gkm_deltasvm_weights", E._p(table), l, wt0, E._p(h), int(kp.rc),
                      int(query_norm and kp.normalize), E._p(out), E.stream_handle())
            weights = out.cpu().numpy()
        else:
            weights = self._deltasvm_weights_from_kernel(query_norm, verbose)
        return DeltaSVM(weights, l, l, include_rc=False, bias=self.bias, device=device)

    _MAX_DELTASVM_L = 14
    _DELTASVM_CHUNK = 65536

    def _deltasvm_weights_from_kernel(self, query_norm: bool, verbose: bool) -> np.ndarray:
        """Weights by scoring every l-mer (DeltaSVM index order); with RC only
        canonical l-mers are scored (svm.py:758-785)."""
        kp = self._params()
        unnorm = not query_norm and kp.normalize and not kp.rbf
        l = kp.l
        idx_all = np.arange(4**l, dtype=np.int64)
        powers = 4 ** np.arange(l - 1, -1, -1)
        digits = (idx_all[:, None] // powers) % 4
        rc = ((3 - digits[:, ::-1]) * powers).sum(axis=1)
        todo = idx_all[idx_all <= rc] if kp.rc else idx_all
        weights = np.empty(4**l, dtype=np.float64)
        starts = range(0, len(todo), self._DELTASVM_CHUNK)
        if verbose:
            from tqdm import tqdm

            starts = tqdm(starts, desc="to_deltasvm")
        sv = self._svs()
        for start in starts:
            idx = todo[start:start + self._DELTASVM_CHUNK]
            codes = np.ascontiguousarray(digits[idx].astype(np.uint8))
            lens = np.full(len(idx), l, dtype=np.int64)
            xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
            out = torch.empty(len(idx), dtype=torch.float64, device=kp.device)
            E.score_rows(sv, xs, 0.0, out, use_index=self.use_lmer_index)
            w = out.cpu().numpy()
            if unnorm:
                diag = E.self_kraw(kp, xs).cpu().numpy()
                w = w * np.sqrt(np.clip(diag, 1e-10, None))
            weights[idx] = w
        if kp.rc:
            weights[rc[todo]] = weights[todo]
        return weights

    # -- scoring ------------------------------------------------------------------

    def _score_codes(self, codes: np.ndarray, lengths: np.ndarray,
                     pairs: int = 0) -> torch.Tensor:
        """[B] decision values of host codes (rows shorter than l included;
        pairs: E.score_rows' allele pairs in the last 4 * pairs rows)."""
        sv = self._svs()
        xs = E.DeviceSeqs(codes, lengths, self._device, rc=self._params().rc)
        out = torch.empty(codes.shape[0], dtype=torch.float64, device=self._device)
        return E.score_rows(sv, xs, self.bias, out, use_index=self.use_lmer_index,
                            table=self._lmer_table_for(), pairs=pairs)

    def __call__(self, x, *, lengths=None, batch_size: int | None = None,
                 verbose: bool = False) -> torch.Tensor:
        """[B, 1] float64 decision values (NaN for sequences shorter than l).

        x: [B, 4, L] one-hot, [B, L] codes, or a list of [4, L_i] arrays, [L_i]
        codes or DNA strings; lengths: [B] lengths of a padded x's rows.
        """
        self._require_device()
        batch = read_batch(x, lengths)
        l = self.kernel.base.l
        short = batch.short(l)
        codes, lens = batch.codes, batch.lengths
        B = batch.n
        if short is not None and short.all():
            return nan_rows(torch.zeros((B, 1), dtype=torch.float64, device=self._device), short)
        if batch.width < l:
            raise ValueError(f"sequences of {batch.width} bases hold no window of l = {l}")
        self._plan_lmer_table(float(batch.n_windows(l).sum()))
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
            r, a = read_pair(ref, alt, ref_lengths, alt_lengths)
            if batch_size is None:
                effects = self._variant_pairs(r, a)
                if effects is not None:
                    return effects
            both = stacked(a, r)
            scores = self(both.codes, lengths=both.given_lengths(), batch_size=batch_size,
                          verbose=verbose)
            return scores[: r.n] - scores[r.n:]
        if method == "gkmexplain":
            from .explain import mutation_impact

            return mutation_impact(self, ref, alt, ref_lengths=ref_lengths,
                                   alt_lengths=alt_lengths, batch_size=batch_size,
                                   verbose=verbose)
        raise ValueError(f"method must be 'kernel' or 'gkmexplain', got {method!r}")

    def _variant_pairs(self, r, a) -> torch.Tensor | None:
        """score(alt) - score(ref) in one call, the alternates that differ
        from their reference in few windows (E.variant_cuts) scored from the
        reference's level sums plus those windows (E.score_rows' pairs layout:
        the same bits as scoring them); None where that does not apply (the
        l-mer table, RBF, position weights, no pair worth it)."""
        self._require_device()
        sv = self._svs()
        l = self.kernel.base.l
        if not E.pairs_apply(sv) or r.width < l:
            return None
        self._plan_lmer_table(float(r.n_windows(l).sum() + a.n_windows(l).sum()))
        if self._lmer_table_for() is not None:
            return None
        width = r.width
        inside = np.arange(width)[None, :] < r.lengths[:, None]
        diff = (r.codes != a.codes) & inside & (r.lengths == a.lengths)[:, None]
        some = diff.any(axis=1)
        lo = np.where(some, diff.argmax(axis=1), -1)
        hi = np.where(some, width - 1 - diff[:, ::-1].argmax(axis=1), -1)
        ok, c0, c1 = E.variant_cuts(r.lengths, lo, hi, l)
        if not ok.any():
            return None
        plain, cut = np.flatnonzero(~ok), np.flatnonzero(ok)
        n_plain, P = plain.size, cut.size
        # rows: plain references, plain alternates, then the pairs layout
        span = c1[cut] - c0[cut]
        k = np.arange(int(span.max()))
        cols = np.minimum(c0[cut][:, None] + k[None, :], width - 1)
        keep = k[None, :] < span[:, None]
        codes = np.concatenate([
            r.codes[plain], a.codes[plain], r.codes[cut],
            np.pad(np.where(keep, a.codes[cut[:, None], cols], 0), ((0, 0), (0, width - k.size))),
            np.pad(np.where(keep, r.codes[cut[:, None], cols], 0), ((0, 0), (0, width - k.size))),
            a.codes[cut]]).astype(np.uint8)
        lens = np.concatenate([r.lengths[plain], a.lengths[plain], r.lengths[cut], span, span,
                               a.lengths[cut]])
        out = self._score_codes(codes, lens, pairs=P).cpu().numpy()
        n = r.n
        ref_row = np.empty(n, dtype=np.int64)
        alt_row = np.empty(n, dtype=np.int64)
        ref_row[plain] = np.arange(n_plain)
        alt_row[plain] = n_plain + np.arange(n_plain)
        ref_row[cut] = 2 * n_plain + np.arange(P)
        alt_row[cut] = 2 * n_plain + 3 * P + np.arange(P)
        effects = out[alt_row] - out[ref_row]
        short = stacked(a, r).short(l)  # NaN (and the warning) where an allele has no window
        if short is not None:
            effects[short[:n] | short[n:]] = np.nan
        return torch.from_numpy(effects[:, None]).to(self._device)

    def _sv_diag(self) -> torch.Tensor:
        """K_raw(sv, sv) [S] of the base kernel, in SV order."""
        return self._svs().diag_original_order()
