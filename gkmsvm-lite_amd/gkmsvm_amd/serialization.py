"""Model files: the native npz format, LS-GKM text and deltaSVM weight files
(gkmsvm-lite serialization.py:15-166, importers/deltasvm.py:21-115).

npz files hold plain arrays and a JSON metadata record (loaded with
``allow_pickle=False``), so a file written here loads in gkmsvm-lite and the
other way round; LS-GKM text files are what lsgkm's gkmtrain writes (rho is
the negative bias; ``norc 1`` only when the reverse complement is off).
"""

from __future__ import annotations

import gzip
import json
from pathlib import Path

import numpy as np

from .codec import decode_codes


def _host(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def save_npz(model, path) -> None:
    """Save a GkmSVM in the native npz format: support sequences (base codes),
    coefficients, sv_lengths when the SVs differ in length, and metadata
    (bias, kernel_type, kernel_params, sv_chunk_size)."""
    metadata = {
        "bias": float(model.bias),
        "kernel_type": model.kernel_type,
        "kernel_params": model.kernel_params,
    }
    if model.sv_chunk_size is not None:
        metadata["sv_chunk_size"] = model.sv_chunk_size
    extra = {} if model.sv_lengths is None else {"sv_lengths": np.asarray(model.sv_lengths)}
    np.savez(path, support_sequences=_host(model.support_sequences),
             coefficients=_host(model.coefficients),
             metadata=np.void(json.dumps(metadata).encode("utf-8")), **extra)


def save_deltasvm_npz(model, path) -> None:
    """Save a DeltaSVM: weights [4^k] and metadata (l, k, include_rc, bias)."""
    metadata = {"model_type": "deltasvm", "l": model.l, "k": model.k,
                "include_rc": model.include_rc, "bias": float(model.bias)}
    np.savez(path, weights=_host(model.weights),
             metadata=np.void(json.dumps(metadata).encode("utf-8")))


def _read_npz(path):
    data = np.load(path, allow_pickle=False)
    return data, json.loads(bytes(data["metadata"]))


def load_npz(path, *, device: str = "rocm"):
    """A GkmSVM from the native npz format (files without sv_lengths hold SVs
    of one length)."""
    from .svm import GkmSVM

    data, meta = _read_npz(path)
    with data:
        return GkmSVM(
            data["support_sequences"], data["coefficients"], meta["bias"], meta["kernel_type"],
            meta["kernel_params"],
            sv_lengths=data["sv_lengths"] if "sv_lengths" in data.files else None,
            sv_chunk_size=meta.get("sv_chunk_size"), device=device)


def load_deltasvm_npz(path, *, device: str = "rocm"):
    from .deltasvm import DeltaSVM

    data, meta = _read_npz(path)
    with data:
        return DeltaSVM(data["weights"], meta["l"], meta["k"], include_rc=meta["include_rc"],
                        bias=meta.get("bias", 0.0), device=device)


def save_lsgkm(model, path) -> None:
    """Save a GkmSVM as an LS-GKM text model (gzip when the name ends in .gz):
    header (kernel_type, L, k, d, gamma, M, H, norc only when RC is off, rho =
    -bias, nr_sv by coefficient sign), then one ``coef SEQUENCE`` line per SV
    without its padding."""
    path = Path(path)
    params = model.kernel_params
    coefs = _host(model.coefficients)
    svs = _host(model.support_sequences)
    lines = ["svm_type c_svc", f"kernel_type {model.kernel_type}", f"L {params['L']}",
             f"k {params['k']}"]
    for key in ("d", "gamma", "M", "H"):
        if key in params:
            lines.append(f"{key} {params[key]}")
    if not params.get("include_rc", True):
        lines.append("norc 1")
    lines += ["nr_class 2", f"total_sv {len(coefs)}", f"rho {-model.bias}", "label 1 -1",
              f"nr_sv {int((coefs > 0).sum())} {int((coefs <= 0).sum())}", "SV"]
    lengths = model.sv_lengths
    for i in range(len(coefs)):
        sv = svs[i] if lengths is None else svs[i][: lengths[i]]
        lines.append(f"{coefs[i]} {decode_codes(sv)}")
    text = "\n".join(lines) + "\n"
    if str(path).endswith(".gz"):
        with gzip.open(path, "wt") as f:
            f.write(text)
    else:
        path.write_text(text)


def load_model(path, *, device: str = "rocm"):
    """A GkmSVM or DeltaSVM by file name: ``.npz`` (native; DeltaSVM when its
    metadata says so), else LS-GKM text (``.txt``, ``.txt.gz``, ``.model``)."""
    path = Path(path)
    if path.name.lower().endswith(".npz"):
        data, meta = _read_npz(path)
        data.close()
        if meta.get("model_type") == "deltasvm":
            return load_deltasvm_npz(path, device=device)
        return load_npz(path, device=device)
    from .importers import load_lsgkm_model

    return load_lsgkm_model(path, device=device)


def load_deltasvm_model(path, l: int, *, include_rc: bool | None = None, bias: float = 0.0,
                        dtype=np.float32, device: str = "rocm"):
    """A DeltaSVM from a deltaSVM weight file: ``kmer<TAB>weight`` lines, or
    ``kmer<TAB>revcomp<TAB>weight`` (RC-collapsed: both k-mers get the weight
    and include_rc defaults to False; a ``# ... include_rc=false`` comment
    does the same).  A first line without ACGT letters is a header.  Weights
    are stored at ``dtype`` (float32 by default, as deltaSVM files print
    them) and scored in float64."""
    from .deltasvm import DeltaSVM, kmer_to_index
    from .importers import _open

    k = None
    weights_of: dict = {}
    has_rc_column = header_rc_false = False
    with _open(path) as f:
        for line_num, line in enumerate(f, 1):
            line = line.strip()
            if not line:
                continue
            if line.startswith("#"):
                if "include_rc=false" in line.lower():
                    header_rc_false = True
                continue
            parts = line.split("\t")
            if len(parts) == 1:
                parts = line.split()
            if len(parts) == 3:
                kmer, revcomp, weight = parts
                has_rc_column = True
            elif len(parts) == 2:
                (kmer, weight), revcomp = parts, None
            else:
                raise ValueError(f"Line {line_num}: expected 2 or 3 columns, got: {line[:80]}")
            kmer = kmer.upper()
            if not all(b in "ACGT" for b in kmer):
                if line_num == 1 and not any(b in "ACGT" for b in kmer):
                    continue
                raise ValueError(f"Line {line_num}: invalid bases in k-mer '{kmer}'")
            if k is None:
                k = len(kmer)
            elif len(kmer) != k:
                raise ValueError(f"Line {line_num}: k-mer length {len(kmer)} != expected {k}")
            weights_of[kmer] = float(weight)
            if revcomp is not None:
                revcomp = revcomp.upper()
                if len(revcomp) == k and all(b in "ACGT" for b in revcomp):
                    weights_of[revcomp] = float(weight)
    if k is None:
        raise ValueError("Empty weight file")
    if l < k:
        raise ValueError(f"l ({l}) must be >= k ({k})")
    if include_rc is None:
        include_rc = not (has_rc_column or header_rc_false)
    weights = np.zeros(4**k, dtype=dtype)
    for kmer, w in weights_of.items():
        weights[kmer_to_index(kmer)] = w
    return DeltaSVM(weights, l, k, include_rc=include_rc, bias=bias, device=device)
