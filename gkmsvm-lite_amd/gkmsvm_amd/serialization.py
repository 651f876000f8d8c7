"""Saving and loading models: the native npz format and LS-GKM text.

npz files hold plain arrays -- ``support_sequences`` (base codes),
``coefficients``, ``sv_lengths`` when the SVs differ in length, or
``weights`` for a DeltaSVM -- and a ``metadata`` entry, a JSON record
stored as raw bytes, so they load with ``allow_pickle=False`` and are the
same files gkmsvm-lite writes and reads (serialization.py:15-166 there).
LS-GKM text is what lsgkm's gkmtrain writes (header order, rho = -bias,
``norc 1`` only when the reverse complement is off), plain or gzipped by
the file name.
"""

from __future__ import annotations

import gzip
import json
from pathlib import Path

import numpy as np

from .codec import BASES
from .importers import load_deltasvm_model  # noqa: F401  (re-exported)


def _numpy(x) -> np.ndarray:
    return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def _meta_entry(record: dict) -> np.void:
    return np.void(json.dumps(record).encode("utf-8"))


class _NpzModel:
    """An open npz model file: its arrays and decoded metadata record."""

    def __init__(self, path):
        self.data = np.load(path, allow_pickle=False)
        if "metadata" not in self.data.files:
            self.data.close()
            raise ValueError(f"{path} is not a model npz file (no metadata entry)")
        self.meta = json.loads(self.data["metadata"].tobytes())

    def optional(self, key):
        return self.data[key] if key in self.data.files else None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.data.close()


def save_npz(model, path) -> None:
    """Write a GkmSVM as native npz (arrays plus bias, kernel type, kernel
    parameters and sv_chunk_size in the metadata record)."""
    record = {"bias": float(model.bias), "kernel_type": model.kernel_type,
              "kernel_params": model.kernel_params}
    if model.sv_chunk_size is not None:
        record["sv_chunk_size"] = model.sv_chunk_size
    arrays = {"support_sequences": _numpy(model.support_sequences),
              "coefficients": _numpy(model.coefficients), "metadata": _meta_entry(record)}
    if model.sv_lengths is not None:
        arrays["sv_lengths"] = _numpy(model.sv_lengths)
    np.savez(path, **arrays)


def save_deltasvm_npz(model, path) -> None:
    """Write a DeltaSVM as npz: weights [4^k] and l, k, include_rc, bias."""
    record = {"model_type": "deltasvm", "l": model.l, "k": model.k,
              "include_rc": model.include_rc, "bias": float(model.bias)}
    np.savez(path, weights=_numpy(model.weights), metadata=_meta_entry(record))


def load_npz(path, *, device: str = "rocm"):
    """GkmSVM of a native npz file (one written without sv_lengths holds SVs
    of one length)."""
    from .svm import GkmSVM

    with _NpzModel(path) as f:
        m = f.meta
        return GkmSVM(f.data["support_sequences"], f.data["coefficients"], m["bias"],
                      m["kernel_type"], m["kernel_params"], sv_lengths=f.optional("sv_lengths"),
                      sv_chunk_size=m.get("sv_chunk_size"), device=device)


def load_deltasvm_npz(path, *, device: str = "rocm"):
    """DeltaSVM of an npz file save_deltasvm_npz wrote."""
    from .deltasvm import DeltaSVM

    with _NpzModel(path) as f:
        m = f.meta
        return DeltaSVM(f.data["weights"], m["l"], m["k"], include_rc=m["include_rc"],
                        bias=m.get("bias", 0.0), device=device)


_LETTERS = np.frombuffer(BASES.encode(), dtype=np.uint8)


def _lsgkm_lines(model):
    """The lines of an LS-GKM model file for model, header first."""
    p = model.kernel_params
    coefs = _numpy(model.coefficients)
    yield "svm_type c_svc"
    yield f"kernel_type {model.kernel_type}"
    yield f"L {p['L']}"
    yield f"k {p['k']}"
    for key in ("d", "gamma", "M", "H"):
        if key in p:
            yield f"{key} {p[key]}"
    if not p.get("include_rc", True):  # builds without the key reject it, so only when set
        yield "norc 1"
    n_pos = int(np.count_nonzero(coefs > 0))
    yield "nr_class 2"
    yield f"total_sv {coefs.size}"
    yield f"rho {-model.bias}"
    yield "label 1 -1"
    yield f"nr_sv {n_pos} {coefs.size - n_pos}"
    yield "SV"
    codes = _numpy(model.support_sequences)
    lens = (np.full(codes.shape[0], codes.shape[1]) if model.sv_lengths is None
            else _numpy(model.sv_lengths))
    for c, row, n in zip(coefs, codes, lens):
        yield f"{c} {_LETTERS[row[: int(n)]].tobytes().decode()}"


def save_lsgkm(model, path) -> None:
    """Write a GkmSVM as an LS-GKM text model (gzipped when path ends in .gz),
    each SV without its padding."""
    path = str(path)
    with (gzip.open(path, "wt") if path.endswith(".gz") else open(path, "w")) as f:
        for line in _lsgkm_lines(model):
            f.write(line)
            f.write("\n")


def load_model(path, *, device: str = "rocm"):
    """A model by file name: ``.npz`` is native (a DeltaSVM when its metadata
    says so), anything else LS-GKM text (``.txt``, ``.txt.gz``, ``.model``)."""
    if not Path(path).name.lower().endswith(".npz"):
        from .importers import load_lsgkm_model

        return load_lsgkm_model(path, device=device)
    with _NpzModel(path) as f:
        is_delta = f.meta.get("model_type") == "deltasvm"
    return load_deltasvm_npz(path, device=device) if is_delta else load_npz(path, device=device)
