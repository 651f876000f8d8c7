"""LS-GKM model files and FASTA (the formats the fixtures and users hold;
gkmsvm-lite importers/lsgkm.py:23-97, importers/__init__.py:15-113,
fasta.py:19-42)."""

from __future__ import annotations

import gzip
import warnings
from pathlib import Path

import numpy as np

from .codec import encode_codes
from .inputs import as_batch
from .kernels import KERNEL_ALIASES

_INT_KEYS = ("L", "k", "d", "M", "nr_class", "total_sv", "norc")
_FLOAT_KEYS = ("gamma", "H")
_LIST_KEYS = {"rho": float, "label": int, "nr_sv": int, "probA": float, "probB": float}


def _open(path):
    path = Path(path)
    with open(path, "rb") as f:
        magic = f.read(2)
    return gzip.open(path, "rt") if magic == b"\x1f\x8b" else open(path, "r")


def read_fasta(path):
    """[(name, sequence)] of a (possibly gzipped) FASTA file."""
    recs, name, seq = [], None, []
    with _open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith(">"):
                if name is not None:
                    recs.append((name, "".join(seq)))
                name, seq = line[1:].strip(), []
            else:
                seq.append(line)
    if name is not None:
        recs.append((name, "".join(seq)))
    return recs


def parse_lsgkm_header(path) -> dict:
    header = {"norc": 0}
    with _open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("SV"):
                break
            parts = line.split(None, 1)
            if len(parts) != 2:
                continue
            key, value = parts
            if key == "svm_type":
                header[key] = value
            elif key == "kernel_type":
                header[key] = KERNEL_ALIASES.get(int(value), value) if value.isdigit() else value
            elif key in _INT_KEYS:
                header[key] = int(value)
            elif key in _FLOAT_KEYS:
                header[key] = float(value)
            elif key in _LIST_KEYS:
                header[key] = [_LIST_KEYS[key](v) for v in value.split()]
    return header


def load_lsgkm_model(path, *, sv_chunk_size: int | None = None, device: str = "rocm"):
    """GkmSVM from an LS-GKM model file (.txt or .txt.gz); bias = -rho."""
    from .svm import GkmSVM

    header = parse_lsgkm_header(path)
    for key in ("kernel_type", "total_sv", "rho"):
        if header.get(key) is None:
            raise ValueError(f"Model file missing {key} field")
    if header.get("nr_class", 2) != 2:
        raise NotImplementedError("Only binary classification (nr_class=2) is supported")
    coefs, seqs, in_sv = [], [], False
    with _open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("SV"):
                in_sv = True
                continue
            if not in_sv:
                continue
            parts = line.split()
            if len(parts) < 2:
                raise ValueError(f"Malformed SV line: {line[:80]}")
            coefs.append(float(parts[0]))
            seqs.append(parts[1])
    if not seqs:
        raise ValueError("No support vector sequences found")
    if len(seqs) != header["total_sv"]:
        raise ValueError(f"Expected {header['total_sv']} support vectors, got {len(seqs)}")
    other = sum(bool(set(s.upper()) - set("ACGTN")) for s in seqs)
    if other:
        warnings.warn(f"{other} support vectors contain letters other than ACGTN; reading "
                      "them as N (scored as A, as LS-GKM does)", stacklevel=2)
    codes, lengths = as_batch([encode_codes(s, allow_other=True) for s in seqs])
    params = {"L": header["L"], "k": header["k"], "include_rc": header.get("norc", 0) == 0}
    for key in ("d", "gamma", "M", "H"):
        if key in header:
            params[key] = header[key]
    return GkmSVM(codes, np.array(coefs, dtype=np.float64), -header["rho"][0],
                  header["kernel_type"], params, sv_lengths=lengths,
                  sv_chunk_size=sv_chunk_size, device=device)
