"""LS-GKM model files, classic / R gkmSVM model files and FASTA (the formats
the fixtures and users hold; gkmsvm-lite importers/lsgkm.py:23-97,
importers/__init__.py:15-113, importers/classic.py, importers/r_gkmsvm.py,
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


# gkmsvm_kernel / gkmsvm_classify -t (R useTgkm) as kernel types: full filter,
# truncated filter (the tools' default) and gkm counts, each cut at d
# mismatches (importers/classic.py:24-27); 3 and 4 have no equivalent.
GKMSVM_FILTERS = {0: "gkm_estfull", 1: "gkm_esttrunc", 2: "gkm_cnt"}


def load_classic_model(alpha_path, svseq_path=None, *, L: int = 10, k: int = 6, d: int = 3,
                       filter_type: int = 1, include_rc: bool = True,
                       sv_chunk_size: int | None = None, device: str = "rocm"):
    """GkmSVM from the two files classic (C++) gkmsvm_train writes
    (importers/classic.py:30-127): PREFIX_svalpha.out (``id <TAB> alpha``) and
    PREFIX_svseq.fa (default: next to it).  The files hold no bias and no
    kernel parameters, so L, k, d, filter_type (-t) and include_rc are
    arguments; SV letters other than ACGT are dropped, as the tools read them
    (with a warning).  Bias 0."""
    import re

    from .svm import GkmSVM

    alpha_path = Path(alpha_path)
    if svseq_path is None:
        name = alpha_path.name
        if not name.endswith("svalpha.out"):
            raise ValueError(f"give svseq_path for {alpha_path}")
        svseq_path = alpha_path.with_name(name[: -len("svalpha.out")] + "svseq.fa")
    ids, coefs = [], []
    with _open(alpha_path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            parts = line.split()
            if len(parts) != 2:
                raise ValueError(f"Expected 'id <TAB> alpha', got: {line[:80]}")
            ids.append(parts[0])
            coefs.append(float(parts[1]))
    by_id = {name.split()[0]: seq for name, seq in read_fasta(svseq_path)}
    missing = [i for i in ids if i not in by_id]
    if missing:
        raise ValueError(f"{len(missing)} SV ids of {alpha_path} are not in {svseq_path}, "
                         f"e.g. {missing[0]}")
    seqs = [re.sub("[^ACGTacgt]", "", by_id[i]).upper() for i in ids]
    dropped = sum(len(by_id[i]) != len(s) for i, s in zip(ids, seqs))
    if dropped:
        warnings.warn(f"{dropped} support vectors hold letters other than ACGT, dropped as "
                      "gkmSVM reads them (queries with N score as in gkmsvm_classify only "
                      "with those letters dropped too)", stacklevel=2)
    if filter_type not in GKMSVM_FILTERS:
        raise NotImplementedError(
            f"gkmSVM filter type {filter_type} (wildcard/mismatch kernels) is not supported; "
            f"use {sorted(GKMSVM_FILTERS)}")
    if not seqs:
        raise ValueError("No support vector sequences found")
    codes, lengths = as_batch([encode_codes(s) for s in seqs])
    params = {"L": L, "k": k, "d": d, "include_rc": bool(include_rc)}
    return GkmSVM(codes, np.array(coefs, dtype=np.float64), 0.0, GKMSVM_FILTERS[filter_type],
                  params, sv_lengths=lengths, sv_chunk_size=sv_chunk_size, device=device)


def load_r_gkmsvm_model(path, svseq_path=None, *, L: int = 10, K: int = 6, maxnmm: int = 3,
                        useTgkm: int = 1, addRC: bool = True, sv_chunk_size: int | None = None,
                        device: str = "rocm"):
    """GkmSVM from the files R gkmSVM's gkmsvm_train writes (the classic
    format, importers/r_gkmsvm.py:17-56) with the R package's argument names."""
    return load_classic_model(path, svseq_path, L=L, k=K, d=maxnmm, filter_type=useTgkm,
                              include_rc=addRC, sv_chunk_size=sv_chunk_size, device=device)
