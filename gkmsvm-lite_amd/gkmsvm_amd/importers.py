"""Readers for the model and sequence files users already hold.

* LS-GKM models (what lsgkm's gkmtrain writes; plain or gzipped): a
  ``key value`` header, a line ``SV``, then ``coefficient SEQUENCE`` lines.
  The decision value is sum_i coef_i K(x, sv_i) - rho, so bias = -rho
  (the reference reads the same format in importers/lsgkm.py:23-97).
* Classic C++ / R gkmSVM models: PREFIX_svalpha.out (``id alpha`` lines) and
  PREFIX_svseq.fa.  They hold neither a bias nor the kernel's parameters,
  which are arguments here (importers/classic.py, importers/r_gkmsvm.py).
* deltaSVM weight files: ``kmer weight`` or ``kmer revcomp weight`` lines
  (importers/deltasvm.py:21-115).
* FASTA (fasta.py:19-42).

Files are read in one pass; gzip is recognised by its magic bytes.
"""

from __future__ import annotations

import gzip
import warnings
from pathlib import Path

import numpy as np

from .codec import encode_strings
from .kernels import KERNEL_ALIASES


def open_text(path):
    """A text handle on path, through gzip when the file starts with its
    magic bytes."""
    with open(path, "rb") as raw:
        gz = raw.read(2) == b"\x1f\x8b"
    return gzip.open(path, "rt") if gz else open(path, "r")


def read_fasta(path) -> list[tuple[str, str]]:
    """[(header without '>', sequence)] of a plain or gzipped FASTA file; a
    record's sequence lines are joined, blank lines skipped."""
    records: list[tuple[str, list[str]]] = []
    with open_text(path) as f:
        for raw in f:
            text = raw.strip()
            if not text:
                continue
            if text[0] == ">":
                records.append((text[1:].strip(), []))
            elif records:
                records[-1][1].append(text)
    return [(name, "".join(parts)) for name, parts in records]


# -- LS-GKM ------------------------------------------------------------------------------


def _kernel_name(text: str):
    """kernel_type as written: a name, or a -t number mapped to its name."""
    return KERNEL_ALIASES.get(int(text), text) if text.isdigit() else text


def _numbers(cast):
    return lambda text: [cast(v) for v in text.split()]


# header fields and how each is typed; unknown keys are skipped
LSGKM_FIELDS = {
    "svm_type": str, "kernel_type": _kernel_name,
    "L": int, "k": int, "d": int, "M": int, "nr_class": int, "total_sv": int, "norc": int,
    "gamma": float, "H": float,
    "rho": _numbers(float), "probA": _numbers(float), "probB": _numbers(float),
    "label": _numbers(int), "nr_sv": _numbers(int),
}


def _lsgkm_sections(path):
    """(header dict, [(coefficient, sequence)]) of an LS-GKM model file."""
    header: dict = {"norc": 0}
    svs: list[tuple[float, str]] = []
    in_svs = False
    with open_text(path) as f:
        for lineno, raw in enumerate(f, 1):
            fields = raw.split()
            if not fields:
                continue
            if in_svs:
                if len(fields) < 2:
                    raise ValueError(f"{path}:{lineno}: an SV line needs a coefficient and a "
                                     f"sequence, got {raw.strip()[:60]!r}")
                svs.append((float(fields[0]), fields[1]))
            elif fields[0] == "SV":
                in_svs = True
            elif len(fields) >= 2 and fields[0] in LSGKM_FIELDS:
                key = fields[0]
                header[key] = LSGKM_FIELDS[key](raw.strip()[len(key):].strip())
    return header, svs


def read_lsgkm_header(path) -> dict:
    """The typed header of an LS-GKM model file (norc defaults to 0)."""
    return _lsgkm_sections(path)[0]


_ACGTN = frozenset("ACGTN")


def _model_from_strings(seqs, coefs, bias, kernel_type, params, *, sv_chunk_size, device):
    """GkmSVM of SV strings (letters other than ACGT read as A, with the
    reverse strand reading T, as LS-GKM packs them)."""
    from .svm import GkmSVM

    codes, lens, _ = encode_strings(seqs)
    given = None if bool((lens == codes.shape[1]).all()) else lens
    return GkmSVM(codes, np.asarray(coefs, dtype=np.float64), bias, kernel_type, params,
                  sv_lengths=given, sv_chunk_size=sv_chunk_size, device=device)


def load_lsgkm_model(path, *, sv_chunk_size: int | None = None, device: str = "rocm"):
    """GkmSVM of an LS-GKM model file (.txt or .txt.gz): kernel type and L,
    k, d, gamma, M, H, norc from the header, bias = -rho.  ValueError for a
    file without kernel_type, total_sv or rho, without SVs, or whose SV
    count differs from total_sv; NotImplementedError unless nr_class is 2."""
    header, svs = _lsgkm_sections(path)
    absent = [key for key in ("kernel_type", "total_sv", "rho") if key not in header]
    if absent:
        raise ValueError(f"{path} is not an LS-GKM model: no {', '.join(absent)} in its header")
    if header.get("nr_class", 2) != 2:
        raise NotImplementedError(
            f"{path} has {header['nr_class']} classes; only two-class models (nr_class 2) load")
    if not svs:
        raise ValueError(f"{path} lists no support vectors")
    if len(svs) != header["total_sv"]:
        raise ValueError(f"{path} declares total_sv {header['total_sv']} but lists {len(svs)} "
                         "support vectors")
    seqs = [s for _, s in svs]
    odd = sum(1 for s in seqs if not set(s.upper()) <= _ACGTN)
    if odd:
        warnings.warn(f"{odd} support vectors of {path} hold letters beyond ACGTN; they are "
                      "read as N (A on the forward strand), as LS-GKM reads them", stacklevel=2)
    params = {"L": header["L"], "k": header["k"], "include_rc": header["norc"] == 0}
    params.update({key: header[key] for key in ("d", "gamma", "M", "H") if key in header})
    return _model_from_strings(seqs, [c for c, _ in svs], -header["rho"][0],
                               header["kernel_type"], params, sv_chunk_size=sv_chunk_size,
                               device=device)


# -- classic C++ and R gkmSVM --------------------------------------------------------------

# gkmsvm_kernel / gkmsvm_classify -t (R: useTgkm) as kernel types, each cut
# at d mismatches: full filter, truncated filter (the tools' default), gkm
# counts.  3 (wildcard) and 4 (mismatch kernel) have no counterpart here.
GKMSVM_FILTERS = {0: "gkm_estfull", 1: "gkm_esttrunc", 2: "gkm_cnt"}

_ACGT = frozenset(b"ACGTacgt")
_NOT_ACGT = bytes(b for b in range(256) if b not in _ACGT)


def _acgt_only(seq: str) -> str:
    """seq without its non-ACGT letters, upper case (how gkmSVM's FASTA
    reader sees it)."""
    return seq.encode("latin-1", "replace").translate(None, _NOT_ACGT).decode().upper()


def _svalpha(path) -> list[tuple[str, float]]:
    pairs = []
    with open_text(path) as f:
        for lineno, raw in enumerate(f, 1):
            fields = raw.split()
            if not fields:
                continue
            if len(fields) != 2:
                raise ValueError(f"{path}:{lineno}: expected an SV id and its alpha, got "
                                 f"{raw.strip()[:60]!r}")
            pairs.append((fields[0], float(fields[1])))
    return pairs


def load_classic_model(alpha_path, svseq_path=None, *, L: int = 10, k: int = 6, d: int = 3,
                       filter_type: int = 1, include_rc: bool = True,
                       sv_chunk_size: int | None = None, device: str = "rocm"):
    """GkmSVM of the two files gkmsvm_train (classic C++, or the R package)
    writes: PREFIX_svalpha.out and PREFIX_svseq.fa (by default the file
    beside alpha_path).  L, k, d, filter_type (-t) and include_rc are the
    kernel's parameters, which the files do not hold; bias 0.  SVs are
    matched to alphas by FASTA id and lose their non-ACGT letters, as the
    tools read them (with a warning)."""
    alpha_path = Path(alpha_path)
    if filter_type not in GKMSVM_FILTERS:
        raise NotImplementedError(
            f"filter type {filter_type} is not a gkm kernel this engine computes "
            f"(supported: {sorted(GKMSVM_FILTERS)})")
    if svseq_path is None:
        stem = alpha_path.name
        if not stem.endswith("svalpha.out"):
            raise ValueError(f"{alpha_path} is not named PREFIX_svalpha.out: pass svseq_path")
        svseq_path = alpha_path.with_name(stem[: -len("svalpha.out")] + "svseq.fa")
    pairs = _svalpha(alpha_path)
    if not pairs:
        raise ValueError(f"{alpha_path} lists no support vectors")
    fasta = {}
    for name, seq in read_fasta(svseq_path):
        fasta[name.split()[0] if name.split() else name] = seq
    absent = [sv_id for sv_id, _ in pairs if sv_id not in fasta]
    if absent:
        raise ValueError(f"{svseq_path} lacks {len(absent)} of the SV ids in {alpha_path} "
                         f"(first: {absent[0]}); they are not in the FASTA file")
    raw = [fasta[sv_id] for sv_id, _ in pairs]
    seqs = [_acgt_only(s) for s in raw]
    changed = sum(len(a) != len(b) for a, b in zip(raw, seqs))
    if changed:
        warnings.warn(f"{changed} support vectors held letters other than ACGT, dropped as "
                      "gkmSVM drops them (queries match gkmsvm_classify only when prepared the "
                      "same way)", stacklevel=2)
    params = {"L": L, "k": k, "d": d, "include_rc": bool(include_rc)}
    return _model_from_strings(seqs, [a for _, a in pairs], 0.0, GKMSVM_FILTERS[filter_type],
                               params, sv_chunk_size=sv_chunk_size, device=device)


def load_r_gkmsvm_model(path, svseq_path=None, *, L: int = 10, K: int = 6, maxnmm: int = 3,
                        useTgkm: int = 1, addRC: bool = True, sv_chunk_size: int | None = None,
                        device: str = "rocm"):
    """load_classic_model under the R gkmSVM package's argument names (L, K,
    maxnmm, useTgkm, addRC)."""
    return load_classic_model(path, svseq_path, L=L, k=K, d=maxnmm, filter_type=useTgkm,
                              include_rc=addRC, sv_chunk_size=sv_chunk_size, device=device)


# -- deltaSVM weight files ---------------------------------------------------------------


def _is_kmer(text: str) -> bool:
    return bool(text) and not text.strip("ACGT")


def load_deltasvm_model(path, l: int, *, include_rc: bool | None = None, bias: float = 0.0,
                        dtype=np.float32, device: str = "rocm"):
    """DeltaSVM of a deltaSVM weight file.  Lines hold ``kmer weight``, or
    ``kmer revcomp weight`` when the file is collapsed by reverse
    complement (both k-mers then get the weight, and include_rc defaults to
    False; so does a ``#`` comment containing ``include_rc=false``).
    Columns split on tabs, else on spaces; a first line with no A/C/G/T
    k-mer is a header.  Weights are kept at `dtype` (deltaSVM prints single
    precision) and scored in float64.  ValueError for a malformed line, a
    k-mer of other letters or another length, an empty file, or l < k."""
    from .deltasvm import DeltaSVM, kmer_to_index

    table: dict[str, float] = {}
    k = None
    collapsed = comment_no_rc = False
    with open_text(path) as f:
        for lineno, raw in enumerate(f, 1):
            text = raw.strip()
            if not text:
                continue
            if text[0] == "#":
                comment_no_rc |= "include_rc=false" in text.lower()
                continue
            cols = text.split("\t") if "\t" in text else text.split()
            if len(cols) not in (2, 3):
                raise ValueError(f"{path}:{lineno}: expected 2 or 3 columns, found {len(cols)}")
            kmer = cols[0].upper()
            if not _is_kmer(kmer):
                if lineno == 1 and not any(c in "ACGT" for c in kmer):
                    continue  # a header line
                raise ValueError(f"{path}:{lineno}: {cols[0]!r} is not a k-mer of A, C, G, T")
            if k is None:
                k = len(kmer)
            if len(kmer) != k:
                raise ValueError(f"{path}:{lineno}: k-mer of length {len(kmer)} in a file of "
                                 f"length-{k} k-mers")
            w = float(cols[-1])
            table[kmer] = w
            if len(cols) == 3:
                collapsed = True
                partner = cols[1].upper()
                if len(partner) == k and _is_kmer(partner):
                    table[partner] = w
    if k is None:
        raise ValueError(f"{path} holds no k-mer weights")
    if l < k:
        raise ValueError(f"window length l = {l} must be >= the file's k = {k}")
    if include_rc is None:
        include_rc = not (collapsed or comment_no_rc)
    weights = np.zeros(4**k, dtype=dtype)
    for kmer, w in table.items():
        weights[kmer_to_index(kmer)] = w
    return DeltaSVM(weights, l, k, include_rc=include_rc, bias=bias, device=device)
