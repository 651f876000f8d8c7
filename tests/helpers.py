"""Shared helpers of the test suite (fixture sets, tolerances)."""

import gzip
import os

import numpy as np

from conftest import LSGKM

# Fixture models and the FASTA sets gkmpredict / gkmexplain were run on
# (the reference's tests/test_explain.py:680-713 pairs them the same way).
WGKM_TYPES = ("t0", "t1", "t2", "t2w2", "t3", "t4", "t4w05", "t5", "t6")
EXPLAIN_SETS = {}
for _t in ("t0", "t1", "t4", "t5", "t6"):
    EXPLAIN_SETS[f"{_t}_85"] = (f"lsgkm_wgkm_{_t}.model.txt.gz", "lsgkm_wgkm_test85.fa",
                                f"lsgkm_wgkm_{_t}_test85")
    EXPLAIN_SETS[f"{_t}_q"] = (f"lsgkm_wgkm_{_t}.model.txt.gz", "lsgkm_ragged_query.fa",
                               f"lsgkm_wgkm_{_t}_query")
for _t, _k in (("t2", "est"), ("t3", "rbf")):
    EXPLAIN_SETS[f"{_t}_85"] = (f"lsgkm_mode5_{_k}.model.txt.gz", "lsgkm_wgkm_test85.fa",
                                f"lsgkm_mode5_{_k}_test85")
    EXPLAIN_SETS[f"{_t}_51"] = (f"lsgkm_mode5_{_k}.model.txt.gz", "lsgkm_mode5_test51.fa",
                                f"lsgkm_mode5_{_k}")
    EXPLAIN_SETS[f"{_t}_q"] = (f"lsgkm_mode5_{_k}.model.txt.gz", "lsgkm_ragged_query.fa",
                               f"lsgkm_mode5_{_k}_query")
for _t in ("t2", "t4", "t5"):
    EXPLAIN_SETS[f"ragged_{_t}_85"] = (f"lsgkm_ragged_{_t}.model.txt.gz", "lsgkm_wgkm_test85.fa",
                                       f"lsgkm_ragged_{_t}_test85")
    EXPLAIN_SETS[f"ragged_{_t}_q"] = (f"lsgkm_ragged_{_t}.model.txt.gz",
                                      "lsgkm_ragged_query.fa", f"lsgkm_ragged_{_t}_query")


def path(name):
    return os.path.join(LSGKM, name)


def read_explain(p):
    """gkmexplain output rows: name, score, then 'a,c,g,t' per position split by ';'."""
    with gzip.open(p, "rt") as f:
        rows = [line.rstrip("\n").split("\t") for line in f]
    return [np.array([[float(v) for v in pos.split(",")] for pos in r[2].split(";")]).T
            for r in rows]


def read_pred(p):
    with open(p) as f:
        return dict((ln.split()[0], float(ln.split()[1])) for ln in f if ln.strip())


def mode5_pairs(seqs, fasta):
    """(ref, alt) lists: each sequence with its centre base set to A, C, G, T."""
    ref, alt = [], []
    for s in seqs:
        c = (len(s) - 1) // 2 if fasta == "lsgkm_ragged_query.fa" else len(s) // 2
        for b in "ACGT":
            ref.append(s)
            alt.append(s[:c] + b + s[c + 1:])
    return ref, alt


def assert_close(got, want, rtol=1e-10, what=""):
    """|got - want| <= rtol * max(|want|, max|want| * 1e-3): relative error,
    with a floor at a thousandth of the array's scale so entries that nearly
    cancel are held to the array's own precision."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, (got.shape, want.shape)
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    assert (nan_g == nan_w).all(), f"{what}: NaN pattern differs"
    g, w = got[~nan_w], want[~nan_w]
    if w.size == 0:
        return
    scale = np.abs(w).max()
    denom = np.maximum(np.abs(w), 1e-3 * scale)
    denom = np.where(denom == 0, 1.0, denom)
    err = np.abs(g - w) / denom
    assert err.max() <= rtol, f"{what}: max rel err {err.max():.3e} > {rtol:.0e}"
    return float(err.max())
