"""The CPU oracle pinned: LS-GKM golden outputs, hand values, closed forms,
independent recounts and cross-algorithm equality (no GPU)."""

import json
import random

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O


@pytest.mark.parametrize("kind", H.WGKM_TYPES)
@pytest.mark.parametrize("n", (100, 85))
def test_gkmpredict_fixtures(kind, n):
    """-t 0..6 (incl. w2 / w05 variants) scores equal LS-GKM gkmpredict."""
    m = O.load_lsgkm(H.path(f"lsgkm_wgkm_{kind}.model.txt.gz"))
    recs = O.read_fasta(H.path(f"lsgkm_wgkm_test{n}.fa"))
    codes, lens = O.batch([s for _, s in recs])
    pred = H.read_pred(H.path(f"lsgkm_wgkm_{kind}_test{n}.pred.txt"))
    want = np.array([pred[k] for k, _ in recs])
    np.testing.assert_allclose(O.score(m, codes, lens), want, rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("kind", ("t2", "t4", "t5"))
def test_gkmpredict_ragged(kind):
    m = O.load_lsgkm(H.path(f"lsgkm_ragged_{kind}.model.txt.gz"))
    recs = O.read_fasta(H.path("lsgkm_ragged_test.fa"))
    codes, lens = O.batch([s for _, s in recs])
    got = O.score(m, codes, lens)
    pred = H.read_pred(H.path(f"lsgkm_ragged_{kind}_test.pred.txt"))
    want = np.array([pred[k] for k, _ in recs])
    ok = ~np.isnan(got)
    assert (lens[~ok] < m.kern.l).all()
    np.testing.assert_allclose(got[ok], want[ok], rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("kind", ("est", "rbf"))
def test_gkmpredict_mode5_models(kind):
    m = O.load_lsgkm(H.path(f"lsgkm_mode5_{kind}.model.txt.gz"))
    recs = O.read_fasta(H.path("lsgkm_mode5_test51.fa"))
    codes, lens = O.batch([s for _, s in recs])
    want = np.array([float(ln.split("\t")[1]) for ln in open(H.path(f"lsgkm_mode5_{kind}.pred.txt"))])
    np.testing.assert_allclose(O.score(m, codes, lens), want, rtol=1e-4, atol=1e-6)


def test_encode_10sv():
    m = O.load_lsgkm(H.path("encode_ENCFF579AOX_10sv.model.txt.gz"))
    gold = json.load(open(H.path("encode_10sv_gkmpredict.json")))["scores"]
    random.seed(42)
    rand = "".join(random.choice("ACGT") for _ in range(300))
    qs = [m.sv_codes[i, : m.sv_lengths[i]] for i in range(3)] + [O.encode(rand)]
    codes, lens = O.batch(qs)
    want = [gold[k] for k in ("query_0", "query_1", "query_2", "random")]
    np.testing.assert_allclose(O.score(m, codes, lens), want, atol=1e-4)


@pytest.mark.parametrize("name", sorted(H.EXPLAIN_SETS))
@pytest.mark.parametrize("mode,lm", [("importance", 0), ("hypothetical", 1), ("perturbation", 4)])
def test_gkmexplain_fixtures(name, mode, lm):
    mf, fa, pre = H.EXPLAIN_SETS[name]
    m = O.load_lsgkm(H.path(mf))
    codes, lens = O.batch([s for _, s in O.read_fasta(H.path(fa))])
    got = O.explain(m, codes, lens, mode)
    exp = H.read_explain(H.path(f"{pre}.explain{lm}.txt.gz"))
    scale = max(np.abs(e).max() for e in exp)
    for i, e in enumerate(exp):
        np.testing.assert_allclose(got[i, :, : e.shape[1]], e, rtol=2e-5, atol=5e-6 * scale)
        assert (got[i, :, e.shape[1]:] == 0).all()


@pytest.mark.parametrize("name", sorted(k for k in H.EXPLAIN_SETS if not k.endswith("_51")))
def test_mode5_fixtures(name):
    mf, fa, pre = H.EXPLAIN_SETS[name]
    m = O.load_lsgkm(H.path(mf))
    seqs = [s for _, s in O.read_fasta(H.path(fa))]
    ref, alt = H.mode5_pairs(seqs, fa)
    rc, rl = O.batch(ref)
    ac, al = O.batch(alt)
    got = O.impact(m, rc, rl, ac, al).reshape(-1, 4)
    exp = np.stack([e[:, 0] for e in H.read_explain(H.path(f"{pre}.explain5.txt.gz"))])
    np.testing.assert_allclose(got, exp, rtol=2e-5, atol=5e-6 * np.abs(exp).max())


@pytest.mark.parametrize("kind", ("est", "rbf"))
def test_mode5_test51(kind):
    m = O.load_lsgkm(H.path(f"lsgkm_mode5_{kind}.model.txt.gz"))
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_mode5_test51.fa"))]
    ref, alt = H.mode5_pairs(seqs, "test51")
    rows = [ln.split("\t") for ln in open(H.path(f"lsgkm_mode5_{kind}.txt"))]
    want = np.array([[float(v) for v in r[2].split(",")] for r in rows])
    rc, rl = O.batch(ref)
    ac, al = O.batch(alt)
    got = O.impact(m, rc, rl, ac, al).reshape(-1, 4)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6)


def _raw_pair(kern, a: str, b: str):
    pa = O.pack_both(kern, *O.batch([a]))
    pb = O.pack_both(kern, *O.batch([b]))
    return O.combine(kern, O.pair_hist(kern, pa, pb))[0, 0]


def _raw_self(kern, a: str):
    return O.combine(kern, O.self_hist(kern, O.pack_both(kern, *O.batch([a]))))[0]


def test_hand_values_l3_k2():
    """tests/test_kernels/test_direct.py:9-62 of the reference: l=3, k=2, no RC."""
    kern = O.Kern.make(0, {"L": 3, "k": 2, "include_rc": False})
    assert kern.table.tolist() == [3, 1, 0, 0]
    assert _raw_self(kern, "AAA") == 3.0
    assert _raw_self(kern, "AAAA") == 12.0
    assert _raw_pair(kern, "AAA", "CCC") == 0.0
    assert _raw_pair(kern, "AAA", "AAC") == 1.0
    assert _raw_pair(kern, "ACGT", "ACGA") == _raw_pair(kern, "ACGA", "ACGT")


def test_count_table_and_d_cut():
    full = O.Kern.make(0, {"L": 11, "k": 7})
    trunc = O.Kern.make(0, {"L": 11, "k": 7, "d": 3})
    np.testing.assert_array_equal(full.table[:5], [330, 120, 36, 8, 1])
    np.testing.assert_array_equal(trunc.table[:5], [330, 120, 36, 8, 0])
    assert full.dmax == 4 and trunc.dmax == 3


@pytest.mark.parametrize("l,k,a_printed", [
    (11, 7, [1.77e-3, -2.65e-3, 7.09e-3, -4.82e-3]),
    (10, 6, [2.78e-3, -7.28e-3, 2.12e-2, -4.84e-2]),
])
def test_estimated_table_masked_join_coefficients(l, k, a_printed):
    """h(m) = sum_t a_t C(l - m, j_t), j_t = l - dmax + t: solving for a_t from
    the -t 2 table must give the coefficients the reference's design notes
    print (docs/design/mask-join-kernel.md:38-42, 3 digits)."""
    from math import comb

    h = O.Kern.make(2, {"L": l, "k": k}).table
    dmax = 3
    j = [l - dmax + t for t in range(dmax + 1)]
    A = np.array([[comb(l - m, jt) for jt in j] for m in range(dmax + 1)], dtype=np.float64)
    a = np.linalg.solve(A, h[: dmax + 1])
    np.testing.assert_allclose(a, a_printed, rtol=6e-3)
    assert (h[dmax + 1:] == 0).all()


@pytest.mark.parametrize("seed", range(6))
def test_histograms_match_onehot_recount(seed):
    """orc_pair_hist equals mismatch counts recomputed from one-hot windows in
    numpy (an independent count: per-position base equality summed over the
    window), fwd and RC, with PAD rows."""
    rng = np.random.default_rng(seed)
    l = int(rng.choice([3, 5, 8, 11]))
    kern = O.Kern.make(0, {"L": l, "k": 1})  # table nonzero at every m
    xs = ["".join(rng.choice(list("ACGT"), int(n))) for n in rng.integers(l, 40, 4)]
    ys = ["".join(rng.choice(list("ACGT"), int(n))) for n in rng.integers(l, 40, 5)]
    px = O.pack_both(kern, *O.batch(xs))
    py = O.pack_both(kern, *O.batch(ys))
    got = O.pair_hist(kern, px, py)
    comp = str.maketrans("ACGT", "TGCA")
    for i, x in enumerate(xs):
        for j, y in enumerate(ys):
            want = np.zeros(kern.dmax + 1, np.uint64)
            for yy in (y, y.translate(comp)[::-1]):
                for a in range(len(x) - l + 1):
                    for b in range(len(yy) - l + 1):
                        m = sum(x[a + q] != yy[b + q] for q in range(l))
                        if m <= kern.dmax:
                            want[m] += 1
            np.testing.assert_array_equal(got[i, j], want)


@pytest.mark.parametrize("kt", [2, 3, 4, 5, 0])
def test_index_equals_brute_force(kt):
    """The l-mer index (the reference's CPU scoring algorithm) and the brute
    force pair loop give one answer."""
    rng = np.random.default_rng(kt)
    p = {"L": 8, "k": 5, "gamma": 1.3, "M": 20, "H": 10.0}
    S = 60
    sv = rng.integers(0, 4, (S, 50)).astype(np.uint8)
    svl = rng.integers(20, 51, S)
    m = O.Model(O.Kern.make(kt, p), sv, svl, rng.normal(size=S), 0.3)
    q = rng.integers(0, 4, (7, 70)).astype(np.uint8)
    ql = rng.integers(5, 71, 7)
    a = O.score(m, q, ql)
    b = O.LmerIndex(m).score(q, ql)
    np.testing.assert_allclose(a, b, rtol=1e-12, equal_nan=True)


@pytest.mark.parametrize("mf", ["lsgkm_wgkm_t0.model.txt.gz", "lsgkm_mode5_est.model.txt.gz",
                                "lsgkm_wgkm_t4.model.txt.gz"])
def test_importance_completeness(mf):
    """Importance sums to score - bias for linear kernels (tests/test_explain.py:271-368)."""
    m = O.load_lsgkm(H.path(mf))
    codes, lens = O.batch([s for _, s in O.read_fasta(H.path("lsgkm_wgkm_test85.fa"))][:4])
    imp = O.explain(m, codes, lens, "importance")
    np.testing.assert_allclose(imp.sum(axis=(1, 2)), O.score(m, codes, lens) - m.bias,
                               rtol=1e-10)


def test_ism_reference_base_and_padding():
    m = O.load_lsgkm(H.path("lsgkm_mode5_rbf.model.txt.gz"))
    seqs = ["ACGTTGCAACGTAGGCTAGCTAGGATCGATCGTTAGC", "ACGT", "TTGACGNNACGTAGCTAGCGATTAGCAT"]
    codes, lens = O.batch(seqs)
    d = O.ism(m, codes, lens)
    assert np.isnan(d[1]).all()
    for b in (0, 2):
        for p in range(lens[b]):
            assert d[b, codes[b, p], p] == 0.0
        assert (d[b, :, lens[b]:] == 0).all()


def test_planted_table_error_is_caught(monkeypatch):
    """A plausible mistake (one pair weight off by 1%) in the oracle's -t 2
    table makes the gkmpredict comparison fail: the fixtures pin the table."""
    m = O.load_lsgkm(H.path("lsgkm_wgkm_t2.model.txt.gz"))
    m.kern.table = m.kern.table.copy()
    m.kern.table[1] *= 1.01
    m.__post_init__()
    recs = O.read_fasta(H.path("lsgkm_wgkm_test100.fa"))
    codes, lens = O.batch([s for _, s in recs])
    pred = H.read_pred(H.path("lsgkm_wgkm_t2_test100.pred.txt"))
    want = np.array([pred[k] for k, _ in recs])
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(O.score(m, codes, lens), want, rtol=2e-5, atol=1e-5)


def test_planted_rc_error_is_caught():
    """Dropping the RC block (a plausible omission) breaks the fixture match."""
    m = O.load_lsgkm(H.path("lsgkm_wgkm_t2.model.txt.gz"))
    m.kern.rc = False
    m.__post_init__()
    recs = O.read_fasta(H.path("lsgkm_wgkm_test100.fa"))
    codes, lens = O.batch([s for _, s in recs])
    pred = H.read_pred(H.path("lsgkm_wgkm_t2_test100.pred.txt"))
    want = np.array([pred[k] for k, _ in recs])
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(O.score(m, codes, lens), want, rtol=2e-5, atol=1e-5)
