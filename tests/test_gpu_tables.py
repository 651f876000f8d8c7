"""GPU parity of the exact l-mer tables (score table, explain table, ISM and
mutation impact through them, DeltaSVM) against the CPU oracle and the LS-GKM
golden files.  Tables are FP64 sums in a different order from the per-SV
path, so values are held to 1e-10 relative (helpers.assert_close)."""

import numpy as np
import pytest
import torch

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TABLE_TYPES = ("t0", "t1", "t2", "t2w2", "t4", "t4w05")


def _load(name, table=True, explain_table=True):
    from gkmsvm_amd import load_lsgkm_model

    m = load_lsgkm_model(H.path(name))
    m.use_lmer_table = table
    m.use_explain_table = explain_table
    return m


def _oracle_model(m):
    from gkmsvm_amd.inputs import row_lengths

    kern = O.Kern.make(m.kernel_type, m.kernel_params)
    return O.Model(kern, m.support_sequences, row_lengths(m.support_sequences, m.sv_lengths),
                   m.coefficients, m.bias)


def _rand_model(kt, l, k, seed, S=40, lo=None, hi=60, extra=None):
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(seed)
    lo = l if lo is None else lo
    lens = rng.integers(lo, hi + 1, S)
    codes = np.zeros((S, hi), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 2 if i % 3 == 0 else 4, lens[i])
    p = {"L": l, "k": k, "include_rc": bool(seed % 2), **(extra or {})}
    m = GkmSVM(codes, rng.normal(size=S), float(rng.normal()), kt, p, sv_lengths=lens)
    m.use_lmer_table = True
    m.use_explain_table = True
    return m


def _rand_queries(rng, n, lo, hi):
    lens = rng.integers(lo, hi + 1, n)
    codes = np.zeros((n, hi), np.uint8)
    for i in range(n):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    return codes, lens.astype(np.int64)


@pytest.mark.parametrize("kind", TABLE_TYPES)
def test_score_table_equals_masked_marginals(gpu, kind):
    """The GPU table (Walsh-Hadamard) equals the oracle's masked marginals
    (lmer_table.py:60-98) on the LS-GKM fixture models (l = 10)."""
    m = _load(f"lsgkm_wgkm_{kind}.model.txt.gz")
    got = m._build_lmer_table().cpu().numpy()
    want = O.lmer_table(_oracle_model(m))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("seed,kt,l,k,extra", [
    (0, 2, 4, 2, None), (1, 0, 5, 3, None), (2, 4, 6, 4, {"M": 11, "H": 5.0}),
    (3, 1, 6, 3, None), (4, 0, 3, 2, {"d": 1}),
])
def test_tables_equal_pairwise_definitions(gpu, seed, kt, l, k, extra):
    """Score and explain tables of random small-l models against their
    definitions over every pair of l-mers."""
    m = _rand_model(kt, l, k, seed, extra=extra)
    om = _oracle_model(m)
    want = O.lmer_table_brute(om)
    got = m._build_lmer_table().cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    if l <= 5:
        from gkmsvm_amd import gkmexplain

        rng = np.random.default_rng(seed + 100)
        codes, lens = _rand_queries(rng, 5, l, 40)
        for pert in (False, True):
            A = m._explain_table_for(1, np.array([l]), pert).cpu().numpy()
            Aw = O.explain_table_brute(om, pert)
            np.testing.assert_allclose(A, Aw, rtol=0, atol=1e-12 * np.abs(Aw).max())
            mode = "perturbation" if pert else "hypothetical"
            got = gkmexplain(m, codes, mode=mode, lengths=lens).cpu().numpy()
            H.assert_close(got, O.explain(om, codes, lens, mode), 1e-10, f"l={l} {mode}")


@pytest.mark.parametrize("kind", TABLE_TYPES)
@pytest.mark.parametrize("n", (100, 85))
def test_table_scores_lsgkm_fixtures(gpu, kind, n):
    """Scoring through the table equals the oracle's per-SV scores (1e-10) and
    gkmpredict (printed digits); test85 holds N stretches."""
    m = _load(f"lsgkm_wgkm_{kind}.model.txt.gz")
    recs = O.read_fasta(H.path(f"lsgkm_wgkm_test{n}.fa"))
    seqs = [s for _, s in recs]
    got = m(seqs).cpu().numpy().ravel()
    assert "score" in m._svs().tables
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.score(_oracle_model(m), codes, lens), 1e-10, kind)
    pred = H.read_pred(H.path(f"lsgkm_wgkm_{kind}_test{n}.pred.txt"))
    np.testing.assert_allclose(got, [pred[nm] for nm, _ in recs], rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("kind", ("t2", "t4"))
def test_table_ragged_models(gpu, kind):
    """Ragged SVs and 30-200 bp queries (some shorter than l: NaN)."""
    m = _load(f"lsgkm_ragged_{kind}.model.txt.gz")
    recs = O.read_fasta(H.path("lsgkm_ragged_test.fa"))
    seqs = [s for _, s in recs]
    got = m(seqs).cpu().numpy().ravel()
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.score(_oracle_model(m), codes, lens), 1e-10, kind)
    pred = H.read_pred(H.path(f"lsgkm_ragged_{kind}_test.pred.txt"))
    want = np.array([pred[nm] for nm, _ in recs])
    ok = ~np.isnan(got)
    np.testing.assert_allclose(got[ok], want[ok], rtol=2e-5, atol=1e-5)


TABLE_EXPLAIN = sorted(k for k, v in H.EXPLAIN_SETS.items()
                       if not any(t in v[0] for t in ("t3", "t5", "t6", "rbf")))


@pytest.mark.parametrize("name", TABLE_EXPLAIN)
@pytest.mark.parametrize("mode,lm", [("importance", 0), ("hypothetical", 1), ("perturbation", 4)])
def test_explain_table_fixtures(gpu, name, mode, lm):
    """GkmExplain from the explain table equals the oracle's per-SV loop (1e-10)
    and lsgkm gkmexplain's output (printed precision)."""
    from gkmsvm_amd import gkmexplain

    mf, fa, pre = H.EXPLAIN_SETS[name]
    m = _load(mf)
    seqs = [s for _, s in O.read_fasta(H.path(fa))]
    got = gkmexplain(m, seqs, mode=mode).cpu().numpy()
    assert ("explain", mode == "perturbation") in m._svs().tables
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.explain(_oracle_model(m), codes, lens, mode), 1e-10, f"{name} {mode}")
    exp = H.read_explain(H.path(f"{pre}.explain{lm}.txt.gz"))
    scale = max(np.abs(e).max() for e in exp)
    for i, e in enumerate(exp):
        np.testing.assert_allclose(got[i, :, : e.shape[1]], e, rtol=2e-5, atol=5e-6 * scale)


@pytest.mark.parametrize("mf", ["lsgkm_wgkm_t0.model.txt.gz", "lsgkm_wgkm_t1.model.txt.gz",
                                "lsgkm_mode5_est.model.txt.gz", "lsgkm_wgkm_t4.model.txt.gz",
                                "lsgkm_ragged_t4.model.txt.gz"])
def test_table_ism_matches_rescoring(gpu, mf):
    """ISM through the table equals f(mutant) - f(x) with every mutant scored
    from scratch by the oracle; the reference base is exactly 0."""
    from gkmsvm_amd import ism

    m = _load(mf)
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_ragged_query.fa"))][:6]
    got = ism(m, seqs).cpu().numpy()
    assert "score" in m._svs().tables
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.ism(_oracle_model(m), codes, lens), 1e-9, mf)
    for b, s in enumerate(seqs):
        for pos in range(len(s)):
            assert got[b, codes[b, pos], pos] == 0.0


@pytest.mark.parametrize("name", [k for k in TABLE_EXPLAIN if not k.endswith("_51")])
def test_table_mutation_impact_fixtures(gpu, name):
    """score_variants(method='gkmexplain') through the table equals the oracle
    and gkmexplain -m 5."""
    mf, fa, pre = H.EXPLAIN_SETS[name]
    m = _load(mf)
    seqs = [s for _, s in O.read_fasta(H.path(fa))]
    ref, alt = H.mode5_pairs(seqs, fa)
    got = m.score_variants(ref, alt, method="gkmexplain").cpu().numpy().ravel()
    rc, rl = O.batch(ref)
    ac, al = O.batch(alt)
    H.assert_close(got, O.impact(_oracle_model(m), rc, rl, ac, al), 1e-10, name)
    exp = np.stack([e[:, 0] for e in H.read_explain(H.path(f"{pre}.explain5.txt.gz"))])
    np.testing.assert_allclose(got.reshape(-1, 4), exp, rtol=2e-5, atol=5e-6 * np.abs(exp).max())


def test_table_and_per_sv_paths_agree_at_l11(gpu):
    """A 300-SV -t 2 l=11 model: table and per-SV scores, ISM and
    hypothetical attributions agree to 1e-10 on 300 bp queries."""
    from gkmsvm_amd import GkmSVM, gkmexplain, ism

    rng = np.random.default_rng(21)
    sv = rng.integers(0, 4, (300, 300)).astype(np.uint8)
    coef = rng.normal(size=300)
    codes, lens = _rand_queries(rng, 24, 200, 300)
    codes[3, 40:60] = 0  # an N stretch (read as A)
    seqs = ["".join("ACGT"[c] for c in codes[i, : lens[i]]) for i in range(len(lens))]
    m = GkmSVM(sv, coef, -0.2, 2, {"L": 11, "k": 7, "d": 3})
    m.use_lmer_table = m.use_explain_table = False
    a = m(seqs).cpu().numpy()
    ea = gkmexplain(m, seqs[:4], mode="hypothetical").cpu().numpy()
    ia = ism(m, seqs[:2]).cpu().numpy()
    m.use_lmer_table = m.use_explain_table = True
    b = m(seqs).cpu().numpy()
    eb = gkmexplain(m, seqs[:4], mode="hypothetical").cpu().numpy()
    ib = ism(m, seqs[:2]).cpu().numpy()
    H.assert_close(b, a, 1e-10, "scores")
    H.assert_close(eb, ea, 1e-10, "explain")
    H.assert_close(ib, ia, 1e-9, "ism")


def test_rbf_models_have_no_table(gpu):
    m = _load("lsgkm_wgkm_t5.model.txt.gz")
    assert m._build_lmer_table() is None
    assert m._explain_table_for(10, np.array([500] * 10), False) is None
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_wgkm_test85.fa"))][:5]
    codes, lens = O.batch(seqs)
    H.assert_close(m(seqs).cpu().numpy().ravel(), O.score(_oracle_model(m), codes, lens), 1e-10)


def test_auto_builds_the_table_when_a_call_pays_for_it(gpu):
    """Under "auto" a small call keeps the per-SV path; a call whose per-SV
    cost exceeds the build cost builds the table, and later calls use it."""
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(4)
    sv = rng.integers(0, 4, (2000, 400)).astype(np.uint8)
    m = GkmSVM(sv, rng.normal(size=2000), 0.0, 2, {"L": 11, "k": 7, "d": 3})
    one = ["".join("ACGT"[c] for c in rng.integers(0, 4, 100))]
    m(one)
    assert "score" not in m._svs().tables
    many = ["".join("ACGT"[c] for c in rng.integers(0, 4, 500)) for _ in range(400)]
    m(many)
    assert "score" in m._svs().tables
    m.clear_caches()
    assert "score" not in m._svs().tables


@pytest.mark.parametrize("kind,qn", [("t2", False), ("t2", True), ("t4", True), ("t0", False)])
def test_to_deltasvm_weights(gpu, kind, qn):
    """to_deltasvm weights equal the oracle's table read in DeltaSVM order
    (svm.py:667-705) and, with query_norm, score the l-mer alone."""
    m = _load(f"lsgkm_wgkm_{kind}.model.txt.gz")
    d = m.to_deltasvm(query_norm=qn)
    om = _oracle_model(m)
    want = O.deltasvm_weights(om, O.lmer_table(om), qn)
    np.testing.assert_allclose(d.weights, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert d.k == d.l == m.kernel.l and not d.include_rc and d.bias == m.bias


def test_to_deltasvm_rbf_scores_every_lmer(gpu):
    """An RBF model (no table) scores every canonical l-mer through the
    kernel; with RC an l-mer and its reverse complement get one weight."""
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(8)
    sv = rng.integers(0, 4, (30, 40)).astype(np.uint8)
    m = GkmSVM(sv, rng.normal(size=30), 0.3, 3, {"L": 5, "k": 3, "d": 2, "gamma": 1.5})
    d = m.to_deltasvm()
    l = 5
    idx = np.arange(4 ** l)
    digits = ((idx[:, None] // 4 ** np.arange(l - 1, -1, -1)) % 4).astype(np.uint8)
    want = O.score(_oracle_model(m), digits, np.full(len(idx), l)) - m.bias
    H.assert_close(d.weights, want, 1e-10, "rbf deltasvm")


@pytest.mark.parametrize("l,k,rc", [(11, 7, True), (6, 6, False), (8, 3, True)])
def test_deltasvm_scoring(gpu, l, k, rc):
    """DeltaSVM scores (gapped k-mer sums on both strands, each row at its own
    length) equal the oracle's restatement of deltasvm.py:22-44."""
    from gkmsvm_amd import DeltaSVM

    rng = np.random.default_rng(l * k)
    w = rng.normal(size=4 ** k)
    d = DeltaSVM(w, l, k, include_rc=rc, bias=0.25)
    codes, lens = _rand_queries(rng, 6, max(1, l - 2), 60)
    got = d(codes, lengths=lens).cpu().numpy().ravel()
    want = O.deltasvm_score(w, l, k, codes, lens, include_rc=rc, bias=0.25)
    H.assert_close(got, want, 1e-12, "deltasvm")
    v = d.score_variants(codes[:3], codes[3:6]).cpu().numpy().ravel()
    assert v.shape == (3,)


def test_table_score_deterministic(gpu):
    """Two builds of the same table are bit-identical, and a row scores the
    same alone and in a batch."""
    m = _load("lsgkm_wgkm_t2.model.txt.gz")
    t1 = m._build_lmer_table().clone()
    m.clear_caches()
    t2 = m._build_lmer_table()
    assert torch.equal(t1, t2)
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_wgkm_test85.fa"))]
    batch = m(seqs).cpu().numpy().ravel()
    alone = np.array([m([s]).cpu().numpy()[0, 0] for s in seqs[:5]])
    np.testing.assert_array_equal(batch[:5], alone)
