"""The oracle's exact l-mer tables and DeltaSVM pinned on the CPU: the masked
marginals of lmer_table.py against the pairwise definition, table scoring
against the per-SV scores (and through them the LS-GKM gkmpredict files), the
explain-table definition against the per-SV GkmExplain loop, and DeltaSVM
weights against scoring each l-mer alone."""

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O


def _rand_model(kt, l, k, seed, S=12, lo=None, hi=30, rc=True, extra=None):
    rng = np.random.default_rng(seed)
    lo = l if lo is None else lo
    lens = rng.integers(lo, hi + 1, S)
    codes = np.zeros((S, hi), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    p = {"L": l, "k": k, "include_rc": rc, **(extra or {})}
    if kt != 0:
        p.setdefault("d", min(3, l - 1))
    kern = O.Kern.make(kt, p)
    return O.Model(kern, codes, lens, rng.normal(size=S), float(rng.normal()))


def _rand_queries(rng, n, lo, hi):
    lens = rng.integers(lo, hi + 1, n)
    codes = np.zeros((n, hi), np.uint8)
    for i in range(n):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    return codes, lens


@pytest.mark.parametrize("kt,l,k,rc,extra", [
    (0, 5, 3, True, None), (0, 6, 3, False, {"d": 2}), (2, 6, 4, True, None),
    (1, 5, 3, True, None), (4, 6, 3, True, {"M": 7, "H": 3.0}), (2, 4, 2, False, None),
])
def test_masked_marginals_equal_pairwise_definition(kt, l, k, rc, extra):
    """w(u) by masked marginals (lmer_table.py:60-98) equals
    sum_v V[v] h[m(u, v)] over every pair of l-mers."""
    m = _rand_model(kt, l, k, seed=l * 10 + kt, rc=rc, extra=extra)
    a = O.lmer_table(m)
    b = O.lmer_table_brute(m)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_lmer_weights_count_windows():
    """V sums to sum_sv c_sv x (number of windows, doubled with RC) for
    unweighted kernels, and a planted l-mer's entry holds its SV's c."""
    m = _rand_model(2, 5, 3, seed=1)
    V = O.lmer_weights(m)
    c = m.coef / np.sqrt(np.clip(m.sv_diag, 1e-10, None))
    nwin = np.clip(m.sv_lengths - 5 + 1, 0, None)
    np.testing.assert_allclose(V.sum(), (c * nwin * 2).sum(), rtol=1e-12)
    codes = np.array([[0, 1, 2, 3, 0]], np.uint8)
    one = O.Model(m.kern, codes, np.array([5]), np.array([2.0]), 0.0)
    V1 = O.lmer_weights(one)
    key = 0 | (1 << 2) | (2 << 4) | (3 << 6) | (0 << 8)
    rc_key = 3 | (0 << 2) | (1 << 4) | (2 << 6) | (3 << 8)  # revcomp of ACGTA = TACGT
    c1 = 2.0 / np.sqrt(one.sv_diag[0])
    assert V1[key] == pytest.approx(c1)
    assert V1[rc_key] == pytest.approx(c1)
    assert np.count_nonzero(V1) == 2


@pytest.mark.parametrize("kind", ("t0", "t1", "t2", "t4"))
def test_table_scores_equal_per_sv_scores_fixtures(kind):
    """Scoring through the table equals the per-SV scores (1e-10) on the LS-GKM
    fixture models, and so gkmpredict's printed scores."""
    from gkmsvm_amd import load_lsgkm_model

    m = load_lsgkm_model(H.path(f"lsgkm_wgkm_{kind}.model.txt.gz"), device="cpu")
    from gkmsvm_amd.inputs import row_lengths

    om = O.Model(O.Kern.make(m.kernel_type, m.kernel_params), m.support_sequences,
                 row_lengths(m.support_sequences, m.sv_lengths), m.coefficients, m.bias)
    recs = O.read_fasta(H.path("lsgkm_wgkm_test100.fa"))
    codes, lens = O.batch([s for _, s in recs])
    w = O.lmer_table(om)
    got = O.score_with_table(om, w, codes, lens)
    H.assert_close(got, O.score(om, codes, lens), 1e-10, kind)
    pred = H.read_pred(H.path(f"lsgkm_wgkm_{kind}_test100.pred.txt"))
    np.testing.assert_allclose(got, [pred[n] for n, _ in recs], rtol=2e-5, atol=1e-5)


def _explain_from_table(model, A, codes, lengths, perturbation):
    """explain/table.py:116-147 (_explain_table_cpu) restated: window u at
    digit k adds A[u, k, d] to base u_k ^ d; d = 0 goes to u_k (hypothetical)
    or the other three bases (perturbation); times 1 / sqrt(K_raw(x, x))."""
    kern = model.kern
    l = kern.l
    px = O.pack_both(kern, codes, lengths)
    dx = O.combine(kern, O.self_hist(kern, px))
    B, L = codes.shape
    out = np.zeros((B, 4, L))
    for b in range(B):
        for i, u in enumerate(px.fwd[b]):
            if u >= 4 ** l:
                continue
            a = 1.0 if px.w is None else float(px.w[b, i])
            for k in range(l):
                uk = (int(u) >> (2 * k)) & 3
                if perturbation:
                    for h in range(4):
                        if h != uk:
                            out[b, h, i + k] += a * A[u, k, 0]
                else:
                    out[b, uk, i + k] += a * A[u, k, 0]
                for d in range(1, 4):
                    out[b, uk ^ d, i + k] += a * A[u, k, d]
        out[b] /= np.sqrt(max(dx[b], 1e-20))
    return O._per_position(out, lengths, l)


@pytest.mark.parametrize("kt,pert,extra", [(2, False, None), (2, True, None), (0, False, None),
                                           (4, False, {"M": 9, "H": 4.0}), (4, True, {"M": 9, "H": 4.0})])
def test_explain_table_definition_equals_per_sv_explain(kt, pert, extra):
    """The explain table's definition, read per window, gives the per-SV
    GkmExplain loop's hypothetical and perturbation attributions."""
    m = _rand_model(kt, 5, 3, seed=7 + kt, S=9, hi=24, extra=extra)
    A = O.explain_table_brute(m, pert)
    rng = np.random.default_rng(3)
    codes, lens = _rand_queries(rng, 4, 5, 20)
    mode = "perturbation" if pert else "hypothetical"
    got = _explain_from_table(m, A, codes, lens, pert)
    H.assert_close(got, O.explain(m, codes, lens, mode), 1e-10, mode)


@pytest.mark.parametrize("kt,rc,extra", [(2, True, None), (2, False, None),
                                         (4, True, {"M": 5, "H": 2.0})])
def test_deltasvm_weights_score_each_lmer(kt, rc, extra):
    """With query_norm each DeltaSVM weight is the model's score of that l-mer
    alone minus the bias (svm.py:707-753); without it, the weight times the
    l-mer's norm."""
    l = 5
    m = _rand_model(kt, l, 3, seed=11, rc=rc, extra=extra)
    w = O.lmer_table_brute(m)
    rng = np.random.default_rng(5)
    idx = rng.integers(0, 4 ** l, 40)
    digits = ((idx[:, None] // 4 ** np.arange(l - 1, -1, -1)) % 4).astype(np.uint8)
    lens = np.full(len(idx), l)
    want = O.score(m, digits, lens) - m.bias
    got = O.deltasvm_weights(m, w, query_norm=True)[idx]
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-14)
    un = O.deltasvm_weights(m, w, query_norm=False)[idx]
    px = O.pack_both(m.kern, digits, lens)
    dx = O.combine(m.kern, O.self_hist(m.kern, px))
    np.testing.assert_allclose(un, want * np.sqrt(dx), rtol=1e-10, atol=1e-14)


def test_deltasvm_score_with_lmer_weights_is_unnormalised_table_sum():
    """A DeltaSVM with k = l, no RC, built from the table (no query_norm)
    scores sum_i wt0 w(x_i): the table sum before normalisation."""
    l = 5
    m = _rand_model(2, l, 3, seed=2)
    w = O.lmer_table_brute(m)
    weights = O.deltasvm_weights(m, w, query_norm=False)
    rng = np.random.default_rng(9)
    codes, lens = _rand_queries(rng, 5, l, 25)
    got = O.deltasvm_score(weights, l, l, codes, lens, include_rc=False, bias=m.bias)
    f = O.score_with_table(m, w, codes, lens)
    px = O.pack_both(m.kern, codes, lens)
    dx = O.combine(m.kern, O.self_hist(m.kern, px))
    np.testing.assert_allclose(got, (f - m.bias) * np.sqrt(dx) + m.bias, rtol=1e-10)


def test_deltasvm_score_gapped_kmers_by_hand():
    """l=3, k=2: a window ACG holds gapped 2-mers AC, AG, CG at combos (0,1),
    (0,2), (1,2); the RC strand CGT holds CG, CT, GT."""
    weights = np.zeros(16)
    for i in range(16):
        weights[i] = i + 1.0
    codes = np.array([[0, 1, 2]], np.uint8)
    fwd = (0 * 4 + 1 + 1) + (0 * 4 + 2 + 1) + (1 * 4 + 2 + 1)
    rc = (1 * 4 + 2 + 1) + (1 * 4 + 3 + 1) + (2 * 4 + 3 + 1)
    got = O.deltasvm_score(weights, 3, 2, codes, np.array([3]), include_rc=True, bias=0.5)
    assert got[0] == fwd + rc + 0.5
