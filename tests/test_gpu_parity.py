"""GPU parity: the HIP engine (through its C-ABI, via the package) against the
CPU oracle on the same inputs -- histograms bit-exact, FP64 values within
1e-10 relative (helpers.assert_close) -- and against the LS-GKM golden files
at their printed precision (rtol 2e-5)."""

import numpy as np
import pytest
import torch

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rand_codes(rng, n, lo, hi, binary_frac=0.3):
    lens = rng.integers(lo, hi + 1, n)
    codes = np.zeros((n, int(lens.max())), np.uint8)
    for i in range(n):
        alphabet = 2 if rng.random() < binary_frac else 4  # low-complexity rows: many hits
        codes[i, : lens[i]] = rng.integers(0, alphabet, lens[i])
    return codes, lens.astype(np.int64)


def _params(kt, l, k, rng):
    p = {"L": l, "k": k, "include_rc": bool(rng.integers(0, 2))}
    if kt in (4, 5):
        p["M"], p["H"] = int(rng.integers(1, 60)), float(rng.uniform(5, 60))
    if kt in (3, 5, 6):
        p["gamma"] = float(rng.uniform(0.5, 3))
    if kt == 0 and rng.integers(0, 2):
        p["d"] = int(rng.integers(1, 4))
    return p


def _per_sv(m):
    """The model with its exact l-mer tables off: these tests hold the per-SV
    kernels (the table paths have tests of their own)."""
    m.use_lmer_table = False
    m.use_explain_table = False
    return m


def _load(name):
    from gkmsvm_amd import load_lsgkm_model

    return _per_sv(load_lsgkm_model(H.path(name)))


def _model(kt, p, codes, lens, coef, bias=0.0):
    from gkmsvm_amd import GkmSVM

    sv_l = None if (lens == codes.shape[1]).all() else lens
    return _per_sv(GkmSVM(codes, coef, bias, kt, p, sv_lengths=sv_l))


def _oracle_model(m):
    from gkmsvm_amd.inputs import row_lengths

    kern = O.Kern.make(m.kernel_type, m.kernel_params)
    return O.Model(kern, m.support_sequences, row_lengths(m.support_sequences, m.sv_lengths),
                   m.coefficients, m.bias)


@pytest.mark.parametrize("seed", range(12))
def test_cross_and_self_histograms_bit_exact(gpu, seed):
    """N[m] of every (query, SV) pair and every self pair equals the oracle's
    brute-force count (weighted: the sum of wx * wy), for l from 3 to 13,
    ragged lengths (some shorter than l), RC on/off, split launch shapes."""
    from gkmsvm_amd import engine as E
    from gkmsvm_amd.kernels import build_kernel

    rng = np.random.default_rng(seed)
    l = int(rng.choice([3, 6, 9, 10, 11, 12, 13]))
    k = max(1, l - int(rng.integers(2, 5)))
    kt = int(rng.choice([0, 1, 2, 4]))
    p = _params(kt, l, k, rng)
    kern = build_kernel(kt, p)
    okern = O.Kern.make(kt, p)
    print(f"[case] seed={seed} l={l} k={k} kt={kt} p={p}", flush=True)
    xc, xl = _rand_codes(rng, int(rng.integers(1, 9)), 1, 300)
    yc, yl = _rand_codes(rng, int(rng.integers(1, 200)), 1, 400)
    if xc.shape[1] < l or yc.shape[1] < l:
        pytest.skip("no windows")
    kp = E.KernelParams(kern, gpu)
    xs = E.DeviceSeqs(xc, xl, gpu, rc=kp.rc)
    ys = E.DeviceSeqs(yc, yl, gpu, rc=kp.rc)
    groups = torch.from_numpy(E.group_bounds(yl)).to(gpu)
    want = O.pair_hist(okern, O.pack_both(okern, xc, xl), O.pack_both(okern, yc, yl))
    for shape in (None, (1, 1), (3, 5)):
        got = E.cross_hist(kp, xs, 0, xs.n, ys, groups, shape=shape)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy().view(np.uint64), want,
                                      err_msg=f"shape={shape}")
    # self-kernels
    hist = torch.zeros((ys.n, kp.nm), dtype=torch.int64, device=gpu)
    wt, wn = kp.wtab(ys.lmax)
    from gkmsvm_amd import _lib

    _lib.call("gkm_self_hist", E._p(ys.planes), E._p(ys.planes_rc) if kp.rc else None,
              E._p(ys.len), ys.n, ys.nw, kp.l, kp.nm, E._p(wt), wn, 3, E._p(hist), None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hist.cpu().numpy().view(np.uint64),
                                  O.self_hist(okern, O.pack_both(okern, yc, yl)))


@pytest.mark.parametrize("d", [4, 5, 6, 7])
def test_histograms_and_ism_5_to_8_levels(gpu, d):
    """-t 0 with d = 4..7 (nm = 5..8 levels: the 5- and 8-level kernel
    instances): pairwise and self histograms equal the oracle's counts, and
    ISM (nm + 1 levels) equals rescoring every mutant."""
    from gkmsvm_amd import engine as E
    from gkmsvm_amd import ism
    from gkmsvm_amd.kernels import build_kernel

    rng = np.random.default_rng(60 + d)
    p = {"L": 12, "k": 12 - d, "d": d}
    kern = build_kernel(0, p)
    okern = O.Kern.make(0, p)
    xc, xl = _rand_codes(rng, 4, 30, 120)
    yc, yl = _rand_codes(rng, 70, 20, 160)
    kp = E.KernelParams(kern, gpu)
    assert kp.nm == d + 1
    xs = E.DeviceSeqs(xc, xl, gpu, rc=kp.rc)
    ys = E.DeviceSeqs(yc, yl, gpu, rc=kp.rc)
    groups = torch.from_numpy(E.group_bounds(yl)).to(gpu)
    got = E.cross_hist(kp, xs, 0, xs.n, ys, groups).cpu().numpy().view(np.uint64)
    np.testing.assert_array_equal(
        got, O.pair_hist(okern, O.pack_both(okern, xc, xl), O.pack_both(okern, yc, yl)))
    from gkmsvm_amd import _lib

    hist = torch.zeros((ys.n, kp.nm), dtype=torch.int64, device=gpu)
    _lib.call("gkm_self_hist", E._p(ys.planes), E._p(ys.planes_rc), E._p(ys.len), ys.n, ys.nw,
              kp.l, kp.nm, None, 0, 3, E._p(hist), None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hist.cpu().numpy().view(np.uint64),
                                  O.self_hist(okern, O.pack_both(okern, yc, yl)))
    m = _model(0, p, yc, yl, rng.normal(size=len(yl)), 0.1)
    qc, ql = _rand_codes(rng, 2, 40, 60)
    got = ism(m, qc, lengths=ql).cpu().numpy()
    H.assert_close(got, O.ism(_oracle_model(m), qc, ql), 1e-9, f"ISM d={d}")


@pytest.mark.parametrize("seed", range(4))
def test_upper_triangle_columns(gpu, seed):
    """col_min restricts row b to columns >= b + row_base; the rest stay 0."""
    from gkmsvm_amd import engine as E
    from gkmsvm_amd.kernels import build_kernel

    rng = np.random.default_rng(100 + seed)
    kern = build_kernel(2, {"L": 11, "k": 7})
    okern = O.Kern.make(2, {"L": 11, "k": 7})
    yc, yl = _rand_codes(rng, 150, 20, 120)
    kp = E.KernelParams(kern, gpu)
    ys = E.DeviceSeqs(yc, yl, gpu)
    groups = torch.from_numpy(E.group_bounds(yl)).to(gpu)
    want = O.pair_hist(okern, O.pack_both(okern, yc, yl), O.pack_both(okern, yc, yl))
    r0, r1 = 37, 121
    got = E.cross_hist(kp, ys, r0, r1, ys, groups, col_min=0).cpu().numpy().view(np.uint64)
    for b in range(r1 - r0):
        np.testing.assert_array_equal(got[b, r0 + b:], want[r0 + b, r0 + b:])
        assert (got[b, : r0 + b] == 0).all()


@pytest.mark.parametrize("kind", H.WGKM_TYPES)
@pytest.mark.parametrize("n", (100, 85))
def test_scores_lsgkm_fixtures(gpu, kind, n):
    """-t 0..6 fixture models: scores equal the oracle to 1e-10 and gkmpredict
    to its printed digits."""
    m = _load(f"lsgkm_wgkm_{kind}.model.txt.gz")
    recs = O.read_fasta(H.path(f"lsgkm_wgkm_test{n}.fa"))
    seqs = [s for _, s in recs]
    got = m(seqs).cpu().numpy().ravel()
    om = _oracle_model(m)
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.score(om, codes, lens), 1e-10, f"{kind}_{n} vs oracle")
    pred = H.read_pred(H.path(f"lsgkm_wgkm_{kind}_test{n}.pred.txt"))
    want = np.array([pred[name] for name, _ in recs])
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("kind", ("t2", "t4", "t5"))
def test_ragged_models_gkmpredict(gpu, kind):
    """Models trained on 60-140 bp sequences (ragged SVs) on 30-200 bp queries."""
    m = _load(f"lsgkm_ragged_{kind}.model.txt.gz")
    recs = O.read_fasta(H.path("lsgkm_ragged_test.fa"))
    seqs = [s for _, s in recs]
    got = m(seqs).cpu().numpy().ravel()
    om = _oracle_model(m)
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.score(om, codes, lens), 1e-10, "vs oracle")
    pred = H.read_pred(H.path(f"lsgkm_ragged_{kind}_test.pred.txt"))
    want = np.array([pred[name] for name, _ in recs])
    ok = ~np.isnan(got)
    np.testing.assert_allclose(got[ok], want[ok], rtol=2e-5, atol=1e-5)


def test_encode_10sv_gkmpredict(gpu):
    import json
    import random

    m = _load("encode_ENCFF579AOX_10sv.model.txt.gz")
    gold = json.load(open(H.path("encode_10sv_gkmpredict.json")))["scores"]
    random.seed(42)
    rand = "".join(random.choice("ACGT") for _ in range(300))
    from gkmsvm_amd.codec import decode_codes

    qs = [decode_codes(m.support_sequences[i, : (m.sv_lengths[i] if m.sv_lengths is not None
                                                    else m.support_sequences.shape[1])])
          for i in range(3)] + [rand]
    got = m(qs).cpu().numpy().ravel()
    want = np.array([gold[k] for k in ("query_0", "query_1", "query_2", "random")])
    np.testing.assert_allclose(got, want, atol=1e-4)


@pytest.mark.parametrize("name", sorted(H.EXPLAIN_SETS))
@pytest.mark.parametrize("mode,lm", [("importance", 0), ("hypothetical", 1), ("perturbation", 4)])
def test_gkmexplain_fixtures(gpu, name, mode, lm):
    """GkmExplain equals the oracle's per-SV loop to 1e-10 and the kundajelab
    lsgkm gkmexplain output at its printed precision."""
    from gkmsvm_amd import gkmexplain

    mf, fa, pre = H.EXPLAIN_SETS[name]
    m = _load(mf)
    seqs = [s for _, s in O.read_fasta(H.path(fa))]
    got = gkmexplain(m, seqs, mode=mode, batch_size=8).cpu().numpy()
    om = _oracle_model(m)
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.explain(om, codes, lens, mode), 1e-10, f"{name} {mode}")
    exp = H.read_explain(H.path(f"{pre}.explain{lm}.txt.gz"))
    scale = max(np.abs(e).max() for e in exp)
    for i, e in enumerate(exp):
        np.testing.assert_allclose(got[i, :, : e.shape[1]], e, rtol=2e-5, atol=5e-6 * scale)


@pytest.mark.parametrize("seed", range(6))
def test_explain_kernels_agree_with_oracle(gpu, seed):
    """The explain kernels (the SVs' l-mer weights, kernels without RBF; the
    l-mer index; lanes = diagonals of one pair) against the oracle's per-SV
    loop on random models: l from 3 to 13 (the l-mer paths to 12), all
    non-RBF and RBF types, weights, RC on/off, ragged SVs and queries (some
    shorter than l), long queries."""
    from gkmsvm_amd import explain as X
    from gkmsvm_amd import gkmexplain

    rng = np.random.default_rng(100 + seed)
    l = int(rng.choice([3, 6, 10, 11, 12, 13]))
    kt = int(rng.integers(0, 7))
    k = int(rng.integers(max(1, l - 5), l))
    p = _params(kt, l, k, rng)
    if kt != 0:
        p["d"] = int(rng.integers(1, min(4, l)))
    yc, yl = _rand_codes(rng, int(rng.integers(70, 200)), 1, 260)
    m = _model(kt, p, yc, yl, rng.normal(size=len(yl)), 0.3)
    xc, xl = _rand_codes(rng, 5, max(1, l - 2), 700)
    om = _oracle_model(m)
    for mode in ("hypothetical", "perturbation"):
        want = O.explain(om, xc, xl, mode)
        for kernel in ("lmer", "index", "sweep"):
            if kernel in ("lmer", "index") and l > 12:
                continue
            if kernel == "lmer" and m._params().rbf:
                continue
            X.EXPLAIN_KERNEL = kernel
            try:
                got = gkmexplain(m, xc, mode=mode, lengths=xl).cpu().numpy()
            finally:
                X.EXPLAIN_KERNEL = "auto"
            H.assert_close(got, want, 1e-10, f"{kernel} {mode} l={l} t={kt}")


@pytest.mark.parametrize("kt", [2, 3, 5])
def test_index_explain_with_repeats(gpu, kt):
    """GkmExplain through the index (and, without RBF, the SVs' l-mer
    weights) on SVs with poly-A and dinucleotide runs (l-mer buckets of
    thousands of entries: saturated tile counts in the l-mer records, long
    runs past the record; weighted: the per-key records) equals the oracle's
    per-SV loop, hypothetical and perturbation."""
    from gkmsvm_amd import explain as X
    from gkmsvm_amd import gkmexplain

    rng = np.random.default_rng(41 + kt)
    L = 300
    yc = rng.integers(0, 4, (150, L)).astype(np.uint8)
    yc[:40] = 0
    yc[40:48] = np.tile([0, 2], L // 2)
    p = {"L": 11, "k": 7, "d": 3}
    if kt == 5:
        p.update(M=50, H=50.0, gamma=1.5)
    if kt == 3:
        p["gamma"] = 1.5
    m = _model(kt, p, yc, np.full(len(yc), L), rng.normal(size=len(yc)), 0.1)
    idx = m._svs().index()
    assert int((idx.offsets[1:] - idx.offsets[:-1]).max()) > 255
    xc = rng.integers(0, 4, (3, 400)).astype(np.uint8)
    xc[0, 50:200] = 0
    xc[1, :] = np.tile([0, 2], 200)
    xl = np.array([400, 400, 350])
    om = _oracle_model(m)
    kernels = ("index",) if m._params().rbf else ("index", "lmer")
    for mode in ("hypothetical", "perturbation"):
        want = O.explain(om, xc, xl, mode)
        for kernel in kernels:
            X.EXPLAIN_KERNEL = kernel
            try:
                got = gkmexplain(m, xc, mode=mode, lengths=xl).cpu().numpy()
            finally:
                X.EXPLAIN_KERNEL = "auto"
            H.assert_close(got, want, 1e-10, f"{kernel} {mode} t={kt}")


@pytest.mark.parametrize("name", sorted(k for k in H.EXPLAIN_SETS if not k.endswith("_51")))
def test_mutation_impact_fixtures(gpu, name):
    """score_variants(method='gkmexplain') equals the oracle and gkmexplain -m 5."""
    mf, fa, pre = H.EXPLAIN_SETS[name]
    m = _load(mf)
    seqs = [s for _, s in O.read_fasta(H.path(fa))]
    ref, alt = H.mode5_pairs(seqs, fa)
    got = m.score_variants(ref, alt, method="gkmexplain", batch_size=16).cpu().numpy().ravel()
    om = _oracle_model(m)
    rc, rl = O.batch(ref)
    ac, al = O.batch(alt)
    H.assert_close(got, O.impact(om, rc, rl, ac, al), 1e-10, name)
    exp = np.stack([e[:, 0] for e in H.read_explain(H.path(f"{pre}.explain5.txt.gz"))])
    np.testing.assert_allclose(got.reshape(-1, 4), exp, rtol=2e-5, atol=5e-6 * np.abs(exp).max())


@pytest.mark.parametrize("mf", ["lsgkm_wgkm_t0.model.txt.gz", "lsgkm_wgkm_t1.model.txt.gz",
                                "lsgkm_mode5_est.model.txt.gz", "lsgkm_mode5_rbf.model.txt.gz",
                                "lsgkm_wgkm_t4.model.txt.gz", "lsgkm_wgkm_t5.model.txt.gz",
                                "lsgkm_wgkm_t6.model.txt.gz", "lsgkm_ragged_t5.model.txt.gz"])
def test_ism_matches_rescoring(gpu, mf):
    """ISM equals f(mutant) - f(x) with every mutant scored from scratch by the
    oracle (ragged queries, N bases, lengths with odd and even window counts)."""
    from gkmsvm_amd import ism

    m = _load(mf)
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_ragged_query.fa"))][:5]
    got = ism(m, seqs).cpu().numpy()
    om = _oracle_model(m)
    codes, lens = O.batch(seqs)
    want = O.ism(om, codes, lens)
    H.assert_close(got, want, 1e-9, mf)
    # the reference base changes nothing
    for b, s in enumerate(seqs):
        for pos in range(len(s)):
            assert got[b, codes[b, pos], pos] == 0.0


def test_score_variants_kernel(gpu):
    m = _load("lsgkm_mode5_rbf.model.txt.gz")
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_mode5_test51.fa"))]
    ref, alt = H.mode5_pairs(seqs, "x")
    got = m.score_variants(ref, alt).cpu().numpy().ravel()
    om = _oracle_model(m)
    rc, rl = O.batch(ref)
    ac, al = O.batch(alt)
    H.assert_close(got, O.score(om, ac, al) - O.score(om, rc, rl), 1e-9, "variants")


@pytest.mark.parametrize("kt", [2, 3, 4, 0])
def test_gram_matches_oracle(gpu, kt):
    """Symmetric Gram (upper triangle computed, mirrored, exactly symmetric) and
    a rectangular one, ragged lengths, through the pairwise kernel and the
    l-mer index, against the oracle; assembled in HBM or on the host, the same
    bits."""
    from gkmsvm_amd import compute_gram
    from gkmsvm_amd.kernels import build_kernel

    rng = np.random.default_rng(kt)
    p = {"L": 10, "k": 6, "gamma": 1.5} if kt == 3 else {"L": 10, "k": 6}
    kern = build_kernel(kt, p)
    okern = O.Kern.make(kt, p)
    xc, xl = _rand_codes(rng, 140, 40, 160)
    yc, yl = _rand_codes(rng, 70, 30, 90)
    want = O.gram(okern, xc, xl)
    want2 = O.gram(okern, xc, xl, yc, yl)
    from gkmsvm_amd import gram as G

    for use_index in (False, True):
        g = compute_gram(kern, xc, lengths=xl, use_index=use_index)
        assert (g == g.T).all()
        H.assert_close(g, want, 1e-10, f"symmetric, index={use_index}")
        g2 = compute_gram(kern, xc, yc, lengths=xl, y_lengths=yl, use_index=use_index)
        H.assert_close(g2, want2, 1e-10, f"rectangular, index={use_index}")
        # the Gram assembled in HBM (mirror, input order) equals the host assembly
        cap = G.GRAM_DEVICE_BYTES
        G.GRAM_DEVICE_BYTES = 0
        try:
            np.testing.assert_array_equal(compute_gram(kern, xc, lengths=xl, use_index=use_index), g)
            np.testing.assert_array_equal(
                compute_gram(kern, xc, yc, lengths=xl, y_lengths=yl, use_index=use_index), g2)
        finally:
            G.GRAM_DEVICE_BYTES = cap


def test_plain_scoring_config_against_index_oracle(gpu):
    """The plain -t 2 l=11 k=7 d=3 workload (500 SVs x 300 bp, 1,000 queries x
    300 bp, sign-balanced N(0,1) coefficients) equals the oracle's l-mer index
    path (the reference's CPU algorithm) to 1e-10."""
    rng = np.random.default_rng(1)
    svc = rng.integers(0, 4, (500, 300)).astype(np.uint8)
    coef = np.abs(np.random.default_rng(2).normal(size=500))
    coef[250:] *= -1
    q = np.random.default_rng(3).integers(0, 4, (1000, 300)).astype(np.uint8)
    p = {"L": 11, "k": 7, "d": 3}
    m = _model(2, p, svc, np.full(500, 300), coef)
    got = m(q).cpu().numpy().ravel()
    om = O.Model(O.Kern.make(2, p), svc, np.full(500, 300), coef, 0.0)
    want = O.LmerIndex(om).score(q, np.full(1000, 300))
    H.assert_close(got, want, 1e-10, "C1")
    # and the brute-force oracle on a slice
    H.assert_close(got[:50], O.score(om, q[:50], np.full(50, 300)), 1e-10, "C1 brute")


def test_short_and_n_rows(gpu):
    """Rows shorter than l score NaN; N bases score as A (RC: T)."""
    m = _load("lsgkm_mode5_est.model.txt.gz")
    seqs = ["ACGT", "ACGTNNNNNACGTACGGATCGATTTACGA" * 3, "NNNNNNNNNNNNNNNNNNNN"]
    got = m(seqs).cpu().numpy().ravel()
    assert np.isnan(got[0]) and np.isfinite(got[1:]).all()
    om = _oracle_model(m)
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.score(om, codes, lens), 1e-10, "N rows")


@pytest.mark.parametrize("seed", range(8))
def test_index_histograms_equal_brute_force(gpu, seed):
    """The l-mer index path gives the brute-force kernel's integers exactly
    (random l, kernel types incl. -t 0 without d and weighted, ragged SVs,
    RC on/off), and model scores agree bit for bit between the two paths."""
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(300 + seed)
    l = int(rng.choice([5, 8, 10, 11, 12]))
    k = max(1, l - int(rng.integers(2, 5)))
    kt = int(rng.choice([0, 2, 4, 5]))
    p = _params(kt, l, k, rng)
    print(f"[case] seed={seed} l={l} k={k} kt={kt} p={p}", flush=True)
    yc, yl = _rand_codes(rng, int(rng.integers(1, 300)), l, 400)
    m = _model(kt, p, yc, yl, rng.normal(size=len(yl)), 0.1)
    sv = m._svs()
    xc, xl = _rand_codes(rng, int(rng.integers(1, 9)), 1, 300)
    if xc.shape[1] < l:
        pytest.skip("no windows")
    xs = E.DeviceSeqs(xc, xl, gpu, rc=sv.kp.rc)
    a = E.index_hist(sv.kp, sv.index(), xs, 0, xs.n, sv.S)
    b = E.cross_hist(sv.kp, xs, 0, xs.n, sv.seqs, sv.groups)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    m.use_lmer_index = True
    s_idx = m(xc, lengths=xl).cpu().numpy()
    m.use_lmer_index = False
    s_bru = m(xc, lengths=xl).cpu().numpy()
    np.testing.assert_array_equal(s_idx, s_bru)


def test_index_mixed_bucket_sizes(gpu):
    """Buckets of a few entries (walked by their lanes) and of thousands (poly-A
    and dinucleotide repeats, walked wave-balanced) in the same launch give the
    brute-force integers, weighted (16-bit counts with the 32-bit fallback, and
    32-bit counts) and unweighted."""
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(17)
    L = 400
    sv = rng.integers(0, 4, (300, L)).astype(np.uint8)
    sv[:180] = 0                     # poly-A: a bucket of 70,200 entries (past a record's count)
    sv[180:190] = np.tile([0, 2], L // 2)  # AGAG...
    q = rng.integers(0, 4, (6, L)).astype(np.uint8)
    q[0, 100:250] = 0
    q[1, :] = np.tile([0, 2], L // 2)
    q[2, 50:120] = 0
    for kt, p in ((2, {"L": 11, "k": 7, "d": 3}), (4, {"L": 10, "k": 6, "d": 3, "M": 50, "H": 50})):
        m = _model(kt, p, sv, np.full(len(sv), L), rng.normal(size=len(sv)), 0.0)
        s = m._svs()
        xs = E.DeviceSeqs(q, np.full(len(q), L), gpu, rc=s.kp.rc)
        idx = s.index()
        assert idx.n_entries / (4 ** s.kp.l) < 1 and int(
            (idx.offsets[1:] - idx.offsets[:-1]).max()) > 65535  # both kinds of bucket
        a = E.index_hist(s.kp, idx, xs, 0, xs.n, s.S)
        b = E.cross_hist(s.kp, xs, 0, xs.n, s.seqs, s.groups)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
        if kt == 4:  # weighted: 16-bit counts (poly-A pairs carry: the 32-bit walk) and 32-bit
            assert idx.pack == 2
            w32 = E.DeviceIndex(s.kp, s.seqs, count_bits=32)
            assert w32.pack == 1
            c = E.index_hist(s.kp, w32, xs, 0, xs.n, s.S)
            np.testing.assert_array_equal(c.cpu().numpy(), b.cpu().numpy())


def test_index_record_bucket_sizes(gpu):
    """Buckets on either side of the bucket records' limits give the brute-force
    integers: unweighted records hold up to 15 ids, a longer bucket 11 of them
    with its start and count (the rest from the ids array: up to 12 taken by
    the lanes, more walked balanced); weighted records hold 8 (id, weight)
    pairs.  A motif copied into n SVs makes a bucket of n entries."""
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(41)
    l, L = 11, 260
    sizes = [1, 8, 9, 14, 15, 16, 17, 20, 21, 22, 23, 24, 25, 40]
    S = 96
    sv = rng.integers(0, 4, (S, L)).astype(np.uint8)
    q = rng.integers(0, 4, (3, L)).astype(np.uint8)
    for j, n in enumerate(sizes):
        motif = rng.integers(0, 4, l).astype(np.uint8)
        pos = 20 + 15 * j  # motifs 15 apart: no two overlap
        sv[rng.choice(S, n, replace=False), pos:pos + l] = motif
        q[j % 3, pos:pos + l] = motif
    for kt, p in ((2, {"L": l, "k": 7, "d": 3}), (4, {"L": l, "k": 7, "d": 3, "M": 50, "H": 50})):
        m = _model(kt, p, sv, np.full(S, L), rng.normal(size=S), 0.0)
        s = m._svs()
        idx = s.index()
        cnt = (idx.offsets[1:].long() - idx.offsets[:-1].long()).cpu().numpy()
        assert all((cnt >= n).any() for n in sizes)
        xs = E.DeviceSeqs(q, np.full(len(q), L), gpu, rc=s.kp.rc)
        a = E.index_hist(s.kp, idx, xs, 0, xs.n, s.S)
        b = E.cross_hist(s.kp, xs, 0, xs.n, s.seqs, s.groups)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5])
def test_index_count_widths(gpu, d):
    """8-bit LDS counts (with the exact 16-bit fallback), 16-bit and 32-bit
    counts all give the brute-force integers: random rows (no count reaches
    256), poly-A and dinucleotide repeats (counts past 255: the fallback, in
    two passes over a 512-SV tile), d + 1 = 2..6 levels, several tiles, and the
    symmetric Gram's upper triangle (col_min)."""
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(40 + d)
    L = 300
    sv = rng.integers(0, 4, (700, L)).astype(np.uint8)
    sv[:40] = 0
    sv[600:610] = np.tile([0, 2], L // 2)
    q = rng.integers(0, 4, (5, L)).astype(np.uint8)
    q[0, 30:200] = 0
    q[1] = np.tile([0, 2], L // 2)
    m = _model(2, {"L": 10, "k": 5, "d": d}, sv, np.full(len(sv), L), rng.normal(size=len(sv)))
    s = m._svs()
    assert s.kp.nm == d + 1
    xs = E.DeviceSeqs(q, np.full(len(q), L), gpu, rc=s.kp.rc)
    want = E.cross_hist(s.kp, xs, 0, xs.n, s.seqs, s.groups).cpu().numpy()
    assert want[:2, :40, 0].max() > 255  # repeats: some counts need the fallback
    g_want = E.cross_hist(s.kp, s.seqs, 0, s.S, s.seqs, s.groups, col_min=0).cpu().numpy()
    upper = np.triu(np.ones((s.S, s.S), bool))
    for bits in (8, 16, 32):
        idx = E.DeviceIndex(s.kp, s.seqs, tile=512, count_bits=bits)
        assert idx.pack == {8: 4 if d >= 2 else 2, 16: 2, 32: 1}[bits]
        print(f"[case] d={d} count_bits={bits} pack={idx.pack} tiles={idx.n_tiles}", flush=True)
        got = E.index_hist(s.kp, idx, xs, 0, xs.n, s.S).cpu().numpy()
        np.testing.assert_array_equal(got, want)
        g = E.index_hist(s.kp, idx, s.seqs, 0, s.S, s.S, col_min=0).cpu().numpy()
        np.testing.assert_array_equal(g[upper], g_want[upper])


@pytest.mark.parametrize("S,tiles,tile", [(42000, 2, 36864), (40000, 1, 40192)])
def test_index_two_full_tiles(gpu, S, tiles, tile):
    """8-bit counts over 2 full SV tiles (42,000 SVs: 36,864 + 5,136) and over
    the largest tile LDS holds (40,000 SVs in one tile of 40,192) give the
    brute-force integers -- random SVs, poly-A SVs against a poly-A query
    stretch (the 16-bit fallback, in two passes over the full tile), rows of
    several lengths."""
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(23)
    L = 60
    sv = rng.integers(0, 4, (S, L)).astype(np.uint8)
    sv[:50] = 0
    m = _model(2, {"L": 11, "k": 7, "d": 3}, sv, np.full(S, L), rng.normal(size=S))
    s = m._svs()
    qc, ql = _rand_codes(rng, 5, 40, 700)
    qc[0, :300] = 0
    ql[0] = max(ql[0], 300)
    xs = E.DeviceSeqs(qc, ql, gpu, rc=s.kp.rc)
    want = E.cross_hist(s.kp, xs, 0, xs.n, s.seqs, s.groups).cpu().numpy()
    assert want[0, :, 0].max() > 255
    idx = s.index()
    assert idx.n_tiles == tiles and idx.pack == 4 and idx.tile == tile
    got = E.index_hist(s.kp, idx, xs, 0, xs.n, S).cpu().numpy()
    np.testing.assert_array_equal(got, want)


def test_index_plain_config(gpu):
    """Plain -t 2 scoring through the l-mer index equals the oracle's index path."""
    rng = np.random.default_rng(1)
    svc = rng.integers(0, 4, (500, 300)).astype(np.uint8)
    coef = np.abs(np.random.default_rng(2).normal(size=500))
    coef[250:] *= -1
    q = np.random.default_rng(3).integers(0, 4, (200, 300)).astype(np.uint8)
    p = {"L": 11, "k": 7, "d": 3}
    m = _model(2, p, svc, np.full(500, 300), coef)
    m.use_lmer_index = True
    got = m(q).cpu().numpy().ravel()
    om = O.Model(O.Kern.make(2, p), svc, np.full(500, 300), coef, 0.0)
    H.assert_close(got, O.LmerIndex(om).score(q, np.full(200, 300)), 1e-10, "index C1")


def test_peak_microbenchmarks_run(gpu):
    """The measurement-only peak kernels run and leave their sink untouched
    (their sentinels never come up on these inputs)."""
    from gkmsvm_amd import _lib
    from gkmsvm_amd import engine as E

    src = torch.ones(1 << 20, dtype=torch.int64, device=gpu)
    sink = torch.zeros(4, dtype=torch.int32, device=gpu)
    _lib.call("gkm_peak_read", src.data_ptr(), src.numel() * 8, sink.data_ptr(), 64,
              E.stream_handle())
    _lib.call("gkm_peak_valu", 16, sink.data_ptr(), 64, E.stream_handle())
    torch.cuda.synchronize()
    assert int(sink.abs().sum()) == 0
    with pytest.raises(_lib.GkmError):
        _lib.call("gkm_peak_read", src.data_ptr(), 15, sink.data_ptr(), 64, E.stream_handle())


def test_index_without_entries(gpu):
    """An index whose SVs are all shorter than l holds no entries (the query's
    loads then read its padding entries): the histograms are zero, equal to
    the pairwise kernel's."""
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(29)
    sv = rng.integers(0, 4, (40, 9)).astype(np.uint8)
    m = _model(2, {"L": 11, "k": 7, "d": 3}, sv, np.full(40, 9), rng.normal(size=40))
    s = m._svs()
    idx = s.index()
    assert idx.n_entries == 0 and idx.ids.numel() == E.INDEX_PAD
    qc, ql = _rand_codes(rng, 3, 20, 300)
    xs = E.DeviceSeqs(qc, ql, gpu, rc=s.kp.rc)
    got = E.index_hist(s.kp, idx, xs, 0, xs.n, s.S).cpu().numpy()
    want = E.cross_hist(s.kp, xs, 0, xs.n, s.seqs, s.groups).cpu().numpy()
    np.testing.assert_array_equal(got, want)
    assert not got.any()
