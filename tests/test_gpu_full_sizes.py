"""Parity at the full sizes of the benchmark workloads, through properties that
hold at any size: an SV's histogram column depends on that SV alone (so a
sample of columns is checked against the oracle's brute-force counts), the two
GPU histogram paths (pairwise kernel, l-mer index) agree bit for bit over every
column, a row scores the same alone and in a batch, the Gram is exactly
symmetric with a unit diagonal, ISM equals the rescored mutants, and
GkmExplain through the index equals the per-pair sweep; the serving mix gives
each request's result alone, and GkmExplain on a slice of the 50,000-SV
model equals the oracle's."""

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P = {"L": 11, "k": 7, "d": 3}


def _model(kt, p, codes, lens, coef):
    from gkmsvm_amd import GkmSVM

    sv_l = None if (lens == codes.shape[1]).all() else lens
    m = GkmSVM(codes, coef, 0.0, kt, p, sv_lengths=sv_l)
    m.use_lmer_table = m.use_explain_table = False
    return m


def _strings(rng, lengths, n_frac=0.0):
    p = [(1 - n_frac) / 4] * 4 + [n_frac]
    return ["".join("ACGTN"[c] for c in rng.choice(5, int(n), p=p)) for n in lengths]


def _gc_codes(rng, n, L, gc):
    return rng.choice(4, size=(n, L), p=[(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]).astype(
        np.uint8)


def _histograms_both_paths(m, qs, gpu, n_cols, rng):
    """Index and pairwise histograms of the rows qs against every SV (equal bit
    for bit), and n_cols sampled SV columns against the oracle."""
    from gkmsvm_amd import engine as E
    from gkmsvm_amd.codec import encode_strings
    from gkmsvm_amd.inputs import row_lengths

    s = m._svs()
    qc, ql, bad = encode_strings(qs)
    assert not bad.any()
    xs = E.DeviceSeqs(qc, ql, gpu, rc=s.kp.rc)
    brute = E.cross_hist(s.kp, xs, 0, xs.n, s.seqs, s.groups).cpu().numpy()
    got = E.index_hist(s.kp, s.index(), xs, 0, xs.n, s.S).cpu().numpy()
    np.testing.assert_array_equal(got, brute)
    assert brute[..., 0].sum() > 0  # exact matches among the pairs
    cols = np.sort(rng.choice(s.S, n_cols, replace=False))
    column = np.argsort(s.order)  # SV (input order) -> histogram column
    svl = row_lengths(m.support_sequences, m.sv_lengths)
    kern = O.Kern.make(m.kernel_type, m.kernel_params)
    oc, ol = O.batch(qs)
    want = O.pair_hist(kern, O.pack_both(kern, oc, ol),
                       O.pack_both(kern, m.support_sequences[cols], svl[cols]))
    np.testing.assert_array_equal(got[:, column[cols], :], want)
    return s


def test_serving_workload_72k_svs(gpu):
    """72,000 SVs x 500 bp, -t 2 (2 SV tiles of 8-bit index counts): rows of
    50-2,000 bp give the same histograms through the index and the pairwise
    kernel, 48 sampled columns equal the oracle's, and every row scores the
    same alone, in the batch and through either path."""
    rng = np.random.default_rng(31)
    S = 72_000
    m = _model(2, P, rng.integers(0, 4, (S, 500)).astype(np.uint8), np.full(S, 500),
               rng.normal(size=S))
    qs = _strings(rng, [1000, 2000, 300, 50, 777])
    s = _histograms_both_paths(m, qs, gpu, 48, rng)
    idx = s.index()
    assert idx.n_tiles == 2 and idx.pack == 4
    m.use_lmer_index = True
    by_index = m(qs).cpu().numpy().ravel()
    m.use_lmer_index = False
    by_pairs = m(qs).cpu().numpy().ravel()
    np.testing.assert_array_equal(by_index, by_pairs)
    m.use_lmer_index = "auto"
    alone = np.array([m([q]).cpu().numpy().ravel()[0] for q in qs])
    np.testing.assert_array_equal(alone, by_index)


def test_gram_workload_10k(gpu):
    """The 10,000 x 10,000 -t 2 Gram of 500 bp sequences (GC 0.42; the first
    5,000 carry a planted 12-bp motif at 1-3 sites): exactly symmetric, unit
    diagonal, and 8 x 8 sampled entries equal the oracle's kernel values."""
    from gkmsvm_amd import compute_gram
    from gkmsvm_amd.kernels import build_kernel

    rng = np.random.default_rng(4)
    X = _gc_codes(rng, 10_000, 500, 0.42)
    pwm = rng.dirichlet(0.3 * np.ones(4), size=12)
    slots = np.argsort(rng.random((5000, 38)), axis=1)[:, :3] * 13  # distinct sites
    n_sites = rng.integers(1, 4, 5000)
    for k in range(3):
        rows = np.flatnonzero(n_sites > k)
        for c in range(12):
            X[rows, slots[rows, k] + c] = rng.choice(4, rows.size, p=pwm[c])
    kern = build_kernel(2, P)
    g = compute_gram(kern, X)
    assert g.shape == (10_000, 10_000) and g.dtype == np.float64
    assert (g == g.T).all()
    np.testing.assert_allclose(np.diag(g), 1.0, rtol=0, atol=1e-14)
    assert g.min() >= 0.0 and g.max() <= 1.0 + 1e-14
    ii = np.sort(np.concatenate([rng.choice(5000, 4, replace=False),
                                 5000 + rng.choice(5000, 4, replace=False)]))
    jj = np.sort(rng.choice(10_000, 8, replace=False))
    lens = np.full(8, 500)
    want = O.gram(O.Kern.make(2, P), X[ii], lens, X[jj], lens)
    H.assert_close(g[np.ix_(ii, jj)], want, 1e-10, "Gram entries")


def test_batched_uneven_workload_20k_svs(gpu):
    """-t 4 (weighted, 16-bit weighted index counts), 20,000 SVs of 100-1,000
    bp, 20,000 of the uneven 100-1,000 bp queries with 1% N (stored as A, as
    the kernels read N): the batch scorer's one persistent launch equals each
    row scored alone bit for bit (12 rows) and the oracle's l-mer index
    scores within 1e-10 (48 rows); the l-mer table path agrees within 1e-10;
    4 rows' histograms agree between the index and the pairwise kernel and
    with the oracle on 48 sampled SVs."""
    from gkmsvm_amd.batching import BatchScorer

    rng = np.random.default_rng(6)
    S = 20_000
    svl = rng.integers(100, 1001, S)
    svc = np.zeros((S, 1000), np.uint8)
    for i, n in enumerate(svl):
        svc[i, :n] = rng.integers(0, 4, n)
    coef = rng.normal(size=S)
    m = _model(4, {**P, "M": 50, "H": 50.0}, svc, svl, coef)
    nq = 20_000
    ql = rng.integers(100, 1001, nq)
    qc = rng.integers(0, 4, (nq, 1000)).astype(np.uint8)
    qc[rng.random(qc.shape) < 0.01] = 0  # N
    sc = BatchScorer(m, rows=256, bucket=128)
    got = sc(qc, lengths=ql).cpu().numpy().ravel()
    assert sc.queue_chunk is not None  # the persistent launch, chunk autotuned
    sample = rng.choice(nq, 12, replace=False)
    alone = np.array([m(qc[i:i + 1, :ql[i]]).cpu().numpy().ravel()[0] for i in sample])
    np.testing.assert_array_equal(got[sample], alone)
    rows48 = np.sort(rng.choice(nq, 48, replace=False))
    om = O.Model(O.Kern.make(4, {**P, "M": 50, "H": 50.0}), svc, svl, coef, 0.0)
    want = O.LmerIndex(om).score(qc[rows48], ql[rows48])
    H.assert_close(got[rows48], want, 1e-10, "persistent launch vs oracle index")
    m.use_lmer_table = True
    H.assert_close(m(qc[:1024], lengths=ql[:1024]).cpu().numpy().ravel(), got[:1024], 1e-10,
                   "table vs per-SV")
    m.use_lmer_table = False
    qs = ["".join("ACGT"[c] for c in qc[i, :ql[i]]) for i in sample[:4]]
    s = _histograms_both_paths(m, qs, gpu, 48, rng)
    idx = s.index()
    # 16-bit weighted counts; the 20,000 SVs fit one tile of the largest size
    # LDS holds (20,224) rather than two of the default 18,432
    assert idx.vals is not None and idx.pack == 2 and idx.n_tiles == 1 and idx.tile == 20_224


def test_explain_ism_workload_50k_svs(gpu):
    """-t 5 (weighted RBF, gamma 2), 50,000 SVs x 500 bp with Laplace(0, 0.5)
    coefficients, one 1 kb query: GkmExplain through the index equals the
    per-pair sweep (1e-10); ISM (through the l-mer index, as the cost model
    picks at this size) equals the sweep's ISM (1e-10), is zero at the
    reference base and equals 12 rescored mutants within 1e-12 of the
    coefficients' absolute sum (the rounding of two 50,000-term score sums)."""
    from gkmsvm_amd import engine as E
    from gkmsvm_amd import explain as X
    from gkmsvm_amd import gkmexplain, ism

    rng = np.random.default_rng(8)
    S = 50_000
    coef = rng.laplace(0, 0.5, S)
    m = _model(5, {**P, "gamma": 2.0, "M": 50, "H": 50.0}, _gc_codes(rng, S, 500, 0.5),
               np.full(S, 500), coef)
    q = _strings(rng, [1000])[0]
    res = {}
    for kernel in ("index", "sweep"):
        X.EXPLAIN_KERNEL = kernel
        try:
            res[kernel] = gkmexplain(m, [q], mode="hypothetical").cpu().numpy()
        finally:
            X.EXPLAIN_KERNEL = "auto"
    H.assert_close(res["index"], res["sweep"], 1e-10, "index vs sweep explain")
    assert np.abs(res["sweep"]).max() > 0
    assert E.ism_use_index(m._svs(), [len(q)], "auto")
    d = ism(m, [q]).cpu().numpy()[0]
    H.assert_close(d, ism(m, [q], use_index=False).cpu().numpy()[0], 1e-10,
                   "index vs sweep ISM")
    codes = O.encode(q)
    assert (d[codes, np.arange(len(q))] == 0.0).all()
    pos = rng.choice(len(q), 12, replace=False)
    alt = [(int(codes[p]) + int(rng.integers(1, 4))) % 4 for p in pos]
    mutants = [q[:p] + "ACGT"[a] + q[p + 1:] for p, a in zip(pos, alt)]
    f = m([q] + mutants).cpu().numpy().ravel()
    tol = 1e-12 * np.abs(coef).sum()
    np.testing.assert_allclose(d[alt, pos], f[1:] - f[0], rtol=0, atol=tol)
    assert np.abs(f[1:] - f[0]).max() > 1e3 * tol


def _serving_requests(rng, n):
    """The serving mix: 70% scoring (50-2,000 bp), 20% SNV pairs (300 bp),
    10% GkmExplain (1 kb)."""
    out = []
    for _ in range(n):
        u = rng.random()
        if u < 0.7:
            out.append(("score", _strings(rng, [int(rng.integers(50, 2001))])))
        elif u < 0.9:
            ref = _strings(rng, [300])[0]
            alt = ref[:150] + "ACGT"[("ACGT".index(ref[150]) + int(rng.integers(1, 4))) % 4] + \
                ref[151:]
            out.append(("variant", [(ref, alt)]))
        else:
            out.append(("explain", _strings(rng, [1000])))
    return out


@pytest.mark.parametrize("tables", [False, True])
def test_serving_mix_72k_svs_equal_alone(gpu, tables):
    """The serving configuration: 72,000 SVs x 500 bp (-t 2) resident, two
    bursts of 96 requests of the serving mix served through GkmServer on the
    per-SV kernels (index query and index explain) or through the l-mer
    tables: every result equals the same call made alone -- scores and
    variant effects bit for bit, attributions within 1e-10."""
    from gkmsvm_amd import gkmexplain
    from gkmsvm_amd.serving import GkmServer

    rng = np.random.default_rng(5)
    S = 72_000
    m = _model(2, P, rng.integers(0, 4, (S, 500)).astype(np.uint8), np.full(S, 500),
               rng.normal(size=S))
    if tables:
        m.use_lmer_table = m.use_explain_table = "auto"
    reqs = _serving_requests(rng, 192)
    with GkmServer(m, hbm_budget=16 << 30, tables=tables) as srv:
        assert (srv.table is not None) == tables
        futs = [srv.submit(kind, payload) for kind, payload in reqs[:96]]
        got = [f.result(timeout=120) for f in futs]
        futs = [srv.submit(kind, payload) for kind, payload in reqs[96:]]
        got += [f.result(timeout=120) for f in futs]
    n_ex = 0
    for (kind, payload), g in zip(reqs, got):
        if kind == "score":
            np.testing.assert_array_equal(g, m(payload).cpu().numpy().ravel())
        elif kind == "variant":
            want = m.score_variants([payload[0][0]], [payload[0][1]]).cpu().numpy().ravel()
            np.testing.assert_array_equal(g, want)
        else:
            n_ex += 1
            H.assert_close(g, gkmexplain(m, payload, mode="hypothetical").cpu().numpy(), 1e-10,
                           "served explain")
    assert n_ex > 0


def test_explain_50k_model_slice_vs_oracle(gpu):
    """GkmExplain of the -t 5 explain configuration (gamma 2, 500 bp SVs with
    Laplace(0, 0.5) coefficients, 1 kb queries) against the oracle's per-SV
    loop on a 2,000-SV slice of the 50,000-SV model -- attributions are sums
    over the SVs, so the slice checks the same arithmetic: through the l-mer
    index and through the per-pair sweep, within 1e-10."""
    from gkmsvm_amd import explain as X
    from gkmsvm_amd import gkmexplain

    rng = np.random.default_rng(8)
    S = 50_000
    codes = _gc_codes(rng, S, 500, 0.5)
    coef = rng.laplace(0, 0.5, S)
    sl = np.sort(rng.choice(S, 2000, replace=False))
    p5 = {**P, "gamma": 2.0, "M": 50, "H": 50.0}
    m = _model(5, p5, codes[sl], np.full(2000, 500), coef[sl])
    qs = _strings(rng, [1000, 1000])
    qc, ql = O.batch(qs)
    om = O.Model(O.Kern.make(5, p5), codes[sl], np.full(2000, 500), coef[sl], 0.0)
    want = O.explain(om, qc, ql, "hypothetical")
    assert np.abs(want).max() > 0
    for kernel in ("index", "sweep"):
        X.EXPLAIN_KERNEL = kernel
        try:
            got = gkmexplain(m, qs, mode="hypothetical").cpu().numpy()
        finally:
            X.EXPLAIN_KERNEL = "auto"
        H.assert_close(got, want, 1e-10, f"{kernel} explain vs oracle")
