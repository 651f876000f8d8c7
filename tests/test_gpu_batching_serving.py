"""Batched scoring (length buckets, autotuned launch plans, HIP-graph replays)
and the continuous-batching server: every result equals the plain call's
(scores bit for bit; GkmExplain, whose FP64 atomics sum in any order, within
1e-10), alone or in any mix."""

import threading

import numpy as np
import pytest
import torch

import helpers as H

pytestmark = pytest.mark.gpu


def _random_model(S=700, lo=150, hi=500, kt=2, seed=0, **extra):
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, S)
    codes = np.zeros((S, hi), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    p = {"L": 11, "k": 7, "d": 3, **extra}
    return GkmSVM(codes, rng.normal(size=S), 0.37, kt, p, sv_lengths=lens)


def _queries(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return ["".join("ACGTN"[c] for c in rng.choice(5, int(L), p=[.2475, .2475, .2475, .2475, .01]))
            for L in rng.integers(lo, hi + 1, n)]


@pytest.mark.parametrize("kt,extra", [(2, {}), (4, {"M": 50, "H": 50.0}), (3, {"gamma": 2.0})])
@pytest.mark.parametrize("table", [False, True])
def test_batch_scorer_matches_plain_calls(gpu, kt, extra, table):
    """Uneven queries (8-700 bp, 1% N, some shorter than l) through buckets,
    the autotuner (or the l-mer table) and graph replays equal model(x)
    exactly."""
    from gkmsvm_amd.batching import BatchScorer

    m = _random_model(kt=kt, **extra)
    m.use_lmer_table = table
    qs = _queries(300, 8, 700, 1)
    want = m(qs).cpu().numpy()
    sc = BatchScorer(m, rows=64, bucket=128, graphs=True, autotune=True)
    for rep in range(3):  # replays of captured graphs, eager work in between
        got = sc(qs).cpu().numpy()
        np.testing.assert_array_equal(got, want, err_msg=f"replay {rep}")
        m(qs[:5])
        torch.cuda.synchronize()
    assert (sc.plans and sc._graphs) or sc.queue_chunk  # buckets, or one persistent launch
    got = BatchScorer(m, rows=32, graphs=False, autotune=False)(qs).cpu().numpy()
    np.testing.assert_array_equal(got, want)
    sb = BatchScorer(m, rows=64, bucket=128, graphs=True, persistent=False)
    for rep in range(2):  # the bucketed path with its graphs, always
        np.testing.assert_array_equal(sb(qs).cpu().numpy(), want, err_msg=f"bucket {rep}")
    assert sb.plans and sb._graphs


@pytest.mark.parametrize("tables", [False, True])
def test_server_results_equal_alone_and_in_the_mix(gpu, tables):
    """Requests submitted concurrently from several threads (scoring of many
    lengths, variants, explain in the three modes with rows shorter than l),
    served in overlapping steps (two in flight), return what each gives alone,
    and empty requests their empty results (per-SV kernels, or the l-mer tables the
    server builds at start)."""
    from gkmsvm_amd import gkmexplain
    from gkmsvm_amd.serving import GkmServer

    m = _random_model(S=900, seed=3)
    if not tables:
        m.use_lmer_table = m.use_explain_table = False
    srv = GkmServer(m, hbm_budget=2 << 30, streams=2, tables=tables)
    assert (srv.table is not None) == tables and srv.explain_tables == tables
    rng = np.random.default_rng(4)
    reqs = []
    for i in range(60):
        u = rng.random()
        if u < 0.6:
            reqs.append(("score", _queries(int(rng.integers(1, 4)), 5, 900, 100 + i)))
        elif u < 0.85:
            ref = _queries(1, 200, 300, 200 + i)[0]
            alt = ref[:100] + ("A" if ref[100] != "A" else "C") + ref[101:]
            reqs.append(("variant", [(ref, alt)]))
        else:  # explain of 1-2 rows (some shorter than l: NaN), each mode
            n = int(rng.integers(1, 3))
            reqs.append(("explain", _queries(n, 5 if n == 2 else 300, 600, 300 + i)))
    modes = ["hypothetical", "perturbation", "importance"]
    alone = []
    for i, (kind, payload) in enumerate(reqs):
        if kind == "score":
            alone.append(m(payload).cpu().numpy().ravel())
        elif kind == "variant":
            alone.append(m.score_variants([p[0] for p in payload],
                                          [p[1] for p in payload]).cpu().numpy().ravel())
        else:
            alone.append(gkmexplain(m, payload, mode=modes[i % 3]).cpu().numpy())
    futs = [None] * len(reqs)
    with srv:
        def client(k):
            for i in range(k, len(reqs), 3):
                kind, payload = reqs[i]
                futs[i] = srv.submit(kind, payload, mode=modes[i % 3])
        ths = [threading.Thread(target=client, args=(k,)) for k in range(3)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        got = [f.result(timeout=120) for f in futs]
        empty = [srv.submit(k, []).result(timeout=60) for k in ("score", "variant", "explain")]
    assert [e.shape for e in empty] == [(0,), (0,), (0, 4, 0)]
    for (kind, _), g, a in zip(reqs, got, alone):
        if kind == "explain":
            H.assert_close(g, a, 1e-10, "explain")  # shapes and NaN rows too
        else:
            np.testing.assert_array_equal(g, a)


def test_server_admission_rejects_oversized_and_small_budget_still_exact(gpu):
    from gkmsvm_amd.serving import GkmServer

    m = _random_model(S=300, seed=5)
    m.use_lmer_table = m.use_explain_table = False
    qs = _queries(20, 50, 400, 9)
    want = m(qs).cpu().numpy().ravel()
    one = m._svs().S * m._params().nm * 8 + 8192
    with GkmServer(m, hbm_budget=one * 3, streams=2, tables=False) as srv:
        big = srv.submit("score", qs)  # 20 rows: over the budget
        with pytest.raises(MemoryError):
            big.result(timeout=60)
        futs = [srv.submit("score", [q]) for q in qs]  # one row each: admitted 3 per step
        got = np.array([f.result(timeout=120)[0] for f in futs])
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("tables", [False, True])
def test_server_bad_request_fails_alone(gpu, tables):
    """A request with an empty string or a letter other than ACGTN (incl. a
    non-ASCII one; scoring, variant or explain in either path) fails with
    ValueError (as model(x) would) while the requests of the same step score
    and explain as alone; short rows give NaN."""
    from gkmsvm_amd.serving import GkmServer

    from gkmsvm_amd import gkmexplain

    m = _random_model(S=300, seed=6)
    m.use_lmer_table = m.use_explain_table = tables  # the path the server takes
    qs = _queries(6, 20, 300, 11) + ["ACGTACG"]  # the last one shorter than l
    want = m(qs).cpu().numpy().ravel()
    ex = _queries(2, 100, 200, 12)
    want_ex = [gkmexplain(m, [q], mode=md).cpu().numpy()
               for q, md in zip(ex, ("hypothetical", "importance"))]
    with GkmServer(m, hbm_budget=1 << 30, tables=tables) as srv:
        srv.stop()  # queue everything first, so the bad requests share a step
        futs = [srv.submit("score", [q]) for q in qs[:3]]
        bad = [srv.submit("score", ["ACGTXACGTACGTACGT"]), srv.submit("variant", [("", "")]),
               srv.submit("score", ["ACGTACGTAC\u00e9GTACGT"]),  # non-ASCII
               srv.submit("explain", [ex[0], "ACGTACGTACGTACGT?"]),
               srv.submit("explain", ["ACGTNNNNacgtRACGT"], mode="importance")]
        futs += [srv.submit("score", [q]) for q in qs[3:]]
        fex = [srv.submit("explain", [q], mode=md) for q, md in zip(ex, ("hypothetical",
                                                                        "importance"))]
        srv.start()
        for f in bad:
            with pytest.raises(ValueError):
                f.result(timeout=60)
        got = np.array([f.result(timeout=60)[0] for f in futs])
        got_ex = [f.result(timeout=60) for f in fex]
    np.testing.assert_array_equal(got, want)
    assert np.isnan(got[-1])
    for g, w in zip(got_ex, want_ex):
        H.assert_close(g, w, 1e-10, "explain next to bad requests")


def test_server_survives_a_failed_step(gpu):
    """An error while a step's results are handed out fails that step's
    requests; the server goes on serving, and its history of completed
    requests stays within its bound."""
    from gkmsvm_amd.serving import GkmServer

    m = _random_model(S=300, seed=7)
    m.use_lmer_table = m.use_explain_table = False
    qs = _queries(4, 50, 300, 13)
    want = m(qs).cpu().numpy().ravel()
    with GkmServer(m, hbm_budget=1 << 30, tables=False, history=2) as srv:
        deliver, calls = srv._deliver, []

        def failing_once(st):
            calls.append(1)
            if len(calls) == 1:
                raise RuntimeError("delivery failed")
            deliver(st)

        srv._deliver = failing_once
        with pytest.raises(RuntimeError, match="delivery failed"):
            srv.submit("score", qs[:1]).result(timeout=60)
        got = np.array([srv.submit("score", [q]).result(timeout=60)[0] for q in qs])
    np.testing.assert_array_equal(got, want)
    assert len(srv.completed) == 2


@pytest.mark.parametrize("tables", [True, False])
def test_server_first_requests_see_finished_tables(gpu, tables):
    """Requests submitted right after GkmServer(...).start(), with no sync of
    the caller's stream in between, read the tables and caches the
    constructor built on that stream (l-mer and explain tables, or the SVs'
    l-mer weights, l-mer index and explain neighbourhoods) whole:
    perturbation explain and scoring equal the same calls made afterwards."""
    from gkmsvm_amd import gkmexplain
    from gkmsvm_amd.serving import GkmServer

    m = _random_model(S=2500, seed=8)
    if not tables:
        m.use_lmer_table = m.use_explain_table = False
        m.use_lmer_index = True
    qs = _queries(8, 200, 900, 21)
    ex = _queries(4, 300, 600, 22)
    srv = GkmServer(m, hbm_budget=2 << 30, tables=tables).start()
    if not tables:
        assert "V" in m._svs().tables  # built by the constructor, not by a step
    f_ex = [srv.submit("explain", [q], mode="perturbation") for q in ex]
    f_sc = [srv.submit("score", [q]) for q in qs]
    got_ex = [f.result(timeout=120) for f in f_ex]
    got_sc = np.array([f.result(timeout=120)[0] for f in f_sc])
    srv.stop()
    torch.cuda.synchronize()
    want_sc = m(qs).cpu().numpy().ravel()
    np.testing.assert_array_equal(got_sc, want_sc)
    for q, g in zip(ex, got_ex):
        H.assert_close(g, gkmexplain(m, [q], mode="perturbation").cpu().numpy(), 1e-10,
                       "explain right after start")
