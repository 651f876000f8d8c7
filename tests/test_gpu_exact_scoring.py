"""Exact scoring of kernels without RBF: the decision value from 128-bit
integer level sums of fixed-point SV weights.  Through the l-mer index with
the counts reduced on the chip (gkm_index_score, every count format and its
carry fallback, several window chunkings), through the pairwise kernel's
histograms (gkm_combine_fixed), for a row alone or in a batch, the bits are
the same; and they agree with the oracle's float64 scores within 1e-10."""

import numpy as np
import pytest
import torch

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _model(kt, extra, seed, S=2500):
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(seed)
    lens = rng.integers(120, 501, S)
    codes = np.zeros((S, 500), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    codes[:40, :90] = 0  # poly-A SVs: with the queries' poly-A runs, counts past 255
    p = {"L": 11, "k": 7, "d": 3, **extra}
    return GkmSVM(codes, rng.normal(size=S), 0.3, kt, p, sv_lengths=lens)


def _queries(n, seed, lo=20, hi=900):
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, n)
    codes = np.zeros((n, hi), np.uint8)
    for i in range(n):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    codes[0, :70] = 0
    codes[1, 10:60] = 0
    lens[2] = 8  # shorter than l: no windows
    return codes, lens


@pytest.mark.parametrize("kt,extra,bits", [
    (2, {}, 8), (2, {}, 16), (0, {}, 8), (1, {}, 16),
    (4, {"M": 50, "H": 50.0}, 16), (4, {"M": 50, "H": 50.0}, 32),
    (2, {"include_rc": False}, 8)])
def test_exact_scores_have_the_same_bits_on_every_path(gpu, kt, extra, bits):
    from gkmsvm_amd import engine as E

    m = _model(kt, extra, 1)
    kp, sv = m._params(), m._svs()
    sv._index = E.DeviceIndex(kp, sv.seqs, count_bits=bits)
    codes, lens = _queries(24, 2)
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    dx = E.self_kraw(kp, xs) if kp.normalize else None
    runs = []
    for tb in (None, 64, 4096):  # window chunks of different sizes
        out = torch.empty(xs.n, dtype=torch.float64, device=kp.device)
        E.index_score(sv, xs, 0, xs.n, dx, m.bias, out, target_blocks=tb)
        runs.append(out.cpu().numpy())
    hist = E.cross_hist(kp, xs, 0, xs.n, sv.seqs, sv.groups)
    out = torch.empty(xs.n, dtype=torch.float64, device=kp.device)
    from gkmsvm_amd import _lib

    _lib.call("gkm_combine_fixed", hist.data_ptr(), xs.n, sv.S, kp.nm, kp.tab.data_ptr(),
              dx.data_ptr() if dx is not None else None, sv.wfix.data_ptr(), sv.fix_shift,
              int(kp.normalize), m.bias, out.data_ptr(), E.stream_handle())
    runs.append(out.cpu().numpy())
    for r in runs[1:]:
        np.testing.assert_array_equal(r, runs[0])
    # each row alone (a batch of one: other widths, other chunks)
    for b in (0, 1, 5, 23):
        x1 = E.DeviceSeqs(codes[b:b + 1, : max(lens[b], 1)], lens[b:b + 1], kp.device, rc=kp.rc)
        d1 = E.self_kraw(kp, x1) if kp.normalize else None
        o1 = torch.empty(1, dtype=torch.float64, device=kp.device)
        E.index_score(sv, x1, 0, 1, d1, m.bias, o1)
        assert o1.item() == runs[0][b], b
    # the oracle's float64 scores (rows with windows)
    om = O.Model(O.Kern.make(m.kernel_type, m.kernel_params), m.support_sequences,
                 np.asarray(m.sv_lengths), m.coefficients, m.bias)
    want = O.score(om, codes, lens)
    keep = lens >= kp.l
    H.assert_close(runs[0][keep], want[keep], 1e-10, f"-t {kt} exact scores vs oracle")


def test_model_scores_through_either_path_are_identical(gpu):
    """model(x) with the index forced on and off: the same bits (and the
    batch scorer's, with its autotuned plans)."""
    from gkmsvm_amd.batching import BatchScorer

    m = _model(2, {}, 3)
    codes, lens = _queries(40, 4)
    m.use_lmer_table = False
    m.use_lmer_index = True
    a = m(codes, lengths=lens).cpu().numpy()
    m.use_lmer_index = False
    b = m(codes, lengths=lens).cpu().numpy()
    np.testing.assert_array_equal(a, b)
    m.use_lmer_index = "auto"
    c = BatchScorer(m, rows=16, bucket=128, graphs=True)(codes, lengths=lens).cpu().numpy()
    np.testing.assert_array_equal(c, a)


@pytest.mark.parametrize("kt,extra,bits", [(2, {}, 8), (2, {}, 16),
                                           (4, {"M": 50, "H": 50.0}, 16),
                                           (4, {"M": 50, "H": 50.0}, 32)])
def test_persistent_launch_has_the_same_bits(gpu, kt, extra, bits):
    """gkm_index_score_queue (every row in one persistent launch, work items
    popped from per-XCD queues) equals the grid launch and the pairwise
    kernel bit for bit, for several window chunks; rows without windows get
    the empty sums."""
    from gkmsvm_amd import _lib
    from gkmsvm_amd import engine as E

    m = _model(kt, extra, 5)
    kp, sv = m._params(), m._svs()
    sv._index = E.DeviceIndex(kp, sv.seqs, count_bits=bits)
    codes, lens = _queries(60, 6, lo=5, hi=1000)
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    dx = E.self_kraw(kp, xs)
    want = torch.empty(xs.n, dtype=torch.float64, device=kp.device)
    E.index_score(sv, xs, 0, xs.n, dx, m.bias, want)
    hist = E.cross_hist(kp, xs, 0, xs.n, sv.seqs, sv.groups)
    brute = torch.empty_like(want)
    _lib.call("gkm_combine_fixed", hist.data_ptr(), xs.n, sv.S, kp.nm, kp.tab.data_ptr(),
              dx.data_ptr(), sv.wfix.data_ptr(), sv.fix_shift, 1, m.bias, brute.data_ptr(),
              E.stream_handle())
    np.testing.assert_array_equal(brute.cpu().numpy(), want.cpu().numpy())
    for chunk in (7, 16, 64):
        got = torch.empty_like(want)
        E.index_score_queue(sv, xs, dx, m.bias, got, chunk=chunk)
        np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy(),
                                      err_msg=f"chunk {chunk}")


def test_persistent_launch_takes_more_rows_than_a_grid(gpu):
    """70,000 rows (past the grid's 65,535 rows per launch) in one persistent
    launch through the batch scorer: equal to model(x) row for row."""
    from gkmsvm_amd.batching import BatchScorer

    m = _model(2, {}, 7, S=300)
    m.use_lmer_table = False
    m.use_lmer_index = True
    rng = np.random.default_rng(8)
    n = 70_000
    lens = rng.integers(12, 41, n)
    codes = rng.integers(0, 4, (n, 40)).astype(np.uint8)
    sc = BatchScorer(m, autotune=False)
    got = sc(codes, lengths=lens).cpu().numpy().ravel()
    assert sc.queue_chunk is not None
    pick = np.arange(0, n, 997)
    want = m(codes[pick], lengths=lens[pick]).cpu().numpy().ravel()
    np.testing.assert_array_equal(got[pick], want)
