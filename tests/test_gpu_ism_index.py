"""ISM through the l-mer index (gkm_index_ism): every SV's K_raw(mutant, s) =
K_raw(x, s) - K_raw(old windows, s) + K_raw(new windows, s), the windows
covering each position probed as they are and substituted in two persistent
launches.  Against the sweep kernel (within 1e-9) and the oracle's rescoring
of every mutant from scratch (within 1e-9), for RBF and linear kernels,
weighted and unweighted, 8- and 16-bit counts with their carry fallbacks
(poly-A runs), several SV tiles and ragged queries down to one window; and
the same bits on every call."""

import numpy as np
import pytest
import torch

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _model(kt, extra, seed, S, tile):
    from gkmsvm_amd import GkmSVM
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(seed)
    lens = rng.integers(60, 301, S)
    codes = np.zeros((S, 300), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    codes[:20, :90] = 0  # poly-A SVs: with the queries' poly-A runs, counts past 255
    lens[:20] = np.maximum(lens[:20], 90)
    p = {"L": 11, "k": 7, "d": 3, **extra}
    m = GkmSVM(codes, rng.normal(size=S), 0.2, kt, p, sv_lengths=lens)
    m.use_lmer_table = False  # the index or the sweep, not the table
    kp, sv = m._params(), m._svs()
    sv._index = E.DeviceIndex(kp, sv.seqs, tile=tile)  # several tiles
    assert sv._index.n_tiles >= 2
    return m


def _queries(seed, lens, l):
    rng = np.random.default_rng(seed)
    lens = np.asarray(lens)
    codes = np.zeros((len(lens), lens.max()), np.uint8)
    for i, n in enumerate(lens):
        codes[i, :n] = rng.integers(0, 4, n)
    codes[0, : min(lens[0], 40)] = 0  # a poly-A run
    return codes, lens


CASES = [
    (5, {"M": 50, "H": 50.0, "gamma": 2.0}),   # wgkmrbf: weighted 16-bit counts
    (3, {"gamma": 1.5}),                         # gkmrbf: 8-bit counts
    (2, {}),                                     # linear, 8-bit
    (4, {"M": 50, "H": 50.0}),                   # weighted linear
    (6, {"gamma": 0.5}),                         # gkm counts + RBF
    (5, {"M": 50, "H": 50.0, "gamma": 2.0, "include_rc": False}),
    (3, {"L": 10, "k": 9, "d": 1, "gamma": 1.0}),  # 2 levels: 16-bit counts
]


@pytest.mark.parametrize("kt,extra", CASES)
def test_index_ism_matches_the_sweep(gpu, kt, extra):
    from gkmsvm_amd import engine as E
    from gkmsvm_amd import ism

    m = _model(kt, extra, 10 + kt, 1500, 512)
    kp, sv = m._params(), m._svs()
    assert E.index_ism_ok(kp, sv.index())
    codes, lens = _queries(20 + kt, [120, kp.l, kp.l + 1, 57, 300], kp.l)
    a = ism(m, codes, lengths=lens, use_index=True).cpu().numpy()
    b = ism(m, codes, lengths=lens, use_index=False).cpu().numpy()
    H.assert_close(a, b, 1e-9, f"-t {kt} {extra} index ISM vs sweep")
    c = ism(m, codes, lengths=lens, use_index=True).cpu().numpy()
    np.testing.assert_array_equal(a, c)  # deterministic
    # the reference base and positions past each length: 0
    for q in range(len(lens)):
        assert (a[q, codes[q, : lens[q]], np.arange(lens[q])] == 0).all()
        assert (a[q, :, lens[q]:] == 0).all()


@pytest.mark.parametrize("kt,extra", [CASES[0], CASES[1], CASES[3]])
def test_index_ism_matches_rescoring(gpu, kt, extra):
    """Every mutant scored from scratch by the oracle."""
    from gkmsvm_amd import ism

    m = _model(kt, extra, 30 + kt, 300, 128)
    kp = m._params()
    codes, lens = _queries(40 + kt, [45, kp.l, 30], kp.l)
    got = ism(m, codes, lengths=lens, use_index=True).cpu().numpy()
    om = O.Model(O.Kern.make(m.kernel_type, m.kernel_params), m.support_sequences,
                 np.asarray(m.sv_lengths), m.coefficients, m.bias)
    H.assert_close(got, O.ism(om, codes, lens), 1e-9, f"-t {kt} index ISM vs oracle")
