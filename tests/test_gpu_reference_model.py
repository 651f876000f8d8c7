"""The reference's own benchmark model: the ENCODE ENCFF579AOX LS-GKM model
(72,145 SVs of 300 bp, l=11 k=7 d=3; a fixture of the reference's test suite)
scoring 300 bp queries as -t 2 (the file's kernel) and -t 3 (gkmrbf, gamma 1,
the same SVs and coefficients) -- the shape of the reference's throughput
benchmark (benchmarks/throughput_sweep.py).  -t 2 equals the oracle's l-mer
index scoring (its C restatement of kernels/lmer_index_cpu.py:70-120) within
1e-10, through the per-SV path and the exact l-mer table.  Under -t 3 the
72,145 terms coef_s exp(k_s - 1) cancel to scores far smaller than their
absolute sum, so the order of a float64 sum shows at ~1e-9 of the score:
the GPU score is held to 1e-10 of the correctly rounded sum (math.fsum) of the
per-SV terms computed from its bit-exact histograms with the oracle's
formulas, and the oracle's sequential sum to 1e-8 of it."""

import math

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODEL = "encode_ENCFF579AOX.model.txt.gz"


@pytest.fixture(scope="module")
def encode():
    from gkmsvm_amd import load_lsgkm_model

    return load_lsgkm_model(H.path(MODEL))


def _oracle(m):
    S, L = m.support_sequences.shape
    lens = np.full(S, L) if m.sv_lengths is None else np.asarray(m.sv_lengths)
    return O.Model(O.Kern.make(m.kernel_type, m.kernel_params), m.support_sequences, lens,
                   m.coefficients, m.bias)


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, 4, (n, 300)).astype(np.uint8)
    codes[0, 100:160] = 0  # a poly-A run: long buckets in the index
    return codes, np.full(n, 300)


@pytest.mark.parametrize("rbf", [False, True])
def test_encode_model_scores_equal_the_oracle(gpu, encode, rbf):
    from gkmsvm_amd import GkmSVM

    m = encode
    if rbf:
        m = GkmSVM(encode.support_sequences, encode.coefficients, encode.bias, "gkmrbf",
                   {**encode.kernel_params, "gamma": 1.0}, sv_lengths=encode.sv_lengths)
    assert len(m.coefficients) == 72145
    codes, lens = _queries(12, 3 if rbf else 2)
    om = _oracle(m)
    want = O.LmerIndex(om).score(codes, lens)
    m.use_lmer_table = False
    got = m(codes, lengths=lens).cpu().numpy().ravel()
    if rbf:
        from gkmsvm_amd import engine as E

        kp, sv = m._params(), m._svs()
        xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
        hist = E.sv_hist(sv, xs, 0, xs.n).cpu().numpy()  # [B, S, nm], the SV model's order
        dx = O.combine(om.kern, O.self_hist(om.kern, O.pack_both(om.kern, codes, lens)))
        k = O.combine(om.kern, hist) / np.maximum(
            np.sqrt(dx[:, None] * sv.ds.cpu().numpy()[None, :]), 1e-10)
        terms = sv.coef.cpu().numpy()[None, :] * O._finish(om.kern, k)
        exact = np.array([math.fsum(t) for t in terms]) + m.bias
        H.assert_close(got, exact, 1e-10, "ENCFF579AOX -t 3 vs the correctly rounded sum")
        H.assert_close(want, exact, 1e-8, "oracle -t 3 vs the correctly rounded sum")
        return
    H.assert_close(got, want, 1e-10, "ENCFF579AOX per-SV")
    m.use_lmer_table = True
    got = m(codes, lengths=lens).cpu().numpy().ravel()
    H.assert_close(got, want, 1e-10, "ENCFF579AOX l-mer table")
