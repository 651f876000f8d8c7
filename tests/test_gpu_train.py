"""Training on the GPU against gkmtrain (LS-GKM and lsgkm-svr): models
trained on the fixture sets have gkmtrain's support vectors, its
coefficients to the solvers' 1e-3 stopping tolerance and gkmpredict's
scores -- through the precomputed-Gram LIBSVM solver and the column-cached
SMO (tests/test_train_oracle.py of gkmsvm-lite holds the same checks)."""

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# gkmtrain's options for each fixture model (-c 1.0; LS-GKM's default d is 3)
PARAMS = {
    "t0": dict(kernel_type=0, l=10, k=6, d=3),
    "t1": dict(kernel_type=1, l=11, k=7, d=3),
    "t2": dict(kernel_type=2, l=10, k=6, d=3),
    "t3": dict(kernel_type=3, l=10, k=6, d=3, gamma=2.0),
    "t4": dict(kernel_type=4, l=11, k=7, d=3, M=20, H=10),
    "t5": dict(kernel_type=5, l=10, k=6, d=3, M=50, H=30, gamma=2.0),
    "t6": dict(kernel_type=6, l=10, k=6, d=3, gamma=2.0),
    "t2w2": dict(kernel_type=2, l=10, k=6, d=3, pos_weight=2.0),
    "t4w05": dict(kernel_type=4, l=11, k=7, d=3, M=20, H=10, pos_weight=0.5),
}
SVR_PARAMS = {
    "t2": dict(kernel_type=2, l=10, k=6, d=3),
    "t3": dict(kernel_type=3, l=10, k=6, d=3, gamma=2.0),
    "t4": dict(kernel_type=4, l=10, k=6, d=3, M=20, H=10),
}


def _seqs(name):
    return [s for _, s in O.read_fasta(H.path(name))]


def _sv_coefficients(model):
    from gkmsvm_amd.codec import decode_codes

    lens = (model.sv_lengths if model.sv_lengths is not None
            else np.full(model.num_support_vectors, model.support_sequences.shape[1]))
    return {decode_codes(sv[:n]): float(c)
            for sv, n, c in zip(model.support_sequences, lens, model.coefficients)}


def _check(model, ref_name, pred_prefix):
    from gkmsvm_amd import load_lsgkm_model

    ref = _sv_coefficients(load_lsgkm_model(H.path(ref_name), device="cpu"))
    got = _sv_coefficients(model)
    assert set(got) == set(ref)
    np.testing.assert_allclose([got[s] for s in ref], list(ref.values()), atol=5e-3)
    for n in (100, 85):
        recs = O.read_fasta(H.path(f"lsgkm_wgkm_test{n}.fa"))
        pred = H.read_pred(H.path(f"{pred_prefix}_test{n}.pred.txt"))
        np.testing.assert_allclose(model([s for _, s in recs]).cpu().numpy().ravel(),
                                   [pred[nm] for nm, _ in recs], atol=5e-4)


@pytest.mark.parametrize("name", PARAMS)
@pytest.mark.parametrize("solver", ["libsvm", "smo"])
def test_training_matches_gkmtrain(gpu, name, solver):
    from gkmsvm_amd import train_gkmsvm

    model = train_gkmsvm(_seqs("lsgkm_wgkm_pos.fa"), _seqs("lsgkm_wgkm_neg.fa"), C=1.0,
                         solver=solver, **PARAMS[name])
    _check(model, f"lsgkm_wgkm_{name}.model.txt.gz", f"lsgkm_wgkm_{name}")


@pytest.mark.parametrize("name", SVR_PARAMS)
@pytest.mark.parametrize("solver", ["libsvm", "smo"])
def test_svr_matches_lsgkm_svr(gpu, name, solver):
    from gkmsvm_amd import train_gkmsvr

    model = train_gkmsvr(_seqs("lsgkm_svr_train.fa"), np.loadtxt(H.path("lsgkm_svr_labels.txt")),
                         C=1.0, epsilon=0.1, solver=solver, **SVR_PARAMS[name])
    _check(model, f"lsgkm_svr_{name}.model.txt.gz", f"lsgkm_svr_{name}")


def test_smo_with_small_cache_and_ragged_training_set(gpu):
    """Ragged training sequences (60-140 bp), a column cache of 24 columns:
    the same model as the Gram solver's, to the stopping tolerance."""
    from gkmsvm_amd import train_gkmsvm

    pos, neg = _seqs("lsgkm_ragged_pos.fa"), _seqs("lsgkm_ragged_neg.fa")
    a = train_gkmsvm(pos, neg, kernel_type=2, l=10, k=6, d=3, solver="libsvm")
    b = train_gkmsvm(pos, neg, kernel_type=2, l=10, k=6, d=3, solver="smo", cache_size=24)
    ca, cb = _sv_coefficients(a), _sv_coefficients(b)
    keys = set(ca) | set(cb)
    np.testing.assert_allclose([ca.get(s, 0.0) for s in keys], [cb.get(s, 0.0) for s in keys],
                               atol=5e-3)
    q = _seqs("lsgkm_ragged_test.fa")
    q = [s for s in q if len(s) >= 10]
    np.testing.assert_allclose(b(q).cpu().numpy(), a(q).cpu().numpy(), atol=5e-3)
