"""The SMO solver (gkm_smo.c) on the CPU: on precomputed kernels from the
oracle it reaches LIBSVM's solution (scikit-learn, the same problem and
stopping tolerance) and the KKT conditions, for C-SVC with and without a
positive-class weight, for epsilon-SVR, and with a column cache far smaller
than N."""

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O


def _kernel_matrix(kt, p, seqs):
    codes, lens = O.batch(seqs)
    return O.gram(O.Kern.make(kt, p), codes, lens)


def _columns(K, calls):
    def f(idx):
        calls.append(len(idx))
        return K[np.asarray(idx)]
    return f


@pytest.fixture(scope="module")
def svc_set():
    pos = [s for _, s in O.read_fasta(H.path("lsgkm_wgkm_pos.fa"))]
    neg = [s for _, s in O.read_fasta(H.path("lsgkm_wgkm_neg.fa"))]
    K = _kernel_matrix(2, {"L": 10, "k": 6, "d": 3}, pos + neg)
    y = np.r_[np.ones(len(pos)), -np.ones(len(neg))]
    return K, y


@pytest.mark.parametrize("pos_weight,cache", [(1.0, None), (2.0, None), (0.5, 16)])
def test_smo_matches_libsvm_csvc(svc_set, pos_weight, cache):
    from sklearn.svm import SVC

    from gkmsvm_amd.train import smo_solve_columns

    K, y = svc_set
    N = len(y)
    C = np.where(y > 0, pos_weight, 1.0)
    calls = []
    alpha, rho, stats = smo_solve_columns(_columns(K, calls), N, y, -np.ones(N), C, tol=1e-3,
                                          cache_size=cache, prefetch=8)
    assert stats["converged"] and stats["gap"] < 1e-3
    if cache is not None:
        assert max(calls) <= cache - 1
    coef = y * alpha
    f = K @ coef - rho
    # KKT at the stopping tolerance: y f = 1 on free SVs, >= 1 at 0, <= 1 at C
    free = (alpha > 1e-8) & (alpha < C - 1e-8)
    yf = y * f
    assert np.abs(yf[free] - 1).max() < 2e-3
    assert (yf[alpha <= 1e-8] > 1 - 2e-3).all() and (yf[alpha >= C - 1e-8] < 1 + 2e-3).all()
    assert abs(y @ alpha) < 1e-9
    ref = SVC(C=1.0, kernel="precomputed", tol=1e-3,
              class_weight=None if pos_weight == 1.0 else {1: pos_weight}).fit(K, y)
    rc = np.zeros(N)
    rc[ref.support_] = ref.dual_coef_[0]
    np.testing.assert_allclose(coef, rc, atol=5e-3)
    np.testing.assert_allclose(f, ref.decision_function(K), atol=5e-3)


def test_smo_matches_libsvm_svr():
    from sklearn.svm import SVR

    from gkmsvm_amd.train import smo_solve_columns

    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_svr_train.fa"))]
    z = np.loadtxt(H.path("lsgkm_svr_labels.txt"))
    K = _kernel_matrix(2, {"L": 10, "k": 6, "d": 3}, seqs)
    N = len(z)
    y = np.r_[np.ones(N), -np.ones(N)]
    p = np.r_[0.1 - z, 0.1 + z]
    alpha, rho, stats = smo_solve_columns(_columns(K, []), N, y, p, 1.0, tol=1e-3, prefetch=16)
    assert stats["converged"]
    coef = alpha[:N] - alpha[N:]
    ref = SVR(C=1.0, epsilon=0.1, kernel="precomputed", tol=1e-3).fit(K, z)
    rc = np.zeros(N)
    rc[ref.support_] = ref.dual_coef_[0]
    np.testing.assert_allclose(coef, rc, atol=5e-3)
    np.testing.assert_allclose(K @ coef - rho, ref.predict(K), atol=5e-3)


def test_smo_column_errors_surface():
    from gkmsvm_amd.train import smo_solve_columns

    def bad(idx):
        raise KeyError("no such column")

    with pytest.raises(RuntimeError, match="column computation failed"):
        smo_solve_columns(bad, 4, np.array([1.0, 1, -1, -1]), -np.ones(4), 1.0)
