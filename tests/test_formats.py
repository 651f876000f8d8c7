"""Model files on the CPU: classic / R gkmSVM models and LS-GKM SVR models
scored by the oracle against their tools' golden outputs, npz and LS-GKM text
round trips, and deltaSVM weight files."""

import gzip
import os
import re

import numpy as np
import pytest

import helpers as H
from conftest import GOLDEN
from oracle import oracle as O

GKMSVM = os.path.join(GOLDEN, "gkmsvm")
# (fixture prefix, tool, gkmsvm_kernel -t / R useTgkm); l=10 k=6 d=3
CLASSIC = [("gkmsvm_classic_t0", "classic", 0), ("gkmsvm_classic_t1", "classic", 1),
           ("gkmsvm_classic_t2", "classic", 2), ("gkmsvm_r_t1", "r", 1), ("gkmsvm_r_t2", "r", 2)]


def load_classic(prefix, tool, filter_type, device="cpu"):
    from gkmsvm_amd import load_classic_model, load_r_gkmsvm_model

    alpha = os.path.join(GKMSVM, f"{prefix}_svalpha.out")
    with pytest.warns(UserWarning, match="other than ACGT, dropped"):
        if tool == "r":
            return load_r_gkmsvm_model(alpha, useTgkm=filter_type, device=device)
        return load_classic_model(alpha, filter_type=filter_type, device=device)


def tool_scores(path):
    return dict((ln.split()[0], float(ln.split()[1])) for ln in open(path) if ln.strip())


def oracle_model(m):
    from gkmsvm_amd.inputs import row_lengths

    return O.Model(O.Kern.make(m.kernel_type, m.kernel_params), m.support_sequences,
                   row_lengths(m.support_sequences, m.sv_lengths), m.coefficients, m.bias)


@pytest.mark.parametrize("prefix,tool,ft", CLASSIC)
@pytest.mark.parametrize("test", ["test100", "test85"])
def test_classic_models_match_gkmsvm_kernel(prefix, tool, ft, test):
    """sum_i alpha_i K(x, sv_i) with the oracle's K equals the tools'
    gkmsvm_kernel output (7 significant digits) for queries read as the tools
    read them (non-ACGT dropped), and gkmsvm_classify within its ~2 %
    (tests/test_importers/test_gkmsvm_tools.py:38-58)."""
    m = load_classic(prefix, tool, ft)
    assert m.bias == 0.0 and m.sv_lengths is not None
    recs = O.read_fasta(H.path(f"lsgkm_wgkm_{test}.fa"))
    queries = [re.sub("[^ACGT]", "", s) for _, s in recs]
    codes, lens = O.batch(queries)
    got = O.score(oracle_model(m), codes, lens)
    want = tool_scores(os.path.join(GKMSVM, f"{prefix}_{test}_kernel.txt"))
    np.testing.assert_allclose(got, [want[n] for n, _ in recs], rtol=1e-5, atol=1e-6)
    cls = tool_scores(os.path.join(GKMSVM, f"{prefix}_{test}_classify.txt"))
    assert np.abs(got - [cls[n] for n, _ in recs]).max() < 0.03


def test_classic_loader_errors(tmp_path):
    from gkmsvm_amd import load_classic_model

    a = tmp_path / "m_svalpha.out"
    a.write_text("p0\t1.0\np1\t-0.5\n")
    (tmp_path / "m_svseq.fa").write_text(">p0\nACGTACGTACGT\n")
    with pytest.raises(ValueError, match="not in"):
        load_classic_model(a, device="cpu")
    (tmp_path / "m_svseq.fa").write_text(">p0\nACGTACGTACGT\n>p1\nTTTTGGGGCCCCAAAA\n")
    with pytest.raises(NotImplementedError):
        load_classic_model(a, filter_type=3, device="cpu")
    m = load_classic_model(a, filter_type=2, L=6, k=4, d=2, include_rc=False, device="cpu")
    assert m.kernel_type == "gkm_cnt" and m.kernel_params["include_rc"] is False
    np.testing.assert_array_equal(m.coefficients, [1.0, -0.5])
    with pytest.raises(ValueError, match="svseq_path"):
        load_classic_model(tmp_path / "other.txt", device="cpu")


@pytest.mark.parametrize("kind", ("t2", "t3", "t4"))
@pytest.mark.parametrize("n", (100, 85))
def test_lsgkm_svr_models_match_gkmpredict(kind, n):
    """epsilon-SVR LS-GKM models (bias = -rho, no labels) score as gkmpredict."""
    from gkmsvm_amd import load_lsgkm_model

    m = load_lsgkm_model(H.path(f"lsgkm_svr_{kind}.model.txt.gz"), device="cpu")
    recs = O.read_fasta(H.path(f"lsgkm_wgkm_test{n}.fa"))
    codes, lens = O.batch([s for _, s in recs])
    got = O.score(oracle_model(m), codes, lens)
    pred = H.read_pred(H.path(f"lsgkm_svr_{kind}_test{n}.pred.txt"))
    np.testing.assert_allclose(got, [pred[nm] for nm, _ in recs], rtol=2e-5, atol=1e-5)


def _same_model(a, b):
    assert a.kernel_type == b.kernel_type
    assert a.kernel_params == b.kernel_params
    assert a.bias == b.bias
    np.testing.assert_array_equal(a.coefficients, b.coefficients)
    la = a.sv_lengths if a.sv_lengths is not None else np.full(len(a.coefficients),
                                                                 a.support_sequences.shape[1])
    lb = b.sv_lengths if b.sv_lengths is not None else np.full(len(b.coefficients),
                                                                 b.support_sequences.shape[1])
    np.testing.assert_array_equal(la, lb)
    for i, n in enumerate(la):
        np.testing.assert_array_equal(a.support_sequences[i, :n], b.support_sequences[i, :n])


@pytest.mark.parametrize("name", ["lsgkm_wgkm_t4.model.txt.gz", "lsgkm_ragged_t5.model.txt.gz",
                                  "lsgkm_mode5_rbf.model.txt.gz"])
def test_npz_round_trip(tmp_path, name):
    from gkmsvm_amd import load_lsgkm_model, load_model, load_npz, save_npz

    m = load_lsgkm_model(H.path(name), device="cpu")
    m.sv_chunk_size = 17
    p = tmp_path / "m.npz"
    save_npz(m, p)
    for back in (load_npz(p, device="cpu"), load_model(p, device="cpu")):
        _same_model(m, back)
        assert back.sv_chunk_size == 17
    # plain arrays and a JSON record: loadable without pickles
    with np.load(p, allow_pickle=False) as d:
        assert set(d.files) >= {"support_sequences", "coefficients", "metadata"}


@pytest.mark.parametrize("gz", [False, True])
@pytest.mark.parametrize("name", ["lsgkm_wgkm_t4.model.txt.gz", "lsgkm_ragged_t2.model.txt.gz",
                                  "lsgkm_wgkm_t6.model.txt.gz"])
def test_lsgkm_text_round_trip(tmp_path, name, gz):
    """save_lsgkm writes what load_lsgkm_model reads back: kernel, params,
    rho = -bias, ragged SVs without padding, N kept as A (the loader's
    reading), nr_sv by coefficient sign."""
    from gkmsvm_amd import load_lsgkm_model, save_lsgkm

    m = load_lsgkm_model(H.path(name), device="cpu")
    p = tmp_path / ("m.model.txt.gz" if gz else "m.model.txt")
    save_lsgkm(m, p)
    back = load_lsgkm_model(p, device="cpu")
    _same_model(m, back)
    text = (gzip.open(p, "rt") if gz else open(p)).read().splitlines()
    assert text[0] == "svm_type c_svc" and "SV" in text
    npos = int((m.coefficients > 0).sum())
    assert f"nr_sv {npos} {len(m.coefficients) - npos}" in text
    assert "norc 1" not in text


def test_lsgkm_norc_written_only_when_off(tmp_path):
    from gkmsvm_amd import GkmSVM, load_lsgkm_model, save_lsgkm

    rng = np.random.default_rng(0)
    m = GkmSVM(rng.integers(0, 4, (5, 30)).astype(np.uint8), rng.normal(size=5), 0.25, 2,
               {"L": 8, "k": 5, "d": 2, "include_rc": False}, device="cpu")
    p = tmp_path / "m.txt"
    save_lsgkm(m, p)
    assert "norc 1" in p.read_text().splitlines()
    back = load_lsgkm_model(p, device="cpu")
    assert back.kernel_params["include_rc"] is False and back.bias == 0.25


def test_deltasvm_npz_round_trip(tmp_path):
    from gkmsvm_amd import DeltaSVM, load_deltasvm_npz, load_model, save_deltasvm_npz

    w = np.random.default_rng(1).normal(size=4 ** 5)
    d = DeltaSVM(w, 8, 5, include_rc=False, bias=0.5, device="cpu")
    p = tmp_path / "d.npz"
    save_deltasvm_npz(d, p)
    for back in (load_deltasvm_npz(p, device="cpu"), load_model(p, device="cpu")):
        assert isinstance(back, DeltaSVM)
        assert (back.l, back.k, back.include_rc, back.bias) == (8, 5, False, 0.5)
        np.testing.assert_array_equal(back.weights, w)


def test_deltasvm_weight_files(tmp_path):
    """2-column files score both strands; 3-column (RC-collapsed) files set
    both k-mers and default to one strand; header lines and comments."""
    from gkmsvm_amd import load_deltasvm_model
    from gkmsvm_amd.deltasvm import kmer_to_index

    two = tmp_path / "w2.txt"
    two.write_text("kmer\tweight\nAAC\t0.5\nGTT\t-1.25\n")
    d = load_deltasvm_model(two, 5, device="cpu")
    assert d.k == 3 and d.l == 5 and d.include_rc
    assert d.weights[kmer_to_index("AAC")] == 0.5 and d.weights[kmer_to_index("GTT")] == -1.25
    assert np.count_nonzero(d.weights) == 2
    three = tmp_path / "w3.txt.gz"
    with gzip.open(three, "wt") as f:
        f.write("AAC\tGTT\t0.75\nACG\tCGT\t2\n")
    d = load_deltasvm_model(three, 4, device="cpu")
    assert not d.include_rc
    for km, v in (("AAC", 0.75), ("GTT", 0.75), ("ACG", 2.0), ("CGT", 2.0)):
        assert d.weights[kmer_to_index(km)] == v
    c = tmp_path / "wc.txt"
    c.write_text("# trained with include_rc=False\nAC\t1\n")
    assert not load_deltasvm_model(c, 3, device="cpu").include_rc
    bad = tmp_path / "bad.txt"
    bad.write_text("ACG\t1\nAC\t2\n")
    with pytest.raises(ValueError, match="length"):
        load_deltasvm_model(bad, 4, device="cpu")
    with pytest.raises(ValueError, match="must be >="):
        load_deltasvm_model(two, 2, device="cpu")
