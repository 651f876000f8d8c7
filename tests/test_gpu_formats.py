"""GPU scoring of the other model formats: classic / R gkmSVM models against
the tools' gkmsvm_kernel output, LS-GKM epsilon-SVR models against
gkmpredict, through the per-SV kernels and the l-mer table; a model written
and read back scores the same."""

import re

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O
from test_formats import CLASSIC, GKMSVM, load_classic, oracle_model, tool_scores

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prefix,tool,ft", CLASSIC)
@pytest.mark.parametrize("table", [False, True])
def test_classic_models_on_gpu(gpu, prefix, tool, ft, table):
    import os

    m = load_classic(prefix, tool, ft, device="rocm")
    m.use_lmer_table = table
    recs = O.read_fasta(H.path("lsgkm_wgkm_test100.fa"))
    queries = [re.sub("[^ACGT]", "", s) for _, s in recs]
    got = m(queries).cpu().numpy().ravel()
    codes, lens = O.batch(queries)
    H.assert_close(got, O.score(oracle_model(m), codes, lens), 1e-10, prefix)
    want = tool_scores(os.path.join(GKMSVM, f"{prefix}_test100_kernel.txt"))
    np.testing.assert_allclose(got, [want[n] for n, _ in recs], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("kind", ("t2", "t3", "t4"))
@pytest.mark.parametrize("table", [False, True])
def test_lsgkm_svr_models_on_gpu(gpu, kind, table):
    from gkmsvm_amd import load_lsgkm_model

    m = load_lsgkm_model(H.path(f"lsgkm_svr_{kind}.model.txt.gz"))
    m.use_lmer_table = table
    recs = O.read_fasta(H.path("lsgkm_wgkm_test85.fa"))
    seqs = [s for _, s in recs]
    got = m(seqs).cpu().numpy().ravel()
    codes, lens = O.batch(seqs)
    H.assert_close(got, O.score(oracle_model(m), codes, lens), 1e-10, kind)
    pred = H.read_pred(H.path(f"lsgkm_svr_{kind}_test85.pred.txt"))
    np.testing.assert_allclose(got, [pred[nm] for nm, _ in recs], rtol=2e-5, atol=1e-5)


def test_saved_models_score_the_same(gpu, tmp_path):
    from gkmsvm_amd import load_lsgkm_model, load_model, save_deltasvm_npz, save_lsgkm, save_npz

    m = load_lsgkm_model(H.path("lsgkm_ragged_t4.model.txt.gz"))
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_wgkm_test85.fa"))]
    want = m(seqs).cpu().numpy()
    save_npz(m, tmp_path / "m.npz")
    save_lsgkm(m, tmp_path / "m.model.txt.gz")
    for p in ("m.npz", "m.model.txt.gz"):
        np.testing.assert_array_equal(load_model(tmp_path / p)(seqs).cpu().numpy(), want)
    d = m.to_deltasvm(query_norm=True)
    save_deltasvm_npz(d, tmp_path / "d.npz")
    d2 = load_model(tmp_path / "d.npz")
    np.testing.assert_array_equal(d2(seqs).cpu().numpy(), d(seqs).cpu().numpy())


def test_deltasvm_weight_file_scores(gpu, tmp_path):
    """A 3-column deltaSVM file (k = 6 within l = 10 windows) scores as the
    oracle's gapped k-mer sum."""
    from gkmsvm_amd import load_deltasvm_model
    from gkmsvm_amd.deltasvm import index_to_kmer

    rng = np.random.default_rng(3)
    k = 6
    comp = str.maketrans("ACGT", "TGCA")
    lines = []
    for idx in rng.choice(4 ** k, 600, replace=False):
        km = index_to_kmer(int(idx), k)
        lines.append(f"{km}\t{km.translate(comp)[::-1]}\t{rng.normal():.6f}")
    p = tmp_path / "w.txt"
    p.write_text("\n".join(lines) + "\n")
    d = load_deltasvm_model(p, 10)
    recs = O.read_fasta(H.path("lsgkm_ragged_test.fa"))
    codes, lens = O.batch([s for _, s in recs])
    got = d(codes, lengths=lens).cpu().numpy().ravel()
    want = O.deltasvm_score(d.weights, 10, k, codes, lens, include_rc=d.include_rc)
    H.assert_close(got, want, 1e-12, "deltasvm file")


def test_strings_decoded_on_the_device(gpu):
    """DeviceSeqs.from_strings (the strings' bytes decoded by gkm_decode_ascii)
    gives the codes, planes and windows of encode_strings + DeviceSeqs, and flags
    exactly the rows encode_strings flags for a letter other than ACGTN (either
    case, non-ASCII included)."""
    import torch

    from gkmsvm_amd import engine as E
    from gkmsvm_amd.codec import encode_strings

    rng = np.random.default_rng(8)
    letters = list("ACGTacgtNn")
    strs = ["".join(rng.choice(letters, int(rng.integers(1, 700)))) for _ in range(40)]
    strs += ["", "A", "ACGTACGTACG", "acgtXacgt", "ACGTéACGT", "ACGT ACGT", "NNNNNNNNNNNN",
             "ACGU", "@ACGT", "acgt[", "ACGTacgt" * 300]
    codes, lens, bad = encode_strings(strs, 11)
    xs, dbad = E.DeviceSeqs.from_strings(strs, gpu, rc=True, min_width=11)
    ref = E.DeviceSeqs(codes, lens, gpu, rc=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(xs.lengths, lens)
    assert xs.lmax == codes.shape[1]
    np.testing.assert_array_equal((dbad.cpu().numpy() != 0) | (lens == 0), bad)
    ok = ~bad
    np.testing.assert_array_equal(xs.codes.cpu().numpy()[ok], codes[ok])
    np.testing.assert_array_equal(xs.planes.cpu().numpy()[ok], ref.planes.cpu().numpy()[ok])
    np.testing.assert_array_equal(xs.planes_rc.cpu().numpy()[ok], ref.planes_rc.cpu().numpy()[ok])
    fw, rw = xs.windows(11)
    fr, rr = ref.windows(11)
    np.testing.assert_array_equal(fw.cpu().numpy()[ok], fr.cpu().numpy()[ok])
    np.testing.assert_array_equal(rw.cpu().numpy()[ok], rr.cpu().numpy()[ok])
