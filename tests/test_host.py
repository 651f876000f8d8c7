"""Host-side logic without a GPU: codecs, batches, kernel tables, window
weights, the LS-GKM loader, the C-ABI library's exports, and the engine's
lane code run on the CPU (libgkm_emu.so) against the oracle."""

import ctypes
import os
import re

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# -- codec / inputs ---------------------------------------------------------


def test_codec_round_trip_and_n():
    from gkmsvm_amd.codec import codes_to_one_hot, decode_codes, encode_codes, one_hot_encode, to_codes

    assert encode_codes("ACGTacgt").tolist() == [0, 1, 2, 3, 0, 1, 2, 3]
    assert encode_codes("ANNT", allow_n=True).tolist() == [0, 0, 0, 3]
    with pytest.raises(ValueError):
        encode_codes("ANT")
    with pytest.raises(ValueError):
        encode_codes("ARN", allow_n=True)
    assert encode_codes("ARN", allow_other=True).tolist() == [0, 0, 0]
    oh = one_hot_encode("ACNT", allow_n=True)
    assert oh.shape == (4, 4) and oh[:, 2].sum() == 0
    assert to_codes(oh[None]).tolist() == [[0, 1, 0, 3]]  # N -> A
    assert decode_codes(encode_codes("GATTACA")) == "GATTACA"
    np.testing.assert_array_equal(codes_to_one_hot(np.array([[3, 0]]))[0].argmax(0), [3, 0])


def test_read_batch_lists_lengths_and_masks():
    """Batches of strings, one-hot and padded codes; validation errors;
    allele pairs widened to one width; NaN rows and zeroed positions."""
    import torch

    from gkmsvm_amd.inputs import mask_positions, nan_rows, read_batch, read_pair, stacked

    b = read_batch(["ACG", "ACGTN", "a"])
    assert b.codes.shape == (3, 5) and b.lengths.tolist() == [3, 5, 1] and b.padded
    assert b.codes[1].tolist() == [0, 1, 2, 3, 0]  # N read as A
    b = read_batch(np.zeros((2, 4, 7)))
    assert b.codes.shape == (2, 7) and not b.padded and b.given_lengths() is None
    assert read_batch(np.zeros((2, 7), np.uint8), lengths=[7, 7]).given_lengths() is None
    for bad in (dict(x=np.zeros((2, 7), np.uint8), lengths=[8, 1]),
                dict(x=np.zeros((2, 7), np.uint8), lengths=[3]),
                dict(x=["ACGT"], lengths=[4]), dict(x=[]), dict(x=["ACGT", ""]),
                dict(x=["ACXT"])):
        with pytest.raises(ValueError):
            read_batch(**bad)
    with pytest.warns(UserWarning, match="2047"):
        read_batch(["A" * 2048])
    r, a = read_pair(["ACGTA"], ["ACGT"])
    assert r.codes.shape == a.codes.shape == (1, 5) and a.lengths.tolist() == [4] and a.padded
    with pytest.raises(ValueError):
        read_pair(["ACGT", "ACGT"], ["ACGT"])
    both = stacked(a, r)
    assert both.given_lengths().tolist() == [4, 5]
    with pytest.warns(UserWarning, match="fewer than l = 4"):
        short = read_batch(["ACG", "ACGTA"]).short(4)
    assert short.tolist() == [True, False]
    out = torch.ones((2, 4, 5), dtype=torch.float64)
    m = mask_positions(out, read_batch(["ACG", "ACGTA"]), short).numpy()
    assert np.isnan(m[0]).all() and (m[1] == 1).all()
    m = mask_positions(out, read_batch(["ACGTAC", "ACGTA"]).rows(1, 2), None).numpy()
    assert (m[0, :, :5] == 1).all()
    m = mask_positions(out[:1], read_batch(["ACGTAC", "AC"]).rows(1, 2), None).numpy()
    assert (m[0, :, :2] == 1).all() and (m[0, :, 2:] == 0).all()
    assert torch.isnan(nan_rows(torch.zeros((3, 1)), np.array([False, True, False]))[1, 0])


# -- kernel tables -------------------------------------------------------------


@pytest.mark.parametrize("kt", range(7))
@pytest.mark.parametrize("l,k", [(3, 2), (6, 4), (8, 5), (10, 6), (11, 7), (12, 8), (15, 9)])
def test_tables_match_oracle(kt, l, k):
    from gkmsvm_amd.kernels import build_kernel

    p = {"L": l, "k": k, "gamma": 2.0}
    got = build_kernel(kt, p).base._mismatch_table
    want = O.Kern.make(kt, p).table
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)


def test_table_known_answers():
    from gkmsvm_amd.kernels import build_kernel

    assert build_kernel(0, {"L": 3, "k": 2}).base._mismatch_table.tolist() == [3, 1, 0, 0]
    t = build_kernel(0, {"L": 11, "k": 7}).base
    np.testing.assert_array_equal(t._mismatch_table[:5], [330, 120, 36, 8, 1])
    assert t.dmax == 4
    t = build_kernel(0, {"L": 11, "k": 7, "d": 3}).base
    assert t._mismatch_table[4] == 0 and t.dmax == 3
    assert build_kernel(6, {"L": 11, "k": 7}).base.dmax == 3  # -t 6 defaults d = 3


@pytest.mark.parametrize("M,H", [(50, 50.0), (1, 1.0), (255, 7.5), (30, 30.0)])
def test_window_weights_formula(M, H):
    """The distance table reproduces the reference's per-length weights for
    every length (the same float64 expression, element by element)."""
    from gkmsvm_amd.kernels import distance_weights, window_weights

    tab = distance_weights(600, M, H)
    for W in list(range(1, 60)) + [290, 291, 990, 1001]:
        want = O.window_weights(W, M, H)
        np.testing.assert_array_equal(window_weights(W, M, H), want)
        if W // 2 < len(tab):
            np.testing.assert_array_equal(tab[np.abs(W // 2 - np.arange(W))], want)


def test_resolve_kernel_type():
    from gkmsvm_amd.kernels import resolve_kernel_type

    assert resolve_kernel_type(2) == "gkm_esttrunc"
    assert resolve_kernel_type("5") == "wgkmrbf"
    assert resolve_kernel_type("weighted") == "wgkm"
    assert resolve_kernel_type("gkm_cnt_rbf") == "gkm_cnt_rbf"
    with pytest.raises(ValueError):
        resolve_kernel_type(9)
    with pytest.raises(NotImplementedError):
        resolve_kernel_type("nope")


def test_explain_tables_match_oracle():
    from gkmsvm_amd.explain import explain_tables
    from gkmsvm_amd.kernels import build_kernel

    for kt, p in [(0, {"L": 11, "k": 7}), (2, {"L": 10, "k": 6}), (0, {"L": 10, "k": 6, "d": 3})]:
        base = build_kernel(kt, p).base
        for pert in (False, True):
            a, b, cmax = explain_tables(base, pert)
            oa, ob, mm = O.explain_tables(O.Kern.make(kt, p), pert)
            np.testing.assert_allclose(a, oa, rtol=1e-14)
            np.testing.assert_allclose(b, ob, rtol=1e-14)
            assert cmax == base.l - mm


# -- loader, device switch ---------------------------------------------------------


def test_load_lsgkm_header_and_svs():
    from gkmsvm_amd import load_lsgkm_model

    m = load_lsgkm_model(H.path("lsgkm_wgkm_t5.model.txt.gz"), device="cpu")
    om = O.load_lsgkm(H.path("lsgkm_wgkm_t5.model.txt.gz"))
    assert m.kernel_type == "wgkmrbf" and m.kernel.gamma == 2.0
    assert m.kernel.base.M == 50 and m.kernel.base.H == 30.0 and m.kernel.base.l == 10
    assert m.num_support_vectors == 160 and m.bias == om.bias
    np.testing.assert_array_equal(m.coefficients, om.coef)
    np.testing.assert_array_equal(m.support_sequences, om.sv_codes)
    r = load_lsgkm_model(H.path("lsgkm_ragged_t4.model.txt.gz"), device="cpu")
    assert r.sv_lengths is not None and r.sv_lengths.min() < r.sv_lengths.max()


def test_device_auto_without_gpu_holds_on_the_host():
    """device="auto" (and the loaders' device argument) picks the ROCm GPU
    when one is visible and otherwise holds the model on the host, as the
    reference's "auto" falls back to the CPU; unknown names raise."""
    import torch

    from gkmsvm_amd import GkmSVM, load_lsgkm_model, load_model
    from gkmsvm_amd import engine as E

    m = load_lsgkm_model(H.path("lsgkm_wgkm_t2.model.txt.gz"), device="auto")
    assert m.device == ("rocm" if torch.cuda.is_available() else "cpu")
    assert m.num_support_vectors > 0
    m2 = load_model(H.path("lsgkm_wgkm_t2.model.txt.gz"), device="auto")
    assert m2.device == m.device
    assert E.placement("cpu") is None
    with pytest.raises(ValueError, match="unknown device"):
        GkmSVM(np.zeros((2, 20), np.uint8), np.ones(2), 0.0, 2, {"L": 5, "k": 3}, device="tpu")


def test_one_hot_api_helpers():
    """encode_batch / one_hot_decode / reverse_complement / validate behave
    as the reference's codec functions (codec.py:155-220)."""
    import torch

    from gkmsvm_amd import encode_batch, one_hot_decode, reverse_complement, validate

    x, mask = encode_batch(["ACGTN", "GA"], allow_n=True, pad_value=-1.0)
    assert x.shape == (2, 4, 5) and x.dtype == np.float32
    assert mask.tolist() == [[True] * 5, [True, True, False, False, False]]
    assert (x[0, :, 4] == 0).all() and (x[1, :, 2:] == -1).all()
    with pytest.raises(ValueError):
        encode_batch(["ACGN"])
    with pytest.raises(ValueError):
        encode_batch([])
    assert one_hot_decode(x[0, :, :4]) == "ACGT"
    rc = reverse_complement(x[:1, :, :4])
    assert one_hot_decode(rc[0]) == "ACGT"  # ACGT is its own reverse complement
    y, _ = encode_batch(["AACG"])
    assert one_hot_decode(reverse_complement(y[0])) == "CGTT"
    assert one_hot_decode(reverse_complement(torch.from_numpy(y))[0]) == "CGTT"
    validate(y)
    with pytest.raises(ValueError):
        validate(x)  # the N column sums to 0
    with pytest.raises(ValueError):
        one_hot_decode(y)


def test_cpu_model_refuses_to_score():
    """No CPU fallback: a model not on ROCm raises instead of computing."""
    from gkmsvm_amd import GkmSVM

    m = GkmSVM(np.zeros((2, 20), np.uint8), np.ones(2), 0.0, 2, {"L": 5, "k": 3}, device="cpu")
    with pytest.raises(RuntimeError, match="ROCm"):
        m(np.zeros((1, 20), np.uint8))


def test_launch_shapes_fit_grid_limits():
    from gkmsvm_amd import engine as E

    for B in (1, 7, 1000, 200_000):
        for S in (1, 64, 500, 72_000):
            rpb, ns = E.launch_shape(B, S, 2)
            assert rpb >= 1 and 1 <= ns <= 16
            assert -(-B // rpb) <= 65535
    g = E.group_bounds(np.array([5, 9, 7] + [100] * 64))
    assert g.tolist() == [[5, 100], [100, 100]]


def test_index_launch_rules():
    """The l-mer index query's counts per LDS word, tile sizes and window chunks:
    8-bit where that saves words (nm >= 3), 16-bit otherwise and for weighted
    counts, tiles filling 144 KiB; chunks of 32 windows for rows of one length
    (16 under 128 blocks), 16 for rows of very different lengths."""
    from gkmsvm_amd import engine as E

    assert [E.index_pack(nm, True) for nm in (1, 2, 3, 4, 5, 8, 9)] == [2, 2, 4, 4, 4, 4, 4]
    assert E.index_pack(4, False) == 2
    assert E.index_tile(4, 4) == 36864 and E.index_tile(4, 2) == 18432
    assert E.index_tile(4, 1) == 9216 and E.index_tile(4, True) == 36864
    assert E.index_tile(4, False) == 18432 and E.index_tile(1, 4) == 36864
    l = 11
    # 16 x 1 kb against 2 tiles: chunks of 32 windows -> 31 per row
    assert E.index_blocks(np.full(16, 1000), l, 2) == 2 * 16 * 31
    # a single 1 kb query: under 128 blocks at 32 windows, so chunks of 16
    assert E.index_blocks(np.full(1, 1000), l, 2) == 2 * 1 * 62
    # ragged rows (max > 1.5 x mean): chunks of 16
    lens = np.array([60, 2000, 100, 150])
    assert E.index_blocks(lens, l, 2) == 2 * 4 * -(-(2000 - l + 1) // 16)
    # an explicit chunk
    assert E.index_blocks(np.full(8, 300), l, 3, 64) == 3 * 8 * -(-290 // 64)


def test_paired_masks():
    """The index query's mask order: a permutation of the neighbourhood (levels
    kept) whose leading entries pair up at lanes 2k, 2k + 1 with masks that
    differ in bit 0 of one digit q < tables -- q in each level byte's high
    nibble, the digit whose record table puts both probes in one 128-byte
    line -- and the unpaired masks last."""
    from gkmsvm_amd import engine as E
    from gkmsvm_amd.kernels import build_kernel

    kp_h = build_kernel(2, {"L": 11, "k": 7, "d": 3})._mismatch_table
    masks, levels = E.neighbor_masks(11, kp_h, 3)
    assert len(masks) == 4984
    want = dict(zip(masks.tolist(), levels.tolist()))
    for tables, n_pairs in ((11, 2432), (1, 872)):
        pm, pl = E.paired_masks(masks, levels, tables)
        assert sorted(pm.tolist()) == sorted(masks.tolist())
        assert all(want[m] == (v & 15) for m, v in zip(pm.tolist(), pl.tolist()))
        q = pl >> 4
        assert q.max() < tables
        k = 0
        while k + 1 < len(pm) and (int(pm[k]) ^ int(pm[k + 1])) == 1 << (2 * int(q[k])):
            assert q[k] == q[k + 1]
            k += 2
        assert k == 2 * n_pairs
        rest = pm[k:].tolist()  # no two unpaired masks could have been paired
        for a in rest:
            assert not any(a ^ (1 << (2 * d)) in set(rest) for d in range(tables))


# -- the C-ABI library -------------------------------------------------------------


def _header_functions():
    src = open(os.path.join(ROOT, "include", "gkmhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gkm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    """libgkmhip.so loads without a GPU and exports every function
    include/gkmhip.h declares; the ctypes table covers them all."""
    from gkmsvm_amd import _lib

    lib = _lib.load()
    names = _header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.SIGNATURES) == names
    assert lib.gkm_version() == 1
    assert lib.gkm_planes_words(300) == 300 // 32 + 4
    # persistent launch workspace: 256 B of queue counters, first[B] rounded
    # to 256 B, one int128 per (pair, tile, level)
    assert lib.gkm_index_queue_bytes(1000, 10, 3, 5) == 256 + 4096 + 10 * 3 * 5 * 16
    assert lib.gkm_index_queue_bytes(1, 0, 3, 5) == 512


def test_smo_library_exports_its_declared_symbol():
    """libgkm_smo.so exports what include/gkm_smo.h declares."""
    from gkmsvm_amd.train import smo_lib

    src = open(os.path.join(ROOT, "include", "gkm_smo.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(gkm_[a-z0-9_]+)\s*\(", src)))
    assert names == ["gkm_smo_solve"]
    lib = smo_lib()
    for n in names:
        assert hasattr(lib, n), n


def test_library_rejects_bad_arguments_without_gpu_work():
    """Argument checks fail before any launch (l out of range)."""
    from gkmsvm_amd import _lib

    with pytest.raises(_lib.GkmError, match="l must be"):
        _lib.call("gkm_cross_hist", None, None, 1, 4, None, None, None, None, 1, 4, 16, 1, None,
                  0, 1, 1, -1, 0, None, None)
    ws = ctypes.create_string_buffer(1024)
    aligned = (ctypes.addressof(ws) + 255) // 256 * 256
    with pytest.raises(_lib.GkmError, match="n_blocks"):
        _lib.call("gkm_index_score_queue", None, None, None, 4, 100, None, None, None, None, 1,
                  None, None, 1, 11, 4, None, 1000, 1000, 1, 16, 2, 1, 4, 0, aligned, 1,
                  None, 0, 0, 0.0, 1, None)


# -- lane code emulated on the CPU -----------------------------------------------------


def _emu():
    from gkmsvm_amd import build_ext

    path = os.path.join(build_ext.LIBDIR, "libgkm_emu.so")
    if not os.path.exists(path):
        build_ext.build_emu()
    E = ctypes.CDLL(path)
    P, I = ctypes.c_void_p, ctypes.c_int
    E.emu_pack_planes.argtypes = [P, P, I, I, I, P, P]
    E.emu_cross_hist.argtypes = [P, P, I, I, P, P, P, I, I, I, I, P, I, ctypes.c_int64, P, I]
    E.emu_self_hist.argtypes = [P, P, P, I, I, I, I, P, P]
    E.emu_sweep.argtypes = [I, P, P, I, I, I, P, P, P, I, I, I, I, I, P, P, P, P, P, P, P]
    return E


def _p(a):
    return None if a is None else a.ctypes.data


def _planes(E, codes, lens):
    n, lmax = codes.shape
    nw = lmax // 32 + 4
    pl = np.zeros((n, nw, 2), np.uint32)
    rc = np.zeros((n, nw, 2), np.uint32)
    E.emu_pack_planes(_p(np.ascontiguousarray(codes)), _p(lens.astype(np.int32)), n, lmax, nw,
                      _p(pl), _p(rc))
    return pl, rc, nw


def _wtab(kern, lmax):
    """Weight by distance from the centre, d = 0..lmax//2, read off the oracle's
    per-length weights of a (2n - 1)-window sequence (centre n - 1)."""
    if not kern.weighted:
        return None
    n = lmax // 2 + 1
    return np.ascontiguousarray(O.window_weights(2 * n - 1, kern.M, kern.H)[n - 1:])


@pytest.mark.parametrize("seed", range(16))
def test_emulated_lane_histograms_bit_exact(seed):
    """The bit-sliced diagonal walk with the cross kernel's 64-SV groups and
    diagonal splits, and the self-kernel walk, equal the oracle's brute force;
    every plane-word index stays inside its row."""
    E = _emu()
    rng = np.random.default_rng(seed)
    l = int(rng.choice([3, 5, 8, 10, 11, 12, 13]))
    k = max(1, l - int(rng.integers(2, 5)))
    kt = int(rng.choice([0, 2, 4]))
    p = {"L": l, "k": k, "include_rc": bool(rng.integers(0, 2))}
    if kt == 4:
        p["M"], p["H"] = int(rng.integers(1, 60)), float(rng.uniform(5, 60))
    kern = O.Kern.make(kt, p)
    B, S = int(rng.integers(1, 5)), int(rng.integers(1, 140))
    lx, ly = rng.integers(1, 140, B), rng.integers(1, 200, S)
    xc = np.zeros((B, lx.max()), np.uint8)
    yc = np.zeros((S, ly.max()), np.uint8)
    for i in range(B):
        xc[i, : lx[i]] = rng.integers(0, 2 if i % 2 else 4, lx[i])
    for i in range(S):
        yc[i, : ly[i]] = rng.integers(0, 2 if i % 3 == 0 else 4, ly[i])
    if xc.shape[1] < l or yc.shape[1] < l:
        pytest.skip("no windows")
    want = O.pair_hist(kern, O.pack_both(kern, xc, lx), O.pack_both(kern, yc, ly))
    xpl, _, nwx = _planes(E, xc, lx)
    ypl, yrc, nwy = _planes(E, yc, ly)
    nm = kern.dmax + 1
    wt = _wtab(kern, max(xc.shape[1], yc.shape[1]))
    for nsplit in (1, 3):
        for generic in (0, 1):
            h = np.zeros((B, S, nm), np.uint64)
            rc = E.emu_cross_hist(_p(xpl), _p(lx.astype(np.int32)), B, nwx, _p(ypl),
                                  _p(yrc) if kern.rc else None, _p(ly.astype(np.int32)), S, nwy,
                                  l, nm, _p(wt), nsplit, -1, _p(h), generic)
            assert rc == 0
            np.testing.assert_array_equal(h, want)
    _, xrc, _ = _planes(E, xc, lx)
    hs = np.zeros((B, nm), np.uint64)
    assert E.emu_self_hist(_p(xpl), _p(xrc) if kern.rc else None, _p(lx.astype(np.int32)), B,
                           nwx, l, nm, _p(wt), _p(hs)) == 0
    np.testing.assert_array_equal(hs, O.self_hist(kern, O.pack_both(kern, xc, lx)))


@pytest.mark.parametrize("mf", ["lsgkm_wgkm_t0.model.txt.gz", "lsgkm_wgkm_t4.model.txt.gz",
                                "lsgkm_mode5_rbf.model.txt.gz", "lsgkm_ragged_t5.model.txt.gz"])
def test_emulated_sweep_explain_and_ism_deltas(mf):
    """The explain / ISM sweep lane code: attributions equal the oracle's
    per-SV loop, and the per-(position, base) kernel deltas equal rescoring
    each mutant from scratch."""
    E = _emu()
    m = O.load_lsgkm(H.path(mf))
    kern = m.kern
    seqs = [s for _, s in O.read_fasta(H.path("lsgkm_ragged_query.fa"))][:4]
    codes, lens = O.batch(seqs)
    B, lmax = codes.shape
    S = len(m.coef)
    xpl, _, nwx = _planes(E, codes, lens)
    ypl, yrc, nwy = _planes(E, m.sv_codes, m.sv_lengths)
    wt = _wtab(kern, max(lmax, m.sv_codes.shape[1]))
    kr, dx, _ = O.raw(m, codes, lens)
    z = np.maximum(np.sqrt(dx[:, None] * m.sv_diag[None, :]), 1e-10)
    ec = m.coef[None, :] / z
    if kern.gamma is not None:
        ec = ec * O.rbf_scale(kr / z, kern.gamma)
    ec = np.ascontiguousarray(ec)
    for mode in ("hypothetical", "perturbation"):
        ta, tb, mm = O.explain_tables(kern, mode == "perturbation")
        cmax = kern.l - mm
        res = np.zeros((B, 4, lmax))
        assert E.emu_sweep(int(mode == "perturbation"), _p(xpl), _p(lens.astype(np.int32)), lmax,
                           B, nwx, _p(ypl), _p(yrc) if kern.rc else None,
                           _p(m.sv_lengths.astype(np.int32)), S, nwy, kern.l, cmax + 1, cmax,
                           _p(wt), _p(ta), _p(tb), _p(ec), _p(res), None, None) == 0
        want = O.explain(m, codes, lens, mode)
        H.assert_close(np.where(np.isnan(want), np.nan, res), want, 1e-12, mode)
    nm = kern.dmax + 2
    tab = np.zeros(kern.l + 2)
    tab[: kern.l + 1] = kern.table
    dk = np.zeros((B, S, 5, lmax))
    hist = np.zeros((B, S, nm), np.uint64)
    assert E.emu_sweep(2, _p(xpl), _p(lens.astype(np.int32)), lmax, B, nwx, _p(ypl),
                       _p(yrc) if kern.rc else None, _p(m.sv_lengths.astype(np.int32)), S, nwy,
                       kern.l, nm, nm - 1, _p(wt), _p(tab), _p(tab), None, None, _p(dk),
                       _p(hist)) == 0
    np.testing.assert_allclose((hist.astype(float) * tab[:nm]).sum(-1), kr, rtol=1e-13)
    for b in range(B):
        for pos in range(0, lens[b], 5):
            for v in range(4):
                if v == codes[b, pos]:
                    continue
                mc = codes[b:b + 1].copy()
                mc[0, pos] = v
                km, _, _ = O.raw(m, mc, lens[b:b + 1])
                got = dk[b, :, 0, pos] + dk[b, :, 1 + v, pos]
                np.testing.assert_allclose(got, km[0] - kr[b], rtol=0,
                                           atol=1e-12 * np.abs(kr[b]).max())


def test_fixed_point_sv_weights():
    """wfix = round(coef / sqrt(K(s, s)) 2^shift): the largest at 2^61 or just
    under, 0 for SVs without windows, each within half a unit of its exact
    value (checked with Python's exact rationals)."""
    from fractions import Fraction

    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(3)
    coef = rng.normal(size=200) * 10.0 ** rng.integers(-6, 3, 200)
    ds = rng.uniform(1, 1e4, 200)
    ds[7] = 0.0
    wfix, shift = E.fixed_weights(coef, ds, "cpu")
    wfix = wfix.numpy()
    w = np.where(ds > 0, coef / np.sqrt(np.where(ds > 0, ds, 1)), 0.0)
    assert wfix[7] == 0 and np.abs(wfix).max() <= 2**61
    assert np.abs(wfix).max() > 2**60
    for i in range(200):
        assert abs(Fraction(int(wfix[i])) - Fraction(w[i]) * 2**shift) <= Fraction(1, 2)
    wf2, s2 = E.fixed_weights(coef, None, "cpu")  # no normalisation: the coefficients
    assert abs(int(wf2.numpy()[0]) - coef[0] * 2.0**s2) <= 0.5
    z, s0 = E.fixed_weights(np.zeros(3), None, "cpu")
    assert s0 == 0 and not z.numpy().any()
