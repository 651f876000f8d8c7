"""gkm_index_hist_noalloc (partials in the library's static scratch) against
gkm_index_hist with a caller workspace, through the C-ABI: identical
histograms on two calls in a row, the second smaller, for the 8-bit, 16-bit
and weighted count formats; a call whose partials exceed the scratch is
refused before anything runs."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _model(kt, S, seed, extra=None):
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(seed)
    lens = rng.integers(150, 501, S)
    codes = np.zeros((S, 500), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    codes[:50, :80] = 0  # poly-A SVs: with the query's poly-A run, counts past 255
    p = {"L": 11, "k": 7, "d": 3, **(extra or {})}
    m = GkmSVM(codes, rng.normal(size=S), 0.1, kt, p, sv_lengths=lens)
    return m


def _queries(kp, n, lo, hi, seed):
    from gkmsvm_amd import engine as E

    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, n)
    codes = np.zeros((n, hi), np.uint8)
    for i in range(n):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    codes[0, :60] = 0  # a poly-A run: counts past 255 in the 8-bit format
    return E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)


def _both(kp, idx, xs, S):
    """(histograms through the workspace entry point, through the no-alloc one)."""
    from gkmsvm_amd import _lib
    from gkmsvm_amd import engine as E

    L = E.IndexLaunch(kp, idx, xs, 0, xs.n, S, -1, None, None)
    st = E.stream_handle()
    want = torch.zeros((xs.n, S, kp.nm), dtype=torch.int64, device=kp.device)
    partial = torch.empty(L.partial_bytes(kp, idx, S), dtype=torch.uint8, device=kp.device)
    _lib.call("gkm_index_hist", *L.args, partial.data_ptr(), want.data_ptr(), st)
    got = torch.zeros_like(want)
    assert L.partial_bytes(kp, idx, S) <= _lib.load().gkm_scratch_bytes()
    _lib.call("gkm_index_hist_noalloc", *L.args, got.data_ptr(), st)
    torch.cuda.synchronize()
    return want.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("kt,extra,bits", [(2, None, 8), (2, None, 16),
                                           (4, {"M": 50, "H": 50.0}, 16)])
def test_noalloc_index_hist_matches_workspace_entry_point(gpu, kt, extra, bits):
    from gkmsvm_amd import engine as E

    m = _model(kt, 3000, 1)
    kp = m._params()
    sv = m._svs()
    idx = E.DeviceIndex(kp, sv.seqs, count_bits=bits)
    for n, lo, hi, seed in ((6, 200, 900, 2), (2, 100, 400, 3)):  # the second call smaller
        xs = _queries(kp, n, lo, hi, seed)
        want, got = _both(kp, idx, xs, sv.S)
        assert want.sum() > 0
        np.testing.assert_array_equal(got, want, err_msg=f"{n} rows, {bits}-bit counts")


def test_noalloc_refuses_partials_beyond_the_scratch(gpu):
    """Too many rows for the scratch: status -2 and a message naming the
    limit, raised before any launch; the entry point still works after."""
    from gkmsvm_amd import _lib
    from gkmsvm_amd import engine as E

    m = _model(2, 3000, 4)
    kp = m._params()
    sv = m._svs()
    idx = sv.index()
    xs = _queries(kp, 4000, 900, 1000, 5)
    L = E.IndexLaunch(kp, idx, xs, 0, xs.n, sv.S, -1, None, None)
    assert L.partial_bytes(kp, idx, sv.S) > _lib.load().gkm_scratch_bytes()
    assert _lib.load().gkm_index_partial_bytes(
        L.B, L.Wx, sv.S, kp.nm, idx.n_tiles, L.cw, 0, L.pack_code) == L.partial_bytes(
        kp, idx, sv.S)
    hist = torch.zeros((xs.n, sv.S, kp.nm), dtype=torch.int64, device=kp.device)
    with pytest.raises(_lib.GkmError, match="scratch"):
        _lib.call("gkm_index_hist_noalloc", *L.args, hist.data_ptr(), E.stream_handle())
    assert int(hist.abs().sum()) == 0
    xs2 = _queries(kp, 3, 200, 600, 6)
    want, got = _both(kp, idx, xs2, sv.S)
    np.testing.assert_array_equal(got, want)
