"""The two forms of GkmExplain through the SVs' l-mer weights V
(explain/lmer.py:31-188, _explain_lmer with V): gkm_lmer_explain (windows in
position order, a block's attributions in LDS) and gkm_lmer_explain_sorted
(the rows' windows sorted by l-mer, each window's attributions added to the
output with FP64 atomics) give the same attributions up to the order of FP64
additions, in both explain modes, with and without normalisation, on ragged
rows (some shorter than l) and a repeat-rich row."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pert", [False, True])
def test_sorted_and_position_order_lmer_explain_agree(gpu, pert):
    from gkmsvm_amd import GkmSVM, _lib
    from gkmsvm_amd import engine as E
    from gkmsvm_amd import explain as X
    from gkmsvm_amd import tables as T

    rng = np.random.default_rng(31 + pert)
    S = 400
    svl = rng.integers(80, 301, S)
    svc = np.zeros((S, 300), np.uint8)
    for i, n in enumerate(svl):
        svc[i, :n] = rng.integers(0, 4, n)
    m = GkmSVM(svc, rng.normal(size=S), 0.1, 2, {"L": 11, "k": 7, "d": 3}, sv_lengths=svl)
    m.use_lmer_table = m.use_explain_table = False
    m.rocm()
    sv, kp = m._svs(), m._params()
    V = T.lmer_weights(sv)
    ta, tb, cmax = X.explain_tables(kp.base, pert)
    ta_d, tb_d = X._explain_ab(sv, ta, tb, pert)
    masks, levels = X._explain_masks(sv, ta, tb, cmax, pert)
    lens = np.array([400, 9, 250, 11, 333, 120])
    codes = np.zeros((len(lens), 400), np.uint8)
    for i, n in enumerate(lens):
        codes[i, :n] = rng.integers(0, 4, n)
    codes[4, 50:200] = 0  # a poly-A stretch: repeated windows, many equal keys
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    dx = E.self_kraw(kp, xs)
    fwd, _ = xs.windows(kp.l)
    B, Wx = fwd.shape
    outs = []
    for sorted_form in (False, True):
        out = torch.zeros((B, 4, xs.lmax), dtype=torch.float64, device=kp.device)
        if sorted_form:
            flat = fwd.reshape(-1)
            keys, order = torch.sort(torch.where(flat < 0, torch.full_like(flat, 2**31 - 1), flat))
            _lib.call("gkm_lmer_explain_sorted", int(pert), E._p(keys),
                      E._p(order.to(torch.int32)), flat.numel(), E._p(xs.len), B, Wx, xs.lmax,
                      E._p(V), E._p(dx), 1, E._p(masks), E._p(levels), len(masks), kp.l, None, 4,
                      E._p(ta_d), E._p(tb_d), E._p(out), None)
        else:
            _lib.call("gkm_lmer_explain", int(pert), E._p(fwd), E._p(xs.len), B, Wx, xs.lmax,
                      E._p(V), E._p(dx), 1, E._p(masks), E._p(levels), len(masks), kp.l, None, 16,
                      E._p(ta_d), E._p(tb_d), E._p(out), None)
        outs.append(out.cpu().numpy())
    a, b = outs
    assert np.abs(a).max() > 0
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-12 * np.abs(a).max())
    assert (a[1] == 0).all() and (b[1] == 0).all()  # a row shorter than l has no window
