"""In silico mutagenesis on the GPU (gkmsvm-lite ism.py:21-320).

For every position p and base v, the score change f(x with x_p = v) - f(x).
Three paths: the l-mer table (kernels without RBF), the l-mer index (each
SV's K_raw(mutant, s) = K_raw(x, s) - K_raw(old windows, s) + K_raw(new
windows, s), the windows covering p probed as they are and substituted), and
the sweep.  The kernel change of a substitution touches only the windows covering p: on a
diagonal of the (x, SV) pair, a window with c mismatches gains one
(h[c+1] - h[c]) where x_p matched the SV, and loses one (h[c-1] - h[c]) for the
base v equal to the SV's where it did not.  One sweep per (x, SV) pair
gathers those deltas for every (p, v) in LDS; the mutant self-kernels come
from gkm_mutant_diag; the normalisation, RBF and coefficient sum follow per SV.
"""

from __future__ import annotations

import torch

from . import _lib
from . import engine as E
from .inputs import mask_positions, read_batch


def _sv_chunk(B: int, S: int) -> int:
    spb = max(1, (B * S) // 8192)
    return max(spb, -(-S // 65535))


def ism(model, x, *, lengths=None, batch_size: int | None = None,
        verbose: bool = False, use_index=None) -> torch.Tensor:
    """[B, 4, L] float64 score deltas for every single-base substitution (0 for
    the reference base and past each sequence's length, NaN rows for sequences
    shorter than l).  Without the l-mer table the deltas come from the l-mer
    index (use_index True, or "auto" -- the model's use_lmer_index by default
    -- where the cost model prefers it) or from the sweep kernel."""
    dev = model._require_device()
    batch = read_batch(x, lengths)
    kp = model._params()
    short = batch.short(kp.l, "ISM deltas")
    codes, lens = batch.codes, batch.lengths
    B, L = codes.shape
    out = torch.zeros((B, 4, L), dtype=torch.float64, device=dev)
    if L >= kp.l and kp.dmax >= 0:
        # without the table each (position, sequence) queries the l windows
        # covering it before and after each substitution (ism.py:49-52)
        model._plan_lmer_table(5.0 * kp.l * B * L)
        table = model._lmer_table_for()
        bs = B if batch_size is None else max(1, int(batch_size))
        bs = min(bs, 65535)
        starts = range(0, B, bs)
        if verbose:
            from tqdm import tqdm

            starts = tqdm(starts, desc="ISM")
        mode = model.use_lmer_index if use_index is None else use_index
        for s in starts:
            e = min(B, s + bs)
            if table is not None:
                _ism_rows_table(model, codes[s:e], lens[s:e], table, out[s:e])
            elif E.ism_use_index(model._svs(), lens[s:e], mode):
                _ism_rows_index(model, codes[s:e], lens[s:e], out[s:e])
            else:
                _ism_rows(model, codes[s:e], lens[s:e], out[s:e])
    return mask_positions(out, batch, short)


def _mutant_diag(kp, xs, dx):
    """K_raw of every single-base mutant [4, lmax, B] (gkm_mutant_diag)."""
    fwd, rcw = xs.windows(kp.l)
    dmut = torch.empty((4, xs.lmax, xs.n), dtype=torch.float64, device=kp.device)
    wt, _ = kp.wtab(xs.lmax)
    _lib.call("gkm_mutant_diag", E._p(fwd), E._p(rcw) if kp.rc else None, E._p(xs.len),
              E._p(xs.codes), xs.lmax, xs.n, kp.l, E._p(kp.tab), E._p(wt), E._p(dx),
              E._p(dmut), E.stream_handle())
    return dmut


def _ism_rows_table(model, codes, lens, table, out):
    """ISM from the l-mer table (ism.py:118-144): per position the table
    weights of the l windows covering it, before and after each substitution."""
    kp = model._params()
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    B = xs.n
    raw = torch.empty(B, dtype=torch.float64, device=kp.device)
    E.table_rows(kp, xs, table, 0.0, raw, normalize=False)
    dx = dmut = None
    if kp.normalize:
        dx = E.self_kraw(kp, xs)
        dmut = _mutant_diag(kp, xs, dx)
    fwd, _ = xs.windows(kp.l)
    wt, _ = kp.wtab(xs.lmax)
    _lib.call("gkm_table_ism", E._p(fwd), E._p(xs.len), E._p(xs.codes), B, fwd.shape[1],
              xs.lmax, kp.l, E._p(wt), E._p(table), E._p(raw), E._p(dx), E._p(dmut),
              int(kp.normalize), E._p(out), E.stream_handle())


def _ism_rows_index(model, codes, lens, out):
    """ISM through the l-mer index (kernels/lmer_index.py:326-376):
    gkm_index_ism per query (its windows probed one by one, then the windows
    covering each position with each substitution, in two persistent
    launches)."""
    kp = model._params()
    sv = model._svs()
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    dx = dmut = None
    if kp.normalize:
        dx = E.self_kraw(kp, xs)
        dmut = _mutant_diag(kp, xs, dx)
    for b in range(xs.n):
        if lens[b] >= kp.l:
            E.index_ism(sv, xs, b, dx, dmut, out[b])


def _ism_rows(model, codes, lens, out):
    kp = model._params()
    sv = model._svs()
    st = E.stream_handle()
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    B = xs.n
    if kp.normalize:
        dx = E.self_kraw(kp, xs)
        dmut = _mutant_diag(kp, xs, dx)
    else:
        dx = torch.ones(B, dtype=torch.float64, device=kp.device)
        dmut = torch.ones((4, xs.lmax, B), dtype=torch.float64, device=kp.device)
    nm = min(kp.dmax + 2, kp.l + 1)
    wt, wn = kp.wtab(max(xs.lmax, sv.seqs.lmax))
    buf = torch.zeros((B, 4, xs.lmax), dtype=torch.float64, device=kp.device)
    _lib.call("gkm_ism", E._p(xs.planes), E._p(xs.len), E._p(xs.codes), xs.lmax, B, xs.nw,
              E._p(sv.seqs.planes), E._p(sv.seqs.planes_rc) if kp.rc else None,
              E._p(sv.seqs.len), sv.S, sv.seqs.nw, kp.l, nm, E._p(wt), wn, _sv_chunk(B, sv.S),
              E._p(kp.tab), E._p(sv.coef), E._p(sv.ds), E._p(dx), E._p(dmut), int(kp.normalize),
              int(kp.rbf), kp.gamma, E._p(buf), st)
    out += buf
