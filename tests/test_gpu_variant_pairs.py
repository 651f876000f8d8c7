"""Allele pairs through their reference's level sums on the GPU
(gkm_index_score_queue_pairs): score_variants, the ABI and the server give
the same bits as scoring both alleles as ordinary rows (svm.py:632-647 scores
score(alt) - score(ref) that way) -- SNVs at the first, middle and last base,
multi-base changes, identical alleles, indels, N and lowercase, short and
long changes (the latter scored as ordinary rows)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _model(S=600, seed=0, kt=2, **extra):
    from gkmsvm_amd import GkmSVM

    rng = np.random.default_rng(seed)
    lens = rng.integers(150, 401, S)
    codes = np.zeros((S, 400), np.uint8)
    for i in range(S):
        codes[i, : lens[i]] = rng.integers(0, 4, lens[i])
    m = GkmSVM(codes, rng.normal(size=S), 0.21, kt, {"L": 11, "k": 7, "d": 3, **extra},
               sv_lengths=lens)
    m.use_lmer_table = m.use_explain_table = False
    m.use_lmer_index = True
    return m


def _pairs(seed):
    """(ref codes, alt codes, ref lengths, alt lengths) of mixed variants."""
    rng = np.random.default_rng(seed)
    refs, alts = [], []
    for t in range(60):
        L = int(rng.choice([300, 120, 37, 11, 8]))
        r = rng.integers(0, 4, L).astype(np.uint8)
        a = r.copy()
        kind = t % 8
        if kind in (0, 1, 2):  # SNV at the first, a middle and the last base
            p = [0, L // 2, L - 1][kind]
            a[p] = (a[p] + 1) % 4
        elif kind == 3:  # a short block substitution
            p = int(rng.integers(0, max(1, L - 5)))
            a[p:p + 5] = (a[p:p + 5] + 2) % 4
        elif kind == 4:  # changes far apart: ordinary rows
            a[0] = (a[0] + 1) % 4
            a[-1] = (a[-1] + 1) % 4
        elif kind == 5:  # a deletion
            a = np.delete(a, L // 2)
        # kinds 6, 7: identical alleles
        refs.append(r)
        alts.append(a)
    W = max(max(len(r) for r in refs), max(len(a) for a in alts))
    rc = np.zeros((len(refs), W), np.uint8)
    ac = np.zeros((len(refs), W), np.uint8)
    for i, (r, a) in enumerate(zip(refs, alts)):
        rc[i, : len(r)] = r
        ac[i, : len(a)] = a
    return rc, ac, np.array([len(r) for r in refs]), np.array([len(a) for a in alts])


def _direct(m, rc, ac, rl, al):
    """score(alt) - score(ref), each allele an ordinary row of model(x)."""
    with np.errstate(invalid="ignore"):
        import warnings

        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            a = m(ac, lengths=al).cpu().numpy()
            r = m(rc, lengths=rl).cpu().numpy()
    return a - r


def test_score_variants_pairs_equal_direct_scoring(gpu):
    from gkmsvm_amd import engine as E

    m = _model()
    rc, ac, rl, al = _pairs(1)
    same = rl == al
    diff = (rc != ac) & (np.arange(rc.shape[1])[None, :] < rl[:, None]) & same[:, None]
    lo = np.where(diff.any(1), diff.argmax(1), -1)
    hi = np.where(diff.any(1), rc.shape[1] - 1 - diff[:, ::-1].argmax(1), -1)
    ok, _, _ = E.variant_cuts(rl, lo, hi, 11)
    assert ok.sum() >= 10 and (~ok).sum() >= 10  # both kinds of rows in one call
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        got = m.score_variants(rc, ac, ref_lengths=rl, alt_lengths=al).cpu().numpy()
    want = _direct(m, rc, ac, rl, al)
    np.testing.assert_array_equal(got, want)
    assert np.isnan(got[rl < 11]).all() and not np.isnan(got[(rl >= 11) & (al >= 11)]).any()


def test_index_pairs_abi_equals_probing_every_row(gpu):
    """gkm_index_score_queue_pairs against gkm_index_score_queue on the same
    rows: the alternates (last block) and every other row, bit for bit."""
    from gkmsvm_amd import engine as E

    m = _model(seed=4)
    sv, kp = m._svs(), m._params()
    rng = np.random.default_rng(5)
    P, L = 9, 260
    ref = rng.integers(0, 4, (P, L)).astype(np.uint8)
    alt = ref.copy()
    pos = rng.integers(0, L - 3, P)
    for i, p in enumerate(pos):
        alt[i, p:p + 1 + i % 3] = (alt[i, p:p + 1 + i % 3] + 1) % 4
    lo = pos
    hi = pos + np.arange(P) % 3
    ok, c0, c1 = E.variant_cuts(np.full(P, L), lo, hi, kp.l)
    assert ok.all()
    other = rng.integers(0, 4, (5, L)).astype(np.uint8)
    olen = np.array([L, 200, 90, 11, 5])
    rows = [other[i] for i in range(5)] + [ref[i] for i in range(P)]
    rows += [alt[i, c0[i]:c1[i]] for i in range(P)] + [ref[i, c0[i]:c1[i]] for i in range(P)]
    rows += [alt[i] for i in range(P)]
    lens = np.array(list(olen) + [L] * P + list(c1 - c0) * 2 + [L] * P)
    codes = np.zeros((len(rows), L), np.uint8)
    for i, r in enumerate(rows):
        codes[i, : lens[i]] = r[: lens[i]]
    xs = E.DeviceSeqs(codes, lens, kp.device, rc=kp.rc)
    dx = E.self_kraw(kp, xs)
    got = torch.empty(len(rows), dtype=torch.float64, device=kp.device)
    want = torch.empty_like(got)
    E.index_score_queue(sv, xs, dx, m.bias, got, pairs=P)
    E.index_score_queue(sv, xs, dx, m.bias, want)
    np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())
    with pytest.raises(Exception, match="allele pairs need 4 rows"):
        E.index_score_queue(sv, xs, dx, m.bias, got, pairs=len(rows))


def test_server_variant_pairs_equal_direct_scoring(gpu):
    """String requests through the server: SNVs, blocks, identical alleles,
    an indel, N and case changes inside the variant, and an allele with a
    letter other than ACGTN (that request fails alone)."""
    from gkmsvm_amd.serving import GkmServer

    m = _model(seed=6)
    rng = np.random.default_rng(7)

    def s(n):
        return "".join("ACGT"[c] for c in rng.integers(0, 4, n))

    reqs = []
    for t in range(24):
        r = s(300)
        k = t % 6
        if k == 0:
            a = r[:150] + ("A" if r[150] != "A" else "C") + r[151:]
        elif k == 1:
            a = r[:10] + r[10:14].lower() + "N" + r[15:]
        elif k == 2:
            a = ("G" if r[0] != "G" else "T") + r[1:]
        elif k == 3:
            a = r[:-1] + ("A" if r[-1] != "A" else "T")
        elif k == 4:
            a = r[:100] + r[101:]  # deletion
        else:
            a = r
        reqs.append([(r, a), (s(40), s(40))])
    bad = [(s(300), "X" + s(299))]
    with GkmServer(m, hbm_budget=2 << 30, tables=False) as srv:
        futs = [srv.submit("variant", q) for q in reqs]
        fbad = srv.submit("variant", bad)
        got = [f.result() for f in futs]
        with pytest.raises(ValueError):
            fbad.result()
    ref = [p[0] for q in reqs for p in q]
    alt = [p[1] for q in reqs for p in q]
    want = (m(alt).cpu().numpy() - m(ref).cpu().numpy()).ravel()
    np.testing.assert_array_equal(np.concatenate(got), want)
