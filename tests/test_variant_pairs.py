"""Allele pairs scored through their reference (host side and the identity
behind it, on the CPU): an alternate allele of the reference's length that
differs in bases lo..hi shares every window outside w0 = lo - l + 1 .. w1 = hi
with the reference, and a pair histogram is a sum over the query's windows,
so N_alt = N_ref + N(alt cut to those windows) - N(ref cut there) exactly for
every SV -- the oracle's brute-force histograms (its C restatement of
kernels/direct.py:23-42) must satisfy it, and a cut one base short must not."""

import numpy as np

from gkmsvm_amd import engine as E
from oracle import oracle as O


def _brute_spans(ref, alt):
    lo, hi = [], []
    for r, a in zip(ref, alt):
        d = [i for i in range(len(r)) if len(r) == len(a) and r[i] != a[i]]
        lo.append(d[0] if d else -1)
        hi.append(d[-1] if d else -1)
    return np.array(lo), np.array(hi)


def test_variant_cut_of_strings_equals_a_per_base_comparison():
    """E.variant_cut (strings, for the server) against the differing bases
    found one by one and E.variant_cuts (code arrays, for score_variants)."""
    rng = np.random.default_rng(0)
    ref, alt = [], []
    for t in range(600):
        L = int(rng.integers(1, 60))
        r = "".join(rng.choice(list("ACGTNacgt"), L))
        a = list(r)
        kind = t % 6
        if kind == 0:  # SNV anywhere, first and last base included
            p = int(rng.integers(0, L))
            a[p] = "T" if a[p] != "T" else "G"
        elif kind == 1:  # several substitutions
            for p in rng.integers(0, L, int(rng.integers(1, 4))):
                a[int(p)] = "acgtN"[int(rng.integers(0, 5))]
        elif kind == 2:  # an indel: lengths differ
            a = a[:-1] if L > 1 else a + ["A"]
        elif kind == 3:  # case only
            p = int(rng.integers(0, L))
            a[p] = a[p].swapcase()
        ref.append(r)
        alt.append("".join(a))  # kind 4, 5: identical
    lo, hi = _brute_spans(ref, alt)
    for l in (3, 7, 11):
        lens = np.array([len(r) for r in ref])
        ok, c0, c1 = E.variant_cuts(lens, lo, hi, l)
        for i, (r, a) in enumerate(zip(ref, alt)):
            got = E.variant_cut(r, a, l)
            want = (int(c0[i]), int(c1[i])) if ok[i] else None
            assert got == want, (l, r, a, got, want)
    assert sum(E.variant_cut(r, a, 7) is not None for r, a in zip(ref, alt)) > 100


def test_variant_cuts_hold_every_changed_window():
    rng = np.random.default_rng(1)
    for _ in range(3000):
        l = int(rng.integers(3, 12))
        L = int(rng.integers(1, 80))
        lo = int(rng.integers(-1, L))
        hi = lo if lo < 0 else int(rng.integers(lo, min(L, lo + 30)))
        ok, c0, c1 = E.variant_cuts(np.array([L]), np.array([lo]), np.array([hi]), l)
        ok, c0, c1 = bool(ok[0]), int(c0[0]), int(c1[0])
        W = L - l + 1
        changed = [w for w in range(max(W, 0)) if lo >= 0 and w <= hi and w + l - 1 >= lo]
        if ok:
            # the cut's windows are exactly the changed ones
            assert changed == list(range(c0, c1 - l + 1)), (l, L, lo, hi, c0, c1)
            assert 0 <= c0 and c1 <= L
            assert len(changed) < E.VARIANT_ALIAS_FRACTION * W
        else:
            assert not changed or len(changed) >= E.VARIANT_ALIAS_FRACTION * W


def _hist(kern, codes_list):
    L = max(len(c) for c in codes_list)
    codes = np.zeros((len(codes_list), L), np.uint8)
    for i, c in enumerate(codes_list):
        codes[i, :len(c)] = c
    lens = np.array([len(c) for c in codes_list])
    return codes, lens


def test_alternate_level_sums_from_the_reference_and_two_cut_rows():
    rng = np.random.default_rng(2)
    kern = O.Kern.make("gkm", {"L": 7, "k": 5, "d": 2})
    l = kern.l
    svs = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(20, 60, 25)]
    sc, sl = _hist(kern, svs)
    psv = O.pack_both(kern, sc, sl)
    refs, alts = [], []
    for t in range(40):
        L = int(rng.integers(30, 70))
        r = rng.integers(0, 4, L).astype(np.uint8)
        a = r.copy()
        p = [0, L - 1, int(rng.integers(0, L))][t % 3]
        n = 1 + (t % 4 == 0) * int(rng.integers(1, 4))
        a[p:p + n] = (a[p:p + n] + 1 + rng.integers(0, 3, len(a[p:p + n]))) % 4
        refs.append(r)
        alts.append(a)
    lens = np.array([len(r) for r in refs])
    diff = [np.flatnonzero(r != a) for r, a in zip(refs, alts)]
    lo = np.array([d[0] for d in diff])
    hi = np.array([d[-1] for d in diff])
    ok, c0, c1 = E.variant_cuts(lens, lo, hi, l)
    assert ok.sum() > 30

    def hist(rows):
        codes, ls = _hist(kern, rows)
        return O.pair_hist(kern, O.pack_both(kern, codes, ls), psv).astype(np.int64)

    i = np.flatnonzero(ok)
    h_ref = hist([refs[k] for k in i])
    h_alt = hist([alts[k] for k in i])
    h_plus = hist([alts[k][c0[k]:c1[k]] for k in i])
    h_minus = hist([refs[k][c0[k]:c1[k]] for k in i])
    np.testing.assert_array_equal(h_ref + h_plus - h_minus, h_alt)
    # a cut that leaves out the last changed window breaks it
    short_plus = hist([alts[k][c0[k]:c1[k] - 1] for k in i])
    short_minus = hist([refs[k][c0[k]:c1[k] - 1] for k in i])
    assert not np.array_equal(h_ref + short_plus - short_minus, h_alt)
