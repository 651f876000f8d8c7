// gkm_core.h -- per-lane primitives of the gapped k-mer engine, shared by the
// HIP kernels (gkm_kernels.hip) and the CPU emulation of their lane logic
// (gkm_emu.cpp).
//
// Sequences are held as two bit-planes, one bit per base: plane 0 holds bit 0
// of each base code (A=0 C=1 G=2 T=3), plane 1 holds bit 1; base t sits in bit
// (t & 31) of word (t >> 5) + 1 (word 0 is a zero front pad).  Two windows
// starting at positions i of x and j of y lie on diagonal delta = j - i.  For a
// diagonal, the mismatch bit string e[t] = [x[t - delta] != y[t]] over y
// positions t is
//     e = (x0 ^ y0) | (x1 ^ y1)     (32 positions per word, x funnel-shifted)
// and the mismatch count of the window pair starting at y position t is the
// sum of e[t .. t+l-1].  gkmsvm-lite computes the same count window pair by
// window pair as popcount(((a ^ b) >> 1 | (a ^ b)) & 0x55555555) on 2-bit
// packed l-mers (kernels/primitives.py:29-32); here the l shifted copies of e
// are added bit-sliced (a carry-save tree), so one 32-bit word carries the
// counts of 32 window pairs at once, and the per-level masks are popcounted
// into the pair's mismatch histogram N[m].
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GKM_HD __host__ __device__ __forceinline__
#else
#define GKM_HD inline
#endif

namespace gkm {

// Low 32 bits of ({hi, lo} >> sh), sh in 0..31.
GKM_HD uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t sh)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
    return (uint32_t)((((uint64_t)hi << 32) | (uint64_t)lo) >> (sh & 31u));
#endif
}

GKM_HD uint32_t popc(uint32_t v)
{
    return (uint32_t)__builtin_popcount(v);
}

// Majority of (a, b, c) given t = a ^ b: (t & c) | (~t & a).
GKM_HD uint32_t maj_t(uint32_t t, uint32_t c, uint32_t a)
{
    return (t & c) | (~t & a);
}

// Carry-save compression of N equal-weight bit-planes v[0..N) into v[0]; the
// carries (one weight up) are appended to cv[nc..].  A full adder takes 3
// operations (two XOR and a bit-select for the majority).
template <int N>
struct Compress {
    GKM_HD static void run(uint32_t *v, uint32_t *cv, int &nc)
    {
        uint32_t a = v[N - 3], b = v[N - 2], c = v[N - 1];
        uint32_t t = a ^ b;
        v[N - 3] = t ^ c;
        cv[nc++] = maj_t(t, c, a);
        Compress<N - 2>::run(v, cv, nc);
    }
};
template <>
struct Compress<2> {
    GKM_HD static void run(uint32_t *v, uint32_t *cv, int &nc)
    {
        uint32_t a = v[0], b = v[1];
        v[0] = a ^ b;
        cv[nc++] = a & b;
    }
};
template <>
struct Compress<1> {
    GKM_HD static void run(uint32_t *, uint32_t *, int &) {}
};
template <>
struct Compress<0> {
    GKM_HD static void run(uint32_t *v, uint32_t *, int &) { v[0] = 0; }
};

// Bit-sliced window sums: for the 32 windows starting at bits 0..31 of e0
// (e1 holds the next 32 positions), c = sum_{p<L} e[t+p] as planes
// b0 + 2 b1 + 4 b2 + 8 b3 (L <= 15, so exact).
template <int L>
GKM_HD void window_counts(uint32_t e0, uint32_t e1, uint32_t &b0, uint32_t &b1,
                          uint32_t &b2, uint32_t &b3)
{
    static_assert(L >= 1 && L <= 15, "l must be 1..15");
    uint32_t w1[L];
    w1[0] = e0;
#pragma unroll
    for (int p = 1; p < L; p++)
        w1[p] = funnel(e1, e0, (uint32_t)p);
    constexpr int N2 = L / 2, N4 = N2 / 2, N8 = N4 / 2;
    uint32_t w2[N2 > 0 ? N2 : 1], w4[N4 > 0 ? N4 : 1], w8[N8 > 0 ? N8 : 1];
    int n2 = 0, n4 = 0, n8 = 0;
    Compress<L>::run(w1, w2, n2);
    Compress<N2>::run(w2, w4, n4);
    Compress<N4>::run(w4, w8, n8);
    b0 = w1[0];
    b1 = N2 > 0 ? w2[0] : 0u;
    b2 = N4 > 0 ? w4[0] : 0u;
    b3 = N8 > 0 ? w8[0] : 0u;
    (void)n2;
    (void)n4;
    (void)n8;
}

// Masks of the windows whose count equals m, for m = 0..NM-1, restricted to
// `valid`.  Counts >= NM fall out.
template <int NM>
GKM_HD void level_masks(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3,
                        uint32_t valid, uint32_t *m)
{
    if (NM <= 4) {
        uint32_t ok = valid & ~(b2 | b3);
        uint32_t t1 = ok & b1;
        uint32_t t0 = ok & ~b1;
        uint32_t m3 = t1 & b0, m1 = t0 & b0;
        m[0] = t0 ^ m1;
        if (NM > 1) m[1] = m1;
        if (NM > 2) m[2] = t1 ^ m3;
        if (NM > 3) m[3] = m3;
    } else {
#pragma unroll
        for (int v = 0; v < NM; v++) {
            uint32_t r = valid;
            r &= (v & 1) ? b0 : ~b0;
            r &= (v & 2) ? b1 : ~b1;
            r &= (v & 4) ? b2 : ~b2;
            r &= (v & 8) ? b3 : ~b3;
            m[v] = r;
        }
    }
}

// Count (0..15) of window bit j from the planes.
GKM_HD int count_at(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, int j)
{
    return (int)(((b0 >> j) & 1u) | (((b1 >> j) & 1u) << 1) |
                 (((b2 >> j) & 1u) << 2) | (((b3 >> j) & 1u) << 3));
}

// Mask of bits j in 0..31 with lo <= j <= hi (lo, hi may lie outside 0..31).
GKM_HD uint32_t range_mask(int lo, int hi)
{
    if (hi < lo || hi < 0 || lo > 31)
        return 0u;
    uint32_t up = hi >= 31 ? 0xFFFFFFFFu : ((2u << hi) - 1u);
    uint32_t dn = lo <= 0 ? 0xFFFFFFFFu : ~((1u << lo) - 1u);
    return up & dn;
}

// Window weight of window j of a sequence with nw windows, from the
// distance table wtab[d] = min(M, floor(M 2^(-d/H) + 1)) (kernels/weighted.py:
// 38-44 gives every window of a length-nw sequence the weight of its distance
// |nw/2 - j| from the centre, nw/2 rounded down).
GKM_HD uint32_t window_weight(const uint8_t *wtab, int nw, int j)
{
    int c = nw >> 1;
    int d = c > j ? c - j : j - c;
    return wtab[d];
}


// A pair of plane words (plane 0, plane 1) of 32 bases.
struct alignas(8) P2 {
    uint32_t x, y;
};

// Mismatch word of 32 y positions: x words xa (lower) and xb funnel-shifted
// right by sh against y words yw.
GKM_HD uint32_t mism_word(P2 xa, P2 xb, P2 yw, uint32_t sh)
{
    return (funnel(xb.x, xa.x, sh) ^ yw.x) | (funnel(xb.y, xa.y, sh) ^ yw.y);
}

// Window counts with the window length fixed at compile time (LT > 0) or given
// at run time (LT == 0: a 15-input tree, inputs past l zero).
template <int LT>
GKM_HD void counts(uint32_t e0, uint32_t e1, int l, uint32_t &b0, uint32_t &b1,
                   uint32_t &b2, uint32_t &b3)
{
    if constexpr (LT > 0) {
        (void)l;
        window_counts<LT>(e0, e1, b0, b1, b2, b3);
    } else {
        uint32_t w1[15];
        w1[0] = e0;
#pragma unroll
        for (int p = 1; p < 15; p++)
            w1[p] = p < l ? funnel(e1, e0, (uint32_t)p) : 0u;
        uint32_t w2[7], w4[3], w8[1];
        int n2 = 0, n4 = 0, n8 = 0;
        Compress<15>::run(w1, w2, n2);
        Compress<7>::run(w2, w4, n4);
        Compress<3>::run(w4, w8, n8);
        b0 = w1[0];
        b1 = w2[0];
        b2 = w4[0];
        b3 = w8[0];
    }
}

// Level masks for levels 0..nm-1 into m[0..NMC) (NMC = 4: nm <= 4, fast
// decode; NMC = 16: any nm).
template <int NMC>
GKM_HD void levels(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, uint32_t valid,
                   int nm, uint32_t *m)
{
    level_masks<NMC>(b0, b1, b2, b3, valid, m);
#pragma unroll
    for (int v = 0; v < NMC; v++)
        if (v >= nm)
            m[v] = 0u;
}

// Walk diagonal delta (window pairs (x window ty - delta, y window ty)) over the
// y words covering window starts ty in [ty_lo, ty_end].  xw(a), yw(a) return
// the plane words at array index a (index 0 is the front pad).  For each y word
// u the callback gets (32 u, e_prev, e_cur, y_prev, y_cur, b0, b1, b2, b3):
// the mismatch bits of positions 32u.. and 32u+32.., the y plane words there,
// and the count planes of the 32 windows starting at 32u.
//
// Index ranges (x of length lx, y of length ly, ty_end <= min(ly, lx + delta) - l):
// x indices ((32 u - delta) >> 5) + 1 .. + 3 lie in [0, lx / 32 + 3]; y indices
// u + 1, u + 2 lie in [1, ly / 32 + 2]; so nw = Lmax / 32 + 4 words suffice.
template <int LT, class XW, class YW, class F>
GKM_HD void walk_diagonal(XW xw, YW yw, int l, int delta, int ty_lo, int ty_end, F &&f)
{
    const int u0 = ty_lo >> 5, u1 = ty_end >> 5;
    const uint32_t sh = (uint32_t)(-delta) & 31u;
    int kx = ((32 * u0 - delta) >> 5) + 1;
    P2 xa = xw(kx), xb = xw(kx + 1);
    P2 yp = yw(u0 + 1);
    uint32_t e_prev = mism_word(xa, xb, yp, sh);
    for (int u = u0; u <= u1; u++) {
        xa = xb;
        xb = xw(kx + 2);
        kx++;
        P2 yc = yw(u + 2);
        uint32_t e_cur = mism_word(xa, xb, yc, sh);
        uint32_t b0, b1, b2, b3;
        counts<LT>(e_prev, e_cur, l, b0, b1, b2, b3);
        f(32 * u, e_prev, e_cur, yp, yc, b0, b1, b2, b3);
        e_prev = e_cur;
        yp = yc;
    }
}

// Bit (j + q) of the 64-bit pair (lo, hi), j + q < 64.
GKM_HD uint32_t bit2(uint32_t lo, uint32_t hi, int i)
{
    return i < 32 ? ((lo >> i) & 1u) : ((hi >> (i - 32)) & 1u);
}

// Per-position work of one hit window (x window tx, y window ty = 32u + j, count
// c) for GkmExplain and ISM: for each of its l positions, whether the bases
// differ and the y base there.  Calls g(pos_in_x, mismatch, y_base).
template <class G>
GKM_HD void hit_positions(int l, int j, int tx, uint32_t e_prev, uint32_t e_cur, P2 yp, P2 yc,
                          G &&g)
{
    for (int q = 0; q < l; q++) {
        const int i = j + q;
        const uint32_t mis = bit2(e_prev, e_cur, i);
        const int yb = (int)(bit2(yp.x, yc.x, i) | (bit2(yp.y, yc.y, i) << 1));
        g(tx + q, mis != 0u, yb);
    }
}


GKM_HD int imin(int a, int b) { return a < b ? a : b; }
GKM_HD int imax(int a, int b) { return a > b ? a : b; }

// Histogram update of one word: levels under `valid`, then counts (or, for
// window weights, sum of wx * wy per level over the hit windows).
template <int NMC, bool WEIGHTED>
GKM_HD void hist_word(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, uint32_t valid,
                      int nm, int base, int delta, int rc, int wx, int wy,
                      const uint8_t *wt, uint32_t *cnt, unsigned long long *wcnt)
{
    uint32_t mk[NMC];
    levels<NMC>(b0, b1, b2, b3, valid, nm, mk);
    if constexpr (!WEIGHTED) {
        (void)base; (void)delta; (void)rc; (void)wx; (void)wy; (void)wt; (void)wcnt;
#pragma unroll
        for (int m = 0; m < NMC; m++)
            cnt[m] += popc(mk[m]);
    } else {
        (void)cnt;
        uint32_t hits = 0;
#pragma unroll
        for (int m = 0; m < NMC; m++)
            hits |= mk[m];
        while (hits) {
            const int j = __builtin_ctz(hits);
            hits &= hits - 1u;
            const int c = count_at(b0, b1, b2, b3, j);
            const int ty = base + j, tx = ty - delta;
            const int jy = rc ? (wy - 1 - ty) : ty;
            const unsigned long long w = (unsigned long long)window_weight(wt, wx, tx) *
                                         (unsigned long long)window_weight(wt, wy, jy);
#pragma unroll
            for (int m = 0; m < NMC; m++)
                wcnt[m] += (c == m) ? w : 0ull;
        }
    }
}

// One lane of the cross kernel: query x (length lx) against one SV of a group
// (length ly; the group's lengths span ly_min..ly_max, every lane of a wave
// walks the same diagonals and words) in one orientation (rc: the SV's reverse
// complement), over diagonals [d0, d1).  Accumulates the lane's N[m].
template <int LT, int NMC, bool WEIGHTED, class XW, class YW>
GKM_HD void cross_lane(XW xw, YW yw, int lx, int ly, bool lane_on, int ly_min, int ly_max,
                       int l, int nm, int rc, int d0, int d1, const uint8_t *wt,
                       uint32_t *cnt, unsigned long long *wcnt)
{
    const int wx = lx - l + 1, wy = ly - l + 1;
    for (int delta = d0; delta < d1; delta++) {
        const int ty_lo = imax(0, delta);
        const int ty_hi_u = imin(ly_max, lx + delta) - l;
        if (ty_hi_u < ty_lo)
            continue;
        const int ty_hi_all = imin(ly_min, lx + delta) - l;
        const int ty_hi_lane = imin(ly, lx + delta) - l;
        walk_diagonal<LT>(xw, yw, l, delta, ty_lo, ty_hi_u,
                          [&](int base, uint32_t, uint32_t, P2, P2, uint32_t b0, uint32_t b1,
                              uint32_t b2, uint32_t b3) {
                              uint32_t valid;
                              if (base >= ty_lo && base + 31 <= ty_hi_all)
                                  valid = lane_on ? 0xFFFFFFFFu : 0u;
                              else
                                  valid = lane_on ? range_mask(ty_lo - base, ty_hi_lane - base) : 0u;
                              hist_word<NMC, WEIGHTED>(b0, b1, b2, b3, valid, nm, base, delta, rc,
                                                       wx, wy, wt, cnt, wcnt);
                          });
    }
}

// One diagonal of a self-kernel: x (length lx) against y = x (rc = 0) or its
// reverse complement (rc = 1).
template <int LT, int NMC, bool WEIGHTED, class XW, class YW>
GKM_HD void self_lane(XW xw, YW yw, int lx, int l, int nm, int delta, int rc,
                      const uint8_t *wt, uint32_t *cnt, unsigned long long *wcnt)
{
    const int ty_lo = imax(0, delta);
    const int ty_hi = imin(lx, lx + delta) - l;
    if (ty_hi < ty_lo)
        return;
    const int w = lx - l + 1;
    walk_diagonal<LT>(xw, yw, l, delta, ty_lo, ty_hi,
                      [&](int base, uint32_t, uint32_t, P2, P2, uint32_t b0, uint32_t b1,
                          uint32_t b2, uint32_t b3) {
                          const uint32_t valid = range_mask(ty_lo - base, ty_hi - base);
                          hist_word<NMC, WEIGHTED>(b0, b1, b2, b3, valid, nm, base, delta, rc,
                                                   w, w, wt, cnt, wcnt);
                      });
}

enum SweepMode { SW_HYP = 0, SW_PERT = 1, SW_ISM = 2 };

// One diagonal of a GkmExplain / ISM sweep of query x against one SV y (or
// its RC).  Hit windows (count <= cmax) distribute over their l positions:
//   SW_HYP:  res(yb, pos) += ew * (mismatch ? tb[c] : ta[c])          (explain/cpu.py:14-65)
//   SW_PERT: mismatch: res(yb, pos) += ew tb[c]; match: ew ta[c] to the other 3 bases
//   SW_ISM:  match: dm(pos) += w (h[c+1] - h[c]); mismatch: da(yb, pos) += w (h[c-1] - h[c])
// (ta = h for SW_ISM); ew = ecs * wx * wy, w = wx * wy (1 without weights).
// SW_ISM also accumulates N[m] (weighted: sum of wx * wy) into cnt; it needs
// cmax = nm - 1 so that every level is a hit level.  add(kind, base, pos, value): kind 0 = res,
// 1 = dm, 2 = da.
template <int LT, int NMC, int MODE, class XW, class YW, class ADD>
GKM_HD void sweep_lane(XW xw, YW yw, int lx, int ly, int l, int nm, int cmax, int delta,
                       int rc, const uint8_t *wt, double ecs, const double *ta,
                       const double *tb, unsigned long long *cnt, ADD &&add)
{
    const int ty_lo = imax(0, delta);
    const int ty_hi = imin(ly, lx + delta) - l;
    if (ty_hi < ty_lo)
        return;
    const int wx = lx - l + 1, wy = ly - l + 1;
    walk_diagonal<LT>(
        xw, yw, l, delta, ty_lo, ty_hi,
        [&](int base, uint32_t ep, uint32_t ec, P2 yp, P2 yc, uint32_t b0, uint32_t b1,
            uint32_t b2, uint32_t b3) {
            const uint32_t valid = range_mask(ty_lo - base, ty_hi - base);
            uint32_t mk[NMC];
            levels<NMC>(b0, b1, b2, b3, valid, nm, mk);
            uint32_t hits = 0;
#pragma unroll
            for (int m = 0; m < NMC; m++) {
                if (MODE == SW_ISM && !wt)
                    cnt[m] += popc(mk[m]);
                if (m <= cmax)
                    hits |= mk[m];
            }
            while (hits) {
                const int j = __builtin_ctz(hits);
                hits &= hits - 1u;
                const int c = count_at(b0, b1, b2, b3, j);
                const int ty = base + j, tx = ty - delta;
                double w = 1.0;
                if (wt) {
                    const unsigned long long wi =
                        (unsigned long long)window_weight(wt, wx, tx) *
                        (unsigned long long)window_weight(wt, wy, rc ? (wy - 1 - ty) : ty);
                    w = (double)wi;
                    if (MODE == SW_ISM) {
                        // weighted N[m]: every level <= dmax is a hit level (cmax = nm - 1)
#pragma unroll
                        for (int m = 0; m < NMC; m++)
                            cnt[m] += (c == m) ? wi : 0ull;
                    }
                }
                double av, bv;
                if (MODE == SW_ISM) {
                    av = w * (ta[c + 1] - ta[c]);
                    bv = c > 0 ? w * (ta[c - 1] - ta[c]) : 0.0;
                } else {
                    double ew = ecs;
                    if (wt)
                        ew = ecs * (double)window_weight(wt, wx, tx) *
                             (double)window_weight(wt, wy, rc ? (wy - 1 - ty) : ty);
                    av = ew * ta[c];
                    bv = ew * tb[c];
                }
                hit_positions(l, j, tx, ep, ec, yp, yc, [&](int pos, bool mis, int yb) {
                    if (MODE == SW_HYP) {
                        add(0, yb, pos, mis ? bv : av);
                    } else if (MODE == SW_PERT) {
                        if (mis) {
                            add(0, yb, pos, bv);
                        } else if (av != 0.0) {
                            for (int h = 0; h < 4; h++)
                                if (h != yb)
                                    add(0, h, pos, av);
                        }
                    } else {
                        if (mis) {
                            if (bv != 0.0)
                                add(2, yb, pos, bv);
                        } else if (av != 0.0) {
                            add(1, 0, pos, av);
                        }
                    }
                });
            }
        });
}


// ---------------------------------------------------------------------------
// Hit records: a window pair with few enough mismatches, reduced to what its
// l positions need (used by the GkmExplain / ISM sweep, whose lanes share the
// per-position work of the hits found by any lane of the wave).
// ---------------------------------------------------------------------------

struct HitRec {
    int tx;         // x window start (positions tx .. tx + l - 1)
    int c;          // mismatch count of the window pair
    uint32_t mis;   // bit q: position q of the windows differs
    uint32_t y0;    // bit q: bit 0 of the SV's base at position q
    uint32_t y1;    // bit q: bit 1 of the SV's base at position q
    uint32_t w;     // wx * wy (1 without window weights)
};

// The record of hit window j (bit j of the word at y position base) on
// diagonal delta.
GKM_HD HitRec make_hit(int l, int j, int base, int delta, int rc, uint32_t ep, uint32_t ec,
                       P2 yp, P2 yc, uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3,
                       const uint8_t *wt, int wx, int wy)
{
    const uint32_t lm = (1u << l) - 1u;
    HitRec h;
    const int ty = base + j;
    h.tx = ty - delta;
    h.c = count_at(b0, b1, b2, b3, j);
    h.mis = funnel(ec, ep, (uint32_t)j) & lm;
    h.y0 = funnel(yc.x, yp.x, (uint32_t)j) & lm;
    h.y1 = funnel(yc.y, yp.y, (uint32_t)j) & lm;
    h.w = wt ? window_weight(wt, wx, h.tx) * window_weight(wt, wy, rc ? (wy - 1 - ty) : ty) : 1u;
    return h;
}

// Position q of a hit: explain adds ew x (mismatch ? tb[c] : ta[c]) to the SV's
// base (hypothetical) or, at a matching position, ew ta[c] to the three other
// bases (perturbation); ISM adds w (h[c+1] - h[c]) to dm (match) or
// w (h[c-1] - h[c]) to da[SV base] (mismatch); ta = h for ISM.  add(kind,
// base, pos, value): kind 0 = res, 1 = dm, 2 = da.
template <int MODE, class ADD>
GKM_HD void apply_hit(const HitRec &h, int q, double ecs, const double *ta, const double *tb,
                      ADD &&add)
{
    const int pos = h.tx + q;
    const bool mis = (h.mis >> q) & 1u;
    const int yb = (int)(((h.y0 >> q) & 1u) | (((h.y1 >> q) & 1u) << 1));
    const int c = h.c;
    if (MODE == SW_ISM) {
        const double w = (double)h.w;
        if (mis) {
            if (c > 0) {
                const double v = w * (ta[c - 1] - ta[c]);
                if (v != 0.0)
                    add(2, yb, pos, v);
            }
        } else {
            const double v = w * (ta[c + 1] - ta[c]);
            if (v != 0.0)
                add(1, 0, pos, v);
        }
        return;
    }
    // the reference multiplies ec x wx x wy (explain/cpu.py:49)
    const double ew = h.w == 1u ? ecs : ecs * (double)h.w;
    if (MODE == SW_HYP) {
        add(0, yb, pos, ew * (mis ? tb[c] : ta[c]));
    } else if (mis) {
        add(0, yb, pos, ew * tb[c]);
    } else {
        const double av = ew * ta[c];
        if (av != 0.0)
            for (int b = 0; b < 4; b++)
                if (b != yb)
                    add(0, b, pos, av);
    }
}

// The walk of one lane's diagonal inside a wave whose lanes take consecutive
// diagonals: the loop runs over the wave's common word range [U0, U1] so all
// lanes stay converged (lanes outside their own range [u0, u1] get an empty
// valid mask and read clamped, in-bounds words: nwx / nwy are the rows' word
// counts).  Words are loaded fresh each step (no carried funnel state), so a
// lane's words are exact from its first word on.  f(base, active, ep, ec, yp,
// yc, b0, b1, b2, b3, valid).
template <int LT, class XW, class YW, class F>
GKM_HD void walk_converged(XW xw, YW yw, int nwx, int nwy, int l, int delta, bool lane_ok,
                           int ty_lo, int ty_hi, int U0, int U1, F &&f)
{
    const uint32_t sh = (uint32_t)(-delta) & 31u;
    const int u0 = ty_lo >> 5, u1 = ty_hi >> 5;
    auto xk = [&](int u) {  // clamped x word index for y word u
        int k = ((32 * u - delta) >> 5) + 1;
        return k < 0 ? 0 : (k > nwx - 2 ? nwx - 2 : k);
    };
    auto yk = [&](int a) { return a < 0 ? 0 : (a > nwy - 1 ? nwy - 1 : a); };
    for (int u = U0; u <= U1; u++) {
        const int ka = xk(u), kb = xk(u + 1);
        const P2 yp = yw(yk(u + 1)), yc = yw(yk(u + 2));
        const uint32_t ep = mism_word(xw(ka), xw(ka + 1), yp, sh);
        const uint32_t ec = mism_word(xw(kb), xw(kb + 1), yc, sh);
        uint32_t b0, b1, b2, b3;
        counts<LT>(ep, ec, l, b0, b1, b2, b3);
        const bool active = lane_ok && u >= u0 && u <= u1;
        const int base = 32 * u;
        const uint32_t valid = active ? range_mask(ty_lo - base, ty_hi - base) : 0u;
        f(base, active, ep, ec, yp, yc, b0, b1, b2, b3, valid);
    }
}


// Word range [U0, U1] of a wave whose lanes take diagonals d0 .. d0 + 63 of an
// (x, y) pair (d = delta + lx - l, nd diagonals): u0 and u1 grow with delta,
// so the first and last valid diagonals bound the range.
GKM_HD void wave_word_range(int lx, int ly, int l, int d0, int nd, int &U0, int &U1)
{
    const int dl = (d0 + 63 < nd ? d0 + 63 : nd - 1);
    const int df = d0 - (lx - l), dl_ = dl - (lx - l);
    U0 = imax(0, df) >> 5;
    U1 = (imin(ly, lx + dl_) - l) >> 5;
}

}  // namespace gkm
