// gkm_emu.cpp -- CPU emulation of the engine's lane code (gkm_core.h) in the
// same shells as the HIP kernels, for checking the bit-sliced algorithm, its
// index ranges and its group / chunk bookkeeping against the CPU oracle without
// a GPU.  Test infrastructure: the product never calls it.
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>
#include <stdexcept>
#include <vector>

#include "gkm_core.h"

using namespace gkm;

#define EMU_API extern "C" __attribute__((visibility("default")))

// Bounds-checked word access: throws (caught below) on an out-of-range index,
// the emulation's stand-in for a GPU memory fault.
struct Words {
    const P2 *p;
    int n;
    P2 operator()(int k) const
    {
        if (k < 0 || k >= n)
            throw std::out_of_range("plane word index");
        return p[k];
    }
};

EMU_API void emu_pack_planes(const uint8_t *codes, const int32_t *lengths, int n, int lmax,
                             int nw, P2 *planes, P2 *planes_rc)
{
    for (int r = 0; r < n; r++) {
        int len = lengths[r];
        const uint8_t *row = codes + (int64_t)r * lmax;
        for (int a = 0; a < nw; a++) {
            uint32_t f0 = 0, f1 = 0, r0 = 0, r1 = 0;
            for (int i = 0; a >= 1 && i < 32; i++) {
                int t = (a - 1) * 32 + i;
                if (t < len) {
                    uint32_t c = row[t] & 3u, rc = 3u - (row[len - 1 - t] & 3u);
                    f0 |= (c & 1u) << i;
                    f1 |= (c >> 1) << i;
                    r0 |= (rc & 1u) << i;
                    r1 |= (rc >> 1) << i;
                }
            }
            planes[(int64_t)r * nw + a] = P2{f0, f1};
            planes_rc[(int64_t)r * nw + a] = P2{r0, r1};
        }
    }
}

template <int LT, int NMC, bool W>
static void cross_t(const P2 *xpl, const int32_t *xlen, int B, int nwx, const P2 *ypl,
                    const P2 *ypl_rc, const int32_t *ylen, int S, int nwy, int l, int nm,
                    const uint8_t *wtab, int nsplit, int64_t col_min, unsigned long long *hist)
{
    const int norient = ypl_rc ? 2 : 1;
    const int ngroups = (S + 63) / 64;
    for (int g = 0; g < ngroups; g++) {
        int ly_min = 1 << 30, ly_max = 0;
        for (int s = g * 64; s < S && s < g * 64 + 64; s++) {
            ly_min = ylen[s] < ly_min ? ylen[s] : ly_min;
            ly_max = ylen[s] > ly_max ? ylen[s] : ly_max;
        }
        if (ly_max < l)
            continue;
        for (int b = 0; b < B; b++)
            for (int o = 0; o < norient; o++)
                for (int split = 0; split < nsplit; split++) {
                    const int lx = xlen[b];
                    if (lx < l)
                        continue;
                    if (col_min >= 0 && (int64_t)g * 64 + 63 < b + col_min)
                        continue;
                    const int dlo = -(lx - l), dhi = ly_max - l;
                    const int nd = dhi - dlo + 1, per = (nd + nsplit - 1) / nsplit;
                    const int d0 = dlo + split * per, d1 = imin(dhi + 1, d0 + per);
                    for (int lane = 0; lane < 64; lane++) {
                        const int s = g * 64 + lane;
                        const int ly = s < S ? ylen[s] : 0;
                        bool on = s < S && ly >= l;
                        if (col_min >= 0)
                            on = on && (int64_t)s >= b + col_min;
                        // lanes past S read the zero tile
                        std::vector<P2> zero(nwy, P2{0u, 0u});
                        const P2 *yrow = s < S ? ((o ? ypl_rc : ypl) + (int64_t)s * nwy) : zero.data();
                        uint32_t cnt[NMC] = {0};
                        unsigned long long wcnt[NMC] = {0};
                        cross_lane<LT, NMC, W>(Words{xpl + (int64_t)b * nwx, nwx}, Words{yrow, nwy}, lx,
                                               ly, on, ly_min, ly_max, l, nm, o, d0, d1, wtab, cnt,
                                               wcnt);
                        if (on)
                            for (int m = 0; m < nm && m < NMC; m++)
                                hist[((int64_t)b * S + s) * nm + m] += W ? wcnt[m] : cnt[m];
                    }
                }
    }
}

template <int LT, int NMC>
static void cross_w(bool w, const P2 *xpl, const int32_t *xlen, int B, int nwx, const P2 *ypl,
                    const P2 *ypl_rc, const int32_t *ylen, int S, int nwy, int l, int nm,
                    const uint8_t *wtab, int nsplit, int64_t col_min, unsigned long long *hist)
{
    if (w)
        cross_t<LT, NMC, true>(xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit,
                               col_min, hist);
    else
        cross_t<LT, NMC, false>(xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit,
                                col_min, hist);
}

// Same arguments as gkm_cross_hist, host pointers; hist accumulated into.
// Returns 0, or -1 when a plane word index fell outside its row.
EMU_API int emu_cross_hist(const P2 *xpl, const int32_t *xlen, int B, int nwx, const P2 *ypl,
                           const P2 *ypl_rc, const int32_t *ylen, int S, int nwy, int l, int nm,
                           const uint8_t *wtab, int nsplit, int64_t col_min,
                           unsigned long long *hist, int generic_l)
{
    try {
        bool w = wtab != nullptr;
        bool small = nm <= 4;
        if (l == 11 && !generic_l) {
            if (small)
                cross_w<11, 4>(w, xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit, col_min, hist);
            else
                cross_w<11, 16>(w, xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit, col_min, hist);
        } else if (l == 10 && !generic_l) {
            if (small)
                cross_w<10, 4>(w, xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit, col_min, hist);
            else
                cross_w<10, 16>(w, xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit, col_min, hist);
        } else {
            if (small)
                cross_w<0, 4>(w, xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit, col_min, hist);
            else
                cross_w<0, 16>(w, xpl, xlen, B, nwx, ypl, ypl_rc, ylen, S, nwy, l, nm, wtab, nsplit, col_min, hist);
        }
    } catch (const std::out_of_range &) {
        return -1;
    }
    return 0;
}

EMU_API int emu_self_hist(const P2 *pl, const P2 *pl_rc, const int32_t *len, int n, int nw, int l,
                          int nm, const uint8_t *wtab, unsigned long long *hist)
{
    try {
        for (int r = 0; r < n; r++) {
            const int lx = len[r];
            if (lx < l)
                continue;
            const int nd = 2 * (lx - l) + 1;
            const int norient = pl_rc ? 2 : 1;
            uint32_t cnt[16] = {0};
            unsigned long long wcnt[16] = {0};
            for (int d = 0; d < nd * norient; d++) {
                const int o = d / nd, delta = d % nd - (lx - l);
                Words xs{pl + (int64_t)r * nw, nw};
                Words ys{(o ? pl_rc : pl) + (int64_t)r * nw, nw};
                if (wtab)
                    self_lane<0, 16, true>(xs, ys, lx, l, nm, delta, o, wtab, cnt, wcnt);
                else
                    self_lane<0, 16, false>(xs, ys, lx, l, nm, delta, o, wtab, cnt, wcnt);
            }
            for (int m = 0; m < nm; m++)
                hist[(int64_t)r * nm + m] += wtab ? wcnt[m] : cnt[m];
        }
    } catch (const std::out_of_range &) {
        return -1;
    }
    return 0;
}

// Explain (mode 0: hypothetical, 1: perturbation) or ISM deltas (mode 2) of
// row b against every SV, with the sweep kernel's bookkeeping: 64-diagonal
// blocks walked over the block's common word range (walk_converged), hit
// records (make_hit) applied position by position (apply_hit).  res
// [B][4][lmax] (explain), or for ISM dk [B][S][5][lmax] (dm then da[4]) plus
// hist [B][S][nm].
EMU_API int emu_sweep(int mode, const P2 *xpl, const int32_t *xlen, int lmax, int B, int nwx,
                      const P2 *ypl, const P2 *ypl_rc, const int32_t *ylen, int S, int nwy, int l,
                      int nm, int cmax, const uint8_t *wtab, const double *ta, const double *tb,
                      const double *ec, double *res, double *dk, unsigned long long *hist)
{
    try {
        for (int b = 0; b < B; b++) {
            const int lx = xlen[b];
            if (lx < l)
                continue;
            const int wx = lx - l + 1;
            for (int s = 0; s < S; s++) {
                const int ly = ylen[s];
                if (ly < l)
                    continue;
                const int wy = ly - l + 1;
                const int norient = ypl_rc ? 2 : 1;
                const int nd = lx + ly - 2 * l + 1;
                const int nblk = (nd + 63) / 64;
                unsigned long long cnt[16] = {0};
                double *dks = dk ? dk + ((int64_t)b * S + s) * 5 * lmax : nullptr;
                double *rb = res ? res + (int64_t)b * 4 * lmax : nullptr;
                auto add = [&](int kind, int yb, int pos, double v) {
                    if (pos < 0 || pos >= lx || yb < 0 || yb > 3)
                        throw std::out_of_range("position");
                    if (kind == 0)
                        rb[yb * lmax + pos] += v;
                    else if (kind == 1)
                        dks[pos] += v;
                    else
                        dks[(1 + yb) * lmax + pos] += v;
                };
                const double ecs = mode == SW_ISM ? 1.0 : ec[(int64_t)b * S + s];
                for (int blk = 0; blk < nblk * norient; blk++) {
                    const int o = blk / nblk, d0 = (blk % nblk) * 64;
                    int U0, U1;
                    wave_word_range(lx, ly, l, d0, nd, U0, U1);
                    for (int lane = 0; lane < 64; lane++) {
                        const int d = d0 + lane;
                        const bool lane_ok = d < nd;
                        const int delta = (lane_ok ? d : nd - 1) - (lx - l);
                        const int ty_lo = imax(0, delta);
                        const int ty_hi = imin(ly, lx + delta) - l;
                        Words xs{xpl + (int64_t)b * nwx, nwx};
                        Words ys{(o ? ypl_rc : ypl) + (int64_t)s * nwy, nwy};
                        walk_converged<0>(
                            xs, ys, nwx, nwy, l, delta, lane_ok, ty_lo, ty_hi, U0, U1,
                            [&](int base, bool, uint32_t ep, uint32_t ecw, P2 yp, P2 yc,
                                uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3,
                                uint32_t valid) {
                                uint32_t mk[16];
                                levels<16>(b0, b1, b2, b3, valid, nm, mk);
                                uint32_t hits = 0;
                                for (int m = 0; m < 16; m++) {
                                    if (mode == SW_ISM && !wtab)
                                        cnt[m] += popc(mk[m]);
                                    if (m <= cmax)
                                        hits |= mk[m];
                                }
                                while (hits) {
                                    const int j = __builtin_ctz(hits);
                                    hits &= hits - 1u;
                                    const HitRec h = make_hit(l, j, base, delta, o, ep, ecw, yp, yc,
                                                              b0, b1, b2, b3, wtab, wx, wy);
                                    if (mode == SW_ISM && wtab)
                                        cnt[h.c] += h.w;
                                    for (int q = 0; q < l; q++) {
                                        if (h.tx + q >= lx)
                                            continue;
                                        if (mode == SW_HYP)
                                            apply_hit<SW_HYP>(h, q, ecs, ta, tb, add);
                                        else if (mode == SW_PERT)
                                            apply_hit<SW_PERT>(h, q, ecs, ta, tb, add);
                                        else
                                            apply_hit<SW_ISM>(h, q, ecs, ta, tb, add);
                                    }
                                }
                            });
                    }
                }
                if (hist)
                    for (int m = 0; m < nm; m++)
                        hist[((int64_t)b * S + s) * nm + m] += cnt[m];
            }
        }
    } catch (const std::out_of_range &) {
        return -1;
    }
    return 0;
}
