/*
 * gkm_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of gkmsvm-lite's gapped k-mer hot path, used as
 * the checker for the HIP engine: only tests/, the smoke check and the bench's
 * CPU-baseline leg load it.  Nothing in the product links, imports or calls
 * this file.
 *
 * Every function follows a reference function (cited as path:line relative to
 * the gkmsvm-lite source tree, src/gkmsvm/...).  Integer parts (packing,
 * mismatch counts, per-pair histograms, window-weight products) are exact.
 *
 * Pinning (what checks these functions, see tests/test_oracle_*.py):
 *   - orc_pack / orc_rc_windows: hand-packed values, RC involution, and the
 *     LS-GKM gkmpredict fixtures (every score passes through them).
 *   - orc_pair_hist / orc_self_hist: brute force by definition; checked
 *     against a numpy one-hot mismatch count on random inputs, the l=3 k=2
 *     hand values, and the gkmpredict fixtures for -t 0..6.
 *   - orc_explain: the gkmexplain modes 0/1/4 fixtures and completeness.
 *   - orc_index_*: equality with orc_pair_hist-based scores (two algorithms,
 *     one answer) and the gkmpredict fixtures.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PAD 0xFFFFFFFFu

/* kernels/primitives.py:29-32 (_packed_mismatches): 2-bit XOR + popcount. */
static inline int orc_mism(uint32_t a, uint32_t b)
{
    uint32_t x = a ^ b;
    return __builtin_popcount(((x >> 1) | x) & 0x55555555u);
}

/*
 * kernels/packed.py:130-169 (PackedGkmKernel.pack): window i of row r packs
 * codes[r, i + k] into bits 2k, 2k+1; windows past a row's length are PAD.
 * codes: [n, Lmax] uint8 (A=0 C=1 G=2 T=3, N already stored as A);
 * lengths: [n]; out: [n, W], W = Lmax - l + 1.
 */
void orc_pack(const uint8_t *codes, const int64_t *lengths, int64_t n,
              int64_t Lmax, int l, uint32_t *out)
{
    int64_t W = Lmax - l + 1;
    for (int64_t r = 0; r < n; r++) {
        int64_t nwin = lengths[r] - l + 1;
        for (int64_t i = 0; i < W; i++) {
            uint32_t w = 0;
            for (int k = 0; k < l; k++)
                w |= ((uint32_t)(codes[r * Lmax + i + k] & 3u)) << (2 * k);
            out[r * W + i] = (i < nwin) ? w : ORC_PAD;
        }
    }
}

/*
 * kernels/packed.py:40-53 (_rc_windows): window j of the reverse complement is
 * window W-1-j with each base complemented (b ^ 3) and the base order reversed;
 * PAD stays PAD (so a padded row's RC windows hold their pads first).
 */
void orc_rc_windows(const uint32_t *fwd, int64_t n, int64_t W, int l,
                    uint32_t *out)
{
    for (int64_t r = 0; r < n; r++) {
        for (int64_t j = 0; j < W; j++) {
            uint32_t u = fwd[r * W + (W - 1 - j)];
            if (u == ORC_PAD) {
                out[r * W + j] = ORC_PAD;
                continue;
            }
            uint32_t v = 0;
            for (int k = 0; k < l; k++) {
                uint32_t b = 3u - ((u >> (2 * k)) & 3u);
                v |= b << (2 * (l - 1 - k));
            }
            out[r * W + j] = v;
        }
    }
}

/* Histogram of one window block pair: hist[m] += wx_i * wy_j for each pair of
 * non-PAD windows with m <= dmax mismatches.  wx / wy may be NULL (weight 1). */
static void block_hist(const uint32_t *x, const uint8_t *wx, int64_t Wx,
                       const uint32_t *y, const uint8_t *wy, int64_t Wy,
                       int dmax, uint64_t *hist)
{
    for (int64_t i = 0; i < Wx; i++) {
        uint32_t u = x[i];
        if (u == ORC_PAD)
            continue;
        uint64_t a = wx ? wx[i] : 1u;
        for (int64_t j = 0; j < Wy; j++) {
            uint32_t v = y[j];
            if (v == ORC_PAD)
                continue;
            int m = orc_mism(u, v);
            if (m <= dmax)
                hist[m] += a * (uint64_t)(wy ? wy[j] : 1u);
        }
    }
}

/*
 * Per-pair mismatch histograms N[b, s, m], m = 0..dmax: the integer content of
 * kernels/direct.py:24-42 (_fused_pairwise_packed_numba) and
 * kernels/weighted.py:47-69 (_brute_weighted_numba), whose K_raw is
 * sum_m table[m] N[m].  The RC block is added as in kernels/packed.py:221-228
 * (raw = brute(bx, by) + brute(bx, by_rc)).
 * wx [B, Wx] / wy, wy_rc [S, Wy] uint8 window weights or NULL; by_rc may be
 * NULL.  hist: [B, S, dmax + 1] uint64, overwritten.
 */
void orc_pair_hist(const uint32_t *bx, const uint8_t *wx, int64_t B, int64_t Wx,
                   const uint32_t *by, const uint8_t *wy, const uint32_t *by_rc,
                   const uint8_t *wy_rc, int64_t S, int64_t Wy, int dmax,
                   uint64_t *hist)
{
    int nm = dmax + 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = 0; p < B * S; p++) {
        int64_t b = p / S, s = p % S;
        uint64_t *h = hist + p * nm;
        memset(h, 0, sizeof(uint64_t) * nm);
        block_hist(bx + b * Wx, wx ? wx + b * Wx : NULL, Wx, by + s * Wy,
                   wy ? wy + s * Wy : NULL, Wy, dmax, h);
        if (by_rc)
            block_hist(bx + b * Wx, wx ? wx + b * Wx : NULL, Wx, by_rc + s * Wy,
                       wy_rc ? wy_rc + s * Wy : NULL, Wy, dmax, h);
    }
}

/*
 * Self-kernel histograms N[b, m] of K_raw(x, x): row b's forward windows
 * against its own forward windows and its RC windows.
 * kernels/direct.py:110-137 (_self_diagonal_packed_numba) and
 * kernels/weighted.py:91-118 (_self_diagonal_weighted_numba) visit each
 * unordered pair of the symmetric blocks once; the sum over every ordered
 * pair is the same integer.  bx_rc may be NULL (no RC).
 */
void orc_self_hist(const uint32_t *bx, const uint8_t *wx, const uint32_t *bx_rc,
                   const uint8_t *wx_rc, int64_t B, int64_t W, int dmax,
                   uint64_t *hist)
{
    int nm = dmax + 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < B; b++) {
        uint64_t *h = hist + b * nm;
        memset(h, 0, sizeof(uint64_t) * nm);
        const uint8_t *w = wx ? wx + b * W : NULL;
        block_hist(bx + b * W, w, W, bx + b * W, w, W, dmax, h);
        if (bx_rc)
            block_hist(bx + b * W, w, W, bx_rc + b * W,
                       wx_rc ? wx_rc + b * W : NULL, W, dmax, h);
    }
}

/*
 * GkmExplain per-SV pair loop, explain/cpu.py:14-65 (_explain_cpu).
 * packed_x [B, W]; packed_y [S, Wy] (an SV's forward windows followed by its
 * RC windows, as explain/core.py:172-175 concatenates them); table_a/table_b
 * [l + 2]; eff_coef [B, S]; wx [B, W] / wy [S, Wy] float64 weights or NULL.
 * result [B, 4, seqlen] float64, accumulated into (caller zeroes it).
 * For a pair of windows (i, j) with at least min_matches matches, each of the
 * l positions gets ew * table_a[m] on the SV's base when the bases agree, else
 * ew * table_b[m] (hypothetical); perturbation spreads table_a over the three
 * other bases when they agree.
 */
void orc_explain(const uint32_t *packed_x, int64_t B, int64_t W,
                 const uint32_t *packed_y, int64_t S, int64_t Wy,
                 const double *table_a, const double *table_b,
                 const double *eff_coef, int l, int min_matches, int64_t seqlen,
                 int perturbation, const double *wx, const double *wy,
                 double *result)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < B; b++) {
        double *res = result + b * 4 * seqlen;
        for (int64_t s = 0; s < S; s++) {
            double ec = eff_coef[b * S + s];
            for (int64_t i = 0; i < W; i++) {
                uint32_t u = packed_x[b * W + i];
                if (u == ORC_PAD)
                    continue;
                for (int64_t j = 0; j < Wy; j++) {
                    uint32_t v = packed_y[s * Wy + j];
                    if (v == ORC_PAD)
                        continue;
                    int mm = orc_mism(u, v);
                    if (l - mm < min_matches)
                        continue;
                    double ew = ec;
                    if (wx)
                        ew = ec * wx[b * W + i] * wy[s * Wy + j];
                    double av = ew * table_a[mm];
                    double bv = ew * table_b[mm];
                    for (int k = 0; k < l; k++) {
                        int xk = (u >> (2 * k)) & 3;
                        int yk = (v >> (2 * k)) & 3;
                        int64_t pos = k + i;
                        if (!perturbation) {
                            res[yk * seqlen + pos] += (xk == yk) ? av : bv;
                        } else if (xk == yk) {
                            for (int h = 0; h < 4; h++)
                                if (h != yk)
                                    res[h * seqlen + pos] += av;
                        } else {
                            res[yk * seqlen + pos] += bv;
                        }
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * The l-mer index: the reference's own CPU algorithm at the scoring configs.
 * ------------------------------------------------------------------------- */

/*
 * kernels/lmer_index.py:67-87 (neighbor_masks): XOR masks of every l-mer within
 * dmax mismatches (levels 0..dmax, skipping levels whose table weight is 0),
 * sorted by mask value.  Returns the count; masks/levels sized by the caller
 * (orc_neighbor_count).
 */
int64_t orc_neighbor_count(int l, int dmax)
{
    int64_t total = 0;
    for (int m = 0; m <= dmax; m++) {
        int64_t c = 1;
        for (int i = 0; i < m; i++)
            c = c * (l - i) / (i + 1);
        for (int i = 0; i < m; i++)
            c *= 3;
        total += c;
    }
    return total;
}

static int cmp_mask(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_neighbor_masks(int l, int dmax, const double *table, uint32_t *masks,
                           int32_t *levels)
{
    /* Enumerate every mask over l 2-bit digits with <= dmax non-zero digits. */
    int64_t n = 0;
    int64_t total = 1;
    for (int i = 0; i < l; i++)
        total *= 4;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) *
                                       (size_t)orc_neighbor_count(l, dmax));
    for (int64_t x = 0; x < total; x++) {
        int m = 0;
        for (int p = 0; p < l; p++)
            m += ((x >> (2 * p)) & 3) != 0;
        if (m > dmax || table[m] == 0.0)
            continue;
        tmp[n++] = ((uint64_t)x << 8) | (uint64_t)m;
    }
    qsort(tmp, (size_t)n, sizeof(uint64_t), cmp_mask);
    for (int64_t i = 0; i < n; i++) {
        masks[i] = (uint32_t)(tmp[i] >> 8);
        levels[i] = (int32_t)(tmp[i] & 0xFF);
    }
    free(tmp);
    return n;
}

/*
 * kernels/lmer_index.py:161-200 (LmerIndex.__init__): CSR over the 4^l l-mers
 * of every non-PAD SV window (forward block, then RC block, per SV), entries of
 * an l-mer in SV order.  offsets [4^l + 1] int64; ids [n_entries] int32;
 * vals [n_entries] uint8 window weights (only when w / w_rc are given).
 * Returns the number of entries.
 */
int64_t orc_index_build(const uint32_t *by, const uint8_t *wy,
                        const uint32_t *by_rc, const uint8_t *wy_rc, int64_t S,
                        int64_t W, int l, int64_t *offsets, int32_t *ids,
                        uint8_t *vals)
{
    int64_t n_lmers = (int64_t)1 << (2 * l);
    memset(offsets, 0, sizeof(int64_t) * (size_t)(n_lmers + 1));
    int nb = by_rc ? 2 : 1;
    for (int64_t s = 0; s < S; s++)
        for (int blk = 0; blk < nb; blk++) {
            const uint32_t *row = (blk ? by_rc : by) + s * W;
            for (int64_t j = 0; j < W; j++)
                if ((int64_t)row[j] < n_lmers)
                    offsets[row[j] + 1]++;
        }
    for (int64_t v = 0; v < n_lmers; v++)
        offsets[v + 1] += offsets[v];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)n_lmers);
    memcpy(fill, offsets, sizeof(int64_t) * (size_t)n_lmers);
    for (int64_t s = 0; s < S; s++)
        for (int blk = 0; blk < nb; blk++) {
            const uint32_t *row = (blk ? by_rc : by) + s * W;
            const uint8_t *wrow = blk ? (wy_rc ? wy_rc + s * W : NULL)
                                      : (wy ? wy + s * W : NULL);
            for (int64_t j = 0; j < W; j++) {
                uint32_t v = row[j];
                if ((int64_t)v >= n_lmers)
                    continue;
                int64_t e = fill[v]++;
                ids[e] = (int32_t)s;
                if (vals)
                    vals[e] = wrow ? wrow[j] : 1;
            }
        }
    free(fill);
    return offsets[n_lmers];
}

/*
 * kernels/lmer_index_cpu.py:23-37 (_add_window) + :71-120 (_query_reduce_cpu):
 * out[b] = sum_s coef_s f(K_raw[b, s] inv_dx[b] inv_ds[s]) with
 * f(k) = exp(gamma (k - 1)) when rbf, else k; each task keeps one row's hits
 * in a length-S buffer.  weights [n_masks] = table[level]; vals NULL when
 * unweighted (wx then NULL too).  Threads: OpenMP.
 */
void orc_index_reduce(const uint32_t *bx, const uint8_t *wx, int64_t B,
                      int64_t W, const int64_t *offsets, const int32_t *ids,
                      const uint8_t *vals, const uint32_t *masks,
                      const double *weights, int64_t n_masks, int l,
                      const double *inv_dx, const double *inv_ds,
                      const double *coef, int64_t S, int rbf, double gamma,
                      double *out)
{
    int64_t n_lmers = (int64_t)1 << (2 * l);
#pragma omp parallel
    {
        double *acc = (double *)calloc((size_t)S, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < B; b++) {
            for (int64_t i = 0; i < W; i++) {
                uint32_t w = bx[b * W + i];
                if ((int64_t)w >= n_lmers)
                    continue;
                double a = wx ? (double)wx[b * W + i] : 1.0;
                for (int64_t t = 0; t < n_masks; t++) {
                    int64_t v = (int64_t)(w ^ masks[t]);
                    double wt = a * weights[t];
                    for (int64_t e = offsets[v]; e < offsets[v + 1]; e++)
                        acc[ids[e]] += vals ? wt * vals[e] : wt;
                }
            }
            double total = 0.0;
            if (rbf) {
                double g = gamma * inv_dx[b];
                for (int64_t j = 0; j < S; j++) {
                    total += coef[j] * exp(g * acc[j] * inv_ds[j]);
                    acc[j] = 0.0;
                }
                total *= exp(-gamma);
            } else {
                for (int64_t j = 0; j < S; j++) {
                    total += coef[j] * acc[j] * inv_ds[j];
                    acc[j] = 0.0;
                }
                total *= inv_dx[b];
            }
            out[b] = total;
        }
        free(acc);
    }
}

/* Threads the OpenMP loops use (for the CPU-baseline report). */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
