// gkm_kernels.hip -- MI355X (gfx950) kernels of the gapped k-mer engine and
// the C-ABI that exposes them (declared in include/gkmhip.h).
//
// Device data (all caller-owned, device pointers):
//   planes     [n][nw] uint2   bit-plane words of each sequence (.x = code bit
//                              0, .y = code bit 1), word 0 a zero pad; base t
//                              in bit t & 31 of word (t >> 5) + 1;
//                              nw >= Lmax / 32 + 4
//   planes_rc  [n][nw] uint2   the same for each sequence's reverse complement
//                              (base t = 3 - code[len - 1 - t]), so an N, read
//                              as A on the forward strand, is T on the reverse
//   lengths    [n] int32
//   wtab       [Wmax/2 + 1] u8 window weight by distance from the centre
//                              (weighted kernels -t 4/-t 5), else NULL
//   hist       [B][S][nm] u64  per-pair mismatch histograms N[m], m < nm
//
// Kernels:
//   k_decode_ascii   ASCII DNA -> base codes, rows with other letters flagged
//   k_pack_planes    codes -> planes, planes_rc
//   k_pack_windows   codes -> 2-bit packed l-mer windows (mutant self-kernels)
//   k_cross_hist     N[m] of query x against SV y (+ RC) -- the hot loop.  One
//                    wave = one query x 64 SVs (lanes); the query's words are
//                    wave-uniform, the SVs' words come from an LDS tile; the
//                    wave walks diagonals, 32 window pairs per lane-step.
//   k_self_hist      N[m] of x against (x, rc x): lanes = diagonals.
//   k_combine_*      FP64: sum_m h[m] N[m], normalisation, RBF, alpha sums.
//   k_mutant_diag    K_raw(mutant, mutant) for every single-base substitution.
//   k_sweep          GkmExplain and ISM: lanes = diagonals of one (x, SV)
//                    pair, per-position FP64 accumulation in LDS.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <math.h>

#include <mutex>
#include <type_traits>

#include "gkm_core.h"
#include "gkm_internal.h"
#include "gkmhip.h"

static int check_l_nm(int l, int nm);

using namespace gkm;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

static thread_local char g_err[1024];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

GKM_API const char *gkm_last_error(void) { return g_err; }

GKM_API int gkm_version(void) { return 1; }

// ---------------------------------------------------------------------------
// Packing
// ---------------------------------------------------------------------------

// ASCII DNA -> base codes (codec.py:40-137 of gkmsvm-lite, one string at a
// time there): row b's bytes raw[starts[b] .. starts[b] + len[b]) become
// codes[b][0 .. lmax) -- A=0 C=1 G=2 T=3 in either case, N and every other
// byte stored as A, A past the row -- and bad[b] = 1 when the row holds a byte
// other than ACGTN / acgtn, else 0.  For those ten letters
// ((c >> 1) ^ (c >> 2)) & 3 is the code (N: 0).  One block per row.
__global__ __launch_bounds__(256) void k_decode_ascii(const uint8_t *__restrict__ raw,
                                                      const int64_t *__restrict__ starts,
                                                      const int32_t *__restrict__ lengths, int lmax,
                                                      uint8_t *__restrict__ codes,
                                                      int32_t *__restrict__ bad)
{
    const int b = blockIdx.x;
    const int len = lengths[b];
    const uint8_t *src = raw + starts[b];
    uint8_t *dst = codes + (int64_t)b * lmax;
    int other = 0;
    for (int p = threadIdx.x; p < lmax; p += 256) {
        uint32_t code = 0;
        if (p < len) {
            const uint32_t c = src[p];
            const uint32_t u = c & ~32u;  // the upper case of a letter
            const bool ok = u == 'A' || u == 'C' || u == 'G' || u == 'T' || u == 'N';
            other |= ok ? 0 : 1;
            code = ok ? ((c >> 1) ^ (c >> 2)) & 3u : 0u;
        }
        dst[p] = (uint8_t)code;
    }
    other = __syncthreads_or(other);
    if (threadIdx.x == 0)
        bad[b] = other ? 1 : 0;
}

// One thread per (sequence, word): 32 codes -> two plane words, forward and
// reverse complement.  Words past the sequence are zero (read as A, and every
// window touching them is masked out by length).
__global__ void k_pack_planes(const uint8_t *__restrict__ codes,
                              const int32_t *__restrict__ lengths, int n, int lmax,
                              int nw, P2 *__restrict__ planes,
                              P2 *__restrict__ planes_rc)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * nw)
        return;
    int r = (int)(idx / nw), a = (int)(idx % nw);
    int len = lengths[r];
    const uint8_t *row = codes + (int64_t)r * lmax;
    uint32_t f0 = 0, f1 = 0, r0 = 0, r1 = 0;
    int t0 = (a - 1) * 32;
    if (a >= 1) {
        for (int i = 0; i < 32; i++) {
            int t = t0 + i;
            if (t < len) {
                uint32_t c = row[t] & 3u;
                f0 |= (c & 1u) << i;
                f1 |= (c >> 1) << i;
                uint32_t rc = 3u - (row[len - 1 - t] & 3u);
                r0 |= (rc & 1u) << i;
                r1 |= (rc >> 1) << i;
            }
        }
    }
    planes[idx] = P2{f0, f1};
    if (planes_rc)
        planes_rc[idx] = P2{r0, r1};
}

// gkmsvm-lite's packed windows (kernels/packed.py:130-169): window i packs
// bases i..i+l-1 at bits 2k; PAD past the last window.  The RC row here holds
// the windows of the reverse complement from its first base (window j of
// rc(x), j < len - l + 1), PAD after them.
__global__ void k_pack_windows(const uint8_t *__restrict__ codes,
                               const int32_t *__restrict__ lengths, int n, int lmax,
                               int l, int W, uint32_t *__restrict__ fwd,
                               uint32_t *__restrict__ rcw)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * W)
        return;
    int r = (int)(idx / W), i = (int)(idx % W);
    int len = lengths[r];
    const uint8_t *row = codes + (int64_t)r * lmax;
    uint32_t f = 0xFFFFFFFFu, v = 0xFFFFFFFFu;
    if (i < len - l + 1) {
        f = 0;
        v = 0;
        for (int k = 0; k < l; k++) {
            f |= (uint32_t)(row[i + k] & 3u) << (2 * k);
            v |= (3u - (uint32_t)(row[len - 1 - (i + k)] & 3u)) << (2 * k);
        }
    }
    fwd[idx] = f;
    if (rcw)
        rcw[idx] = v;
}

// ---------------------------------------------------------------------------
// k_cross_hist
// ---------------------------------------------------------------------------

constexpr int CROSS_THREADS = 256;  // 4 waves
constexpr int WAVE = 64;

struct CrossArgs {
    const P2 *xplanes;       // [B][nwx]
    const int32_t *xlen;     // [B]
    const P2 *yplanes;       // [S][nwy]
    const P2 *yplanes_rc;    // [S][nwy] or NULL (no RC)
    const int32_t *ylen;     // [S]
    const int32_t *ygrp;     // [n_groups][2]: min, max length of each 64-SV group
    const uint8_t *wtab;     // weight by distance, or NULL
    int wtab_len;
    unsigned long long *hist;  // [B][S][nm] (row b = x row b, column s)
    int B, S, nwx, nwy, l, nm;
    int rows_per_block;      // x rows handled by one block
    int nsplit;              // diagonal chunks per (row, orientation)
    int64_t col_min;         // >= 0: column s counts only when s >= row + col_min
    int64_t row_base;        // global index of x row 0 (for col_min)
};

template <int LT, int NMC, bool WEIGHTED>
__global__ __launch_bounds__(CROSS_THREADS) void k_cross_hist(CrossArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    const int g = blockIdx.x;  // SV group
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int norient = a.yplanes_rc ? 2 : 1;
    const int nwy = a.nwy;
    // LDS: SV tile [orient][word][lane] plane words, then the weight table.
    P2 *tile = (P2 *)smem;
    uint8_t *wt = smem + (size_t)norient * nwy * WAVE * sizeof(P2);

    const int s0 = g * WAVE;
    for (int idx = threadIdx.x; idx < norient * nwy * WAVE; idx += CROSS_THREADS) {
        int o = idx / (nwy * WAVE);
        int rem = idx % (nwy * WAVE);
        int sl = rem / nwy, w = rem % nwy;  // coalesced over (sv, word)
        int s = s0 + sl;
        P2 v = {0u, 0u};
        if (s < a.S)
            v = (o ? a.yplanes_rc : a.yplanes)[(int64_t)s * nwy + w];
        tile[((size_t)o * nwy + w) * WAVE + sl] = v;
    }
    if (WEIGHTED)
        for (int i = threadIdx.x; i < a.wtab_len; i += CROSS_THREADS)
            wt[i] = a.wtab[i];
    __syncthreads();

    const int l = LT > 0 ? LT : a.l;
    const int s = s0 + lane;
    const int ly = s < a.S ? a.ylen[s] : 0;
    const int ly_min = a.ygrp[2 * g], ly_max = a.ygrp[2 * g + 1];
    if (ly_max < l)
        return;
    const int row0 = blockIdx.y * a.rows_per_block;
    const int nrows = min(a.rows_per_block, a.B - row0);
    const int units = nrows * norient * a.nsplit;

    for (int unit = wave; unit < units; unit += CROSS_THREADS / WAVE) {
        const int split = unit % a.nsplit;
        const int o = (unit / a.nsplit) % norient;
        const int b = row0 + unit / (a.nsplit * norient);
        const int lx = a.xlen[b];
        if (lx < l)
            continue;
        bool lane_on = s < a.S && ly >= l;
        if (a.col_min >= 0) {
            int64_t lo = a.row_base + b + a.col_min;
            if ((int64_t)s0 + WAVE - 1 < lo)
                continue;  // the whole group lies below the diagonal
            lane_on = lane_on && (int64_t)s >= lo;
        }
        const P2 *xp = a.xplanes + (int64_t)b * a.nwx;
        const P2 *yt = tile + (size_t)o * nwy * WAVE + lane;
        // Diagonals delta = ty - tx in [-(lx - l), ly_max - l], split evenly.
        const int dlo = -(lx - l), dhi = ly_max - l;
        const int nd = dhi - dlo + 1;
        const int per = (nd + a.nsplit - 1) / a.nsplit;
        const int d0 = dlo + split * per, d1 = min(dhi + 1, d0 + per);

        uint32_t cnt[NMC];
        unsigned long long wcnt[NMC];
#pragma unroll
        for (int m = 0; m < NMC; m++) {
            cnt[m] = 0;
            wcnt[m] = 0;
        }
        cross_lane<LT, NMC, WEIGHTED>([&](int k) { return xp[k]; },
                                      [&](int u) { return yt[(size_t)u * WAVE]; }, lx, ly, lane_on,
                                      ly_min, ly_max, l, a.nm, o, d0, d1, wt, cnt, wcnt);
        if (lane_on) {
            unsigned long long *h = a.hist + ((int64_t)b * a.S + s) * a.nm;
#pragma unroll
            for (int m = 0; m < NMC; m++) {
                if (m < a.nm) {
                    unsigned long long v = WEIGHTED ? wcnt[m] : (unsigned long long)cnt[m];
                    if (v)
                        atomicAdd(h + m, v);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_self_hist: K_raw(x, x) histograms, lanes = diagonals
// ---------------------------------------------------------------------------

constexpr int SELF_THREADS = 256;

struct SelfArgs {
    const P2 *planes, *planes_rc;  // [n][nw]
    const int32_t *len;
    const uint8_t *wtab;
    int wtab_len;
    unsigned long long *hist;  // [n][nm]
    int n, nw, l, nm, nchunks;
};

template <int LT, int NMC, bool WEIGHTED>
__global__ __launch_bounds__(SELF_THREADS) void k_self_hist(SelfArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    const int r = blockIdx.x;
    const int nw = a.nw;
    P2 *xs = (P2 *)smem;
    P2 *rs = xs + nw;
    unsigned long long *red = (unsigned long long *)(rs + nw);  // [16][SELF_THREADS/64]
    uint8_t *wt = (uint8_t *)(red + 16 * (SELF_THREADS / WAVE));
    const int norient = a.planes_rc ? 2 : 1;
    for (int i = threadIdx.x; i < nw; i += SELF_THREADS) {
        xs[i] = a.planes[(int64_t)r * nw + i];
        rs[i] = norient > 1 ? a.planes_rc[(int64_t)r * nw + i] : P2{0u, 0u};
    }
    if (WEIGHTED)
        for (int i = threadIdx.x; i < a.wtab_len; i += SELF_THREADS)
            wt[i] = a.wtab[i];
    __syncthreads();
    const int l = LT > 0 ? LT : a.l;
    const int lx = a.len[r];
    uint32_t cnt[NMC];
    unsigned long long wcnt[NMC];
#pragma unroll
    for (int m = 0; m < NMC; m++) {
        cnt[m] = 0;
        wcnt[m] = 0;
    }
    if (lx >= l) {
        const int nd = 2 * (lx - l) + 1;
        const int total = nd * norient;
        for (int d = blockIdx.y * SELF_THREADS + threadIdx.x; d < total;
             d += SELF_THREADS * a.nchunks) {
            const int o = d / nd;
            const int delta = d % nd - (lx - l);
            const P2 *ys = o ? rs : xs;
            self_lane<LT, NMC, WEIGHTED>([&](int k) { return xs[k]; }, [&](int u) { return ys[u]; },
                                         lx, l, a.nm, delta, o, wt, cnt, wcnt);
        }
    }
    // Block reduction (integers: exact in any order).
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int m = 0; m < NMC; m++) {
        unsigned long long v = WEIGHTED ? wcnt[m] : (unsigned long long)cnt[m];
        for (int off = WAVE / 2; off > 0; off >>= 1)
            v += __shfl_down(v, off, WAVE);
        if (lane == 0)
            red[m * (SELF_THREADS / WAVE) + wave] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < NMC && (int)threadIdx.x < a.nm) {
        unsigned long long v = 0;
        for (int w = 0; w < SELF_THREADS / WAVE; w++)
            v += red[threadIdx.x * (SELF_THREADS / WAVE) + w];
        if (v)
            atomicAdd(a.hist + (int64_t)r * a.nm + threadIdx.x, v);
    }
}

// ---------------------------------------------------------------------------
// FP64 combine kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ double kraw_of(const unsigned long long *h, const double *tab, int nm)
{
    double k = 0.0;
    for (int m = 0; m < nm; m++)
        k += tab[m] * (double)h[m];
    return k;
}

// K_raw [n] = sum_m h[m] N[m].
__global__ void k_combine_kraw(const unsigned long long *__restrict__ hist, int64_t n,
                               int nm, const double *__restrict__ tab, double *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = kraw_of(hist + i * nm, tab, nm);
}

// Normalised, finished kernel values [B][S]: F(K / max(sqrt(dx ds), 1e-10)),
// F = exp(gamma (k - 1)) when rbf, else the identity (gram.py:94-97).
__global__ void k_combine_gram(const unsigned long long *__restrict__ hist, int B, int S,
                               int nm, const double *__restrict__ tab,
                               const double *__restrict__ dx, const double *__restrict__ ds,
                               int normalize, int rbf, double gamma,
                               double *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * S)
        return;
    int b = (int)(i / S), s = (int)(i % S);
    double k = kraw_of(hist + i * nm, tab, nm);
    if (normalize)
        k = k / fmax(sqrt(dx[b] * ds[s]), 1e-10);
    if (rbf)
        k = exp(gamma * (k - 1.0));
    out[i] = k;
}

// Block-wide sum in a fixed order (deterministic).
__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    for (int off = WAVE / 2; off > 0; off >>= 1)
        v += __shfl_down(v, off, WAVE);
    if (lane == 0)
        red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x / WAVE); w++)
            t += red[w];
    __syncthreads();
    return t;
}

// Decision values: out[b] = sum_s coef_s F(K[b,s] / max(sqrt(dx ds), 1e-10)) + bias
// (svm.py:43-61, _normalized_dot), one block per row.
// One block of SCORE_THREADS per row; each thread keeps SCORE_UNROLL
// independent sums (their loads in flight together) added in a fixed order,
// then the fixed-order block sum: a row's score depends only on that row.
constexpr int SCORE_THREADS = 1024;
constexpr int SCORE_UNROLL = 4;

__global__ __launch_bounds__(SCORE_THREADS) void k_combine_score(
    const unsigned long long *__restrict__ hist, int S, int nm, const double *__restrict__ tab,
    const double *__restrict__ dx, const double *__restrict__ ds,
    const double *__restrict__ coef, int normalize, int rbf, double gamma, double bias,
    double *__restrict__ out)
{
    __shared__ double red[SCORE_THREADS / WAVE];
    __shared__ double stab[16];
    if (threadIdx.x < 16)
        stab[threadIdx.x] = (int)threadIdx.x < nm ? tab[threadIdx.x] : 0.0;
    __syncthreads();
    const int b = blockIdx.x;
    double acc[SCORE_UNROLL];
#pragma unroll
    for (int u = 0; u < SCORE_UNROLL; ++u)
        acc[u] = 0.0;
    const double dxb = dx ? dx[b] : 1.0;
    for (int s0 = threadIdx.x; s0 < S; s0 += SCORE_THREADS * SCORE_UNROLL) {
#pragma unroll
        for (int u = 0; u < SCORE_UNROLL; ++u) {
            const int s = s0 + u * SCORE_THREADS;
            if (s >= S)
                continue;
            double k = kraw_of(hist + ((int64_t)b * S + s) * nm, stab, nm);
            if (normalize)
                k = k / fmax(sqrt(dxb * ds[s]), 1e-10);
            if (rbf)
                k = exp(gamma * (k - 1.0));
            acc[u] += coef[s] * k;
        }
    }
    double v = acc[0];
#pragma unroll
    for (int u = 1; u < SCORE_UNROLL; ++u)
        v += acc[u];
    double t = block_sum(v, red);
    if (threadIdx.x == 0)
        out[b] = t + bias;
}

// Mutation impact (explain/impact.py:22-48): rows [ref; alt] of hist (2n rows),
// impact[b] = sum_s coef_s dk (RBF: dk (exp(g(k-1)) - exp(-g)) / k for k > 0).
__global__ __launch_bounds__(SCORE_THREADS) void k_combine_impact(
    const unsigned long long *__restrict__ hist, int n, int S, int nm,
    const double *__restrict__ tab, const double *__restrict__ dref,
    const double *__restrict__ ds, const double *__restrict__ coef, int normalize, int rbf,
    double gamma, double *__restrict__ out)
{
    __shared__ double red[SCORE_THREADS / WAVE];
    __shared__ double stab[16];
    if (threadIdx.x < 16)
        stab[threadIdx.x] = (int)threadIdx.x < nm ? tab[threadIdx.x] : 0.0;
    __syncthreads();
    const int b = blockIdx.x;
    const double floor_ = rbf ? exp(-gamma) : 0.0;
    double acc[SCORE_UNROLL];
#pragma unroll
    for (int u = 0; u < SCORE_UNROLL; ++u)
        acc[u] = 0.0;
    for (int s0 = threadIdx.x; s0 < S; s0 += SCORE_THREADS * SCORE_UNROLL) {
#pragma unroll
        for (int u = 0; u < SCORE_UNROLL; ++u) {
            const int s = s0 + u * SCORE_THREADS;
            if (s >= S)
                continue;
            double k = kraw_of(hist + ((int64_t)b * S + s) * nm, stab, nm);
            double dk = kraw_of(hist + ((int64_t)(n + b) * S + s) * nm, stab, nm) - k;
            if (normalize) {
                double z = fmax(sqrt(dref[b] * ds[s]), 1e-10);
                k /= z;
                dk /= z;
            }
            if (rbf) {
                if (k <= 0.0)
                    continue;
                dk *= (exp(gamma * (k - 1.0)) - floor_) / k;
            }
            acc[u] += coef[s] * dk;
        }
    }
    double v = acc[0];
#pragma unroll
    for (int u = 1; u < SCORE_UNROLL; ++u)
        v += acc[u];
    double t = block_sum(v, red);
    if (threadIdx.x == 0)
        out[b] = t;
}

// GkmExplain effective coefficients (explain/core.py:203-216):
// ec[b,s] = coef_s / max(sqrt(dx ds), 1e-10), times _rbf_scale(k) for RBF
// (explain/common.py:27-32).
__global__ void k_explain_coef(const unsigned long long *__restrict__ hist, int B, int S,
                               int nm, const double *__restrict__ tab,
                               const double *__restrict__ dx, const double *__restrict__ ds,
                               const double *__restrict__ coef, int normalize, int rbf,
                               double gamma, double *__restrict__ ec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * S)
        return;
    int b = (int)(i / S), s = (int)(i % S);
    double e = coef[s];
    double z = normalize ? fmax(sqrt(dx[b] * ds[s]), 1e-10) : 1.0;
    if (normalize)
        e = e / z;
    if (rbf) {
        double k = kraw_of(hist + i * nm, tab, nm) / z;
        e = e * (k > 0.0 ? (exp(gamma * (k - 1.0)) - exp(-gamma)) / k : 0.0);
    }
    ec[i] = e;
}

// ---------------------------------------------------------------------------
// k_mutant_diag: K_raw(mutant, mutant) [4][Lmax][n]
// ---------------------------------------------------------------------------

constexpr int MUT_THREADS = 128;

__device__ __forceinline__ int mism2(uint32_t a, uint32_t b)
{
    uint32_t x = a ^ b;
    return __builtin_popcount(((x >> 1) | x) & 0x55555555u);
}

// One block per (row, position).  The substitution at p changes the forward
// windows i in C = [p-l+1, p] and the RC windows j in C' = [q-l+1, q],
// q = len-1-p (base 3 - v there).  K_raw(x', x') - K_raw(x, x) is
//   forward block: 2 sum_{i in C, j not in C} (g(x'_i, x_j) - g(x_i, x_j))
//                  + sum_{i, j in C} (g(x'_i, x'_j) - g(x_i, x_j))
//   RC block:      sum_{i in C, all j} (g(x'_i, r'_j) - g(x_i, r_j))
//                  + sum_{i not in C, j in C'} (g(x_i, r'_j) - g(x_i, r_j))
// with g(u, v) = w_u w_v h[m(u, v)] -- the pairs kernels/direct.py:141-201
// (_block_delta, _mutant_diag_numba) and kernels/weighted.py:122-197 visit.
__global__ __launch_bounds__(MUT_THREADS) void k_mutant_diag(
    const uint32_t *__restrict__ fwd, const uint32_t *__restrict__ rcw,
    const int32_t *__restrict__ len, const uint8_t *__restrict__ codes, int lmax, int n,
    int W, int l, const double *__restrict__ tab, const uint8_t *__restrict__ wtab,
    const double *__restrict__ diag, double *__restrict__ out)
{
    __shared__ double stab[16];
    __shared__ double red[4][MUT_THREADS / WAVE];
    const int r = blockIdx.y, p = blockIdx.x;
    if (threadIdx.x < 16)
        stab[threadIdx.x] = (int)threadIdx.x <= l ? tab[threadIdx.x] : 0.0;
    __syncthreads();
    const int L = len[r];
    const int w = L - l + 1;
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    const uint8_t xb = p < L ? (codes[(int64_t)r * lmax + p] & 3u) : 0u;
    if (p < L && w > 0) {
        const uint32_t *f = fwd + (int64_t)r * W;
        const uint32_t *rv = rcw ? rcw + (int64_t)r * W : nullptr;
        const int c0 = max(0, p - l + 1), c1 = min(p, w - 1);
        const int q = L - 1 - p;
        const int r0 = max(0, q - l + 1), r1 = min(q, w - 1);
        auto wgt = [&](int nwin, int j) -> double {
            return wtab ? (double)window_weight(wtab, nwin, j) : 1.0;
        };
        for (int j = threadIdx.x; j < w; j += MUT_THREADS) {
            const uint32_t xj = f[j];
            const double wj = wgt(w, j);
            const bool jin = j >= c0 && j <= c1;
            const uint32_t rj = rv ? rv[j] : 0u;
            const double wrj = wgt(w, w - 1 - j);
            const bool jrin = rv && j >= r0 && j <= r1;
            for (int i = c0; i <= c1; i++) {
                const uint32_t xi = f[i];
                const double wi = wgt(w, i);
                const int shi = 2 * (p - i);
                const double g_old = wi * wj * stab[mism2(xi, xj)];
                const double gr_old = rv ? wi * wrj * stab[mism2(xi, rj)] : 0.0;
                for (int v = 0; v < 4; v++) {
                    if (v == xb)
                        continue;
                    const uint32_t xi2 = (xi & ~(3u << shi)) | ((uint32_t)v << shi);
                    if (!jin) {
                        d[v] += 2.0 * (wi * wj * stab[mism2(xi2, xj)] - g_old);
                    } else {
                        const int shj = 2 * (p - j);
                        const uint32_t xj2 = (xj & ~(3u << shj)) | ((uint32_t)v << shj);
                        d[v] += wi * wj * stab[mism2(xi2, xj2)] - g_old;
                    }
                    if (rv) {
                        uint32_t rj2 = rj;
                        if (jrin) {
                            const int shr = 2 * (q - j);
                            rj2 = (rj & ~(3u << shr)) | ((uint32_t)(3 - v) << shr);
                        }
                        d[v] += wi * wrj * stab[mism2(xi2, rj2)] - gr_old;
                    }
                }
            }
            if (rv && !jin) {
                // row j (not in C) against the changed RC windows
                for (int jr = r0; jr <= r1; jr++) {
                    const uint32_t rr = rv[jr];
                    const double wr = wgt(w, w - 1 - jr);
                    const int shr = 2 * (q - jr);
                    const double g_old = wj * wr * stab[mism2(xj, rr)];
                    for (int v = 0; v < 4; v++) {
                        if (v == xb)
                            continue;
                        const uint32_t rr2 = (rr & ~(3u << shr)) | ((uint32_t)(3 - v) << shr);
                        d[v] += wj * wr * stab[mism2(xj, rr2)] - g_old;
                    }
                }
            }
        }
    }
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    for (int v = 0; v < 4; v++) {
        double t = d[v];
        for (int off = WAVE / 2; off > 0; off >>= 1)
            t += __shfl_down(t, off, WAVE);
        if (lane == 0)
            red[v][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int k = 0; k < MUT_THREADS / WAVE; k++)
            t += red[threadIdx.x][k];
        out[((int64_t)threadIdx.x * lmax + p) * n + r] = diag[r] + t;
    }
}

// ---------------------------------------------------------------------------
// k_sweep: GkmExplain (hypothetical / perturbation) and ISM deltas
// ---------------------------------------------------------------------------

constexpr int SWEEP_THREADS = 256;

struct SweepArgs {
    const P2 *xplanes;  // [B][nwx]
    const int32_t *xlen;
    const uint8_t *xcodes;  // [B][lmax] (ISM: the reference base at each position)
    int lmax;
    const P2 *yplanes, *yplanes_rc;  // [S][nwy]
    const int32_t *ylen;
    const uint8_t *wtab;
    int wtab_len;
    int B, S, nwx, nwy, l, nm, cmax;  // hits are windows with count <= cmax
    int sv_per_block;
    const double *ta, *tb;  // explain tables [l + 2]
    const double *ec;       // explain: [B][S] effective coefficients
    const double *tab;      // ISM: h[m] [l + 2] (zero past dmax)
    const double *coef, *ds, *dx;  // ISM: coefficients, K_raw(sv, sv), K_raw(x, x)
    const double *dmut;     // ISM: K_raw(mutant, mutant) [4][lmax][B]
    int normalize, rbf;
    double gamma;
    double *out;  // [B][4][lmax], accumulated with FP64 atomics
};

template <int LT, int NMC, int MODE>
__global__ __launch_bounds__(SWEEP_THREADS) void k_sweep(SweepArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    const int b = blockIdx.x;
    const int lx = a.xlen[b];
    const int l = LT > 0 ? LT : a.l;
    const int norient = a.yplanes_rc ? 2 : 1;
    const int nwx = a.nwx, nwy = a.nwy;
    const int lmax = a.lmax;
    // LDS layout
    double *res = (double *)smem;                   // [4][lmax] per-query sums
    double *dm = res + 4 * (size_t)lmax;            // ISM: [lmax] match deltas
    double *da = dm + (MODE == SW_ISM ? lmax : 0);  // ISM: [4][lmax] mismatch deltas
    double *stab = da + (MODE == SW_ISM ? 4 * (size_t)lmax : 0);  // [2][16]
    unsigned long long *hred = (unsigned long long *)(stab + 32);  // [16][4]
    P2 *xs = (P2 *)(hred + 64);
    P2 *ys = xs + nwx;  // [norient][nwy]
    uint8_t *wt = (uint8_t *)(ys + 2 * nwy);

    for (int i = threadIdx.x; i < 4 * lmax; i += SWEEP_THREADS)
        res[i] = 0.0;
    if (MODE == SW_ISM)
        for (int i = threadIdx.x; i < 5 * lmax; i += SWEEP_THREADS)
            dm[i] = 0.0;
    if (threadIdx.x < 16) {
        int t = threadIdx.x;
        stab[t] = t < a.l + 2 ? (MODE == SW_ISM ? a.tab[t] : a.ta[t]) : 0.0;
        stab[16 + t] = (MODE != SW_ISM && t < a.l + 2) ? a.tb[t] : 0.0;
    }
    for (int i = threadIdx.x; i < nwx; i += SWEEP_THREADS)
        xs[i] = a.xplanes[(int64_t)b * nwx + i];
    if (a.wtab)
        for (int i = threadIdx.x; i < a.wtab_len; i += SWEEP_THREADS)
            wt[i] = a.wtab[i];
    __syncthreads();
    if (lx < l)
        return;
    const int sv0 = blockIdx.y * a.sv_per_block;
    const int sv1 = min(a.S, sv0 + a.sv_per_block);
    const uint8_t *xc = MODE == SW_ISM ? a.xcodes + (int64_t)b * lmax : nullptr;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const uint8_t *wtp = a.wtab ? wt : nullptr;
    const int wx = lx - l + 1;

    auto add = [&](int kind, int yb, int pos, double v) {
        if (kind == 0)
            atomicAdd(&res[yb * lmax + pos], v);
        else if (kind == 1)
            atomicAdd(&dm[pos], v);
        else
            atomicAdd(&da[yb * lmax + pos], v);
    };
    // Hits per distribution round: lane = (group, position), groups of l lanes.
    const int G = min(WAVE / l, 32);
    const int grp = lane / l, qpos = lane - grp * l;
    const uint64_t below = (1ull << lane) - 1ull;

    for (int s = sv0; s < sv1; s++) {
        const int ly = a.ylen[s];
        if (ly < l)
            continue;
        const int wy = ly - l + 1;
        for (int i = threadIdx.x; i < norient * nwy; i += SWEEP_THREADS) {
            int o = i / nwy, k = i % nwy;
            ys[o * nwy + k] = (o ? a.yplanes_rc : a.yplanes)[(int64_t)s * nwy + k];
        }
        __syncthreads();
        const double ecs = MODE == SW_ISM ? 1.0 : a.ec[(int64_t)b * a.S + s];
        unsigned long long cnt[NMC];
#pragma unroll
        for (int m = 0; m < NMC; m++)
            cnt[m] = 0;
        const int nd = lx + ly - 2 * l + 1;
        const int nblk = (nd + WAVE - 1) / WAVE;
        // waves take 64 consecutive diagonals of one strand; lanes stay converged
        for (int blk = wave; blk < nblk * norient; blk += SWEEP_THREADS / WAVE) {
            const int o = blk / nblk;
            const int d0 = (blk % nblk) * WAVE;
            const int d = d0 + lane;
            const bool lane_ok = d < nd;
            const int delta = (lane_ok ? d : nd - 1) - (lx - l);
            const int ty_lo = max(0, delta);
            const int ty_hi = min(ly, lx + delta) - l;
            int U0, U1;
            wave_word_range(lx, ly, l, d0, nd, U0, U1);
            const P2 *yo = ys + o * nwy;
            walk_converged<LT>(
                [&](int k) { return xs[k]; }, [&](int u) { return yo[u]; }, nwx, nwy, l, delta,
                lane_ok, ty_lo, ty_hi, U0, U1,
                [&](int base, bool, uint32_t ep, uint32_t ec, P2 yp, P2 yc, uint32_t b0,
                    uint32_t b1, uint32_t b2, uint32_t b3, uint32_t valid) {
                    uint32_t mk[NMC];
                    levels<NMC>(b0, b1, b2, b3, valid, a.nm, mk);
                    uint32_t hits = 0;
#pragma unroll
                    for (int m = 0; m < NMC; m++) {
                        if (MODE == SW_ISM && !wtp)
                            cnt[m] += popc(mk[m]);
                        if (m <= a.cmax)
                            hits |= mk[m];
                    }
                    uint64_t any = __ballot(hits != 0u);
                    while (any) {  // wave-uniform
                        const int rank = __popcll(any & below);
                        const bool chosen = hits != 0u && rank < G;
                        const int nchosen = min(__popcll(any), G);
                        const int j = hits ? __builtin_ctz(hits) : 0;
                        // (lanes without a hit build a record nobody reads; no weights
                        // for them, whose window indexes may lie outside the rows)
                        const HitRec h = make_hit(l, j, base, delta, o, ep, ec, yp, yc, b0, b1,
                                                  b2, b3, hits ? wtp : nullptr, wx, wy);
                        if (MODE == SW_ISM && wtp && chosen) {
#pragma unroll
                            for (int m = 0; m < NMC; m++)
                                cnt[m] += (h.c == m) ? (unsigned long long)h.w : 0ull;
                        }
                        // compact the chosen lanes' records to lanes 0..nchosen-1
                        const int dst = (chosen ? rank : WAVE - 1) * 4;
                        const uint32_t r0 = (uint32_t)h.tx | ((uint32_t)h.c << 16);
                        const uint32_t r1 = h.mis | (h.y0 << 15);
                        const uint32_t r2 = h.y1 | (h.w << 15);
                        const uint32_t c0 = __builtin_amdgcn_ds_permute(dst, (int)r0);
                        const uint32_t c1 = __builtin_amdgcn_ds_permute(dst, (int)r1);
                        const uint32_t c2 = __builtin_amdgcn_ds_permute(dst, (int)r2);
                        // each lane takes one position of one compacted hit
                        const uint32_t s0 = __builtin_amdgcn_ds_bpermute(grp * 4, (int)c0);
                        const uint32_t s1 = __builtin_amdgcn_ds_bpermute(grp * 4, (int)c1);
                        const uint32_t s2 = __builtin_amdgcn_ds_bpermute(grp * 4, (int)c2);
                        if (grp < nchosen) {
                            HitRec g;
                            g.tx = (int)(s0 & 0xFFFFu);
                            g.c = (int)(s0 >> 16);
                            g.mis = s1 & 0x7FFFu;
                            g.y0 = (s1 >> 15) & 0x7FFFu;
                            g.y1 = s2 & 0x7FFFu;
                            g.w = s2 >> 15;
                            if (g.tx + qpos < lx)
                                apply_hit<MODE>(g, qpos, ecs, stab, stab + 16, add);
                        }
                        if (chosen)
                            hits &= hits - 1u;
                        any = __ballot(hits != 0u);
                    }
                });
        }
        if (MODE == SW_ISM) {
            // K_raw(x, s) from the block's histogram, then the F-sweep.
#pragma unroll
            for (int m = 0; m < NMC; m++) {
                unsigned long long v = cnt[m];
                for (int off = WAVE / 2; off > 0; off >>= 1)
                    v += __shfl_down(v, off, WAVE);
                if (lane == 0)
                    hred[m * 4 + wave] = v;
            }
            __syncthreads();
            double kraw = 0.0;
            for (int m = 0; m < a.nm; m++) {
                unsigned long long v = 0;
                for (int w = 0; w < SWEEP_THREADS / WAVE; w++)
                    v += hred[m * 4 + w];
                kraw += stab[m] * (double)v;
            }
            const double cs = a.coef[s];
            const double dss = a.ds[s];
            const double zr = a.normalize ? fmax(sqrt(a.dx[b] * dss), 1e-10) : 1.0;
            double fref = kraw / zr;
            if (a.rbf)
                fref = exp(a.gamma * (fref - 1.0));
            for (int p = threadIdx.x; p < lx; p += SWEEP_THREADS) {
                const int xb_ = xc[p] & 3;
                const double dmp = dm[p];
                for (int v = 0; v < 4; v++) {
                    if (v == xb_)
                        continue;
                    const double dk = dmp + da[v * lmax + p];
                    double k = kraw + dk;
                    if (a.normalize)
                        k = k / fmax(sqrt(a.dmut[((int64_t)v * lmax + p) * a.B + b] * dss), 1e-10);
                    if (a.rbf)
                        k = exp(a.gamma * (k - 1.0));
                    res[v * lmax + p] += cs * (k - fref);
                    da[v * lmax + p] = 0.0;
                }
                dm[p] = 0.0;
            }
        }
        __syncthreads();
    }
    double *ob = a.out + (int64_t)b * 4 * lmax;
    for (int i = threadIdx.x; i < 4 * lx; i += SWEEP_THREADS) {
        int v = i / lx, p = i % lx;
        double r = res[v * lmax + p];
        if (r != 0.0)
            atomicAdd(&ob[v * lmax + p], r);
    }
}

// ---------------------------------------------------------------------------
// Exact scoring of kernels without RBF.  A decision value is linear in the
// mismatch histograms: f(x) = sum_m h[m] S_m / sqrt(K_raw(x, x)) + b with
// S_m = sum_s w_s N_s[m] and w_s = coef_s / sqrt(K_raw(s, s)) (svm.py:43-61
// normalises each pair by sqrt(dx ds), the same product).  The SV weights are
// fixed point, wfix_s = w_s 2^shift rounded (|wfix_s| <= 2^61), and S_m is an
// exact 128-bit integer sum, so a row's value does not depend on the order in
// which its pairs, window chunks and SV tiles are added: the l-mer index with
// any chunking, the pairwise kernel, a row alone or in any batch give the same
// bits.  fixed_score is the one place the sums become a float64.
// ---------------------------------------------------------------------------

typedef __int128 i128;

__device__ __forceinline__ double i128_to_double(i128 v)
{
    const long long hi = (long long)(v >> 64);
    const unsigned long long lo = (unsigned long long)v;
    return fma((double)hi, 18446744073709551616.0, (double)lo);
}

__device__ __forceinline__ i128 wave_sum_i128(i128 v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const long long hi = __shfl_xor((long long)(v >> 64), o, WAVE);
        const unsigned long long lo = __shfl_xor((unsigned long long)v, o, WAVE);
        v += ((i128)hi << 64) + (i128)lo;
    }
    return v;
}

__device__ __noinline__ double fixed_score(const i128 *S, int nm, const double *tab, int shift,
                                           double dx, int normalize, double bias)
{
    double acc = 0.0;
    for (int m = 0; m < nm; ++m)
        acc = fma(tab[m], i128_to_double(S[m]), acc);
    acc = ldexp(acc, -shift);
    if (normalize)
        acc = dx > 0.0 ? __ddiv_rn(acc, __dsqrt_rn(dx)) : 0.0;
    return acc + bias;
}

// bsum[lv] += sum_{r < n} wfix[r] * count(r, lv) for lv < nm over a block (all
// its threads call it), counts laid out [lv / PER][n] in 32-bit words of PER
// counts (4: 8-bit, 2: 16-bit, 1: 32-bit).  One pass over the SVs per 4 levels.
// Counts below 2^16 (PER 4, 2) go through float64 FMAs, exactly: wfix =
// whi 2^31 + wlo with |whi| <= 2^30, wlo < 2^31, so each product with a count is an
// integer below 2^47 and a thread's at most 64 of them sum below 2^53 without
// rounding; the two sums become int64 and then the 128-bit level sum.  32-bit
// counts take 32 x 32 -> 64-bit integer products (|wfix >> 32| <= 2^29).  It runs once per block after
// the walk, whose registers it reuses.  U: SV weights loaded per round.
template <int U, int PER>
__device__ __forceinline__ void fold_level_sums(const uint32_t *cnt, int n,
                                                const long long *wfix, int nm, i128 *bsum,
                                                i128 *red)
{
    constexpr int bits = 32 / PER;
    constexpr uint32_t mask = PER == 1 ? 0xFFFFFFFFu : (1u << bits) - 1u;
    constexpr bool fp = PER > 1;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    for (int l0 = 0; l0 < nm; l0 += 4) {
        i128 acc[4] = {0, 0, 0, 0};
        double fhi[4] = {0.0, 0.0, 0.0, 0.0}, flo[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r0 = threadIdx.x; r0 < n; r0 += U * (int)blockDim.x) {
            long long wv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = r0 + u * (int)blockDim.x;
                wv[u] = wfix[r < n ? r : 0];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = r0 + u * (int)blockDim.x;
                if (r >= n)
                    break;
                const long long w = wv[u];
                if constexpr (fp) {
                    const double dhi = (double)(int)(w >> 31);
                    const double dlo = (double)(uint32_t)(w & 0x7FFFFFFFll);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int lv = l0 + k;
                        if (lv < nm) {
                            const double c = (double)((cnt[(lv / PER) * n + r] >>
                                                       (bits * (lv % PER))) & mask);
                            fhi[k] = fma(dhi, c, fhi[k]);
                            flo[k] = fma(dlo, c, flo[k]);
                        }
                    }
                } else {
                    const uint32_t wlo = (uint32_t)w;
                    const long long whi = w >> 32;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int lv = l0 + k;
                        if (lv < nm) {
                            const uint32_t c = cnt[lv * n + r];
                            const unsigned long long plo = (unsigned long long)wlo * c;
                            const long long phi = whi * (long long)c;
                            acc[k] += ((i128)phi << 32) + (i128)plo;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (l0 + k < nm) {
                i128 v = acc[k];
                if constexpr (fp)
                    v = ((i128)(long long)fhi[k] << 31) + (i128)(long long)flo[k];
                v = wave_sum_i128(v);
                if (lane == 0)
                    red[wave] = v;
                __syncthreads();
                if (threadIdx.x == 0) {
                    i128 t = bsum[l0 + k];
                    for (int q = 0; q < (int)(blockDim.x / WAVE); ++q)
                        t += red[q];
                    bsum[l0 + k] = t;
                }
                __syncthreads();
            }
        }
    }
}

// The same, out of line: for the kernel variants where two inlined copies
// would add to the walk's register pressure (weighted 16-bit counts).
template <int PER>
__device__ __noinline__ void fold_level_sums_call(const uint32_t *cnt, int n,
                                                  const long long *wfix, int nm, i128 *bsum,
                                                  i128 *red)
{
    fold_level_sums<2, PER>(cnt, n, wfix, nm, bsum, red);
}

// ---------------------------------------------------------------------------
// k_index_tile: the same histograms through an l-mer index of the SV windows
// ---------------------------------------------------------------------------

constexpr int INDEX_THREADS = 1024;  // 16 waves share one LDS tile (one block per CU)
constexpr int INDEX_PROBES = 1;      // probes per lane per group (64 INDEX_PROBES masks)
constexpr int INDEX_WALK = 12;       // entries per probe per round of the lane walk
static_assert(INDEX_WALK <= GKM_INDEX_PAD, "the first round reads INDEX_WALK entries ahead");
constexpr int INDEX_PASSES = 4;      // 64-entry passes of a balanced walk issued together
constexpr int RECORD_IDS = 15;       // ids a 32-byte bucket record holds (unweighted, short)
constexpr int RECORD_IDS_BIG = 11;   // ids it holds beside the start and count (long)
constexpr int RECORD_IDS_W = 8;      // (id, weight) pairs held in one (weighted index)
constexpr int RECORD_SINKS = WAVE;   // sink words past an index tile's LDS counts (one a lane)

// Bucket records, 8 words (16 halfwords), one per key.
// RK 1 (unweighted): halfword 0 = min(count, 0xFFFF); a bucket of at most
// RECORD_IDS entries has its ids in halfwords 1..count (no start: nothing past
// them) and 0xFFFF in halfwords count + 1..15 (the sink mark: a lane adds
// every slot unconditionally, at min(slot, ns + lane) -- ns the tile's SV count,
// the lane's own LDS word past the tile's counts, so the padding adds of a
// wave meet on no word and no bank); a longer one its first RECORD_IDS_BIG ids
// in halfwords 1..11, its start in word 6 and its count in word 7 (the rest at
// ids[start + 11 ..]).
// RK 2 (weighted): word 0 start, word 1 count, RECORD_IDS_W uint16 ids in
// halfwords 4..11 and their uint8 weights in bytes 24..31.
// RK 3: one 64-byte record per l-mer for unweighted indexes of at most 4 SV
// tiles, whose buckets are consecutive in ids: word 0 the start of the l-mer's
// run, word 1 the tiles' counts as bytes (255: saturated, read the offsets),
// then the run's first LREC_IDS ids (uint16; wd = words 2..15).
constexpr int LREC_IDS = 28;
constexpr int LREC_MAX_TILES = 4;

// ids a record can hold (the lane walk's unrolled reads)
template <int RK>
constexpr int record_ids() { return RK == 2 ? RECORD_IDS_W : RK == 1 ? RECORD_IDS : 0; }
__device__ __forceinline__ uint32_t lrec_id(const uint32_t *wd, int k)
{
    return (wd[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
}
// id k of a record, w = its 8 words
template <int RK>
__device__ __forceinline__ uint32_t rec_id(const uint32_t *w, int k)
{
    constexpr int H = RK == 2 ? 4 : 1;  // halfword of id 0
    return (w[(k + H) >> 1] >> (16 * ((k + H) & 1))) & 0xFFFFu;
}
__device__ __forceinline__ uint32_t rec_val(const uint32_t *w, int k)
{
    return (w[6 + (k >> 2)] >> (8 * (k & 3))) & 0xFFu;
}
// A record's bucket: its entry count, the ids it holds, and where the rest
// begin in the ids array (e0 + n_in; e0 = the bucket's start).
struct RecBucket {
    int c, n_in;
    uint32_t e0;
};
template <int RK>
__device__ __forceinline__ RecBucket rec_bucket(uint4 ra, uint4 rb)
{
    if constexpr (RK == 1) {
        const uint32_t h0 = ra.x & 0xFFFFu;
        const bool big = h0 > (uint32_t)RECORD_IDS;
        return {big ? (int)rb.w : (int)h0, big ? RECORD_IDS_BIG : (int)h0, big ? rb.z : 0u};
    } else {
        return {(int)ra.y, min((int)ra.y, RECORD_IDS_W), ra.x};
    }
}

// The layout of the bucket records gkm_index_explain is passed: none (0), 32
// bytes per key (1: unweighted, 2: weighted) or, unweighted with at most 4
// tiles, 64 bytes per l-mer (3).
static inline int record_kind(const uint32_t *records, const uint8_t *vals, int n_tiles)
{
    return records == nullptr ? 0 : vals ? 2 : n_tiles <= LREC_MAX_TILES ? 3 : 1;
}

// A bucket's [start, end) offsets in one 8-byte load (4-byte aligned).
struct __attribute__((aligned(4))) bounds2 {
    uint32_t x, y;
};

// In silico mutagenesis through the index (gkm_index_ism): the rows are the
// query's windows one by one (window rows) or the windows covering position p
// with its base substituted (new rows 3 p + alt), and each block's counts
// finish per SV instead of going out.
struct IsmArgs {
    const double *tab;    // h[m], m < nm
    const double *svp;    // [S][4] per SV: K_raw(x, s), its finished value, coefficient,
                          // K_raw(s, s)
    double *R;            // window rows: C [windows][S] out (each window's K_raw part);
                          // new rows: R [positions][S] in (the windows covering p)
    const double *dmut;   // K_raw(mutant, mutant) at ((v lmax + p) stride), or NULL
    const uint8_t *codes; // the query's base codes
    double *part;         // new rows: [pairs][n_tiles] sums over the tile's SVs
    int lmax, dmut_stride, wx;  // wx: the query's windows (window weights)
    int old_rows, normalize, rbf;  // old_rows: the window rows
    double gamma;
};

struct IndexArgs {
    const uint32_t *xwin;    // [B][Wx] packed query windows (PAD past each row's)
    const int32_t *xpos;     // [B][Wx] position of each window in its row, or NULL
    const int32_t *xlen;     // [B]
    const uint32_t *offsets; // [4^l * n_tiles + 1], key = lmer * n_tiles + tile
    const uint16_t *ids;     // [entries] SV of each entry, relative to its tile
    const uint8_t *vals;     // [entries] window weight of each entry, or NULL
    const uint32_t *records; // [record_tables][keys][8] bucket records (RK 1 or 2), or NULL
    const uint32_t *masks;   // [n_masks] XOR masks of the neighbourhood
    const uint8_t *levels;   // [n_masks] mismatch count of each mask | its record table << 4
    int record_tables;       // tables in records (table q: digits 0 and q of the key swapped)
    const uint8_t *wtab;     // weight by distance, or NULL
    void *partial;           // [nchunks][B][S][nm] per-chunk counts (PACK 2, 4: uint16),
                             // or with wfix [nchunks][B][n_tiles][nm] int128 level sums
    const long long *wfix;   // [S] fixed-point SV weights (exact scoring), or NULL
    const int32_t *pair_first;  // persistent form: [B] each row's first pair (exclusive
                                // prefix sums of the rows' chunk counts)
    int *queue;                       // 8 item counters (zero at launch)
    int n_pairs;
    int B, Wx, n_masks, l, nm, S, tile, n_tiles, nchunks, chunk_windows;
    int sub_tile;            // PACK 4: SVs per pass of the 16-bit walk
    int64_t col_min, row_base;
    IsmArgs ism;             // the ISM form's epilogue (ISM instances only)
};

// ISM epilogue over n SVs of a block's tile from SV s_first (counts laid out
// [lv / PER][n], PER counts per word): K_raw's part from the block's windows,
// k = sum_m h[m] N[m] in level order.  Window rows store it as C[i][s]; new rows
// (p, v) finish K_raw(x, s) - R[p][s] + k like a score (normalised by the
// mutant's self-kernel zmut and the SV's, RBF) and return this thread's sum
// of coef_s (f(mutant, s) - f(x, s)) (ism.py:49-52 rescoring, restated per SV).
// Arguments by value: no copy of the kernel's argument block to the stack.
template <int PER>
__device__ __noinline__ double ism_fold(const double *tab, const double *svp, double *R,
                                        int old_rows, int normalize, int rbf, double gamma,
                                        int p, double zmut, const uint32_t *cnt, int n,
                                        int s_first, int nm, int S)
{
    constexpr int BITS = 32 / PER;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += INDEX_THREADS) {
        double k = 0.0;
        for (int lv = 0; lv < nm; ++lv) {
            uint32_t c = cnt[(lv / PER) * n + i];
            if (PER > 1)
                c = (c >> (BITS * (lv % PER))) & ((1u << BITS) - 1u);
            k += tab[lv] * (double)c;
        }
        const int s = s_first + i;
        if (old_rows) {
            R[(int64_t)p * S + s] = k;
            continue;
        }
        const double *q = svp + (int64_t)s * 4;
        double kn = q[0] - R[(int64_t)p * S + s] + k;
        if (normalize)
            kn = kn / fmax(sqrt(zmut * q[3]), 1e-10);
        if (rbf)
            kn = exp(gamma * (kn - 1.0));
        acc += q[2] * (kn - q[1]);
    }
    return acc;
}

// Block = (SV tile, chunk of query windows, query row).  Its LDS holds the
// tile's counts [tile][nm]; each wave takes a query window and its lanes
// 64 * INDEX_PROBES neighbour masks at a time: v = w ^ mask is an l-mer within
// the mask's mismatch count of w, and every window of a tile SV equal to v
// adds (its weight times the query window's) to N[level] of that pair -- the
// pairs kernels/lmer_index_cpu.py:23-37 (_add_window) visits.  The tile's
// counts go out with plain stores to the chunk's slice of `partial`.
// PACK: counts per 32-bit LDS word.  1: uint32 counts (weighted).  2: two
// 16-bit counts (unweighted; the host caps the window chunk so no count of a
// block can reach 2^16), which doubles the SVs a tile holds and halves the
// probes.  4: four 8-bit counts -- twice the SVs again -- made exact by a
// check: an add that carries out of a byte lowers the word's byte sum below
// the number of adds (a carry turns 255 + 1 into 0 and adds at most 1 to the
// next byte), so the block compares the byte sum of its words with its count
// of adds, and where they differ (repeats) it walks its chunk again with 16-bit
// counts, a part of the tile per pass.
// Memory: the kernel is bound by random cache lines -- per neighbour l-mer the
// bucket bounds and the bucket's ids.  RK: the bucket records (start, count,
// the first ids and weights) put both in one 32-byte load, and the n_tiles
// records of an l-mer share a line, which the tile blocks of a chunk fetch
// together (same XCD, see below).  One block fills a CU's LDS, so a CU runs 16
// waves; each wave keeps the records (or bounds) of its next two groups in
// flight while it walks the current group's buckets (two register buffers,
// the loop unrolled by two, so a buffer is read two steps after its load).
// When every bucket of the group ends within INDEX_WALK entries past its
// record, each lane walks its own bucket: the record's ids, then the rest with
// loads issued before the next fetch.  Larger buckets (repeats) are walked
// wave-balanced: the wave's buckets laid end to end, each pass's entries
// located by a binary search over the prefix sums.
// Block order: the XCDs take blocks round-robin by linear id (gridDim.x is a
// multiple of 8), so block p runs on XCD p % 8.  The blocks of one window
// chunk -- one per SV tile, probing the same offsets lines -- are given ids
// congruent mod 8 and so share an XCD's L2: p -> q = p / 8, tile = q % n_tiles,
// chunk = (q / n_tiles) * 8 + p % 8.  Ids past the last chunk, and chunks past
// the row's last window, exit at once.
// One work item of the index query: window chunk `chunk` of row b against SV
// tile tile_id; with wfix its level sums go to slot `slot` of the psum output.
// The LDS count of one level in the index query: word offset of the level's
// counts (level / counts per word times the SVs counted) and the increment
// (or shift) that adds to its field.
struct CntKey {
    int off;
    uint32_t inc;
};

template <int PACK, int RK, bool W16, bool ISM = false, bool DIAG = false>
__device__ __forceinline__ void index_block(const IndexArgs &a, int b, int chunk, int tile_id,
                                            int64_t slot)
{
    extern __shared__ __align__(16) uint8_t smem[];
    uint32_t *cnt = (uint32_t *)smem;
    const int wx = a.xlen[b] - a.l + 1;
    const int wxw = ISM ? a.ism.wx : wx;  // the window count the window weights take
    const int i0 = chunk * a.chunk_windows;
    if (i0 >= wx)
        return;  // no windows of this row: the reduce skips this chunk
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int wps = (a.nm + PACK - 1) / PACK;  // LDS words per SV
    const int s_base = tile_id * a.tile;
    const int ns = min(a.tile, a.S - s_base);  // this tile's SVs
    const int ncnt = ns * wps;
    for (int i = threadIdx.x; i < ncnt; i += INDEX_THREADS)
        cnt[i] = 0u;
    // exact scoring (wfix): the block's level sums S_m over its tile's SVs
    __shared__ i128 bsum[16], red128[INDEX_THREADS / WAVE];
    if (threadIdx.x < 16)
        bsum[threadIdx.x] = 0;
    __syncthreads();
    // bsum[lv] += sum over SVs s0 .. s0 + n - 1 of the tile of wfix times the
    // count of level lv, counts laid out [lv / per][n] with per counts per word
    double iacc = 0.0;  // ISM: this thread's sum over its SVs
    int ism_p = b;       // ISM: the row's position and mutant self-kernel
    double ism_z = 1.0;
    if constexpr (ISM) {
        if (!a.ism.old_rows) {
            ism_p = b / 3;
            const int alt = b - 3 * ism_p, ref = a.ism.codes[ism_p] & 3;
            const int v = alt + (alt >= ref ? 1 : 0);
            if (a.ism.normalize)
                ism_z = a.ism.dmut[((int64_t)v * a.ism.lmax + ism_p) * a.ism.dmut_stride];
        }
    }
    auto fold = [&](int n, int s0, auto per) {  // per: std::integral_constant
        constexpr int PER = decltype(per)::value;
        if constexpr (ISM)
            iacc += ism_fold<PER>(a.ism.tab, a.ism.svp, a.ism.R, a.ism.old_rows,
                                  a.ism.normalize, a.ism.rbf, a.ism.gamma, ism_p, ism_z, cnt, n,
                                  s_base + s0, a.nm, a.S);
        else if constexpr (W16)
            fold_level_sums_call<PER>(cnt, n, a.wfix + s_base + s0, a.nm, bsum, red128);
        else
            fold_level_sums<12, PER>(cnt, n, a.wfix + s_base + s0, a.nm, bsum, red128);
    };
    // exact scoring and ISM finish the counts in the block
    const bool fused = ISM || a.wfix != nullptr;
    using per4 = std::integral_constant<int, 4>;
    using per2 = std::integral_constant<int, 2>;
    using per1 = std::integral_constant<int, 1>;
    auto emit = [&]() {  // the level sums out: slot * nm + level
        if constexpr (ISM) {  // new rows: the tile's sum to part[slot] (fixed order)
            if (a.ism.old_rows)
                return;
            __shared__ double redd[INDEX_THREADS / WAVE];
            double v = iacc;
            for (int o = WAVE / 2; o > 0; o >>= 1)
                v += __shfl_xor(v, o, WAVE);
            if (lane == 0)
                redd[wave] = v;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0.0;
                for (int k = 0; k < INDEX_THREADS / WAVE; ++k)
                    t += redd[k];
                a.ism.part[slot] = t;
            }
        } else if (threadIdx.x < a.nm) {
            ((i128 *)a.partial)[slot * a.nm + threadIdx.x] = bsum[threadIdx.x];
        }
    };
    const uint32_t nl = 1u << (2 * a.l);
    const int i1 = min(wx, i0 + a.chunk_windows);
    // DIAG (col_min >= 0: the Gram's upper triangle): entries of SVs below
    // lo_col (relative to the tile, clamped to -1 .. 2^16) are skipped
    const int lo_col = DIAG ? (int)min(max(a.row_base + b + a.col_min - s_base, (int64_t)-1),
                                       (int64_t)0x10000)
                            : -1;
    const int64_t off = (((int64_t)chunk * a.B + b) * a.S + s_base) * a.nm;
    // One walk of the block's (window, mask group) pairs, every bucket entry
    // handed to count(sid, key, weight), key = keyf(level) computed once per
    // probe (the LDS word offset and increment of that level).  Probe p of
    // lane t in group g takes mask m = (INDEX_PROBES g + p) 64 + t, neighbour
    // v = w ^ mask[m], whose bucket is [e0, e0 + c).  Returns this lane's
    // number of entries walked (every one of them was handed to count once).
    // plain (1): count(sid, key, w) is atomicAdd(&cnt[key.off + sid], key.inc),
    // so an entry slot a lane does not hold can add 0 to a word of its own
    // instead of being branched around: the record and first-round counts
    // then run without exec masks (fewer scalar instructions per step).
    // sink (2, one count word per SV, key.off 0): every record slot is added,
    // with the probe's increment (0 for a probe past the masks), at
    // min(slot, ns + lane) -- the slots past a short bucket's ids hold
    // 0xFFFF, so they land on the lane's own word past the counts (which
    // nothing reads), and a long bucket's start and count slots are forced
    // there: two instructions per slot, no compare
    const int safe_word = min(lane, ncnt - 1);
    const uint32_t sink_lane = (uint32_t)(ns + lane);
    auto walk = [&](auto keyf, auto count, auto plain) -> uint32_t {
        constexpr int PMODE = (int)decltype(plain)::value;
        constexpr bool PLAIN = PMODE == 1;
        using Key = decltype(keyf(0));
        const int n_groups = (a.n_masks + INDEX_PROBES * WAVE - 1) / (INDEX_PROBES * WAVE);
        uint32_t walked = 0;
        // two (window, mask group) pairs' records in flight: the pair walked
        // now was fetched two steps ago, into the buffer (A or B) it is read from.
        // ISM rows hold 1..l windows: their pairs are dealt to the waves in turn
        // (pair k: window k / n_groups, group k % n_groups), so every wave has
        // work; otherwise a wave walks every group of one window, then the
        // window 16 on.
        constexpr bool FLAT = ISM;
        constexpr int WSTEP = FLAT ? 1 : INDEX_THREADS / WAVE;  // a wave's next window
        int ci = FLAT ? i0 + wave / n_groups : i0 + wave;  // the pair being walked
        int cg = FLAT ? wave % n_groups : 0;
        int ni = ci, ng = cg;  // the pair being fetched
        // (i, g) to this wave's next pair; true when the window changed
        auto advance = [&](int &i, int &g) -> bool {
            if constexpr (FLAT) {
                g += INDEX_THREADS / WAVE;
                bool moved = false;
                while (g >= n_groups) {
                    g -= n_groups;
                    ++i;
                    moved = true;
                }
                return moved;
            } else {
                if (++g == n_groups) {
                    g = 0;
                    i += INDEX_THREADS / WAVE;
                    return true;
                }
                return false;
            }
        };
        uint4 nra[INDEX_PROBES][2], nrc[INDEX_PROBES][2];  // record (RK 1, 2), or (start, end)
        int nlva[INDEX_PROBES], nlvc[INDEX_PROBES];
        // Loads complete in order, so a wait for a load also waits for every
        // older one: what a fetch needs at once -- its masks and levels, its
        // window's l-mer -- is loaded a fetch (a window) ahead, before the
        // records then issued, and waiting for it leaves those in flight.
        uint32_t pm[INDEX_PROBES], pl[INDEX_PROBES];  // the next fetch's masks, levels
        // (loads without branches around them -- past the end they read the last
        // entry, which no probe uses -- so the compiler can count them)
        auto load_masks = [&](int g) {
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p) {
                const int t = min((g * INDEX_PROBES + p) * WAVE + lane, a.n_masks - 1);
                pm[p] = a.masks[t];
                pl[p] = a.levels[t];
            }
        };
        auto load_window = [&](int i) -> uint32_t {  // windows past the chunk: no probes
            return a.xwin[(int64_t)b * a.Wx + min(i, a.Wx - 1)];
        };
        uint32_t fw = load_window(ni), fw_next = load_window(ni + WSTEP);
        // the walked window's position in its row (window weights), a window ahead
        auto load_pos = [&](int i) -> int {
            return a.xpos[(int64_t)b * a.Wx + min(i, a.Wx - 1)];
        };
        int cpos = 0, cpos_next = 0;
        if (a.xpos) {
            cpos = load_pos(ci);
            cpos_next = load_pos(ci + WSTEP);
        }
        load_masks(cg);
        // record byte offsets in 32 bits (the host keeps the record tables
        // below 4 GiB): table q at q * tbytes, key (l-mer, tile) at
        // l-mer * n_tiles * 32 + tile * 32
        const uint32_t tbytes = nl * (uint32_t)a.n_tiles * 32u;
        const uint32_t vbytes = (uint32_t)a.n_tiles * 32u;
        const uint32_t tile_bytes = (uint32_t)tile_id * 32u;
        const uint32_t last_table = (uint32_t)(a.record_tables - 1);
        auto fetch = [&](uint4 (&nr)[INDEX_PROBES][2], int (&nlv)[INDEX_PROBES]) {
            const uint32_t w = ni < i1 ? fw : nl;  // PAD past the chunk
            uint32_t mk[INDEX_PROBES], lb[INDEX_PROBES];
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p) {
                mk[p] = pm[p];
                lb[p] = pl[p];
            }
            if constexpr (FLAT) {
                int ti = ni, tg = ng;
                advance(ti, tg);
                load_masks(tg);
            } else {
                load_masks(ng + 1 == n_groups ? 0 : ng + 1);
            }
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p) {
                const int t = (ng * INDEX_PROBES + p) * WAVE + lane;
                // every lane loads (a lane without a probe reads key 0 and is
                // marked by level -1): no branch around the loads, so the
                // compiler can count them and wait for this fetch two steps on
                // without draining the fetches after it
                const bool ok = w < nl && t < a.n_masks;
                {
                    const uint32_t v = ok ? w ^ mk[p] : 0u;
                    const uint32_t lvb = lb[p];  // level | record table << 4
                    nlv[p] = ok ? (int)(lvb & 15u) : -1;
                    if (RK) {
                        // record table q holds the records with digits 0 and q of
                        // the l-mer swapped, so the two masks of a pair (differing
                        // in bit 0 of digit q) read one 64-byte-aligned pair of
                        // records, a 128-byte line, in one instruction
                        const uint32_t q = min(lvb >> 4, last_table);
                        const uint32_t x = (v ^ (v >> (2u * q))) & 3u;
                        const uint32_t vq = v ^ x ^ (x << (2u * q));
                        const uint32_t boff = q * tbytes + vq * vbytes + tile_bytes;
                        const uint4 *r = (const uint4 *)((const char *)a.records + boff);
                        nr[p][0] = r[0];
                        nr[p][1] = r[1];
                    } else {  // the bucket's bounds, its count taken when it is walked
                        const uint32_t key = v * (uint32_t)a.n_tiles + (uint32_t)tile_id;
                        const bounds2 r = *(const bounds2 *)(a.offsets + key);
                        nr[p][0].x = r.x;
                        nr[p][0].y = r.y;
                    }
                }
            }
            if constexpr (FLAT) {
                const int prev = ni;
                if (advance(ni, ng)) {
                    fw = ni == prev + 1 ? fw_next : load_window(ni);
                    fw_next = load_window(ni + 1);
                }
            } else if (++ng == n_groups) {
                ng = 0;
                ni += INDEX_THREADS / WAVE;
                fw = fw_next;
                fw_next = load_window(ni + INDEX_THREADS / WAVE);
            }
        };
        fetch(nra, nlva);
        fetch(nrc, nlvc);
        auto step = [&](uint4 (&nrb)[INDEX_PROBES][2], int (&nlvb)[INDEX_PROBES]) {
            const int i = ci;
            uint32_t rw[INDEX_PROBES][8], rb0[INDEX_PROBES];
            int c[INDEX_PROBES], lv[INDEX_PROBES], n_in[INDEX_PROBES];
            Key key[INDEX_PROBES];
            uint32_t sinc[INDEX_PROBES];  // sink mode: the probe's increment
            bool small = true;
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p) {
                rw[p][0] = nrb[p][0].x;
                rw[p][1] = nrb[p][0].y;
                rw[p][2] = nrb[p][0].z;
                rw[p][3] = nrb[p][0].w;
                rw[p][4] = nrb[p][1].x;
                rw[p][5] = nrb[p][1].y;
                rw[p][6] = nrb[p][1].z;
                rw[p][7] = nrb[p][1].w;
                RecBucket bk;
                if constexpr (RK != 0)
                    bk = rec_bucket<RK>(nrb[p][0], nrb[p][1]);
                else
                    bk = {(int)(nrb[p][0].y - nrb[p][0].x), 0, nrb[p][0].x};
                c[p] = nlvb[p] < 0 ? 0 : bk.c;
                n_in[p] = nlvb[p] < 0 ? 0 : bk.n_in;  // ids in the record
                rb0[p] = bk.e0 + (uint32_t)n_in[p];    // the rest of the bucket from here
                lv[p] = max(nlvb[p], 0);  // (nlvb is the next fetch's once fetched)
                key[p] = keyf(lv[p]);
                if constexpr (PMODE == 2)
                    sinc[p] = nlvb[p] < 0 ? 0u : key[p].inc;
                walked += (uint32_t)c[p];
                // a group whose buckets' rests all fit INDEX_WALK entries takes
                // them from the loads below; a longer rest (repeats) is walked
                // balanced
                small = small && c[p] - n_in[p] <= INDEX_WALK;
            }
            // the rests, issued before the fetch below, so waiting for them
            // leaves that fetch in flight.  A lane with a rest reads INDEX_WALK
            // entries from its first (up to INDEX_WALK - 1 past the bucket's end:
            // the ids array holds INDEX_WALK padding entries past its last
            // bucket), the others the array's first entries -- one address per
            // probe, the entries at immediate offsets from it, no branch around
            // a load
            int sid0[INDEX_PROBES][INDEX_WALK];
            uint32_t wv0[INDEX_PROBES][INDEX_WALK];
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p) {
                const uint32_t ob = small && n_in[p] < c[p] ? rb0[p] : 0u;
                const uint16_t *op = a.ids + ob;
                const uint8_t *ov = a.vals ? a.vals + ob : nullptr;
#pragma unroll
                for (int k = 0; k < INDEX_WALK; ++k) {
                    sid0[p][k] = (int)op[k];
                    if (PACK == 1 || W16)
                        wv0[p][k] = ov ? ov[k] : 1u;
                }
            }
            fetch(nrb, nlvb);  // the pair two ahead, into the registers just read
            const int pos = a.xpos ? cpos : i;  // loaded a window ahead
            if constexpr (FLAT) {
                const int prev = ci;
                if (advance(ci, cg) && a.xpos) {
                    cpos = ci == prev + 1 ? cpos_next : load_pos(ci);
                    cpos_next = load_pos(ci + 1);
                }
            } else if (++cg == n_groups) {
                cg = 0;
                ci += INDEX_THREADS / WAVE;
                if (a.xpos) {
                    cpos = cpos_next;
                    cpos_next = load_pos(ci + INDEX_THREADS / WAVE);
                }
            }
            const uint32_t aw = a.wtab ? window_weight(a.wtab, wxw, pos) : 1u;
            // the ids in the records
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p)
#pragma unroll
                for (int k = 0; k < record_ids<RK>(); ++k) {
                    if constexpr (PMODE == 2) {
                        static_assert(RK == 1, "sink slots are a property of RK 1 records");
                        uint32_t id = rec_id<RK>(rw[p], k);
                        if (k >= RECORD_IDS_BIG)
                            id |= c[p] > RECORD_IDS ? 0xFFFFu : 0u;
                        count((int)min(id, sink_lane), Key{0, sinc[p]}, 1u);
                    } else if constexpr (PLAIN) {
                        const bool in = k < n_in[p];
                        count(in ? (int)rec_id<RK>(rw[p], k) : safe_word,
                              in ? key[p] : Key{0, 0u}, 1u);
                    } else if (k < n_in[p]) {
                        count((int)rec_id<RK>(rw[p], k), key[p],
                              PACK > 1 && !W16 ? 1u : aw * (RK == 2 ? rec_val(rw[p], k) : 1u));
                    }
                }
            if (__ballot(!small) == 0) {
                // the rests, from the loads above: entry k of probe p's rest for
                // k < c - n_in; the round stops at the first k past every
                // lane's rest (few buckets reach past their records).  No loads
                // here beyond those: a loop of loads in this path would make
                // the compiler drain the fetches in flight at every step.
                int d[INDEX_PROBES], dtop = 0;
#pragma unroll
                for (int p = 0; p < INDEX_PROBES; ++p) {
                    d[p] = c[p] - n_in[p];
                    dtop = max(dtop, d[p]);
                }
#pragma unroll
                for (int k = 0; k < INDEX_WALK; ++k) {
                    if (__ballot(k < dtop) == 0)
                        break;
#pragma unroll
                    for (int p = 0; p < INDEX_PROBES; ++p) {
                        if constexpr (PLAIN) {
                            const bool in = k < d[p];
                            count(in ? sid0[p][k] : safe_word, in ? key[p] : Key{0, 0u}, 1u);
                        } else if (k < d[p]) {
                            count(sid0[p][k], key[p], PACK > 1 && !W16 ? 1u : aw * wv0[p][k]);
                        }
                    }
                }
                return;
            }
#pragma unroll
            for (int p = 0; p < INDEX_PROBES; ++p) {
                // balanced walk of this probe's rests: laid end to end (exclusive
                // prefix sums) and walked 64 entries per pass: lane j of a pass takes
                // entry base + j, in the rest of the last lane r with excl[r] <= base + j
                const int rest = c[p] - n_in[p];
                int incl = rest;
                for (int o = 1; o < WAVE; o <<= 1) {
                    const int u = __shfl_up(incl, o, WAVE);
                    incl += lane >= o ? u : 0;
                }
                const int excl = incl - rest;
                const int total = __shfl(incl, WAVE - 1, WAVE);
                // INDEX_PASSES passes at a time: their searches, gathers and id
                // loads are independent, so their latencies overlap
                for (int base = 0; base < total; base += INDEX_PASSES * WAVE) {
                    int pr[INDEX_PASSES];
#pragma unroll
                    for (int u = 0; u < INDEX_PASSES; ++u)
                        pr[u] = 0;
#pragma unroll
                    for (int step = WAVE / 2; step > 0; step >>= 1) {
#pragma unroll
                        for (int u = 0; u < INDEX_PASSES; ++u) {
                            const int ex = __shfl(excl, pr[u] + step, WAVE);
                            pr[u] += ex <= base + u * WAVE + lane ? step : 0;
                        }
                    }
                    uint32_t e[INDEX_PASSES];
                    int plv[INDEX_PASSES], sid[INDEX_PASSES];
#pragma unroll
                    for (int u = 0; u < INDEX_PASSES; ++u) {
                        const int j = base + u * WAVE + lane;
                        const uint32_t pe0 = (uint32_t)__shfl((int)rb0[p], pr[u], WAVE);
                        plv[u] = __shfl(lv[p], pr[u], WAVE);
                        const int pex = __shfl(excl, pr[u], WAVE);
                        e[u] = pe0 + (uint32_t)(j - pex);
                        sid[u] = j < total ? (int)a.ids[e[u]] : -1;
                    }
#pragma unroll
                    for (int u = 0; u < INDEX_PASSES; ++u)
                        if (sid[u] >= 0)
                            count(sid[u], keyf(plv[u]),
                                  PACK > 1 && !W16 ? 1u : aw * (a.vals ? a.vals[e[u]] : 1u));
                }
            }
        };
        while (ci < i1) {
            step(nra, nlva);
            if (ci >= i1)
                break;
            step(nrc, nlvc);
        }
        return walked;
    };
    // 16-bit counts [(nm + 1) / 2][n] of n SVs from SV s0 of the tile out as
    // the [n][nm] uint16 partials
    auto write16 = [&](int s0, int n) {
        const int w16 = (a.nm + 1) / 2;
        if ((a.nm & 1) == 0) {
            // word j of SV s holds levels 2j, 2j + 1
            uint32_t *dst = (uint32_t *)((uint16_t *)a.partial + off + (int64_t)s0 * a.nm);
            for (int i = threadIdx.x; i < n * w16; i += INDEX_THREADS) {
                const int sid = i / w16, j = i - sid * w16;
                dst[i] = cnt[j * n + sid];
            }
        } else {
            uint16_t *dst = (uint16_t *)a.partial + off + (int64_t)s0 * a.nm;
            for (int i = threadIdx.x; i < n * a.nm; i += INDEX_THREADS) {
                const int sid = i / a.nm, lv = i - sid * a.nm;
                dst[i] = (uint16_t)(cnt[(lv >> 1) * n + sid] >> (16 * (lv & 1)));
            }
        }
    };
    if constexpr (W16) {
        // weighted, two 16-bit counts per word: an add (at most 255^2 < 2^16)
        // that carries out of its field lowers the sum of the block's fields
        // below the weight added (by 65,535, or 65,536 out of the word), so the
        // block compares the two (64-bit sums) and, where they differ, walks
        // its chunk again with 32-bit counts, sub SVs per pass
        unsigned long long adds = 0;
        walk([&](int lv) { return CntKey{(lv >> 1) * ns, 16u * (uint32_t)(lv & 1)}; },
             [&](int sid, CntKey k, uint32_t add) {
                 if (DIAG && sid < lo_col)
                     return;
                 adds += add;
                 atomicAdd(&cnt[k.off + sid], add << k.inc);
             },
             std::false_type{});
        __syncthreads();
        unsigned long long d = adds;
        for (int i = threadIdx.x; i < ncnt; i += INDEX_THREADS) {
            const uint32_t w = cnt[i];
            d -= (unsigned long long)((w & 0xFFFFu) + (w >> 16));
        }
        __shared__ unsigned long long red64[INDEX_THREADS / WAVE];
        for (int o = WAVE / 2; o > 0; o >>= 1)
            d += __shfl_xor(d, o, WAVE);
        if (lane == 0)
            red64[wave] = d;
        __syncthreads();
        unsigned long long carried = 0;
#pragma unroll
        for (int k = 0; k < INDEX_THREADS / WAVE; ++k)
            carried += red64[k];
        uint32_t *dst = (uint32_t *)a.partial + off;
        if (carried == 0) {
            if (fused) {
                fold(ns, 0, per2{});
                emit();
                return;
            }
            for (int i = threadIdx.x; i < ns * a.nm; i += INDEX_THREADS) {
                const int sid = i / a.nm, lv = i - sid * a.nm;
                dst[i] = (cnt[(lv >> 1) * ns + sid] >> (16 * (lv & 1))) & 0xFFFFu;
            }
            return;
        }
        const int sub = (int)((int64_t)a.tile * ((a.nm + 1) / 2) / a.nm);  // LDS words / nm
        for (int s0 = 0; s0 < ns; s0 += sub) {
            const int n = min(sub, ns - s0);
            __syncthreads();
            for (int i = threadIdx.x; i < n * a.nm; i += INDEX_THREADS)
                cnt[i] = 0u;
            __syncthreads();
            walk([&](int lv) { return CntKey{lv * n, 0u}; },
                 [&](int sid, CntKey k, uint32_t add) {
                     const int r = sid - s0;
                     if ((DIAG && sid < lo_col) || (unsigned)r >= (unsigned)n)
                         return;
                     atomicAdd(&cnt[k.off + r], add);
                 },
                 std::false_type{});
            __syncthreads();
            if (fused) {
                fold(n, s0, per1{});
                continue;
            }
            for (int i = threadIdx.x; i < n * a.nm; i += INDEX_THREADS) {
                const int sid = i / a.nm, lv = i - sid * a.nm;
                dst[(int64_t)s0 * a.nm + i] = cnt[lv * n + sid];
            }
        }
        if (fused)
            emit();
        return;
    } else if constexpr (PACK == 4) {
        // every entry walked is one add, except (DIAG) those skipped
        uint32_t adds = 0;
        auto key4 = [&](int lv) { return CntKey{(lv >> 2) * ns, 1u << (8 * (lv & 3))}; };
        auto count4 = [&](int sid, CntKey k, uint32_t) {
            if constexpr (DIAG) {
                if (sid < lo_col)
                    return;
                ++adds;
            }
            atomicAdd(&cnt[k.off + sid], k.inc);
        };
        uint32_t walked;
        if constexpr (RK == 1 && !DIAG) {
            // up to 4 levels: one word per SV, the records' sink slots apply
            if (wps == 1)
                walked = walk(key4, count4, std::integral_constant<int, 2>{});
            else
                walked = walk(key4, count4, std::true_type{});
        } else {
            walked = walk(key4, count4, std::integral_constant<bool, !DIAG>{});
        }
        if constexpr (!DIAG)
            adds = walked;
        __syncthreads();
        // adds - byte sums over the block: 0 exactly when no add carried (the
        // block's adds stay below 2^32: at most 65,535 per SV, see the host cap)
        uint32_t d = adds;
        for (int i = threadIdx.x; i < ncnt; i += INDEX_THREADS) {
            const uint32_t w = cnt[i];
            d -= (w & 0xFFu) + ((w >> 8) & 0xFFu) + ((w >> 16) & 0xFFu) + (w >> 24);
        }
        __shared__ uint32_t red[INDEX_THREADS / WAVE];
        for (int o = WAVE / 2; o > 0; o >>= 1)
            d += (uint32_t)__shfl_xor((int)d, o, WAVE);
        if (lane == 0)
            red[wave] = d;
        __syncthreads();
        uint32_t carried = 0;
#pragma unroll
        for (int k = 0; k < INDEX_THREADS / WAVE; ++k)
            carried += red[k];
        // partials: the low bytes of every count [nchunks][B][S][nm], then the
        // high bytes (written by blocks that took the 16-bit walk only), then
        // one flag per block [nchunks][B][n_tiles] (1: high bytes written)
        const int64_t n_all = (int64_t)a.nchunks * a.B * a.S * a.nm;
        uint8_t *lo = (uint8_t *)a.partial + off;
        uint8_t *hi = lo + n_all;
        uint8_t *flag = (uint8_t *)a.partial + 2 * n_all +
                        ((int64_t)chunk * a.B + b) * a.n_tiles + tile_id;
        if (carried == 0 && fused) {
            fold(ns, 0, per4{});
            emit();
            return;
        }
        if (carried == 0) {
            if (a.nm == 4) {
                // the word of an SV is its 4 counts in level order: a plain copy
                uint32_t *d4 = (uint32_t *)lo;
                for (int s = threadIdx.x; s < ns; s += INDEX_THREADS)
                    d4[s] = cnt[s];
            } else {
                for (int i = threadIdx.x; i < ns * a.nm; i += INDEX_THREADS) {
                    const int sid = i / a.nm, lv = i - sid * a.nm;
                    lo[i] = (uint8_t)(cnt[(lv >> 2) * ns + sid] >> (8 * (lv & 3)));
                }
            }
            if (threadIdx.x == 0)
                *flag = 0;
            return;
        }
        // a count reached 256: the chunk again with 16-bit counts, sub_tile
        // SVs per pass (what the LDS holds at (nm + 1) / 2 words per SV)
        for (int s0 = 0; s0 < ns; s0 += a.sub_tile) {
            const int n = min(a.sub_tile, ns - s0);
            __syncthreads();
            for (int i = threadIdx.x; i < n * ((a.nm + 1) / 2); i += INDEX_THREADS)
                cnt[i] = 0u;
            __syncthreads();
            walk([&](int lv) { return CntKey{(lv >> 1) * n, 1u << (16 * (lv & 1))}; },
                 [&](int sid, CntKey k, uint32_t) {
                     const int r = sid - s0;
                     if ((DIAG && sid < lo_col) || (unsigned)r >= (unsigned)n)
                         return;
                     atomicAdd(&cnt[k.off + r], k.inc);
                 },
                 std::false_type{});
            __syncthreads();
            if (fused) {
                fold(n, s0, per2{});
                continue;
            }
            for (int i = threadIdx.x; i < n * a.nm; i += INDEX_THREADS) {
                const int sid = i / a.nm, lv = i - sid * a.nm;
                const uint32_t c = (cnt[(lv >> 1) * n + sid] >> (16 * (lv & 1))) & 0xFFFFu;
                const int64_t e = (int64_t)s0 * a.nm + i;
                lo[e] = (uint8_t)c;
                hi[e] = (uint8_t)(c >> 8);
            }
        }
        if (fused) {
            emit();
            return;
        }
        if (threadIdx.x == 0)
            *flag = 1;
        return;
    } else {
        // counts laid out [word][SV], so that a wave's atomics spread over the banks
        walk(
            [&](int lv) {
                return PACK == 2 ? CntKey{(lv >> 1) * ns, 1u << (16 * (lv & 1))}
                                 : CntKey{lv * ns, 0u};
            },
            [&](int sid, CntKey k, uint32_t add) {
                if (DIAG && sid < lo_col)
                    return;
                atomicAdd(&cnt[k.off + sid], PACK == 2 ? k.inc : add);
            },
            std::integral_constant<bool, PACK == 2 && !DIAG>{});
        __syncthreads();
        if (fused) {
            fold(ns, 0, std::integral_constant<int, PACK == 2 ? 2 : 1>{});
            emit();
        } else if (PACK == 2) {
            write16(0, ns);
        } else {
            uint32_t *dst = (uint32_t *)a.partial + off;
            for (int i = threadIdx.x; i < ns * a.nm; i += INDEX_THREADS) {
                const int sid = i / a.nm, lv = i - sid * a.nm;
                dst[i] = cnt[lv * ns + sid];
            }
        }
    }
}

template <int PACK, int RK, bool W16 = false>
__global__ __launch_bounds__(INDEX_THREADS) void k_index_tile(IndexArgs a)
{
    const int q = blockIdx.x >> 3;
    const int tile_id = q % a.n_tiles;
    const int chunk = (q / a.n_tiles) * 8 + (int)(blockIdx.x & 7);
    if (chunk >= a.nchunks)
        return;
    const int b = blockIdx.y;
    const int64_t slot = ((int64_t)chunk * a.B + b) * a.n_tiles + tile_id;
    if (a.col_min >= 0)  // the Gram's upper triangle
        index_block<PACK, RK, W16, false, true>(a, b, chunk, tile_id, slot);
    else
        index_block<PACK, RK, W16>(a, b, chunk, tile_id, slot);
}

// The persistent form: as many blocks as the chip holds at once (one per CU),
// each popping work items -- (row, window chunk) pairs, times the SV tiles --
// until none is left; pair rc is chunk rc - first[b] of the row b with the
// last first[b] <= rc (a binary search: rows without windows share the next
// row's first, or n_pairs at the end).  The pairs are dealt round-robin to 8
// queues, one per XCD (blocks are dispatched to the XCDs round-robin by
// linear id, so block p runs on XCD p % 8), and a pair's tile items follow each
// other in its queue, so the blocks probing one pair's record lines share an
// L2; a block whose queue is empty takes items from the others.  Item k of
// queue x is pair x + 8 (k / n_tiles), tile k % n_tiles; its level sums go to
// slot pair * n_tiles + tile.
template <int PACK, int RK, bool W16 = false, bool ISM = false>
__global__ __launch_bounds__(INDEX_THREADS) void k_index_tile_queue(IndexArgs a)
{
    __shared__ int pick[4];
    const int home = (int)(blockIdx.x & 7);
    for (;;) {
        if (threadIdx.x == 0) {
            int rc = -1, tile = 0;
            for (int t = 0; t < 8 && rc < 0; ++t) {
                const int qx = (home + t) & 7;
                const int k = atomicAdd(&a.queue[qx], 1);
                const int r = qx + 8 * (k / a.n_tiles);
                if (r < a.n_pairs) {
                    rc = r;
                    tile = k - (k / a.n_tiles) * a.n_tiles;
                }
            }
            int lo = 0;
            if (rc >= 0) {
                int hi = a.B - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (a.pair_first[mid] <= rc)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
            }
            pick[0] = rc;
            pick[1] = tile;
            pick[2] = lo;
            pick[3] = rc >= 0 ? rc - a.pair_first[lo] : 0;
        }
        __syncthreads();
        const int rc = pick[0], tile = pick[1], row = pick[2], chunk = pick[3];
        __syncthreads();  // every thread has its item before thread 0 picks again
        if (rc < 0)
            return;
        index_block<PACK, RK, W16, ISM>(a, row, chunk, tile, (int64_t)rc * a.n_tiles + tile);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_index_explain: GkmExplain through the l-mer index
// ---------------------------------------------------------------------------

constexpr int IXEXP_THREADS = 256;
// SV tiles whose buckets a lane gathers together, entries per bucket and round
// (records: a first round of RECORD_IDS from the record, then IXEXP_ROUND)
constexpr int IXEXP_TILES_REC = 2;
constexpr int IXEXP_TILES = 4;
constexpr int IXEXP_ROUND = 4;

struct IndexExplainArgs {
    const uint32_t *xwin;    // [B][Wx] packed query windows (PAD past each row's)
    const int32_t *xlen;     // [B]
    const uint32_t *offsets; // [4^l * n_tiles + 1], key = lmer * n_tiles + tile
    const uint16_t *ids;     // [entries] SV of each entry, relative to its tile
    const uint8_t *vals;     // [entries] window weight of each entry, or NULL
    const uint32_t *records; // [keys][8] bucket records (unweighted), or NULL
    const uint32_t *masks;   // [n_masks] XOR neighbourhood (levels 0..cmax)
    const uint8_t *levels;   // [n_masks]
    const uint8_t *wtab;     // query window weight by distance, or NULL
    const double *ta, *tb;   // explain tables [l + 2]
    const double *ec;        // [B][S] effective coefficients (SV order of the index)
    const double *lw;        // RK 4: [4^l] the SVs' l-mer weights V, in place of the index
    const double *dx;        // RK 4: [B] K_raw(x, x) (normalize), or NULL
    double *out;             // [B][4][lmax], added with FP64 atomics
    int B, Wx, lmax, n_masks, l, S, tile, n_tiles, chunk_windows, pert, normalize;
};

// Block = (chunk of query windows, query row); a wave takes a window u and its
// lanes 64 neighbourhood masks at a time.  For v = u ^ mask (mismatch count m)
// the bucket of l-mer v in every SV tile gives s = sum_e ec[sv_e] * weight_e
// -- the SV windows equal to v, each a pair with m mismatches -- and every one
// of them adds the same pattern to the window's l positions: at a matching
// position ta[m] to base u_k, at a mismatching one tb[m] to base v_k
// (explain/cpu.py:68-86, 146-207).  A lane sums its probes' patterns in
// registers (am[k]: matches; ad[k][d - 1]: mismatches with digit delta d);
// after the window's last mask the wave adds them up and 4 l lanes put them
// into the block's [4][lmax] LDS sums, which go out with FP64 atomics.
// Probe walk: a lane takes the buckets of XT tiles at a time -- their
// records (RK: start, count, the first ids and weights) or bounds -- and gathers
// the first round of coefficients of all of them with independent loads; a
// bucket's entries past the first round follow IXEXP_ROUND at a time.
// RK 4 (kernels without RBF): s is V[v], the SVs' l-mer weights -- the sum
// over SVs of coef / sqrt(K(sv, sv)) times the windows equal to v (with their
// window weights) -- one load per probe, and the row's 1 / sqrt(K(x, x))
// applied at the end (explain/cpu.py:89-143, explain/lmer.py:31-254).
template <int LT, int RK>
__device__ __forceinline__ void index_explain_block(const IndexExplainArgs &a)
{
    constexpr int LM = LT > 0 ? LT : 15;
    extern __shared__ __align__(16) uint8_t smem[];
    double *res = (double *)smem;          // [4][span]
    double *stab = res;                    // set below, after res
    const int l = LT > 0 ? LT : a.l;
    // Block order: the XCDs take blocks round-robin by id (block p runs on XCD
    // p % 8), so XCD x is given the consecutive (row, chunk) pairs
    // k = x * per + p / 8 (row-major): its L2 then holds the coefficients of
    // about B / 8 rows instead of every row's.  Ids past the last pair exit.
    const int nx = (a.Wx + a.chunk_windows - 1) / a.chunk_windows;  // chunks per row
    const int64_t n_pairs = (int64_t)nx * a.B;
    const int64_t per = (n_pairs + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (k >= n_pairs)
        return;
    const int b = (int)(k / nx);
    const int wx = a.xlen[b] - l + 1;
    const int i0 = (int)(k - (int64_t)b * nx) * a.chunk_windows;
    const int i1 = min(wx, i0 + a.chunk_windows);
    const int span = a.chunk_windows + LM - 1;  // positions i0 .. i0 + span - 1
    stab = res + 4 * (size_t)span;
    for (int i = threadIdx.x; i < 4 * span; i += IXEXP_THREADS)
        res[i] = 0.0;
    if (threadIdx.x < 16) {
        const int t = threadIdx.x;
        stab[t] = t < a.l + 2 ? a.ta[t] : 0.0;
        stab[16 + t] = t < a.l + 2 ? a.tb[t] : 0.0;
    }
    __syncthreads();
    const uint32_t nl = 1u << (2 * l);
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const double *ecb = a.ec + (int64_t)b * a.S;
    const int nt = a.n_tiles;
    for (int i = i0 + wave; i < i1; i += IXEXP_THREADS / WAVE) {
        const uint32_t u = a.xwin[(int64_t)b * a.Wx + i];
        if (u >= nl)
            continue;  // PAD
        double am[LM], ad[LM][3];
#pragma unroll
        for (int k = 0; k < LM; k++) {
            am[k] = 0.0;
            ad[k][0] = ad[k][1] = ad[k][2] = 0.0;
        }
        for (int t0 = 0; t0 < a.n_masks; t0 += WAVE) {
            const int t = t0 + lane;
            if (t < a.n_masks) {
                const uint32_t mask = a.masks[t];
                const uint32_t v = u ^ mask;
                double sv = 0.0;
                bool walked = RK == 4;
                if (RK == 4)
                    sv = a.lw[v];
                if (RK == 3) {
                    // the l-mer's record: its run over all tiles, the first
                    // LREC_IDS ids inline; run entry k is in the tile whose
                    // prefix end pe first exceeds k
                    const uint4 *r = (const uint4 *)(a.records + (size_t)v * 16);
                    const uint4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
                    const uint32_t wd[14] = {r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x,
                                             r2.y, r2.z, r2.w, r3.x, r3.y, r3.z, r3.w};
                    int pe[LREC_MAX_TILES], tot = 0;
                    bool sat = false;
#pragma unroll
                    for (int u = 0; u < LREC_MAX_TILES; ++u) {
                        const int cu = u < nt ? (int)((r0.y >> (8 * u)) & 0xFFu) : 0;
                        sat = sat || cu == 0xFF;
                        tot += cu;
                        pe[u] = tot;
                    }
                    if (!sat) {
                        const int n0 = min(tot, LREC_IDS);
                        double cv[LREC_IDS];
#pragma unroll
                        for (int k = 0; k < LREC_IDS; ++k) {
                            const int tk = (k >= pe[0]) + (k >= pe[1]) + (k >= pe[2]);
                            cv[k] = k < n0 ? ecb[(size_t)tk * a.tile + lrec_id(wd, k)] : 0.0;
                        }
#pragma unroll
                        for (int k = 0; k < LREC_IDS; ++k)
                            sv += cv[k];
                        for (int kb = LREC_IDS; kb < tot; kb += IXEXP_ROUND) {
                            double cw[IXEXP_ROUND];
#pragma unroll
                            for (int k = 0; k < IXEXP_ROUND; ++k) {
                                const int e = kb + k;
                                const int tk = (e >= pe[0]) + (e >= pe[1]) + (e >= pe[2]);
                                cw[k] = e < tot ? ecb[(size_t)tk * a.tile + a.ids[r0.x + (uint32_t)e]]
                                                : 0.0;
                            }
#pragma unroll
                            for (int k = 0; k < IXEXP_ROUND; ++k)
                                sv += cw[k];
                        }
                        walked = true;
                    }
                }
                // per tile: 32-byte records (RK 1, 2) or bounds (RK 0; RK 3 saturated)
                constexpr int RT = RK >= 3 ? 0 : RK;
                constexpr int R0 = RT ? record_ids<RT>() : IXEXP_ROUND;  // first round
                constexpr int XT = RT ? IXEXP_TILES_REC : IXEXP_TILES;
                for (int tl0 = 0; tl0 < (walked ? 0 : nt); tl0 += XT) {
                    uint32_t e0[XT], rid[XT][R0], rwt[XT][R0];
                    int c[XT], n0[XT];  // entries, and those of the first round
#pragma unroll
                    for (int x = 0; x < XT; ++x) {
                        const uint32_t key = v * (uint32_t)nt + (uint32_t)(tl0 + x);
                        e0[x] = 0u;
                        c[x] = n0[x] = 0;
                        if (tl0 + x < nt) {
                            if (RT) {
                                const uint4 *r = (const uint4 *)(a.records + (size_t)key * 8);
                                const uint4 ra = r[0], rb = r[1];
                                const uint32_t w[8] = {ra.x, ra.y, ra.z, ra.w,
                                                       rb.x, rb.y, rb.z, rb.w};
                                const RecBucket bk = rec_bucket<RT>(ra, rb);
                                e0[x] = bk.e0;
                                c[x] = bk.c;
                                n0[x] = bk.n_in;
#pragma unroll
                                for (int k = 0; k < R0; ++k) {
                                    rid[x][k] = rec_id<RT>(w, k);
                                    rwt[x][k] = RT == 2 ? rec_val(w, k) : 1u;
                                }
                            } else {
                                const bounds2 r = *(const bounds2 *)(a.offsets + key);
                                e0[x] = r.x;
                                c[x] = (int)(r.y - r.x);
                                n0[x] = min(c[x], R0);
                            }
                        }
                    }
                    double cv[XT][R0];
#pragma unroll
                    for (int x = 0; x < XT; ++x) {
                        const double *ect = ecb + (size_t)(tl0 + x) * a.tile;
#pragma unroll
                        for (int k = 0; k < R0; ++k) {
                            const bool in = k < n0[x];
                            if (!RT) {
                                rid[x][k] = in ? a.ids[e0[x] + k] : 0u;
                                rwt[x][k] = a.vals && in ? a.vals[e0[x] + k] : 1u;
                            }
                            cv[x][k] = in ? ect[rid[x][k]] * (double)rwt[x][k] : 0.0;
                        }
                    }
#pragma unroll
                    for (int x = 0; x < XT; ++x)
#pragma unroll
                        for (int k = 0; k < R0; ++k)
                            sv += cv[x][k];
#pragma unroll
                    for (int x = 0; x < XT; ++x) {
                        const double *ect = ecb + (size_t)(tl0 + x) * a.tile;
                        for (int kb = n0[x]; kb < c[x]; kb += IXEXP_ROUND) {
                            double cw[IXEXP_ROUND];
#pragma unroll
                            for (int k = 0; k < IXEXP_ROUND; ++k) {
                                const bool in = kb + k < c[x];
                                const uint32_t e = e0[x] + (uint32_t)(kb + k);
                                const double wv = a.vals && in ? (double)a.vals[e] : 1.0;
                                cw[k] = in ? ect[a.ids[e]] * wv : 0.0;
                            }
#pragma unroll
                            for (int k = 0; k < IXEXP_ROUND; ++k)
                                sv += cw[k];
                        }
                    }
                }
                if (sv != 0.0) {
                    const int m = a.levels[t];
                    const double sa = sv * stab[m], sb = sv * stab[16 + m];
#pragma unroll
                    for (int k = 0; k < LM; k++) {
                        // digits past l are 0: with LT == 0 am[k >= l] gathers a
                        // sum nobody reads (no dynamic indexes, no scratch)
                        const uint32_t d = (mask >> (2 * k)) & 3u;
                        am[k] += d == 0 ? sa : 0.0;
                        ad[k][0] += d == 1 ? sb : 0.0;
                        ad[k][1] += d == 2 ? sb : 0.0;
                        ad[k][2] += d == 3 ? sb : 0.0;
                    }
                }
            }
        }
        // wave totals (every lane ends with them), then lane j < 4 l places
        // accumulator j: position k = j / 4, slot j % 4 (0: matches, 1..3: delta)
        double aw = a.wtab ? (double)window_weight(a.wtab, wx, i) : 1.0;
        if (RK == 4 && a.normalize)
            aw *= 1.0 / sqrt(fmax(a.dx[b], 1e-20));
#pragma unroll
        for (int k = 0; k < LM; k++) {
            for (int off = WAVE / 2; off > 0; off >>= 1) {
                am[k] += __shfl_xor(am[k], off, WAVE);
                ad[k][0] += __shfl_xor(ad[k][0], off, WAVE);
                ad[k][1] += __shfl_xor(ad[k][1], off, WAVE);
                ad[k][2] += __shfl_xor(ad[k][2], off, WAVE);
            }
        }
        if (lane < 4 * l) {
            const int k = lane >> 2, slot = lane & 3;
            double val = 0.0;
#pragma unroll
            for (int kk = 0; kk < LM; kk++)  // static indexes only (no scratch)
                if (kk == k)
                    val = slot == 0 ? am[kk] : slot == 1 ? ad[kk][0] : slot == 2 ? ad[kk][1]
                                                                            : ad[kk][2];
            val *= aw;
            const uint32_t uk = (u >> (2 * k)) & 3u;
            const int pos = i - i0 + k;
            if (val != 0.0) {
                if (slot > 0) {
                    atomicAdd(&res[(uk ^ (uint32_t)slot) * span + pos], val);
                } else if (!a.pert) {
                    atomicAdd(&res[uk * span + pos], val);
                } else {
                    for (uint32_t h = 0; h < 4; h++)
                        if (h != uk)
                            atomicAdd(&res[h * span + pos], val);
                }
            }
        }
    }
    __syncthreads();
    if (i1 <= i0)
        return;
    const int npos = i1 - i0 + l - 1;
    double *ob = a.out + (int64_t)b * 4 * a.lmax + i0;
    for (int j = threadIdx.x; j < 4 * npos; j += IXEXP_THREADS) {
        const int v = j / npos, p = j - v * npos;
        const double r = res[v * span + p];
        if (r != 0.0)
            atomicAdd(&ob[(int64_t)v * a.lmax + p], r);
    }
}

template <int LT, int RK>
__global__ __launch_bounds__(IXEXP_THREADS) void k_index_explain(IndexExplainArgs a)
{
    index_explain_block<LT, RK>(a);
}

// the l-mer weights form (RK 4), under its own name
template <int LT>
__global__ __launch_bounds__(IXEXP_THREADS) void k_lmer_explain(IndexExplainArgs a)
{
    index_explain_block<LT, 4>(a);
}

// The l-mer weights form over the windows of all rows sorted by l-mer: xwin
// holds n_win keys in ascending order (PAD last), xloc[w] = b * Wx + i the
// row and window index of key w.  A block takes chunk_windows consecutive
// keys, which share their high digits, so the V entries their neighbourhoods
// reach lie in few lines that stay in the L2 between them (in position order a
// window's 4,984 gathers are lines of its own); each window adds its 4 l
// values to out with FP64 atomics.  Consecutive chunks go to one XCD.
template <int LT>
__global__ __launch_bounds__(IXEXP_THREADS) void k_lmer_explain_sorted(IndexExplainArgs a,
                                                                      const int32_t *xloc,
                                                                      int n_win)
{
    constexpr int LM = LT > 0 ? LT : 15;
    __shared__ double stab[32];
    const int l = LT > 0 ? LT : a.l;
    const int nchunk = (n_win + a.chunk_windows - 1) / a.chunk_windows;
    const int per = (nchunk + 7) / 8;
    const int c = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (c >= nchunk)
        return;
    if (threadIdx.x < 16) {
        const int t = threadIdx.x;
        stab[t] = t < a.l + 2 ? a.ta[t] : 0.0;
        stab[16 + t] = t < a.l + 2 ? a.tb[t] : 0.0;
    }
    __syncthreads();
    const int w0 = c * a.chunk_windows, w1 = min(n_win, w0 + a.chunk_windows);
    const uint32_t nl = 1u << (2 * l);
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    for (int w = w0 + wave; w < w1; w += IXEXP_THREADS / WAVE) {
        const uint32_t u = a.xwin[w];
        if (u >= nl)
            continue;  // PAD
        const int loc = xloc[w];
        const int b = loc / a.Wx, i = loc - b * a.Wx;
        double am[LM], ad[LM][3];
#pragma unroll
        for (int k = 0; k < LM; k++) {
            am[k] = 0.0;
            ad[k][0] = ad[k][1] = ad[k][2] = 0.0;
        }
        for (int t0 = 0; t0 < a.n_masks; t0 += WAVE) {
            const int t = t0 + lane;
            if (t < a.n_masks) {
                const uint32_t mask = a.masks[t];
                const double sv = a.lw[u ^ mask];
                if (sv != 0.0) {
                    const int m = a.levels[t];
                    const double sa = sv * stab[m], sb = sv * stab[16 + m];
#pragma unroll
                    for (int k = 0; k < LM; k++) {
                        const uint32_t d = (mask >> (2 * k)) & 3u;
                        am[k] += d == 0 ? sa : 0.0;
                        ad[k][0] += d == 1 ? sb : 0.0;
                        ad[k][1] += d == 2 ? sb : 0.0;
                        ad[k][2] += d == 3 ? sb : 0.0;
                    }
                }
            }
        }
        const int wx = a.xlen[b] - l + 1;
        double aw = a.wtab ? (double)window_weight(a.wtab, wx, i) : 1.0;
        if (a.normalize)
            aw *= 1.0 / sqrt(fmax(a.dx[b], 1e-20));
#pragma unroll
        for (int k = 0; k < LM; k++) {
            for (int off = WAVE / 2; off > 0; off >>= 1) {
                am[k] += __shfl_xor(am[k], off, WAVE);
                ad[k][0] += __shfl_xor(ad[k][0], off, WAVE);
                ad[k][1] += __shfl_xor(ad[k][1], off, WAVE);
                ad[k][2] += __shfl_xor(ad[k][2], off, WAVE);
            }
        }
        if (lane < 4 * l) {
            const int k = lane >> 2, slot = lane & 3;
            double val = 0.0;
#pragma unroll
            for (int kk = 0; kk < LM; kk++)  // static indexes only (no scratch)
                if (kk == k)
                    val = slot == 0 ? am[kk] : slot == 1 ? ad[kk][0] : slot == 2 ? ad[kk][1]
                                                                            : ad[kk][2];
            val *= aw;
            const uint32_t uk = (u >> (2 * k)) & 3u;
            double *ob = a.out + (int64_t)b * 4 * a.lmax + i + k;
            if (val != 0.0) {
                if (slot > 0) {
                    atomicAdd(&ob[(int64_t)(uk ^ (uint32_t)slot) * a.lmax], val);
                } else if (!a.pert) {
                    atomicAdd(&ob[(int64_t)uk * a.lmax], val);
                } else {
                    for (uint32_t h = 0; h < 4; h++)
                        if (h != uk)
                            atomicAdd(&ob[(int64_t)h * a.lmax], val);
                }
            }
        }
    }
}

// hist[b][s][m] (+)= sum over chunks of partial[chunk][b][s][m], over the
// chunks that hold windows of row b (k_index_tile writes no others).
template <typename T>
__global__ void k_index_reduce(const T *__restrict__ partial, int64_t n, int64_t row_n,
                               const int32_t *__restrict__ xlen, int l, int chunk_windows,
                               int nchunks, unsigned long long *__restrict__ hist)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int wx = xlen[i / row_n] - l + 1;
    const int used = wx > 0 ? min(nchunks, (wx + chunk_windows - 1) / chunk_windows) : 0;
    unsigned long long v = 0;
    for (int c = 0; c < used; c++)
        v += partial[(int64_t)c * n + i];
    hist[i] += v;
}

// The same sum, 4 consecutive counts per thread (one 8- or 16-byte load per
// chunk, 4 chunks' loads in flight); n % 4 == 0.  The 4 counts lie in at most
// two rows (row_n >= 4), each summed over the chunks of its own row.
template <typename T> struct vec4_of;
template <> struct vec4_of<uint16_t> {
    typedef ushort4 type;
};
template <> struct vec4_of<uint32_t> {
    typedef uint4 type;
};

template <typename T>
__global__ __launch_bounds__(256) void k_index_reduce4(const T *__restrict__ partial, int64_t n,
                                                       int64_t row_n,
                                                       const int32_t *__restrict__ xlen, int l,
                                                       int chunk_windows, int nchunks,
                                                       unsigned long long *__restrict__ hist)
{
    typedef typename vec4_of<T>::type V;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= n)
        return;
    const int64_t r0 = i / row_n;
    const int64_t split = (r0 + 1) * row_n - i;  // counts k < split are in row r0
    auto used_of = [&](int64_t r) {
        const int wx = xlen[r] - l + 1;
        return wx > 0 ? min(nchunks, (wx + chunk_windows - 1) / chunk_windows) : 0;
    };
    const int ua = used_of(r0);
    const int ub = split < 4 ? used_of(r0 + 1) : ua;
    const int um = max(ua, ub);
    unsigned long long v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    const V *src = (const V *)(partial + i);
    const int64_t step = n / 4;  // one chunk, in vectors
#pragma unroll 4
    for (int c = 0; c < um; c++) {
        const V p = src[(int64_t)c * step];
        v0 += (0 < split ? c < ua : c < ub) ? (unsigned long long)p.x : 0ull;
        v1 += (1 < split ? c < ua : c < ub) ? (unsigned long long)p.y : 0ull;
        v2 += (2 < split ? c < ua : c < ub) ? (unsigned long long)p.z : 0ull;
        v3 += (3 < split ? c < ua : c < ub) ? (unsigned long long)p.w : 0ull;
    }
    hist[i] += v0;
    hist[i + 1] += v1;
    hist[i + 2] += v2;
    hist[i + 3] += v3;
}

// The same sum over the byte partials of the 8-bit count kernel: low bytes,
// plus the high bytes of the blocks (chunk, row, SV tile) whose flag is set;
// 4 consecutive counts per thread (one 4-byte load per chunk when n % 4 == 0).
__global__ __launch_bounds__(256) void k_index_reduce8(const uint8_t *__restrict__ partial,
                                                       int64_t n, int64_t row_n, int nm, int tile,
                                                       int n_tiles, int B,
                                                       const int32_t *__restrict__ xlen, int l,
                                                       int chunk_windows, int nchunks,
                                                       unsigned long long *__restrict__ hist)
{
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= n)
        return;
    const int64_t n_all = (int64_t)nchunks * n;
    const uint8_t *hi = partial + n_all;
    const uint8_t *flags = partial + 2 * n_all;
    // each count's flag column (row, SV tile) and number of chunks (its row's)
    int fr[4], uk[4];
    int um = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        fr[k] = 0;
        uk[k] = 0;
        const int64_t e = i + k;
        if (e < n) {
            const int64_t r = e / row_n;
            const int sv = (int)((e - r * row_n) / nm);
            const int wx = xlen[r] - l + 1;
            fr[k] = (int)r * n_tiles + sv / tile;
            uk[k] = wx > 0 ? min(nchunks, (wx + chunk_windows - 1) / chunk_windows) : 0;
            um = max(um, uk[k]);
        }
    }
    const bool vec = (n & 3) == 0;
    unsigned long long v[4] = {0ull, 0ull, 0ull, 0ull};
    if (vec && fr[0] == fr[3] && uk[0] == uk[3]) {
        // the 4 counts share a row and an SV tile (nearly always): one flag
        // and one 4-byte load per chunk, 4 chunks in flight
        uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // < 2^32: at most 2^16 per chunk
        int c = 0;
        for (; c < um; c++) {
            const uchar4 p = *(const uchar4 *)(partial + (int64_t)c * n + i);
            uint32_t x0 = p.x, x1 = p.y, x2 = p.z, x3 = p.w;
            if (flags[(int64_t)c * B * n_tiles + fr[0]]) {
                const uchar4 h = *(const uchar4 *)(hi + (int64_t)c * n + i);
                x0 |= (uint32_t)h.x << 8;
                x1 |= (uint32_t)h.y << 8;
                x2 |= (uint32_t)h.z << 8;
                x3 |= (uint32_t)h.w << 8;
            }
            s0 += x0;
            s1 += x1;
            s2 += x2;
            s3 += x3;
            if ((c & 0xFFFF) == 0xFFFF) {  // flush before 2^16 chunks could overflow
                v[0] += s0, v[1] += s1, v[2] += s2, v[3] += s3;
                s0 = s1 = s2 = s3 = 0;
            }
        }
        v[0] += s0, v[1] += s1, v[2] += s2, v[3] += s3;
        um = 0;  // done
    }
    for (int c = 0; c < um; c++) {
        const uint8_t *pc = partial + (int64_t)c * n + i;
        uint32_t pv[4];
        if (vec) {
            const uchar4 p = *(const uchar4 *)pc;
            pv[0] = p.x;
            pv[1] = p.y;
            pv[2] = p.z;
            pv[3] = p.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                pv[k] = c < uk[k] ? pc[k] : 0u;
        }
        const uint8_t *fc = flags + (int64_t)c * B * n_tiles;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (c < uk[k]) {
                uint32_t x = pv[k];
                if (fc[fr[k]])
                    x |= (uint32_t)hi[(int64_t)c * n + i + k] << 8;
                v[k] += x;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i + k < n)
            hist[i + k] += v[k];
}

// Launch k_index_tile (shared by gkm_index_hist and gkm_index_score): wfix NULL
// writes the per-chunk counts to `partial`, else the blocks' exact level sums.
// *count_bytes: the format of those counts (1, 2 or 4) for gkm_index_reduce.
static int launch_index_tile(const char *who, const uint32_t *xwin, const int32_t *xpos,
                             const int32_t *xlen, int B, int Wx, const uint32_t *offsets,
                             const uint16_t *ids, const uint8_t *vals, const uint32_t *records,
                             int record_tables, const uint32_t *masks, const uint8_t *levels,
                             int n_masks, int l, int nm, const uint8_t *wtab, int S, int tile,
                             int n_tiles, int chunk_windows, int pack_counts, int64_t col_min,
                             int64_t row_base, void *partial, const long long *wfix,
                             int *count_bytes, void *stream, const int32_t *pair_first = nullptr,
                             int n_pairs = 0, int *queue = nullptr, int n_blocks = 0,
                             const IsmArgs *ism = nullptr)
{
    int rc = check_l_nm(l, nm);
    if (rc)
        return rc;
    if (l > 13)
        return fail(-1, "%s: the l-mer index needs l <= 13, got %d", who, l);
    if (tile < 1 || tile > 65536 || n_tiles != (S + tile - 1) / tile || chunk_windows < 1)
        return fail(-1, "%s: bad tiling (tile=%d n_tiles=%d S=%d chunk=%d)", who, tile,
                    n_tiles, S, chunk_windows);
    const bool persistent = pair_first != nullptr;
    if (ism && !persistent)
        return fail(-1, "%s: the ISM form is a persistent launch", who);
    if (B > 65535 && !persistent)
        return fail(-1, "%s: at most 65535 rows per call, got %d", who, B);
    if (pack_counts < 0 || pack_counts > 2)
        return fail(-1, "%s: pack_counts must be 0, 1 or 2, got %d", who, pack_counts);
    if (records != nullptr && (record_tables < 1 || record_tables > l))
        return fail(-1, "%s: record_tables must be 1..l (%d), got %d", who, l, record_tables);
    // the kernels address the record tables with 32-bit byte offsets
    if (records != nullptr &&
        (uint64_t)record_tables * ((uint64_t)1 << (2 * l)) * (uint64_t)n_tiles * 32u >
            ((uint64_t)1 << 32))
        return fail(-1, "%s: record tables over 4 GiB (%d tables, l=%d, %d tiles)", who,
                    record_tables, l, n_tiles);
    if (persistent && col_min >= 0)
        return fail(-1, "%s: the persistent launch takes no col_min", who);
    if (wfix != nullptr && nm > 16)
        return fail(-1, "%s: exact scoring takes at most 16 levels, got %d", who, nm);
    // counts per LDS word: 1 (uint32), 2 (16-bit), 4 (8-bit); weighted
    // indexes take 16-bit counts (with the carry check) when pack_counts is 1
    const bool w16 = vals != nullptr && pack_counts == 1;
    const int pack = vals != nullptr ? (w16 ? 2 : 1) : pack_counts == 2 ? 4 : pack_counts == 1 ? 2 : 1;
    const bool pack16 = pack > 1 && !w16;  // uint16 partials
    *count_bytes = pack == 4 ? 1 : pack16 ? 2 : 4;
    if (B <= 0 || Wx <= 0 || S <= 0 || n_masks <= 0)
        return 0;
    const size_t words = (size_t)tile * ((nm + pack - 1) / pack);
    size_t lds = (words + RECORD_SINKS) * sizeof(uint32_t);  // + the records' sink words
    if (lds > 159 * 1024)
        return fail(-2, "%s: a tile of %d SVs x %d levels exceeds LDS", who, tile, nm);
    IndexArgs a;
    a.xwin = xwin;
    a.xpos = xpos;
    a.xlen = xlen;
    a.offsets = offsets;
    a.ids = ids;
    a.vals = vals;
    a.masks = masks;
    a.levels = levels;
    a.record_tables = records != nullptr ? record_tables : 1;
    a.wtab = wtab;
    a.partial = partial;
    a.wfix = wfix;
    a.pair_first = pair_first;
    a.queue = queue;
    a.ism = ism ? *ism : IsmArgs{};
    a.n_pairs = n_pairs;
    a.B = B;
    a.Wx = Wx;
    a.n_masks = n_masks;
    a.l = l;
    a.nm = nm;
    a.S = S;
    a.tile = tile;
    a.n_tiles = n_tiles;
    a.chunk_windows = chunk_windows;
    a.nchunks = (Wx + chunk_windows - 1) / chunk_windows;
    a.col_min = col_min;
    a.row_base = row_base;
    a.sub_tile = (int)(words / ((nm + 1) / 2));
    const int rk = records == nullptr ? 0 : vals ? 2 : 1;  // 32-byte records per key
    a.records = records;
    typedef void (*tile_fn)(IndexArgs);
    tile_fn kfn;
    if (!persistent)
        kfn = w16       ? (rk ? k_index_tile<2, 2, true> : k_index_tile<2, 0, true>)
            : pack == 4 ? (rk ? k_index_tile<4, 1> : k_index_tile<4, 0>)
            : pack == 2 ? (rk ? k_index_tile<2, 1> : k_index_tile<2, 0>)
            : rk == 2   ? k_index_tile<1, 2>
            : rk == 1   ? k_index_tile<1, 1> : k_index_tile<1, 0>;
    else if (ism) {  // the index formats of the ISM form: records, 8- or 16-bit counts
        kfn = w16 && rk == 2             ? k_index_tile_queue<2, 2, true, true>
            : !w16 && pack == 4 && rk == 1 ? k_index_tile_queue<4, 1, false, true>
            : !w16 && pack == 2 && rk == 1 ? k_index_tile_queue<2, 1, false, true>
                                          : nullptr;
        if (kfn == nullptr)
            return fail(-2, "%s: ISM through the index takes an index with bucket records "
                            "and 8- or 16-bit counts", who);
    } else
        kfn = w16       ? (rk ? k_index_tile_queue<2, 2, true> : k_index_tile_queue<2, 0, true>)
            : pack == 4 ? (rk ? k_index_tile_queue<4, 1> : k_index_tile_queue<4, 0>)
            : pack == 2 ? (rk ? k_index_tile_queue<2, 1> : k_index_tile_queue<2, 0>)
            : rk == 2   ? k_index_tile_queue<1, 2>
            : rk == 1   ? k_index_tile_queue<1, 1> : k_index_tile_queue<1, 0>;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        return fail(-2, "could not raise the LDS limit to %zu bytes", lds);
    if (!persistent) {
        dim3 grid((unsigned)(n_tiles * ((a.nchunks + 7) / 8) * 8), (unsigned)B);
        hipLaunchKernelGGL(kfn, grid, dim3(INDEX_THREADS), lds, S_(stream), a);
    } else if (n_pairs > 0) {
        hipLaunchKernelGGL(kfn, dim3((unsigned)n_blocks), dim3(INDEX_THREADS), lds, S_(stream),
                           a);
    }
    GKM_CHECK_LAUNCH(who);
    return 0;
}

GKM_API int gkm_index_hist(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B, int Wx,
                           const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                           const uint32_t *records, int record_tables, const uint32_t *masks,
                           const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab,
                           int S, int tile, int n_tiles, int chunk_windows, int pack_counts,
                           int64_t col_min, int64_t row_base, void *partial, void *hist,
                           void *stream)
{
    int count_bytes = 4;
    int rc = launch_index_tile("gkm_index_hist", xwin, xpos, xlen, B, Wx, offsets, ids, vals,
                               records, record_tables, masks, levels, n_masks, l, nm, wtab, S,
                               tile, n_tiles, chunk_windows, pack_counts, col_min, row_base,
                               partial, nullptr, &count_bytes, stream);
    if (rc || hist == nullptr || B <= 0 || Wx <= 0 || S <= 0 || n_masks <= 0)
        return rc;  // hist NULL: the caller sums the partials with gkm_index_reduce
    return gkm_index_reduce(partial, count_bytes, xlen, B, Wx, l, nm, S, tile, chunk_windows,
                            hist, stream);
}

// out[b] = the row's decision value from the level sums of its blocks (the
// chunks that hold its windows, every SV tile).
// pair_first NULL: slots [chunk][B][n_tiles] of gkm_index_score; else the
// persistent form's slots (pair_first[b] + chunk) * n_tiles + tile.
// alias P > 0 (persistent form): the last P rows are alternate alleles that
// were not probed; row b = B - P + p takes the level sums of rows
// B - 4P + p (its reference allele) + B - 3P + p (its windows over the
// variant) - B - 2P + p (the reference's windows there): the counts are
// sums over windows, so these are its own level sums exactly.
__global__ __launch_bounds__(WAVE) void k_index_score_reduce(
    const i128 *__restrict__ psum, int B, int n_tiles, int nm, int nchunks,
    const int32_t *__restrict__ xlen, int l, int chunk_windows, const double *__restrict__ tab,
    const double *__restrict__ dx, int shift, int normalize, double bias,
    double *__restrict__ out, const int32_t *__restrict__ pair_first, int alias)
{
    __shared__ i128 sums[16];
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    const bool aliased = alias > 0 && b >= B - alias;
    const int n_src = aliased ? 3 : 1;
    for (int lv = 0; lv < nm; ++lv) {
        i128 acc = 0;
        for (int j = 0; j < n_src; ++j) {
            const int r = aliased ? b - (3 - j) * alias : b;
            const int wx = xlen[r] - l + 1;
            const int u = wx > 0 ? min(nchunks, (wx + chunk_windows - 1) / chunk_windows) : 0;
            i128 part = 0;
            for (int k = lane; k < u * n_tiles; k += WAVE) {
                const int c = k / n_tiles, t = k - c * n_tiles;
                const int64_t slot = pair_first ? ((int64_t)pair_first[r] + c) * n_tiles + t
                                                : ((int64_t)c * B + r) * n_tiles + t;
                part += psum[slot * nm + lv];
            }
            acc = j == 2 ? acc - part : acc + part;
        }
        acc = wave_sum_i128(acc);
        if (lane == 0)
            sums[lv] = acc;
    }
    if (lane == 0)
        out[b] = fixed_score(sums, nm, tab, shift, dx ? dx[b] : 1.0, normalize, bias);
}

// out[b] = the row's decision value from full histograms [B][S][nm] (the
// pairwise kernel's): the same exact level sums, then fixed_score.
__global__ __launch_bounds__(256) void k_combine_fixed(
    const unsigned long long *__restrict__ hist, int B, int S, int nm,
    const double *__restrict__ tab, const double *__restrict__ dx,
    const long long *__restrict__ wfix, int shift, int normalize, double bias,
    double *__restrict__ out)
{
    __shared__ i128 red[256 / WAVE], sums[16];
    const int b = blockIdx.x;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const unsigned long long *h = hist + (int64_t)b * S * nm;
    for (int lv = 0; lv < nm; ++lv) {
        i128 acc = 0;
        for (int sv = threadIdx.x; sv < S; sv += 256)
            acc += (i128)wfix[sv] * (i128)h[(int64_t)sv * nm + lv];
        acc = wave_sum_i128(acc);
        if (lane == 0)
            red[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            i128 t = 0;
#pragma unroll
            for (int k = 0; k < 256 / WAVE; ++k)
                t += red[k];
            sums[lv] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[b] = fixed_score(sums, nm, tab, shift, dx ? dx[b] : 1.0, normalize, bias);
}

GKM_API int gkm_combine_fixed(const void *hist, int B, int S, int nm, const double *tab,
                              const double *dx, const long long *wfix, int shift, int normalize,
                              double bias, double *out, void *stream)
{
    if (nm < 1 || nm > 16)
        return fail(-1, "gkm_combine_fixed: 1..16 levels, got %d", nm);
    if (normalize && dx == nullptr)
        return fail(-1, "gkm_combine_fixed: normalize needs dx");
    if (B <= 0)
        return 0;
    hipLaunchKernelGGL(k_combine_fixed, dim3((unsigned)B), dim3(256), 0, S_(stream),
                       (const unsigned long long *)hist, B, S, nm, tab, dx, wfix, shift,
                       normalize, bias, out);
    GKM_CHECK_LAUNCH("gkm_combine_fixed");
    return 0;
}

GKM_API int64_t gkm_index_psum_bytes(int B, int Wx, int nm, int n_tiles, int chunk_windows)
{
    if (B <= 0 || Wx <= 0 || nm <= 0 || chunk_windows < 1)
        return 0;
    return (int64_t)((Wx + chunk_windows - 1) / chunk_windows) * B * n_tiles * nm * 16;
}

GKM_API int gkm_index_score(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B,
                            int Wx, const uint32_t *offsets, const uint16_t *ids,
                            const uint8_t *vals, const uint32_t *records, int record_tables,
                            const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                            int nm, const uint8_t *wtab, int S, int tile, int n_tiles,
                            int chunk_windows, int pack_counts, const long long *wfix,
                            void *psum, const double *tab, const double *dx, int shift,
                            int normalize, double bias, double *out, void *stream)
{
    if (wfix == nullptr || psum == nullptr || out == nullptr || tab == nullptr)
        return fail(-1, "gkm_index_score: wfix, psum, tab and out are required");
    if (normalize && dx == nullptr)
        return fail(-1, "gkm_index_score: normalize needs dx");
    if ((uintptr_t)psum % 16)
        return fail(-1, "gkm_index_score: psum must be 16-byte aligned");
    int count_bytes = 4;
    int rc = launch_index_tile("gkm_index_score", xwin, xpos, xlen, B, Wx, offsets, ids, vals,
                               records, record_tables, masks, levels, n_masks, l, nm, wtab, S,
                               tile, n_tiles, chunk_windows, pack_counts, -1, 0, psum, wfix,
                               &count_bytes, stream);
    if (rc || B <= 0)
        return rc;
    // rows with no probe (Wx, S or n_masks 0) get the value of empty sums
    const int nchunks = Wx > 0 && S > 0 && n_masks > 0 ? (Wx + chunk_windows - 1) / chunk_windows : 0;
    hipLaunchKernelGGL(k_index_score_reduce, dim3((unsigned)B), dim3(WAVE), 0, S_(stream),
                       (const i128 *)psum, B, n_tiles, nm, nchunks, xlen, l, chunk_windows, tab,
                       dx, shift, normalize, bias, out, nullptr, 0);
    GKM_CHECK_LAUNCH("gkm_index_score");
    return 0;
}

// The persistent launch's plan, in one block: the 8 queue counters zeroed
// and first[b] = the exclusive prefix sum of the rows' window-chunk counts
// (min(nchunks, ceil(windows / chunk_windows)), 0 without windows), each
// thread summing a strip of consecutive rows.
constexpr int PLAN_THREADS = 1024;
__global__ __launch_bounds__(PLAN_THREADS) void k_queue_plan(
    const int32_t *__restrict__ xlen, int B, int l, int chunk_windows, int nchunks,
    int *__restrict__ queue, int32_t *__restrict__ first)
{
    __shared__ int part[PLAN_THREADS];
    const int t = threadIdx.x;
    if (t < 64)
        queue[t] = 0;
    const int per = (B + PLAN_THREADS - 1) / PLAN_THREADS;
    const int b0 = min(B, t * per), b1 = min(B, b0 + per);
    auto chunks = [&](int b) {
        const int wx = xlen[b] - l + 1;
        return wx > 0 ? min(nchunks, (wx + chunk_windows - 1) / chunk_windows) : 0;
    };
    int sum = 0;
    for (int b = b0; b < b1; ++b)
        sum += chunks(b);
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < PLAN_THREADS; d <<= 1) {  // inclusive scan of the strip sums
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int b = b0; b < b1; ++b) {
        first[b] = run;
        run += chunks(b);
    }
}

// bytes of the persistent launch's plan: 8 queue counters (256 B) and first[B]
static int64_t queue_plan_bytes(int B)
{
    return 256 + ((int64_t)(B > 0 ? B : 0) * 4 + 255) / 256 * 256;
}

GKM_API int64_t gkm_index_queue_bytes(int B, int n_pairs, int n_tiles, int nm)
{
    return queue_plan_bytes(B) + (int64_t)(n_pairs > 0 ? n_pairs : 0) * n_tiles * nm * 16;
}

GKM_API int gkm_index_score_queue_pairs(const uint32_t *xwin, const int32_t *xpos,
                                        const int32_t *xlen, int B, int Wx,
                                        const uint32_t *offsets, const uint16_t *ids,
                                        const uint8_t *vals, const uint32_t *records,
                                        int record_tables, const uint32_t *masks,
                                        const uint8_t *levels, int n_masks, int l, int nm,
                                        const uint8_t *wtab, int S, int tile, int n_tiles,
                                        int chunk_windows, int pack_counts,
                                        const long long *wfix, int n_pairs, int n_blocks,
                                        void *workspace, const double *tab, const double *dx,
                                        int shift, int normalize, double bias, double *out,
                                        int pairs, void *stream)
{
    if (wfix == nullptr || workspace == nullptr || out == nullptr || tab == nullptr)
        return fail(-1, "gkm_index_score_queue: wfix, workspace, tab and out are required");
    if (normalize && dx == nullptr)
        return fail(-1, "gkm_index_score_queue: normalize needs dx");
    if ((uintptr_t)workspace % 256)
        return fail(-1, "gkm_index_score_queue: workspace must be 256-byte aligned");
    if (n_blocks < 1 || n_blocks > 65536)
        return fail(-1, "gkm_index_score_queue: n_blocks must be 1..65536, got %d", n_blocks);
    if (pairs < 0 || (int64_t)pairs * 4 > B)
        return fail(-1, "gkm_index_score_queue_pairs: %d allele pairs need 4 rows each, %d rows",
                    pairs, B);
    if (pairs > 0 && (wtab != nullptr || xpos != nullptr))
        return fail(-1, "gkm_index_score_queue_pairs: windows weighted by position cannot move "
                        "to rows of their own");
    if (B <= 0)
        return 0;
    if (chunk_windows < 1)
        return fail(-1, "gkm_index_score_queue: chunk_windows must be >= 1");
    // the rows probed: all but the last `pairs` (alternate alleles, summed in the reduce)
    const int Bq = B - pairs;
    // workspace: 8 queue counters | first[Bq] | level sums [n_pairs][n_tiles][nm]
    int *queue = (int *)workspace;
    int32_t *pair_first = (int32_t *)((uint8_t *)workspace + 256);
    void *psum = (uint8_t *)workspace + queue_plan_bytes(Bq);
    const int nchunks = Wx > 0 && S > 0 && n_masks > 0 ? (Wx + chunk_windows - 1) / chunk_windows : 0;
    hipLaunchKernelGGL(k_queue_plan, dim3(1), dim3(PLAN_THREADS), 0, S_(stream), xlen, Bq, l,
                       chunk_windows, nchunks, queue, pair_first);
    GKM_CHECK_LAUNCH("gkm_index_score_queue");
    int count_bytes = 4;
    int rc = launch_index_tile("gkm_index_score_queue", xwin, xpos, xlen, Bq, Wx, offsets, ids,
                               vals, records, record_tables, masks, levels, n_masks, l, nm, wtab,
                               S, tile, n_tiles, chunk_windows, pack_counts, -1, 0, psum, wfix,
                               &count_bytes, stream, pair_first, n_pairs, queue, n_blocks);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_index_score_reduce, dim3((unsigned)B), dim3(WAVE), 0, S_(stream),
                       (const i128 *)psum, B, n_tiles, nm, nchunks, xlen, l, chunk_windows, tab,
                       dx, shift, normalize, bias, out, pair_first, pairs);
    GKM_CHECK_LAUNCH("gkm_index_score_queue");
    return 0;
}

GKM_API int gkm_index_score_queue(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen,
                                  int B, int Wx, const uint32_t *offsets, const uint16_t *ids,
                                  const uint8_t *vals, const uint32_t *records, int record_tables,
                                  const uint32_t *masks, const uint8_t *levels, int n_masks,
                                  int l, int nm, const uint8_t *wtab, int S, int tile,
                                  int n_tiles, int chunk_windows, int pack_counts,
                                  const long long *wfix, int n_pairs, int n_blocks,
                                  void *workspace, const double *tab, const double *dx,
                                  int shift, int normalize, double bias, double *out,
                                  void *stream)
{
    return gkm_index_score_queue_pairs(xwin, xpos, xlen, B, Wx, offsets, ids, vals, records,
                                       record_tables, masks, levels, n_masks, l, nm, wtab, S,
                                       tile, n_tiles, chunk_windows, pack_counts, wfix, n_pairs,
                                       n_blocks, workspace, tab, dx, shift, normalize, bias, out,
                                       0, stream);
}

GKM_API int gkm_index_hist_queue(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen,
                                 int B, int Wx, const uint32_t *offsets, const uint16_t *ids,
                                 const uint8_t *vals, const uint32_t *records, int record_tables,
                                 const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                                 int nm, const uint8_t *wtab, int S, int tile, int n_tiles,
                                 int chunk_windows, int pack_counts, int n_pairs, int n_blocks,
                                 void *workspace, void *partial, void *hist, void *stream)
{
    if (workspace == nullptr || partial == nullptr)
        return fail(-1, "gkm_index_hist_queue: workspace and partial are required");
    if ((uintptr_t)workspace % 256)
        return fail(-1, "gkm_index_hist_queue: workspace must be 256-byte aligned");
    if (n_blocks < 1 || n_blocks > 65536)
        return fail(-1, "gkm_index_hist_queue: n_blocks must be 1..65536, got %d", n_blocks);
    if (chunk_windows < 1)
        return fail(-1, "gkm_index_hist_queue: chunk_windows must be >= 1");
    if (B <= 0 || Wx <= 0 || S <= 0 || n_masks <= 0)
        return 0;
    // workspace: 8 queue counters | first[B] (the level-sum part is not used)
    int *queue = (int *)workspace;
    int32_t *pair_first = (int32_t *)((uint8_t *)workspace + 256);
    const int nchunks = (Wx + chunk_windows - 1) / chunk_windows;
    hipLaunchKernelGGL(k_queue_plan, dim3(1), dim3(PLAN_THREADS), 0, S_(stream), xlen, B, l,
                       chunk_windows, nchunks, queue, pair_first);
    GKM_CHECK_LAUNCH("gkm_index_hist_queue");
    int count_bytes = 4;
    int rc = launch_index_tile("gkm_index_hist_queue", xwin, xpos, xlen, B, Wx, offsets, ids,
                               vals, records, record_tables, masks, levels, n_masks, l, nm, wtab,
                               S, tile, n_tiles, chunk_windows, pack_counts, -1, 0, partial,
                               nullptr, &count_bytes, stream, pair_first, n_pairs, queue,
                               n_blocks);
    if (rc || hist == nullptr)
        return rc;
    return gkm_index_reduce(partial, count_bytes, xlen, B, Wx, l, nm, S, tile, chunk_windows,
                            hist, stream);
}

// ---------------------------------------------------------------------------
// ISM through the l-mer index
// ---------------------------------------------------------------------------

// The rows of one query's ISM.  Window row i: the query's window i alone
// (its position i kept for the window weight).  New row 3 p + alt [l]: the
// windows covering position p (window i at slot i - max(0, p - l + 1); PAD
// past the last) with base p replaced by the alt-th base other than the
// reference (digit p - i of window i).  Row lengths are windows + l - 1.
__global__ void k_ism_rows(const uint32_t *__restrict__ xwin, const uint8_t *__restrict__ codes,
                           int lx, int l, int32_t *__restrict__ wpos, int32_t *__restrict__ wlen,
                           uint32_t *__restrict__ nwin, int32_t *__restrict__ npos,
                           int32_t *__restrict__ nlen)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int wx = lx - l + 1;
    const int64_t n_new = 3 * (int64_t)lx * l;
    if (t < wx) {
        wpos[t] = (int32_t)t;
        wlen[t] = l;
    }
    if (t >= n_new)
        return;
    const int row = (int)(t / l), k = (int)(t - (int64_t)row * l);
    const int p = row / 3;
    const int lo = max(0, p - l + 1), hi = min(p, wx - 1);
    const int cnt = hi - lo + 1;
    uint32_t w = 0xFFFFFFFFu;
    int pos = 0;
    if (k < cnt) {
        const int i = lo + k;
        const int alt = row - 3 * p, ref = codes[p] & 3;
        const uint32_t v = (uint32_t)(alt + (alt >= ref ? 1 : 0));
        const int j = p - i;
        w = (xwin[i] & ~(3u << (2 * j))) | (v << (2 * j));
        pos = i;
    }
    nwin[t] = w;
    npos[t] = pos;
    if (k == 0)
        nlen[row] = cnt + l - 1;
}

// From the windows' parts C [wx][S]: R[p][s] = the sum over the windows
// covering p for the positions of this block's segment.
constexpr int ISM_SEG = 64;  // positions per block row
__global__ __launch_bounds__(256) void k_ism_cover(const double *__restrict__ C, int wx, int lx,
                                                   int l, int S, double *__restrict__ R)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S)
        return;
    const int p0 = blockIdx.y * ISM_SEG, p1 = min(lx, p0 + ISM_SEG);
    // R at p0 summed directly, then slid along the segment: + the window
    // entering, - the window leaving (a fixed order: deterministic)
    int lo = max(0, p0 - l + 1), hi = min(p0, wx - 1);
    double r = 0.0;
    for (int i = lo; i <= hi; ++i)
        r += C[(int64_t)i * S + s];
    R[(int64_t)p0 * S + s] = r;
#pragma unroll 4
    for (int p = p0 + 1; p < p1; ++p) {
        const int nlo = max(0, p - l + 1), nhi = min(p, wx - 1);
        const double add = nhi > hi ? C[(int64_t)nhi * S + s] : 0.0;
        const double sub = nlo > lo ? C[(int64_t)lo * S + s] : 0.0;
        r = r + add - sub;
        lo = nlo;
        hi = nhi;
        R[(int64_t)p * S + s] = r;
    }
}

// Per SV into svp [S][4]: K_raw(x, s) -- the windows fall into the disjoint
// groups [k l, k l + l) that R[k l + l - 1] covers -- its finished value
// (normalised by K_raw(x, x) and K_raw(s, s), RBF), the coefficient and
// K_raw(s, s).
__global__ __launch_bounds__(256) void k_ism_ref(const double *__restrict__ R, int wx, int l,
                                                 int S, const double *__restrict__ coef,
                                                 const double *__restrict__ ds,
                                                 const double *__restrict__ dx, int normalize,
                                                 int rbf, double gamma, double *__restrict__ svp)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S)
        return;
    double k = 0.0;
#pragma unroll 8
    for (int g = 0; g < wx; g += l)
        k += R[(int64_t)(g + l - 1) * S + s];
    double f = k;
    if (normalize)
        f = k / fmax(sqrt(dx[0] * ds[s]), 1e-10);
    if (rbf)
        f = exp(gamma * (f - 1.0));
    double *q = svp + (int64_t)s * 4;
    q[0] = k;
    q[1] = f;
    q[2] = coef[s];
    q[3] = ds[s];
}

// out[v][p] = the sum over tiles (in order) of new row (p, v)'s tile sums.
__global__ void k_ism_reduce(const double *__restrict__ part, const int32_t *__restrict__ first,
                             int n_tiles, const uint8_t *__restrict__ codes, int lx, int lmax,
                             double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * lx)
        return;
    const int p = r / 3, alt = r - 3 * p, ref = codes[p] & 3;
    const int v = alt + (alt >= ref ? 1 : 0);
    double t = 0.0;
    for (int k = 0; k < n_tiles; ++k)
        t += part[(int64_t)first[r] * n_tiles + k];
    out[(int64_t)v * lmax + p] = t;
}

static int64_t align256(int64_t n)
{
    return (n + 255) / 256 * 256;
}

GKM_API int64_t gkm_index_ism_bytes(int lx, int l, int S, int n_tiles)
{
    if (lx < l || l < 1 || S < 1 || n_tiles < 1)
        return 0;
    const int64_t wx = lx - l + 1, rows_new = 3 * (int64_t)lx;
    return align256(wx * S * 8) + align256((int64_t)lx * S * 8) +  // C, R
           align256((int64_t)S * 32) +                              // svp
           align256(wx * 4) * 2 +                                   // window rows
           align256(rows_new * l * 4) * 2 + align256(rows_new * 4) +  // new rows
           queue_plan_bytes((int)wx) + queue_plan_bytes((int)rows_new) +
           align256(rows_new * n_tiles * 8);                        // tile sums
}

GKM_API int gkm_index_ism(const uint32_t *xwin, const uint8_t *codes, int lx, int lmax,
                          const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                          const uint32_t *records, int record_tables, const uint32_t *masks,
                          const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab,
                          int S, int tile, int n_tiles, int pack_counts, const double *tab,
                          const double *coef, const double *ds, const double *dx,
                          const double *dmut, int dmut_stride, int normalize, int rbf,
                          double gamma, int n_blocks, void *workspace, double *out, void *stream)
{
    if (xwin == nullptr || codes == nullptr || tab == nullptr || coef == nullptr ||
        ds == nullptr || workspace == nullptr || out == nullptr)
        return fail(-1, "gkm_index_ism: xwin, codes, tab, coef, ds, workspace and out are "
                        "required");
    if (normalize && (dmut == nullptr || dx == nullptr))
        return fail(-1, "gkm_index_ism: normalize needs dx and dmut");
    if ((uintptr_t)workspace % 256)
        return fail(-1, "gkm_index_ism: workspace must be 256-byte aligned");
    if (n_blocks < 1 || n_blocks > 65536)
        return fail(-1, "gkm_index_ism: n_blocks must be 1..65536, got %d", n_blocks);
    if (lx > lmax)
        return fail(-1, "gkm_index_ism: length %d exceeds lmax %d", lx, lmax);
    if (lx < l || S <= 0 || n_masks <= 0)
        return 0;  // no windows (or no SVs): every delta is 0
    if ((int64_t)lx * 3 * l > (1 << 30) || (lx + ISM_SEG - 1) / ISM_SEG > 65535)
        return fail(-1, "gkm_index_ism: a query of %d bp is too long", lx);
    uint8_t *w = (uint8_t *)workspace;
    const int wx = lx - l + 1, rows_new = 3 * lx;
    double *C = (double *)w;
    w += align256((int64_t)wx * S * 8);
    double *R = (double *)w;
    w += align256((int64_t)lx * S * 8);
    double *svp = (double *)w;
    w += align256((int64_t)S * 32);
    int32_t *wpos = (int32_t *)w;
    w += align256((int64_t)wx * 4);
    int32_t *wlen = (int32_t *)w;
    w += align256((int64_t)wx * 4);
    uint32_t *nwin = (uint32_t *)w;
    w += align256((int64_t)rows_new * l * 4);
    int32_t *npos = (int32_t *)w;
    w += align256((int64_t)rows_new * l * 4);
    int32_t *nlen = (int32_t *)w;
    w += align256((int64_t)rows_new * 4);
    uint8_t *plan_w = w;
    w += queue_plan_bytes(wx);
    uint8_t *plan_new = w;
    w += queue_plan_bytes(rows_new);
    double *part = (double *)w;
    const int64_t nth = 3 * (int64_t)lx * l;
    hipLaunchKernelGGL(k_ism_rows, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0,
                       S_(stream), xwin, codes, lx, l, wpos, wlen, nwin, npos, nlen);
    GKM_CHECK_LAUNCH("gkm_index_ism");
    IsmArgs m{};
    m.tab = tab;
    m.svp = svp;
    m.dmut = dmut;
    m.codes = codes;
    m.part = part;
    m.lmax = lmax;
    m.dmut_stride = dmut_stride;
    m.wx = wx;
    m.normalize = normalize;
    m.rbf = rbf;
    m.gamma = gamma;
    // pass 0: the window rows (one window each: xwin itself, one window per
    // row) into C; pass 1: the new rows (1..l windows each) against R and svp.
    // One work item (pair) per row and tile.
    for (int pass = 0; pass < 2; ++pass) {
        const bool win = pass == 0;
        const int B = win ? wx : rows_new;
        const int W = win ? 1 : l;
        uint8_t *plan = win ? plan_w : plan_new;
        int *queue = (int *)plan;
        int32_t *first = (int32_t *)(plan + 256);
        const int32_t *len = win ? wlen : nlen;
        hipLaunchKernelGGL(k_queue_plan, dim3(1), dim3(PLAN_THREADS), 0, S_(stream), len, B, l, W,
                           1, queue, first);
        GKM_CHECK_LAUNCH("gkm_index_ism");
        m.old_rows = win ? 1 : 0;
        m.R = win ? C : R;
        int count_bytes = 4;
        int rc = launch_index_tile("gkm_index_ism", win ? xwin : nwin, win ? wpos : npos, len, B,
                                   W, offsets, ids, vals, records, record_tables, masks, levels,
                                   n_masks, l, nm, wtab, S, tile, n_tiles, W, pack_counts, -1, 0,
                                   nullptr, nullptr, &count_bytes, stream, first, B, queue,
                                   n_blocks, &m);
        if (rc)
            return rc;
        if (win) {
            dim3 grid((unsigned)((S + 255) / 256), (unsigned)((lx + ISM_SEG - 1) / ISM_SEG));
            hipLaunchKernelGGL(k_ism_cover, grid, dim3(256), 0, S_(stream), C, wx, lx, l, S, R);
            GKM_CHECK_LAUNCH("gkm_index_ism");
            hipLaunchKernelGGL(k_ism_ref, dim3((unsigned)((S + 255) / 256)), dim3(256), 0,
                               S_(stream), R, wx, l, S, coef, ds, dx, normalize, rbf, gamma, svp);
            GKM_CHECK_LAUNCH("gkm_index_ism");
        }
    }
    hipLaunchKernelGGL(k_ism_reduce, dim3((unsigned)((rows_new + 255) / 256)), dim3(256), 0,
                       S_(stream), part, (const int32_t *)(plan_new + 256), n_tiles, codes, lx,
                       lmax, out);
    GKM_CHECK_LAUNCH("gkm_index_ism");
    return 0;
}

GKM_API int gkm_index_reduce(const void *partial, int count_bytes, const int32_t *xlen, int B,
                             int Wx, int l, int nm, int S, int tile, int chunk_windows, void *hist,
                             void *stream)
{
    if (count_bytes != 1 && count_bytes != 2 && count_bytes != 4)
        return fail(-1, "gkm_index_reduce: count_bytes must be 1, 2 or 4, got %d", count_bytes);
    if (chunk_windows < 1 || tile < 1)
        return fail(-1, "gkm_index_reduce: chunk_windows and tile must be >= 1 (%d, %d)",
                    chunk_windows, tile);
    if (B <= 0 || Wx <= 0 || S <= 0 || nm <= 0)
        return 0;
    const int nchunks = (Wx + chunk_windows - 1) / chunk_windows;
    const int64_t n = (int64_t)B * S * nm;
    if (count_bytes == 1) {
        const int64_t blocks = ((n + 3) / 4 + 255) / 256;
        hipLaunchKernelGGL(k_index_reduce8, dim3((unsigned)blocks), dim3(256), 0, S_(stream),
                           (const uint8_t *)partial, n, (int64_t)S * nm, nm, tile,
                           (S + tile - 1) / tile, B, xlen, l, chunk_windows, nchunks,
                           (unsigned long long *)hist);
        GKM_CHECK_LAUNCH("gkm_index_reduce");
        return 0;
    }
    if (n % 4 == 0 && (int64_t)S * nm >= 4) {
        const int64_t blocks = (n / 4 + 255) / 256;
        if (count_bytes == 2)
            hipLaunchKernelGGL(k_index_reduce4<uint16_t>, dim3((unsigned)blocks), dim3(256), 0,
                               S_(stream), (const uint16_t *)partial, n, (int64_t)S * nm, xlen, l,
                               chunk_windows, nchunks, (unsigned long long *)hist);
        else
            hipLaunchKernelGGL(k_index_reduce4<uint32_t>, dim3((unsigned)blocks), dim3(256), 0,
                               S_(stream), (const uint32_t *)partial, n, (int64_t)S * nm, xlen, l,
                               chunk_windows, nchunks, (unsigned long long *)hist);
        GKM_CHECK_LAUNCH("gkm_index_reduce");
        return 0;
    }
    if (count_bytes == 2)
        hipLaunchKernelGGL(k_index_reduce<uint16_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           S_(stream), (const uint16_t *)partial, n, (int64_t)S * nm, xlen, l,
                           chunk_windows, nchunks, (unsigned long long *)hist);
    else
        hipLaunchKernelGGL(k_index_reduce<uint32_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           S_(stream), (const uint32_t *)partial, n, (int64_t)S * nm, xlen, l,
                           chunk_windows, nchunks, (unsigned long long *)hist);
    GKM_CHECK_LAUNCH("gkm_index_reduce");
    return 0;
}

// ---------------------------------------------------------------------------
// gkm_index_hist without a workspace argument: the partial counts live in a
// file-scope device array, one copy per GPU, which the host addresses only
// through the pointer hipGetSymbolAddress gives on that GPU (the array's name
// in host code is a host-side stand-in, not a device address).

static __device__ __align__(256) uint8_t scratch[GKM_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // the array's address on each device, once looked up
}  // namespace

// The scratch array's device address on the current device (looked up on
// that device's first call); 0 or a negative status.
static int scratch_address(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return fail(-3, "gkm_index_hist_noalloc: no current device: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(-1, "gkm_index_hist_noalloc: device %d beyond the %d supported", dev,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(-3, "gkm_index_hist_noalloc: hipGetSymbolAddress(scratch) failed on "
                        "device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

GKM_API int64_t gkm_scratch_bytes(void) { return (int64_t)GKM_SCRATCH_BYTES; }

GKM_API int64_t gkm_index_partial_bytes(int B, int Wx, int S, int nm, int n_tiles,
                                        int chunk_windows, int weighted, int pack_counts)
{
    if (B <= 0 || Wx <= 0 || S <= 0 || nm <= 0 || chunk_windows < 1)
        return 0;
    const int64_t counts = (int64_t)((Wx + chunk_windows - 1) / chunk_windows) * B * S * nm;
    if (weighted)
        return counts * 4;  // uint32 partials (16-bit LDS counts or not)
    if (pack_counts == 2)   // low bytes, high bytes, one flag per block
        return 2 * counts + (int64_t)((Wx + chunk_windows - 1) / chunk_windows) * B * n_tiles;
    return counts * (pack_counts == 1 ? 2 : 4);
}

GKM_API int gkm_index_hist_noalloc(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen,
                                   int B, int Wx, const uint32_t *offsets, const uint16_t *ids,
                                   const uint8_t *vals, const uint32_t *records,
                                   int record_tables, const uint32_t *masks,
                                   const uint8_t *levels, int n_masks, int l, int nm,
                                   const uint8_t *wtab, int S, int tile, int n_tiles,
                                   int chunk_windows, int pack_counts, int64_t col_min,
                                   int64_t row_base, void *hist, void *stream)
{
    if (hist == nullptr)
        return fail(-1, "gkm_index_hist_noalloc: hist is required (the partial counts live in "
                        "the library's scratch and are summed within the call)");
    if (B <= 0 || Wx <= 0 || S <= 0 || n_masks <= 0)
        return 0;
    const int64_t need = gkm_index_partial_bytes(B, Wx, S, nm, n_tiles, chunk_windows,
                                                 vals != nullptr, pack_counts);
    if (need > (int64_t)GKM_SCRATCH_BYTES)
        return fail(-2, "gkm_index_hist_noalloc: the partial counts of this call take %lld "
                        "bytes, over the %lld-byte scratch: use fewer rows or SVs, or larger "
                        "chunk_windows (or gkm_index_hist with a workspace)",
                    (long long)need, (long long)GKM_SCRATCH_BYTES);
    void *partial = nullptr;
    int rc = scratch_address(&partial);
    if (rc)
        return rc;
    return gkm_index_hist(xwin, xpos, xlen, B, Wx, offsets, ids, vals, records, record_tables,
                          masks, levels, n_masks, l, nm, wtab, S, tile, n_tiles, chunk_windows,
                          pack_counts, col_min, row_base, partial, hist, stream);
}

// ---------------------------------------------------------------------------
// Dispatch helpers
// ---------------------------------------------------------------------------

#define GKM_L_CASES(X) X(10) X(11) X(12)

template <template <int, int, bool> class K>
struct Unused;

typedef void (*cross_fn)(CrossArgs);
typedef void (*self_fn)(SelfArgs);
typedef void (*sweep_fn)(SweepArgs);

template <int LT, int NMC>
static cross_fn pick_cross_w(bool weighted)
{
    return weighted ? (cross_fn)k_cross_hist<LT, NMC, true> : (cross_fn)k_cross_hist<LT, NMC, false>;
}
template <int LT>
static cross_fn pick_cross_n(int nm, bool weighted)
{
    return nm <= 4 ? pick_cross_w<LT, 4>(weighted)
         : nm <= 8 ? pick_cross_w<LT, 8>(weighted) : pick_cross_w<LT, 16>(weighted);
}
static cross_fn pick_cross(int l, int nm, bool weighted)
{
    switch (l) {
    case 10: return pick_cross_n<10>(nm, weighted);
    case 11: return pick_cross_n<11>(nm, weighted);
    case 12: return pick_cross_n<12>(nm, weighted);
    default: return pick_cross_n<0>(nm, weighted);
    }
}

template <int LT, int NMC>
static self_fn pick_self_w(bool weighted)
{
    return weighted ? (self_fn)k_self_hist<LT, NMC, true> : (self_fn)k_self_hist<LT, NMC, false>;
}
template <int LT>
static self_fn pick_self_n(int nm, bool weighted)
{
    return nm <= 4 ? pick_self_w<LT, 4>(weighted)
         : nm <= 8 ? pick_self_w<LT, 8>(weighted) : pick_self_w<LT, 16>(weighted);
}
static self_fn pick_self(int l, int nm, bool weighted)
{
    switch (l) {
    case 10: return pick_self_n<10>(nm, weighted);
    case 11: return pick_self_n<11>(nm, weighted);
    case 12: return pick_self_n<12>(nm, weighted);
    default: return pick_self_n<0>(nm, weighted);
    }
}

template <int LT, int NMC>
static sweep_fn pick_sweep_m(int mode)
{
    switch (mode) {
    case SW_HYP: return (sweep_fn)k_sweep<LT, NMC, SW_HYP>;
    case SW_PERT: return (sweep_fn)k_sweep<LT, NMC, SW_PERT>;
    default: return (sweep_fn)k_sweep<LT, NMC, SW_ISM>;
    }
}
template <int LT>
static sweep_fn pick_sweep_n(int nm, int mode)
{
    // 5 levels: ISM at the default d = 3 (a mutation moves a pair by one level)
    return nm <= 4 ? pick_sweep_m<LT, 4>(mode) : nm == 5 ? pick_sweep_m<LT, 5>(mode)
         : nm <= 8 ? pick_sweep_m<LT, 8>(mode) : pick_sweep_m<LT, 16>(mode);
}
static sweep_fn pick_sweep(int l, int nm, int mode)
{
    switch (l) {
    case 10: return pick_sweep_n<10>(nm, mode);
    case 11: return pick_sweep_n<11>(nm, mode);
    case 12: return pick_sweep_n<12>(nm, mode);
    default: return pick_sweep_n<0>(nm, mode);
    }
}

static int check_l_nm(int l, int nm)
{
    if (l < 1 || l > 15)
        return fail(-1, "l must be in 1..15, got %d", l);
    if (nm < 1 || nm > 16 || nm > l + 1)
        return fail(-1, "histogram levels must be in 1..min(16, l + 1), got %d", nm);
    return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

GKM_API int gkm_planes_words(int lmax) { return lmax / 32 + 4; }

GKM_API int gkm_pack_planes(const uint8_t *codes, const int32_t *lengths, int n, int lmax,
                            int nw, void *planes, void *planes_rc, void *stream)
{
    if (n < 0 || lmax < 1 || nw < lmax / 32 + 4)
        return fail(-1, "gkm_pack_planes: need n >= 0, lmax >= 1, nw >= lmax/32 + 4 (n=%d lmax=%d nw=%d)",
                    n, lmax, nw);
    if (n == 0)
        return 0;
    int64_t total = (int64_t)n * nw;
    int threads = 256;
    int64_t blocks = (total + threads - 1) / threads;
    hipLaunchKernelGGL(k_pack_planes, dim3((unsigned)blocks), dim3(threads), 0, S_(stream), codes,
                       lengths, n, lmax, nw, (P2 *)planes, (P2 *)planes_rc);
    GKM_CHECK_LAUNCH("gkm_pack_planes");
    return 0;
}

GKM_API int gkm_decode_ascii(const uint8_t *raw, const int64_t *starts, const int32_t *lengths,
                             int n, int lmax, uint8_t *codes, int32_t *bad, void *stream)
{
    if (n < 0 || lmax < 1)
        return fail(-1, "gkm_decode_ascii: need n >= 0 and lmax >= 1 (n=%d lmax=%d)", n, lmax);
    if (n == 0)
        return 0;
    hipLaunchKernelGGL(k_decode_ascii, dim3((unsigned)n), dim3(256), 0, S_(stream), raw, starts,
                       lengths, lmax, codes, bad);
    GKM_CHECK_LAUNCH("gkm_decode_ascii");
    return 0;
}

GKM_API int gkm_pack_windows(const uint8_t *codes, const int32_t *lengths, int n, int lmax,
                             int l, uint32_t *fwd, uint32_t *rcw, void *stream)
{
    if (l < 1 || l > 15 || lmax < l)
        return fail(-1, "gkm_pack_windows: need 1 <= l <= 15 and lmax >= l (l=%d lmax=%d)", l, lmax);
    if (n == 0)
        return 0;
    int W = lmax - l + 1;
    int64_t total = (int64_t)n * W;
    int threads = 256;
    int64_t blocks = (total + threads - 1) / threads;
    hipLaunchKernelGGL(k_pack_windows, dim3((unsigned)blocks), dim3(threads), 0, S_(stream), codes,
                       lengths, n, lmax, l, W, fwd, rcw);
    GKM_CHECK_LAUNCH("gkm_pack_windows");
    return 0;
}

// Fill a u64 buffer with zeros with a kernel (graph-capture friendly).
__global__ void k_zero64(unsigned long long *p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

GKM_API int gkm_zero_u64(void *p, int64_t n, void *stream)
{
    if (n <= 0)
        return 0;
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero64, dim3((unsigned)blocks), dim3(256), 0, S_(stream),
                       (unsigned long long *)p, n);
    GKM_CHECK_LAUNCH("gkm_zero_u64");
    return 0;
}

GKM_API int gkm_cross_hist(const void *xplanes, const int32_t *xlen, int B, int nwx,
                           const void *yplanes, const void *yplanes_rc, const int32_t *ylen,
                           const int32_t *ygroups, int S, int nwy, int l, int nm,
                           const uint8_t *wtab, int wtab_len, int rows_per_block, int nsplit,
                           int64_t col_min, int64_t row_base, void *hist, void *stream)
{
    int rc = check_l_nm(l, nm);
    if (rc)
        return rc;
    if (B < 0 || S < 0 || rows_per_block < 1 || nsplit < 1)
        return fail(-1, "gkm_cross_hist: bad sizes (B=%d S=%d rows_per_block=%d nsplit=%d)", B, S,
                    rows_per_block, nsplit);
    if (B == 0 || S == 0)
        return 0;
    CrossArgs a;
    a.xplanes = (const P2 *)xplanes;
    a.xlen = xlen;
    a.yplanes = (const P2 *)yplanes;
    a.yplanes_rc = (const P2 *)yplanes_rc;
    a.ylen = ylen;
    a.ygrp = ygroups;
    a.wtab = wtab;
    a.wtab_len = wtab ? wtab_len : 0;
    a.hist = (unsigned long long *)hist;
    a.B = B;
    a.S = S;
    a.nwx = nwx;
    a.nwy = nwy;
    a.l = l;
    a.nm = nm;
    a.rows_per_block = rows_per_block;
    a.nsplit = nsplit;
    a.col_min = col_min;
    a.row_base = row_base;
    int norient = yplanes_rc ? 2 : 1;
    size_t lds = (size_t)norient * nwy * WAVE * sizeof(P2) + (wtab ? (size_t)wtab_len : 0);
    if (lds > 160 * 1024)
        return fail(-2, "gkm_cross_hist: SV tile needs %zu bytes of LDS (SVs too long)", lds);
    dim3 grid((unsigned)((S + WAVE - 1) / WAVE), (unsigned)((B + rows_per_block - 1) / rows_per_block));
    cross_fn f = pick_cross(l, nm, wtab != nullptr);
    if (lds > 64 * 1024)
        if (hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return fail(-2, "could not raise the LDS limit to %zu bytes", lds);
    hipLaunchKernelGGL(f, grid, dim3(CROSS_THREADS), lds, S_(stream), a);
    GKM_CHECK_LAUNCH("gkm_cross_hist");
    return 0;
}

GKM_API int gkm_self_hist(const void *planes, const void *planes_rc, const int32_t *len, int n,
                          int nw, int l, int nm, const uint8_t *wtab, int wtab_len, int nchunks,
                          void *hist, void *stream)
{
    int rc = check_l_nm(l, nm);
    if (rc)
        return rc;
    if (n == 0)
        return 0;
    if (nchunks < 1)
        nchunks = 1;
    SelfArgs a;
    a.planes = (const P2 *)planes;
    a.planes_rc = (const P2 *)planes_rc;
    a.len = len;
    a.wtab = wtab;
    a.wtab_len = wtab ? wtab_len : 0;
    a.hist = (unsigned long long *)hist;
    a.n = n;
    a.nw = nw;
    a.l = l;
    a.nm = nm;
    a.nchunks = nchunks;
    size_t lds = 2 * (size_t)nw * sizeof(P2) + 16 * (SELF_THREADS / WAVE) * sizeof(unsigned long long) +
                 (wtab ? (size_t)wtab_len : 0);
    if (lds > 160 * 1024)
        return fail(-2, "gkm_self_hist: needs %zu bytes of LDS", lds);
    self_fn f = pick_self(l, nm, wtab != nullptr);
    if (lds > 64 * 1024)
        if (hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return fail(-2, "could not raise the LDS limit to %zu bytes", lds);
    hipLaunchKernelGGL(f, dim3((unsigned)n, (unsigned)nchunks), dim3(SELF_THREADS), lds, S_(stream), a);
    GKM_CHECK_LAUNCH("gkm_self_hist");
    return 0;
}

GKM_API int gkm_combine_kraw(const void *hist, int64_t n, int nm, const double *tab, double *out,
                             void *stream)
{
    if (n <= 0)
        return 0;
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_combine_kraw, dim3((unsigned)blocks), dim3(256), 0, S_(stream),
                       (const unsigned long long *)hist, n, nm, tab, out);
    GKM_CHECK_LAUNCH("gkm_combine_kraw");
    return 0;
}

GKM_API int gkm_combine_gram(const void *hist, int B, int S, int nm, const double *tab,
                             const double *dx, const double *ds, int normalize, int rbf,
                             double gamma, double *out, void *stream)
{
    int64_t n = (int64_t)B * S;
    if (n <= 0)
        return 0;
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_combine_gram, dim3((unsigned)blocks), dim3(256), 0, S_(stream),
                       (const unsigned long long *)hist, B, S, nm, tab, dx, ds, normalize, rbf,
                       gamma, out);
    GKM_CHECK_LAUNCH("gkm_combine_gram");
    return 0;
}

GKM_API int gkm_combine_score(const void *hist, int B, int S, int nm, const double *tab,
                              const double *dx, const double *ds, const double *coef,
                              int normalize, int rbf, double gamma, double bias, double *out,
                              void *stream)
{
    if (nm > 16)
        return fail(-1, "gkm_combine_score: nm > 16");
    if (B <= 0)
        return 0;
    hipLaunchKernelGGL(k_combine_score, dim3((unsigned)B), dim3(SCORE_THREADS), 0, S_(stream),
                       (const unsigned long long *)hist, S, nm, tab, dx, ds, coef, normalize, rbf,
                       gamma, bias, out);
    GKM_CHECK_LAUNCH("gkm_combine_score");
    return 0;
}

GKM_API int gkm_combine_impact(const void *hist, int n, int S, int nm, const double *tab,
                               const double *dref, const double *ds, const double *coef,
                               int normalize, int rbf, double gamma, double *out, void *stream)
{
    if (nm > 16)
        return fail(-1, "gkm_combine_impact: nm > 16");
    if (n <= 0)
        return 0;
    hipLaunchKernelGGL(k_combine_impact, dim3((unsigned)n), dim3(SCORE_THREADS), 0, S_(stream),
                       (const unsigned long long *)hist, n, S, nm, tab, dref, ds, coef, normalize,
                       rbf, gamma, out);
    GKM_CHECK_LAUNCH("gkm_combine_impact");
    return 0;
}

GKM_API int gkm_explain_coef(const void *hist, int B, int S, int nm, const double *tab,
                             const double *dx, const double *ds, const double *coef, int normalize,
                             int rbf, double gamma, double *ec, void *stream)
{
    int64_t n = (int64_t)B * S;
    if (n <= 0)
        return 0;
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_explain_coef, dim3((unsigned)blocks), dim3(256), 0, S_(stream),
                       (const unsigned long long *)hist, B, S, nm, tab, dx, ds, coef, normalize,
                       rbf, gamma, ec);
    GKM_CHECK_LAUNCH("gkm_explain_coef");
    return 0;
}

GKM_API int gkm_mutant_diag(const uint32_t *fwd, const uint32_t *rcw, const int32_t *len,
                            const uint8_t *codes, int lmax, int n, int l, const double *tab,
                            const uint8_t *wtab, const double *diag, double *out, void *stream)
{
    if (l < 1 || l > 15 || lmax < l)
        return fail(-1, "gkm_mutant_diag: need 1 <= l <= 15 and lmax >= l");
    if (n == 0)
        return 0;
    if (n > 65535)
        return fail(-1, "gkm_mutant_diag: at most 65535 rows per call, got %d", n);
    int W = lmax - l + 1;
    hipLaunchKernelGGL(k_mutant_diag, dim3((unsigned)lmax, (unsigned)n), dim3(MUT_THREADS), 0,
                       S_(stream), fwd, rcw, len, codes, lmax, n, W, l, tab, wtab, diag, out);
    GKM_CHECK_LAUNCH("gkm_mutant_diag");
    return 0;
}

static int sweep_launch(int mode, const void *xplanes, const int32_t *xlen, const uint8_t *xcodes,
                        int lmax, int B, int nwx, const void *yplanes, const void *yplanes_rc,
                        const int32_t *ylen, int S, int nwy, int l, int nm, int cmax,
                        const uint8_t *wtab, int wtab_len, int sv_per_block, const double *ta,
                        const double *tb, const double *ec, const double *tab, const double *coef,
                        const double *ds, const double *dx, const double *dmut, int normalize,
                        int rbf, double gamma, double *out, void *stream)
{
    int rc = check_l_nm(l, nm);
    if (rc)
        return rc;
    if (cmax >= nm)
        return fail(-1, "sweep: cmax (%d) must be < nm (%d)", cmax, nm);
    if (B == 0 || S == 0)
        return 0;
    if (sv_per_block < 1)
        sv_per_block = 1;
    SweepArgs a;
    a.xplanes = (const P2 *)xplanes;
    a.xlen = xlen;
    a.xcodes = xcodes;
    a.lmax = lmax;
    a.yplanes = (const P2 *)yplanes;
    a.yplanes_rc = (const P2 *)yplanes_rc;
    a.ylen = ylen;
    a.wtab = wtab;
    a.wtab_len = wtab ? wtab_len : 0;
    a.B = B;
    a.S = S;
    a.nwx = nwx;
    a.nwy = nwy;
    a.l = l;
    a.nm = nm;
    a.cmax = cmax;
    a.sv_per_block = sv_per_block;
    a.ta = ta;
    a.tb = tb;
    a.ec = ec;
    a.tab = tab;
    a.coef = coef;
    a.ds = ds;
    a.dx = dx;
    a.dmut = dmut;
    a.normalize = normalize;
    a.rbf = rbf;
    a.gamma = gamma;
    a.out = out;
    size_t dbl = 4 * (size_t)lmax + (mode == SW_ISM ? 5 * (size_t)lmax : 0) + 32;
    size_t lds = dbl * sizeof(double) + 64 * sizeof(unsigned long long) +
                 ((size_t)nwx + 2 * (size_t)nwy) * sizeof(P2) + (wtab ? (size_t)wtab_len : 0);
    if (lds > 160 * 1024)
        return fail(-2, "sweep: needs %zu bytes of LDS (sequence too long)", lds);
    sweep_fn f = pick_sweep(l, nm, mode);
    if (lds > 64 * 1024)
        if (hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return fail(-2, "could not raise the LDS limit to %zu bytes", lds);
    dim3 grid((unsigned)B, (unsigned)((S + sv_per_block - 1) / sv_per_block));
    hipLaunchKernelGGL(f, grid, dim3(SWEEP_THREADS), lds, S_(stream), a);
    GKM_CHECK_LAUNCH("sweep");
    return 0;
}

GKM_API int gkm_explain(int perturbation, const void *xplanes, const int32_t *xlen, int lmax,
                        int B, int nwx, const void *yplanes, const void *yplanes_rc,
                        const int32_t *ylen, int S, int nwy, int l, int nm, int cmax,
                        const uint8_t *wtab, int wtab_len, int sv_per_block, const double *ta,
                        const double *tb, const double *ec, double *out, void *stream)
{
    return sweep_launch(perturbation ? SW_PERT : SW_HYP, xplanes, xlen, nullptr, lmax, B, nwx,
                        yplanes, yplanes_rc, ylen, S, nwy, l, nm, cmax, wtab, wtab_len,
                        sv_per_block, ta, tb, ec, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                        0, 0.0, out, stream);
}

GKM_API int gkm_index_explain(int perturbation, const uint32_t *xwin, const int32_t *xlen, int B,
                              int Wx, int lmax, const uint32_t *offsets, const uint16_t *ids,
                              const uint8_t *vals, const uint32_t *records,
                              const uint32_t *masks, const uint8_t *levels,
                              int n_masks, int l, const uint8_t *wtab, int S, int tile,
                              int n_tiles, int chunk_windows, const double *ta, const double *tb,
                              const double *ec, double *out, void *stream)
{
    if (l < 1 || l > 13)
        return fail(-1, "gkm_index_explain: the l-mer index needs 1 <= l <= 13, got %d", l);
    if (tile < 1 || tile > 65536 || n_tiles != (S + tile - 1) / tile || chunk_windows < 1)
        return fail(-1, "gkm_index_explain: bad tiling (tile=%d n_tiles=%d S=%d chunk=%d)", tile,
                    n_tiles, S, chunk_windows);
    if (B < 0 || B > 65535 || Wx < 0 || lmax < Wx + l - 1)
        return fail(-1, "gkm_index_explain: need 0 <= B <= 65535 and lmax >= Wx + l - 1");
    if (B == 0 || Wx == 0 || S == 0 || n_masks <= 0)
        return 0;
    IndexExplainArgs a;
    a.xwin = xwin;
    a.xlen = xlen;
    a.offsets = offsets;
    a.ids = ids;
    a.vals = vals;
    a.records = records;
    a.masks = masks;
    a.levels = levels;
    a.wtab = wtab;
    a.ta = ta;
    a.tb = tb;
    a.ec = ec;
    a.lw = nullptr;
    a.dx = nullptr;
    a.normalize = 0;
    a.out = out;
    a.B = B;
    a.Wx = Wx;
    a.lmax = lmax;
    a.n_masks = n_masks;
    a.l = l;
    a.S = S;
    a.tile = tile;
    a.n_tiles = n_tiles;
    a.chunk_windows = chunk_windows;
    a.pert = perturbation != 0;
    const int lm = (l == 10 || l == 11 || l == 12) ? l : 15;
    size_t lds = (4 * ((size_t)chunk_windows + lm - 1) + 32) * sizeof(double);
    if (lds > 64 * 1024)
        return fail(-2, "gkm_index_explain: chunk_windows %d too large", chunk_windows);
    // one dimension, 8 ids per XCD slot (the kernel maps ids to (row, chunk))
    const int64_t n_pairs = (int64_t)((Wx + chunk_windows - 1) / chunk_windows) * B;
    if (n_pairs > ((int64_t)1 << 30))
        return fail(-1, "gkm_index_explain: %lld blocks exceed the launch limit",
                    (long long)n_pairs);
    dim3 grid((unsigned)(8 * ((n_pairs + 7) / 8)));
    typedef void (*ixexp_fn)(IndexExplainArgs);
    const int rk = record_kind(records, vals, n_tiles);
#define GKM_IXEXP(LT) (rk == 3 ? k_index_explain<LT, 3> : rk == 2 ? k_index_explain<LT, 2> \
                       : rk == 1 ? k_index_explain<LT, 1> : k_index_explain<LT, 0>)
    const ixexp_fn fn = l == 10 ? GKM_IXEXP(10) : l == 11 ? GKM_IXEXP(11)
                      : l == 12 ? GKM_IXEXP(12) : GKM_IXEXP(0);
#undef GKM_IXEXP
    hipLaunchKernelGGL(fn, grid, dim3(IXEXP_THREADS), lds, S_(stream), a);
    GKM_CHECK_LAUNCH("gkm_index_explain");
    return 0;
}

GKM_API int gkm_lmer_explain(int perturbation, const uint32_t *xwin, const int32_t *xlen, int B,
                             int Wx, int lmax, const double *lw, const double *dx, int normalize,
                             const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                             const uint8_t *wtab, int chunk_windows, const double *ta,
                             const double *tb, double *out, void *stream)
{
    if (l < 1 || l > 13)
        return fail(-1, "gkm_lmer_explain: the l-mer weights need 1 <= l <= 13, got %d", l);
    if (B < 0 || B > 65535 || Wx < 0 || lmax < Wx + l - 1 || chunk_windows < 1)
        return fail(-1, "gkm_lmer_explain: need 0 <= B <= 65535, lmax >= Wx + l - 1 and "
                        "chunk_windows >= 1");
    if (lw == nullptr || (normalize && dx == nullptr))
        return fail(-1, "gkm_lmer_explain: needs the l-mer weights (and dx when normalising)");
    if (B == 0 || Wx == 0 || n_masks <= 0)
        return 0;
    IndexExplainArgs a = {};
    a.xwin = xwin;
    a.xlen = xlen;
    a.masks = masks;
    a.levels = levels;
    a.wtab = wtab;
    a.ta = ta;
    a.tb = tb;
    a.lw = lw;
    a.dx = dx;
    a.normalize = normalize != 0;
    a.out = out;
    a.B = B;
    a.Wx = Wx;
    a.lmax = lmax;
    a.n_masks = n_masks;
    a.l = l;
    a.chunk_windows = chunk_windows;
    a.pert = perturbation != 0;
    const int lm = (l == 10 || l == 11 || l == 12) ? l : 15;
    size_t lds = (4 * ((size_t)chunk_windows + lm - 1) + 32) * sizeof(double);
    if (lds > 64 * 1024)
        return fail(-2, "gkm_lmer_explain: chunk_windows %d too large", chunk_windows);
    const int64_t n_pairs = (int64_t)((Wx + chunk_windows - 1) / chunk_windows) * B;
    if (n_pairs > ((int64_t)1 << 30))
        return fail(-1, "gkm_lmer_explain: %lld blocks exceed the launch limit",
                    (long long)n_pairs);
    dim3 grid((unsigned)(8 * ((n_pairs + 7) / 8)));
    typedef void (*ixexp_fn)(IndexExplainArgs);
    const ixexp_fn fn = l == 10 ? k_lmer_explain<10> : l == 11 ? k_lmer_explain<11>
                      : l == 12 ? k_lmer_explain<12> : k_lmer_explain<0>;
    hipLaunchKernelGGL(fn, grid, dim3(IXEXP_THREADS), lds, S_(stream), a);
    GKM_CHECK_LAUNCH("gkm_lmer_explain");
    return 0;
}

GKM_API int gkm_lmer_explain_sorted(int perturbation, const uint32_t *xwin, const int32_t *xloc,
                                    int n_win, const int32_t *xlen, int B, int Wx, int lmax,
                                    const double *lw, const double *dx, int normalize,
                                    const uint32_t *masks, const uint8_t *levels, int n_masks,
                                    int l, const uint8_t *wtab, int chunk_windows,
                                    const double *ta, const double *tb, double *out, void *stream)
{
    if (l < 1 || l > 13)
        return fail(-1, "gkm_lmer_explain_sorted: the l-mer weights need 1 <= l <= 13, got %d", l);
    if (B < 0 || B > 65535 || Wx < 0 || lmax < Wx + l - 1 || chunk_windows < 1 || n_win < 0 ||
        (int64_t)n_win > (int64_t)B * Wx)
        return fail(-1, "gkm_lmer_explain_sorted: need 0 <= B <= 65535, lmax >= Wx + l - 1, "
                        "chunk_windows >= 1 and n_win <= B * Wx");
    if (lw == nullptr || xloc == nullptr || (normalize && dx == nullptr))
        return fail(-1, "gkm_lmer_explain_sorted: needs the l-mer weights, xloc (and dx when "
                        "normalising)");
    if (B == 0 || Wx == 0 || n_win == 0 || n_masks <= 0)
        return 0;
    IndexExplainArgs a = {};
    a.xwin = xwin;
    a.xlen = xlen;
    a.masks = masks;
    a.levels = levels;
    a.wtab = wtab;
    a.ta = ta;
    a.tb = tb;
    a.lw = lw;
    a.dx = dx;
    a.normalize = normalize != 0;
    a.out = out;
    a.B = B;
    a.Wx = Wx;
    a.lmax = lmax;
    a.n_masks = n_masks;
    a.l = l;
    a.chunk_windows = chunk_windows;
    a.pert = perturbation != 0;
    const int64_t nchunk = ((int64_t)n_win + chunk_windows - 1) / chunk_windows;
    dim3 grid((unsigned)(8 * ((nchunk + 7) / 8)));
    typedef void (*sorted_fn)(IndexExplainArgs, const int32_t *, int);
    const sorted_fn fn = l == 10 ? k_lmer_explain_sorted<10> : l == 11 ? k_lmer_explain_sorted<11>
                       : l == 12 ? k_lmer_explain_sorted<12> : k_lmer_explain_sorted<0>;
    hipLaunchKernelGGL(fn, grid, dim3(IXEXP_THREADS), 0, S_(stream), a, xloc, n_win);
    GKM_CHECK_LAUNCH("gkm_lmer_explain_sorted");
    return 0;
}

GKM_API int gkm_ism(const void *xplanes, const int32_t *xlen, const uint8_t *xcodes, int lmax,
                    int B, int nwx, const void *yplanes, const void *yplanes_rc,
                    const int32_t *ylen, int S, int nwy, int l, int nm, const uint8_t *wtab,
                    int wtab_len, int sv_per_block, const double *tab, const double *coef,
                    const double *ds, const double *dx, const double *dmut, int normalize, int rbf,
                    double gamma, double *out, void *stream)
{
    // Hits: windows within one mismatch of the cutoff (count <= dmax + 1).
    return sweep_launch(SW_ISM, xplanes, xlen, xcodes, lmax, B, nwx, yplanes, yplanes_rc, ylen, S,
                        nwy, l, nm, nm - 1, wtab, wtab_len, sv_per_block, nullptr, nullptr, nullptr,
                        tab, coef, ds, dx, dmut, normalize, rbf, gamma, out, stream);
}

GKM_API int gkm_sync(void *stream)
{
    hipError_t e = hipStreamSynchronize(S_(stream));
    if (e != hipSuccess)
        return fail(-4, "gkm_sync: %s", hipGetErrorString(e));
    return 0;
}
