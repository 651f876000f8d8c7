// gkm_tables.hip -- exact l-mer tables of non-RBF gkm models (MI355X, gfx950).
//
// For a kernel without RBF the decision value is linear in the query's
// windows apart from the query's own normalisation:
//
//   f(x) = sum_i wt_i w(x_i) / sqrt(K_raw(x, x)) + b,
//   w(u) = sum_v V[v] h[m(u, v)],
//   V[v] = sum_s c_s (window-weighted occurrences of l-mer v in SV s and its RC)
//
// with c_s = alpha_s / sqrt(K_raw(s, s)).  m(u, v) is the number of non-zero
// 2-bit digits of u ^ v, so w = V (*) g is an XOR convolution over the 4^l
// l-mers with g(z) = h[m(z)].  The radix-4 Walsh-Hadamard transform H
// (H H = 4^l I) diagonalises it: w = H(H(V) . H(g)) / 4^l, and H(g)(y)
// depends only on the number of non-zero digits of y (a Krawtchouk sum the
// host evaluates exactly).  The GkmExplain table is the same convolution with
// per-(digit k, delta d) weights G_kd(z) = [z_k == d] t_d(m(z)):
// A[u][k][d] = H(H(V) . H(G_kd)) / 4^l, H(G_kd)(y) = chi(y_k, d) lam_d(m(y) - [y_k != 0]).
// gkmsvm-lite builds the score table from masked marginals on the CPU
// (lmer_table.py:60-98) and the explain table with Walsh-Hadamard transforms
// on the CPU (explain/table.py:14-114); here both are a handful of HBM passes.
//
// Data (device pointers, caller-owned):
//   V, T      [4^l] f64       SV l-mer weights, score table w
//   A         [4^l][l][4] f64 explain table: window u, digit k, delta d
//   xwin      [B][Wx] u32     packed query windows (base k at bits 2k; PAD
//                             0xFFFFFFFF past a row's windows)
//   out       per entry point
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "gkm_core.h"
#include "gkm_internal.h"
#include "gkmhip.h"

using namespace gkm;

namespace {

constexpr int TAB_THREADS = 256;
constexpr int TAB_WAVE = 64;
constexpr int TABLE_MAX_L = 15;

__device__ __forceinline__ int digit_weight(uint64_t y)
{
    return __popcll(((y >> 1) | y) & 0x5555555555555555ull);
}

__device__ __forceinline__ double wave_sum(double s)
{
    for (int off = TAB_WAVE / 2; off > 0; off >>= 1)
        s += __shfl_down(s, off, TAB_WAVE);
    return s;
}

// Sum over a TAB_THREADS block in a fixed order (deterministic); the result is
// valid in thread 0.
__device__ __forceinline__ double block_sum(double s, double *red)
{
    s = wave_sum(s);
    const int lane = threadIdx.x & (TAB_WAVE - 1), wave = threadIdx.x / TAB_WAVE;
    if (lane == 0)
        red[wave] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < TAB_THREADS / TAB_WAVE; k++)
            t += red[k];
    return t;
}

__device__ __forceinline__ void add_to(double &c0, double &c1, double &c2, double &c3, uint32_t j,
                                       double v)
{
    c0 += j == 0 ? v : 0.0;
    c1 += j == 1 ? v : 0.0;
    c2 += j == 2 ? v : 0.0;
    c3 += j == 3 ? v : 0.0;
}

}  // namespace

// V[v] = sum over the CSR bucket of l-mer v of c[sv] * weight; entries of a
// bucket are in SV order, so the sum order is fixed.
__global__ void k_lmer_weights(const int32_t *__restrict__ off, const int32_t *__restrict__ sv,
                               const uint8_t *__restrict__ vals, const double *__restrict__ c,
                               int64_t n, double *__restrict__ V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    double s = 0.0;
    const int e1 = off[v + 1];
    for (int e = off[v]; e < e1; e++)
        s += vals ? c[sv[e]] * (double)vals[e] : c[sv[e]];
    V[v] = s;
}

// The radix-4 butterfly of explain/table.py:42-57 on four values.
__device__ __forceinline__ void bfly4(double &a0, double &a1, double &a2, double &a3)
{
    const double s01 = a0 + a1, d01 = a0 - a1, s23 = a2 + a3, d23 = a2 - a3;
    a0 = s01 + s23;
    a1 = d01 + d23;
    a2 = s01 - s23;
    a3 = d01 - d23;
}

// One radix-4 pass (digit with stride h) down the rows of a [n][C] array.
__global__ void k_wht4_pass(double *__restrict__ a, int64_t nq, int64_t h, int C)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * C)
        return;
    const int64_t j = t / C;
    const int c = (int)(t - j * C);
    const int64_t i = (j / h) * 4 * h + j % h;
    const int64_t hs = h * C;
    double *p = a + i * C + c;
    double a0 = p[0], a1 = p[hs], a2 = p[2 * hs], a3 = p[3 * hs];
    bfly4(a0, a1, a2, a3);
    p[0] = a0;
    p[hs] = a1;
    p[2 * hs] = a2;
    p[3 * hs] = a3;
}

// Two radix-4 passes at once (digits of strides h and 4h): the same butterflies
// in the same order as two single passes, one HBM round trip.
__global__ void k_wht16_pass(double *__restrict__ a, int64_t n16, int64_t h, int C)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n16 * C)
        return;
    const int64_t j = t / C;
    const int c = (int)(t - j * C);
    const int64_t i = (j / h) * 16 * h + j % h;
    const int64_t hs = h * C;
    double *p = a + i * C + c;
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; q++)
        v[q] = p[q * hs];
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++)
        bfly4(v[4 * q2], v[4 * q2 + 1], v[4 * q2 + 2], v[4 * q2 + 3]);
#pragma unroll
    for (int q1 = 0; q1 < 4; q1++)
        bfly4(v[q1], v[q1 + 4], v[q1 + 8], v[q1 + 12]);
#pragma unroll
    for (int q = 0; q < 16; q++)
        p[q * hs] = v[q];
}

// f[y] *= lam[m(y)] * scale: the spectrum of g(z) = h[m(z)].
__global__ void k_table_spectrum(double *__restrict__ f, int64_t n,
                                 const double *__restrict__ lam, double scale)
{
    const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= n)
        return;
    f[y] = f[y] * (lam[digit_weight((uint64_t)y)] * scale);
}

// A'[y][k][d] = fv[y] scale chi(y_k, d) lam_d(m(y) - [y_k != 0]), chi(a, b) =
// (-1)^popcount(a & b); lam_0 = lamA, lam_1..3 = lamB.
__global__ void k_explain_fill(const double *__restrict__ fv, int64_t n, int l,
                               const double *__restrict__ lamA, const double *__restrict__ lamB,
                               double scale, double *__restrict__ A)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * l)
        return;
    const int64_t y = t / l;
    const int k = (int)(t - y * l);
    const uint32_t yk = (uint32_t)(y >> (2 * k)) & 3u;
    const int w = digit_weight((uint64_t)y) - (yk != 0);
    const double f = fv[y] * scale;
    const double fa = f * lamA[w], fb = f * lamB[w];
    double2 *o = reinterpret_cast<double2 *>(A + t * 4);
    o[0] = make_double2(fa, (yk & 1u) ? -fb : fb);
    o[1] = make_double2((yk & 2u) ? -fb : fb, (__popc(yk) & 1) ? -fb : fb);
}

// out[b] = sum_i wt_i T[x_i] (/ sqrt(max(dx, 1e-10)) when normalising) + bias:
// svm.py:559-582 (_table_raw, _score_lmer_table).  One block per row.
__global__ __launch_bounds__(TAB_THREADS) void k_table_score(
    const uint32_t *__restrict__ xwin, const int32_t *__restrict__ xlen, int Wx, int l,
    const uint8_t *__restrict__ wtab, const double *__restrict__ T,
    const double *__restrict__ dx, int normalize, double bias, double *__restrict__ out)
{
    __shared__ double red[TAB_THREADS / TAB_WAVE];
    const int b = blockIdx.x;
    const int W = xlen[b] - l + 1;
    const uint32_t nl = 1u << (2 * l);
    const uint32_t *row = xwin + (int64_t)b * Wx;
    double s = 0.0;
    for (int i = threadIdx.x; i < W; i += TAB_THREADS) {
        const uint32_t u = row[i];
        if (u < nl)
            s += wtab ? (double)window_weight(wtab, W, i) * T[u] : T[u];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        if (normalize)
            s = s / sqrt(fmax(dx[b], 1e-10));
        out[b] = s + bias;
    }
}

// GkmExplain from the explain table (explain/table.py:116-147): position p of
// row b gathers the windows covering it; window u at digit k adds A[u][k][d]
// to base u_k ^ d (d = 1..3) and A[u][k][0] to u_k (hypothetical) or to the
// three other bases (perturbation), times its window weight; the sum is scaled
// by 1 / sqrt(max(K_raw(x, x), 1e-20)).  A block covers EXPLAIN_TILE positions
// of one row with one thread per (position, digit k): every table row is
// fetched at once, and the l partial sums of a position are added in k order
// through LDS.  Every output entry is written.
constexpr int EXPLAIN_TILE = 64;

__global__ __launch_bounds__(EXPLAIN_TILE * 16) void k_table_explain(
    const uint32_t *__restrict__ xwin, const int32_t *__restrict__ xlen, int Wx, int lmax, int l,
    const uint8_t *__restrict__ wtab, const double *__restrict__ A,
    const double *__restrict__ dx, int normalize, int pert, double *__restrict__ out)
{
    __shared__ double part[16][EXPLAIN_TILE][4];
    const int b = blockIdx.y;
    const int pl = threadIdx.x, k = threadIdx.y;
    const int p = blockIdx.x * EXPLAIN_TILE + pl;
    const int L = xlen[b], W = L - l + 1;
    const uint32_t nl = 1u << (2 * l);
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    const int i = p - k;
    if (k < l && p < L && i >= 0 && i < W) {
        const uint32_t u = xwin[(int64_t)b * Wx + i];
        if (u < nl) {
            const double a = wtab ? (double)window_weight(wtab, W, i) : 1.0;
            const uint32_t uk = (u >> (2 * k)) & 3u;
            const double2 *r = reinterpret_cast<const double2 *>(A + ((int64_t)u * l + k) * 4);
            const double2 r01 = r[0], r23 = r[1];
            const double v0 = a * r01.x;
            add_to(c0, c1, c2, c3, uk ^ 1u, a * r01.y);
            add_to(c0, c1, c2, c3, uk ^ 2u, a * r23.x);
            add_to(c0, c1, c2, c3, uk ^ 3u, a * r23.y);
            if (pert) {
                c0 += uk != 0 ? v0 : 0.0;
                c1 += uk != 1 ? v0 : 0.0;
                c2 += uk != 2 ? v0 : 0.0;
                c3 += uk != 3 ? v0 : 0.0;
            } else {
                add_to(c0, c1, c2, c3, uk, v0);
            }
        }
    }
    part[k][pl][0] = c0;
    part[k][pl][1] = c1;
    part[k][pl][2] = c2;
    part[k][pl][3] = c3;
    __syncthreads();
    // thread (pl, j < 4) sums base j of position pl over the digits in order
    // (blockDim.y = max(l, 4); rows k >= l hold zeros)
    if (k < 4 && p < lmax) {
        double t = 0.0;
        for (int kk = 0; kk < l; kk++)
            t += part[kk][pl][k];
        const double inv = normalize ? 1.0 / sqrt(fmax(dx[b], 1e-20)) : 1.0;
        out[((int64_t)b * 4 + k) * lmax + p] = p < L ? t * inv : 0.0;
    }
}

// ISM from the score table (ism.py:118-144): substituting base v at p changes
// the windows covering p; their table weights give the raw change, and the
// mutant's own K_raw (dmut [4][lmax][B]) its normalisation.  0 for the
// reference base and past the row.  Every output entry is written.
__global__ __launch_bounds__(TAB_THREADS) void k_table_ism(
    const uint32_t *__restrict__ xwin, const int32_t *__restrict__ xlen,
    const uint8_t *__restrict__ codes, int B, int Wx, int lmax, int l,
    const uint8_t *__restrict__ wtab, const double *__restrict__ T,
    const double *__restrict__ raw, const double *__restrict__ dx,
    const double *__restrict__ dmut, int normalize, double *__restrict__ out)
{
    const int b = blockIdx.y;
    const int p = blockIdx.x * TAB_THREADS + threadIdx.x;
    if (p >= lmax)
        return;
    const int L = xlen[b], W = L - l + 1;
    const uint32_t *row = xwin + (int64_t)b * Wx;
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    double *o = out + (int64_t)b * 4 * lmax + p;
    if (p >= L || W <= 0) {
        for (int v = 0; v < 4; v++)
            o[v * (int64_t)lmax] = 0.0;
        return;
    }
    const uint32_t xb = codes[(int64_t)b * lmax + p] & 3u;
    const int i0 = max(0, p - l + 1), i1 = min(p, W - 1);
    for (int i = i0; i <= i1; i++) {
        const uint32_t u = row[i];
        const double a = wtab ? (double)window_weight(wtab, W, i) : 1.0;
        const int sh = 2 * (p - i);
        const double t_old = T[u];
        const uint32_t base = u & ~(3u << sh);
#pragma unroll
        for (uint32_t v = 0; v < 4; v++)
            if (v != xb)
                d[v] += a * (T[base | (v << sh)] - t_old);
    }
    const double r = raw[b];
    const double ref = normalize ? r / sqrt(fmax(dx[b], 1e-10)) : r;
#pragma unroll
    for (uint32_t v = 0; v < 4; v++) {
        double val = 0.0;
        if (v != xb) {
            if (normalize) {
                const double dm = dmut[((int64_t)v * lmax + p) * B + b];
                val = (r + d[v]) / sqrt(fmax(dm, 1e-10)) - ref;
            } else {
                val = d[v];
            }
        }
        o[v * (int64_t)lmax] = val;
    }
}

// DeltaSVM weights from the score table (svm.py:667-705): index idx has the
// first base most significant, the table's packed index the first base in the
// low digit.  Times the single window's weight wt0; with query_norm divided by
// sqrt(max(K_raw(u, u), 1e-10)), K_raw(u, u) = (h[0] + [rc] h[m(u, rc u)]) wt0^2.
__global__ void k_deltasvm_weights(const double *__restrict__ T, int64_t n, int l, double wt0,
                                   const double *__restrict__ h, int rc, int query_norm,
                                   double *__restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
    uint32_t u = 0;
    int m = 0;
    for (int k = 0; k < l; k++) {
        const uint32_t dk = (uint32_t)(idx >> (2 * (l - 1 - k))) & 3u;
        const uint32_t dm = (uint32_t)(idx >> (2 * k)) & 3u;  // base at position l-1-k
        u |= dk << (2 * k);
        m += dk != 3u - dm;
    }
    double w = T[u] * wt0;
    if (query_norm) {
        const double selfk = (h[0] + (rc ? h[m] : 0.0)) * wt0 * wt0;
        w = w / sqrt(fmax(selfk, 1e-10));
    }
    out[idx] = w;
}

// DeltaSVM scoring (deltasvm.py:22-44, 107-147): every window of each strand
// adds the weight of each gapped k-mer combo (first base most significant);
// the RC strand reads 3 - x[L - 1 - t].  out[b] = fwd + rc + bias.
__global__ __launch_bounds__(TAB_THREADS) void k_deltasvm_score(
    const uint8_t *__restrict__ codes, const int32_t *__restrict__ xlen, int lmax, int l, int k,
    const int32_t *__restrict__ combos, int n_combos, const double *__restrict__ weights, int rc,
    double bias, double *__restrict__ out)
{
    __shared__ double red[TAB_THREADS / TAB_WAVE];
    const int b = blockIdx.x;
    const int L = xlen[b];
    const int W = L - l + 1;
    const uint8_t *x = codes + (int64_t)b * lmax;
    double sf = 0.0, sr = 0.0;
    for (int w = threadIdx.x; w < W; w += TAB_THREADS) {
        for (int c = 0; c < n_combos; c++) {
            const int32_t *cb = combos + (int64_t)c * k;
            uint32_t fi = 0, ri = 0;
            for (int j = 0; j < k; j++) {
                const int t = w + cb[j];
                fi = fi * 4u + (x[t] & 3u);
                if (rc)
                    ri = ri * 4u + (3u - (x[L - 1 - t] & 3u));
            }
            sf += weights[fi];
            if (rc)
                sr += weights[ri];
        }
    }
    sf = block_sum(sf, red);
    __syncthreads();
    sr = block_sum(sr, red);
    if (threadIdx.x == 0)
        out[b] = sf + sr + bias;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

static int check_table_l(const char *what, int l)
{
    if (l < 1 || l > TABLE_MAX_L)
        return fail(-1, "%s: l must be in 1..%d, got %d", what, TABLE_MAX_L, l);
    return 0;
}

static inline unsigned blocks_for(int64_t n, int threads)
{
    return (unsigned)((n + threads - 1) / threads);
}

GKM_API int gkm_lmer_weights(const int32_t *offsets, const int32_t *svid, const uint8_t *vals,
                             const double *c, int l, double *V, void *stream)
{
    if (int rc = check_table_l("gkm_lmer_weights", l))
        return rc;
    const int64_t n = (int64_t)1 << (2 * l);
    hipLaunchKernelGGL(k_lmer_weights, dim3(blocks_for(n, 256)), dim3(256), 0, S_(stream),
                       offsets, svid, vals, c, n, V);
    GKM_CHECK_LAUNCH("gkm_lmer_weights");
    return 0;
}

GKM_API int gkm_wht4(double *a, int l, int C, void *stream)
{
    if (int rc = check_table_l("gkm_wht4", l))
        return rc;
    if (C < 1)
        return fail(-1, "gkm_wht4: need C >= 1, got %d", C);
    const int64_t n = (int64_t)1 << (2 * l);
    int dig = 0;
    for (; dig + 1 < l; dig += 2) {
        const int64_t h = (int64_t)1 << (2 * dig);
        hipLaunchKernelGGL(k_wht16_pass, dim3(blocks_for(n / 16 * C, 256)), dim3(256), 0,
                           S_(stream), a, n / 16, h, C);
        GKM_CHECK_LAUNCH("gkm_wht4");
    }
    if (dig < l) {
        const int64_t h = (int64_t)1 << (2 * dig);
        hipLaunchKernelGGL(k_wht4_pass, dim3(blocks_for(n / 4 * C, 256)), dim3(256), 0,
                           S_(stream), a, n / 4, h, C);
        GKM_CHECK_LAUNCH("gkm_wht4");
    }
    return 0;
}

GKM_API int gkm_table_spectrum(double *f, int l, const double *lam, double scale, void *stream)
{
    if (int rc = check_table_l("gkm_table_spectrum", l))
        return rc;
    const int64_t n = (int64_t)1 << (2 * l);
    hipLaunchKernelGGL(k_table_spectrum, dim3(blocks_for(n, 256)), dim3(256), 0, S_(stream), f, n,
                       lam, scale);
    GKM_CHECK_LAUNCH("gkm_table_spectrum");
    return 0;
}

GKM_API int gkm_explain_fill(const double *fv, int l, const double *lamA, const double *lamB,
                             double scale, double *A, void *stream)
{
    if (int rc = check_table_l("gkm_explain_fill", l))
        return rc;
    const int64_t n = (int64_t)1 << (2 * l);
    hipLaunchKernelGGL(k_explain_fill, dim3(blocks_for(n * l, 256)), dim3(256), 0, S_(stream), fv,
                       n, l, lamA, lamB, scale, A);
    GKM_CHECK_LAUNCH("gkm_explain_fill");
    return 0;
}

GKM_API int gkm_table_score(const uint32_t *xwin, const int32_t *xlen, int B, int Wx, int l,
                            const uint8_t *wtab, const double *table, const double *dx,
                            int normalize, double bias, double *out, void *stream)
{
    if (int rc = check_table_l("gkm_table_score", l))
        return rc;
    if (B < 0 || Wx < 0 || (normalize && !dx))
        return fail(-1, "gkm_table_score: need B, Wx >= 0 and dx when normalising");
    if (B == 0)
        return 0;
    hipLaunchKernelGGL(k_table_score, dim3((unsigned)B), dim3(TAB_THREADS), 0, S_(stream), xwin,
                       xlen, Wx, l, wtab, table, dx, normalize, bias, out);
    GKM_CHECK_LAUNCH("gkm_table_score");
    return 0;
}

GKM_API int gkm_table_explain(const uint32_t *xwin, const int32_t *xlen, int B, int Wx, int lmax,
                              int l, const uint8_t *wtab, const double *A, const double *dx,
                              int normalize, int perturbation, double *out, void *stream)
{
    if (int rc = check_table_l("gkm_table_explain", l))
        return rc;
    if (B < 0 || B > 65535 || lmax < l || Wx < lmax - l + 1 || (normalize && !dx))
        return fail(-1, "gkm_table_explain: need 0 <= B <= 65535, lmax >= l, Wx >= lmax - l + 1 "
                    "and dx when normalising (B=%d lmax=%d Wx=%d)", B, lmax, Wx);
    if (B == 0)
        return 0;
    // one thread per (position, digit); at least 4 digit rows for the 4 bases
    hipLaunchKernelGGL(k_table_explain, dim3(blocks_for(lmax, EXPLAIN_TILE), (unsigned)B),
                       dim3(EXPLAIN_TILE, l < 4 ? 4 : l), 0, S_(stream), xwin, xlen, Wx, lmax, l, wtab, A, dx,
                       normalize, perturbation, out);
    GKM_CHECK_LAUNCH("gkm_table_explain");
    return 0;
}

GKM_API int gkm_table_ism(const uint32_t *xwin, const int32_t *xlen, const uint8_t *codes, int B,
                          int Wx, int lmax, int l, const uint8_t *wtab, const double *table,
                          const double *raw, const double *dx, const double *dmut, int normalize,
                          double *out, void *stream)
{
    if (int rc = check_table_l("gkm_table_ism", l))
        return rc;
    if (B < 0 || B > 65535 || lmax < l || Wx < lmax - l + 1 || (normalize && (!dx || !dmut)))
        return fail(-1, "gkm_table_ism: need 0 <= B <= 65535, lmax >= l, Wx >= lmax - l + 1 "
                    "and dx, dmut when normalising (B=%d lmax=%d Wx=%d)", B, lmax, Wx);
    if (B == 0)
        return 0;
    hipLaunchKernelGGL(k_table_ism, dim3(blocks_for(lmax, TAB_THREADS), (unsigned)B),
                       dim3(TAB_THREADS), 0, S_(stream), xwin, xlen, codes, B, Wx, lmax, l, wtab,
                       table, raw, dx, dmut, normalize, out);
    GKM_CHECK_LAUNCH("gkm_table_ism");
    return 0;
}

GKM_API int gkm_deltasvm_weights(const double *table, int l, double wt0, const double *h, int rc,
                                 int query_norm, double *out, void *stream)
{
    if (int e = check_table_l("gkm_deltasvm_weights", l))
        return e;
    const int64_t n = (int64_t)1 << (2 * l);
    hipLaunchKernelGGL(k_deltasvm_weights, dim3(blocks_for(n, 256)), dim3(256), 0, S_(stream),
                       table, n, l, wt0, h, rc, query_norm, out);
    GKM_CHECK_LAUNCH("gkm_deltasvm_weights");
    return 0;
}

GKM_API int gkm_deltasvm_score(const uint8_t *codes, const int32_t *xlen, int B, int lmax, int l,
                               int k, const int32_t *combos, int n_combos, const double *weights,
                               int rc, double bias, double *out, void *stream)
{
    if (l < 1 || k < 1 || k > l || k > 15 || n_combos < 1 || B < 0 || lmax < 1)
        return fail(-1, "gkm_deltasvm_score: need 1 <= k <= min(l, 15), n_combos >= 1, B >= 0 "
                    "(l=%d k=%d n_combos=%d B=%d)", l, k, n_combos, B);
    if (B == 0)
        return 0;
    hipLaunchKernelGGL(k_deltasvm_score, dim3((unsigned)B), dim3(TAB_THREADS), 0, S_(stream),
                       codes, xlen, lmax, l, k, combos, n_combos, weights, rc, bias, out);
    GKM_CHECK_LAUNCH("gkm_deltasvm_score");
    return 0;
}
