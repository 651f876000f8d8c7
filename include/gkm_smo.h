/*
 * gkm_smo.h -- C-ABI of the SMO training solver (libgkm_smo.so, built from
 * gkmsvm-lite_amd/csrc/gkm_smo.c).  Host code: every pointer is host memory.
 * The kernel columns come from the caller through a callback (in the Python
 * package: computed on the GPU by gkmsvm_amd.train.KernelColumns).
 */
#ifndef GKM_SMO_H
#define GKM_SMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fill out[b * N + s] = K(idx[b], s), b < count: normalised, finished kernel
 * columns (K(i, i) = 1) over the N training sequences. */
typedef void (*gkm_columns_fn)(int count, const int *idx, int N, double *out, void *ctx);

/* Solve min 1/2 a'Qa + p'a s.t. y'a = 0, 0 <= a_t <= C_t, with Q_ts =
 * y_t y_s K(t mod N, s mod N) over n dual variables (C-SVC: n = N, p = -1;
 * epsilon-SVR: n = 2N, y = (+1, -1), p = (eps - z, eps + z)): second-order
 * working-set selection (Fan, Chen & Lin 2005), an LRU cache of cache_cols
 * columns, misses served in batches of up to `prefetch` columns (the missing
 * one and the most violating variables' columns).  Stops when the maximal
 * violation is below tol or after max_iter iterations.  alpha_out [n];
 * rho_out: f(x) = sum_t y_t a_t K(t mod N, x) - rho; stats_out [4] (may be
 * NULL): iterations, cache misses, columns computed, final violation.
 * Returns 0 (converged), 1 (max_iter reached) or -1 (out of memory).
 * Replaces csmo_solve (_csmo.c) as solver.py:261-392 (_smo_csolver) calls it. */
int gkm_smo_solve(int n, int N, const double *y, const double *p, const double *C, double tol,
                  int64_t max_iter, int cache_cols, int prefetch, gkm_columns_fn fn, void *ctx,
                  double *alpha_out, double *rho_out, double *stats_out);

#ifdef __cplusplus
}
#endif

#endif /* GKM_SMO_H */
