/*
 * gkm_smo.c -- SMO for gkm-SVM training over kernel columns computed on
 * demand (host code; the columns come from the GPU through a callback).
 *
 * Solves  min_a  1/2 a'Qa + p'a   s.t.  y'a = 0,  0 <= a_t <= C_t,
 * Q_ts = y_t y_s K(t mod N, s mod N), for n dual variables over N training
 * sequences (C-SVC: n = N, p = -1; epsilon-SVR: n = 2N, y = (+1, -1),
 * p = (eps - z, eps + z)).  The kernel is normalised (K(i, i) = 1).
 *
 * Working-set selection is the second-order rule of Fan, Chen & Lin (JMLR
 * 6:1889-1918, 2005): i maximises -y_t G_t over I_up, j minimises
 * -b_it^2 / a_it over t in I_low with b_it = -y_i G_i + y_t G_t > 0,
 * a_it = K_ii + K_tt - 2 K_it (tau when not positive).  The pair moves along
 * a_i += y_i d, a_j -= y_j d (keeping y'a), d = b / a clipped to both boxes,
 * and every gradient changes by y_t d (K_ti - K_tj).  The solver stops when
 * max_{I_up} -y G - min_{I_low} -y G < tol.  Columns live in an LRU cache of
 * `cache_cols` columns; a miss asks the callback for the missing column and
 * up to `prefetch` - 1 more (the most violating variables' columns that are
 * not cached), so the GPU computes columns in batches.  No shrinking.
 * gkmsvm-lite solver.py:88-392 and _csmo.c drive the same problem through
 * the same kind of callback.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SMO_API __attribute__((visibility("default")))

/* Fill out[b * N + s] = K(idx[b], s) for b < count (normalised, finished). */
typedef void (*gkm_columns_fn)(int count, const int *idx, int N, double *out, void *ctx);

typedef struct {
    int N, cap, used;
    int *slot_of;     /* [N] slot of each column, -1 when not cached */
    int *col_in;      /* [cap] column held by each slot, -1 when empty */
    int64_t *stamp;   /* [cap] last use */
    int64_t clock;
    double *data;     /* [cap][N] */
    int64_t misses, computed;
} Cache;

static int cache_init(Cache *c, int N, int cap)
{
    memset(c, 0, sizeof(*c));
    c->N = N;
    c->cap = cap;
    c->slot_of = (int *)malloc(sizeof(int) * (size_t)N);
    c->col_in = (int *)malloc(sizeof(int) * (size_t)cap);
    c->stamp = (int64_t *)calloc((size_t)cap, sizeof(int64_t));
    c->data = (double *)malloc(sizeof(double) * (size_t)cap * (size_t)N);
    if (!c->slot_of || !c->col_in || !c->stamp || !c->data)
        return -1;
    for (int i = 0; i < N; i++)
        c->slot_of[i] = -1;
    for (int s = 0; s < cap; s++)
        c->col_in[s] = -1;
    return 0;
}

static void cache_free(Cache *c)
{
    free(c->slot_of);
    free(c->col_in);
    free(c->stamp);
    free(c->data);
}

/* A slot for a new column: an empty one, else the least recently used one
 * whose column is not `keep` (the other column of the current pair). */
static int cache_victim(Cache *c, int keep)
{
    if (c->used < c->cap)
        return c->used++;
    int best = -1;
    for (int s = 0; s < c->cap; s++) {
        if (c->col_in[s] == keep)
            continue;
        if (best < 0 || c->stamp[s] < c->stamp[best])
            best = s;
    }
    return best;
}

typedef struct {
    int n, N;
    const double *y, *p, *C;
    double *a, *G;
    Cache cache;
    gkm_columns_fn fn;
    void *ctx;
    int prefetch;
    int *batch;      /* [prefetch] column ids of one callback */
    double *buf;     /* [prefetch][N] */
    double *score;   /* [n] scratch for prefetch ranking */
    int *order;      /* [n] */
} Smo;

static int in_up(const Smo *s, int t)
{
    return s->y[t] > 0 ? s->a[t] < s->C[t] : s->a[t] > 0;
}

static int in_low(const Smo *s, int t)
{
    return s->y[t] > 0 ? s->a[t] > 0 : s->a[t] < s->C[t];
}

/* Reorder the first cnt candidates so the k with the largest scores come
 * first (quickselect). */
static void select_top(double *score, int *col, int cnt, int k)
{
    int lo = 0, hi = cnt - 1;
    while (lo < hi && k > lo && k <= hi) {
        const double pivot = score[(lo + hi) / 2];
        int a = lo, b = hi;
        while (a <= b) {
            while (score[a] > pivot) a++;
            while (score[b] < pivot) b--;
            if (a <= b) {
                double ts = score[a]; score[a] = score[b]; score[b] = ts;
                int tc = col[a]; col[a] = col[b]; col[b] = tc;
                a++;
                b--;
            }
        }
        if (k <= b + 1)
            hi = b;
        else if (k >= a)
            lo = a;
        else
            break;
    }
}

/* The batch of a miss on column `col`: it and up to k - 1 uncached columns
 * whose variables violate the optimality conditions most (the largest -yG
 * in I_up, the smallest in I_low), the likeliest next working-set members. */
static int prefetch_list(Smo *s, int col, int k)
{
    int m = 0;
    s->batch[m++] = col;
    if (k <= 1)
        return m;
    int cnt = 0;
    for (int t = 0; t < s->n; t++) {
        const int c = t % s->N;
        if (c == col || s->cache.slot_of[c] >= 0)
            continue;
        const double v = -s->y[t] * s->G[t];
        double sc = -INFINITY;
        if (in_up(s, t))
            sc = v;
        if (in_low(s, t) && -v > sc)
            sc = -v;
        if (sc == -INFINITY)
            continue;
        s->score[cnt] = sc;
        s->order[cnt] = c;
        cnt++;
    }
    const int want = k - 1 < cnt ? k - 1 : cnt;
    select_top(s->score, s->order, cnt, want);
    for (int r = 0; r < want; r++) {
        const int c = s->order[r];
        int dup = 0;
        for (int q = 1; q < m; q++)  /* an SVR sequence appears twice */
            if (s->batch[q] == c)
                dup = 1;
        if (!dup)
            s->batch[m++] = c;
    }
    return m;
}

/* Column `col` of K (cached), keeping column `keep` resident. */
static const double *column(Smo *s, int col, int keep)
{
    Cache *c = &s->cache;
    int slot = c->slot_of[col];
    if (slot < 0) {
        c->misses++;
        const int m = prefetch_list(s, col, s->prefetch < c->cap - 1 ? s->prefetch : c->cap - 1);
        s->fn(m, s->batch, s->N, s->buf, s->ctx);
        c->computed += m;
        for (int b = 0; b < m; b++) {
            const int cb = s->batch[b];
            int sl = cache_victim(c, keep);
            if (c->col_in[sl] >= 0)
                c->slot_of[c->col_in[sl]] = -1;
            c->col_in[sl] = cb;
            c->slot_of[cb] = sl;
            c->stamp[sl] = ++c->clock - (b > 0 ? 1 : 0); /* prefetched: older */
            memcpy(c->data + (size_t)sl * s->N, s->buf + (size_t)b * s->N,
                   sizeof(double) * (size_t)s->N);
        }
        slot = c->slot_of[col];
    }
    c->stamp[slot] = ++c->clock;
    return c->data + (size_t)slot * s->N;
}

/* rho from the final gradient: the mean of y G over free variables, else the
 * middle of the bounds the variables at their box ends put on it. */
static double compute_rho(const Smo *s)
{
    double sum = 0.0, ub = INFINITY, lb = -INFINITY;
    int nfree = 0;
    for (int t = 0; t < s->n; t++) {
        const double yg = s->y[t] * s->G[t];
        if (s->a[t] > 0 && s->a[t] < s->C[t]) {
            sum += yg;
            nfree++;
        } else if (s->a[t] <= 0) {
            if (s->y[t] > 0)
                ub = yg < ub ? yg : ub;
            else
                lb = yg > lb ? yg : lb;
        } else {
            if (s->y[t] > 0)
                lb = yg > lb ? yg : lb;
            else
                ub = yg < ub ? yg : ub;
        }
    }
    if (nfree > 0)
        return sum / nfree;
    return (ub + lb) / 2;
}

/*
 * Solve the dual.  y, p, C: [n]; alpha_out: [n] (a_t; the caller forms
 * y_t a_t); rho_out: the decision offset (f(x) = sum_t y_t a_t K - rho);
 * stats_out (may be NULL): [4] iterations, column misses, columns computed,
 * final violation.  Returns 0, 1 when max_iter was reached, or -1 when
 * memory ran out.
 */
SMO_API int gkm_smo_solve(int n, int N, const double *y, const double *p, const double *C,
                          double tol, int64_t max_iter, int cache_cols, int prefetch,
                          gkm_columns_fn fn, void *ctx, double *alpha_out, double *rho_out,
                          double *stats_out)
{
    const double tau = 1e-12;
    Smo s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.N = N;
    s.y = y;
    s.p = p;
    s.C = C;
    s.fn = fn;
    s.ctx = ctx;
    if (cache_cols < 2)
        cache_cols = 2;
    if (cache_cols > N)
        cache_cols = N;
    s.prefetch = prefetch < 1 ? 1 : prefetch;
    if (s.prefetch > cache_cols - 1 && cache_cols > 1)
        s.prefetch = cache_cols - 1;
    s.a = (double *)calloc((size_t)n, sizeof(double));
    s.G = (double *)malloc(sizeof(double) * (size_t)n);
    s.batch = (int *)malloc(sizeof(int) * (size_t)(s.prefetch + 1));
    s.buf = (double *)malloc(sizeof(double) * (size_t)(s.prefetch + 1) * (size_t)N);
    s.score = (double *)malloc(sizeof(double) * (size_t)n);
    s.order = (int *)malloc(sizeof(int) * (size_t)n);
    if (!s.a || !s.G || !s.batch || !s.buf || !s.score || !s.order ||
        cache_init(&s.cache, N, cache_cols) != 0) {
        free(s.a); free(s.G); free(s.batch); free(s.buf); free(s.score); free(s.order);
        cache_free(&s.cache);
        return -1;
    }
    for (int t = 0; t < n; t++)
        s.G[t] = p[t]; /* a = 0 */

    int64_t it = 0;
    double gap = INFINITY;
    int status = 1;
    for (; it < max_iter; it++) {
        /* i: the most violating variable in I_up */
        int i = -1;
        double gmax = -INFINITY;
        for (int t = 0; t < n; t++)
            if (in_up(&s, t) && -y[t] * s.G[t] >= gmax) {
                gmax = -y[t] * s.G[t];
                i = t;
            }
        if (i < 0) {
            status = 0;
            gap = 0.0;
            break;
        }
        const int ci = i % N;
        const double *Ki = column(&s, ci, -1);
        /* j: second-order choice in I_low; the stopping gap on the way */
        int j = -1;
        double gmin = INFINITY, best = INFINITY;
        for (int t = 0; t < n; t++) {
            if (!in_low(&s, t))
                continue;
            const double v = -y[t] * s.G[t];
            if (v < gmin)
                gmin = v;
            const double b = gmax - v;
            if (b > 0) {
                const int ct = t % N;
                double a = 2.0 - 2.0 * Ki[ct]; /* K_ii = K_tt = 1 */
                if (a <= 0)
                    a = tau;
                const double obj = -(b * b) / a;
                if (obj <= best) {
                    best = obj;
                    j = t;
                }
            }
        }
        gap = gmax - gmin;
        if (gap < tol || j < 0) {
            status = 0;
            break;
        }
        const int cj = j % N;
        const double *Kj = column(&s, cj, ci);
        Ki = column(&s, ci, cj); /* still cached: refresh the pointer after evictions */
        double a = 2.0 - 2.0 * Ki[cj];
        if (a <= 0)
            a = tau;
        double d = (gmax - (-y[j] * s.G[j])) / a;
        const double room_i = y[i] > 0 ? C[i] - s.a[i] : s.a[i];
        const double room_j = y[j] > 0 ? s.a[j] : C[j] - s.a[j];
        if (d > room_i)
            d = room_i;
        if (d > room_j)
            d = room_j;
        /* a_i += y_i d, a_j -= y_j d, snapped to the box ends they reach */
        s.a[i] += y[i] * d;
        s.a[j] -= y[j] * d;
        if (d == room_i)
            s.a[i] = y[i] > 0 ? C[i] : 0.0;
        if (d == room_j)
            s.a[j] = y[j] > 0 ? 0.0 : C[j];
        if (s.a[i] < 0) s.a[i] = 0;
        if (s.a[i] > C[i]) s.a[i] = C[i];
        if (s.a[j] < 0) s.a[j] = 0;
        if (s.a[j] > C[j]) s.a[j] = C[j];
        for (int t = 0; t < n; t++) {
            const int ct = t % N;
            s.G[t] += y[t] * d * (Ki[ct] - Kj[ct]);
        }
    }
    memcpy(alpha_out, s.a, sizeof(double) * (size_t)n);
    *rho_out = compute_rho(&s);
    if (stats_out) {
        stats_out[0] = (double)it;
        stats_out[1] = (double)s.cache.misses;
        stats_out[2] = (double)s.cache.computed;
        stats_out[3] = gap;
    }
    free(s.a); free(s.G); free(s.batch); free(s.buf); free(s.score); free(s.order);
    cache_free(&s.cache);
    return status;
}
