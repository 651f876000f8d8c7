/*
 * gkmhip.h -- C-ABI of the MI355X (gfx950) gapped k-mer engine
 * (libgkmhip.so, built from gkmsvm-lite_amd/csrc/gkm_kernels.hip and
 * gkm_tables.hip).
 *
 * Every pointer argument is a DEVICE pointer owned by the caller (in the
 * Python package: torch tensors on the ROCm device), except where noted.
 * `stream` is a hipStream_t (NULL = the default stream).  Every call is
 * asynchronous on `stream`; it returns 0 or a negative status, with the
 * message in gkm_last_error() (thread-local).  No call allocates device
 * memory.
 *
 * Data layout:
 *   planes / planes_rc  [n][nw] pairs of uint32 (bit-plane 0, bit-plane 1):
 *                       base t of a sequence (or of its reverse complement)
 *                       at bit t & 31 of word (t >> 5) + 1, word 0 zero;
 *                       nw = gkm_planes_words(Lmax).
 *   hist                [rows][cols][nm] uint64 mismatch histograms N[m].
 *   tab                 h[m], m = 0..l (+1 zero), float64: the kernel type's
 *                       pair weight by mismatch count.
 *   wtab                [Wmax/2 + 1] uint8 window weight by distance from a
 *                       sequence's centre (-t 4/-t 5), or NULL.
 *
 * Citations name the gkmsvm-lite function (src/gkmsvm/...) each entry point
 * replaces; the Python wrappers in gkmsvm-lite_amd/gkmsvm_amd keep the
 * reference's public API on top of these.
 */
#ifndef GKMHIP_H
#define GKMHIP_H

#include <stdint.h>

/* Entries the l-mer index's ids (and vals) arrays hold past their last
 * bucket: the index query reads up to this many entries ahead of a
 * bucket's overflow without branching around the loads. */
#define GKM_INDEX_PAD 16

#ifdef __cplusplus
extern "C" {
#endif

/* Thread-local message of the last failed call. */
const char *gkm_last_error(void);
/* ABI version (1). */
int gkm_version(void);
/* Plane words per sequence for sequences up to lmax bases: lmax / 32 + 4. */
int gkm_planes_words(int lmax);
/* Wait for `stream`. */
int gkm_sync(void *stream);
/* Zero n uint64 with a kernel (usable under stream capture). */
int gkm_zero_u64(void *p, int64_t n, void *stream);

/* ASCII DNA -> base codes: row b is raw[starts[b] .. starts[b] + lengths[b])
 * (starts int64, lengths int32, every row's bytes in raw); codes [n][lmax]
 * uint8 gets A=0 C=1 G=2 T=3 (either case), N and any other byte as A, A past
 * the row; bad [n] int32 = 1 where a row holds a byte other than ACGTN/acgtn.
 * Replaces encode_codes / to_codes over a batch (codec.py:40-137). */
int gkm_decode_ascii(const uint8_t *raw, const int64_t *starts, const int32_t *lengths, int n,
                     int lmax, uint8_t *codes, int32_t *bad, void *stream);

/* Base codes [n][lmax] uint8 (A=0 C=1 G=2 T=3; N stored as A) -> planes and
 * reverse-complement planes (planes_rc may be NULL).
 * Replaces PackedGkmKernel.pack / pack_both (kernels/packed.py:130-187). */
int gkm_pack_planes(const uint8_t *codes, const int32_t *lengths, int n, int lmax, int nw,
                    void *planes, void *planes_rc, void *stream);

/* Base codes -> 2-bit packed l-mer windows [n][lmax-l+1] (PAD 0xFFFFFFFF past
 * each row's windows); rcw: windows of the reverse complement from its first
 * base (may be NULL).  kernels/packed.py:40-53,130-169. */
int gkm_pack_windows(const uint8_t *codes, const int32_t *lengths, int n, int lmax, int l,
                     uint32_t *fwd, uint32_t *rcw, void *stream);

/* Mismatch histograms of rows x against columns y and their reverse
 * complements, ADDED into hist [B][S][nm] (zero it first): N[m] = number (or,
 * with wtab, weight sum wx*wy) of window pairs with m mismatches, m < nm.
 * ygroups [ceil(S/64)][2] int32 = min and max length of each 64-column group.
 * rows_per_block, nsplit: launch shape (rows per block, diagonal chunks).
 * col_min >= 0: only columns s >= row_base + b + col_min (self-Gram upper
 * triangle).  The integer content of K_raw in PackedGkmKernel.raw_packed
 * (kernels/packed.py:191-228; kernels/direct.py:24-42; kernels/weighted.py:47-69)
 * and LmerIndex.pairwise (kernels/lmer_index.py:239-268). */
int gkm_cross_hist(const void *xplanes, const int32_t *xlen, int B, int nwx,
                   const void *yplanes, const void *yplanes_rc, const int32_t *ylen,
                   const int32_t *ygroups, int S, int nwy, int l, int nm, const uint8_t *wtab,
                   int wtab_len, int rows_per_block, int nsplit, int64_t col_min,
                   int64_t row_base, void *hist, void *stream);

/* Self-kernel histograms of x against x and rc(x), ADDED into hist [n][nm].
 * raw_diag_packed (kernels/packed.py:234-238; kernels/direct.py:110-137;
 * kernels/weighted.py:91-118). */
int gkm_self_hist(const void *planes, const void *planes_rc, const int32_t *len, int n, int nw,
                  int l, int nm, const uint8_t *wtab, int wtab_len, int nchunks, void *hist,
                  void *stream);

/* out[i] = sum_m tab[m] hist[i][m] (K_raw), i < n. */
int gkm_combine_kraw(const void *hist, int64_t n, int nm, const double *tab, double *out,
                     void *stream);

/* out[b][s] = F(K_raw / max(sqrt(dx[b] ds[s]), 1e-10)), F = exp(gamma(k-1))
 * when rbf: one Gram tile (gram.py:82-106 tile normalise + finish). */
int gkm_combine_gram(const void *hist, int B, int S, int nm, const double *tab, const double *dx,
                     const double *ds, int normalize, int rbf, double gamma, double *out,
                     void *stream);

/* out[b] = sum_s coef[s] F(K_raw[b][s] / max(sqrt(dx ds), 1e-10)) + bias, a
 * fixed-order reduction per row.  svm.py:43-61 (_normalized_dot),
 * svm.py:368-405 (_sv_scores). */
int gkm_combine_score(const void *hist, int B, int S, int nm, const double *tab,
                      const double *dx, const double *ds, const double *coef, int normalize,
                      int rbf, double gamma, double bias, double *out, void *stream);

/* Mutation impact from hist rows [ref (n); alt (n)]: explain/impact.py:22-48
 * (_impact_cpu). */
int gkm_combine_impact(const void *hist, int n, int S, int nm, const double *tab,
                       const double *dref, const double *ds, const double *coef, int normalize,
                       int rbf, double gamma, double *out, void *stream);

/* GkmExplain effective coefficients ec [B][S] = coef / max(sqrt(dx ds),1e-10)
 * (x _rbf_scale for RBF): explain/core.py:203-216, explain/common.py:27-32. */
int gkm_explain_coef(const void *hist, int B, int S, int nm, const double *tab, const double *dx,
                     const double *ds, const double *coef, int normalize, int rbf, double gamma,
                     double *ec, void *stream);

/* K_raw(mutant, mutant) [4][lmax][n] for every single-base substitution, from
 * the packed windows (gkm_pack_windows), the codes and diag = K_raw(x, x).
 * raw_diag_mutants (kernels/packed.py:244-253; kernels/direct.py:141-201;
 * kernels/weighted.py:122-197).  n <= 65535. */
int gkm_mutant_diag(const uint32_t *fwd, const uint32_t *rcw, const int32_t *len,
                    const uint8_t *codes, int lmax, int n, int l, const double *tab,
                    const uint8_t *wtab, const double *diag, double *out, void *stream);

/* GkmExplain per-SV attributions ADDED into out [B][4][lmax]: hypothetical
 * (perturbation = 0: ta = alpha, tb = kappa) or perturbation (ta = gamma,
 * tb = kappa deltas) over window pairs with count <= cmax; ec from
 * gkm_explain_coef.  explain/cpu.py:14-65 (_explain_cpu), explain/cuda.py:22-171
 * (fused_explain).  S / sv_per_block <= 65535. */
int gkm_explain(int perturbation, const void *xplanes, const int32_t *xlen, int lmax, int B,
                int nwx, const void *yplanes, const void *yplanes_rc, const int32_t *ylen, int S,
                int nwy, int l, int nm, int cmax, const uint8_t *wtab, int wtab_len,
                int sv_per_block, const double *ta, const double *tb, const double *ec,
                double *out, void *stream);

/* ISM: out [B][4][lmax] ADDED with sum_s coef_s (F(k_mut) - F(k_ref)) for every
 * substitution (0 for the reference base); nm = dmax + 2; tab = h (l + 2
 * entries, zero past dmax); ds, dx = K_raw self-kernels of SVs and rows;
 * dmut from gkm_mutant_diag.  ism.py:194-251 (_ism_kernel) and
 * kernels/lmer_index_cpu.py:203-285 (_query_mutants_cpu). */
int gkm_ism(const void *xplanes, const int32_t *xlen, const uint8_t *xcodes, int lmax, int B,
            int nwx, const void *yplanes, const void *yplanes_rc, const int32_t *ylen, int S,
            int nwy, int l, int nm, const uint8_t *wtab, int wtab_len, int sv_per_block,
            const double *tab, const double *coef, const double *ds, const double *dx,
            const double *dmut, int normalize, int rbf, double gamma, double *out, void *stream);

/* ISM of one query through the l-mer index: out [4][lmax] (the query's
 * slice of [B][4][lmax]) SET to sum_s coef_s (F(k_mut) - F(k_ref)) at every
 * substitution of positions 0..lx-1 (the reference base's entries and those
 * past lx are not written).  K_raw(mutant, s) = K_raw(x, s) - R[p][s] +
 * K_raw(new, s): R[p][s] sums the parts C[i][s] of the windows covering p,
 * new holds those windows with the base substituted.  Two persistent index
 * queries (the format of gkm_index_score_queue; indexes with bucket records
 * and 8- or 16-bit counts, else -2) whose blocks finish their counts per SV:
 * the lx - l + 1 windows one by one into C [windows][S], then the 3 lx new
 * rows of at most l windows into per-tile sums of coef_s (F(k_mut) -
 * F(k_ref)), summed over tiles in order; every sum has a fixed order
 * (deterministic).  xwin: the query's packed windows; coef, ds: the SVs'
 * coefficients and K_raw(s, s) in index order; dx: the query's K_raw(x, x);
 * dmut: K_raw(mutant, mutant) of this query, entry (v lmax + p) dmut_stride
 * (gkm_mutant_diag's [4][lmax][B] from column b: stride B); dx and dmut may
 * be NULL without normalize; workspace gkm_index_ism_bytes bytes, 256-byte
 * aligned.  kernels/lmer_index_cpu.py:203-285 (_query_mutants_cpu) and
 * kernels/lmer_index.py:326-376 (LmerIndex.mutants). */
int64_t gkm_index_ism_bytes(int lx, int l, int S, int n_tiles);
int gkm_index_ism(const uint32_t *xwin, const uint8_t *codes, int lx, int lmax,
                  const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                  const uint32_t *records, int record_tables, const uint32_t *masks,
                  const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab, int S,
                  int tile, int n_tiles, int pack_counts, const double *tab, const double *coef,
                  const double *ds, const double *dx, const double *dmut, int dmut_stride,
                  int normalize, int rbf, double gamma, int n_blocks, void *workspace,
                  double *out, void *stream);

/* GkmExplain through the l-mer index, ADDED into out [B][4][lmax]: every
 * query window u (xwin [B][Wx], PAD past each row's windows) visits the l-mers
 * v = u ^ masks[t] (levels[t] mismatches, the levels 0..cmax whose explain
 * weight is non-zero); the index buckets of v in every SV tile (the
 * gkm_index_hist CSR) give s = sum_e ec[b][sv_e] * vals_e, which adds
 * s * ta[m] at matching positions (to base u_k; perturbation: to the three
 * other bases) and s * tb[m] at mismatching ones (to base v_k), times the
 * query window's weight (wtab, or 1).  records: NULL, the per-key records of
 * gkm_index_hist or -- unweighted with n_tiles <= 4, where an l-mer's buckets
 * are consecutive across tiles -- [4^l][16] uint32 l-mer records: the run's
 * start, the tiles' counts as bytes (255 = saturated: read the offsets), the
 * run's first 28 ids (uint16).  chunk_windows: windows per
 * block.
 * l <= 13, B <= 65535.  _explain_index_cpu, _distribute
 * (explain/cpu.py:68-86, 146-207); explain_lmer (explain/cuda.py:174-351). */
int gkm_index_explain(int perturbation, const uint32_t *xwin, const int32_t *xlen, int B, int Wx,
                      int lmax, const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                      const uint32_t *records, const uint32_t *masks, const uint8_t *levels,
                      int n_masks, int l,
                      const uint8_t *wtab, int S, int tile, int n_tiles, int chunk_windows,
                      const double *ta, const double *tb, const double *ec, double *out,
                      void *stream);

/* GkmExplain of a kernel without RBF through the SVs' l-mer weights, ADDED
 * into out [B][4][lmax]: lw [4^l] = V, V[v] = sum over SVs of
 * coef / sqrt(K_raw(sv, sv)) times the SV windows (forward and RC) equal to
 * v, each with its window weight (gkm_lmer_weights); every query window u
 * visits v = u ^ masks[t] as gkm_index_explain does, with s = V[v], and each
 * row is scaled by 1 / sqrt(max(dx[b], 1e-20)) when normalize (dx [B]
 * K_raw(x, x)).  l <= 13, B <= 65535.  _explain_lmer_cpu, _distribute
 * (explain/cpu.py:68-143); _explain_lmer, _explain_V (explain/lmer.py:31-150,
 * explain/common.py:87-98). */
int gkm_lmer_explain(int perturbation, const uint32_t *xwin, const int32_t *xlen, int B, int Wx,
                     int lmax, const double *lw, const double *dx, int normalize,
                     const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                     const uint8_t *wtab, int chunk_windows, const double *ta, const double *tb,
                     double *out, void *stream);

/* gkm_lmer_explain over the windows of every row sorted by l-mer: xwin
 * [n_win] ascending (PAD, 0xFFFFFFFF, last), xloc [n_win] each key's
 * b * Wx + i in the rows' [B][Wx] windows, xlen / dx / out as there.  A block
 * takes chunk_windows consecutive keys -- which share their high digits, so
 * the V entries of their neighbourhoods share cache lines -- and each window
 * adds its attributions to out [B][4][lmax] with FP64 atomics.  The same
 * values as gkm_lmer_explain up to the order of FP64 additions.
 * explain/lmer.py:31-188 (_explain_lmer_cpu / _explain_lmer_cuda with V). */
int gkm_lmer_explain_sorted(int perturbation, const uint32_t *xwin, const int32_t *xloc,
                            int n_win, const int32_t *xlen, int B, int Wx, int lmax,
                            const double *lw, const double *dx, int normalize,
                            const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                            const uint8_t *wtab, int chunk_windows, const double *ta,
                            const double *tb, double *out, void *stream);

/* The same histograms through an l-mer index of the SV windows, ADDED into
 * hist [B][S][nm]: xwin [B][Wx] packed query windows (gkm_pack_windows), in any
 * order within a row (PAD last), with xpos [B][Wx] their positions in the row
 * (needed for window weights; NULL: in order); the
 * index is a CSR over keys lmer * n_tiles + tile (offsets [4^l n_tiles + 1]
 * uint32; ids uint16 SV within its tile of `tile` SVs; vals uint8 window
 * weights or NULL) of every SV window's l-mer, forward and RC; ids and vals
 * hold GKM_INDEX_PAD entries past the last bucket (read ahead, never
 * counted), and at least that many in all; records: NULL
 * or [record_tables][4^l n_tiles][8] uint32 bucket records, one per key
 * (at most 4 GiB of tables in all), 16 uint16 halfwords each -- one
 * 32-byte load per probe.  Without vals: halfword 0 = min(count, 0xFFFF); a
 * bucket of at most 15 entries has its ids in halfwords 1..count and 0xFFFF
 * in halfwords count + 1..15 (the kernel adds those slots to LDS words past
 * the tile's counts, which nothing reads), a longer one its first 11 ids in halfwords 1..11, its start (uint32) in word 6
 * and its count in word 7.  With vals: word 0 start, word 1 count, the first 8
 * ids in halfwords 4..11 and their uint8 weights in bytes 24..31.  Table q (1 <=
 * record_tables <= l) holds them at the key of the l-mer with its digits 0
 * and q swapped (table 0: in key order).  masks / levels [n_masks]: the XOR
 * neighbourhood of mismatch counts < nm with non-zero weight; levels[t] =
 * mismatch count | (record table of mask t) << 4, the table clamped to
 * record_tables - 1 (any table gives the same counts; a table q shared by
 * masks that differ in bit 0 of digit q puts their two records in one
 * 128-byte line).  partial: caller scratch of ceil(Wx / chunk_windows) * B * S * nm
 * counts, uint16 with pack_counts, else uint32.  With vals (weighted),
 * pack_counts 1 = two 16-bit counts per LDS word with the same kind of exact
 * check (field sums against the weight added; a block that carried walks its
 * chunk again with 32-bit counts), partials uint32 either way.  pack_counts
 * (unweighted; the caller keeps chunk_windows * windows per SV below 2^16): 1 =
 * 16-bit counts, two per LDS word; 2 = 8-bit counts, four per LDS word, with
 * an exact fallback -- a block whose byte sums show a carry walks its chunk
 * again with 16-bit counts, tile * ceil(nm/4) / ceil(nm/2) SVs per pass.
 * With 2 the partials are bytes: [nchunks][B][S][nm] low bytes, the same
 * again for the high bytes (written by fallback blocks only), then
 * [nchunks][B][n_tiles] flags (1: that block wrote high bytes) -- 2 * counts
 * + nchunks * B * n_tiles bytes in all.
 * l <= 13, B <= 65535, tile * ceil(nm / p) * 4 bytes <= 160 KiB (p = 1, 2, 4
 * counts per word).  hist NULL: only the partials (see gkm_index_reduce).
 * LmerIndex.pairwise (kernels/lmer_index.py:239-268, 400-441;
 * kernels/lmer_index_cpu.py:23-67). */
int gkm_index_hist(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B, int Wx,
                   const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                   const uint32_t *records, int record_tables, const uint32_t *masks,
                   const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab,
                   int S, int tile, int n_tiles, int chunk_windows, int pack_counts,
                   int64_t col_min, int64_t row_base, void *partial, void *hist, void *stream);

/* hist [B][S][nm] uint64 += the sum over window chunks of gkm_index_hist's
 * partial counts, chunks past a row's last window skipped.  count_bytes 4:
 * uint32, 2: uint16, 1: the 8-bit kernel's bytes (low bytes, high bytes, block
 * flags; S SVs in tiles of `tile`).  gkm_index_hist runs it itself unless its
 * hist is NULL; called on its own it lets the caller time the index query and
 * its sum apart.  The second half of LmerIndex.pairwise
 * (kernels/lmer_index.py:239-268), which accumulates its counts in place. */
int gkm_index_reduce(const void *partial, int count_bytes, const int32_t *xlen, int B, int Wx,
                     int l, int nm, int S, int tile, int chunk_windows, void *hist, void *stream);

/* ---- Exact scoring of kernels without RBF -----------------------------------
 * The decision value is linear in the histograms: out[b] = 2^-shift
 * sum_m tab[m] S_m / sqrt(dx[b]) + bias (no division without normalize; 0
 * where dx[b] <= 0), S_m = sum_s wfix[s] N_bs[m] summed exactly in 128-bit
 * integers, wfix[s] = round(coef_s / sqrt(K_raw(s, s)) * 2^shift) (coef_s
 * without normalize; 0 for an SV without windows), |wfix[s]| <= 2^61.  The
 * sums do not depend on the order of the pairs, window chunks or SV tiles
 * they are made of, so every path and launch shape gives the same bits.
 * svm.py:43-61 (_normalized_dot), svm.py:368-405 (_sv_scores). */

/* From full histograms hist [B][S][nm] uint64 (gkm_cross_hist or
 * gkm_index_reduce), nm <= 16. */
int gkm_combine_fixed(const void *hist, int B, int S, int nm, const double *tab, const double *dx,
                      const long long *wfix, int shift, int normalize, double bias, double *out,
                      void *stream);

/* Bytes of gkm_index_score's psum workspace: one int128 per (window chunk,
 * row, SV tile, level). */
int64_t gkm_index_psum_bytes(int B, int Wx, int nm, int n_tiles, int chunk_windows);

/* Decision values straight from the l-mer index (gkm_index_hist's arguments,
 * without col_min): each block reduces its tile's counts to the exact level
 * sums of its window chunk in psum (caller scratch of gkm_index_psum_bytes,
 * 16-byte aligned), and a second kernel adds them per row -- no histogram
 * or partial counts leave the chip.  LmerIndex.reduce
 * (kernels/lmer_index.py:275-298; kernels/lmer_index_cpu.py:70-120). */
int gkm_index_score(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B, int Wx,
                    const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                    const uint32_t *records, int record_tables, const uint32_t *masks,
                    const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab,
                    int S, int tile, int n_tiles, int chunk_windows, int pack_counts,
                    const long long *wfix, void *psum, const double *tab, const double *dx,
                    int shift, int normalize, double bias, double *out, void *stream);

/* gkm_index_score as one persistent launch over any number of rows: the
 * work items are (row, window chunk) pairs -- each row's ceil((xlen - l + 1) /
 * chunk_windows) chunks (none without windows), n_pairs in all, which the
 * caller counts from the lengths it uploaded -- times the SV tiles.  A
 * one-block plan kernel turns xlen into each row's first pair on the device;
 * then n_blocks blocks (one per CU: the tile fills its LDS) pop items from 8
 * per-XCD queues and take from the other queues when theirs is empty.
 * Equal-sized items make the list its own length bucketing: rows of every
 * length share one launch.  workspace: gkm_index_queue_bytes(B, n_pairs, ...)
 * bytes, 256-byte aligned (8 queue counters, first[B], then one int128 per
 * (pair, tile, level)).  The values are those of gkm_index_score, bit for
 * bit.  Replaces the row-block loop of GkmSVM._score_unbatched
 * (svm.py:584-598). */
int64_t gkm_index_queue_bytes(int B, int n_pairs, int n_tiles, int nm);
int gkm_index_score_queue(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B,
                          int Wx, const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                          const uint32_t *records, int record_tables, const uint32_t *masks,
                          const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab,
                          int S, int tile, int n_tiles, int chunk_windows, int pack_counts,
                          const long long *wfix, int n_pairs, int n_blocks, void *workspace,
                          const double *tab, const double *dx, int shift, int normalize,
                          double bias, double *out, void *stream);

/* gkm_index_score_queue over B rows whose last 4 * pairs are allele pairs
 * laid out in four blocks of `pairs` rows: reference alleles, then each
 * alternate allele's windows over its variant (the rows cut to windows
 * w0..w1, those covering a base that differs), then the reference's windows
 * at the same positions, then the alternate alleles themselves.  Only the
 * first B - pairs rows are probed (xwin and xpos hold those, n_pairs counts
 * their items, workspace is gkm_index_queue_bytes(B - pairs, ...)); alternate
 * allele p gets the level sums reference + its windows - the reference's
 * windows, which are its own level sums exactly (sums over windows of exact
 * integers), and its own dx.  Windows without position weights only (wtab
 * and xpos NULL when pairs > 0).  out holds all B rows (the cut rows score as
 * the short sequences they are).  Replaces the two full scorings of
 * GkmSVM.score_variants (svm.py:632-647: score(alt) - score(ref) over the
 * concatenated alleles) with one scoring of the reference and of the
 * windows a variant changes. */
int gkm_index_score_queue_pairs(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen,
                                int B, int Wx, const uint32_t *offsets, const uint16_t *ids,
                                const uint8_t *vals, const uint32_t *records, int record_tables,
                                const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                                int nm, const uint8_t *wtab, int S, int tile, int n_tiles,
                                int chunk_windows, int pack_counts, const long long *wfix,
                                int n_pairs, int n_blocks, void *workspace, const double *tab,
                                const double *dx, int shift, int normalize, double bias,
                                double *out, int pairs, void *stream);

/* gkm_index_hist as one persistent launch (the work list and queues of
 * gkm_index_score_queue): the same per-chunk partial counts in `partial`
 * (layout and size of gkm_index_hist's), then, hist non-NULL, ADDED into hist
 * [B][S][nm] -- the same integers.  workspace: gkm_index_queue_bytes(B, 0,
 * n_tiles, nm) bytes, 256-byte aligned.  For what the per-SV histograms feed
 * (RBF kernels, the explain coefficients); kernels/lmer_index.py:239-268
 * (LmerIndex.pairwise). */
int gkm_index_hist_queue(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B,
                         int Wx, const uint32_t *offsets, const uint16_t *ids, const uint8_t *vals,
                         const uint32_t *records, int record_tables, const uint32_t *masks,
                         const uint8_t *levels, int n_masks, int l, int nm, const uint8_t *wtab,
                         int S, int tile, int n_tiles, int chunk_windows, int pack_counts,
                         int n_pairs, int n_blocks, void *workspace, void *partial, void *hist,
                         void *stream);

/* ---- No-allocation variant --------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory or pass a workspace (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  gkm_index_hist_noalloc is
 * gkm_index_hist with hist required and the partial counts kept in a
 * GKM_SCRATCH_BYTES static device array that the library holds on each GPU
 * it runs on until the process ends; it calls no allocator.  A call whose
 * partials (gkm_index_partial_bytes) exceed the array fails with status -2
 * before launching anything.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race, so
 * serialise them.  The array needs no clearing: every partial a call reads
 * was written by that call. */
#define GKM_SCRATCH_BYTES (64u << 20)

/* GKM_SCRATCH_BYTES. */
int64_t gkm_scratch_bytes(void);

/* Bytes of gkm_index_hist's partial counts for these arguments (weighted:
 * vals != NULL). */
int64_t gkm_index_partial_bytes(int B, int Wx, int S, int nm, int n_tiles, int chunk_windows,
                                int weighted, int pack_counts);

int gkm_index_hist_noalloc(const uint32_t *xwin, const int32_t *xpos, const int32_t *xlen, int B,
                           int Wx, const uint32_t *offsets, const uint16_t *ids,
                           const uint8_t *vals, const uint32_t *records, int record_tables,
                           const uint32_t *masks, const uint8_t *levels, int n_masks, int l,
                           int nm, const uint8_t *wtab, int S, int tile, int n_tiles,
                           int chunk_windows, int pack_counts, int64_t col_min, int64_t row_base,
                           void *hist, void *stream);

/* ---- Exact l-mer tables of non-RBF models (gkm_tables.hip) ----------------
 * V, table  [4^l] float64 indexed by the packed l-mer (base k in bits 2k).
 * A         [4^l][l][4] float64 explain table: window u, digit k, delta d.
 * l <= 15 for every table entry point (memory: 8 * 4^l bytes per table,
 * 32 * l * 4^l for the explain table). */

/* V[v] = sum over entries e of the CSR bucket of l-mer v (offsets [4^l + 1]
 * int32, svid [n] int32 SV of each entry, vals [n] uint8 window weights or
 * NULL) of c[svid[e]] * vals[e].  Buckets hold their entries in SV order, so
 * V is deterministic.  sv_lmer_weights (lmer_table.py:35-57),
 * _explain_V (explain/common.py:87-96). */
int gkm_lmer_weights(const int32_t *offsets, const int32_t *svid, const uint8_t *vals,
                     const double *c, int l, double *V, void *stream);

/* In-place unnormalised radix-4 Walsh-Hadamard transform down the rows of a
 * [4^l][C] float64 array (C columns at once).  _wht4, _wht4_rows
 * (explain/table.py:41-75). */
int gkm_wht4(double *a, int l, int C, void *stream);

/* f[y] *= lam[m(y)] * scale, m(y) the number of non-zero 2-bit digits of y
 * (lam [l + 1]): the spectrum of the pair weight h[m].  With two gkm_wht4
 * calls around it this builds the score table w(u) = sum_v V[v] h[m(u, v)]
 * (build_lmer_table, lmer_table.py:60-98, by transforms instead of masked
 * marginals). */
int gkm_table_spectrum(double *f, int l, const double *lam, double scale, void *stream);

/* A[y][k][d] = fv[y] * scale * chi(y_k, d) * lam_d(m(y) - [y_k != 0]) with
 * chi(a, b) = (-1)^popcount(a & b), lam_0 = lamA, lam_1..3 = lamB ([l]
 * each); gkm_wht4 over the 4l columns then gives the explain table.
 * _fill_products, _build_explain_table (explain/table.py:77-113). */
int gkm_explain_fill(const double *fv, int l, const double *lamA, const double *lamB,
                     double scale, double *A, void *stream);

/* out[b] = sum_i wt_i table[xwin[b][i]] (over the row's xlen[b] - l + 1
 * windows; wt_i from wtab or 1), divided by sqrt(max(dx[b], 1e-10)) when
 * normalize, plus bias.  _table_raw, _score_lmer_table (svm.py:559-582). */
int gkm_table_score(const uint32_t *xwin, const int32_t *xlen, int B, int Wx, int l,
                    const uint8_t *wtab, const double *table, const double *dx, int normalize,
                    double bias, double *out, void *stream);

/* GkmExplain from the explain table into out [B][4][lmax] (every entry
 * written; 0 past a row): hypothetical (perturbation = 0) or perturbation
 * attributions, scaled by 1 / sqrt(max(dx[b], 1e-20)) when normalize.
 * B <= 65535.  _explain_table_cpu (explain/table.py:116-147). */
int gkm_table_explain(const uint32_t *xwin, const int32_t *xlen, int B, int Wx, int lmax, int l,
                      const uint8_t *wtab, const double *A, const double *dx, int normalize,
                      int perturbation, double *out, void *stream);

/* ISM deltas from the score table into out [B][4][lmax] (every entry
 * written): raw [B] = unnormalised table sums of the rows (gkm_table_score
 * with normalize 0, bias 0), codes [B][lmax] the reference bases, dx [B] and
 * dmut [4][lmax][B] (gkm_mutant_diag) the self-kernels when normalize.
 * B <= 65535.  _ism_table (ism.py:118-144). */
int gkm_table_ism(const uint32_t *xwin, const int32_t *xlen, const uint8_t *codes, int B, int Wx,
                  int lmax, int l, const uint8_t *wtab, const double *table, const double *raw,
                  const double *dx, const double *dmut, int normalize, double *out, void *stream);

/* DeltaSVM weights [4^l] (first base most significant) from the score table:
 * table[u] * wt0, divided by sqrt(max((h[0] + rc * h[m(u, rc u)]) wt0^2,
 * 1e-10)) when query_norm (h [l + 1]).  _deltasvm_weights_from_table,
 * _lmer_self_kernels (svm.py:667-705). */
int gkm_deltasvm_weights(const double *table, int l, double wt0, const double *h, int rc,
                         int query_norm, double *out, void *stream);

/* DeltaSVM scores out[b] = sum over the row's windows and the gapped k-mer
 * combos [n_combos][k] (int32 positions within a window of l) of
 * weights[k-mer] (first base most significant), plus the same over the
 * reverse complement when rc, plus bias.  k <= 15.
 * _score_strand_numba, DeltaSVM.__call__ (deltasvm.py:22-44, 107-147). */
int gkm_deltasvm_score(const uint8_t *codes, const int32_t *xlen, int B, int lmax, int l, int k,
                       const int32_t *combos, int n_combos, const double *weights, int rc,
                       double bias, double *out, void *stream);

/* ---- Peak microbenchmarks (gkm_peaks.hip) ---------------------------------
 * Measurement only, no counterpart in the reference: the HBM read bandwidth
 * and int32 VALU rate the hot kernels' rooflines are priced against.
 * gkm_peak_read streams `bytes` (a multiple of 16) from src with `blocks`
 * blocks of 256 threads; gkm_peak_valu runs `rounds` rounds of 24 int32 lane
 * operations (8 chains of XOR, add, popcount) per thread on `blocks` blocks of
 * 256 threads.  sink: 4 device bytes, written only if a sentinel comes up. */
int gkm_peak_read(const void *src, int64_t bytes, void *sink, int blocks, void *stream);
int gkm_peak_valu(int rounds, void *sink, int blocks, void *stream);
/* gkm_peak_gather reads `rounds` random 128-byte lines of the first `bytes`
 * (a multiple of 128) of src per lane pair, 32 bytes per lane, two lines per
 * lane pair in flight, with `blocks` blocks of 1,024 threads (rounds even). */
int gkm_peak_gather(const void *src, int64_t bytes, void *sink, int blocks, int rounds,
                    void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GKMHIP_H */
