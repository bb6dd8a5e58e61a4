/* pgld -- MI355X-native two-locus LD library (C ABI).
 *
 * Drop-in boundary for pg_gpu's binned two-locus LD path.  The reference has
 * no C ABI: its "operator API" is the Python surface in pg_gpu/ld_pipeline.py,
 * pg_gpu/ld_statistics.py, pg_gpu/haplotype_kernels.py,
 * pg_gpu/genotype_kernels.py and pg_gpu/moments_ld.py working on CuPy arrays.
 * Each entry point below names the reference function(s) whose device work it
 * replaces; the Python package pg_gpu_amd binds them with ctypes and keeps the
 * reference's Python signatures.
 *
 * Conventions
 *   - Every pointer argument named d_* is device memory owned by the caller.
 *     The library allocates no device memory; its only device memory is the
 *     static 1 KiB scheduler array of pgld_binned_fused_noalloc.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *     Calls are asynchronous unless stated otherwise.
 *   - Return value: PGLD_OK or an error code; pgld_last_error() returns a
 *     thread-local message for the last failing call on the calling thread.
 *   - Genotype/haplotype matrices are int8, row-major (n_rows x n_sites,
 *     row stride `ld`), sample-major as in HaplotypeMatrix/GenotypeMatrix:
 *     haplotypes take 0/1 with -1 missing, genotypes 0/1/2 with -1 missing.
 */
#ifndef PGLD_H
#define PGLD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PGLD_OK = 0,
  PGLD_EINVAL = 1,   /* bad argument -> ValueError */
  PGLD_EDEVICE = 2,  /* HIP runtime failure -> RuntimeError */
  PGLD_EUNSUPPORTED = 3
};

enum { PGLD_MODE_HAP = 0, PGLD_MODE_GENO = 1 };

#define PGLD_MAX_POPS 4
#define PGLD_MAX_BINS 64
#define PGLD_SCHED_WORDS 256   /* words of the tile-scheduler buffer */

/* Kinds of pgld_pair_block. */
#define PGLD_PAIR_SIGMA_D2 0   /* haplotypes: unbiased D^2 / pi^2 ratio */
#define PGLD_PAIR_R2_NAIVE 1   /* haplotypes: frequency r^2 (per-site frequencies) */
#define PGLD_PAIR_R2_DOSAGE 2  /* genotypes: r^2 of mean-centred dosages */
#define PGLD_PAIR_R2 3         /* haplotypes: r^2 of the pair's counts (NaN if undefined) */

/* Library version (major*10000 + minor*100 + patch). */
int pgld_version(void);
/* Message for the last error on this thread ("" if none). */
const char* pgld_last_error(void);

/* Description of the packed sample panels shared by the counting kernels.
 * Sites are grouped in blocks of 32; per block, per population, per feature,
 * per 64-sample k-step there is one 1 KiB fragment of 4-bit E2M1 indicators
 * (1.0 or 0) in the operand layout of the gfx950 32x32x64 FP4 MFMA
 * (v_mfma_scale_f32_32x32x64_f8f6f4).  Features per population:
 *   haplotype: H = [x==1] and, if the population has missing calls, V = [x>=0]
 *   genotype:  I1 = [x==1], I2 = [x==2] and, with missing calls, V = [x>=0]. */
typedef struct {
  int32_t mode;                       /* PGLD_MODE_HAP / PGLD_MODE_GENO */
  int32_t n_pops;                     /* 1..4 */
  int64_t n_sites;
  int32_t pop_size[PGLD_MAX_POPS];    /* samples (haplotypes / individuals) */
  int32_t has_missing[PGLD_MAX_POPS]; /* 0/1, decides the V feature */
} pgld_panel_desc;

/* Bytes of the packed panel buffer for `desc` (host-side, synchronous). */
int64_t pgld_panel_bytes(const pgld_panel_desc* desc);

/* Per-site, per-population tallies and value validation.
 * Replaces the per-chunk `cp.any(haplotypes == -1)` / `cp.any(g < 0)` probes
 * (pg_gpu/ld_pipeline.py:158, :253) and the allele counts the no-missing
 * fast path needs.  Writes d_site_stats[p][k][site] (int32) with k = 0: count
 * of value 1, k = 1: count of value 2, k = 2: count of missing (-1); and
 * d_flags[0] |= 1 if any selected value is outside the mode's alphabet.
 * Both outputs must be zeroed by the caller. */
int pgld_scan_sites(const int8_t* d_mat, int64_t n_rows, int64_t n_sites, int64_t ld,
                    const int32_t* d_rows, const int32_t* pop_offsets /* host, n_pops+1 */,
                    const pgld_panel_desc* desc, int32_t* d_site_stats, int32_t* d_flags,
                    void* stream);

/* Pack the selected rows into MFMA operand panels.
 * Replaces the gathers `haplotypes[pop_indices, :]`, `haplotypes[:, idx_i]`
 * (pg_gpu/ld_pipeline.py:150-156) and `cp.ascontiguousarray(genotypes)`
 * (pg_gpu/ld_pipeline.py:242-247): the data is re-laid out once, not per chunk. */
int pgld_pack(const int8_t* d_mat, int64_t n_rows, int64_t n_sites, int64_t ld,
              const int32_t* d_rows, const int32_t* pop_offsets /* host */,
              const pgld_panel_desc* desc, int8_t* d_panels, void* stream);

/* Arguments of the binned LD kernels. */
typedef struct {
  const int8_t* d_panels;      /* from pgld_pack (fused kernel) */
  const int32_t* d_site_stats; /* from pgld_scan_sites */
  const int8_t* d_mat;         /* raw matrix (plain kernel only) */
  int64_t ld;                  /* raw matrix row stride (plain kernel only) */
  const int32_t* d_rows;       /* raw matrix rows per pop, concatenated (plain) */
  int32_t row_offset[PGLD_MAX_POPS + 1]; /* pop p uses d_rows[row_offset[p]..) */
  const double* d_pos;         /* n_sites, sorted, the cutoff coordinate (bp) */
  const double* d_coord;       /* n_sites, the binning coordinate (bp or Morgans),
                                  non-decreasing along the sites */
  int32_t coord_is_pos;        /* 1: d_coord == d_pos, distances are pos_j - pos_i */
  double max_dist;             /* pairs with pos_j <= pos_i + max_dist (float64) */
  const double* d_edges;       /* n_bins + 1 increasing bin edges */
  int32_t n_bins;
  const double* d_rtab;        /* 4*(max pop size + 1) reciprocal falling factorials */
  const int32_t* d_tiles;      /* (row block, col block) pairs of 32-site blocks, row <= col */
  int64_t n_tiles;
  long long* d_out;            /* [n_bins][n_stats + 1][3] fixed-point limbs, zeroed */
  int32_t grid;                /* workgroups for the persistent fused kernel (0 = auto) */
  int32_t n_lo[PGLD_MAX_POPS]; /* per pop: a lower bound on any pair's valid-sample count
                                  (e.g. size - 2 * max missing per site); 0 if unknown */
  int32_t with_r2;             /* 1 (one population only): two more sums per bin, the
                                  sum of r^2 over pairs where it is defined and the
                                  number of such pairs (mean r^2 = ratio) */
  uint32_t* d_sched;           /* PGLD_SCHED_WORDS zeroed words of tile-scheduler state
                                  for this launch (fused kernel), or NULL for a static
                                  tile schedule; launches sharing one buffer must not
                                  overlap.  The result does not depend on it. */
  int32_t form;                /* fused kernel form: 0 = every wave counts its own
                                  tiles and evaluates them; 1 = producer/consumer
                                  (counting waves hand tiles to statistics waves
                                  through LDS; one population, two haplotype
                                  populations or one genotype population, when its
                                  LDS plan fits -- otherwise form 0).  The result
                                  does not depend on it. */
} pgld_binned_args;

/* Number of statistics produced per bin for (mode, n_pops). */
int pgld_n_stats(int32_t mode, int32_t n_pops);

/* Host-side planner (no device work, synchronous): the band tile list of a
 * binned job.  Writes to `tiles` (int32 [n_tiles][2], host memory, `cap`
 * entries) the 32 x 32 site tiles (r, c), c >= r, holding at least one pair
 * i < j with pos[j] <= pos[i] + max_dist, in squares of `square` x `square`
 * tiles (row-major squares, row-major inside), and to *n_tiles their number;
 * with tiles == NULL only the counts are computed.  *n_pairs (optional)
 * receives the number of such pairs (one pass over the sites; pass NULL to
 * skip it).  pos: host float64, sorted.  Replaces the
 * band enumeration of iter_pairs_within_distance (pg_gpu/ld_pipeline.py:85-88)
 * at tile granularity, and its total pair count (:85-90). */
int pgld_band_tiles(const double* pos, int64_t n_sites, double max_dist, int32_t square,
                    int32_t* tiles, int64_t cap, int64_t* n_tiles, int64_t* n_pairs);

/* Fused binned LD: banded MFMA counting (FP4 indicators, exact FP32 sums) +
 * estimator epilogue + bin
 * reduction in one kernel.  Replaces the chunk loop of
 * pg_gpu/moments_ld.py:_compute_ld_sums (:248-343) -- pair enumeration
 * (ld_pipeline.py:57-122), counting (:130-263), the estimator batch
 * (ld_pipeline.py:271-349, haplotype_kernels.py:584-640,
 * genotype_kernels.py:881-936), digitize and add.at (:313-339) -- and the same
 * loop in HaplotypeMatrix.compute_ld_statistics_gpu_{single_pop,two_pops}
 * (haplotype_matrix.py:1592-1866).  Output: d_out[b][s][0..2] fixed-point
 * limbs of the bin sum of statistic s (value = (L2*2^64 + L1*2^32 + L0)/2^72),
 * d_out[b][n_out][0] = number of pairs in bin b, where n_out = n_stats, or
 * n_stats + 2 with with_r2 (r^2 sum and defined-pair count at n_stats and
 * n_stats + 1; haplotype r^2 as ld_statistics.r_squared,
 * pg_gpu/ld_statistics.py:190-217, genotype r^2 the pairwise-complete dosage
 * Pearson r^2, equal to Rogers-Huff r^2 without missing calls, :492-521).
 * Where the per-bin limbs of all bins do not fit in LDS (four populations,
 * more than ~50 bins) the bins are covered by consecutive launches over bin
 * ranges, with the static tile schedule. */
int pgld_binned_fused(const pgld_panel_desc* desc, const pgld_binned_args* args, void* stream);

/* pgld_binned_fused for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and output as
 * pgld_binned_fused, except: d_sched must be NULL -- the PGLD_SCHED_WORDS
 * tile-scheduler words (1 KiB, fixed whatever the job size) live in a
 * file-scope device array of the library, one copy per device held until
 * the process ends -- and d_out need not be zeroed: this call clears it and
 * the scheduler words with a fill kernel on `stream` (no memset node, no
 * allocation, no host synchronisation).  Calls that overlap in time on one
 * device share the scheduler words and race; serialise them (one stream). */
int pgld_binned_fused_noalloc(const pgld_panel_desc* desc, const pgld_binned_args* args,
                              void* stream);

/* The counting stage of pgld_binned_fused by itself (haplotypes): the same
 * tiles (d_tiles, n_tiles), schedule (d_sched, grid) and MFMA counting, with
 * the estimator epilogue replaced by a checksum.  d_sums (uint64 [n_pops][4],
 * zeroed) receives per population the sums over every slot of every listed
 * 32 x 32 tile of its four count fields: (n11, pA, pB, n) for two to four
 * populations (pA, pB, n the valid-both marginals, or with no missing call
 * the site tallies and the population size), (n11, n10, n01, n00) for one.
 * Replaces, as a measurement of its arithmetic, the counting of
 * compute_counts_for_pairs over the band (pg_gpu/ld_pipeline.py:130-175);
 * the other fields of args are ignored. */
int pgld_count_band(const pgld_panel_desc* desc, const pgld_binned_args* args,
                    unsigned long long* d_sums, void* stream);

/* Arguments of the ragged-window kernels (one population). */
typedef struct {
  const int8_t* d_panels;      /* from pgld_pack (fused kernel) */
  const int32_t* d_site_stats; /* from pgld_scan_sites */
  const int8_t* d_mat;         /* raw matrix (plain kernel only) */
  int64_t ld;                  /* raw matrix row stride (plain kernel only) */
  const int32_t* d_rows;       /* raw matrix rows of the population (plain kernel) */
  const double* d_rtab;        /* 4*(pop size + 1) reciprocal falling factorials */
  const int64_t* d_windows;    /* [n_windows][2] site ranges [lo, hi), lo <= hi */
  int64_t n_windows;
  const int32_t* d_work;       /* fused: [n_work][3] (window, row block, col block) with
                                  `tile`-site blocks covering each window's upper triangle */
  int64_t n_work;
  long long* d_out;            /* [n_windows][n_stats + 3][3] limbs, zeroed */
  int32_t grid;                /* workgroups (0 = auto) */
  int32_t n_lo[PGLD_MAX_POPS]; /* as in pgld_binned_args */
  uint32_t* d_sched;           /* not used: the window kernel cuts each XCD's eighth of
                                  the work list into equal turns for its waves */
  int32_t tile;                /* sites per side of the work items' blocks: 32 (or 0) for
                                  32 x 32 tiles, 16 for 16 x 16 tiles (haplotypes; the
                                  v_mfma_scale_f32_16x16x128_f8f6f4 path) */
  int32_t group;               /* work items per wave turn, 1..4096 (0 = 16 for 32-site
                                  items, 32 for 16-site items): each XCD's eighth of the
                                  work list is cut into R * m equal turns (R waves on the
                                  XCD, m its items / (R * group), rounded, at least 1);
                                  the result does not depend on it */
} pgld_window_args;

/* Per-window sums over every i < j pair of sites inside each window (no
 * distance cutoff): the single-population DD, Dz, pi2 (as pgld_binned_fused),
 * then the r^2 sum and the number of pairs where r^2 is defined, then the
 * pair count at column n_stats + 2.  Replaces, per window, the reference's
 * pair loop over a window's sites (ld_pipeline.py:57-175 with
 * max_dist >= window span) followed by ld_statistics.dd / r_squared
 * (ld_statistics.py:13-65, :190-217); all windows in one persistent launch
 * over the (window, tile) work list. */
int pgld_windows_fused(const pgld_panel_desc* desc, const pgld_window_args* args, void* stream);

/* Plain comparator of pgld_windows_fused: one workgroup per anchor site, one
 * thread per partner in the anchor's window, per-pair loops over samples of
 * the raw matrix.  d_work is not used.  Same output. */
int pgld_windows_plain(const pgld_panel_desc* desc, const pgld_window_args* args, void* stream);

/* The plain comparator: one thread per pair, a loop over samples of the raw
 * int8 matrix, FP64 estimators, per-workgroup bin partials.  Same output. */
int pgld_binned_plain(const pgld_panel_desc* desc, const pgld_binned_args* args, void* stream);

/* Dense counts for a rectangle of sites [i0, i0+ni) x [j0, j0+nj) of one
 * population through the MFMA path (used to check counts bit-exactly and to
 * time the counting stage).  d_counts is int32 [q][ni][nj] with q = 4
 * (n11, n10, n01, n00) for haplotypes or 9 in moments class order (n22, n21,
 * n20, n12, n11, n10, n02, n01, n00) for genotypes; i0 and
 * j0 must be multiples of 32. */
int pgld_count_block(const pgld_panel_desc* desc, const int8_t* d_panels,
                     const int32_t* d_site_stats, int32_t pop, int64_t i0, int64_t ni,
                     int64_t j0, int64_t nj, int32_t* d_counts, void* stream);

/* Dense per-pair LD values for the site rectangle [i0, i0+ni) x [j0, j0+nj)
 * of one population (i0, j0 multiples of 32), from the MFMA counts:
 * d_out float64 [2][ni][nj], plane 0 the value, plane 1 a 0/1 flag.
 * PGLD_PAIR_SIGMA_D2: D^2-numerator / pi2-numerator where n >= 4 and the
 *   latter is not 0 (else 0), flag = (n >= 4); replaces _tile_sigma_d2 /
 *   _build_sigma_d2_matrix (pg_gpu/ld_statistics.py:311-341, :734-758).
 * PGLD_PAIR_R2_NAIVE: frequency r^2 with per-site frequencies over each
 *   site's valid samples, 0 where undefined, flag = defined; replaces
 *   _tile_r2_naive and HaplotypeMatrix.pairwise_r2 (ld_statistics.py:301-309,
 *   haplotype_matrix.py:1092-1170).
 * PGLD_PAIR_R2_DOSAGE: r^2 of mean-centred dosages, missing calls zero,
 *   0 where a site has no variance; replaces _r2_matrix_diploid
 *   (ld_statistics.py:973-1021).
 * PGLD_PAIR_R2: r^2 from the pair's own counts (samples valid at both
 *   sites), NaN where undefined, flag = defined; replaces
 *   tally_gpu_haplotypes + ld_statistics.r_squared in windowed_r_squared
 *   (haplotype_matrix.py:1235-1317).
 * These feed zns / omega (ld_statistics.py:534-862). */
int pgld_pair_block(const pgld_panel_desc* desc, const int8_t* d_panels,
                    const int32_t* d_site_stats, int32_t pop, int32_t kind, int64_t i0,
                    int64_t ni, int64_t j0, int64_t nj, double* d_out, void* stream);

/* Heterozygosity statistics H_i_j (i <= j, het_names order) summed over
 * sites from the pgld_scan_sites tallies, as fixed-point limbs
 * d_out[e][0..2] (zeroed by the caller); d_nonfinite[e] counts sites whose
 * term was not finite.  Replaces _compute_heterozygosity
 * (pg_gpu/moments_ld.py:346-388). */
int pgld_heterozygosity(const pgld_panel_desc* desc, const int32_t* d_site_stats,
                        long long* d_out, int32_t* d_nonfinite, void* stream);

/* ---------------------------------------------------------------------------
 * Container passes (once per dataset)
 * ------------------------------------------------------------------------- */

#define PGLD_FILTER_HAP_BIALLELIC 0   /* no value >= 2, a 0 and a 1 present */
#define PGLD_FILTER_GENO_BIALLELIC 1  /* 0 < alt < 2 n_valid, n_valid >= 2 */

/* d_keep[s] (uint8, n_sites) = 1 if site s passes `rule` over the selected
 * rows d_rows[0..n_sel) (NULL: all n_rows rows; entries must lie in
 * [0, n_rows)), else 0.  PGLD_FILTER_HAP_BIALLELIC replaces the site test of
 * HaplotypeMatrix.apply_biallelic_filter (pg_gpu/haplotype_matrix.py:751-765);
 * PGLD_FILTER_GENO_BIALLELIC the one of GenotypeMatrix.apply_biallelic_filter
 * (pg_gpu/genotype_matrix.py:516-523) and, over the union of the selected
 * populations' rows, the genotype union filter of moments_ld._compute_ld_sums
 * (pg_gpu/moments_ld.py:256-276).  The caller compacts the kept columns. */
int pgld_biallelic_mask(const int8_t* d_mat, int64_t n_rows, int64_t n_sites, int64_t ld,
                        const int32_t* d_rows, int32_t n_sel, int32_t rule, uint8_t* d_keep,
                        void* stream);

/* Diploid genotypes from consecutive haplotype pairs (2i, 2i+1):
 * g = max(h0, 0) + max(h1, 0), -1 if either is missing.  d_geno is
 * (n_hap / 2) x n_sites with row stride ld_geno.  Replaces the conversion
 * loop of GenotypeMatrix.from_haplotype_matrix
 * (pg_gpu/genotype_matrix.py:263-291). */
int pgld_hap_to_geno(const int8_t* d_hap, int64_t n_hap, int64_t n_sites, int64_t ld_hap,
                     int8_t* d_geno, int64_t ld_geno, void* stream);

/* Two haplotypes per individual: rows 2i = [g >= 1], 2i+1 = [g >= 2], both -1
 * where g is missing.  Replaces GenotypeMatrix.to_haplotype_matrix
 * (pg_gpu/genotype_matrix.py:307-340). */
int pgld_geno_to_hap(const int8_t* d_geno, int64_t n_ind, int64_t n_sites, int64_t ld_geno,
                     int8_t* d_hap, int64_t ld_hap, void* stream);

/* ---------------------------------------------------------------------------
 * Per-pair API (not the binned hot path)
 * ------------------------------------------------------------------------- */

/* d_jmax[i] = number of sites j with pos[j] <= pos[i] + max_dist (float64),
 * i.e. the exclusive right end of anchor i's partners; positions sorted.
 * Replaces the searchsorted step of iter_pairs_within_distance
 * (pg_gpu/ld_pipeline.py:85-88). */
int pgld_pair_bounds(const double* d_pos, int64_t n_sites, double max_dist, int64_t* d_jmax,
                     void* stream);

/* Writes the n_pairs (i, j) pairs of anchors v_lo .. v_lo+n_anchors-1 in
 * (i, then j) order; d_offsets[k] is the exclusive prefix count of pairs of
 * the chunk's first k anchors (n_anchors + 1 entries).  Replaces the
 * per-chunk pair expansion of pg_gpu/ld_pipeline.py:103-121. */
int pgld_emit_pairs(const int64_t* d_offsets, int64_t v_lo, int64_t n_anchors, int64_t n_pairs,
                    int32_t* d_idx_i, int32_t* d_idx_j, void* stream);

/* Counts of population `pop` for an explicit pair list, from packed panels:
 * haplotype int32 [n][4] = (n11, n10, n01, n00), genotype int32 [n][9] =
 * (n00, n01, ..., n22), and the valid-sample count per pair (d_nvalid may be
 * NULL).  Replaces compute_counts_for_pairs (pg_gpu/ld_pipeline.py:130-175)
 * and compute_genotype_counts_for_pairs with _GENO_COUNTS_KERN
 * (pg_gpu/ld_pipeline.py:183-263).  Indices must be in [0, n_sites). */
int pgld_count_pairs(const pgld_panel_desc* desc, const int8_t* d_panels,
                     const int32_t* d_site_stats, int32_t pop, const int32_t* d_idx_i,
                     const int32_t* d_idx_j, int64_t n_pairs, int32_t* d_counts,
                     int32_t* d_nvalid, void* stream);

/* Raw haplotype estimator calls per pair.  d_counts: float64 [n][stride],
 * population p in columns 4p..4p+3 as (n11, n10, n01, n00); d_nvalid:
 * float64 [n][n_pops] or NULL; where it is NULL or an entry is negative the
 * sample size is the sum of that population's counts.  calls: host
 * int32 [n_calls][5] = (kind, p0, p1, p2, p3), kind 0 = DD(p0,p1),
 * 1 = Dz(p0,p1,p2), 2 = pi2(p0,p1,p2,p3).  d_out: float64 [n][n_calls].
 * Replaces ld_statistics.dd / dz / pi2 (pg_gpu/ld_statistics.py:13-129,
 * :1116-1617) and the per-case kernels of pg_gpu/haplotype_kernels.py:69-304. */
int pgld_hap_calls(const double* d_counts, int64_t stride, int32_t n_pops,
                   const double* d_nvalid, int64_t n_pairs, const int32_t* calls,
                   int32_t n_calls, double* d_out, void* stream);

/* Raw genotype estimator calls per pair.  d_counts: float64 [n][stride],
 * population p in columns 9p..9p+8 as the nine counts n00, n01, ..., n22
 * (first digit the genotype at locus i, as compute_genotype_counts_for_pairs
 * returns them); d_nvalid, calls and d_out as in pgld_hap_calls.  Replaces
 * compute_all_dd_geno / compute_all_dz_geno / compute_all_pi2_geno
 * (pg_gpu/genotype_kernels.py:561-878) and the per-case functions of
 * pg_gpu/ld_statistics_genotype.py:37-1505.  Where a population has too few
 * valid individuals for an estimator's falling factorial the value is 0. */
int pgld_geno_calls(const double* d_counts, int64_t stride, int32_t n_pops,
                    const double* d_nvalid, int64_t n_pairs, const int32_t* calls,
                    int32_t n_calls, double* d_out, void* stream);

/* The full genotype statistic set per pair (pgld_n_stats(PGLD_MODE_GENO,
 * n_pops) columns, ld_names order).  Replaces
 * compute_multi_pop_statistics_batch_geno (pg_gpu/genotype_kernels.py:881-936). */
int pgld_geno_stat_set(const double* d_counts, int64_t stride, int32_t n_pops,
                       const double* d_nvalid, int64_t n_pairs, double* d_out, void* stream);

/* sigma_D^2 = DD / pi2 per pair from one population's nine counts; NaN where
 * fewer than 4 individuals are valid or pi2 <= 0.  Replaces sigma_d2_geno
 * (pg_gpu/ld_statistics_genotype.py:1518-1598, genotype_kernels.py:91-169). */
int pgld_sigma_d2_geno(const double* d_counts, int64_t stride, int64_t n_pairs, double* d_out,
                       void* stream);

/* The full moments statistic set (pgld_n_stats(PGLD_MODE_HAP, n_pops)
 * columns, ld_names order) per pair.  Replaces
 * compute_single_pop_statistics_batch / compute_two_pop_statistics_batch
 * (pg_gpu/ld_pipeline.py:271-349) and compute_multi_pop_statistics_batch_hap
 * (pg_gpu/haplotype_kernels.py:584-640). */
int pgld_hap_stat_set(const double* d_counts, int64_t stride, int32_t n_pops,
                      const double* d_nvalid, int64_t n_pairs, double* d_out, void* stream);

/* r (which = 0), r^2 (1) or Lewontin's D' (2) per pair from 4-way counts;
 * NaN where undefined.  Replaces ld_statistics.r / r_squared / d_prime
 * (pg_gpu/ld_statistics.py:167-246, haplotype_kernels.py:20-62). */
int pgld_r_family(const double* d_counts, int64_t stride, const double* d_nvalid,
                  int64_t n_pairs, int32_t which, double* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PGLD_H */
