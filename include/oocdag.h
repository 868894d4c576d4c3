/*
 * oocdag.h -- C ABI of liboocdag, the MI355X out-of-core Cholesky library.
 *
 * Every entry point is plain C: pointers, sizes, a HIP stream passed as
 * void*.  Each one names the reference interface it replaces.  All return
 * an int status (OOCDAG_OK or a negative OOCDAG_E*) unless noted; the
 * factorization entry points additionally report LAPACK-style *info.
 *
 * Storage conventions
 *   host A      column-major N x N, leading dimension LDA >= N, lower
 *               triangle used (the reference's test_A, oocdag.cpp:20,251).
 *   tiles       "packed lower tiles": the reference's device panel Y
 *               (oocquark.cpp:256) for the whole matrix: tile (i,j), i >= j,
 *               of edge NB at element offset (i + j*bb - j*(j+1)/2) * TS*TS,
 *               each tile column-major with ld = TS, bb = ceil(N/NB), where
 *               TS = oocdag_tile_edge(NB) is NB rounded up to a multiple of
 *               64 (TS = NB when NB % 64 == 0).  Rows and columns past N, and
 *               past NB inside a TS x TS slot, are padding (0, with 1 on the
 *               diagonal of a diagonal tile).
 */
#ifndef OOCDAG_H
#define OOCDAG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OOCDAG_OK 0
#define OOCDAG_EARG (-1)    /* invalid argument */
#define OOCDAG_EBUDGET (-2) /* budget below two tile columns (oocdag.cpp:249-250) */
#define OOCDAG_ENOMEM (-3)  /* device or pinned host allocation failed */
#define OOCDAG_EHIP (-4)    /* HIP runtime error */
#define OOCDAG_ENODEV (-5)  /* no usable gfx950 device */

/* ---------------------------------------------------------------------------
 * Factorization, host array in / host array out.
 * Replaces: double Cholesky(Quark *quark, double *A, int N, int NB, int LDA,
 *                           size_t memsize)                  oocdag.cpp:18-112
 * Same arguments and meaning (the QUARK handle is gone: the library runs its
 * own HIP streams).  The lower triangle of A is overwritten with L, the strict
 * upper triangle is untouched.  memsize is the device budget in bytes; the
 * panel widths follow the reference planner (find_Yn, oocdag.cpp:478-492)
 * with room for two staged X tile columns instead of one when the budget
 * allows it (oocdag_cholesky_nx with nx = 1 runs the reference's schedule
 * exactly).  Returns the wall-clock seconds of the factorization
 * (oocdag.cpp:51,90,110), or a negative value on error.  *info: 0 success;
 * k > 0 the leading minor of order k is not positive definite (LAPACK
 * dpotrf convention; the reference only printed it, oocquark.cpp:95); < 0 an
 * OOCDAG_E* code.  Any NB >= 1, as the reference takes: the tile grid,
 * memBlock = memsize / (8 NB^2) tiles and the panel schedule are the
 * reference's; a tile whose edge is not a multiple of 64 is held on the
 * device in a slot of edge oocdag_tile_edge(NB) (NB rounded up to 64,
 * padded with an identity block), so the device footprint is
 * memBlock * 8 * oocdag_tile_edge(NB)^2 bytes.  The out-of-core path drives
 * four HIP streams; run the
 * process with GPU_MAX_HW_QUEUES >= 5 (HIP's default is 4) so that none of
 * them shares a hardware queue -- shared, a finished column's copy to the
 * host holds back the next update (about 4 % at n = 196608).
 * ------------------------------------------------------------------------- */
double oocdag_cholesky(double *A, int N, int NB, int LDA, size_t memsize, int *info);
/* The same with the number of staged X tile columns fixed: nx = 1 is the
 * reference's schedule (find_Yn, oocdag.cpp:478-492: one X column, the same
 * panels and A2Y/A2X copy volume, oocquark.cpp:253-319), nx = 2 double
 * buffers X so each copy overlaps the previous update, nx = 0 is the
 * default (2 when the budget leaves a panel column beside two X columns). */
double oocdag_cholesky_nx(double *A, int N, int NB, int LDA, size_t memsize, int nx, int *info);

/* ---------------------------------------------------------------------------
 * Factorization, host packed tiles in / out (out of core).
 * Replaces: the same Cholesky() panel loop (oocdag.cpp:18-112; staging
 *   A2Y/A2X/D2H oocquark.cpp:16-63,253-319) for a caller that keeps A on the
 *   host as packed lower tiles -- tile (i,j), i >= j, in a column-major slot
 *   of edge TS = oocdag_tile_edge(NB) at element offset
 *   (i + j*bb - j(j+1)/2)*TS*TS, bb = ceil(N/NB), the NB x NB tile in its
 *   top-left corner, padded past N and past NB with zeros and a unit
 *   diagonal (the device layout; oocdag_tiles_elems gives the size).  Every
 *   tile-column copy is then one contiguous DMA and nothing is packed on the
 *   host: T in pinned memory (oocdag_host_alloc) is
 *   copied directly; pageable T is staged through the library's pinned
 *   double-buffered slots with a parallel memcpy.  Same panel schedule,
 *   budget, return value and *info as oocdag_cholesky; the tiles are
 *   overwritten with L.
 * ------------------------------------------------------------------------- */
double oocdag_cholesky_host_tiles(double *T, int N, int NB, size_t memsize, int *info);
double oocdag_cholesky_host_tiles_nx(double *T, int N, int NB, size_t memsize, int nx,
                                     int *info);

/* PDMATGEN('S'/'H'/'F', DIAG, ISEED) straight into host packed lower tiles
 * (the layout above), generated on the GPU tile column group by group and
 * copied down; values bit-identical to oocdag_pdmatgen.
 * Replaces: pdmatgen (oocdag.h:85-90) for the host-tiles entry point. */
int oocdag_pdmatgen_host_tiles(double *T, int N, int NB, char aform, char diag, int iseed);

/* Pinned (page-locked) host memory for the host-tiles entry points:
 * hipHostMalloc / hipHostFree.  NULL on failure (oocdag_last_error). */
void *oocdag_host_alloc(size_t bytes);
int oocdag_host_free(void *p);

/* Panel widths (tile columns per Y panel) that oocdag_cholesky will use for
 * (N, NB, memsize): widths[0..return-1].  Returns the panel count, or < 0. */
int oocdag_panel_plan(int N, int NB, size_t memsize, int *widths, int maxp);
/* Same with the number of staged X tile columns fixed: nx = 1 is the
 * reference's find_Yn schedule (oocdag.cpp:478-492) exactly, nx = 2 the
 * double-buffered one; nx = 0 lets the library choose. */
int oocdag_panel_plan_nx(int N, int NB, size_t memsize, int nx, int *widths, int maxp);

/* ---------------------------------------------------------------------------
 * Input generator.
 * Replaces: pdmatgen(ICTXT, AFORM, DIAG, M, N, MB, NB, A, LDA, IAROW, IACOL,
 *   ISEED, IROFF, IRNUM, ICOFF, ICNUM, MYROW, MYCOL, NPROW, NPCOL) BIND(C)
 *   oocdag.h:85-90, LIN/pdmatgen.f:1-3, called at oocdag.cpp:298-302
 * Specialised to the 1x1 grid call of the reference (offsets 0, whole
 * matrix).  Values are bit-identical to PDMATGEN.  Computed on the GPU and
 * copied to A.  Returns the PDMATGEN INFO (0 ok, > 0 bad argument position)
 * or an OOCDAG_E* code.  AFORM: 'S'/'H' symmetric, 'F' hard-coded SPD form,
 * 'T'/'C' transposed, other general; DIAG: 'D' or 'N'.
 * ------------------------------------------------------------------------- */
int oocdag_pdmatgen(char aform, char diag, int M, int N, int MB, int NB, double *A, int LDA,
                    int iseed);

/* ---------------------------------------------------------------------------
 * HBM-resident factorization on packed lower tiles (no host traffic).
 * Replaces the in-core part of Cholesky() when the whole lower triangle fits
 * in the device budget (one Y panel: ooc_incore, oocquark.cpp:358-421).
 * dY: device pointer to oocdag_tiles_elems(N, NB) doubles.  stream: a
 * hipStream_t or NULL for the library's own stream.  Synchronous: returns
 * after the factorization finished; *info as for oocdag_cholesky.
 * ------------------------------------------------------------------------- */
int oocdag_potrf_tiles(double *dY, int N, int NB, void *stream, int *info);

/* Number of doubles of the packed lower tile storage for (N, NB). */
size_t oocdag_tiles_elems(int N, int NB);
/* Edge TS of the slot that holds one NB x NB tile (NB rounded up to a
 * multiple of 64), or OOCDAG_EARG for NB < 1. */
int oocdag_tile_edge(int NB);

/* PDMATGEN('S'|'H'|'F', DIAG, ISEED) written straight into packed lower tiles
 * in HBM (values identical to oocdag_pdmatgen; padding as described above). */
int oocdag_pdmatgen_tiles(double *dY, int N, int NB, char aform, char diag, int iseed,
                          void *stream);

/* Copy between packed lower tiles (device) and a column-major host array
 * (lower triangle incl. diagonal blocks' full NB x NB squares, like the
 * reference's A2Y / Y2A tile copies, oocquark.cpp:253-299). */
int oocdag_tiles_from_host(double *dY, const double *A, int N, int NB, int LDA);
int oocdag_tiles_to_host(const double *dY, double *A, int N, int NB, int LDA);

/* ---------------------------------------------------------------------------
 * Solve with the factor: A X = B, A = L L^T.
 * Replaces: dpotrs('L', N, nrhs, L, LDA, B, LDB, info) of the VERIFY block
 *           oocdag.cpp:431-441 (host MKL in the reference)
 * oocdag_potrs: host L (lower triangle, as left by oocdag_cholesky) and host
 * B (N x nrhs, overwritten by X); the factor must fit in device memory.
 * oocdag_potrs_tiles: device packed factor dY and device B (ldb >= N).
 * ------------------------------------------------------------------------- */
int oocdag_potrs(const double *L, int N, int LDL, double *B, int nrhs, int LDB);
int oocdag_potrs_tiles(const double *dY, int N, int NB, double *dB, int nrhs, int ldb,
                       void *stream);

/* ---------------------------------------------------------------------------
 * LIN scaled residual ||A - L L^T||_inf / (N * eps * ||A||_inf), eps = 2^-53.
 * Replaces: PDPOTRRV('L') + PDLAFCHK(AFORM, DIAG, ..., IASEED, ANORM, FRESID)
 *           LIN/pdpotrrv.f:240-314, LIN/pdlafchk.f:181-225
 * A is regenerated on the device from (aform, diag, iseed) exactly as
 * oocdag_pdmatgen produces it.  oocdag_llt_resid takes the host array left
 * by oocdag_cholesky (L in the lower triangle, A's strict upper triangle
 * kept; both count); oocdag_llt_resid_tiles the HBM-resident packed factor.
 * ------------------------------------------------------------------------- */
int oocdag_llt_resid(const double *A, int N, int LDA, char aform, char diag, int iseed,
                     double *resid);
int oocdag_llt_resid_tiles(const double *dY, int N, int NB, char aform, char diag, int iseed,
                           double *resid);

/* ---------------------------------------------------------------------------
 * Solve residual max_k ||A x_k - b_k||_inf / (||A||_inf ||x_k||_inf eps N),
 * eps = 2^-53, the maximum over the right-hand sides.
 * Replaces: PDLASCHK(SYMM, DIAG, N, NRHS, X, ..., ANORM, RESID)
 *           LIN/pdlaschk.f:22-26,186,295 (ANORM = ||A||_inf of the symmetric
 *           matrix, formed here).
 * oocdag_solve_resid: host arrays; the lower triangle of A (column-major,
 * LDA) defines the symmetric matrix.  oocdag_solve_resid_tiles: device A in
 * packed lower tiles with full (symmetric) diagonal tiles, as the tile
 * generators write them; device X and B (ldx, ldb >= N).  Every reduction
 * runs in a fixed order: the same inputs give the same bits.
 * ------------------------------------------------------------------------- */
int oocdag_solve_resid(const double *A, int N, int LDA, const double *X, int nrhs, int LDX,
                       const double *B, int LDB, double *resid);
int oocdag_solve_resid_tiles(const double *dA, int N, int NB, const double *dX, int nrhs, int ldx,
                             const double *dB, int ldb, double *resid);

/* ---------------------------------------------------------------------------
 * Mixed precision + iterative refinement (no reference counterpart: the
 * reference factors in FP64 only).  Factor a copy of the packed FP64 tiles
 * dA with the trailing updates in FP64 (precision 0), FP32 (1: FP32
 * accuracy, the super-column updates on BF16 MFMAs from three BF16 pieces
 * per element) or BF16 with per-tile power-of-two scaling (2), FP32
 * accumulation; diagonal POTRF, TRSM,
 * residual r = b - A x and the update x += d stay in FP64.  Refine until the
 * normwise backward error max_k ||b_k - A x_k|| / (||A|| ||x_k|| + ||b_k||)
 * (infinity norms) is <= tol; a low-precision factor that is not positive
 * definite, or refinement that stalls (error not halving) or exceeds
 * max_iter, falls back one level (BF16 -> FP32 -> FP64; a factor that meets
 * a bad pivot skips its remaining low-precision updates).  dA must hold the full
 * diagonal tiles (as oocdag_gram_tiles / oocdag_pdmatgen_tiles write them).
 * NB % 128 == 0.  dB, dX: device, column-major N x nrhs.
 * ------------------------------------------------------------------------- */
typedef struct {
    int precision_used;    /* update precision of the accepted factor */
    int fell_back;         /* 1: the low-precision attempt was abandoned */
    int iterations;        /* refinement steps over all attempts */
    int info;              /* > 0: A itself is not positive definite */
    double backward_error; /* of the returned solution */
    double factor_seconds, solve_seconds, total_seconds;
    double solve_resid;    /* PDLASCHK residual of the returned solution
                              (LIN/pdlaschk.f:186,295; oocdag_solve_resid) */
} oocdag_mpir_stats;

int oocdag_mpir_solve_tiles(const double *dA, int N, int NB, const double *dB, int nrhs, int ldb,
                            double *dX, int ldx, int precision, double tol, int max_iter,
                            oocdag_mpir_stats *stats);

/* Gram test matrices of the mixed-precision workloads: A = (1/N) G G^T + mu I
 * with G an N x N matrix of N(0,1) variates from a counter-based generator
 * (splitmix64 + Box-Muller, keyed by seed), written as packed lower tiles
 * with full diagonal tiles.  Needs a temporary N x N buffer on the device. */
int oocdag_gram_tiles(double *dY, int N, int NB, double mu, unsigned long long seed, void *stream);
/* N(0,1) variates (same generator, stream_id selects an independent stream)
 * into a device block of rows x cols, leading dimension ld. */
int oocdag_normal_fill(double *dB, int rows, int cols, int ld, unsigned long long seed,
                       unsigned long long stream_id);

/* ---------------------------------------------------------------------------
 * Batched small SPD factor + solve (no reference counterpart).  nsys
 * independent systems A_s x_s = b_s, n_s in [1, 1024], nrhs_s in [1, 64], in
 * device arenas laid out by oocdag_batch_layout (A_s n x n and B_s/X_s
 * n x nrhs, column-major with leading dimension n rounded up to even).  One
 * persistent kernel (work queue over the systems, largest first; each
 * workgroup factors and solves one system with LDS-staged panels), after
 * block-column launches that factor the largest systems across the whole
 * GPU when that is faster (oocdag_batch_set_split), replayed through a
 * captured hipGraph `reps` times; its block size and register tiling are
 * picked by timing the compiled variants on a sample of the batch.  precision 0: FP64 factor; 1: FP32 factor + FP64 refinement to the
 * backward error tol.  Per system (optional arrays): berr, the normwise
 * backward error max_k ||r_k|| / (||A|| ||x_k|| + ||b_k||); solve_resid,
 * PDLASCHK's max_k ||r_k|| / (||A|| ||x_k|| eps n) (LIN/pdlaschk.f:186,295);
 * iters; status: 0 ok, k > 0 not positive definite at column k, -2
 * refinement did not reach tol (rerun it with precision 0).
 * ------------------------------------------------------------------------- */
typedef struct {
    double seconds;             /* one batch, averaged over reps */
    double systems_per_second;
    double max_backward_error;  /* over systems with status 0 */
    int failed;                 /* systems with status != 0 */
    int variant_T, variant_RB;  /* chosen block size and register tile */
    double tune_seconds;
    double max_solve_resid;     /* PDLASCHK residual, max over systems with status 0 */
    int split_min_n;            /* systems with n >= this (0: none) were factored by */
    int split_systems;          /* the block-column launches; how many */
} oocdag_batch_stats;

long long oocdag_batch_layout(int nsys, const int *n, const int *nrhs, long long *a_off,
                              long long *b_off);
int oocdag_batch_generate(int nsys, const int *n, const int *nrhs, const long long *a_off,
                          const long long *b_off, double *dA, double *dB, double mu,
                          unsigned long long seed);
int oocdag_batched_solve(int nsys, const int *n, const int *nrhs, const long long *a_off,
                         const long long *b_off, const double *dA, const double *dB, double *dX,
                         int precision, double tol, int max_iter, int reps, double *berr,
                         double *solve_resid, int *iters, int *status, oocdag_batch_stats *stats);
/* Fixes the kernel variant oocdag_batched_solve uses for a precision
 * (0 FP64, 1 FP32+IR) instead of autotuning; variant -1 restores the
 * autotuner.  Returns the number of variants, or OOCDAG_EARG.  Variants
 * (T, RB): right-looking with T x T diagonal blocks and RB x RB register
 * tiles, and the left-looking MFMA kernel (T = 64, RB = 0). */
int oocdag_batch_set_variant(int precision, int variant);
/* Fixes which systems oocdag_batched_solve factors by block-column launches
 * instead of inside one workgroup each: those with n >= min_n (0: none).
 * -1 restores the default: a call with reps > 0 times the whole batch with
 * no split and with each candidate threshold (769, 513, 385, 257) and keeps
 * the fastest; a call with reps <= 0 splits the systems of n >= 513 when
 * there are at most two per CU of them.  Split systems take one launch per
 * step of 64 columns for the update, the diagonal block and the rows below,
 * across all of them at once (so a few large systems use the whole GPU);
 * the left-looking kernel then solves them.  The factor is bit-identical
 * either way.  Only the left-looking variant splits.  Returns OOCDAG_OK or
 * OOCDAG_EARG. */
int oocdag_batch_set_split(int precision, int min_n);

/* Chooses how the mixed-precision factorization (oocdag_mpir_solve_tiles and
 * the server) runs its low-precision trailing updates, as bits: 1 = the
 * LDS-staged FP32 update kernel instead of the one that loads operands
 * straight into MFMA registers; 2 = each super-column's bulk update runs
 * after the next super-column's FP64 factorization instead of beside it on a
 * second stream; 4 = the FP32 route's super-column updates on FP32 MFMA
 * instead of from three BF16 pieces per element.  Variants 0-3 give
 * bit-identical results; bit 4 changes the FP32 factor's rounding (same
 * accuracy).  A test and tuning hook with no reference counterpart; 0
 * restores the default; returns the number of variants (8) or OOCDAG_EARG. */
int oocdag_mp_set_update_variant(int variant);

/* ---------------------------------------------------------------------------
 * Continuous-batching solve server (no reference counterpart).  Requests are
 * served by `slots` worker threads with their own HIP streams, admitted
 * against an HBM byte budget; small systems (n <= 1024) that are
 * pending together run in one launch of the batched kernel; larger ones are
 * factored in mixed precision (from n = 8192 on; FP64 below, where the
 * refinement costs more than it saves) + FP64 refinement, their factor cached by
 * matrix_id (least recently used evicted under the budget); a request whose
 * working set exceeds the whole budget takes the out-of-core FP64 route.
 * ------------------------------------------------------------------------- */
typedef struct {
    long long matrix_id;         /* same id = same matrix (its factor may be cached) */
    int n, nrhs;                 /* nrhs in [1, 64] */
    double mu;                   /* A = (1/n) G G^T + mu I, G from `seed` */
    unsigned long long seed;
    unsigned long long rhs_seed; /* B ~ N(0,1) from this seed */
    int precision;               /* factor precision: 0 FP64, 1 FP32, 2 BF16 (+ refinement) */
    double tol;                  /* target normwise backward error */
} oocdag_request;

typedef struct {
    int status;                  /* 0 ok, > 0 not positive definite, < 0 OOCDAG_E*
                                    (OOCDAG_EBUDGET: needs more than the whole budget) */
    int route;                   /* 0 batched, 1 HBM mixed precision, 2 out-of-core */
    int cache_hit;
    int batch_size;              /* requests in the same launch (route 0) */
    int precision_used, fell_back, iterations;
    double backward_error;
    double queue_seconds, service_seconds, latency_seconds;
    double x_checksum;           /* sum of the solution's entries */
    double input_seconds;        /* part of service_seconds spent materializing the
                                    request's matrix (A = G G^T / n + mu I, n^3 flops)
                                    and right-hand sides; 0 on a factor-cache hit */
    double solve_resid;          /* PDLASCHK residual of the solution,
                                    max_k ||A x_k - b_k|| / (||A|| ||x_k|| eps n)
                                    (LIN/pdlaschk.f:186,295) */
} oocdag_result;

typedef struct oocdag_server oocdag_server;
oocdag_server *oocdag_server_create(size_t hbm_budget_bytes, int slots);
long long oocdag_server_submit(oocdag_server *srv, const oocdag_request *rq); /* ticket > 0 */
/* Blocks until the ticket is done; x (optional) receives the n x nrhs
 * solution, column-major, up to x_len doubles. */
int oocdag_server_wait(oocdag_server *srv, long long ticket, oocdag_result *res, double *x,
                       size_t x_len);
int oocdag_server_stats(oocdag_server *srv, size_t *used_bytes, int *cached, int *pending,
                        int *running);
/* The HBM bytes admission charges for a request: on a factor-cache hit, on a
 * miss, the resident factor it leaves in the cache, and the least the
 * out-of-core route needs (a miss above the budget runs out of core inside
 * max(that, budget / 2), capped at the budget; a request that cannot fit
 * even then completes with OOCDAG_EBUDGET). */
int oocdag_server_estimate(const oocdag_request *rq, size_t *hit_bytes, size_t *miss_bytes,
                           size_t *cached_bytes, size_t *ooc_min_bytes);
int oocdag_server_destroy(oocdag_server *srv);

/* ---------------------------------------------------------------------------
 * Tile kernels (device pointers, column-major, explicit leading dimensions).
 * lda/ldb/ldc must be even and pointers 16-byte aligned.
 * Replaces: cublasDgemm_v2(N,T)   oocquark.cpp:222  (CORE_incore_dgemm)
 *           cublasDsyrk_v2(L,N)   oocquark.cpp:176  (CORE_incore_dsyrk)
 *           cublasDtrsm_v2(R,L,T,N) oocquark.cpp:130 (CORE_incore_dtrsm)
 *           dpotrf_('L')          oocquark.cpp:87   (CORE_incore_dpotrf)
 * ------------------------------------------------------------------------- */
int oocdag_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *C, int ldc, void *stream);
int oocdag_dsyrk_ln(int n, int k, double alpha, const double *A, int lda, double beta,
                    double *C, int ldc, void *stream);
/* B(m x n) <- B * L^-T, L n x n lower.  Any n; n % 64 == 0 with even
 * leading dimensions runs in place, other shapes through padded copies. */
int oocdag_dtrsm_rltn(int m, int n, const double *L, int ldl, double *B, int ldb, void *stream);
/* In-place lower Cholesky of an n x n matrix, any n (n % 64 == 0 with an
 * even lda in place, otherwise through a padded copy); strict upper
 * triangle untouched.  Synchronous; *info LAPACK-style. */
int oocdag_dpotrf_l(int n, double *A, int lda, int *info, void *stream);
/* oocdag_dpotrf_l that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (the 64 x 64 diagonal-block inverses and the info word, 8 MiB) is
 * a fixed device array of the library, one per GPU, held until the process
 * ends; the caller passes only A.  Limits: n % 64 == 0, n <= 16384, even
 * lda (other shapes: OOCDAG_EARG; oocdag_dpotrf_l takes them through a
 * padded copy, which would need scratch).  Every call on one GPU shares that
 * array: calls that overlap on a GPU race.  Same results as
 * oocdag_dpotrf_l, bit for bit. */
int oocdag_dpotrf_l_noalloc(int n, double *A, int lda, int *info, void *stream);

/* ---------------------------------------------------------------------------
 * Device memory helpers for bindings without a GPU array library.
 * ------------------------------------------------------------------------- */
void *oocdag_dev_alloc(size_t bytes);
int oocdag_dev_free(void *p);
int oocdag_memcpy_h2d(void *dst, const void *src, size_t bytes);
int oocdag_memcpy_d2h(void *dst, const void *src, size_t bytes);
int oocdag_memcpy_d2d(void *dst, const void *src, size_t bytes);
int oocdag_memset_d(void *dst, int value, size_t bytes);
int oocdag_device_sync(void);
/* Number of visible HIP devices (0 without a GPU); does not fail. */
int oocdag_device_count(void);
/* Last error message of this thread ("" if none). */
const char *oocdag_last_error(void);

/* Timing of the last oocdag_cholesky / oocdag_potrf_tiles call on this
 * thread: seconds in the factorization, bytes moved host->device and
 * device->host, and seconds the copy streams were busy (0 if no copies). */
int oocdag_last_stats(double *seconds, double *h2d_bytes, double *d2h_bytes,
                      double *h2d_seconds, double *d2h_seconds);

/* Instrumentation of the dominant kernel, the panel-update GEMM (replaces
 * the per-task timestamps of oocquark.cpp:2-8 and the CUDA_PROFILE switch of
 * run_gpu.sh:27): when enabled, every update launch is bracketed by HIP
 * events on the stream it runs on.  profile_read returns the launches, the
 * summed kernel seconds and the algorithmic flops (2*K per useful lower
 * element) since enable.  Process-wide, not thread-safe. */
int oocdag_profile_enable(int on);
int oocdag_profile_read(long long *launches, double *seconds, double *flops);
/* Algorithmic HBM bytes of the same launches since enable: every C element
 * of the lower trapezoid read and written once (16 B) plus the K-deep
 * operand strips read once (8 B per element); the roofline's byte count. */
int oocdag_profile_bytes(double *bytes);

/* The out-of-core entry points stage host data through page-locked buffers
 * that they keep in a process-wide pool for the next call (at most 10 idle
 * buffers); this frees the idle ones and returns how many it freed. */
int oocdag_release_staging(void);

/* Progress lines on stderr every this many factored tile columns (0 = off),
 * the successor of the reference's per-panel printf (oocdag.cpp:67). */
int oocdag_set_progress(int every_columns);

#ifdef __cplusplus
}
#endif
#endif
