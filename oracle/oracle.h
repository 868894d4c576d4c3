/*
 * oracle.h -- CPU restatement of the reference out-of-core Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (oocdag_amd/)
 * links, loads or calls this code; only tests/, the smoke check and the
 * benchmark's CPU-baseline leg use it, and only as the checker.
 *
 * Each function names the reference file:line it restates.  Parity status
 * of every function is recorded in the header comment of oracle.c.
 */
#ifndef OOCDAG_ORACLE_H
#define OOCDAG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generator: LIN/pdmatgen.f + LIN/pmatgeninc.f, 1x1 process grid ---- */

/* x_{k+1} = (a*x_k + c) mod 2^31 with a = 1103515245, c = 12345
 * (LIN/pdmatgen.f:117-119,210-213; LMUL/LADD in LIN/pmatgeninc.f:5-78). */
uint32_t ora_lcg_skip(uint32_t x, uint64_t n);

/* PDMATGEN(ICTXT, AFORM, DIAG, M, N, MB, NB, A, LDA, 0, 0, ISEED, 0, M, 0, N,
 * 0, 0, 1, 1) -- the call made by oocdag.cpp:298-302.  Supported AFORM:
 * 'S'/'H' (symmetric), 'T'/'C' (transposed), 'F' (hard-coded 10 / 0.1f form,
 * LIN/pdmatgen.f:546-626), anything else = general.  DIAG 'D' adds max(M,N)
 * to |a_ii| (LIN/pdmatgen.f:516-543).  Returns the PDMATGEN INFO code
 * (0 = ok). */
int ora_pdmatgen(char aform, char diag, int M, int N, int MB, int NB,
                 double *A, int LDA, int iseed);

/* ---- panel planner: oocdag.cpp:478-492 ---- */

/* find_Yn with its function-static 'Yn' made explicit: *state carries the
 * previous panel width (start at 0) exactly as the reference's static. */
int ora_find_Yn(int bb, int memBlock, int jb, int *state);

/* ---- tile kernels (column-major, explicit leading dimensions) ---- */
/* dpotrf_('L') as called at oocquark.cpp:87; returns LAPACK info. */
int  ora_dpotrf_l(int n, double *A, int lda);
/* cublasDtrsm(R,L,T,N, m, n, 1.0, L, ldl, B, ldb): B <- B * L^-T
 * (oocquark.cpp:130). */
void ora_dtrsm_rltn(int m, int n, const double *L, int ldl, double *B, int ldb);
/* cublasDsyrk(L,N, n, k, alpha, X, ldx, beta, C, ldc), lower only
 * (oocquark.cpp:176). */
void ora_dsyrk_ln(int n, int k, double alpha, const double *X, int ldx,
                  double beta, double *C, int ldc);
/* cublasDgemm(N,T, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
 * (oocquark.cpp:222). */
void ora_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc);

/* ---- the out-of-core left-looking driver: oocdag.cpp:18-112 ---- */

/* Tile-kernel backend: 0 = the oracle's own loops above, 1 = MKL resolved at
 * run time (dlopen of libmkl_rt), which is what the reference linked for its
 * host dpotrf (Makefile:17,45).  Returns the backend actually in effect. */
int ora_set_backend(int backend);
int ora_get_backend(void);

/* Worker threads for independent tile tasks (the QUARK pool's role,
 * oocdag.cpp:330).  0 = all available cores. */
void ora_set_threads(int nthreads);
int  ora_get_threads(void);

/* Cholesky(quark, A, N, NB, LDA, memsize) restated op-for-op: the "device"
 * is a host buffer of memBlock = memsize/8/NB^2 tiles; H2D/D2H are memcpy.
 * Lower triangle of A <- L, strict upper triangle untouched.  Returns the
 * wall-clock seconds of the factorization (timed like oocdag.cpp:51,90).
 * *info: 0 ok; k>0 first non-positive pivot (global 1-based column);
 * -1 bad args; -2 budget below 2 tile columns (oocdag.cpp:249-250). */
double ora_cholesky(double *A, int N, int NB, int LDA, size_t memsize,
                    int *info);

/* Panel schedule the driver above would run: widths[p] = Yn of panel p.
 * Returns the number of panels (<= maxp), or -1 if the budget is too small. */
int ora_panel_plan(int N, int NB, size_t memsize, int *widths, int maxp);

/* ---- residual oracles ---- */

/* PDPOTRRV('L') + PDLAFCHK('S', DIAG) semantics (LIN/pdpotrrv.f:240-314,
 * LIN/pdlafchk.f:181-225): rebuild L*L^T in the lower triangle, keep the
 * strict upper triangle of Afact, subtract the regenerated A and return
 * ||.||_inf / (N * eps * ANORM) with eps = 2^-53 and ANORM = ||A||_inf.
 * aform/diag/iseed regenerate A exactly as ora_pdmatgen. */
double ora_llt_resid(const double *Afact, int N, int LDA, char aform,
                     char diag, int iseed, double *anorm_out);

/* Same scaled residual against an explicit original matrix A0 (lda0). */
double ora_llt_resid_explicit(const double *Afact, int N, int LDA,
                              const double *A0, int lda0);

/* dpotrs('L') with the factor in the lower triangle of L (oocdag.cpp:439):
 * B (N x nrhs, ldb) <- A^-1 B. */
void ora_dpotrs_l(int N, int nrhs, const double *L, int ldl, double *B, int ldb);

/* PDLASCHK semantics (LIN/pdlaschk.f:180-300): max over right-hand sides of
 * ||b - A x||_inf / (||x||_inf * ANORM * eps * N), A symmetric given in
 * full (lda), eps = 2^-53. */
double ora_solve_resid(int N, int nrhs, const double *A, int lda,
                       const double *X, int ldx, const double *B, int ldb);

/* ||A||_inf of a full N x N matrix. */
double ora_norm_inf(int N, const double *A, int lda);

#ifdef __cplusplus
}
#endif
#endif
