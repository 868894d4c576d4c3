/*
 * oracle.c -- CPU restatement of the reference out-of-core Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the thing
 * measured or shipped.  Only tests/, the smoke check and the benchmark's
 * CPU-baseline leg load it.
 *
 * Parity pinning of each part (what a plausible mistake would fail):
 *   ora_lcg_skip / ora_pdmatgen -- bit-exact against the reference's own
 *       Fortran PDMATGEN built from LIN/pdmatgen.f + LIN/pmatgeninc.f
 *       (oracle/Makefile -> oracle/_ref/) and against golden fixtures dumped
 *       from that build (tests/golden/).
 *   ora_find_Yn / ora_panel_plan -- brute-force feasibility/maximality of
 *       every width plus a committed table of schedules.
 *   ora_dpotrf_l, ora_dtrsm_rltn, ora_dsyrk_ln, ora_dgemm_nt -- against
 *       numpy/LAPACK (numpy.linalg.cholesky, solve_triangular, matmul) and
 *       against MKL through the run-time backend.
 *   ora_cholesky -- against the closed-form factor of the AFORM='F' matrix,
 *       against numpy.linalg.cholesky, and against the same driver run on MKL
 *       tile kernels (the reference's host dpotrf library).
 *   ora_llt_resid -- against a numpy restatement on small n, and by planting
 *       a perturbation of known size into L and checking the predicted change.
 *   ora_dpotrs_l / ora_solve_resid -- against numpy.linalg.solve.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------ */
/* LCG: LIN/pmatgeninc.f LMUL (:40-78) and LADD (:5-34) compute        */
/* (k*j) mod 2^31 and (k+j) mod 2^31 on 16/15-bit halves.  Since 2^31  */
/* divides 2^32, uint32 wrap-around followed by a 31-bit mask gives    */
/* the same residues.                                                  */
/* ------------------------------------------------------------------ */
#define LCG_MULT 1103515245u /* MULT0 + MULT1*2^16 = 20077 + 16838*65536 */
#define LCG_IADD 12345u      /* IADD0 + IADD1*2^16 = 12345 + 0            */
#define MASK31 0x7fffffffu

typedef struct {
    uint32_t a, c; /* x -> (a*x + c) mod 2^31 */
} lcg_map;

static lcg_map lcg_then(lcg_map f, lcg_map g) /* g(f(x)) */
{
    lcg_map r;
    r.a = (g.a * f.a) & MASK31;
    r.c = (g.a * f.c + g.c) & MASK31;
    return r;
}

static lcg_map lcg_pow(uint64_t n)
{
    /* XJUMPM (LIN/pmatgeninc.f:84-135) forms the n-step map by repeated
     * composition; binary powering gives the same map. */
    lcg_map r = {1u, 0u};
    lcg_map b = {LCG_MULT, LCG_IADD};
    while (n) {
        if (n & 1u) r = lcg_then(r, b);
        b = lcg_then(b, b);
        n >>= 1;
    }
    return r;
}

uint32_t ora_lcg_skip(uint32_t x, uint64_t n)
{
    lcg_map m = lcg_pow(n);
    return (m.a * (x & MASK31) + m.c) & MASK31;
}

/* PDRAND (LIN/pmatgeninc.f:285-313) returns x/2^31 and then advances;
 * the element is ONE - TWO*PDRAND(0) (LIN/pdmatgen.f:261).  Both steps are
 * exact in binary64. */
static inline double lcg_value(uint32_t x)
{
    return 1.0 - 2.0 * ((double)x / 2147483648.0);
}

/* ------------------------------------------------------------------ */
/* PDMATGEN on a 1x1 grid with IROFF=ICOFF=0, IRNUM=M, ICNUM=N.        */
/* SETRAN starts the stream at x1 = one step past ISEED                */
/* (LIN/pdmatgen.f:231,240).  Stream position of element (i,j):        */
/*  - symmetric lower pass (:248-303): i >= j  ->  j*M + i; the strict */
/*    upper triangle of diagonal blocks copies it (:273-276) and the   */
/*    upper pass (:340-379) reads position i*N + j, which for M == N   */
/*    is the same number, so A is exactly symmetric;                   */
/*  - transposed (:383-440): i*N + j;  general (:441-505): j*M + i.    */
/* ------------------------------------------------------------------ */
int ora_pdmatgen(char aform, char diag, int M, int N, int MB, int NB,
                 double *A, int LDA, int iseed)
{
    char af = (char)(aform & ~0x20); /* upper-case, like LSAME */
    char dg = (char)(diag & ~0x20);
    int symm = (af == 'S' || af == 'H');
    int tran = (af == 'T' || af == 'C');
    int info = 0;
    if (dg != 'D' && dg != 'N')
        info = 3;
    else if (symm) {
        if (M != N)
            info = 5;
        else if (MB != NB)
            info = 7;
    } else if (M < 0)
        info = 4;
    else if (N < 0)
        info = 5;
    else if (MB < 1)
        info = 6;
    else if (NB < 1)
        info = 7;
    else if (LDA < 0)
        info = 9;
    if (info) return info; /* PXERBLA: nothing generated */
    if (M == 0 || N == 0) return 0;

    uint32_t x1 = ora_lcg_skip((uint32_t)iseed, 1);

    if (symm) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int j = 0; j < N; j++) {
            uint32_t x = ora_lcg_skip(x1, (uint64_t)j * (uint64_t)M + (uint64_t)j);
            double *col = A + (size_t)j * LDA;
            for (int i = j; i < M; i++) {
                col[i] = lcg_value(x);
                x = (LCG_MULT * x + LCG_IADD) & MASK31;
            }
        }
        /* mirror into the strict upper triangle */
#pragma omp parallel for schedule(dynamic, 16)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < j; i++) A[i + (size_t)j * LDA] = A[j + (size_t)i * LDA];
    } else if (tran) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < M; i++) {
            uint32_t x = ora_lcg_skip(x1, (uint64_t)i * (uint64_t)N);
            for (int j = 0; j < N; j++) {
                A[i + (size_t)j * LDA] = lcg_value(x);
                x = (LCG_MULT * x + LCG_IADD) & MASK31;
            }
        }
    } else {
#pragma omp parallel for schedule(dynamic, 16)
        for (int j = 0; j < N; j++) {
            uint32_t x = ora_lcg_skip(x1, (uint64_t)j * (uint64_t)M);
            double *col = A + (size_t)j * LDA;
            for (int i = 0; i < M; i++) {
                col[i] = lcg_value(x);
                x = (LCG_MULT * x + LCG_IADD) & MASK31;
            }
        }
    }

    if (dg == 'D') { /* LIN/pdmatgen.f:516-543 */
        if (MB != NB) return 0; /* the reference prints and returns */
        double maxmn = (double)MAX(M, N);
        int nd = MIN(M, N);
        for (int i = 0; i < nd; i++) {
            double *a = A + i + (size_t)i * LDA;
            *a = fabs(*a) + maxmn;
        }
    }

    if (af == 'F') { /* LIN/pdmatgen.f:546-626: REAL literal 0.1, integer 10 */
        const double c = (double)0.1f;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < M; i++) A[i + (size_t)j * LDA] = (i == j) ? 10.0 : c;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* find_Yn (oocdag.cpp:478-492), static made explicit.                 */
/* ------------------------------------------------------------------ */
int ora_find_Yn(int bb, int memBlock, int jb, int *state)
{
    int Yn = *state;
    int Ym = bb - jb;
    if (Ym * (Ym + 1) <= (memBlock - Ym) * 2) {
        Yn = Ym;
    } else {
        while ((memBlock - Ym) * 2 >= (2 * Ym - Yn) * (Yn + 1)) {
            if (Yn == Ym) break;
            Yn++;
        }
    }
    *state = Yn;
    return Yn;
}

static int budget_tiles(size_t memsize, int NB)
{
    /* int memBlock = memsize/sizeof(double)/NB/NB;  (oocdag.cpp:31) */
    return (int)(memsize / sizeof(double) / (size_t)NB / (size_t)NB);
}

static int budget_ok(int N, int NB, size_t memsize)
{
    /* w = memsize/sizeof(double)/BB/NB/NB - 1; assert(w > 0) (oocdag.cpp:248-250) */
    int BB = (N + NB - 1) / NB;
    long w = (long)(memsize / sizeof(double) / (size_t)BB / (size_t)NB / (size_t)NB) - 1;
    return w > 0;
}

int ora_panel_plan(int N, int NB, size_t memsize, int *widths, int maxp)
{
    if (N <= 0 || NB <= 0 || !budget_ok(N, NB, memsize)) return -1;
    int bb = (N + NB - 1) / NB;
    int memBlock = budget_tiles(memsize, NB);
    int state = 0, np = 0;
    for (int JB = 0, jb = 0, Yn = 0; JB < N; jb += Yn) {
        Yn = ora_find_Yn(bb, memBlock, jb, &state);
        int YN = MIN((jb + Yn) * NB, N) - jb * NB;
        if (np < maxp) widths[np] = Yn;
        np++;
        JB += YN;
    }
    return np;
}

/* ------------------------------------------------------------------ */
/* Tile kernels: the oracle's own loops.                               */
/* ------------------------------------------------------------------ */
int ora_dpotrf_l(int n, double *A, int lda)
{
    /* Left-looking column Cholesky (the unblocked LAPACK dpotf2 'L'). */
    for (int j = 0; j < n; j++) {
        double *cj = A + (size_t)j * lda;
        for (int k = 0; k < j; k++) {
            const double *ck = A + (size_t)k * lda;
            double t = ck[j];
            for (int i = j; i < n; i++) cj[i] -= t * ck[i];
        }
        double ajj = cj[j];
        if (!(ajj > 0.0)) return j + 1; /* also catches NaN */
        ajj = sqrt(ajj);
        cj[j] = ajj;
        for (int i = j + 1; i < n; i++) cj[i] /= ajj;
    }
    return 0;
}

void ora_dtrsm_rltn(int m, int n, const double *L, int ldl, double *B, int ldb)
{
    /* X * L^T = B, column by column. */
    for (int c = 0; c < n; c++) {
        double *bc = B + (size_t)c * ldb;
        for (int d = 0; d < c; d++) {
            double t = L[c + (size_t)d * ldl];
            const double *bd = B + (size_t)d * ldb;
            for (int i = 0; i < m; i++) bc[i] -= t * bd[i];
        }
        double dcc = L[c + (size_t)c * ldl];
        for (int i = 0; i < m; i++) bc[i] /= dcc;
    }
}

void ora_dsyrk_ln(int n, int k, double alpha, const double *X, int ldx,
                  double beta, double *C, int ldc)
{
    for (int j = 0; j < n; j++) {
        double *cj = C + (size_t)j * ldc;
        if (beta == 0.0)
            for (int i = j; i < n; i++) cj[i] = 0.0;
        else if (beta != 1.0)
            for (int i = j; i < n; i++) cj[i] *= beta;
        for (int p = 0; p < k; p++) {
            const double *xp = X + (size_t)p * ldx;
            double t = alpha * xp[j];
            for (int i = j; i < n; i++) cj[i] += t * xp[i];
        }
    }
}

void ora_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc)
{
    int j = 0;
    for (; j < n; j++) {
        double *cj = C + (size_t)j * ldc;
        if (beta == 0.0)
            for (int i = 0; i < m; i++) cj[i] = 0.0;
        else if (beta != 1.0)
            for (int i = 0; i < m; i++) cj[i] *= beta;
    }
    /* four columns of C per sweep so every column of A is reused 4x */
    for (j = 0; j + 4 <= n; j += 4) {
        double *c0 = C + (size_t)j * ldc, *c1 = c0 + ldc, *c2 = c1 + ldc, *c3 = c2 + ldc;
        for (int p = 0; p < k; p++) {
            const double *ap = A + (size_t)p * lda;
            const double *bp = B + (size_t)p * ldb + j;
            double t0 = alpha * bp[0], t1 = alpha * bp[1], t2 = alpha * bp[2], t3 = alpha * bp[3];
            for (int i = 0; i < m; i++) {
                double a = ap[i];
                c0[i] += t0 * a;
                c1[i] += t1 * a;
                c2[i] += t2 * a;
                c3[i] += t3 * a;
            }
        }
    }
    for (; j < n; j++) {
        double *cj = C + (size_t)j * ldc;
        for (int p = 0; p < k; p++) {
            const double *ap = A + (size_t)p * lda;
            double t = alpha * B[j + (size_t)p * ldb];
            for (int i = 0; i < m; i++) cj[i] += t * ap[i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Optional MKL backend, resolved at run time.                         */
/* ------------------------------------------------------------------ */
typedef void (*f_dgemm)(const char *, const char *, const int *, const int *, const int *,
                        const double *, const double *, const int *, const double *, const int *,
                        const double *, double *, const int *);
typedef void (*f_dsyrk)(const char *, const char *, const int *, const int *, const double *,
                        const double *, const int *, const double *, double *, const int *);
typedef void (*f_dtrsm)(const char *, const char *, const char *, const char *, const int *,
                        const int *, const double *, const double *, const int *, double *,
                        const int *);
typedef void (*f_dpotrf)(const char *, const int *, double *, const int *, int *);
typedef int (*f_setlayer)(int);

static struct {
    int backend;
    int loaded;
    f_dgemm dgemm;
    f_dsyrk dsyrk;
    f_dtrsm dtrsm;
    f_dpotrf dpotrf;
} g_blas;

static int g_threads = 0;

static int mkl_load(void)
{
    if (g_blas.loaded) return g_blas.loaded > 0;
    const char *names[] = {"libmkl_rt.so.2", "libmkl_rt.so.1", "libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", NULL};
    void *h = NULL;
    for (int i = 0; names[i] && !h; i++) h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        g_blas.loaded = -1;
        return 0;
    }
    f_setlayer setlayer = (f_setlayer)dlsym(h, "MKL_Set_Threading_Layer");
    if (setlayer) setlayer(1); /* MKL_THREADING_SEQUENTIAL: one BLAS thread per tile task */
    g_blas.dgemm = (f_dgemm)dlsym(h, "dgemm_");
    g_blas.dsyrk = (f_dsyrk)dlsym(h, "dsyrk_");
    g_blas.dtrsm = (f_dtrsm)dlsym(h, "dtrsm_");
    g_blas.dpotrf = (f_dpotrf)dlsym(h, "dpotrf_");
    g_blas.loaded = (g_blas.dgemm && g_blas.dsyrk && g_blas.dtrsm && g_blas.dpotrf) ? 1 : -1;
    return g_blas.loaded > 0;
}

int ora_set_backend(int backend)
{
    g_blas.backend = (backend == 1 && mkl_load()) ? 1 : 0;
    return g_blas.backend;
}

int ora_get_backend(void) { return g_blas.backend; }

void ora_set_threads(int nthreads) { g_threads = nthreads; }

int ora_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

static int k_potrf(int n, double *A, int lda)
{
    if (g_blas.backend == 1) {
        int info = 0;
        g_blas.dpotrf("L", &n, A, &lda, &info);
        return info;
    }
    return ora_dpotrf_l(n, A, lda);
}

static void k_trsm(int m, int n, const double *L, int ldl, double *B, int ldb)
{
    if (g_blas.backend == 1) {
        double one = 1.0;
        g_blas.dtrsm("R", "L", "T", "N", &m, &n, &one, L, &ldl, B, &ldb);
        return;
    }
    ora_dtrsm_rltn(m, n, L, ldl, B, ldb);
}

static void k_syrk(int n, int k, const double *X, int ldx, double *C, int ldc)
{
    double alpha = -1.0, beta = 1.0;
    if (g_blas.backend == 1) {
        g_blas.dsyrk("L", "N", &n, &k, &alpha, X, &ldx, &beta, C, &ldc);
        return;
    }
    ora_dsyrk_ln(n, k, alpha, X, ldx, beta, C, ldc);
}

static void k_gemm(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                   double *C, int ldc)
{
    double alpha = -1.0, beta = 1.0;
    if (g_blas.backend == 1) {
        g_blas.dgemm("N", "T", &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
        return;
    }
    ora_dgemm_nt(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ------------------------------------------------------------------ */
/* The driver.  H2D/D2H = cublasSetMatrix/GetMatrix (oocquark.cpp:27,  */
/* 57) become 2-D memcpy between ld=LDA and ld=NB.                     */
/* ------------------------------------------------------------------ */
static void copy2d(int m, int n, const double *src, int lds, double *dst, int ldd)
{
    for (int j = 0; j < n; j++)
        memcpy(dst + (size_t)j * ldd, src + (size_t)j * lds, (size_t)m * sizeof(double));
}

/* Packed trapezoid index of the device panel (oocquark.cpp:256,326,361). */
static inline size_t ypack(int ib, int jb, int Ym) { return (size_t)(ib + jb * Ym - (jb * (jb + 1)) / 2); }

/* A2Y (oocquark.cpp:253-275) */
static void a2y(const double *A, double *Y, int LDA, int NB, int M, int N)
{
    int Ym = (M + NB - 1) / NB;
    size_t nb2 = (size_t)NB * NB;
#pragma omp parallel for schedule(dynamic) collapse(2) num_threads(ora_get_threads())
    for (int jb = 0; jb < (N + NB - 1) / NB; jb++)
        for (int ib = 0; ib < Ym; ib++) {
            if (ib < jb) continue;
            int JB = jb * NB, IB = ib * NB;
            int LN = MIN(JB + NB, N) - JB, LM = MIN(IB + NB, M) - IB;
            copy2d(LM, LN, A + (size_t)JB * LDA + IB, LDA, Y + ypack(ib, jb, Ym) * nb2, NB);
        }
}

/* A2X (oocquark.cpp:301-319) */
static void a2x(const double *A, int LDA, double *X, int NB, int K)
{
    size_t nb2 = (size_t)NB * NB;
    int nt = (K + NB - 1) / NB;
#pragma omp parallel for schedule(static) num_threads(ora_get_threads())
    for (int ib = 0; ib < nt; ib++) {
        int IB = ib * NB, LM = MIN(IB + NB, K) - IB;
        copy2d(LM, NB, A + IB, LDA, X + (size_t)ib * nb2, NB);
    }
}

/* ooc_syrk (oocquark.cpp:321-356): Y(j,j) -= X(j)X(j)^T, Y(i,j) -= X(i)X(j)^T */
static void ooc_syrk(const double *X, double *Y, int H, int K, int NB)
{
    int Ym = (H + NB - 1) / NB;
    int Kt = (K + NB - 1) / NB;
    size_t nb2 = (size_t)NB * NB;
#pragma omp parallel for schedule(dynamic) collapse(2) num_threads(ora_get_threads())
    for (int jb = 0; jb < Kt; jb++)
        for (int ib = 0; ib < Ym; ib++) {
            if (ib < jb) continue;
            int JB = jb * NB, IB = ib * NB;
            int LN = MIN(JB + NB, K) - JB;
            if (ib == jb) {
                k_syrk(LN, NB, X + (size_t)ib * nb2, NB, Y + ypack(ib, jb, Ym) * nb2, NB);
            } else {
                int LM = MIN(IB + NB, H) - IB;
                k_gemm(LM, LN, NB, X + (size_t)ib * nb2, NB, X + (size_t)jb * nb2, NB,
                       Y + ypack(ib, jb, Ym) * nb2, NB);
            }
        }
}

/* ooc_incore (oocquark.cpp:358-421), USE_HOST_DPOTRF flavour (Makefile:8). */
static int ooc_incore(double *A, double *Y, int LDA, int NB, int M, int N, int col0)
{
    int Ym = (M + NB - 1) / NB;
    size_t nb2 = (size_t)NB * NB;
    for (int JB = 0, jb = 0; JB < N; JB += NB, jb++) {
        int LN = MIN(JB + NB, N) - JB;
        double *Ajj = A + (size_t)JB * LDA + JB;
        double *Yjj = Y + ypack(jb, jb, Ym) * nb2;
        copy2d(LN, LN, Yjj, NB, Ajj, LDA);   /* D2H, :394 */
        int info = k_potrf(LN, Ajj, LDA);    /* host dpotrf_, :395 */
        if (info) return col0 + JB + info;   /* the reference only prints it (:95) */
        copy2d(LN, LN, Ajj, LDA, Yjj, NB);   /* H2D, :396 */
        int nrow = Ym - jb - 1;
#pragma omp parallel for schedule(dynamic) num_threads(ora_get_threads())
        for (int r = 0; r < nrow; r++) {     /* :405-410 */
            int ib = jb + 1 + r;
            int IB = ib * NB, LM = MIN(IB + NB, M) - IB;
            double *Yij = Y + ypack(ib, jb, Ym) * nb2;
            k_trsm(LM, LN, Yjj, NB, Yij, NB);
            copy2d(LM, LN, Yij, NB, A + (size_t)JB * LDA + IB, LDA); /* D2H */
        }
        if (JB + NB < N) /* :414-418 */
            ooc_syrk(Y + ypack(jb + 1, jb, Ym) * nb2, Y + ypack(jb + 1, jb + 1, Ym) * nb2,
                     M - JB - NB, N - JB - NB, NB);
    }
    return 0;
}

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

double ora_cholesky(double *A, int N, int NB, int LDA, size_t memsize, int *info)
{
    *info = 0;
    if (N < 0 || NB < 1 || LDA < MAX(1, N)) {
        *info = -1;
        return 0.0;
    }
    if (N == 0) return 0.0;
    if (!budget_ok(N, NB, memsize)) {
        *info = -2;
        return 0.0;
    }
    int bb = (N + NB - 1) / NB;
    int memBlock = budget_tiles(memsize, NB);
    size_t nb2 = (size_t)NB * NB;
    double *Y = (double *)malloc((size_t)memBlock * nb2 * sizeof(double));
    if (!Y) {
        *info = -3;
        return 0.0;
    }
    double t0 = wall();
    int state = 0;
    for (int JB = 0, jb = 0, Yn = 0, YN = 0; JB < N; JB += YN, jb += Yn) {
        int Ym = bb - jb;
        Yn = ora_find_Yn(bb, memBlock, jb, &state);
        int YM = N - JB;
        YN = MIN((jb + Yn) * NB, N) - jb * NB;
        double *X = Y + (size_t)(memBlock - Ym) * nb2;
        double *Ajb = A + (size_t)JB * LDA + JB;
        a2y(Ajb, Y, LDA, NB, YM, YN);
        for (int jjb = 0; jjb < jb; jjb++) {
            a2x(A + (size_t)jjb * NB * LDA + JB, LDA, X, NB, YM);
            ooc_syrk(X, Y, YM, YN, NB);
        }
        int bad = ooc_incore(Ajb, Y, LDA, NB, YM, YN, JB);
        if (bad) {
            *info = bad;
            break;
        }
    }
    double t = wall() - t0;
    free(Y);
    return t;
}

/* ------------------------------------------------------------------ */
/* Residuals                                                           */
/* ------------------------------------------------------------------ */
double ora_norm_inf(int N, const double *A, int lda)
{
    double best = 0.0;
    double *rs = (double *)calloc((size_t)N, sizeof(double));
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) rs[i] += fabs(A[i + (size_t)j * lda]);
    for (int i = 0; i < N; i++) best = MAX(best, rs[i]);
    free(rs);
    return best;
}

double ora_llt_resid_explicit(const double *Af, int N, int LDA, const double *A0, int lda0)
{
    const double eps = ldexp(1.0, -53); /* DLAMCH('E') */
    int nt = ora_get_threads();
    double *rows = (double *)calloc((size_t)N * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *rs = rows + (size_t)tid * N;
        double *acc = (double *)malloc((size_t)N * sizeof(double));
#pragma omp for schedule(dynamic, 4)
        for (int j = 0; j < N; j++) {
            /* (L L^T)(i,j) for i >= j, the lower triangle PDPOTRRV rebuilds */
            for (int i = j; i < N; i++) acc[i] = 0.0;
            for (int k = 0; k <= j; k++) {
                const double *lk = Af + (size_t)k * LDA;
                double t = lk[j];
                for (int i = j; i < N; i++) acc[i] += t * lk[i];
            }
            const double *a0 = A0 + (size_t)j * lda0;
            const double *af = Af + (size_t)j * LDA;
            for (int i = j; i < N; i++) rs[i] += fabs(acc[i] - a0[i]);
            /* strict upper triangle: PDPOTRRV restores it from A, so any
             * change the factorization made to it shows up here */
            for (int i = 0; i < j; i++) rs[i] += fabs(af[i] - a0[i]);
        }
        free(acc);
    }
    double worst = 0.0;
    for (int i = 0; i < N; i++) {
        double s = 0.0;
        for (int t = 0; t < nt; t++) s += rows[(size_t)t * N + i];
        worst = MAX(worst, s);
    }
    free(rows);
    double anorm = ora_norm_inf(N, A0, lda0);
    return worst / ((double)N * eps * anorm);
}

double ora_llt_resid(const double *Afact, int N, int LDA, char aform, char diag, int iseed,
                     double *anorm_out)
{
    double *A0 = (double *)malloc((size_t)N * N * sizeof(double));
    if (!A0) return -1.0;
    ora_pdmatgen(aform, diag, N, N, 1, 1, A0, N, iseed);
    if (anorm_out) *anorm_out = ora_norm_inf(N, A0, N);
    double r = ora_llt_resid_explicit(Afact, N, LDA, A0, N);
    free(A0);
    return r;
}

void ora_dpotrs_l(int N, int nrhs, const double *L, int ldl, double *B, int ldb)
{
#pragma omp parallel for schedule(dynamic) num_threads(ora_get_threads())
    for (int r = 0; r < nrhs; r++) {
        double *b = B + (size_t)r * ldb;
        /* L y = b (column-oriented forward substitution) */
        for (int j = 0; j < N; j++) {
            double yj = b[j] / L[j + (size_t)j * ldl];
            b[j] = yj;
            const double *lj = L + (size_t)j * ldl;
            for (int i = j + 1; i < N; i++) b[i] -= yj * lj[i];
        }
        /* L^T x = y (row-oriented backward substitution) */
        for (int j = N - 1; j >= 0; j--) {
            const double *lj = L + (size_t)j * ldl;
            double s = b[j];
            for (int i = j + 1; i < N; i++) s -= lj[i] * b[i];
            b[j] = s / lj[j];
        }
    }
}

double ora_solve_resid(int N, int nrhs, const double *A, int lda, const double *X, int ldx,
                       const double *B, int ldb)
{
    const double eps = ldexp(1.0, -53);
    double anorm = ora_norm_inf(N, A, lda);
    double worst = 0.0;
    double *r = (double *)malloc((size_t)N * sizeof(double));
    for (int c = 0; c < nrhs; c++) {
        const double *x = X + (size_t)c * ldx;
        const double *b = B + (size_t)c * ldb;
        for (int i = 0; i < N; i++) r[i] = b[i];
        for (int j = 0; j < N; j++) {
            double xj = x[j];
            const double *aj = A + (size_t)j * lda;
            for (int i = 0; i < N; i++) r[i] -= aj[i] * xj;
        }
        double rn = 0.0, xn = 0.0;
        for (int i = 0; i < N; i++) {
            rn = MAX(rn, fabs(r[i]));
            xn = MAX(xn, fabs(x[i]));
        }
        double q = rn / (xn * anorm * eps * (double)N);
        worst = MAX(worst, q);
    }
    free(r);
    return worst;
}
