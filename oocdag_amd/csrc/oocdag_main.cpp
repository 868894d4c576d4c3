// oocdag N NB memsize_MB [--verify] [--resid] [--nx K]
//
// Command-line driver with the reference's argument order and output lines
// (main, oocdag.cpp:223-475): generates the PDMATGEN('S','D', ISEED=100)
// matrix, factors it out of core with the given device budget and prints
// "N NB memsize(MB) quark_pthreads time Gflops" with the reference's rate
// formula (N/3 + 0.5) N^2 / t / 2^30 (oocdag.cpp:360-365).  The thread
// column reports the library's host threads (the reference printed its
// QUARK pool size there).  N <= 32 prints A and L like do_io
// (oocdag.cpp:253,307-310,375-378).  --verify runs the VERIFY block's
// identity check dpotrs(L, A) = I on the GPU path's factor (oocdag.cpp:379-464)
// through oocdag_potrs; --resid prints the LIN scaled residual.  The panel
// schedule is the reference's (find_Yn with one staged X tile column,
// oocdag.cpp:478-492) unless --nx 2 (two X columns, copies overlapped with
// the updates) or --nx 0 (the library's choice) is given.
#include "oocdag.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

static double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

static void matprint(const double *A, int N, int LDA, char AL) {
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) {
            double v = (AL == 'L' && j > i) ? 0.0 : A[(size_t)j * LDA + i];
            printf("%lf ", v);
        }
        printf("\n");
    }
}

int main(int argc, char **argv) {
    // four busy HIP streams plus the null stream: a pool of at least 5
    // hardware queues keeps them apart (HIP reads this at its first call)
    const char *q = getenv("GPU_MAX_HW_QUEUES");
    if (!q || atoi(q) < 5) setenv("GPU_MAX_HW_QUEUES", "8", 1);
    if (argc < 4) {
        printf("invalid arguments N NB memsize(M)\n");
        return 1;
    }
    const int N = atoi(argv[1]);
    const int NB = atoi(argv[2]);
    const size_t memsize = (size_t)atoi(argv[3]) * 1024 * 1024;
    bool verify = false, resid = false;
    int nx = 1;
    for (int i = 4; i < argc; ++i) {
        if (!strcmp(argv[i], "--verify")) verify = true;
        if (!strcmp(argv[i], "--resid")) resid = true;
        if (!strcmp(argv[i], "--nx") && i + 1 < argc) nx = atoi(argv[++i]);
    }
    if (N <= 0 || NB <= 0) {
        printf("invalid arguments N NB memsize(M)\n");
        return 1;
    }
    const int BB = (N + NB - 1) / NB;
    const long long w = (long long)(memsize / sizeof(double) / BB / NB / NB) - 1;
    if (w <= 0) {
        printf("memsize %zu MB is below two tile columns for N=%d NB=%d\n", memsize >> 20, N, NB);
        return 1;
    }
    const int LDA = N;
    const bool do_io = N <= 32;
    // left uninitialized: the generator writes every element
    std::unique_ptr<double[]> Abuf(new double[(size_t)LDA * N]);
    double *Ap = Abuf.get();

    printf("Initialize A ... ");
    fflush(stdout);
    double t = wall();
    int inf = oocdag_pdmatgen('S', 'D', N, N, NB, NB, Ap, LDA, 100);
    if (inf) {
        printf("pdmatgen failed: %d %s\n", inf, oocdag_last_error());
        return 1;
    }
    printf("time %lf\n", wall() - t);
    if (do_io) {
        printf("Original A=\n\n");
        matprint(Ap, N, LDA, 'A');
    }

    int widths[4096];
    int np = oocdag_panel_plan_nx(N, NB, memsize, nx, widths, 4096);
    int jb = 0;
    for (int p = 0; p < np && p < 4096; ++p) {
        int Ym = BB - jb, Yn = widths[p];
        int YM = N - jb * NB;
        int YN = std::min((jb + Yn) * NB, N) - jb * NB;
        printf("bb %d jb %d YM %d YN %d Ym %d Yn %d\n", BB, jb, YM, YN, Ym, Yn);
        jb += Yn;
    }

    printf("LLT start %lf\n", wall());
    int info = 0;
    double llttime = oocdag_cholesky_nx(Ap, N, NB, LDA, memsize, nx, &info);
    printf("LLT end %lf\n", wall());
    if (llttime < 0 || info < 0) {
        printf("oocdag_cholesky_nx failed: info %d %s\n", info, oocdag_last_error());
        return 1;
    }
    if (info > 0) printf("dpotrf info %d: leading minor not positive definite\n", info);
    printf("llt time %lf %lf\n", llttime, wall());
    printf("%lf %lf\n", Ap[(size_t)(N - 1) * LDA + N - 1], wall());

    double h2d = 0, d2h = 0, th = 0, td = 0;
    oocdag_last_stats(nullptr, &h2d, &d2h, &th, &td);
    double gflops = (double)N;
    gflops = gflops / 3.0 + 0.5;
    gflops = gflops * (double)N * (double)N;
    gflops = gflops / llttime / 1024.0 / 1024.0 / 1024.0;
    printf("N NB memsize(MB) quark_pthreads time Gflops\n%d %d %lf %d %lf %lf\n", N, NB,
           (double)memsize / 1024 / 1024, 1, llttime, gflops);
    printf("TFLOP/s(n^3/3) %.4f  H2D %.3f GB in %.3f s (%.2f GB/s)  D2H %.3f GB in %.3f s "
           "(%.2f GB/s)\n",
           (double)N * N * N / 3.0 / llttime / 1e12, h2d / 1e9, th, th > 0 ? h2d / th / 1e9 : 0.0,
           d2h / 1e9, td, td > 0 ? d2h / td / 1e9 : 0.0);
    if (do_io) {
        printf("L:\n\n");
        matprint(Ap, N, LDA, 'L');
    }
    if (resid) {
        double r = 0;
        int rc = oocdag_llt_resid(Ap, N, LDA, 'S', 'D', 100, &r);
        if (rc) {
            printf("residual failed: %s\n", oocdag_last_error());
            return 1;
        }
        printf("scaled residual ||A-LL^T||/(n eps ||A||) %e\n", r);
    }
    if (verify) {
        // VERIFY (oocdag.cpp:379-464): regenerate A, solve L L^T X = A, X = I
        printf("Verify... ");
        double tv = wall();
        std::vector<double> B((size_t)LDA * N);
        oocdag_pdmatgen('S', 'D', N, N, NB, NB, B.data(), LDA, 100);
        int rc = oocdag_potrs(Ap, N, LDA, B.data(), N, LDA);
        if (rc) {
            printf("potrs failed: %s\n", oocdag_last_error());
            return 1;
        }
        double maxerr = 0, l2 = 0;
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                double e = std::fabs(B[(size_t)j * LDA + i] - (i == j ? 1.0 : 0.0));
                maxerr = std::max(maxerr, e);
                l2 += e * e;
            }
        printf("time %lf\n", wall() - tv);
        printf("max error %lf \n", maxerr);
        printf("max L2 error %lf \n", std::sqrt(l2));
    }
    return info > 0 ? 2 : 0;
}
