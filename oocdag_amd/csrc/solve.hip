// Triangular solves with the packed factor (dpotrs) and the LIN scaled
// residual on the device.
//
// Solve: the reference only has the compiled-out VERIFY block that calls
// dpotrs('L', N, nrhs, L, A2) on the host (oocdag.cpp:431-441).  Here the
// right-hand sides are kept transposed (X^T, nrhs x N) so both sweeps are
// right-side triangular solves on NB-wide tile columns:
//   forward  L Y = B     <=>  Y^T = B^T L^-T  (tile by tile, then
//                             B^T[:, rest] -= Y^T_k L[rest, k]^T)
//   backward L^T X = Y   <=>  X^T = Y^T L^-1  (last tile first, then
//                             Y^T[:, :k] -= X^T_k L[k, :k])
// with 64 x 64 diagonal-block inverses from the LDS kernel and every
// product on the MFMA GEMM.
//
// Residual: PDPOTRRV('L') + PDLAFCHK (LIN/pdpotrrv.f:240-314,
// LIN/pdlafchk.f:181-225): R = L L^T is rebuilt one tile column at a time
// (GEMM over the previous tile columns plus the lower-triangular diagonal
// tile), compared with the regenerated PDMATGEN matrix, and
//   ||A - L L^T||_inf / (N * eps * ||A||_inf),  eps = 2^-53
// is formed from per-row sums accumulated on the device.
#include "ooc_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace ooc {

namespace {

constexpr int PB = 64;

struct LcgMap {
    unsigned a, c;
};

LcgMap map_then(LcgMap f, LcgMap g) {
    LcgMap r;
    r.a = (g.a * f.a) & 0x7fffffffu;
    r.c = (g.a * f.c + g.c) & 0x7fffffffu;
    return r;
}

LcgMap map_pow(unsigned long long n) {
    LcgMap r{1u, 0u}, b{1103515245u, 12345u};
    while (n) {
        if (n & 1ull) r = map_then(r, b);
        b = map_then(b, b);
        n >>= 1;
    }
    return r;
}

__device__ __forceinline__ unsigned map_apply(LcgMap m, unsigned x) {
    return (m.a * x + m.c) & 0x7fffffffu;
}

// Regenerates PDMATGEN('S'|'F', DIAG) element values along one row, column
// by column (stream position min(i,j)*N + max(i,j)).
struct RowGen {
    LcgMap pow2[40];
    LcgMap stepN;
    unsigned x1;
    int N, symm_F, diagD;
};

struct RowGenState {
    unsigned x;
    bool have;
};

__device__ __forceinline__ double gen_next(const RowGen &g, RowGenState &st, int gi, int gj) {
    if (g.symm_F) return (gi == gj) ? 10.0 : (double)0.1f;
    if (!st.have) {
        unsigned long long p = (gi >= gj) ? (unsigned long long)gj * g.N + gi
                                          : (unsigned long long)gi * g.N + gj;
        unsigned x = g.x1;
        for (int k = 0; p; ++k, p >>= 1)
            if (p & 1ull) x = map_apply(g.pow2[k], x);
        st.x = x;
        st.have = true;
    } else {
        st.x = (gi >= gj) ? map_apply(g.stepN, st.x) : map_apply(g.pow2[0], st.x);
    }
    double v = 1.0 - 2.0 * ((double)st.x / 2147483648.0);
    if (gi == gj && g.diagD) v = fabs(v) + (double)g.N;
    return v;
}

RowGen make_rowgen(int N, char aform, char diag, int iseed) {
    RowGen g;
    for (int k = 0; k < 40; ++k) g.pow2[k] = map_pow(1ull << k);
    g.stepN = map_pow((unsigned long long)N);
    g.x1 = (1103515245u * ((unsigned)iseed & 0x7fffffffu) + 12345u) & 0x7fffffffu;
    g.N = N;
    g.symm_F = ((aform & ~0x20) == 'F');
    g.diagD = ((diag & ~0x20) == 'D');
    return g;
}

// Row sums of |A - R| for tile column jt; one thread per slot row, so
// each row's sum is added by one thread in a fixed column order and the
// launches (one per tile column, ordered on one stream) accumulate in a
// fixed order too: the residual is reproducible bit for bit.
//   R      : L L^T for tile rows jt..bb-1 of this column (rect slots)
//   Y      : packed factor (its diagonal tile's strict upper part holds A)
// Tiles of logical edge nb sit in slots of edge ts; rows and columns past
// nb in a slot, or past N, are padding and skipped.
__global__ __launch_bounds__(256) void k_resid_column(RowGen g, const double *R, const double *Y,
                                                      int nb, int ts, int bb, int jt,
                                                      double *rowsum) {
    const int r = blockIdx.x * 256 + threadIdx.x;  // slot row offset below the column top
    const int rows = (bb - jt) * ts;
    if (r >= rows) return;
    const int ti = r / ts, rt = r - ti * ts;
    const int gr = (jt + ti) * nb + rt;
    if (rt >= nb || gr >= g.N) return;
    const long long ts2 = (long long)ts * ts;
    const long long ydiag = ((long long)jt * bb - ((long long)jt * (jt - 1)) / 2) * ts2;
    double rs = 0.0;
    RowGenState st{0u, false};
    for (int c = 0; c < nb; ++c) {
        const int gc = jt * nb + c;
        if (gc >= g.N) break;
        const double a = gen_next(g, st, gr, gc);
        if (gr >= gc)
            rs += fabs(a - R[ti * ts2 + rt + (long long)c * ts]);
        else
            rs += fabs(a - Y[ydiag + rt + (long long)c * ts]);   // ti == 0 here
    }
    rowsum[gr] += rs;
}

// ||A||_inf row sums of the regenerated symmetric matrix: one thread per
// row walks the whole row left to right (stream positions j*N + i below
// the diagonal, i*N + j above it), a fixed order, no atomics.
__global__ __launch_bounds__(256) void k_row_norms(RowGen g, double *anorm) {
    const int gr = blockIdx.x * 256 + threadIdx.x;
    if (gr >= g.N) return;
    RowGenState st{0u, false};
    double s = 0.0;
    for (int gc = 0; gc < g.N; ++gc) s += fabs(gen_next(g, st, gr, gc));
    anorm[gr] = s;
}

// Row sums of |A_host - A| over the strict upper tiles above tile (jt, jt):
// U holds rows [0, jt*nb) of columns [jt*nb, jt*nb + nb) (ld = ldu).  One
// thread per row; launches ordered on one stream.
__global__ __launch_bounds__(256) void k_resid_upper(RowGen g, const double *U, long long ldu,
                                                     int nb, int jt, double *rowsum) {
    const int gr = blockIdx.x * 256 + threadIdx.x;
    if (gr >= jt * nb || gr >= g.N) return;
    RowGenState st{0u, false};
    double rs = 0.0;
    for (int c = 0; c < nb; ++c) {
        const int gc = jt * nb + c;
        if (gc >= g.N) break;
        double a = gen_next(g, st, gr, gc);
        rs += fabs(U[gr + (long long)c * ldu] - a);
    }
    rowsum[gr] += rs;
}

// Copy tile column jt (tiles jt..bb-1, packed) to a rect scratch column with
// the strict upper triangle of the diagonal tile zeroed.
__global__ void k_masked_column(const double *Y, double *S, int nb, int bb, int jt) {
    const long long n = (long long)(bb - jt) * nb * nb;
    const long long base = ((long long)jt * bb - ((long long)jt * (jt - 1)) / 2) * nb * nb;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        long long t = idx / ((long long)nb * nb);
        long long w = idx - t * nb * nb;
        int r = (int)(w % nb), c = (int)(w / nb);
        S[idx] = (t == 0 && r < c) ? 0.0 : Y[base + idx];
    }
}

// out(c, r) = in(r, c): in is rows x cols (ld_in), out is cols x rows (ld_out).
// With slots (ts > 0) a logical index g of the tiled dimension becomes the
// slot index (g / nbl) * ts + g % nbl: mode 1 maps in's row index (= out's
// column index), mode 2 in's column index.
__global__ void k_transpose(const double *in, long long ld_in, double *out, long long ld_out,
                            int rows, int cols, int nbl, int ts, int mode) {
    __shared__ double t[32][33];
    int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    auto map = [&](long long g) { return ts > 0 ? (g / nbl) * ts + g % nbl : g; };
    for (int k = 0; k < 32; k += 8) {
        int r = bx + tx, c = by + ty + k;
        const long long ci = mode == 2 ? map(c) : c;
        t[ty + k][tx] = (r < rows && c < cols) ? in[r + ci * ld_in] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < 32; k += 8) {
        int c = by + tx, r = bx + ty + k;
        const long long ro = mode == 1 ? map(r) : r;
        if (c < cols && r < rows) out[c + ro * ld_out] = t[tx][ty + k];
    }
}

}  // namespace

hipError_t diag_inverses(const Opnd &L, int n, double *dinv, hipStream_t s) {
    for (int c = 0; c < n / PB; ++c) {
        Opnd D = L;
        D.r0 += c * PB;
        D.c0 += c * PB;
        hipError_t e = launch_potrf64(D, 0, dinv + (size_t)c * PB * PB, nullptr, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t trsm_right_lower_t(const Opnd &B, int m, int n, const Opnd &L, const double *dinv,
                              hipStream_t s) {
    const int nblk = n / PB;
    for (int c = 0; c < nblk; ++c) {
        hipError_t e;
        if (c > 0) {
            Gemm g;
            g.A = B;
            g.B = L; g.B.r0 += c * PB;
            g.C = B; g.C.c0 += c * PB;
            g.M = m; g.N = PB; g.K = c * PB;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if ((e = launch_gemm(g, GEMM_NARROW, s)) != hipSuccess) return e;
        }
        Gemm g;
        g.A = B; g.A.c0 += c * PB;
        g.C = g.A;
        g.B = plain((double *)dinv + (size_t)c * PB * PB, PB);
        g.M = m; g.N = PB; g.K = PB;
        g.alpha = 1.0; g.beta = 0.0; g.lower = 0;
        if ((e = launch_gemm(g, GEMM_APPLY, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t trsm_right_lower_n(const Opnd &B, int m, int n, const Opnd &L, const double *dinv,
                              hipStream_t s) {
    const int nblk = n / PB;
    for (int c = nblk - 1; c >= 0; --c) {
        hipError_t e;
        if (c < nblk - 1) {
            // B_c -= B[:, c+1:] * L[c+1:, c]
            Gemm g;
            g.A = B; g.A.c0 += (c + 1) * PB;
            g.B = L; g.B.r0 += (c + 1) * PB; g.B.c0 += c * PB;  // stored, read transposed
            g.C = B; g.C.c0 += c * PB;
            g.M = m; g.N = PB; g.K = n - (c + 1) * PB;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if ((e = launch_gemm(g, GEMM_NARROW_TB, s)) != hipSuccess) return e;
        }
        // B_c <- B_c * D_c  (D_c = L_cc^-1)
        Gemm g;
        g.A = B; g.A.c0 += c * PB;
        g.C = g.A;
        g.B = plain((double *)dinv + (size_t)c * PB * PB, PB);
        g.M = m; g.N = PB; g.K = PB;
        g.alpha = 1.0; g.beta = 0.0; g.lower = 0;
        if ((e = launch_gemm(g, GEMM_APPLY_TB, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t potrs_tiles_t(const double *dY, int N, int NB, const double *dinv_all, double *Xt,
                         int nrhs, long long ldx, hipStream_t s) {
    const int bb = (N + NB - 1) / NB;
    const Opnd Lp = panel((double *)dY, NB, bb, 1);
    const size_t per_tile = (size_t)(NB / PB) * PB * PB;
    hipError_t e;
    // forward: Y^T = B^T L^-T
    for (int k = 0; k < bb; ++k) {
        Opnd Bk = plain(Xt, ldx, 0, k * NB);
        Opnd Lk = Lp; Lk.r0 = k * NB; Lk.c0 = k * NB;
        if ((e = trsm_right_lower_t(Bk, nrhs, NB, Lk, dinv_all + k * per_tile, s)) != hipSuccess)
            return e;
        if (k + 1 < bb) {
            Gemm g;
            g.A = plain(Xt, ldx, 0, k * NB);
            g.B = Lp; g.B.r0 = (k + 1) * NB; g.B.c0 = k * NB;
            g.C = plain(Xt, ldx, 0, (k + 1) * NB);
            g.M = nrhs; g.N = (bb - k - 1) * NB; g.K = NB;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if ((e = launch_gemm(g, nrhs <= 64 ? GEMM_RHS : GEMM_GENERAL, s)) != hipSuccess) return e;
        }
    }
    // backward: X^T = Y^T L^-1
    for (int k = bb - 1; k >= 0; --k) {
        Opnd Bk = plain(Xt, ldx, 0, k * NB);
        Opnd Lk = Lp; Lk.r0 = k * NB; Lk.c0 = k * NB;
        if ((e = trsm_right_lower_n(Bk, nrhs, NB, Lk, dinv_all + k * per_tile, s)) != hipSuccess)
            return e;
        if (k > 0) {
            Gemm g;
            g.A = plain(Xt, ldx, 0, k * NB);
            g.B = Lp; g.B.r0 = k * NB; g.B.c0 = 0;  // L[k, :k], read transposed
            g.C = plain(Xt, ldx, 0, 0);
            g.M = nrhs; g.N = k * NB; g.K = NB;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if ((e = launch_gemm(g, nrhs <= 64 ? GEMM_RHS_TB : GEMM_GENERAL_TB, s)) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

hipError_t launch_transpose(const double *in, long long ld_in, double *out, long long ld_out,
                            int rows, int cols, hipStream_t s, int nbl, int ts, int mode) {
    if (ts == nbl) ts = 0;   // slots without padding: indices map to themselves
    dim3 grid((rows + 31) / 32, (cols + 31) / 32), block(32, 8);
    hipLaunchKernelGGL(k_transpose, grid, block, 0, s, in, ld_in, out, ld_out, rows, cols, nbl,
                       ts, mode);
    return hipGetLastError();
}

hipError_t llt_resid_tiles(const double *dY, int N, int NB, char aform, char diag, int iseed,
                           double *R, double *S, double *rowsum, double *anorm, hipStream_t s) {
    const int bb = (N + NB - 1) / NB;
    const int ts = tile_edge(NB);   // slot edge: L L^T is rebuilt on the padded slots
    RowGen gen = make_rowgen(N, aform, diag, iseed);
    const Opnd Lp = panel((double *)dY, ts, bb, 1);
    hipError_t e;
    hipLaunchKernelGGL(k_row_norms, dim3((N + 255) / 256), dim3(256), 0, s, gen, anorm);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int j = 0; j < bb; ++j) {
        const int rows = (bb - j) * ts;
        long long n = (long long)rows * ts;
        int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_masked_column, dim3(blocks), dim3(256), 0, s, dY, S, ts, bb, j);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (j > 0) {
            Gemm g;
            g.A = Lp; g.A.r0 = j * ts; g.A.c0 = 0;
            g.B = g.A;
            g.C = panel(R, ts, bb - j, 0);
            g.M = rows; g.N = ts; g.K = j * ts;
            g.alpha = 1.0; g.beta = 0.0; g.lower = 0;
            if ((e = launch_gemm(g, GEMM_GENERAL, s)) != hipSuccess) return e;
        }
        Gemm g;
        g.A = panel(S, ts, bb - j, 0);
        g.B = g.A;
        g.C = panel(R, ts, bb - j, 0);
        g.M = rows; g.N = ts; g.K = ts;
        g.alpha = 1.0; g.beta = j > 0 ? 1.0 : 0.0; g.lower = 0;
        if ((e = launch_gemm(g, GEMM_GENERAL, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_resid_column, dim3((rows + 255) / 256), dim3(256), 0, s, gen, R, dY,
                           NB, ts, bb, j, rowsum);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t resid_upper_block(const double *U, long long ldu, int N, int NB, int jt, char aform,
                             char diag, int iseed, double *rowsum, hipStream_t s) {
    if (jt == 0) return hipSuccess;
    RowGen gen = make_rowgen(N, aform, diag, iseed);
    int rows = jt * NB;
    hipLaunchKernelGGL(k_resid_upper, dim3((rows + 255) / 256), dim3(256), 0, s, gen, U, ldu, NB,
                       jt, rowsum);
    return hipGetLastError();
}

}  // namespace ooc
