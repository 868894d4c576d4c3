// Mixed-precision Cholesky with FP64 iterative refinement, and the Gram-
// matrix generator of the mixed-precision workloads.  No reference
// counterpart (the reference factors in FP64 only); every result is checked
// against the FP64 path.
//
// Factor: the working copy W of A (packed FP64 tiles) is factored tile
// column by tile column; the diagonal POTRF and the TRSM of each column run
// in FP64 (factor_column), the trailing update in FP64, FP32 or BF16 (per-
// tile scaled) on MFMA with FP32 accumulation into the FP64 tiles.
// Refinement, all in FP64 with right-hand sides kept transposed (nrhs x N):
//   x = W^-1 b;  repeat  r = b - A x,  eta = max_k ||r_k|| / (||A|| ||x_k|| + ||b_k||),
//                        stop if eta <= tol,  x += W^-1 r
// A low-precision factor that is not positive definite, or refinement that
// stalls (eta not halving) or runs out of iterations, falls back one level:
// BF16 -> FP32 -> FP64.  A factor that meets a bad pivot stops its queued
// low-precision updates at once, so a failed attempt costs little more than
// the columns up to that pivot.
#include "ooc_internal.h"
#include "../../include/oocdag.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <vector>

namespace ooc {

int mfma_rowmap(int kind, hipStream_t s);
hipError_t convert_super(const double *W, int nb, int ym, int j0, int sw, int i0, int nrt, int mode,
                         unsigned *amax, float *scale, void *X, hipStream_t s);
hipError_t launch_update_lp(int mode, double *C, int c_ym, int nb, int r0, const void *X,
                            long long ldx, int K, const float *scale, int M, int N, int rowmap,
                            hipStream_t s, const int *stop);
hipError_t convert_region_f32(const double *W, int nb, int ym, int gr0, int gc0, int R, int K,
                              float *X, hipStream_t s);

namespace {

// counter-based normal variates: splitmix64 (Steele, Lea, Flood 2014) of
// (seed, stream, index) -> two uniforms -> Box-Muller
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double normal_at(unsigned long long seed, unsigned long long stream,
                                            unsigned long long idx) {
    unsigned long long h = splitmix64(seed ^ splitmix64(stream * 0x632be59bd9b4e019ull + idx));
    unsigned long long h2 = splitmix64(h);
    double u1 = ((h >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
    double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);         // [0, 1)
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void k_fill_normal(double *G, long long rows, long long cols, long long ld,
                              unsigned long long seed, unsigned long long stream, long long nvalid_r,
                              long long nvalid_c) {
    const long long total = rows * cols;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx % rows, c = idx / rows;
        G[r + c * ld] = (r < nvalid_r && c < nvalid_c)
                            ? normal_at(seed, stream, (unsigned long long)(c * nvalid_r + r))
                            : 0.0;
    }
}

// Diagonal tiles: copy the lower triangle into the upper one and add mu on
// the diagonal (1 on the padded diagonal past N).
__global__ void k_finish_diag(double *Y, int N, int nb, int bb, double mu) {
    const int t = blockIdx.y;
    const long long base = ((long long)t * bb - ((long long)t * (t - 1)) / 2) * nb * nb;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nb * nb;
         idx += gridDim.x * blockDim.x) {
        int r = idx % nb, c = idx / nb;
        int g = t * nb + r;
        if (r == c)
            Y[base + idx] = (g < N) ? Y[base + idx] + mu : 1.0;
        else if (r < c)
            Y[base + idx] = Y[base + c + (long long)r * nb];
    }
}

// The same for one diagonal tile T of tile row/column t.
__global__ void k_finish_diag_at(double *T, int N, int nb, int t, double mu) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nb * nb;
         idx += gridDim.x * blockDim.x) {
        int r = idx % nb, c = idx / nb;
        int g = t * nb + r;
        if (r == c)
            T[idx] = (g < N) ? T[idx] + mu : 1.0;
        else if (r < c)
            T[idx] = T[c + (long long)r * nb];
    }
}

// Row sums of |A| over the full symmetric matrix from its packed lower tiles
// (diagonal tiles stored full).  One block per 64 rows r0..r0+63 of tile row
// ti reads exactly those rows of A: the row part from tiles (ti, 0..ti), a
// wave per tile column with lane = row (512 contiguous bytes per load), and
// the column part from tiles (ti+1.., ti) -- rows of A above their diagonal
// are columns of the lower tiles below it -- a wave per column, summed by a
// wave reduction.  The four waves' partial sums are combined in a fixed
// order: no atomics, the same result on every run.
__global__ __launch_bounds__(256) void k_tile_rowabs(const double *Y, int nb, int bb,
                                                     double *rowsum) {
    const int ti = blockIdx.y, r0 = blockIdx.x * 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long per = (long long)nb * nb;
    auto tile = [&](int i, int j) {
        return Y + (i + (long long)j * bb - ((long long)j * (j + 1)) / 2) * per;
    };
    double acc = 0.0;   // this lane's row r0 + lane (row part)
    for (int tj = 0; tj <= ti; ++tj) {
        const double *T = tile(ti, tj) + r0 + lane;
        for (int c = w; c < nb; c += 4) acc += fabs(T[(long long)c * nb]);
    }
    __shared__ double part[4][64];
    __shared__ double colsum[64];
    if (threadIdx.x < 64) colsum[threadIdx.x] = 0.0;
    __syncthreads();
    for (int cc = w; cc < 64; cc += 4) {   // column r0 + cc of the tiles below
        double cs = 0.0;
        for (int tk = ti + 1; tk < bb; ++tk) {
            const double *col = tile(tk, ti) + (long long)(r0 + cc) * nb;
            for (int r = lane; r < nb; r += 64) cs += fabs(col[r]);
        }
        for (int o = 32; o > 0; o >>= 1) cs += __shfl_down(cs, o);
        if (lane == 0) colsum[cc] = cs;
    }
    part[w][lane] = acc;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int t = threadIdx.x;
        rowsum[(long long)ti * nb + r0 + t] =
            ((part[0][t] + part[1][t]) + (part[2][t] + part[3][t])) + colsum[t];
    }
}

// Per right-hand side (row k of the transposed block) max |.| over N columns.
__global__ void k_rows_absmax(const double *Xt, long long ldx, int nrhs, int N, double *out) {
    const int k = blockIdx.x;
    double m = 0.0;
    for (int c = threadIdx.x; c < N; c += blockDim.x) m = fmax(m, fabs(Xt[k + (long long)c * ldx]));
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
}

__global__ void k_axpy(double *y, const double *x, long long n, double a) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

// Right-hand sides (nrhs x Np, ld ldt, element (j, i) at j + i*ldt) to and
// from FP32 for the correction solves.  Column j is scaled by 2^-e, e the
// binary exponent of its largest magnitude cmax[j] (exact scaling), so the
// narrowed values neither overflow nor lose range to subnormals; rows
// j >= nrhs (padding) stay unscaled.
__device__ __forceinline__ int col_exp(const double *cmax, int j, int nrhs) {
    const double m = j < nrhs ? cmax[j] : 0.0;
    return (m > 0.0 && m < INFINITY) ? ilogb(m) : 0;
}

__global__ void k_to_f32_cols(const double *x, float *y, long long n, int ldt, int nrhs,
                              const double *cmax) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = (float)ldexp(x[i], -col_exp(cmax, (int)(i % ldt), nrhs));
}

// y = beta*y + 2^e * x (beta 0 or 1)
__global__ void k_from_f32_cols(double *y, const float *x, long long n, int ldt, int nrhs,
                                const double *cmax, double beta) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double v = ldexp((double)x[i], col_exp(cmax, (int)(i % ldt), nrhs));
        y[i] = beta != 0.0 ? y[i] + v : v;
    }
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

hipError_t gram_tiles(double *dY, int N, int NB, double mu, unsigned long long seed,
                      hipStream_t s) {
    const int bb = (N + NB - 1) / NB;
    const long long Np = (long long)bb * NB;
    double *G = nullptr;
    hipError_t e = hipMallocAsync((void **)&G, (size_t)(Np * Np) * sizeof(double), s);
    if (e != hipSuccess) return e;
    int blocks = 65536;
    hipLaunchKernelGGL(k_fill_normal, dim3(blocks), dim3(256), 0, s, G, Np, Np, Np, seed, 1ull,
                       (long long)N, (long long)N);
    if ((e = hipGetLastError()) == hipSuccess) {
        Gemm g;
        g.A = plain(G, Np);
        g.B = g.A;
        g.C = panel(dY, NB, bb, 1);
        g.M = (int)Np;
        g.N = (int)Np;
        g.K = (int)Np;
        g.alpha = 1.0 / (double)N;
        g.beta = 0.0;
        g.lower = 1;
        e = launch_gemm(g, GEMM_UPDATE, s);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_finish_diag, dim3(64, bb), dim3(256), 0, s, dY, N, NB, bb, mu);
        e = hipGetLastError();
    }
    (void)hipFreeAsync(G, s);
    hipError_t e2 = hipStreamSynchronize(s);
    return e != hipSuccess ? e : e2;
}

// G[r0 + r, c0 + c] of the Gram generator's N x N normal matrix (column
// index c*N + r of its stream; 0 past N) into a rows x cols block (ld).
__global__ void k_fill_normal_at(double *D, long long rows, long long cols, long long ld,
                                 long long r0, long long c0, unsigned long long seed, long long N) {
    const long long total = rows * cols;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long r = idx % rows, c = idx / rows, gr = r0 + r, gc = c0 + c;
        D[r + c * ld] = (gr < N && gc < N)
                            ? normal_at(seed, 1ull, (unsigned long long)(gc * N + gr))
                            : 0.0;
    }
}

// Tile column j of the Gram matrix A = G G^T / N + mu I (the matrix
// gram_tiles writes) into `col`: tiles j..bb-1 of edge NB, consecutive,
// ld = NB, the diagonal tile stored full.  G is streamed in blocks of kw
// columns through Gp ((bb - j) * NB x kw), so neither G nor A has to fit in
// device memory; the K sum runs block by block (values equal gram_tiles'
// up to the rounding of that blocking).
hipError_t gram_tile_column(double *col, int N, int NB, int j, double mu, unsigned long long seed,
                            double *Gp, int kw, hipStream_t s) {
    const int bb = (N + NB - 1) / NB;
    const long long rows = (long long)(bb - j) * NB, Np = (long long)bb * NB;
    hipError_t e;
    for (long long k0 = 0; k0 < Np; k0 += kw) {
        const int w = (int)std::min<long long>(kw, Np - k0);
        hipLaunchKernelGGL(k_fill_normal_at, dim3(4096), dim3(256), 0, s, Gp, rows, (long long)w,
                           rows, (long long)j * NB, k0, seed, (long long)N);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        Gemm g;
        g.A = plain(Gp, rows);
        g.B = g.A;   // the top NB rows: G rows of tile j
        g.C = panel(col, NB, bb - j, 0);
        g.M = (int)rows;
        g.N = NB;
        g.K = w;
        g.alpha = 1.0 / (double)N;
        g.beta = k0 == 0 ? 0.0 : 1.0;
        g.lower = 0;
        if ((e = launch_gemm(g, GEMM_GENERAL, s)) != hipSuccess) return e;
    }
    // the diagonal tile: lower triangle mirrored, mu on the diagonal
    // (1 on the padded diagonal past N); k_finish_diag on a one-tile layout
    hipLaunchKernelGGL(k_finish_diag_at, dim3(64), dim3(256), 0, s, col, N, NB, j, mu);
    return hipGetLastError();
}

hipError_t normal_block(double *dB, int rows, int cols, long long ld, unsigned long long seed,
                        unsigned long long stream, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_normal, dim3(4096), dim3(256), 0, s, dB, (long long)rows,
                       (long long)cols, ld, seed, stream, (long long)rows, (long long)cols);
    return hipGetLastError();
}

// Tile columns per super-step of the solves and the residual: the updates
// reaching past a super-step go out as one launch with K = W*NB (~4096).
int solve_step_width(int NB) { return std::max(1, 4096 / NB); }

// Y^T (nrhs x Np) += alpha * X^T A for symmetric A in packed lower tiles
// with full diagonal tiles.  Tile row i of the stored triangle gives
//   Y^T[:, j] += X^T[:, i] A[i, j]     (j <= i, the tile read transposed)
// and tile column j gives
//   Y^T[:, i] += X^T[:, j] A[i, j]^T   (i > j).
// In super-blocks of W tile rows [s0, s1): one launch for the rows' parts
// left of the block (K = (s1-s0)*NB across the W tile rows), one for the
// columns' parts below it, and per tile the parts inside the block.
hipError_t symm_t(const double *dA, int N, int NB, const double *Xt, double *Yt, long long ld,
                  int nrhs, double alpha, double beta, hipStream_t s) {
    const int bb = (N + NB - 1) / NB, W = solve_step_width(NB);
    const Opnd Ap = panel((double *)dA, NB, bb, 1);
    const GemmKind nt = nrhs <= 64 ? GEMM_RHS : GEMM_GENERAL;
    const GemmKind tt = nrhs <= 64 ? GEMM_RHS_TB : GEMM_GENERAL_TB;
    hipError_t e;
    if (beta != 1.0) {
        // scale Y^T by beta via a zero-K GEMM
        Gemm g;
        g.A = plain((double *)Xt, ld);
        g.B = Ap;
        g.C = plain(Yt, ld);
        g.M = nrhs; g.N = bb * NB; g.K = 0;
        g.alpha = 0.0; g.beta = beta; g.lower = 0;
        if ((e = launch_gemm(g, GEMM_GENERAL, s)) != hipSuccess) return e;
    }
    // Y^T[:, c0 .. c0+n*NB) += alpha X^T[:, a0 .. a0+k*NB) op(A block at (br, bc))
    auto product = [&](int a0, int c0, int br, int bc, int n, int k, GemmKind kind) -> hipError_t {
        Gemm g;
        g.A = plain((double *)Xt, ld, 0, a0 * NB);
        g.B = Ap; g.B.r0 = br * NB; g.B.c0 = bc * NB;
        g.C = plain(Yt, ld, 0, c0 * NB);
        g.M = nrhs; g.N = n * NB; g.K = k * NB;
        g.alpha = alpha; g.beta = 1.0; g.lower = 0;
        return launch_gemm(g, kind, s);
    };
    for (int s0 = 0; s0 < bb; s0 += W) {
        const int s1 = std::min(bb, s0 + W);
        // rows s0..s1-1, columns j < s0
        if (s0 > 0 && (e = product(s0, 0, s0, 0, s0, s1 - s0, tt)) != hipSuccess) return e;
        // columns s0..s1-1, rows i >= s1
        if (s1 < bb && (e = product(s0, s1, s1, s0, bb - s1, s1 - s0, nt)) != hipSuccess) return e;
        for (int i = s0; i < s1; ++i) {
            // row i, columns s0..i; column i, rows i+1..s1-1
            if ((e = product(i, s0, i, s0, i + 1 - s0, 1, tt)) != hipSuccess) return e;
            if (i + 1 < s1 && (e = product(i, i + 1, i + 1, i, s1 - i - 1, 1, nt)) != hipSuccess)
                return e;
        }
    }
    return hipSuccess;
}

// Step s of the blocked inverse X = L_kk^-1 of every diagonal tile k at once
// (64 x 64 blocks, P = NB/64 per tile row): X_jj = D_j, the inverse of the
// 64 x 64 diagonal block that the factorization left in dinv, and for
// i = j + s > j
//     X_ij = -D_i * sum_{m=j}^{i-1} L_im X_mj ,
// which needs only blocks of earlier steps.  Z_k = X^T is what the solves
// use, so X_ij is stored as Z block (j, i), X_mj read back from Z block
// (j, m).  Step 0 writes the diagonal blocks and zeros below them.  Block
// (x, y) = (tile k, j); 256 threads, thread t owns row t % 64 and columns
// t / 64 + 4u of the block, operands staged in LDS.
constexpr int TI = 64, TIS = 65;
__global__ __launch_bounds__(256) void k_tile_inv_step(const double *W, int NB, int bb,
                                                       const double *dinv_all,
                                                       long long inv_per_tile, double *Z,
                                                       int s) {
    const int k = blockIdx.x, j = blockIdx.y, i = j + s;
    const int P = NB / TI;
    if (i >= P) return;
    const int t = threadIdx.x, r = t & 63, c0 = t >> 6;
    const double *Lk = W + (k + (long long)k * bb - ((long long)k * (k + 1)) / 2) *
                               (long long)NB * NB;
    const double *D = dinv_all + (long long)k * inv_per_tile;
    double *Zk = Z + (long long)k * NB * NB;
    __shared__ double La[TI * TIS], Xb[TI * TIS];
    if (s == 0) {
        // Z block (j, j) = D_j^T; Z blocks (a, j), a > j, = 0
        const double *Dj = D + (long long)j * TI * TI;
        for (int u = 0; u < 16; ++u) {
            const int cc = c0 + 4 * u;
            // Z(j*64 + r, j*64 + cc) = D_j(cc, r)
            Zk[(long long)(j * TI + r) + (long long)(j * TI + cc) * NB] = Dj[cc + r * TI];
            for (int a = j + 1; a < P; ++a)
                Zk[(long long)(a * TI + r) + (long long)(j * TI + cc) * NB] = 0.0;
        }
        return;
    }
    double acc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = 0.0;
    for (int m = j; m < i; ++m) {
        // La = L_im (rows i*64.., cols m*64.. of the tile), Xb = X_mj
        for (int u = 0; u < 16; ++u) {
            const int cc = c0 + 4 * u;
            La[r + cc * TIS] = Lk[(long long)(i * TI + r) + (long long)(m * TI + cc) * NB];
            // X_mj(cc, r) = Z(j*64 + r, m*64 + cc): consecutive r, consecutive addresses
            if (m == j)
                Xb[r + cc * TIS] = D[(long long)j * TI * TI + r + cc * TI];
            else
                Xb[cc + r * TIS] = Zk[(long long)(j * TI + r) + (long long)(m * TI + cc) * NB];
        }
        __syncthreads();
        for (int q = 0; q < TI; ++q) {
            const double l = La[r + q * TIS];
#pragma unroll
            for (int u = 0; u < 16; ++u) acc[u] += l * Xb[q + (c0 + 4 * u) * TIS];
        }
        __syncthreads();
    }
    // X_ij = -D_i * T
    const double *Di = D + (long long)i * TI * TI;
    for (int u = 0; u < 16; ++u) {
        const int cc = c0 + 4 * u;
        Xb[r + cc * TIS] = acc[u];
        La[r + cc * TIS] = Di[r + cc * TI];
    }
    __syncthreads();
    double x[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = 0.0;
    for (int q = 0; q < TI; ++q) {
        const double d = La[r + q * TIS];
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] -= d * Xb[q + (c0 + 4 * u) * TIS];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) Xb[r + (c0 + 4 * u) * TIS] = x[u];
    __syncthreads();
    // Z(j*64 + c, i*64 + rr) = X_ij(rr, c): consecutive threads, consecutive c
    for (int u = 0; u < 16; ++u) {
        const int rr = c0 + 4 * u;
        Zk[(long long)(j * TI + r) + (long long)(i * TI + rr) * NB] = Xb[rr + r * TIS];
    }
}

// Z_k = L_kk^-T for every diagonal tile (NB x NB each, ld NB), from the
// 64-block inverses the factorization left in dinv_all: P = NB/64 launches,
// each over all tiles.
hipError_t tile_inverses(const double *dY, int N, int NB, const double *dinv_all, double *Z,
                         hipStream_t s) {
    const int bb = (N + NB - 1) / NB, P = NB / TI;
    if (NB % TI) return hipErrorInvalidValue;
    const long long inv_per_tile = (long long)P * TI * TI;
    for (int st = 0; st < P; ++st) {
        hipLaunchKernelGGL(k_tile_inv_step, dim3(bb, P - st), dim3(256), 0, s, dY, NB, bb,
                           dinv_all, inv_per_tile, Z, st);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// X^T <- B^T A^-1 with the diagonal tiles' inverses Z.  The forward sweep
// writes Y^T into Yt (nrhs x N, ld ldx) while it updates the rest of B^T in
// Xt; the backward sweep writes X^T back into Xt while it updates the rest
// of Y^T in Yt: each diagonal product lands in the other buffer, so no copy
// follows it.  Each sweep goes in super-steps of W tiles: inside one, a
// diagonal product per tile and a narrow update of the super-step's later
// tiles; after it, one update of everything beyond with K = W*NB (the
// factor's W tile columns, or tile rows, at once).
//
// f32 != 0: the whole solve in FP32 -- dY and Z the FP32 copy of a
// low-precision factor, Xt and Yt float, FP32 MFMAs (result rows per
// rowmap) -- for the correction solves of iterative refinement.
hipError_t potrs_inv_t(const void *dY, int N, int NB, const void *Z, int f32, int rowmap,
                       void *Xt, int nrhs, long long ldx, void *Yt, hipStream_t s) {
    const int bb = (N + NB - 1) / NB, W = solve_step_width(NB);
    const Opnd Lp = panel((double *)dY, NB, bb, 1);
    const size_t per = (size_t)NB * NB;
    // the FP32 factor only feeds the right-hand-side kernels (nrhs <= 64)
    if (f32 && nrhs > 64) return hipErrorInvalidValue;
    const GemmKind nt = nrhs <= 64 ? GEMM_RHS : GEMM_GENERAL;
    const GemmKind tt = nrhs <= 64 ? GEMM_RHS_TB : GEMM_GENERAL_TB;
    hipError_t e;
    auto diag = [&](const void *src, void *dst, int k, GemmKind kind) -> hipError_t {
        Gemm g;
        g.A = plain((double *)src, ldx, 0, k * NB);
        g.B = plain(f32 ? (double *)((const float *)Z + k * per) : (double *)Z + k * per, NB);
        g.C = plain((double *)dst, ldx, 0, k * NB);
        g.M = nrhs; g.N = NB; g.K = NB;
        g.alpha = 1.0; g.beta = 0.0; g.lower = 0;
        g.f32 = f32;
        g.rowmap = rowmap;
        return launch_gemm(g, kind, s);
    };
    // C[:, c0 .. c0+n*NB) -= src[:, a0 .. a0+k*NB) * op(L block at (br, bc))
    auto update = [&](const void *src, int a0, void *dst, int c0, int br, int bc, int n,
                      int k, GemmKind kind) -> hipError_t {
        Gemm g;
        g.A = plain((double *)src, ldx, 0, a0 * NB);
        g.B = Lp; g.B.r0 = br * NB; g.B.c0 = bc * NB;
        g.C = plain((double *)dst, ldx, 0, c0 * NB);
        g.M = nrhs; g.N = n * NB; g.K = k * NB;
        g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
        g.f32 = f32;
        g.rowmap = rowmap;
        return launch_gemm(g, kind, s);
    };
    for (int s0 = 0; s0 < bb; s0 += W) {   // Y^T_k = B^T_k L_kk^-T = B^T_k Z_k
        const int s1 = std::min(bb, s0 + W);
        for (int k = s0; k < s1; ++k) {
            if ((e = diag(Xt, Yt, k, tt)) != hipSuccess) return e;
            // B^T_i -= Y^T_k L_ik^T, k < i < s1
            if (k + 1 < s1 && (e = update(Yt, k, Xt, k + 1, k + 1, k, s1 - k - 1, 1, nt)) != hipSuccess)
                return e;
        }
        // B^T_i -= sum_{s0 <= k < s1} Y^T_k L_ik^T, i >= s1
        if (s1 < bb && (e = update(Yt, s0, Xt, s1, s1, s0, bb - s1, s1 - s0, nt)) != hipSuccess)
            return e;
    }
    for (int s1 = bb; s1 > 0; s1 -= W) {   // X^T_k = Y^T_k L_kk^-1 = Y^T_k Z_k^T
        const int s0 = std::max(0, s1 - W);
        for (int k = s1 - 1; k >= s0; --k) {
            if ((e = diag(Yt, Xt, k, nt)) != hipSuccess) return e;
            // Y^T_j -= X^T_k L_kj, s0 <= j < k (tile row k read transposed)
            if (k > s0 && (e = update(Xt, k, Yt, s0, k, s0, k - s0, 1, tt)) != hipSuccess) return e;
        }
        // Y^T_j -= sum_{s0 <= k < s1} X^T_k L_kj, j < s0
        if (s0 > 0 && (e = update(Xt, s0, Yt, 0, s0, 0, s0, s1 - s0, tt)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// FP64 trailing updates for the plain path of the look-ahead panel loop.
struct Fp64PanelUpdater : PanelUpdater {
    double *W;
    int NB, bb;
    hipError_t prepare(int, hipStream_t) override { return hipSuccess; }
    hipError_t update(int j, int c0, int nc, hipStream_t s) override {
        const Opnd P = panel(W, NB, bb, 1);
        Gemm g;
        g.A = P; g.A.r0 = c0 * NB; g.A.c0 = j * NB;
        g.B = g.A;
        g.C = P; g.C.r0 = c0 * NB; g.C.c0 = c0 * NB;
        g.M = (bb - c0) * NB; g.N = nc * NB; g.K = NB;
        g.alpha = -1.0; g.beta = 1.0; g.lower = 1;
        return launch_gemm(g, GEMM_UPDATE, s);
    }
};

// Tile columns per super-column: low-precision updates of depth ~2048.
int mp_super_width(int NB, int bb) { return std::max(1, std::min(bb, 2048 / NB)); }

// Factor W (packed tiles) with the given update precision; 64-block inverses
// of every diagonal tile land in dinv_all for the solves.
//
// precision 0: the FP64 look-ahead panel loop (one tile column at a time).
// precision 1/2 (FP32/BF16): super-columns of sw tile columns.  Inside a
// super-column every tile column is factored in FP64 (factor_column) and
// applied to the rest of the super-column with the FP64 GEMM; the finished
// super-column is converted once (convert_super) and drives
//   narrow: the next super-column        (low precision, K = sw*NB, stream s)
//   bulk  : everything to its right      (low precision, stream s1)
// so the next super-column's FP64 work overlaps the bulk update.  lpbuf
// holds two converted super-columns (by parity), scale and amax 2*bb*sw and
// bb*sw entries (BF16: one per tile).  Once a diagonal block reports a bad
// pivot (dinfo < INT_MAX) the low-precision updates still queued return at
// once: the factor is discarded.
int factor_mixed(double *W, int N, int NB, int precision, double *dinv_all, int *dinfo,
                 void *lpbuf, float *scale, unsigned *amax, float *colbuf, hipStream_t s,
                 hipStream_t s1, hipEvent_t w_ready) {
    const int bb = (N + NB - 1) / NB;
    const size_t inv_per_tile = (size_t)(NB / 64) * 64 * 64;
    const Opnd P = panel(W, NB, bb, 1);
    if (precision == 0) {
        Fp64PanelUpdater up;
        up.W = W;
        up.NB = NB;
        up.bb = bb;
        int rc = factor_panel_la(P, NB, bb, bb, 0, dinv_all, inv_per_tile, dinfo, s, s1, up,
                                 [](int) { return 0; });
        return rc ? -2 : 0;
    }
    // super-column update operand: 0 FP32, 1 scaled BF16, 2 FP32-accurate
    // three BF16 pieces (the FP32 route unless variant bit 4)
    const int mode = precision == 2 ? 1 : (mp_variant() & 4) ? 0 : 2;
    // overlap: the bulk update of super-column J runs on s1 beside the FP64
    // factorization of super-column J+1 (else after it on s)
    const bool overlap = (mp_variant() & 2) == 0;
    const int rowmap = mfma_rowmap(mode == 0 ? 0 : 1, s);
    const int rowmap32 = mfma_rowmap(0, s);
    if (rowmap < 0 || rowmap32 < 0) return -1;
    const int sw = mp_super_width(NB, bb);
    const int nsc = (bb + sw - 1) / sw;
    const size_t esz = mode == 0 ? 4 : mode == 1 ? 2 : 6;
    const size_t lp_bytes = (size_t)bb * NB * sw * NB * esz;
    void *lp[2] = {lpbuf, (char *)lpbuf + lp_bytes};
    float *sc[2] = {scale, scale + (size_t)bb * sw};   // BF16: per tile of a super-column

    hipEvent_t ev_fact, ev_bulk[2];
    if (hipEventCreateWithFlags(&ev_fact, hipEventDisableTiming) != hipSuccess) return -4;
    if (hipEventCreateWithFlags(&ev_bulk[0], hipEventDisableTiming) != hipSuccess) return -4;
    if (hipEventCreateWithFlags(&ev_bulk[1], hipEventDisableTiming) != hipSuccess) return -4;
    struct Cleanup {
        hipEvent_t e[3];
        ~Cleanup() {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } cl{{ev_fact, ev_bulk[0], ev_bulk[1]}};
    bool any_bulk = false;
#define MP_TRY(x)                           \
    do {                                    \
        if ((x) != hipSuccess) return -4;   \
    } while (0)
    // Tile column t in FP64 by factor_column's 64-wide left-looking steps,
    // split recursively into halves while a half is a multiple of 128 wide
    // (NB = 512: 256 + 256, each 128 + 128): a right half's update from its
    // left half (K = the half's width) runs on the FP32 update kernel, so the
    // FP64 steps stream the tall panel far less often.
    std::function<hipError_t(int, int, double *)> factor_cols = [&](int c0, int w,
                                                                   double *dinv_c) -> hipError_t {
        const int h = w / 2;
        if (w < 256 || h % 128)
            return factor_column(P, c0, bb * NB - c0, w, c0, dinv_c, dinfo, s);
        hipError_t e = factor_cols(c0, h, dinv_c);
        if (e != hipSuccess) return e;
        const int r0 = c0 + h, M = bb * NB - r0;
        if ((e = convert_region_f32(W, NB, bb, r0, c0, M, h, colbuf, s)) != hipSuccess ||
            (e = launch_update_lp(0, W, bb, NB, r0, colbuf, M, h, nullptr, M, h, rowmap32, s,
                                  dinfo)) !=
                hipSuccess)
            return e;
        return factor_cols(r0, h, dinv_c + (size_t)(h / 64) * 64 * 64);
    };
    auto factor_column_halves = [&](int t) -> hipError_t {
        return factor_cols(t * NB, NB, dinv_all + (size_t)t * inv_per_tile);
    };
    for (int J = 0; J < nsc; ++J) {
        const int j0 = J * sw, j1 = std::min(j0 + sw, bb);
        // super-column J received bulk(J-2) on s1 (bulks run in order)
        if (J >= 2) MP_TRY(hipStreamWaitEvent(s, ev_bulk[J & 1], 0));
        for (int t = j0; t < j1; ++t) {
            MP_TRY(factor_column_halves(t));
            if (t + 1 < j1) {
                // the rest of the super-column from tile column t, on the
                // FP32 update kernel (FP32 copy of the column, FP32 sums
                // subtracted from the FP64 tiles): its error is far below
                // the low-precision bulk updates'
                const int nrt = bb - t - 1;
                MP_TRY(convert_super(W, NB, bb, t, 1, t + 1, nrt, 0, nullptr, nullptr, colbuf, s));
                MP_TRY(launch_update_lp(0, W, bb, NB, (t + 1) * NB, colbuf, (long long)nrt * NB, NB,
                                        nullptr, nrt * NB, (j1 - t - 1) * NB, rowmap32, s, dinfo));
            }
        }
        if (j1 >= bb) break;
        const int nrt = bb - j1;
        const int K = (j1 - j0) * NB;
        const long long ldx = mode == 0 ? (long long)nrt * NB : mode == 1 ? (long long)K : 3LL * K;
        // lp[J & 1] was last read by bulk(J-2), finished above
        MP_TRY(convert_super(W, NB, bb, j0, j1 - j0, j1, nrt, mode, amax, sc[J & 1], lp[J & 1], s));
        // the next super-column also received bulk(J-1); past the first
        // super-column W is copied in on s1 (w_ready)
        if (J >= 1) MP_TRY(hipStreamWaitEvent(s, ev_bulk[(J - 1) & 1], 0));
        if (J == 0 && w_ready) MP_TRY(hipStreamWaitEvent(s, w_ready, 0));
        const int k1 = std::min(j1 + sw, bb);
        MP_TRY(launch_update_lp(mode, W, bb, NB, j1 * NB, lp[J & 1], ldx, K, sc[J & 1],
                                (bb - j1) * NB, (k1 - j1) * NB, rowmap, s, dinfo));
        if (k1 < bb) {
            const long long roff = (long long)(k1 - j1) * NB;   // X row of global row k1*NB
            const void *Xb = mode == 0 ? (const void *)((const float *)lp[J & 1] + roff)
                                       : (const void *)((const unsigned short *)lp[J & 1] +
                                                        roff * ldx);
            if (overlap) {
                MP_TRY(hipEventRecord(ev_fact, s));
                MP_TRY(hipStreamWaitEvent(s1, ev_fact, 0));
            }
            hipStream_t sb = overlap ? s1 : s;
            // BF16 scales: one row of (j1 - j0) per tile row, from row j1
            MP_TRY(launch_update_lp(mode, W, bb, NB, k1 * NB, Xb, ldx, K,
                                    sc[J & 1] + (size_t)(k1 - j1) * (j1 - j0), (bb - k1) * NB,
                                    (bb - k1) * NB, rowmap, sb, dinfo));
            MP_TRY(hipEventRecord(ev_bulk[J & 1], sb));
            any_bulk = overlap;
        }
    }
    if (any_bulk) {
        MP_TRY(hipEventRecord(ev_fact, s1));
        MP_TRY(hipStreamWaitEvent(s, ev_fact, 0));
    }
#undef MP_TRY
    return 0;
}

}  // namespace ooc

namespace ooc {

// ---------------------------------------------------------------------------
// Reusable pieces: a resident factorization (kept by the server's factor
// cache) and the refinement solve against it.
// ---------------------------------------------------------------------------

namespace {
template <class T>
hipError_t salloc(T **p, size_t bytes, hipStream_t s) {
    return hipMallocAsync((void **)p, bytes, s);
}
}  // namespace

__global__ void k_to_f32(const double *x, float *y, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

// W, Z, the 64-block inverses, their FP32 copies (low-precision factors)
// and the two low-precision super-column buffers
size_t mp_factor_bytes(int N, int NB) {
    const size_t bb = (size_t)((N + NB - 1) / NB);
    const size_t tiles = bb * (bb + 1) / 2 * NB * NB;
    const size_t sw = (size_t)mp_super_width(NB, (int)bb);
    return (tiles + bb * NB * NB + bb * NB * 64) * 8 + (tiles + bb * NB * NB) * 4 +
           2 * bb * NB * sw * NB * 6 + bb * NB * NB * 4;
}

void mp_release(MpFactor &f, hipStream_t s) {
    if (f.W) (void)hipFreeAsync(f.W, s);
    if (f.Z) (void)hipFreeAsync(f.Z, s);
    if (f.dinv) (void)hipFreeAsync(f.dinv, s);
    if (f.W32) (void)hipFreeAsync(f.W32, s);
    if (f.Z32) (void)hipFreeAsync(f.Z32, s);
    f.W = f.Z = f.dinv = nullptr;
    f.W32 = f.Z32 = nullptr;
}

int mp_factor(MpFactor &f, const double *dA, int N, int NB, int precision, hipStream_t s0,
              hipStream_t s1) {
    const double t0 = now();
    const int bb = (N + NB - 1) / NB;
    const long long Np = (long long)bb * NB;
    const size_t tiles = (size_t)bb * (bb + 1) / 2 * NB * NB;
    f.A = dA;
    f.N = N;
    f.NB = NB;
    f.precision = precision;
    f.info = 0;
    f.bytes = mp_factor_bytes(N, NB);
    void *lp = nullptr;
    float *scale = nullptr, *colbuf = nullptr;
    unsigned *amax = nullptr;
    int *info = nullptr;
    double *anorm = nullptr;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    if (!f.W && (!ok(salloc(&f.W, tiles * 8, s0)) || !ok(salloc(&f.Z, (size_t)bb * NB * NB * 8, s0)) ||
                 !ok(salloc(&f.dinv, (size_t)Np * 64 * 8, s0))))
        return OOCDAG_ENOMEM;
    const size_t sw = (size_t)mp_super_width(NB, bb);
    // FP32 route: three BF16 pieces (6 B) per element unless variant bit 4
    const size_t lp_bytes = precision > 0 ? 2 * (size_t)bb * NB * sw * NB * (precision == 1 ? 6 : 2) : 8;
    if (!ok(salloc(&lp, lp_bytes, s0)) || !ok(salloc(&scale, (size_t)2 * bb * sw * 4, s0)) ||
        (precision > 0 && !ok(salloc(&colbuf, (size_t)bb * NB * NB * 4, s0))) ||
        !ok(salloc(&amax, (size_t)bb * sw * 4, s0)) || !ok(salloc(&info, sizeof(int), s0)) ||
        !ok(salloc(&anorm, (size_t)Np * 8, s0)))
        return OOCDAG_ENOMEM;
    hipEvent_t w_ready = nullptr;
    if (!ok(hipEventCreateWithFlags(&w_ready, hipEventDisableTiming))) return OOCDAG_EHIP;
    int rc = OOCDAG_OK;
    do {
        std::vector<double> hrow(Np);
        int big = INT_MAX, hinfo = INT_MAX;
        // W <- A and ||A||_inf.  Low precision: the first super-column's tiles
        // (contiguous, first in the packed order) on s0, the rest of the copy
        // and then the row sums on s1, beside that super-column's FP64 work;
        // the first update past it waits for the copy (w_ready), the row sums
        // come back with the final synchronization.
        const int sw0 = std::min(bb, mp_super_width(NB, bb));
        const size_t head =
            precision > 0 ? (size_t)tile_index(sw0, sw0, bb, 1) * NB * NB : tiles;   // tiles of columns < sw0
        if (!ok(hipMemcpyAsync(f.W, dA, std::min(head, tiles) * 8, hipMemcpyDeviceToDevice, s0)) ||
            !ok(hipMemcpyAsync(info, &big, sizeof(int), hipMemcpyHostToDevice, s0))) {
            rc = OOCDAG_EHIP;
            break;
        }
        hipStream_t sn = s0;
        if (head < tiles) {
            if (!ok(hipMemcpyAsync(f.W + head, dA + head, (tiles - head) * 8,
                                   hipMemcpyDeviceToDevice, s1)) ||
                !ok(hipEventRecord(w_ready, s1))) {
                rc = OOCDAG_EHIP;
                break;
            }
            sn = s1;
        }
        hipLaunchKernelGGL(k_tile_rowabs, dim3(NB / 64, bb), dim3(256), 0, sn, dA, NB, bb, anorm);
        int r = factor_mixed(f.W, N, NB, precision, f.dinv, info, lp, scale, amax, colbuf, s0, s1,
                             head < tiles ? w_ready : nullptr);
        if (r) { rc = OOCDAG_EHIP; break; }
        if (!ok(tile_inverses(f.W, N, NB, f.dinv, f.Z, s0))) {
            rc = OOCDAG_EHIP;
            break;
        }
        if (precision > 0) {
            // the correction solves read the factor in FP32: its error is at
            // the low precision's level already, and the sweeps stream half
            // the bytes
            const size_t zn = (size_t)bb * NB * NB;
            if ((!f.W32 && !ok(salloc(&f.W32, tiles * 4, s0))) ||
                (!f.Z32 && !ok(salloc(&f.Z32, zn * 4, s0)))) {
                rc = OOCDAG_ENOMEM;
                break;
            }
            hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, s0, f.W, f.W32, (long long)tiles);
            hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, s0, f.Z, f.Z32, (long long)zn);
        }
        if (!ok(hipGetLastError()) ||
            !ok(hipMemcpyAsync(hrow.data(), anorm, Np * 8, hipMemcpyDeviceToHost, sn)) ||
            !ok(hipStreamSynchronize(sn)) ||
            !ok(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s0)) ||
            !ok(hipStreamSynchronize(s0))) {
            rc = OOCDAG_EHIP;
            break;
        }
        f.Anorm = *std::max_element(hrow.begin(), hrow.end());
        f.info = hinfo == INT_MAX ? 0 : std::min(hinfo, N);
    } while (0);
    (void)hipEventDestroy(w_ready);
    (void)hipFreeAsync(lp, s0);
    (void)hipFreeAsync(scale, s0);
    if (colbuf) (void)hipFreeAsync(colbuf, s0);
    (void)hipFreeAsync(amax, s0);
    (void)hipFreeAsync(info, s0);
    (void)hipFreeAsync(anorm, s0);
    f.factor_seconds = now() - t0;
    return rc;
}

// PDLASCHK's residual (LIN/pdlaschk.f:22-26,186,295) from per-right-hand-side
// norms: max_k ||b_k - A x_k||_inf / (||x_k||_inf * ||A||_inf * eps * N),
// eps = 2^-53 (PDLAMCH('eps')).
double pdlaschk_resid(const double *rnorm, const double *xnorm, int nrhs, double anorm, int N) {
    const double div = anorm * std::ldexp(1.0, -53) * (double)N;
    double r = 0.0;
    for (int k = 0; k < nrhs; ++k) {
        const double q = rnorm[k] == 0.0 ? 0.0 : rnorm[k] / (xnorm[k] * div);
        r = (q > r || !(q == q)) ? q : r;
    }
    return r;
}

int mp_solve(const MpFactor &f, const double *dB, int nrhs, int ldb, double *dX, int ldx,
             double tol, int max_iter, hipStream_t s, int *iters, double *eta_out,
             double *lres_out) {
    const int N = f.N, NB = f.NB, bb = (N + NB - 1) / NB;
    const long long Np = (long long)bb * NB;
    const long long ldt = nrhs + (nrhs & 1);
    double *Bt = nullptr, *Xt = nullptr, *Rt = nullptr, *T = nullptr, *nrm = nullptr;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    const size_t vb = (size_t)(ldt * Np) * 8;
    if (!ok(salloc(&Bt, vb, s)) || !ok(salloc(&Xt, vb, s)) || !ok(salloc(&Rt, vb, s)) ||
        !ok(salloc(&T, vb, s)) || !ok(salloc(&nrm, (size_t)2 * ldt * 8, s)))
        return OOCDAG_ENOMEM;
    int rc = OOCDAG_OK;
    int it = 0;
    double eta = INFINITY, prev = INFINITY;
    // a low-precision factor's solves run in FP32 on its FP32 copy (few
    // right-hand sides): the factor's own error is far above FP32 rounding
    int lp32 = f.precision > 0 && f.W32 && f.Z32 && nrhs <= 64;
    const int rowmap = lp32 ? mfma_rowmap(0, s) : 0;
    float *X32 = nullptr, *T32 = nullptr;
    const size_t vb32 = (size_t)(ldt * Np) * 4;
    if (rowmap < 0 || (lp32 && (!ok(salloc(&X32, vb32, s)) || !ok(salloc(&T32, vb32, s))))) {
        for (void *p : {(void *)Bt, (void *)Xt, (void *)Rt, (void *)T, (void *)nrm, (void *)X32})
            if (p) (void)hipFreeAsync(p, s);
        return rowmap < 0 ? OOCDAG_EHIP : OOCDAG_ENOMEM;
    }
    const long long nel = ldt * Np;
    // x (FP64) <- solve(r), or x += solve(r); cmax: r's column maxima
    auto solve = [&](const double *r, double *x, const double *cmax, double beta) -> bool {
        if (!lp32) {
            if (beta == 0.0) {
                if (x != r && !ok(hipMemcpyAsync(x, r, vb, hipMemcpyDeviceToDevice, s))) return false;
                return ok(potrs_inv_t(f.W, N, NB, f.Z, 0, 0, x, nrhs, ldt, T, s));
            }
            // r is overwritten by its solution, then added to x
            if (!ok(potrs_inv_t(f.W, N, NB, f.Z, 0, 0, (void *)r, nrhs, ldt, T, s))) return false;
            hipLaunchKernelGGL(k_axpy, dim3(1024), dim3(256), 0, s, x, r, nel, 1.0);
            return ok(hipGetLastError());
        }
        hipLaunchKernelGGL(k_to_f32_cols, dim3(1024), dim3(256), 0, s, r, X32, nel, (int)ldt, nrhs, cmax);
        if (!ok(hipGetLastError()) ||
            !ok(potrs_inv_t(f.W32, N, NB, f.Z32, 1, rowmap, X32, nrhs, ldt, T32, s)))
            return false;
        hipLaunchKernelGGL(k_from_f32_cols, dim3(1024), dim3(256), 0, s, x, X32, nel, (int)ldt, nrhs,
                           cmax, beta);
        return ok(hipGetLastError());
    };
    std::vector<double> bnorm(nrhs), hx(nrhs), hr(nrhs);
    do {
        if (!ok(hipMemsetAsync(Bt, 0, vb, s)) || !ok(launch_transpose(dB, ldb, Bt, ldt, N, nrhs, s))) {
            rc = OOCDAG_EHIP;
            break;
        }
        hipLaunchKernelGGL(k_rows_absmax, dim3(nrhs), dim3(256), 0, s, Bt, ldt, nrhs, N, nrm);
        if (!ok(hipMemcpyAsync(bnorm.data(), nrm, nrhs * 8, hipMemcpyDeviceToHost, s)) ||
            !solve(Bt, Xt, nrm, 0.0)) {
            rc = OOCDAG_EHIP;
            break;
        }
        for (;;) {
            if (!ok(hipMemcpyAsync(Rt, Bt, vb, hipMemcpyDeviceToDevice, s)) ||
                !ok(symm_t(f.A, N, NB, Xt, Rt, ldt, nrhs, -1.0, 1.0, s))) {
                rc = OOCDAG_EHIP;
                break;
            }
            hipLaunchKernelGGL(k_rows_absmax, dim3(nrhs), dim3(256), 0, s, Xt, ldt, nrhs, N, nrm);
            hipLaunchKernelGGL(k_rows_absmax, dim3(nrhs), dim3(256), 0, s, Rt, ldt, nrhs, N, nrm + ldt);
            if (!ok(hipMemcpyAsync(hx.data(), nrm, nrhs * 8, hipMemcpyDeviceToHost, s)) ||
                !ok(hipMemcpyAsync(hr.data(), nrm + ldt, nrhs * 8, hipMemcpyDeviceToHost, s)) ||
                !ok(hipStreamSynchronize(s))) {
                rc = OOCDAG_EHIP;
                break;
            }
            prev = eta;
            eta = 0;
            for (int k = 0; k < nrhs; ++k) eta = std::max(eta, hr[k] / (f.Anorm * hx[k] + bnorm[k]));
            if (!(eta == eta)) eta = INFINITY;
            if (lres_out) *lres_out = pdlaschk_resid(hr.data(), hx.data(), nrhs, f.Anorm, N);
            if (eta <= tol) break;
            if (it >= max_iter || !(eta < 0.5 * prev)) {
                rc = 1;  // stalled
                break;
            }
            if (!solve(Rt, Xt, nrm + ldt, 1.0)) {
                rc = OOCDAG_EHIP;
                break;
            }
            ++it;
        }
        if (rc < 0) break;
        if (!ok(launch_transpose(Xt, ldt, dX, ldx, nrhs, N, s)) || !ok(hipStreamSynchronize(s)))
            rc = OOCDAG_EHIP;
    } while (0);
    for (double *p : {Bt, Xt, Rt, T, nrm})
        if (p) (void)hipFreeAsync(p, s);
    for (float *p : {X32, T32})
        if (p) (void)hipFreeAsync(p, s);
    if (iters) *iters = it;
    if (eta_out) *eta_out = eta;
    return rc;
}

int mp_factor_solve(const double *dA, int N, int NB, const double *dB, int nrhs, int ldb,
                    double *dX, int ldx, int precision, double tol, int max_iter, hipStream_t s,
                    hipStream_t s1, oocdag_mpir_stats *st, MpFactor *keep) {
    const double t0 = now();
    MpFactor f;
    int prec = precision, rc = OOCDAG_OK;
    for (int attempt = 0; attempt < 3; ++attempt) {
        rc = mp_factor(f, dA, N, NB, prec, s, s1);
        if (rc) break;
        st->factor_seconds += f.factor_seconds;
        if (f.info && prec == 0) {
            st->info = f.info;
            st->precision_used = 0;
            break;
        }
        int it = 0;
        double eta = INFINITY, lres = INFINITY;
        if (!f.info) {
            const double ts = now();
            rc = mp_solve(f, dB, nrhs, ldb, dX, ldx, tol, max_iter, s, &it, &eta, &lres);
            st->solve_seconds += now() - ts;
            if (rc < 0) break;
        }
        st->iterations += it;
        st->backward_error = eta;
        st->solve_resid = lres;
        if ((!f.info && rc == 0) || prec == 0) {
            st->precision_used = prec;
            rc = OOCDAG_OK;
            break;
        }
        st->fell_back = 1;
        prec = prec == 2 ? 1 : 0;   // BF16 -> FP32 -> FP64
    }
    if (keep && rc == OOCDAG_OK && !st->info)
        *keep = f;
    else
        mp_release(f, s);
    (void)hipStreamSynchronize(s);
    st->total_seconds = now() - t0;
    return rc;
}

namespace {

// Copies the lower triangle of every diagonal slot (edge ts) of packed
// lower tiles to its upper triangle: the matrix is then defined by its lower
// triangle alone.  Writes only strict-upper entries, reads only lower ones.
__global__ void k_symm_diag(double *Y, int ts, int bb) {
    const int t = blockIdx.y;
    double *T = Y + (t + (long long)t * bb - ((long long)t * (t + 1)) / 2) * (long long)ts * ts;
    const long long n = (long long)ts * ts;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(idx % ts), c = (int)(idx / ts);
        if (r < c) T[idx] = T[c + (long long)r * ts];
    }
}

}  // namespace

// PDLASCHK (LIN/pdlaschk.f:22-26,186,295) on the device: R = B - A X with A
// symmetric in packed lower tiles (logical edge NB in slots of tile_edge(NB),
// diagonal tiles full), then the per-right-hand-side infinity norms of R
// and X and ||A||_inf, all reductions in a fixed order.
int solve_resid_dev(const double *dA, int N, int NB, const double *dX, int nrhs, long long ldx,
                    const double *dB, long long ldb, double *resid, hipStream_t s) {
    const int ts = tile_edge(NB), bb = (N + NB - 1) / NB;
    const long long Np = (long long)bb * ts, ldt = nrhs + (nrhs & 1);
    const size_t vb = (size_t)(ldt * Np) * 8;
    double *Xt = nullptr, *Rt = nullptr, *nrm = nullptr, *an = nullptr;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    int rc = OOCDAG_OK;
    do {
        if (!ok(hipMalloc(&Xt, vb)) || !ok(hipMalloc(&Rt, vb)) ||
            !ok(hipMalloc(&nrm, (size_t)2 * ldt * 8)) || !ok(hipMalloc(&an, (size_t)Np * 8))) {
            rc = OOCDAG_ENOMEM;
            break;
        }
        std::vector<double> hx(nrhs), hr(nrhs), ha(Np);
        if (!ok(hipMemsetAsync(Xt, 0, vb, s)) || !ok(hipMemsetAsync(Rt, 0, vb, s)) ||
            !ok(launch_transpose(dX, ldx, Xt, ldt, N, nrhs, s, NB, ts, 1)) ||
            !ok(launch_transpose(dB, ldb, Rt, ldt, N, nrhs, s, NB, ts, 1)) ||
            !ok(symm_t(dA, (int)Np, ts, Xt, Rt, ldt, nrhs, -1.0, 1.0, s))) {
            rc = OOCDAG_EHIP;
            break;
        }
        hipLaunchKernelGGL(k_rows_absmax, dim3(nrhs), dim3(256), 0, s, Xt, ldt, nrhs, (int)Np, nrm);
        hipLaunchKernelGGL(k_rows_absmax, dim3(nrhs), dim3(256), 0, s, Rt, ldt, nrhs, (int)Np,
                           nrm + ldt);
        hipLaunchKernelGGL(k_tile_rowabs, dim3(ts / 64, bb), dim3(256), 0, s, dA, ts, bb, an);
        if (!ok(hipGetLastError()) ||
            !ok(hipMemcpyAsync(hx.data(), nrm, nrhs * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipMemcpyAsync(hr.data(), nrm + ldt, nrhs * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipMemcpyAsync(ha.data(), an, (size_t)Np * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipStreamSynchronize(s))) {
            rc = OOCDAG_EHIP;
            break;
        }
        double anorm = 0.0;   // over the matrix's rows only, not the slot padding
        for (int g = 0; g < N; ++g) anorm = std::max(anorm, ha[(size_t)(g / NB) * ts + g % NB]);
        *resid = pdlaschk_resid(hr.data(), hx.data(), nrhs, anorm, N);
    } while (0);
    for (double *p : {Xt, Rt, nrm, an})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace ooc

// ===========================================================================
// C ABI
// ===========================================================================
#include "../../include/oocdag.h"

using namespace ooc;

extern "C" int oocdag_solve_resid_tiles(const double *dA, int N, int NB, const double *dX,
                                        int nrhs, int ldx, const double *dB, int ldb,
                                        double *resid) {
    if (!dA || !dX || !dB || !resid || N <= 0 || NB <= 0 || nrhs <= 0 || ldx < N || ldb < N)
        return OOCDAG_EARG;
    return solve_resid_dev(dA, N, NB, dX, nrhs, ldx, dB, ldb, resid, nullptr);
}

extern "C" int oocdag_solve_resid(const double *A, int N, int LDA, const double *X, int nrhs,
                                  int LDX, const double *B, int LDB, double *resid) {
    if (!A || !X || !B || !resid || N <= 0 || nrhs <= 0 || LDA < N || LDX < N || LDB < N)
        return OOCDAG_EARG;
    const int NB = N >= 4096 ? 512 : 128;
    const int bb = (N + NB - 1) / NB;
    double *dY = nullptr, *dX = nullptr, *dB = nullptr;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    int rc = OOCDAG_OK;
    do {
        if (!ok(hipMalloc(&dY, oocdag_tiles_elems(N, NB) * 8)) ||
            !ok(hipMalloc(&dX, (size_t)N * nrhs * 8)) || !ok(hipMalloc(&dB, (size_t)N * nrhs * 8))) {
            rc = OOCDAG_ENOMEM;
            break;
        }
        if ((rc = oocdag_tiles_from_host(dY, A, N, NB, LDA))) break;
        hipLaunchKernelGGL(k_symm_diag, dim3(64, bb), dim3(256), 0, 0, dY, NB, bb);
        if (!ok(hipGetLastError()) ||
            !ok(hipMemcpy2D(dX, (size_t)N * 8, X, (size_t)LDX * 8, (size_t)N * 8, nrhs,
                            hipMemcpyHostToDevice)) ||
            !ok(hipMemcpy2D(dB, (size_t)N * 8, B, (size_t)LDB * 8, (size_t)N * 8, nrhs,
                            hipMemcpyHostToDevice))) {
            rc = OOCDAG_EHIP;
            break;
        }
        rc = solve_resid_dev(dY, N, NB, dX, nrhs, N, dB, N, resid, nullptr);
    } while (0);
    for (double *p : {dY, dX, dB})
        if (p) (void)hipFree(p);
    return rc;
}

extern "C" int oocdag_gram_tiles(double *dY, int N, int NB, double mu, unsigned long long seed,
                                 void *stream) {
    if (!dY || N <= 0 || NB < 64 || NB % 64) return OOCDAG_EARG;
    return gram_tiles(dY, N, NB, mu, seed, (hipStream_t)stream) == hipSuccess ? OOCDAG_OK
                                                                              : OOCDAG_EHIP;
}

extern "C" int oocdag_normal_fill(double *dB, int rows, int cols, int ld, unsigned long long seed,
                                  unsigned long long stream_id) {
    if (!dB || rows < 0 || cols < 0 || ld < rows) return OOCDAG_EARG;
    if (normal_block(dB, rows, cols, ld, seed, stream_id, 0) != hipSuccess) return OOCDAG_EHIP;
    return hipDeviceSynchronize() == hipSuccess ? OOCDAG_OK : OOCDAG_EHIP;
}

extern "C" int oocdag_mpir_solve_tiles(const double *dA, int N, int NB, const double *dB, int nrhs,
                                       int ldb, double *dX, int ldx, int precision, double tol,
                                       int max_iter, oocdag_mpir_stats *st) {
    oocdag_mpir_stats dummy;
    if (!st) st = &dummy;
    *st = oocdag_mpir_stats{};
    if (!dA || !dB || !dX || N <= 0 || nrhs <= 0 || ldb < N || ldx < N || NB < 128 || NB % 128 ||
        precision < 0 || precision > 2)
        return OOCDAG_EARG;
    hipStream_t s = nullptr, s1 = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return OOCDAG_EHIP;
    if (hipStreamCreateWithFlags(&s1, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return OOCDAG_EHIP;
    }
    int rc = mp_factor_solve(dA, N, NB, dB, nrhs, ldb, dX, ldx, precision, tol, max_iter, s, s1, st,
                             nullptr);
    (void)hipStreamDestroy(s1);
    (void)hipStreamDestroy(s);
    return rc;
}
