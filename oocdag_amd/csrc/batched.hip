// Batched small SPD factor + solve in one persistent kernel (no reference
// counterpart: the reference factors one large matrix).
//
// Each workgroup (256 threads) pulls system ids from an atomic work queue
// (systems pre-sorted largest first) and runs the whole system inside the
// workgroup:
//   1. W = A (FT = double or float working copy, column-major, ld n),
//   2. right-looking blocked Cholesky with block T: the T x T diagonal block
//      is factored and inverted by one wave in LDS; the panel below is
//      solved row by row against the inverse and staged in LDS; the trailing
//      lower triangle is updated with RB x RB register blocks reading the
//      staged panel,
//   3. x = W^-1 b by blocked forward/backward substitution with the stored
//      diagonal-block inverses, right-hand sides resident in LDS,
//   4. FP64 residual r = b - A x, normwise backward error
//      max_k ||r_k|| / (||A|| ||x_k|| + ||b_k||), and (always for FT = float,
//      else if above tol) refinement x += W^-1 r until the error is <= tol.
// Systems whose factor is not positive definite or whose refinement stalls
// report status != 0 so the caller can rerun them in FP64.
//
// A left-looking variant (64-wide block columns, MFMA updates, two
// workgroups per CU) covers the same steps for larger systems; its phases
// also run as block-column launches across all the largest systems at once
// (a few large systems then use the whole GPU instead of one CU each).
//
// Variants (T, RB, FT) are compiled ahead; the host autotuner times them on
// a sample of the batch, and the block-column threshold on the whole batch,
// and launches the fastest through a captured hipGraph.
#include "ooc_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/oocdag.h"

namespace ooc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NTB = 256;
constexpr int PMAX = 1024;        // max staged panel rows
constexpr int NMAX = PMAX + 64;   // max system order
constexpr int RMAX = 8;           // right-hand sides resident at a time
constexpr int RTOTAL = 64;        // max right-hand sides per system
constexpr double kEps = 1.1102230246251565e-16;   // 2^-53, PDLAMCH('eps')

struct SysDesc {
    int n, nrhs, ld;       // ld: even leading dimension >= n of A, B, X and W
    long long a, b, w, d;  // offsets: A/B/X (doubles), W and inverses (FT elements)
};

struct BatchArgs {
    const SysDesc *sys;
    const int *order;
    int nsys;
    const double *A, *B;
    double *X;
    void *W, *Dinv;
    double *berr;
    double *lres;   // PDLASCHK residual per system (LIN/pdlaschk.f:186,295)
    int *iters, *status;
    int *counter;
    double tol;
    int max_iter;
    int rowmap;   // FP32 MFMA result rows: 0: 4*(lane/16) + r, 1: lane/16 + 4r
    int prefact;  // leading queue entries already factored (k_batched_ll solves them only)
    int diag_k0;  // the diagonal-block step's column (k_batched_ll<FT, false>)
};

template <typename FT, int T>
struct Smem {
    // panel (factorization) and solve vectors share storage
    union {
        FT P[PMAX * (T + 1)];
        struct {
            double y[RMAX * NMAX];
            double x[RMAX * NMAX];
        } v;
    } u;
    FT D[T * (T + 1)];
    FT Di[T * (T + 1)];
    double red[NTB];
    double tmp[RMAX * 64];
    int bad;
    int sid;
};

// Orders LDS traffic between the lanes of one wave (the wave runs in lock
// step; this only stops the compiler from moving LDS accesses across it).
// lane l's value of v, broadcast to the whole wave (two v_readlane_b32)
__device__ __forceinline__ double bcast_lane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void tri_lower_decode(int idx, int &bi, int &bj) {
    int r = (int)((sqrtf(8.0f * idx + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= idx) ++r;
    while (r * (r + 1) / 2 > idx) --r;
    bi = r;
    bj = idx - r * (r + 1) / 2;
}

// x <- W^-1 y for all right-hand sides held in S.u.v.y (overwritten with x)
template <typename FT, int T>
__device__ void tri_solve(Smem<FT, T> &S, const FT *W, const FT *Dinv, int n, int ld, int nrhs) {
    const int tid = threadIdx.x;
    double *y = S.u.v.y;
    const int nbk = (n + T - 1) / T;
    // forward: y_blk = Di y_blk ; y[rest] -= L[rest, blk] y_blk
    for (int kb = 0; kb < nbk; ++kb) {
        const int k0 = kb * T, w = min(T, n - k0);
        const FT *Di = Dinv + (long long)kb * T * T;
        if (tid < w * nrhs) {
            int c = tid % w, k = tid / w;
            double s = 0.0;
            for (int d = 0; d <= c; ++d) s += (double)Di[c + d * T] * y[k * NMAX + k0 + d];
            S.tmp[k * 64 + c] = s;
        }
        __syncthreads();
        if (tid < w * nrhs) {
            int c = tid % w, k = tid / w;
            y[k * NMAX + k0 + c] = S.tmp[k * 64 + c];
        }
        __syncthreads();
        for (int idx = tid; idx < (n - k0 - w) * nrhs; idx += NTB) {
            int i = k0 + w + idx % (n - k0 - w), k = idx / (n - k0 - w);
            double s = 0.0;
            for (int c = 0; c < w; ++c) s += (double)W[i + (long long)(k0 + c) * ld] * y[k * NMAX + k0 + c];
            y[k * NMAX + i] -= s;
        }
        __syncthreads();
    }
    // backward: x_blk = Di^T y_blk ; y[above] -= L[blk, above]^T x_blk
    for (int kb = nbk - 1; kb >= 0; --kb) {
        const int k0 = kb * T, w = min(T, n - k0);
        const FT *Di = Dinv + (long long)kb * T * T;
        if (tid < w * nrhs) {
            int c = tid % w, k = tid / w;
            double s = 0.0;
            for (int d = c; d < w; ++d) s += (double)Di[d + c * T] * y[k * NMAX + k0 + d];
            S.tmp[k * 64 + c] = s;
        }
        __syncthreads();
        if (tid < w * nrhs) {
            int c = tid % w, k = tid / w;
            y[k * NMAX + k0 + c] = S.tmp[k * 64 + c];
        }
        __syncthreads();
        for (int idx = tid; idx < k0 * nrhs; idx += NTB) {
            int i = idx % k0, k = idx / k0;
            double s = 0.0;
            for (int r = 0; r < w; ++r) s += (double)W[(k0 + r) + (long long)i * ld] * y[k * NMAX + k0 + r];
            y[k * NMAX + i] -= s;
        }
        __syncthreads();
    }
}

template <typename FT, int T, int RB>
__global__ __launch_bounds__(NTB, 1) void k_batched(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem<FT, T> &S = *reinterpret_cast<Smem<FT, T> *>(smem_raw);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (;;) {
        if (tid == 0) S.sid = atomicAdd(a.counter, 1);
        __syncthreads();
        const int q = S.sid;
        __syncthreads();
        if (q >= a.nsys) break;
        const int sidx = a.order[q];
        const SysDesc sd = a.sys[sidx];
        const int n = sd.n, nrhs_all = sd.nrhs, ld = sd.ld;
        const double *A = a.A + sd.a;
        const double *B0 = a.B + sd.b;
        double *X0 = a.X + sd.b;
        FT *W = (FT *)a.W + sd.w;
        FT *Dinv = (FT *)a.Dinv + sd.d;
        if (n > PMAX || nrhs_all > RTOTAL) {
            if (tid == 0) a.status[sidx] = -1;
            continue;
        }
        if (tid == 0) S.bad = 0;
        // 1. working copy (lower triangle)
        for (long long idx = tid; idx < (long long)n * n; idx += NTB) {
            int r = (int)(idx % n), c = (int)(idx / n);
            if (r >= c) W[r + (long long)c * ld] = (FT)A[r + (long long)c * ld];
        }
        __syncthreads();
        // 2. blocked right-looking Cholesky
        for (int k0 = 0; k0 < n; k0 += T) {
            const int w = min(T, n - k0), m = n - k0 - w;
            for (int idx = tid; idx < T * T; idx += NTB) {
                int r = idx % T, c = idx / T;
                S.D[c * (T + 1) + r] = (r < w && c < w && r >= c) ? W[(k0 + r) + (long long)(k0 + c) * ld]
                                                                   : (FT)(r == c ? 1 : 0);
            }
            __syncthreads();
            if (wave == 0) {
                // factor (lane = row) then invert (lane = column), lanes >= T idle
                for (int j = 0; j < T; ++j) {
                    FT v = 0;
                    if (lane >= j && lane < T) {
                        v = S.D[j * (T + 1) + lane];
                        for (int k = 0; k < j; ++k) v -= S.D[k * (T + 1) + lane] * S.D[k * (T + 1) + j];
                    }
                    if (lane == j) {
                        if (!(v > (FT)0)) S.bad = k0 + j + 1;
                        S.D[j * (T + 1) + j] = sqrt(v);
                    }
                    wave_sync();
                    if (lane > j && lane < T) S.D[j * (T + 1) + lane] = v / S.D[j * (T + 1) + j];
                    wave_sync();
                }
                if (lane < T) {
                    const int col = lane;
                    for (int i = 0; i < col; ++i) S.Di[col * (T + 1) + i] = 0;
                    S.Di[col * (T + 1) + col] = (FT)1 / S.D[col * (T + 1) + col];
                    for (int i = col + 1; i < T; ++i) {
                        FT sum = 0;
                        for (int d = col; d < i; ++d) sum += S.D[d * (T + 1) + i] * S.Di[col * (T + 1) + d];
                        S.Di[col * (T + 1) + i] = -sum / S.D[i * (T + 1) + i];
                    }
                }
            }
            __syncthreads();
            if (S.bad) break;
            for (int idx = tid; idx < T * T; idx += NTB) {
                int r = idx % T, c = idx / T;
                if (r < w && c < w && r >= c) W[(k0 + r) + (long long)(k0 + c) * ld] = S.D[c * (T + 1) + r];
                Dinv[(long long)(k0 / T) * T * T + idx] = S.Di[c * (T + 1) + r];
            }
            // panel: P[r] = W[k0+w+r, k0:k0+w] * Di^T, staged in LDS
            for (int r = tid; r < m; r += NTB) {
                FT row[T];
#pragma unroll
                for (int c = 0; c < T; ++c) row[c] = (c < w) ? W[(k0 + w + r) + (long long)(k0 + c) * ld] : (FT)0;
#pragma unroll
                for (int c = 0; c < T; ++c) {
                    FT s = 0;
#pragma unroll
                    for (int d = 0; d <= c; ++d) s += row[d] * S.Di[d * (T + 1) + c];  // Di(c, d)
                    S.u.P[r * (T + 1) + c] = s;
                    if (c < w) W[(k0 + w + r) + (long long)(k0 + c) * ld] = s;
                }
            }
            __syncthreads();
            // trailing lower update with RB x RB register blocks
            const int nbr = (m + RB - 1) / RB;
            const int total = nbr * (nbr + 1) / 2;
            for (int idx = tid; idx < total; idx += NTB) {
                int bi, bj;
                tri_lower_decode(idx, bi, bj);
                FT acc[RB][RB];
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < RB; ++j) acc[i][j] = 0;
                for (int k = 0; k < w; ++k) {
                    FT pa[RB], pb[RB];
#pragma unroll
                    for (int i = 0; i < RB; ++i) {
                        int r = bi * RB + i;
                        pa[i] = r < m ? S.u.P[r * (T + 1) + k] : (FT)0;
                    }
#pragma unroll
                    for (int j = 0; j < RB; ++j) {
                        int c = bj * RB + j;
                        pb[j] = c < m ? S.u.P[c * (T + 1) + k] : (FT)0;
                    }
#pragma unroll
                    for (int i = 0; i < RB; ++i)
#pragma unroll
                        for (int j = 0; j < RB; ++j) acc[i][j] += pa[i] * pb[j];
                }
#pragma unroll
                for (int j = 0; j < RB; ++j) {
                    int c = bj * RB + j;
                    if (c >= m) continue;
                    FT *col = W + (long long)(k0 + w + c) * ld + (k0 + w);
#pragma unroll
                    for (int i = 0; i < RB; ++i) {
                        int r = bi * RB + i;
                        if (r < m && r >= c) col[r] -= acc[i][j];
                    }
                }
            }
            __syncthreads();
        }
        if (S.bad) {
            if (tid == 0) {
                a.status[sidx] = S.bad;
                a.berr[sidx] = INFINITY;
                a.iters[sidx] = 0;
            }
            __syncthreads();
            continue;
        }
        // 3./4. solve and refine, RMAX right-hand sides at a time
        // (x in S.u.v.x, work vector in S.u.v.y)
        double eta_all = 0.0, lres_all = 0.0;
        int it_all = 0;
        for (int r0 = 0; r0 < nrhs_all; r0 += RMAX) {
        const int nrhs = min(RMAX, nrhs_all - r0);
        const double *B = B0 + (long long)r0 * ld;
        double *X = X0 + (long long)r0 * ld;
        double *y = S.u.v.y, *x = S.u.v.x;
        for (int idx = tid; idx < n * nrhs; idx += NTB) {
            int i = idx % n, k = idx / n;
            y[k * NMAX + i] = B[i + (long long)k * ld];
            x[k * NMAX + i] = 0.0;
        }
        __syncthreads();
        int it = 0;
        double eta = INFINITY, prev = INFINITY, lres = INFINITY;
        for (;;) {
            tri_solve<FT, T>(S, W, Dinv, n, ld, nrhs);
            for (int idx = tid; idx < n * nrhs; idx += NTB) {
                int i = idx % n, k = idx / n;
                x[k * NMAX + i] += y[k * NMAX + i];
            }
            __syncthreads();
            // r = b - A x (FP64), row sums of |A|, norms
            double rmax[RMAX], xmax[RMAX], bmax[RMAX];
#pragma unroll
            for (int k = 0; k < RMAX; ++k) rmax[k] = xmax[k] = bmax[k] = 0.0;
            double amax = 0.0;
            for (int i = tid; i < n; i += NTB) {
                double rs[RMAX], as = 0.0;
#pragma unroll
                for (int k = 0; k < RMAX; ++k) rs[k] = (k < nrhs) ? B[i + (long long)k * ld] : 0.0;
                for (int j = 0; j < n; ++j) {
                    const double aij = A[i + (long long)j * ld];
                    as += fabs(aij);
#pragma unroll
                    for (int k = 0; k < RMAX; ++k)
                        if (k < nrhs) rs[k] -= aij * x[k * NMAX + j];
                }
                amax = fmax(amax, as);
#pragma unroll
                for (int k = 0; k < RMAX; ++k) {
                    if (k >= nrhs) break;
                    y[k * NMAX + i] = rs[k];
                    rmax[k] = fmax(rmax[k], fabs(rs[k]));
                    xmax[k] = fmax(xmax[k], fabs(x[k * NMAX + i]));
                    bmax[k] = fmax(bmax[k], fabs(B[i + (long long)k * ld]));
                }
            }
            // block max-reductions: amax, then per rhs the three norms
            double e = 0.0, lr = 0.0;
            S.red[tid] = amax;
            __syncthreads();
            for (int s2 = NTB / 2; s2 > 0; s2 >>= 1) {
                if (tid < s2) S.red[tid] = fmax(S.red[tid], S.red[tid + s2]);
                __syncthreads();
            }
            const double An = S.red[0];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RMAX; ++k) {
                if (k >= nrhs) break;
                double vals[3] = {rmax[k], xmax[k], bmax[k]};
                double out[3];
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    S.red[tid] = vals[v];
                    __syncthreads();
                    for (int s2 = NTB / 2; s2 > 0; s2 >>= 1) {
                        if (tid < s2) S.red[tid] = fmax(S.red[tid], S.red[tid + s2]);
                        __syncthreads();
                    }
                    out[v] = S.red[0];
                    __syncthreads();
                }
                e = fmax(e, out[0] / (An * out[1] + out[2]));
                    lr = fmax(lr, out[0] == 0.0 ? 0.0 : out[0] / (An * out[1] * kEps * n));
            }
            prev = eta;
            eta = e;
            lres = lr;
            if (eta <= a.tol) break;
            if (it >= a.max_iter || !(eta < 0.5 * prev)) break;
            ++it;  // y holds r: next pass solves W d = r and adds d to x
        }
        for (int idx = tid; idx < n * nrhs; idx += NTB) {
            int i = idx % n, k = idx / n;
            X[i + (long long)k * ld] = x[k * NMAX + i];
        }
        eta_all = fmax(eta_all, eta);
        lres_all = fmax(lres_all, lres);
        it_all = max(it_all, it);
        __syncthreads();
        }
        if (tid == 0) {
            a.berr[sidx] = eta_all;
            a.lres[sidx] = lres_all;
            a.iters[sidx] = it_all;
            a.status[sidx] = (eta_all <= a.tol) ? 0 : -2;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Left-looking variant for larger systems.  Block columns of LT = 64:
//   U. C_j = A[k0:, k0:k0+64] - L[k0:, :k0] L[k0:k0+64, :k0]^T as one MFMA
//      GEMM per 128-row tile (4 waves as 2 x 2, operands streamed from
//      global memory straight into MFMA registers with a ring of loads in
//      flight; every element of the trailing matrix is read and written
//      once over the whole factorization, instead of once per block step as
//      in the right-looking kernel),
//   D. the 64 x 64 diagonal block of C_j factored in LDS (16-column sub-
//      blocks: one wave factors, all threads solve and update) and inverted
//      by 16 x 16 blocks (diagonal blocks first, then one block row of
//      products per barrier),
//   T. the rows below solved as X = C_j Di^T with MFMA (Di from LDS),
// then blocked substitution with one thread per row (each element of L read
// once per pass, y in LDS, x in the caller's X), FP64 residual staged through
// LDS and refinement, 8 right-hand sides per pass.
// ---------------------------------------------------------------------------
constexpr int LT = 64;
constexpr int LS = LT + 1;
constexpr int LRT = 128;       // rows per update tile
constexpr int LRMAX = 8;

template <typename FT>
struct SmemLL {
    union {
        struct {
            double D[LT * LS];
            double Di[LT * LS];
            double T[3 * 256];   // blocked inverse: one block row of products
        } f;
        struct {
            double y[LRMAX * NMAX];   // x lives in the caller's X (global)
        } v;
    } u;
    double red[NTB];
    double tmp[LRMAX * LT];
    int bad;
    int sid;
};

template <typename FT> struct AccT;
template <> struct AccT<double> { typedef d4 type; };
template <> struct AccT<float> { typedef f4 type; };

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// row of accumulator element r of this lane inside its 16 x 16 block
template <typename FT>
__device__ __forceinline__ int acc_row(int lane, int r, int rowmap) {
    if (sizeof(FT) == 8) return (lane >> 4) + 4 * r;   // FP64 MFMA
    return rowmap ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

// x <- W^-1 y with 64-wide blocks and the stored block inverses
template <typename FT>
__device__ void tri_solve_ll(SmemLL<FT> &S, const FT *W, const FT *Dinv, int n, int ld, int nrhs) {
    const int tid = threadIdx.x;
    double *y = S.u.v.y;
    const int nbk = (n + LT - 1) / LT;
    for (int kb = 0; kb < nbk; ++kb) {
        const int k0 = kb * LT, w = min(LT, n - k0);
        const FT *Di = Dinv + (long long)kb * LT * LT;   // Di(c, d) at c + d*LT
        for (int idx = tid; idx < w * nrhs; idx += NTB) {
            const int c = idx % w, k = idx / w;
            const double *yk = y + k * NMAX + k0;
            double s = 0.0;
            int d = 0;
            for (; d + 8 <= c + 1; d += 8) {   // 8 independent loads, then the products
                double dv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) dv[u] = (double)Di[c + (d + u) * LT];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += dv[u] * yk[d + u];
            }
            for (; d <= c; ++d) s += (double)Di[c + d * LT] * yk[d];
            S.tmp[k * LT + c] = s;
        }
        __syncthreads();
        // the block's new values (S.tmp) go to y while the rows below are
        // updated from S.tmp: disjoint parts of y, one barrier
        for (int idx = tid; idx < w * nrhs; idx += NTB) {
            const int c = idx % w, k = idx / w;
            y[k * NMAX + k0 + c] = S.tmp[k * LT + c];
        }
        // rows below: one thread per row, all right-hand sides (L read once)
        for (int i = k0 + w + tid; i < n; i += NTB) {
            double sk[LRMAX];
#pragma unroll
            for (int k = 0; k < LRMAX; ++k) sk[k] = 0.0;
            int c = 0;
            for (; c + 16 <= w; c += 16) {   // 16 independent loads, then the products
                double wv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) wv[u] = (double)W[i + (long long)(k0 + c + u) * ld];
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int k = 0; k < LRMAX; ++k)
                        if (k < nrhs) sk[k] += wv[u] * S.tmp[k * LT + c + u];
            }
            for (; c < w; ++c) {
                const double wv = (double)W[i + (long long)(k0 + c) * ld];
#pragma unroll
                for (int k = 0; k < LRMAX; ++k)
                    if (k < nrhs) sk[k] += wv * S.tmp[k * LT + c];
            }
#pragma unroll
            for (int k = 0; k < LRMAX; ++k)
                if (k < nrhs) y[k * NMAX + i] -= sk[k];
        }
        __syncthreads();
    }
    for (int kb = nbk - 1; kb >= 0; --kb) {
        const int k0 = kb * LT, w = min(LT, n - k0);
        const FT *Di = Dinv + (long long)kb * LT * LT;
        for (int idx = tid; idx < w * nrhs; idx += NTB) {
            const int c = idx % w, k = idx / w;
            const double *yk = y + k * NMAX + k0;
            const FT *dc = Di + c * LT;   // Di(d, c), d = c..w-1, contiguous
            double s = 0.0;
            int d = c;
            for (; d + 8 <= w; d += 8) {
                double dv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) dv[u] = (double)dc[d + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += dv[u] * yk[d + u];
            }
            for (; d < w; ++d) s += (double)dc[d] * yk[d];
            S.tmp[k * LT + c] = s;
        }
        __syncthreads();
        for (int idx = tid; idx < w * nrhs; idx += NTB) {
            const int c = idx % w, k = idx / w;
            y[k * NMAX + k0 + c] = S.tmp[k * LT + c];
        }
        for (int i = tid; i < k0; i += NTB) {
            const FT *col = W + (long long)i * ld + k0;   // L(k0.., i), contiguous
            double sk[LRMAX];
#pragma unroll
            for (int k = 0; k < LRMAX; ++k) sk[k] = 0.0;
            int r = 0;
            for (; r + 16 <= w; r += 16) {
                double wv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) wv[u] = (double)col[r + u];
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int k = 0; k < LRMAX; ++k)
                        if (k < nrhs) sk[k] += wv[u] * S.tmp[k * LT + r + u];
            }
            for (; r < w; ++r) {
                const double wv = (double)col[r];
#pragma unroll
                for (int k = 0; k < LRMAX; ++k)
                    if (k < nrhs) sk[k] += wv * S.tmp[k * LT + r];
            }
#pragma unroll
            for (int k = 0; k < LRMAX; ++k)
                if (k < nrhs) y[k * NMAX + i] -= sk[k];
        }
        __syncthreads();
    }
}

// The 16 x 16 diagonal sub-block at (s0, s0) of the LDS block D (column
// stride LS; rows past sw are padding) factored in registers by one wave:
// lane r holds row s0 + r (padded with an identity), right-looking, the
// pivot and column j read from the owning lanes (v_readlane, no LDS round
// trips): the same subtraction order per element as a left-looking dot
// product, the column scaled by the reciprocal pivot as LAPACK dpotf2 does.
// The reciprocal pivots are kept, so the inverse needs no divisions: lane c
// forms column c of the factor's inverse by forward substitution into Li
// (Li(i, c) at i + 16c) and into the diagonal block of the 64 x 64 inverse
// Di (column stride LS).  (Reading the factor back from LDS instead of
// v_readlane made the compiler keep the loads live: 256 VGPRs and a stack
// of ~700 bytes per lane.)  Returns the
// first bad pivot (1-based inside the sub-block) or 0.  Kept out of line:
// its 48 live doubles must not add to the register pressure of the
// caller's other phases.  The blocks come as LDS (address space 3)
// pointers: from a generic pointer into the kernel's dynamic LDS an
// out-of-line function finds the block's address through a per-kernel
// offset table, one scalar memory load per access (~4x the factor's time).
typedef __attribute__((address_space(3))) double lds_double;

__device__ __forceinline__ lds_double *to_lds(double *p) { return (lds_double *)p; }

__device__ __noinline__ int factor16(lds_double *D, int s0, int sw, lds_double *Li, lds_double *Di,
                                     int lane) {
    double av[16], rdv[16];
#pragma unroll
    for (int c = 0; c < 16; ++c)
        av[c] = (lane < sw && c < sw && c <= lane) ? D[(s0 + c) * LS + s0 + lane]
                                                   : (c == lane ? 1.0 : 0.0);
    int bad = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double d = bcast_lane(av[j], j);
        if (bad == 0 && !(d > 0.0)) bad = j + 1;
        const double sd = sqrt(d), rd = 1.0 / sd;
        rdv[j] = rd;
        av[j] = lane > j ? av[j] * rd : (lane == j ? sd : 0.0);
#pragma unroll
        for (int k = j + 1; k < 16; ++k) av[k] -= av[j] * bcast_lane(av[j], k);
    }
    double yv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) yv[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        yv[k] *= rdv[k];
#pragma unroll
        for (int i = k + 1; i < 16; ++i) yv[i] -= bcast_lane(av[k], i) * yv[k];
    }
    if (lane < sw) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c <= lane) D[(s0 + c) * LS + s0 + lane] = av[c];
    }
    if (lane < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) Li[i + 16 * lane] = yv[i];
    }
    if (lane < sw) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < sw) Di[(s0 + lane) * LS + s0 + i] = yv[i];
    }
    return bad;
}

// ---- U: rows r0 .. r0 + 127 (one 128-row tile) of block column
// [k0, k0 + w): C = A[:, k0:k0+w] - L[:, :k0] L[k0:k0+w, :k0]^T, the lower
// part only in the diagonal block; 4 waves as 2 x 2, each 64 rows x 32
// columns.  With to_d the diagonal block (waves wm = 0 of the tile at
// r0 = k0) goes straight to the LDS block D of the factorization (as the FT
// value W would hold) instead of W.
template <typename FT>
__device__ __forceinline__ void ll_update_tile(const double *A, FT *W, int n, int ld, int k0, int w,
                                               int r0, bool to_d, double *D, int rowmap) {
    typedef typename AccT<FT>::type acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane & 15, kl = lane >> 4;
    acc_t acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    // Operands straight from global memory into MFMA registers, no LDS and
    // no barriers: lane (il, kl) of wave (wm, wn) feeds MFMA block i with
    // row r0 + wm*64 + i*16 + il of L and block j with row
    // k0 + wn*32 + j*16 + il, at column k + kl of the current 4-wide K slab.
    // A ring of LRING slabs keeps LAHEAD slabs of loads in flight ahead of
    // the MFMAs; rows past n are clamped to row n-1 (in bounds; their
    // products land in rows the epilogue does not write) and past the last
    // slab the loads repeat it (never used).
    if (k0 > 0) {
        constexpr int LRING = 4, LAHEAD = 3;
        int ra_[4], rb_[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) ra_[i] = min(r0 + wm * 64 + i * 16 + il, n - 1);
#pragma unroll
        for (int j = 0; j < 2; ++j) rb_[j] = min(k0 + wn * 32 + j * 16 + il, n - 1);
        const FT *colp = W + (long long)kl * ld;   // column kl of the current slab
        const long long step = 4LL * ld;
        const int ns = k0 >> 2;                     // slabs (a multiple of 16)
        FT va[LRING][4], vb[LRING][2];
        auto ld_slab = [&](FT (&xa)[4], FT (&xb)[2]) {
#pragma unroll
            for (int i = 0; i < 4; ++i) xa[i] = colp[ra_[i]];
#pragma unroll
            for (int j = 0; j < 2; ++j) xb[j] = colp[rb_[j]];
        };
#pragma unroll
        for (int q = 0; q < LAHEAD; ++q) {
            ld_slab(va[q], vb[q]);
            colp += step;
            __builtin_amdgcn_sched_barrier(0);
        }
        for (int s0 = 0; s0 < ns; s0 += LRING) {
#pragma unroll
            for (int q = 0; q < LRING; ++q) {
                const int nx = (q + LAHEAD) & (LRING - 1);
                ld_slab(va[nx], vb[nx]);
                colp += (s0 + q + LAHEAD + 1 < ns) ? step : 0;
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = mfma4(va[q][i], vb[q][j], acc[i][j]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // C = A - acc (lower part only in the diagonal block).  All loads of A
    // first: a store to W may alias A for the compiler, so interleaved they
    // would be 32 round trips.
    double aval[4][2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = wn * 32 + j * 16 + il;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = wm * 64 + i * 16 + acc_row<FT>(lane, r, rowmap);
                const int gr = r0 + rr;
                const bool ok = gr < n && c < w && gr >= k0 + c;
                aval[i][j][r] = ok ? A[gr + (long long)(k0 + c) * ld] : 0.0;
            }
        }
    const bool dd = to_d && wm == 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = wn * 32 + j * 16 + il;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = wm * 64 + i * 16 + acc_row<FT>(lane, r, rowmap);
                const int gr = r0 + rr;
                if (gr >= n || c >= w || gr < k0 + c) continue;
                const FT v = (FT)(aval[i][j][r] - (double)acc[i][j][r]);
                if (dd)
                    D[c * LS + rr] = (double)v;
                else
                    W[gr + (long long)(k0 + c) * ld] = v;
            }
        }
}

// ---- T: rows r0 .. r0 + 63 below the diagonal block of block column k0
// (64 wide), X = C Di^T on MFMA.  Di(c, d) comes from the LDS block
// (DiL[d * LS + c], double) or from the block's slot of the inverse store
// (DiG[c + d * LT], FT): the same FT operand values either way.
template <typename FT, bool DI_GLOBAL>
__device__ __forceinline__ void ll_trsm_tile(FT *W, const double *DiL, const FT *DiG, int n, int ld,
                                             int k0, int r0, int rowmap) {
    typedef typename AccT<FT>::type acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane & 15, kl = lane >> 4;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    int rows[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) rows[i] = min(r0 + wm * 32 + i * 16 + il, n - 1);
    // all 32 loads of the tile's rows first (one round trip), then the
    // MFMAs with Di
    FT fa[LT / 4][2];
#pragma unroll
    for (int ks = 0; ks < LT / 4; ++ks)
#pragma unroll
        for (int i = 0; i < 2; ++i) fa[ks][i] = W[rows[i] + (long long)(k0 + ks * 4 + kl) * ld];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < LT / 4; ++ks) {
        const int kk = ks * 4 + kl;
        FT fb[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = wn * 32 + j * 16 + il;
            fb[j] = DI_GLOBAL ? DiG[c + kk * LT] : (FT)DiL[kk * LS + c];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma4(fa[ks][i], fb[j], acc[i][j]);
    }
    __syncthreads();   // every wave has read its rows before any is overwritten
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = wn * 32 + j * 16 + il;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = r0 + wm * 32 + i * 16 + acc_row<FT>(lane, r, rowmap);
                if (gr >= n) continue;
                W[gr + (long long)(k0 + c) * ld] = (FT)acc[i][j][r];
            }
        }
    __syncthreads();
}

// FULL: the persistent kernel.  !FULL: the diagonal-block step of the
// block-column launches (k_bl_update, k_bl_trsm below): workgroup b factors
// and inverts the diagonal block at column a.diag_k0 of system a.order[b],
// whose update is already in W, records a bad pivot in a.status and returns
// -- the same D phase code as the persistent kernel's.
template <typename FT, bool FULL = true>
__global__ __launch_bounds__(NTB, 2) void k_batched_ll(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemLL<FT> &S = *reinterpret_cast<SmemLL<FT> *>(smem_raw);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (;;) {
        int q = blockIdx.x;
        if (FULL) {
            if (tid == 0) S.sid = atomicAdd(a.counter, 1);
            __syncthreads();
            q = S.sid;
            __syncthreads();
            if (q >= a.nsys) break;
        } else if (a.status[a.order[q]] != 0) {
            return;
        }
        const int sidx = a.order[q];
        const SysDesc sd = a.sys[sidx];
        const int n = sd.n, nrhs_all = sd.nrhs, ld = sd.ld;
        const double *A = a.A + sd.a;
        const double *B0 = a.B + sd.b;
        double *X0 = a.X + sd.b;
        FT *W = (FT *)a.W + sd.w;
        FT *Dinv = (FT *)a.Dinv + sd.d;
        if (n > PMAX || nrhs_all > RTOTAL) {
            if (tid == 0) a.status[sidx] = -1;
            if (!FULL) return;   // one system per workgroup: never loop back
            continue;
        }
        // the first a.prefact queue entries were factored by the block-column
        // launches: their status holds 0 or the first bad pivot
        const bool factored = FULL && q < a.prefact;
        if (tid == 0) S.bad = factored ? a.status[sidx] : 0;
        const int jb0 = FULL ? 0 : a.diag_k0 / LT;
        const int jb1 = FULL ? (factored ? 0 : (n + LT - 1) / LT) : jb0 + 1;
        if (!FULL) {
            // the updated diagonal block's lower part from W (the FT values
            // the update stored); the D phase sets the rest
            const int k0 = a.diag_k0, w = min(LT, n - k0);
            for (int idx = tid; idx < LT * LT; idx += NTB) {
                const int r = idx & (LT - 1), c = idx >> 6;
                if (r < w && c < w && r >= c)
                    S.u.f.D[c * LS + r] = (double)W[(k0 + r) + (long long)(k0 + c) * ld];
            }
        }
        __syncthreads();
        for (int jb = jb0; jb < jb1; ++jb) {
            const int k0 = jb * LT, w = min(LT, n - k0), m = n - k0;
            if (FULL) {
                // ---- U: left-looking update of block column jb, 128-row
                // tiles; the diagonal block goes straight to the LDS block D
                const int nrt = (m + LRT - 1) / LRT;
                for (int rt = 0; rt < nrt; ++rt) {
                    ll_update_tile<FT>(A, W, n, ld, k0, w, k0 + rt * LRT, rt == 0, S.u.f.D, a.rowmap);
                    __syncthreads();
                }
            }
            // ---- D: factor and invert the diagonal block in LDS (double).
            // D's lower part came from the update; the rest is the identity.
            // On exit L_D is in W and the inverse Di in the block's slot of
            // Dinv, or S.bad holds the first bad pivot (1-based, global row).
            // (Inline here: moved into a function of its own, this phase
            // pushes the kernel past 256 VGPRs into scratch spills.)
            double *D = S.u.f.D, *Di = S.u.f.Di;
            for (int idx = tid; idx < LT * LT; idx += NTB) {
                const int r = idx & (LT - 1), c = idx >> 6;
                if (!(r < w && c < w && r >= c)) D[c * LS + r] = (r == c) ? 1.0 : 0.0;
                Di[c * LS + r] = (r == c) ? 1.0 : 0.0;
            }
            __syncthreads();
            // trailing update of the 64-block from the panel at columns
            // p0 .. p0+pw-1 (lower part, rows and columns from p0+pw):
            // either only the next sub-block's 16 x 16 diagonal block, or
            // everything else, by threads t0, t0 + nt, ...
            // Only lower elements (i >= k) are enumerated: element t of the
            // packed lower triangle, t = i (i + 1) / 2 + k, from t = 136 (past
            // the leading 16 x 16 triangle) unless diag_only.  A full panel
            // (pw = 16) loads its 32 operands before the products, which keep
            // the order of the one-at-a-time loop.
            auto trail = [&](int p0, int pw, bool diag_only, int t0, int nt) {
                const int b0 = p0 + pw, mt = w - b0, md = diag_only ? min(16, mt) : mt;
                const int tend = md * (md + 1) / 2;
                for (int t = (diag_only ? 0 : 136) + t0; t < tend; t += nt) {
                    int i, k;
                    tri_lower_decode(t, i, k);
                    double v = D[(b0 + k) * LS + b0 + i];
                    if (pw == 16) {
                        double x[16], y[16];
#pragma unroll
                        for (int c = 0; c < 16; ++c) {
                            x[c] = D[(p0 + c) * LS + b0 + i];
                            y[c] = D[(p0 + c) * LS + b0 + k];
                        }
#pragma unroll
                        for (int c = 0; c < 16; ++c) v -= x[c] * y[c];
                    } else {
                        for (int c = 0; c < pw; ++c) v -= D[(p0 + c) * LS + b0 + i] * D[(p0 + c) * LS + b0 + k];
                    }
                    D[(b0 + k) * LS + b0 + i] = v;
                }
            };
            for (int s0 = 0; s0 < w; s0 += 16) {
                const int sw = min(16, w - s0);
                double *Li = S.u.f.T;   // inverse of the sub-block's factor, Li(i, c) at i + 16c
                if (wave == 0) {
                    // also the diagonal block of Di
                    const int bad = factor16(to_lds(D), s0, sw, to_lds(Li), to_lds(Di), lane);
                    if (lane == 0 && bad && S.bad == 0) S.bad = k0 + s0 + bad;
                } else if (s0 > 0) {
                    // meanwhile the rest of the previous panel's trailing update
                    // (disjoint from this sub-block, which it already updated)
                    trail(s0 - 16, 16, false, tid - 64, NTB - 64);
                }
                __syncthreads();
                if (S.bad) break;
                // rows of the 64-block below the sub-block: x = row L_ss^-T,
                // x_c = sum_{d <= c} row_d Li(c, d); four threads per row,
                // four columns each, written back after every row is read
                {
                    const int nrow = w - s0 - sw;   // <= 48: at most 192 threads
                    const int i = s0 + sw + (tid >> 2), g = tid & 3;
                    const bool act = tid < 4 * nrow;
                    double xo[4];
                    if (act) {
                        double rv[16];
        #pragma unroll
                        for (int d = 0; d < 16; ++d) rv[d] = (d < sw && d < 4 * g + 4) ? D[(s0 + d) * LS + i] : 0.0;
        #pragma unroll
                        for (int cc = 0; cc < 4; ++cc) {
                            const int c = 4 * g + cc;
                            double x = 0.0;
        #pragma unroll
                            for (int d = 0; d < 16; ++d)
                                if (d <= c) x += rv[d] * Li[c + 16 * d];
                            xo[cc] = x;
                        }
                    }
                    __syncthreads();
                    if (act) {
        #pragma unroll
                        for (int cc = 0; cc < 4; ++cc)
                            if (4 * g + cc < sw) D[(s0 + 4 * g + cc) * LS + i] = xo[cc];
                    }
                }
                __syncthreads();
                // the next sub-block's diagonal block now; the rest of this
                // update beside the next factor16 (a sub-block narrower than
                // 16 is the last: nothing is left to update after it)
                trail(s0, sw, true, tid, NTB);
                __syncthreads();
            }
            if (S.bad) break;
            // inverse, blocked by 16: the diagonal blocks X_ii are the
            // sub-block inverses factor16 left in Di; block row bi = 1..3 from
            // X_(bi,bj) = -X_(bi,bi) sum_{kb=bj}^{bi-1} L_(bi,kb) X_(kb,bj), one
            // block per wave (bj = wave), the 16 x 16 products on MFMA.  The
            // sum T as the accumulators hold it (row 4r + lane/16 of column
            // lane%16 in t[r]) is the next product's B operand for k-step r
            // as it stands.  One barrier per block row.
            for (int bi = 1; bi * 16 < w; ++bi) {
                if (wave < bi) {
                    const int bj = wave, il = lane & 15, kl = lane >> 4;
                    d4 t = d4{0, 0, 0, 0};
                    for (int kb = bj; kb < bi; ++kb) {
#pragma unroll
                        for (int k0 = 0; k0 < 16; k0 += 4) {
                            const int k = kb * 16 + k0 + kl;
                            t = mfma4(D[k * LS + bi * 16 + il], Di[(bj * 16 + il) * LS + k], t);
                        }
                    }
                    d4 x = d4{0, 0, 0, 0};
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        x = mfma4(Di[(bi * 16 + 4 * r + kl) * LS + bi * 16 + il], t[r], x);
#pragma unroll
                    for (int r = 0; r < 4; ++r) Di[(bj * 16 + il) * LS + bi * 16 + kl + 4 * r] = -x[r];
                }
                __syncthreads();
            }
            // L_D into W, Di into the inverse store
            for (int idx = tid; idx < LT * LT; idx += NTB) {
                const int r = idx & (LT - 1), c = idx >> 6;
                if (r < w && c < w && r >= c) W[(k0 + r) + (long long)(k0 + c) * ld] = (FT)D[c * LS + r];
                Dinv[(long long)jb * LT * LT + idx] = (FT)((r >= c) ? Di[c * LS + r] : 0.0);
            }
            __syncthreads();
            if (S.bad) break;
            if (FULL) {
                // ---- T: rows below, X = C Di^T
                const int mb = m - w;
                for (int r0 = k0 + w; r0 < k0 + w + mb; r0 += LT)
                    ll_trsm_tile<FT, false>(W, S.u.f.Di, nullptr, n, ld, k0, r0, a.rowmap);
            }
        }
        if (!FULL) {
            if (tid == 0 && S.bad) a.status[sidx] = S.bad;
            return;
        }
        if (S.bad) {
            if (tid == 0) {
                a.status[sidx] = S.bad;
                a.berr[sidx] = INFINITY;
                a.iters[sidx] = 0;
            }
            __syncthreads();
            continue;
        }
        // ---- solve and refine, LRMAX right-hand sides at a time; the
        // correction d in LDS, x accumulated in the caller's X
        double eta_all = 0.0, lres_all = 0.0;
        int it_all = 0;
        for (int c0 = 0; c0 < nrhs_all; c0 += LRMAX) {
            const int nrhs = min(LRMAX, nrhs_all - c0);
            const double *B = B0 + (long long)c0 * ld;
            double *X = X0 + (long long)c0 * ld;
            double *y = S.u.v.y;
            // two loops: a store to X may alias B for the compiler
            for (int idx = tid; idx < n * nrhs; idx += NTB) {
                const int i = idx % n, k = idx / n;
                y[k * NMAX + i] = B[i + (long long)k * ld];
            }
            for (int idx = tid; idx < n * nrhs; idx += NTB) {
                const int i = idx % n, k = idx / n;
                X[i + (long long)k * ld] = 0.0;
            }
            __syncthreads();
            int it = 0;
            double eta = INFINITY, prev = INFINITY, lres = INFINITY;
            for (;;) {
                tri_solve_ll<FT>(S, W, Dinv, n, ld, nrhs);
                // x += d, eight loads of x before their stores
                for (int idx0 = tid; idx0 < n * nrhs; idx0 += 8 * NTB) {
                    double xv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int idx = idx0 + u * NTB;
                        const int i = idx % n, k = idx / n;
                        xv[u] = idx < n * nrhs ? X[i + (long long)k * ld] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int idx = idx0 + u * NTB;
                        const int i = idx % n, k = idx / n;
                        if (idx < n * nrhs) X[i + (long long)k * ld] = xv[u] + y[k * NMAX + i];
                    }
                }
                __syncthreads();
                // r = b - A x: x staged through LDS 64 entries at a time, each
                // chunk once for two of this thread's rows (tid + g NTB, g in
                // a pair), 2 x 8 loads before their products
                double amax = 0.0;
                for (int g0 = 0; g0 * NTB < n; g0 += 2) {
                    const int i0 = g0 * NTB + tid, i1 = i0 + NTB;
                    const bool v0 = i0 < n, v1 = i1 < n;
                    double rs0[LRMAX], rs1[LRMAX], as0 = 0.0, as1 = 0.0;
#pragma unroll
                    for (int k = 0; k < LRMAX; ++k) {
                        rs0[k] = (v0 && k < nrhs) ? B[i0 + (long long)k * ld] : 0.0;
                        rs1[k] = (v1 && k < nrhs) ? B[i1 + (long long)k * ld] : 0.0;
                    }
                    for (int j0 = 0; j0 < n; j0 += LT) {
                        const int jw = min(LT, n - j0);
                        for (int idx = tid; idx < LT * LRMAX; idx += NTB) {
                            const int jj = idx & (LT - 1), k = idx >> 6;
                            S.tmp[k * LT + jj] = (jj < jw && k < nrhs) ? X[j0 + jj + (long long)k * ld] : 0.0;
                        }
                        __syncthreads();
                        if (v0) {
                            const double *a0 = A + i0 + (long long)j0 * ld, *a1 = A + i1 + (long long)j0 * ld;
                            int jj = 0;
                            for (; jj + 8 <= jw; jj += 8) {
                                double av0[8], av1[8];
#pragma unroll
                                for (int u = 0; u < 8; ++u) {
                                    av0[u] = a0[(long long)(jj + u) * ld];
                                    av1[u] = v1 ? a1[(long long)(jj + u) * ld] : 0.0;
                                }
#pragma unroll
                                for (int u = 0; u < 8; ++u) {
                                    as0 += fabs(av0[u]);
                                    as1 += fabs(av1[u]);
#pragma unroll
                                    for (int k = 0; k < LRMAX; ++k) {
                                        const double t = S.tmp[k * LT + jj + u];
                                        rs0[k] -= av0[u] * t;
                                        rs1[k] -= av1[u] * t;
                                    }
                                }
                            }
                            for (; jj < jw; ++jj) {
                                const double b0 = a0[(long long)jj * ld];
                                const double b1 = v1 ? a1[(long long)jj * ld] : 0.0;
                                as0 += fabs(b0);
                                as1 += fabs(b1);
#pragma unroll
                                for (int k = 0; k < LRMAX; ++k) {
                                    const double t = S.tmp[k * LT + jj];
                                    rs0[k] -= b0 * t;
                                    rs1[k] -= b1 * t;
                                }
                            }
                        }
                        __syncthreads();
                    }
                    if (v0) {
                        amax = fmax(amax, as0);
#pragma unroll
                        for (int k = 0; k < LRMAX; ++k)
                            if (k < nrhs) y[k * NMAX + i0] = rs0[k];
                    }
                    if (v1) {
                        amax = fmax(amax, as1);
#pragma unroll
                        for (int k = 0; k < LRMAX; ++k)
                            if (k < nrhs) y[k * NMAX + i1] = rs1[k];
                    }
                }
                double rmax[LRMAX], xmax[LRMAX], bmax[LRMAX];
#pragma unroll
                for (int k = 0; k < LRMAX; ++k) rmax[k] = xmax[k] = bmax[k] = 0.0;
                for (int i = tid; i < n; i += NTB) {   // this thread's own rows of y
#pragma unroll
                    for (int k = 0; k < LRMAX; ++k) {
                        if (k >= nrhs) break;
                        rmax[k] = fmax(rmax[k], fabs(y[k * NMAX + i]));
                        xmax[k] = fmax(xmax[k], fabs(X[i + (long long)k * ld]));
                        bmax[k] = fmax(bmax[k], fabs(B[i + (long long)k * ld]));
                    }
                }
                // the 1 + 3 * LRMAX maxima: a butterfly inside each wave, one
                // value at a time, then the four waves' partials from LDS
                // (max is exact: every thread gets the same values)
                auto wave_max = [&](double x, int slot) {
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
                    if (lane == 0) S.red[wave * 32 + slot] = x;
                };
                wave_max(amax, 0);
#pragma unroll
                for (int k = 0; k < LRMAX; ++k) {
                    wave_max(rmax[k], 1 + 3 * k);
                    wave_max(xmax[k], 2 + 3 * k);
                    wave_max(bmax[k], 3 + 3 * k);
                }
                __syncthreads();
                auto red4 = [&](int v) {
                    return fmax(fmax(S.red[v], S.red[32 + v]), fmax(S.red[64 + v], S.red[96 + v]));
                };
                const double An = red4(0);
                double e = 0.0, lr = 0.0;
                for (int k = 0; k < nrhs; ++k) {
                    const double rn = red4(1 + 3 * k), xn = red4(2 + 3 * k), bn = red4(3 + 3 * k);
                    e = fmax(e, rn / (An * xn + bn));
                    lr = fmax(lr, rn == 0.0 ? 0.0 : rn / (An * xn * kEps * n));
                }
                __syncthreads();
                prev = eta;
                eta = e;
                lres = lr;
                if (eta <= a.tol) break;
                if (it >= a.max_iter || !(eta < 0.5 * prev)) break;
                ++it;
            }
            eta_all = fmax(eta_all, eta);
            lres_all = fmax(lres_all, lres);
            it_all = max(it_all, it);
            __syncthreads();
        }
        if (tid == 0) {
            a.berr[sidx] = eta_all;
            a.lres[sidx] = lres_all;
            a.iters[sidx] = it_all;
            a.status[sidx] = (eta_all <= a.tol) ? 0 : -2;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Block-column launches for the largest systems of a batch (queue entries
// a.order[0 .. cnt), sorted largest first, so the systems still active at a
// step are a prefix).  Inside one workgroup a system's 64-wide block columns
// form a serial chain (update, diagonal factor, rows below); here step k0 of
// every large system runs as three launches -- one workgroup per 128-row
// update tile (k_bl_update), one per system for the diagonal block
// (k_batched_ll<FT, false>), one per 64-row tile below it (k_bl_trsm) -- and
// k_batched_ll then only solves and refines them (a.prefact = cnt).  The
// phases are k_batched_ll's own code, so the factor is the same bit for bit.
// A system whose diagonal block meets a bad pivot records it in a.status and
// is skipped by the later steps.
// ---------------------------------------------------------------------------
template <typename FT>
__global__ __launch_bounds__(NTB, 2) void k_bl_update(BatchArgs a, const int2 *items, int k0) {
    const int2 it = items[xcd_remap(blockIdx.x, gridDim.x)];   // (system, 128-row tile)
    if (a.status[it.x] != 0) return;
    const SysDesc sd = a.sys[it.x];
    ll_update_tile<FT>(a.A + sd.a, (FT *)a.W + sd.w, sd.n, sd.ld, k0, min(LT, sd.n - k0),
                       k0 + it.y * LRT, false, nullptr, a.rowmap);
}

template <typename FT>
__global__ __launch_bounds__(NTB, 2) void k_bl_trsm(BatchArgs a, const int2 *items, int k0) {
    const int2 it = items[xcd_remap(blockIdx.x, gridDim.x)];   // (system, 64-row tile below)
    if (a.status[it.x] != 0) return;
    const SysDesc sd = a.sys[it.x];
    ll_trsm_tile<FT, true>((FT *)a.W + sd.w, nullptr,
                           (const FT *)a.Dinv + sd.d + (long long)(k0 / LT) * LT * LT, sd.n, sd.ld,
                           k0, k0 + LT + it.y * LT, a.rowmap);
}

// status of queue entries order[0 .. cnt) <- 0 (the block-column launches
// record bad pivots there)
__global__ void k_zero_status(BatchArgs a, int cnt) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x)
        a.status[a.order[q]] = 0;
}

__global__ void k_set_int(int *p, int v) { *p = v; }

__global__ void k_add_diag_one(double *A, int n, int ld, double mu) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) A[i + (long long)i * ld] += mu;
}

__global__ void k_add_diag(double *A, const SysDesc *sys, int nsys, double mu) {
    const int s = blockIdx.x;
    if (s >= nsys) return;
    const SysDesc sd = sys[s];
    for (int i = threadIdx.x; i < sd.n; i += blockDim.x) A[sd.a + i + (long long)i * sd.ld] += mu;
}

template <typename FT, int T, int RB>
size_t smem_bytes() {
    return sizeof(Smem<FT, T>);
}

typedef void (*BatchKernel)(BatchArgs);

struct Variant {
    BatchKernel k;
    size_t smem;
    int T, RB, fp32;
    int wgs;   // workgroups per CU
};

// RB = 0 marks the left-looking MFMA kernel (T = 64)
Variant variants_fp64[] = {
    {k_batched<double, 16, 4>, smem_bytes<double, 16, 4>(), 16, 4, 0, 1},
    {k_batched<double, 16, 8>, smem_bytes<double, 16, 8>(), 16, 8, 0, 1},
    {k_batched_ll<double>, sizeof(SmemLL<double>), 64, 0, 0, 2},
};
Variant variants_fp32[] = {
    {k_batched<float, 16, 4>, smem_bytes<float, 16, 4>(), 16, 4, 1, 1},
    {k_batched<float, 16, 8>, smem_bytes<float, 16, 8>(), 16, 8, 1, 1},
    {k_batched<float, 32, 4>, smem_bytes<float, 32, 4>(), 32, 4, 1, 1},
    {k_batched<float, 32, 8>, smem_bytes<float, 32, 8>(), 32, 8, 1, 1},
    {k_batched_ll<float>, sizeof(SmemLL<float>), 64, 0, 1, 2},
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

typedef void (*BlTileKernel)(BatchArgs, const int2 *, int);
struct BlKernels {
    BlTileKernel update, trsm;
    BatchKernel diag;
};
const BlKernels bl_kernels[2] = {
    {k_bl_update<double>, k_bl_trsm<double>, k_batched_ll<double, false>},
    {k_bl_update<float>, k_bl_trsm<float>, k_batched_ll<float, false>},
};

// Work lists of the block-column launches for queue entries order[0 .. nl)
// (n descending): per step k0 = 64 j the systems still active (n > k0, a
// prefix of the order), their 128-row update tiles from row k0 and their
// 64-row tiles below the diagonal block, as (system, tile) pairs.
struct BlPlan {
    int nl = 0;
    std::vector<int> active;        // systems per step
    std::vector<int> u_off, t_off;  // per step (+ 1 end entry) into items
    std::vector<int2> items;
};

BlPlan bl_plan(const std::vector<int> &order, const int *n, int nl) {
    BlPlan p;
    p.nl = nl;
    if (nl == 0) return p;
    const int steps = (n[order[0]] + LT - 1) / LT;
    for (int j = 0; j < steps; ++j) {
        const int k0 = j * LT;
        int act = 0;
        p.u_off.push_back((int)p.items.size());
        for (; act < nl && n[order[act]] > k0; ++act)
            for (int t = 0; t < (n[order[act]] - k0 + LRT - 1) / LRT; ++t)
                p.items.push_back(make_int2(order[act], t));
        p.active.push_back(act);
    }
    p.u_off.push_back((int)p.items.size());
    for (int j = 0; j < steps; ++j) {
        const int k0 = j * LT;
        p.t_off.push_back((int)p.items.size());
        for (int q = 0; q < p.active[j]; ++q) {
            const int mb = n[order[q]] - k0 - LT;   // rows below the diagonal block
            for (int t = 0; t < (mb > 0 ? (mb + LT - 1) / LT : 0); ++t)
                p.items.push_back(make_int2(order[q], t));
        }
    }
    p.t_off.push_back((int)p.items.size());
    return p;
}

// Leading entries of the order (n descending) with n >= min_n.
int count_at_least(const std::vector<int> &order, const int *n, int min_n) {
    if (min_n <= 0) return 0;
    int c = 0;
    while (c < (int)order.size() && n[order[c]] >= min_n) ++c;
    return c;
}

}  // namespace
}  // namespace ooc

using namespace ooc;

extern "C" long long oocdag_batch_layout(int nsys, const int *n, const int *nrhs, long long *a_off,
                                         long long *b_off) {
    // A blocks n x n and B/X blocks n x nrhs with leading dimension
    // ld = n rounded up to even, each 32-double aligned;
    // returns the total doubles of the A arena, B arena size in b_off[nsys]
    long long a = 0, b = 0;
    for (int s = 0; s < nsys; ++s) {
        const long long ld = n[s] + (n[s] & 1);
        a_off[s] = a;
        b_off[s] = b;
        a += (ld * n[s] + 31) / 32 * 32;
        b += (ld * nrhs[s] + 31) / 32 * 32;
    }
    b_off[nsys] = b;
    return a;
}

extern "C" int oocdag_batch_generate(int nsys, const int *n, const int *nrhs, const long long *a_off,
                                     const long long *b_off, double *dA, double *dB, double mu,
                                     unsigned long long seed) {
    // A_s = (1/n) G_s G_s^T + mu I (full symmetric), B_s ~ N(0,1)
    int nmax = 0;
    for (int s = 0; s < nsys; ++s) nmax = std::max(nmax, n[s]);
    double *G = nullptr;
    if (hipMalloc(&G, (size_t)nmax * nmax * sizeof(double)) != hipSuccess) return OOCDAG_ENOMEM;
    std::vector<SysDesc> sd(nsys);
    for (int s = 0; s < nsys; ++s) {
        sd[s].n = n[s];
        sd[s].nrhs = nrhs[s];
        sd[s].ld = n[s] + (n[s] & 1);
        sd[s].a = a_off[s];
        sd[s].b = b_off[s];
    }
    int rc = OOCDAG_OK;
    for (int s = 0; s < nsys && rc == OOCDAG_OK; ++s) {
        const int m = n[s], ld = m + (m & 1);
        sd[s].ld = ld;
        if (oocdag_normal_fill(G, ld, m, ld, seed, 1000000ull + s) != OOCDAG_OK) rc = OOCDAG_EHIP;
        // zero the padding row of G so G G^T has zeros there
        if (rc == OOCDAG_OK && (m & 1))
            (void)hipMemset2D(G + m, ld * sizeof(double), 0, sizeof(double), m);
        Gemm g;
        g.A = plain(G, ld);
        g.B = g.A;
        g.C = plain(dA + a_off[s], ld);
        g.M = m; g.N = m; g.K = m;
        g.alpha = 1.0 / m; g.beta = 0.0; g.lower = 0;
        if (rc == OOCDAG_OK && launch_gemm(g, GEMM_GENERAL, 0) != hipSuccess) rc = OOCDAG_EHIP;
        if (rc == OOCDAG_OK &&
            oocdag_normal_fill(dB + b_off[s], m, nrhs[s], ld, seed, 2000000ull + s) != OOCDAG_OK)
            rc = OOCDAG_EHIP;
    }
    SysDesc *dsd = nullptr;
    if (rc == OOCDAG_OK && hipMalloc(&dsd, nsys * sizeof(SysDesc)) == hipSuccess) {
        (void)hipMemcpy(dsd, sd.data(), nsys * sizeof(SysDesc), hipMemcpyHostToDevice);
        hipLaunchKernelGGL(k_add_diag, dim3(nsys), dim3(256), 0, 0, dA, dsd, nsys, mu);
        if (hipDeviceSynchronize() != hipSuccess) rc = OOCDAG_EHIP;
        (void)hipFree(dsd);
    }
    (void)hipFree(G);
    return rc;
}

namespace {
std::mutex g_tune_mu;
int g_tuned[2] = {-1, -1};
int g_forced[2] = {-1, -1};
int g_split_forced[2] = {-1, -1};
constexpr int kVariants[2] = {3, 5};   // entries of variants_fp64 / variants_fp32
}  // namespace

extern "C" int oocdag_batch_set_variant(int precision, int variant) {
    if (precision != 0 && precision != 1) return OOCDAG_EARG;
    if (variant < -1 || variant >= kVariants[precision]) return OOCDAG_EARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_forced[precision] = variant;
    return kVariants[precision];
}

extern "C" int oocdag_batch_set_split(int precision, int min_n) {
    if (precision != 0 && precision != 1) return OOCDAG_EARG;
    if (min_n < -1) return OOCDAG_EARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_split_forced[precision] = min_n;
    return OOCDAG_OK;
}

namespace ooc {
// Generates one A = (1/n) G G^T + mu I (full, ld even) with G from the
// counter-based normal generator keyed by (seed, stream_id); Gtmp holds at
// least ld * n doubles.
hipError_t gen_small_gram(double *A, int n, int ld, double mu, unsigned long long seed,
                          unsigned long long stream_id, double *Gtmp, hipStream_t s) {
    hipError_t e = normal_block(Gtmp, n, n, ld, seed, stream_id, s);
    if (e != hipSuccess) return e;
    if (ld != n && (e = hipMemset2DAsync(Gtmp + n, ld * sizeof(double), 0, sizeof(double), n, s)) != hipSuccess)
        return e;
    Gemm g;
    g.A = plain(Gtmp, ld);
    g.B = g.A;
    g.C = plain(A, ld);
    g.M = n; g.N = n; g.K = n;
    g.alpha = 1.0 / n; g.beta = 0.0; g.lower = 0;
    if ((e = launch_gemm(g, GEMM_GENERAL, s)) != hipSuccess) return e;
    SysDesc sd;
    sd.n = n;
    sd.ld = ld;
    sd.a = 0;
    // add mu on the diagonal through a one-system descriptor passed by value
    hipLaunchKernelGGL(k_add_diag_one, dim3(1), dim3(256), 0, s, A, n, ld, mu);
    (void)sd;
    return hipGetLastError();
}
}  // namespace ooc

extern "C" int oocdag_batched_solve(int nsys, const int *n, const int *nrhs, const long long *a_off,
                                    const long long *b_off, const double *dA, const double *dB,
                                    double *dX, int precision, double tol, int max_iter, int reps,
                                    double *berr, double *lres, int *iters, int *status,
                                    oocdag_batch_stats *st) {
    oocdag_batch_stats dummy;
    if (!st) st = &dummy;
    *st = oocdag_batch_stats{};
    if (nsys <= 0 || !n || !nrhs || !dA || !dB || !dX || (precision != 0 && precision != 1))
        return OOCDAG_EARG;
    for (int s = 0; s < nsys; ++s)
        if (n[s] < 1 || n[s] > PMAX || nrhs[s] < 1 || nrhs[s] > RTOTAL) return OOCDAG_EARG;
    // descriptors and workspace offsets (FT elements), largest systems first
    std::vector<SysDesc> sd(nsys);
    const int Tmax = 32;
    long long w = 0, d = 0;
    for (int s = 0; s < nsys; ++s) {
        sd[s].n = n[s];
        sd[s].nrhs = nrhs[s];
        sd[s].ld = n[s] + (n[s] & 1);
        sd[s].a = a_off[s];
        sd[s].b = b_off[s];
        sd[s].w = w;
        sd[s].d = d;
        w += ((long long)sd[s].ld * n[s] + 31) / 32 * 32;
        d += std::max((long long)((n[s] + 15) / 16) * Tmax * Tmax,
                      (long long)((n[s] + LT - 1) / LT) * LT * LT);
    }
    std::vector<int> order(nsys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
    const size_t fts = precision == 0 ? 8 : 4;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return OOCDAG_EHIP;
    struct SS {
        hipStream_t s;
        ~SS() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } ss{s};
    SysDesc *dsd = nullptr;
    int *dorder = nullptr, *dint = nullptr, *dcounter = nullptr;
    double *dberr = nullptr;
    void *dW = nullptr, *dD = nullptr;
    int2 *ditems = nullptr;   // work lists of the block-column launches
    struct F {
        hipStream_t s;
        void **p[8];
        ~F() {
            for (auto q : p)
                if (*q) (void)hipFreeAsync(*q, s);
        }
    } fr{s, {(void **)&dsd, (void **)&dorder, (void **)&dint, (void **)&dcounter, (void **)&dberr, &dW,
          &dD, (void **)&ditems}};
    if (hipMallocAsync((void **)&dsd, nsys * sizeof(SysDesc), s) != hipSuccess ||
        hipMallocAsync((void **)&dorder, nsys * sizeof(int), s) != hipSuccess ||
        hipMallocAsync((void **)&dint, 2 * (size_t)nsys * sizeof(int), s) != hipSuccess ||
        hipMallocAsync((void **)&dcounter, sizeof(int), s) != hipSuccess ||
        hipMallocAsync((void **)&dberr, 2 * (size_t)nsys * sizeof(double), s) != hipSuccess ||
        hipMallocAsync(&dW, w * fts, s) != hipSuccess || hipMallocAsync(&dD, d * fts, s) != hipSuccess)
        return OOCDAG_ENOMEM;
    if (hipMemcpyAsync(dsd, sd.data(), nsys * sizeof(SysDesc), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dorder, order.data(), nsys * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return OOCDAG_EHIP;
    int ncu = 256;
    {
        int dev = 0;
        hipDeviceProp_t pr;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess)
            ncu = pr.multiProcessorCount;
    }
    // FP32 MFMA result row order (the FP64 one is fixed)
    const int rowmap = precision == 0 ? 1 : mfma_rowmap(0, s);
    if (rowmap < 0) return OOCDAG_EHIP;
    auto make_args = [&](const int *ord, int cnt) {
        BatchArgs a;
        a.sys = dsd;
        a.order = ord;
        a.nsys = cnt;
        a.A = dA;
        a.B = dB;
        a.X = dX;
        a.W = dW;
        a.Dinv = dD;
        a.berr = dberr;
        a.lres = dberr + nsys;
        a.iters = dint;
        a.status = dint + nsys;
        a.counter = dcounter;
        a.tol = tol;
        a.max_iter = max_iter;
        a.rowmap = rowmap;
        a.prefact = 0;
        a.diag_k0 = 0;
        return a;
    };
    // the systems of the sorted order with n >= min_n (0: none) go through
    // the block-column launches; their work lists are uploaded to ditems
    BlPlan plan;
    auto set_plan = [&](int min_n) -> hipError_t {
        plan = bl_plan(order, n, count_at_least(order, n, min_n));
        if (ditems) {
            (void)hipFreeAsync(ditems, s);
            ditems = nullptr;
        }
        if (plan.items.empty()) return hipSuccess;
        hipError_t e = hipMallocAsync((void **)&ditems, plan.items.size() * sizeof(int2), s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ditems, plan.items.data(), plan.items.size() * sizeof(int2),
                               hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e;
    };
    // one batch: the block-column steps of the plan's systems (when split,
    // with the left-looking kernel only: the one that can solve systems
    // factored elsewhere), then the persistent kernel over the queue
    auto run = [&](const Variant &v, const int *ord, int cnt, bool split) -> hipError_t {
        BatchArgs a = make_args(ord, cnt);
        if (split && v.RB == 0 && plan.nl > 0) {
            const BlKernels &bk = bl_kernels[precision];
            hipLaunchKernelGGL(k_zero_status, dim3((plan.nl + 255) / 256), dim3(256), 0, s, a,
                               plan.nl);
            for (size_t j = 0; j < plan.active.size(); ++j) {
                const int k0 = (int)j * LT;
                const int nu = plan.u_off[j + 1] - plan.u_off[j];
                const int nt = plan.t_off[j + 1] - plan.t_off[j];
                hipLaunchKernelGGL(bk.update, dim3(nu), dim3(NTB), 0, s, a,
                                   (const int2 *)ditems + plan.u_off[j], k0);
                a.diag_k0 = k0;
                hipLaunchKernelGGL(bk.diag, dim3(plan.active[j]), dim3(NTB), sizeof(SmemLL<double>),
                                   s, a);
                if (nt > 0)
                    hipLaunchKernelGGL(bk.trsm, dim3(nt), dim3(NTB), 0, s, a,
                                       (const int2 *)ditems + plan.t_off[j], k0);
            }
            a.prefact = plan.nl;
        }
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, dcounter, 0);
        hipLaunchKernelGGL(v.k, dim3(ncu * v.wgs), dim3(NTB), v.smem, s, a);
        return hipGetLastError();
    };
    Variant *vs = precision == 0 ? variants_fp64 : variants_fp32;
    const int nv = kVariants[precision];
    for (int i = 0; i < nv; ++i)
        if (vs[i].smem > 0 && vs[i].smem <= 160 * 1024)
            (void)hipFuncSetAttribute((const void *)vs[i].k,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)vs[i].smem);
    (void)hipFuncSetAttribute((const void *)bl_kernels[precision].diag,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(SmemLL<double>));
    int best = -1, split_n = -1;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_forced[precision] >= 0) best = g_forced[precision];
        else if (reps <= 0) best = g_tuned[precision];
        split_n = g_split_forced[precision];
    }
    // --- autotune on a sample of the sorted order (skipped when a tuned
    // variant is cached and the caller asked for a plain run)
    const double tt0 = now_s();
    if (best < 0) {
    std::vector<int> sample;
    for (int i = 0; i < nsys; i += std::max(1, nsys / 256)) sample.push_back(order[i]);
    int *dsample = nullptr;
    if (hipMallocAsync((void **)&dsample, sample.size() * sizeof(int), s) != hipSuccess) return OOCDAG_ENOMEM;
    (void)hipMemcpyAsync(dsample, sample.data(), sample.size() * sizeof(int), hipMemcpyHostToDevice, s);
    double best_t = 1e30;
    for (int i = 0; i < nv; ++i) {
        if (vs[i].smem == 0 || vs[i].smem > 160 * 1024) continue;
        double tbest = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            double t0 = now_s();
            if (run(vs[i], dsample, (int)sample.size(), false) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                tbest = 1e30;
                break;
            }
            tbest = std::min(tbest, now_s() - t0);
        }
        if (tbest < best_t) {
            best_t = tbest;
            best = i;
        }
    }
    (void)hipFreeAsync(dsample, s);
    if (best < 0) return OOCDAG_EHIP;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[precision] = best;
    }
    }
    if (vs[best].smem == 0 || vs[best].smem > 160 * 1024) return OOCDAG_EARG;
    // --- the block-column threshold.  A plain run (reps <= 0, the server)
    // takes the measured rule: systems of n >= 513 split when there are at
    // most two per CU of them (more keep every CU busy inside the persistent
    // kernel, which then overlaps their phases better than the launches do);
    // otherwise timed on the whole batch, no split and each candidate that
    // moves a different number of systems
    if (split_n < 0 && vs[best].RB == 0 && reps <= 0) {
        const int nl = count_at_least(order, n, 513);
        split_n = (nl > 0 && nl <= 2 * ncu) ? 513 : 0;
    }
    if (split_n < 0 && vs[best].RB == 0) {
        double best_t = 1e30;
        int last_nl = -1;
        for (int cand : {0, 769, 513, 385, 257}) {
            const int nl = count_at_least(order, n, cand);
            if (nl == last_nl) continue;
            last_nl = nl;
            if (set_plan(cand) != hipSuccess) return OOCDAG_EHIP;
            double tbest = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                const double t0 = now_s();
                if (run(vs[best], dorder, nsys, true) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)
                    return OOCDAG_EHIP;
                tbest = std::min(tbest, now_s() - t0);
            }
            if (tbest < best_t) {
                best_t = tbest;
                split_n = cand;
            }
        }
    }
    if (split_n < 0 || vs[best].RB != 0) split_n = 0;
    if (set_plan(split_n) != hipSuccess) return OOCDAG_EHIP;
    st->tune_seconds = now_s() - tt0;
    st->split_min_n = split_n;
    st->split_systems = plan.nl;
    st->variant_T = vs[best].T;
    st->variant_RB = vs[best].RB;
    int rc = OOCDAG_OK;
    if (reps <= 0) {
        // one direct launch on this call's stream (used by the server)
        const double t0 = now_s();
        if (run(vs[best], dorder, nsys, true) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            return OOCDAG_EHIP;
        st->seconds = now_s() - t0;
    } else {
    // --- capture the full batch (queue reset + persistent kernel) in a graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return OOCDAG_EHIP;
    hipError_t ce = run(vs[best], dorder, nsys, true);
    if (hipStreamEndCapture(s, &graph) != hipSuccess || ce != hipSuccess) return OOCDAG_EHIP;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return OOCDAG_EHIP;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    double tsum = 0;
    for (int r = 0; r < std::max(1, reps); ++r) {
        (void)hipEventRecord(e0, s);
        if (hipGraphLaunch(exec, s) != hipSuccess) rc = OOCDAG_EHIP;
        (void)hipEventRecord(e1, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = OOCDAG_EHIP;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        tsum += ms * 1e-3;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    st->seconds = tsum / std::max(1, reps);
    }
    st->systems_per_second = nsys / st->seconds;
    std::vector<double> hb(2 * (size_t)nsys);
    std::vector<int> hi(2 * (size_t)nsys);
    if (hipMemcpyAsync(hb.data(), dberr, 2 * (size_t)nsys * sizeof(double), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipMemcpyAsync(hi.data(), dint, 2 * (size_t)nsys * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return OOCDAG_EHIP;
    for (int i = 0; i < nsys; ++i) {
        if (berr) berr[i] = hb[i];
        if (lres) lres[i] = hb[nsys + i];
        if (iters) iters[i] = hi[i];
        if (status) status[i] = hi[nsys + i];
        if (hi[nsys + i] != 0) {
            st->failed += 1;
        } else {
            st->max_backward_error = std::max(st->max_backward_error, hb[i]);
            st->max_solve_resid = std::max(st->max_solve_resid, hb[nsys + i]);
        }
    }
    return OOCDAG_OK;
}
