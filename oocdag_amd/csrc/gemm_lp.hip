// Low-precision panel updates for mixed-precision Cholesky with iterative
// refinement (no reference counterpart: the reference is FP64 only).
//
//   C64 -= s_i * s_j * (X_i X_j^T)     over the lower trapezoid of a panel
//
// C stays in the FP64 packed tiles; X is a low-precision copy of the tile
// column that drives the update, produced by k_convert_column:
//   FP32 : column-major tiles (ld = NB), s = 1, v_mfma_f32_16x16x4_f32
//   BF16 : row-major tiles (each row's NB values contiguous, so one 16-byte
//          LDS read feeds a lane of v_mfma_f32_16x16x32_bf16), one power-of-
//          two scale per tile so every tile is O(1) before rounding.
// Accumulation is FP32 in both cases; the diagonal POTRF, the TRSM, the
// residual and the solution update remain FP64 (factor.hip, solve.hip).
//
// The accumulator row order of the FP32-output MFMAs is read from a one-wave
// probe the first time the library runs (k_probe_mfma), which also checks
// the operand lane layouts assumed below; a mismatch is reported as an error.
#include "ooc_internal.h"

#include <cmath>
#include <mutex>

namespace ooc {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int BM = 128;
constexpr int NT = 256;

__device__ __forceinline__ long long tix(int ti, int tk, int ym, int tri) {
    return ti + (long long)tk * ym - (tri ? ((long long)tk * (tk + 1)) / 2 : 0);
}

__device__ __forceinline__ long long xcd_remap(long long bid, long long G) {
    const int NX = 8;
    long long per = G / NX, rem = G % NX;
    long long x = bid % NX, i = bid / NX;
    return (x < rem) ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

__device__ __forceinline__ void tri_decode(long long L, int Mb, int Nb, int &mb, int &nb) {
    double a = (double)Mb + 0.5;
    double disc = a * a - 2.0 * (double)L;
    int n = (int)floor(a - sqrt(fmax(disc, 0.0)));
    n = max(0, min(n, Nb - 1));
    auto S = [&](int c) { return (long long)c * Mb - ((long long)c * (c - 1)) / 2; };
    while (n > 0 && S(n) > L) --n;
    while (n + 1 < Nb && S(n + 1) <= L) ++n;
    nb = n;
    mb = n + (int)(L - S(n));
}

__device__ __forceinline__ unsigned short f2bf(float f) {
    // round to nearest even
    unsigned u = __float_as_uint(f);
    unsigned r = 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)((u + r) >> 16);
}

// ---- probe --------------------------------------------------------------
// A[i][k] = (k == 0 ? i + 1 : 0), B[k][j] = (k == 0 ? j + 1 : 0), so
// D[i][j] = (i + 1)(j + 1): out[lane*4 + r] = D element held in slot r.
__global__ void k_probe_mfma(int kind, float *out) {
    const int lane = threadIdx.x;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    if (kind == 0) {
        // f32 16x16x4: lane holds A[lane%16][lane/16], B[lane/16][lane%16]
        float a = (lane / 16 == 0) ? (float)(lane % 16 + 1) : 0.f;
        float b = (lane / 16 == 0) ? (float)(lane % 16 + 1) : 0.f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    } else {
        // bf16 16x16x32: lane holds A[lane%16][8*(lane/16) + e], e = 0..7
        s8 a, b;
        for (int e = 0; e < 8; ++e) {
            bool k0 = (lane / 16 == 0) && e == 0;
            a[e] = (short)f2bf(k0 ? (float)(lane % 16 + 1) : 0.f);
            b[e] = (short)f2bf(k0 ? (float)(lane % 16 + 1) : 0.f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                      __builtin_bit_cast(bf8, b), acc, 0, 0, 0);
    }
    for (int r = 0; r < 4; ++r) out[lane * 4 + r] = acc[r];
}

// ---- conversion ------------------------------------------------------------
// Tiles jt..bb-1 of tile column jt of the packed FP64 panel -> rect
// low-precision buffer.  mode 0: FP32 column-major; mode 1: BF16 row-major
// with per-tile scale 2^e (scale[t] holds it), e = exponent of the tile's
// max |x|.
__global__ void k_tile_absmax(const double *Y, int nb, int ym, int jt, int ntiles,
                              float *scale) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const double *tile = Y + tix(jt + t, jt, ym, 1) * (long long)nb * nb;
    double m = 0.0;
    for (long long i = threadIdx.x; i < (long long)nb * nb; i += blockDim.x) m = fmax(m, fabs(tile[i]));
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int e = 0;
        if (red[0] > 0.0) frexp(red[0], &e);
        scale[t] = ldexpf(1.0f, e);
    }
}

__global__ void k_convert_column(const double *Y, int nb, int ym, int jt, int ntiles, int mode,
                                 const float *scale, void *out) {
    const long long per = (long long)nb * nb;
    const long long total = per * ntiles;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long t = idx / per, w = idx - t * per;
        const double *tile = Y + tix(jt + (int)t, jt, ym, 1) * per;
        if (mode == 0) {
            ((float *)out)[idx] = (float)tile[w];
        } else {
            // w enumerates the output row-major: (row, k) = (w / nb, w % nb)
            int row = (int)(w / nb), k = (int)(w - (long long)row * nb);
            float v = (float)(tile[row + (long long)k * nb] / (double)scale[t]);
            ((unsigned short *)out)[idx] = f2bf(v);
        }
    }
}

// ---- FP32 update ------------------------------------------------------------
struct LpArgs {
    double *C;          // packed FP64 panel
    int c_ym;           // tile rows of the panel
    int r0;             // global row/col of the region's top-left (diagonal)
    const void *X;      // low-precision column: tile t = rows [r0 + t*nb, ...)
    const float *scale; // BF16: per-tile scale
    int nb;
    int M, N, K;        // region rows, cols, K = nb
    int mb, nbk;        // block grid
    long long nblk;
    int rowmap;         // 0: row = 4*(lane/16) + r, 1: row = lane/16 + 4*r
};

constexpr int BK32 = 32;            // floats per K step
constexpr int S32 = BM + 4;         // LDS row stride (floats)

__global__ __launch_bounds__(NT, 2) void k_update_f32(LpArgs a) {
    __shared__ float As[2][BK32 * S32];
    __shared__ float Bs[2][BK32 * S32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    long long L = xcd_remap(blockIdx.x, a.nblk);
    int mbi, nbi;
    tri_decode(L, a.mb, a.nbk, mbi, nbi);
    const int m0 = mbi * BM, n0 = nbi * BM;
    const float *X = (const float *)a.X;
    const long long per = (long long)a.nb * a.nb;
    // each thread streams 4 consecutive rows (16 B) of 4 columns per step:
    // q = tid + 256 s -> column q / 32, rows 4 * (q % 32)
    const int quad = tid & 31, col0 = tid >> 5;  // col0 in 0..7, columns col0 + 8 s
    auto rowptr = [&](int r) -> const float * {
        int t = r / a.nb;
        return X + t * per + (r - t * a.nb);
    };
    const float *pa = rowptr(m0 + 4 * quad);
    const float *pb = rowptr(n0 + 4 * quad);
    const bool oka = (m0 + 4 * quad) < a.M, okb = (n0 + 4 * quad) < a.N;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    f4 ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            int c = k0 + col0 + 8 * s;
            ra[s] = oka ? *reinterpret_cast<const f4 *>(pa + (long long)c * a.nb) : f4{0, 0, 0, 0};
            rb[s] = okb ? *reinterpret_cast<const f4 *>(pb + (long long)c * a.nb) : f4{0, 0, 0, 0};
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            int c = col0 + 8 * s;
            *reinterpret_cast<f4 *>(&As[buf][c * S32 + 4 * quad]) = ra[s];
            *reinterpret_cast<f4 *>(&Bs[buf][c * S32 + 4 * quad]) = rb[s];
        }
    };
    const int nk = a.K / BK32;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) gload((kt + 1) * BK32);
#pragma unroll
        for (int ks = 0; ks < BK32 / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            float fa[4], fb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[i] = As[cur][kk * S32 + wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = Bs[cur][kk * S32 + wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) lstore(cur ^ 1);
        __syncthreads();
    }
    // epilogue: C64 -= acc over the lower part
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int mb16 = m0 + wm * 64 + i * 16;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn * 64 + j * 16 + (lane & 15);
            if (n >= a.N) continue;
            const int gr0 = a.r0 + mb16, gc = a.r0 + n;
            const int ti = gr0 / a.nb, tk = gc / a.nb;
            double *c = a.C + tix(ti, tk, a.c_ym, 1) * per + (gr0 - ti * a.nb) +
                        (long long)(gc - tk * a.nb) * a.nb;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = a.rowmap ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                if (mb16 + rr >= a.M || gr0 + rr < gc) continue;
                c[rr] -= (double)acc[i][j][r];
            }
        }
    }
}

// ---- BF16 update ------------------------------------------------------------
constexpr int BK16 = 64;            // bf16 per K step
constexpr int S16 = BK16 + 8;       // LDS row stride (bf16), rows are m

__global__ __launch_bounds__(NT, 2) void k_update_bf16(LpArgs a) {
    __shared__ unsigned short As[2][BM * S16];
    __shared__ unsigned short Bs[2][BM * S16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    long long L = xcd_remap(blockIdx.x, a.nblk);
    int mbi, nbi;
    tri_decode(L, a.mb, a.nbk, mbi, nbi);
    const int m0 = mbi * BM, n0 = nbi * BM;
    const unsigned short *X = (const unsigned short *)a.X;
    const long long per = (long long)a.nb * a.nb;
    // row-major tiles: element (r, k) at tile(r)*per + (r % nb)*nb + k.
    // q = tid + 256 s -> row q / 8 (32 rows per pass), 8-wide k chunk q % 8
    const int kc = tid & 7, rr0 = tid >> 3;
    auto rowbase = [&](int r) -> long long {
        int t = r / a.nb;
        return t * per + (long long)(r - t * a.nb) * a.nb;
    };
    long long ba[4], bb_[4];
    bool oka[4], okb[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        int ra_ = m0 + rr0 + 32 * s, rb_ = n0 + rr0 + 32 * s;
        oka[s] = ra_ < a.M;
        okb[s] = rb_ < a.N;
        ba[s] = oka[s] ? rowbase(ra_) : 0;
        bb_[s] = okb[s] ? rowbase(rb_) : 0;
    }
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    s8 ra[4], rb[4];
    const s8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    auto gload = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            ra[s] = oka[s] ? *reinterpret_cast<const s8 *>(X + ba[s] + k0 + 8 * kc) : z;
            rb[s] = okb[s] ? *reinterpret_cast<const s8 *>(X + bb_[s] + k0 + 8 * kc) : z;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            int r = rr0 + 32 * s;
            *reinterpret_cast<s8 *>(&As[buf][r * S16 + 8 * kc]) = ra[s];
            *reinterpret_cast<s8 *>(&Bs[buf][r * S16 + 8 * kc]) = rb[s];
        }
    };
    const int nk = a.K / BK16;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) gload((kt + 1) * BK16);
#pragma unroll
        for (int ks = 0; ks < BK16 / 32; ++ks) {
            const int kof = ks * 32 + 8 * (lane >> 4);
            s8 fa[4], fb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                fa[i] = *reinterpret_cast<const s8 *>(&As[cur][(wm * 64 + i * 16 + (lane & 15)) * S16 + kof]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                fb[j] = *reinterpret_cast<const s8 *>(&Bs[cur][(wn * 64 + j * 16 + (lane & 15)) * S16 + kof]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, fa[i]), __builtin_bit_cast(bf8, fb[j]), acc[i][j], 0, 0, 0);
        }
        if (more) lstore(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int mb16 = m0 + wm * 64 + i * 16;
        const float si = a.scale[mb16 / a.nb];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn * 64 + j * 16 + (lane & 15);
            if (n >= a.N) continue;
            const double s = (double)si * (double)a.scale[n / a.nb];
            const int gr0 = a.r0 + mb16, gc = a.r0 + n;
            const int ti = gr0 / a.nb, tk = gc / a.nb;
            double *c = a.C + tix(ti, tk, a.c_ym, 1) * per + (gr0 - ti * a.nb) +
                        (long long)(gc - tk * a.nb) * a.nb;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = a.rowmap ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                if (mb16 + rr >= a.M || gr0 + rr < gc) continue;
                c[rr] -= s * (double)acc[i][j][r];
            }
        }
    }
}

int g_rowmap[2] = {-1, -1};
std::mutex g_probe_mu;

}  // namespace

// Runs the one-wave probe for kind (0 FP32, 1 BF16) and returns the row
// order (0 or 1), or -1 if the operand/result layouts differ from both.
int mfma_rowmap(int kind, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_probe_mu);
    if (g_rowmap[kind] >= 0) return g_rowmap[kind];
    float *d = nullptr;
    if (hipMalloc(&d, 256 * sizeof(float)) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_probe_mfma, dim3(1), dim3(64), 0, s, kind, d);
    float h[256];
    bool ok = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s) == hipSuccess &&
              hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(d);
    if (!ok) return -1;
    for (int map = 0; map < 2; ++map) {
        bool good = true;
        for (int lane = 0; lane < 64 && good; ++lane)
            for (int r = 0; r < 4 && good; ++r) {
                int row = map ? lane / 16 + 4 * r : 4 * (lane / 16) + r;
                int col = lane % 16;
                good = h[lane * 4 + r] == (float)((row + 1) * (col + 1));
            }
        if (good) {
            g_rowmap[kind] = map;
            return map;
        }
    }
    return -1;
}

hipError_t convert_column(const double *Y, int nb, int ym, int jt, int ntiles, int mode,
                          float *scale, void *out, hipStream_t s) {
    if (mode == 1) {
        hipLaunchKernelGGL(k_tile_absmax, dim3(ntiles), dim3(256), 0, s, Y, nb, ym, jt, ntiles,
                           scale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    long long total = (long long)nb * nb * ntiles;
    int blocks = (int)std::min<long long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_convert_column, dim3(blocks), dim3(256), 0, s, Y, nb, ym, jt, ntiles,
                       mode, (const float *)scale, out);
    return hipGetLastError();
}

hipError_t launch_update_lp(int mode, double *C, int c_ym, int r0, const void *X,
                            const float *scale, int nb, int M, int N, int rowmap,
                            hipStream_t s) {
    if (M <= 0 || N <= 0) return hipSuccess;
    if (nb % BK16 || nb % BM || M % BM || N % BM || r0 % BM) return hipErrorInvalidValue;
    LpArgs a;
    a.C = C;
    a.c_ym = c_ym;
    a.r0 = r0;
    a.X = X;
    a.scale = scale;
    a.nb = nb;
    a.M = M;
    a.N = N;
    a.K = nb;
    a.mb = M / BM;
    a.nbk = N / BM;
    long long n = 0;
    for (int c = 0; c < a.nbk; ++c) n += a.mb - c;
    a.nblk = n;
    a.rowmap = rowmap;
    if (mode == 0)
        hipLaunchKernelGGL(k_update_f32, dim3((unsigned)n), dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL(k_update_bf16, dim3((unsigned)n), dim3(NT), 0, s, a);
    return hipGetLastError();
}

}  // namespace ooc
