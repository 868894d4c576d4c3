// Diagonal-block factorization, triangular-solve apply, padding and the
// on-device PDMATGEN for the out-of-core Cholesky.
//
// The reference factors each diagonal NB x NB tile on the host with MKL
// dpotrf_ (oocquark.cpp:86-87,393-396) and then solves every sub-diagonal
// tile with cublasDtrsm(R,L,T,N) (oocquark.cpp:130,405-408).  Here a whole
// tile column (diagonal tile plus everything below it in the panel) is
// factored on the device by a blocked left-looking sweep over 64-wide block
// columns:
//   1. P[c:, c] -= P[c:, :c] * P[c, :c]^T       (GEMM, 128x64 MFMA tiles)
//   2. L_cc = chol(P[c, c]); D_c = L_cc^-1      (one workgroup, LDS-resident)
//   3. P[c+1:, c] <- P[c+1:, c] * D_c^T          (GEMM apply, 128x64 tiles)
// which is POTRF of the diagonal tile followed by the TRSM of the tiles
// below it, with no host round trip.
#include "ooc_internal.h"

#include <cmath>
#include <functional>

namespace ooc {

namespace {

constexpr int PB = 64;        // diagonal block edge
constexpr int PS = PB + 1;    // LDS column stride (odd: conflict-free column walks)

__device__ __forceinline__ long long opnd_off(const Opnd &o, int r, int c) {
    int gr = o.r0 + r, gc = o.c0 + c;
    if (o.ts <= 0) return gr + (long long)gc * o.ld;
    int ti = gr / o.ts, tk = gc / o.ts;
    long long t = ti + (long long)tk * o.ym - (o.tri ? ((long long)tk * (tk + 1)) / 2 : 0);
    return t * o.te + (gr - ti * o.ts) + (long long)(gc - tk * o.ts) * o.ld;
}

// Lane l's value of v, broadcast to the whole wave (two v_readlane_b32 into
// scalar registers: no LDS round trip).
__device__ __forceinline__ double lane_value(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// Columns [K0, K1) of the trailing update of factorization step J: lane i
// subtracts l_ij * l_kj from its a_ik, with l_kj read from lane k.
template <int J, int K0, int K1>
__device__ __forceinline__ void chol_cols(double (&a)[PB]) {
    if constexpr (K1 > K0) {
        double cv[K1 - K0];
#pragma unroll
        for (int k = K0; k < K1; ++k) cv[k - K0] = lane_value(a[J], k);
#pragma unroll
        for (int k = K0; k < K1; ++k) {
            a[k] -= a[J] * cv[k - K0];
            // pin the update here: otherwise the compiler sinks it to the
            // next read of a[k] and keeps every earlier column alive
            asm volatile("" : "+v"(a[k]));
        }
    }
    __builtin_amdgcn_sched_barrier(0);   // bounded register use
}

// Step J of the register-resident factorization (template recursion keeps
// every register index a compile-time constant).  bad: first failed pivot
// (1-based), uniform.
template <int J>
__device__ __forceinline__ void chol_step(double (&a)[PB], int lane, int &bad) {
    if constexpr (J < PB) {
        // an opaque copy of the lane index keeps the compiler from forming
        // all 64 steps' lane masks up front (they would not fit in SGPRs)
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const double d = lane_value(a[J], J);
        if (bad == 0 && !(d > 0.0)) bad = J + 1;
        const double sd = sqrt(d);
        const double rd = 1.0 / sd;
        // lanes above the pivot keep 0 so later steps leave them unchanged
        a[J] = ln > J ? a[J] * rd : (ln == J ? sd : 0.0);
        __builtin_amdgcn_sched_barrier(0);
        chol_cols<J, J + 1, (J + 1 > 16 ? J + 1 : 16)>(a);
        chol_cols<J, (J + 1 > 16 ? J + 1 : 16), (J + 1 > 32 ? J + 1 : 32)>(a);
        chol_cols<J, (J + 1 > 32 ? J + 1 : 32), (J + 1 > 48 ? J + 1 : 48)>(a);
        chol_cols<J, (J + 1 > 48 ? J + 1 : 48), PB>(a);
        chol_step<J + 1>(a, lane, bad);
    }
}

// Rows [I0, I1) of step K of the inversion: l_iK is lane i's a[K].
template <int K, int I0, int I1>
__device__ __forceinline__ void inv_rows(double (&y)[PB], const double (&a)[PB]) {
    if constexpr (I1 > I0) {
        double lv[I1 - I0];
#pragma unroll
        for (int i = I0; i < I1; ++i) lv[i - I0] = lane_value(a[K], i);
#pragma unroll
        for (int i = I0; i < I1; ++i) y[i] -= lv[i - I0] * y[K];
    }
    __builtin_amdgcn_sched_barrier(0);
}

// Step K of the inversion: lane c holds column c of L^-1 being formed;
// column K of L comes from the lanes' registers (row i in lane i).
template <int K>
__device__ __forceinline__ void inv_step(double (&y)[PB], const double (&a)[PB]) {
    if constexpr (K < PB) {
        y[K] *= 1.0 / lane_value(a[K], K);
        inv_rows<K, K + 1, (K + 1 > 16 ? K + 1 : 16)>(y, a);
        inv_rows<K, (K + 1 > 16 ? K + 1 : 16), (K + 1 > 32 ? K + 1 : 32)>(y, a);
        inv_rows<K, (K + 1 > 32 ? K + 1 : 32), (K + 1 > 48 ? K + 1 : 48)>(y, a);
        inv_rows<K, (K + 1 > 48 ? K + 1 : 48), PB>(y, a);
        inv_step<K + 1>(y, a);
    }
}

// One wave64 factors a 64 x 64 diagonal block and inverts the triangle,
// register-resident.  Factorization (right-looking, lane i owns row i in
// registers): at step j the pivot and column j are read from the owning
// lanes' registers (v_readlane), column j is scaled by 1/l_jj (as LAPACK
// dpotf2 scales by the reciprocal) and every lane subtracts l_ij * l_kj from
// its a_ik, k > j -- the same subtraction order per element as the
// left-looking dot products.  Inversion (lane c owns column c of L^-1):
// y = e_c, then for k = 0..63 y_k *= 1/l_kk and y_i -= l_ik y_k for i > k,
// l_ik read from lane i.  LDS only transposes L^-1 for the coalesced store.
// do_factor = 0: the block already holds L; only the inverse is formed and
// the block is left untouched.
__global__ __launch_bounds__(64) void k_potrf64(Opnd P, int gcol0, double *dinv, int *info,
                                                int do_factor) {
    __shared__ double S[PB * PS];   // L^-1, column-major, column stride PS
    const int lane = threadIdx.x;
    const long long base = opnd_off(P, 0, 0);  // block never straddles a tile
    const long long ld = P.ld;
    double *g = P.p + base;

    double a[PB];
#pragma unroll
    for (int c = 0; c < PB; ++c) {
        const double v = g[lane + c * ld];   // whole block is in bounds
        a[c] = lane >= c ? v : 0.0;
    }

    if (do_factor) {
        int bad = 0;
        chol_step<0>(a, lane, bad);
        if (bad && lane == 0) atomicMin(info, gcol0 + bad);
        // opaque base: the write-back recomputes its addresses instead of
        // keeping the 64 load addresses alive through the factorization
        double *gs = g;
        asm volatile("" : "+s"(gs));
#pragma unroll
        for (int c = 0; c < PB; ++c)
            if (lane >= c) gs[lane + c * ld] = a[c];
    }

    double y[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) y[i] = (i == lane) ? 1.0 : 0.0;
    inv_step<0>(y, a);
#pragma unroll
    for (int i = 0; i < PB; ++i) S[lane * PS + i] = y[i];   // column lane of L^-1
    __syncthreads();
#pragma unroll
    for (int c = 0; c < PB; ++c) dinv[lane + c * PB] = S[c * PS + lane];
}

__global__ void k_pad_tile(double *t, int nb, int vr, int vc, int diag) {
    long long n = (long long)nb * nb;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        int r = (int)(idx % nb), c = (int)(idx / nb);
        if (r >= vr || c >= vc) t[idx] = (diag && r == c) ? 1.0 : 0.0;
    }
}

// ---- PDMATGEN on the device ----------------------------------------------
// Stream position of element (i, j) of the symmetric form (LIN/pdmatgen.f:
// 248-379, 1x1 grid): min(i,j)*N + max(i,j) steps past x1 = a*ISEED + c.
// The per-element value is 1 - 2*x/2^31 (:261); DIAG 'D' turns the diagonal
// into |a_ii| + N (:516-543); AFORM 'F' is 10 on the diagonal and the REAL
// literal 0.1 elsewhere (:546-626).
struct LcgMap {
    unsigned a, c;
};
struct GenArgs {
    LcgMap pow2[40];   // 2^k-step maps
    LcgMap stepN;      // N-step map
    unsigned x1;
    double *Y;
    int N, nb, ts, bb, jt0, jt1;   // logical tile edge nb, slot edge ts >= nb
    int symm_F;        // 1: AFORM 'F'
    int diagD;
    long long ntiles;  // tiles in the column range
    int runs;          // column runs per tile
    long long tbase;   // packed index of the tile at Y (a window of the layout)
};

constexpr int GEN_RUN = 32;

__device__ __forceinline__ unsigned lcg_apply(LcgMap m, unsigned x) {
    return (m.a * x + m.c) & 0x7fffffffu;
}

// One thread per slot row r of a run of GEN_RUN slot columns of one tile.
// Element (r, c) of tile (ti, tj) is matrix element (ti*nb + r, tj*nb + c)
// when r, c < nb and both indices are below N; every other slot entry is
// padding.  Consecutive valid columns of a row advance the LCG by one
// N-step (lower triangle, stream position j*N + i) or one step (upper).
__global__ __launch_bounds__(256) void k_gen_tiles(GenArgs g) {
    const int chunks = (g.ts + 255) / 256;
    long long b = blockIdx.x;
    int run = (int)(b % g.runs);
    b /= g.runs;
    int chunk = (int)(b % chunks);
    long long t = b / chunks;
    if (t >= g.ntiles) return;
    int tj = g.jt0;
    long long rem = t;
    while (rem >= (long long)(g.bb - tj)) {
        rem -= (g.bb - tj);
        ++tj;
    }
    int ti = tj + (int)rem;
    int r = chunk * 256 + threadIdx.x;
    if (r >= g.ts) return;
    long long tidx = ti + (long long)tj * g.bb - ((long long)tj * (tj + 1)) / 2;
    double *tile = g.Y + (tidx - g.tbase) * (long long)g.ts * g.ts;
    const int gi = ti * g.nb + r;
    const bool row_ok = r < g.nb && gi < g.N;
    const int cbeg = run * GEN_RUN;
    const int cend = min(cbeg + GEN_RUN, g.ts);
    const double c01 = (double)0.1f;
    const double fN = (double)g.N;
    int gj = tj * g.nb + cbeg;
    unsigned x = 0;
    bool have = false;
    for (int c = cbeg; c < cend; ++c, ++gj) {
        double v;
        if (!row_ok || c >= g.nb || gj >= g.N) {
            v = (ti == tj && r == c) ? 1.0 : 0.0;
        } else if (g.symm_F) {
            v = (gi == gj) ? 10.0 : c01;
        } else {
            if (!have) {
                unsigned long long p = (gi >= gj) ? (unsigned long long)gj * g.N + gi
                                                  : (unsigned long long)gi * g.N + gj;
                x = g.x1;
                for (int k = 0; p; ++k, p >>= 1)
                    if (p & 1ull) x = lcg_apply(g.pow2[k], x);
                have = true;
            } else {
                x = (gi >= gj) ? lcg_apply(g.stepN, x) : lcg_apply(g.pow2[0], x);
            }
            v = 1.0 - 2.0 * ((double)x / 2147483648.0);
            if (gi == gj && g.diagD) v = fabs(v) + fN;
        }
        tile[r + (long long)c * g.ts] = v;
    }
}

LcgMap lcg_then(LcgMap f, LcgMap g) {  // g after f
    LcgMap r;
    r.a = (g.a * f.a) & 0x7fffffffu;
    r.c = (g.a * f.c + g.c) & 0x7fffffffu;
    return r;
}

LcgMap lcg_pow(unsigned long long n) {
    LcgMap r{1u, 0u}, b{1103515245u, 12345u};
    while (n) {
        if (n & 1ull) r = lcg_then(r, b);
        b = lcg_then(b, b);
        n >>= 1;
    }
    return r;
}

}  // namespace

hipError_t launch_potrf64(const Opnd &P, int gcol0, double *dinv, int *info, hipStream_t s) {
    hipLaunchKernelGGL(k_potrf64, dim3(1), dim3(64), 0, s, P, gcol0, dinv, info,
                       info != nullptr ? 1 : 0);
    return hipGetLastError();
}

hipError_t factor_column(const Opnd &P, int row0, int mrows, int ncols, int gcol_base,
                         double *dinv_ws, int *info, hipStream_t s) {
    if (ncols % PB) return hipErrorInvalidValue;
    const int nblk = ncols / PB;
    for (int c = 0; c < nblk; ++c) {
        const int r = row0 + c * PB;
        hipError_t e;
        if (c > 0) {
            Gemm g;
            g.A = P; g.A.r0 = r; g.A.c0 = row0;
            g.B = g.A;
            g.C = P; g.C.r0 = r; g.C.c0 = r;
            g.M = mrows - c * PB;
            g.N = PB;
            g.K = c * PB;
            g.alpha = -1.0;
            g.beta = 1.0;
            g.lower = 1;
            if ((e = launch_gemm(g, GEMM_NARROW, s)) != hipSuccess) return e;
        }
        Opnd D = P;
        D.r0 = r;
        D.c0 = r;
        double *dinv = dinv_ws + (size_t)c * PB * PB;
        if ((e = launch_potrf64(D, gcol_base + c * PB, dinv, info, s)) != hipSuccess) return e;
        const int below = mrows - (c + 1) * PB;
        if (below > 0) {
            Gemm g;
            g.A = P; g.A.r0 = r + PB; g.A.c0 = r;
            g.C = g.A;
            g.B = plain(dinv, PB);
            g.M = below;
            g.N = PB;
            g.K = PB;
            g.alpha = 1.0;
            g.beta = 0.0;
            g.lower = 0;
            if ((e = launch_gemm(g, GEMM_APPLY, s)) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

int factor_panel_la(const Opnd &P, int NB, int Ym, int Yn, int gcol0, double *dinv,
                    size_t dinv_stride, int *dinfo, hipStream_t s0, hipStream_t s1,
                    PanelUpdater &up, const std::function<int(int)> &on_column,
                    const std::function<int(int)> *after_updates) {
    hipEvent_t ev_fact, ev_bulk[2];
    if (hipEventCreateWithFlags(&ev_fact, hipEventDisableTiming) != hipSuccess) return -4;
    if (hipEventCreateWithFlags(&ev_bulk[0], hipEventDisableTiming) != hipSuccess) return -4;
    if (hipEventCreateWithFlags(&ev_bulk[1], hipEventDisableTiming) != hipSuccess) return -4;
    struct Cleanup {
        hipEvent_t *e;
        ~Cleanup() {
            (void)hipEventDestroy(e[0]);
            (void)hipEventDestroy(e[1]);
            (void)hipEventDestroy(e[2]);
        }
    };
    hipEvent_t all[3] = {ev_fact, ev_bulk[0], ev_bulk[1]};
    Cleanup cl{all};
    int rc = 0;
    bool any_bulk = false;
#define LA_TRY(x)                       \
    do {                                \
        if ((x) != hipSuccess) return -4; \
    } while (0)
    for (int j = 0; j < Yn; ++j) {
        // column j received bulk(j-2) on s1
        if (j >= 2) LA_TRY(hipStreamWaitEvent(s0, ev_bulk[j & 1], 0));
        LA_TRY(factor_column(P, j * NB, (Ym - j) * NB, NB, gcol0 + j * NB,
                             dinv + (size_t)j * dinv_stride, dinfo, s0));
        if ((rc = on_column(j)) != 0) return rc;
        if (j + 1 >= Yn) {
            if (after_updates && (rc = (*after_updates)(j)) != 0) return rc;
            break;
        }
        LA_TRY(up.prepare(j, s0));
        // column j+1 also received bulk(j-1)
        if (j >= 1) LA_TRY(hipStreamWaitEvent(s0, ev_bulk[(j - 1) & 1], 0));
        LA_TRY(up.update(j, j + 1, 1, s0));
        if (j + 2 < Yn) {
            LA_TRY(hipEventRecord(ev_fact, s0));
            LA_TRY(hipStreamWaitEvent(s1, ev_fact, 0));
            LA_TRY(up.update(j, j + 2, Yn - j - 2, s1));
            LA_TRY(hipEventRecord(ev_bulk[j & 1], s1));
            any_bulk = true;
        }
        if (after_updates && (rc = (*after_updates)(j)) != 0) return rc;
    }
    if (any_bulk) {
        LA_TRY(hipEventRecord(ev_fact, s1));
        LA_TRY(hipStreamWaitEvent(s0, ev_fact, 0));
    }
#undef LA_TRY
    return 0;
}

hipError_t launch_pad_tile(double *tile, int nb, int vr, int vc, int diag, hipStream_t s) {
    if (vr >= nb && vc >= nb) return hipSuccess;
    long long n = (long long)nb * nb;
    int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_pad_tile, dim3(blocks), dim3(256), 0, s, tile, nb, vr, vc, diag);
    return hipGetLastError();
}

hipError_t launch_gen_tiles(double *Y, int N, int nb, int bb, int jt0, int jt1, char aform,
                            char diag, int iseed, hipStream_t s, long long tbase) {
    if (jt1 <= jt0) return hipSuccess;
    GenArgs g;
    g.tbase = tbase;
    for (int k = 0; k < 40; ++k) g.pow2[k] = lcg_pow(1ull << k);
    g.stepN = lcg_pow((unsigned long long)N);
    g.x1 = (1103515245u * ((unsigned)iseed & 0x7fffffffu) + 12345u) & 0x7fffffffu;
    g.Y = Y;
    g.N = N;
    g.nb = nb;
    g.ts = tile_edge(nb);
    g.bb = bb;
    g.jt0 = jt0;
    g.jt1 = jt1;
    char af = (char)(aform & ~0x20), dg = (char)(diag & ~0x20);
    g.symm_F = (af == 'F');
    g.diagD = (dg == 'D');
    long long nt = 0;
    for (int j = jt0; j < jt1; ++j) nt += bb - j;
    g.ntiles = nt;
    g.runs = (g.ts + GEN_RUN - 1) / GEN_RUN;
    const int chunks = (g.ts + 255) / 256;
    long long blocks = nt * chunks * g.runs;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gen_tiles, dim3((unsigned)blocks), dim3(256), 0, s, g);
    return hipGetLastError();
}

}  // namespace ooc
