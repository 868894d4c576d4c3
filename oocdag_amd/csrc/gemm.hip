// FP64 MFMA GEMM-NT for the Cholesky panel updates on CDNA4 (gfx950).
//
//   C = beta*C + alpha * A * B^T       (A: M x K, B: N x K, C: M x N)
//
// This one kernel family carries every dense update of the factorization:
//  * the left-looking update of a Y panel from a staged X tile column
//    (ooc_syrk, oocquark.cpp:321-356: SYRK on diagonal tiles, GEMM below),
//  * the right-looking update inside a panel (oocquark.cpp:413-418),
//  * the updates and triangular-solve applies of the blocked diagonal-tile
//    factorization (oocquark.cpp:393-410).
// The reference issued one cuBLAS call per NB x NB tile
// (oocquark.cpp:130,176,222); here one launch covers the whole lower
// trapezoid of a panel, and SYRK is the same kernel with blocks above the
// diagonal skipped and the diagonal blocks stored lower-only.
//
// Work decomposition: 256 threads (4 wave64s as 2 x 2) per 128 x BN output
// block; each wave owns a 64 x (BN/2) sub-block held as 16x16 FP64
// accumulators of v_mfma_f64_16x16x4_f64.  Whole 128 x 128 blocks (every
// panel update) load their operands straight into MFMA registers, 3 K slabs
// ahead, with no LDS and no barriers (k_dgemm_dtv); ragged and 64-wide
// blocks step K 16 at a time through double-buffered LDS (k_dgemm_nt).
// Block order is remapped so that each XCD (8 per MI355X, round-robin
// dispatch) receives a contiguous range of blocks that share operand rows in
// its own L2.
#include "ooc_internal.h"

namespace ooc {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

namespace {

constexpr int BM = 128;
constexpr int BK = 16;
constexpr int NT = 256;
constexpr int LPAD = 16;  // doubles of padding per LDS row (row stride = 16 mod 32 doubles)

struct KArgs {
    Opnd A, B, C;
    int M, N, K;
    double alpha, beta;
    int lower;
    int mb, nb;       // block-grid extent
    long long nblk;   // blocks launched
};

__device__ __forceinline__ long long tix(int ti, int tk, int ym, int tri) {
    return ti + (long long)tk * ym - (tri ? ((long long)tk * (tk + 1)) / 2 : 0);
}

// Streams one operand's rows [rb + 2*rp, +2) along K for this thread.
template <int BX>
struct Stream {
    static constexpr int PPC = BX / 2;       // row pairs per column
    static constexpr int CPP = NT / PPC;     // columns per pass
    static constexpr int NPASS = BK / CPP;   // passes per K step

    const double *p;
    long long ld, te, tbase, roff;
    int ts, ym, tri, ti, tk, kc, col0;
    bool ok0, ok1;

    __device__ __forceinline__ void init(const Opnd &o, int rb, int rows, int tid) {
        int rp = tid % PPC;
        col0 = tid / PPC;
        int r = rb + 2 * rp;
        ok0 = r < rows;
        ok1 = (r + 1) < rows;
        p = o.p;
        ld = o.ld;
        te = o.te;
        ym = o.ym;
        tri = o.tri;
        int gr = o.r0 + r;
        int gk = o.c0;
        if (o.ts > 0) {
            ts = o.ts;
            ti = gr / ts;
            roff = gr - (long long)ti * ts;
            tk = gk / ts;
            kc = gk - tk * ts;
            tbase = tix(ti, tk, ym, tri) * te + roff;
        } else {
            ts = 0x40000000;
            ti = 0;
            roff = gr;
            tk = 0;
            kc = gk;
            tbase = roff;
        }
    }

    __device__ __forceinline__ void load(d2 (&v)[NPASS], int k0, int K) const {
#pragma unroll
        for (int s = 0; s < NPASS; ++s) {
            int c = col0 + s * CPP;
            d2 x = {0.0, 0.0};
            if (ok0 && (k0 + c) < K) {
                const double *q = p + tbase + (long long)(kc + c) * ld;
                if (ok1) {
                    x = *reinterpret_cast<const d2 *>(q);
                } else {
                    x.x = q[0];
                }
            }
            v[s] = x;
        }
    }

    __device__ __forceinline__ void advance() {
        kc += BK;
        if (kc >= ts) {
            kc -= ts;
            ++tk;
            tbase = tix(ti, tk, ym, tri) * te + roff;
        }
    }

    __device__ __forceinline__ void store_lds(double *lds, int stride, const d2 (&v)[NPASS],
                                              int tid) const {
        int rp = tid % PPC;
#pragma unroll
        for (int s = 0; s < NPASS; ++s) {
            int c = col0 + s * CPP;
            *reinterpret_cast<d2 *>(lds + c * stride + 2 * rp) = v[s];
        }
    }
};

// Streams a transposed operand: operand element (r, k) is element
// (o.r0 + k, o.c0 + r) of the stored matrix o, so K runs down stored
// columns.  Each thread loads a pair of consecutive k (16 bytes) for
// BX/32 operand rows per step and scatters them into the k-major LDS tile.
template <int BX>
struct StreamT {
    static constexpr int KP = BK / 2;       // k pairs per step
    static constexpr int RPP = NT / KP;     // operand rows per pass
    static constexpr int NPASS = BX / RPP;  // passes per K step

    const double *p;
    long long ld, te;
    long long coloff[NPASS];
    int tkc[NPASS];
    bool ok[NPASS];
    int ts, ym, tri, kp, rr, gk;

    __device__ __forceinline__ void init(const Opnd &o, int rb, int rows, int tid) {
        kp = tid % KP;
        rr = tid / KP;
        p = o.p;
        ld = o.ld;
        te = o.te;
        ym = o.ym;
        tri = o.tri;
        ts = o.ts > 0 ? o.ts : 0x40000000;
        gk = o.r0;
#pragma unroll
        for (int s = 0; s < NPASS; ++s) {
            int r = rb + rr + s * RPP;
            ok[s] = r < rows;
            int gc = o.c0 + r;
            tkc[s] = gc / ts;
            coloff[s] = (long long)(gc - tkc[s] * ts) * ld;
        }
    }

    __device__ __forceinline__ void load(d2 (&v)[NPASS], int k0, int K) const {
        const int k = k0 + 2 * kp;
        const int g = gk + 2 * kp;
        const int ti = g / ts;
        const long long roff = g - (long long)ti * ts;
#pragma unroll
        for (int s = 0; s < NPASS; ++s) {
            d2 x = {0.0, 0.0};
            if (ok[s] && k < K) {
                const double *q = p + tix(ti, tkc[s], ym, tri) * te + roff + coloff[s];
                if (k + 1 < K)
                    x = *reinterpret_cast<const d2 *>(q);
                else
                    x.x = q[0];
            }
            v[s] = x;
        }
    }

    __device__ __forceinline__ void advance() { gk += BK; }

    __device__ __forceinline__ void store_lds(double *lds, int stride, const d2 (&v)[NPASS],
                                              int) const {
#pragma unroll
        for (int s = 0; s < NPASS; ++s) {
            int r = rr + s * RPP;
            lds[(2 * kp) * stride + r] = v[s].x;
            lds[(2 * kp + 1) * stride + r] = v[s].y;
        }
    }
};

// Full-block operand stream: the block's BX rows lie inside one tile (or
// the operand is a plain matrix) and K is a multiple of BK, so the address
// of the block's current K step is uniform across the workgroup (scalar
// registers) and each thread adds a fixed 32-bit offset.  No bounds checks.
template <int BX>
struct StreamF {
    static constexpr int PPC = BX / 2;
    static constexpr int CPP = NT / PPC;
    static constexpr int NPASS = BK / CPP;

    const double *p, *base;
    long long ld, te, rowoff;
    int ts, ym, tri, ti, tk, kc;
    int voff, rp, col0;

    __device__ __forceinline__ void init(const Opnd &o, int rb, int, int tid) {
        rp = tid % PPC;
        col0 = tid / PPC;
        p = o.p;
        ld = o.ld;
        te = o.te;
        ym = o.ym;
        tri = o.tri;
        const int gr = o.r0 + rb;
        if (o.ts > 0) {
            ts = o.ts;
            ti = gr / ts;
            rowoff = gr - (long long)ti * ts;
            tk = o.c0 / ts;
            kc = o.c0 - tk * ts;
        } else {
            ts = 0x40000000;
            ti = 0;
            rowoff = gr;
            tk = 0;
            kc = o.c0;
        }
        base = p + tix(ti, tk, ym, tri) * te + rowoff + (long long)kc * ld;
        voff = 2 * rp + col0 * (int)ld;
    }

    __device__ __forceinline__ void load(d2 (&v)[NPASS], int, int) const {
#pragma unroll
        for (int s = 0; s < NPASS; ++s)
            v[s] = *reinterpret_cast<const d2 *>(base + (voff + s * CPP * (int)ld));
    }

    __device__ __forceinline__ void advance() {
        kc += BK;
        if (kc >= ts) {
            kc -= ts;
            ++tk;
            base = p + tix(ti, tk, ym, tri) * te + rowoff + (long long)kc * ld;
        } else {
            base += BK * ld;
        }
    }

    __device__ __forceinline__ void store_lds(double *lds, int stride, const d2 (&v)[NPASS],
                                              int) const {
#pragma unroll
        for (int s = 0; s < NPASS; ++s)
            *reinterpret_cast<d2 *>(lds + (col0 + s * CPP) * stride + 2 * rp) = v[s];
    }
};

template <int BX, int TR, int FULL>
struct StreamSel {
    typedef Stream<BX> type;
};
template <int BX>
struct StreamSel<BX, 0, 1> {
    typedef StreamF<BX> type;
};
template <int BX, int FULL>
struct StreamSel<BX, 1, FULL> {
    typedef StreamT<BX> type;
};

// Blocks in rows [0, R) of a lower trapezoid with Nb block columns.
__device__ __forceinline__ long long trap_rows(long long R, int Nb) {
    return R <= Nb ? R * (R + 1) / 2 : (long long)Nb * (Nb + 1) / 2 + (R - Nb) * Nb;
}

// Linear block id -> (mb, nb), in groups of GM block rows, column-major
// inside a group, so the blocks an XCD runs at once form a GM x GM patch
// that shares GM operand strips of each kind in its L2.
// MODE 0: rectangular grid.  MODE 1: lower trapezoid (mb >= nb).
template <int MODE>
__device__ __forceinline__ void decode_block(long long L, int Mb, int Nb, int &mb, int &nb) {
    const int GM = 8;
    if (MODE == 0) {
        long long per_group = (long long)GM * Nb;
        int g = (int)(L / per_group);
        int first = g * GM;
        int gsz = min(Mb - first, GM);
        int w = (int)(L - (long long)g * per_group);
        mb = first + w % gsz;
        nb = w / gsz;
    } else {
        // group g holds rows [g GM, g GM + gs); trap_rows(g GM) blocks precede it
        int lo = 0, hi = (Mb + GM - 1) / GM - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (trap_rows((long long)mid * GM, Nb) <= L) lo = mid; else hi = mid - 1;
        }
        const int R = lo * GM, gs = min(GM, Mb - R);
        long long w = L - trap_rows(R, Nb);
        const int n1 = min(R + 1, Nb);   // columns 0..n1-1: all gs rows of the group
        if (w < (long long)n1 * gs) {
            nb = (int)(w / gs);
            mb = R + (int)(w - (long long)nb * gs);
        } else {
            w -= (long long)n1 * gs;     // column n > R: rows n .. R+gs-1
            int n = n1;
            while (w >= R + gs - n) {
                w -= R + gs - n;
                ++n;
            }
            nb = n;
            mb = n + (int)w;
        }
    }
}

template <int BMT, int BN, int MODE, int KIND, int TB, int FULL>
__global__ __launch_bounds__(NT, 2) void k_dgemm_nt(KArgs a) {
    constexpr int BM = BMT;         // block rows (128, or 32 for few right-hand sides)
    constexpr int WM = BM / 2;      // wave tile rows
    constexpr int WN = BN / 2;      // wave tile columns
    constexpr int FM = WM / 16;
    constexpr int FN = WN / 16;
    constexpr int SA = BM + LPAD;   // LDS row strides (doubles)
    constexpr int SB = BN + LPAD;

    __shared__ double As[2][BK * SA];
    __shared__ double Bs[2][BK * SB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    long long L = xcd_remap(blockIdx.x, a.nblk);
    int mbi, nbi;
    decode_block<MODE>(L, a.mb, a.nb, mbi, nbi);
    const int m0 = mbi * BM, n0 = nbi * BN;
    if (m0 >= a.M || n0 >= a.N) return;
    if (a.lower && (a.C.r0 + m0 + BM - 1) < (a.C.c0 + n0)) return;

    typedef typename StreamSel<BM, 0, FULL>::type SA_t;
    typedef typename StreamSel<BN, TB, FULL>::type SB_t;
    SA_t sa;
    SB_t sb;
    sa.init(a.A, m0, a.M, tid);
    sb.init(a.B, n0, a.N, tid);

    d4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    d2 ra[SA_t::NPASS], rb[SB_t::NPASS];
    const int nk = (a.K + BK - 1) / BK;

    if (nk > 0) {
        sa.load(ra, 0, a.K);
        sb.load(rb, 0, a.K);
        sa.store_lds(As[0], SA, ra, tid);
        sb.store_lds(Bs[0], SB, rb, tid);
        sa.advance();
        sb.advance();
    }
    __syncthreads();

    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = (kt + 1) < nk;
        if (more) {
            sa.load(ra, (kt + 1) * BK, a.K);
            sb.load(rb, (kt + 1) * BK, a.K);
        }
        const double *as = As[cur];
        const double *bs = Bs[cur];
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            double fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) fa[i] = as[kk * SA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < FN; ++j) fb[j] = bs[kk * SB + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            sa.store_lds(As[cur ^ 1], SA, ra, tid);
            sb.store_lds(Bs[cur ^ 1], SB, rb, tid);
            sa.advance();
            sb.advance();
        }
        __syncthreads();
    }

    // Epilogue.  v_mfma_f64_16x16x4_f64 leaves element r (0..3) of a lane's
    // accumulator at row (lane/16) + 4*r, column lane%16 of its 16x16 block
    // (checked on gfx950 against a CPU GEMM).
    if constexpr (FULL != 0) {
        // the whole BM x BN block lies in one tile of C (or C is plain)
        const Opnd &Co = a.C;
        const int gr0 = Co.r0 + m0, gc0 = Co.c0 + n0;
        const double *cb;
        if (Co.ts > 0) {
            const int ti = gr0 / Co.ts, tk = gc0 / Co.ts;
            cb = Co.p + tix(ti, tk, Co.ym, Co.tri) * Co.te + (gr0 - ti * Co.ts) +
                 (long long)(gc0 - tk * Co.ts) * Co.ld;
        } else {
            cb = Co.p + gr0 + (long long)gc0 * Co.ld;
        }
        const bool diag = a.lower && gr0 < gc0 + BN;
        const int ldc = (int)Co.ld;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int rl = wm * WM + i * 16 + (lane >> 4);
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int cl = wn * WN + j * 16 + (lane & 15);
                double *c = const_cast<double *>(cb) + (rl + cl * ldc);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (diag && gr0 + rl + 4 * r < gc0 + cl) continue;
                    double v = a.alpha * acc[i][j][r];
                    if (a.beta != 0.0) v += a.beta * c[4 * r];
                    c[4 * r] = v;
                }
            }
        }
        return;
    }
    const Opnd &Co = a.C;
    const int cts = Co.ts > 0 ? Co.ts : 0x40000000;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int mb16 = m0 + wm * WM + i * 16;          // first row of the 16x16 block
        const int m = mb16 + (lane >> 4);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn * WN + j * 16 + (lane & 15);
            if (n >= a.N || m >= a.M) continue;
            // the 16 rows of the block lie in one tile (row offsets are
            // multiples of 16 and tiles are multiples of 16 tall)
            const int gr = Co.r0 + mb16, gc = Co.c0 + n;
            long long off;
            if (Co.ts > 0) {
                int ti = gr / cts, tk = gc / cts;
                off = tix(ti, tk, Co.ym, Co.tri) * Co.te + (gr - ti * cts) +
                      (long long)(gc - tk * cts) * Co.ld;
            } else {
                off = gr + (long long)gc * Co.ld;
            }
            double *c = Co.p + off + (lane >> 4);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m + 4 * r;
                if (row >= a.M) break;
                if (a.lower && (Co.r0 + row) < gc) continue;
                double v = a.alpha * acc[i][j][r];
                if (a.beta != 0.0) v += a.beta * c[4 * r];
                c[4 * r] = v;
            }
        }
    }
}

// ---- full 128 x 128 blocks: operands straight into MFMA registers ---------
// Used when every block of C is a whole 128 x 128 block inside one tile and
// K is a multiple of 16.  No LDS and no barriers: each wave (wm, wn) owns a
// 64 x 64 block of C and loads its own operands.  Lane (r = l%16, k = l/16)
// loads 4 consecutive rows 4r..4r+3 of its 64 A rows (two 16-byte loads) at
// column k of the current 4-wide K slab, and the same for B, so MFMA block t
// of the wave holds the rows 4R + t (R = row inside the 16 x 16 MFMA block):
// a row permutation that the epilogue undoes.  A ring of 4 slabs keeps
// DPD slabs of loads in flight ahead of the MFMAs.
constexpr int DPD = 3;

// One operand's uniform slab address along K (crossing tiles of a packed
// operand) plus this lane's fixed offset.
struct DStream {
    const double *p, *base;
    long long ld, te, rowoff;
    int ts, ym, tri, ti, tk, kc, voff;

    __device__ __forceinline__ void init(const Opnd &o, int rb, int lane_off) {
        p = o.p;
        ld = o.ld;
        te = o.te;
        ym = o.ym;
        tri = o.tri;
        const int gr = o.r0 + rb;
        if (o.ts > 0) {
            ts = o.ts;
            ti = gr / ts;
            rowoff = gr - (long long)ti * ts;
            tk = o.c0 / ts;
            kc = o.c0 - tk * ts;
        } else {
            ts = 0x40000000;
            ti = 0;
            rowoff = gr;
            tk = 0;
            kc = o.c0;
        }
        base = p + tix(ti, tk, ym, tri) * te + rowoff + (long long)kc * ld;
        voff = lane_off;
    }
    __device__ __forceinline__ void load(d2 &x0, d2 &x1) const {
        const double *q = base + voff;
        x0 = *reinterpret_cast<const d2 *>(q);
        x1 = *reinterpret_cast<const d2 *>(q + 2);
    }
    // FLAT: the K range never leaves the current tile (no tile crossing)
    template <bool FLAT>
    __device__ __forceinline__ void advance(bool go) {
        if (FLAT) {
            base += go ? 4 * ld : 0;
            return;
        }
        if (!go) return;
        kc += 4;
        if (kc >= ts) {
            kc -= ts;
            ++tk;
            base = p + tix(ti, tk, ym, tri) * te + rowoff + (long long)kc * ld;
        } else {
            base += 4 * ld;
        }
    }
};

template <int MODE, int KIND, bool FLAT>
__global__ __launch_bounds__(NT, 2) void k_dgemm_dtv(KArgs a) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    long long L = xcd_remap(blockIdx.x, a.nblk);
    int mbi, nbi;
    decode_block<MODE>(L, a.mb, a.nb, mbi, nbi);
    const int m0 = mbi * 128, n0 = nbi * 128;
    if (a.lower && (a.C.r0 + m0 + 127) < (a.C.c0 + n0)) return;

    const int r = lane & 15, kq = lane >> 4;
    DStream la, lb;
    la.init(a.A, m0, wm * 64 + 4 * r + kq * (int)a.A.ld);
    lb.init(a.B, n0, wn * 64 + 4 * r + kq * (int)a.B.ld);

    d4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    d2 ra[4][2], rb[4][2];
    const int ns = a.K >> 2;   // K slabs (a multiple of 4, at least 4)
#pragma unroll
    for (int q = 0; q < DPD; ++q) {
        la.load(ra[q][0], ra[q][1]);
        lb.load(rb[q][0], rb[q][1]);
        la.advance<FLAT>(true);
        lb.advance<FLAT>(true);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int s0 = 0; s0 < ns; s0 += 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // the load stream is at slab min(s + DPD, ns - 1)
            const int nx = (q + DPD) & 3;
            la.load(ra[nx][0], ra[nx][1]);
            lb.load(rb[nx][0], rb[nx][1]);
            const bool go = s0 + q + DPD + 1 < ns;
            la.advance<FLAT>(go);
            lb.advance<FLAT>(go);
            __builtin_amdgcn_sched_barrier(0);
            const double fa[4] = {ra[q][0].x, ra[q][0].y, ra[q][1].x, ra[q][1].y};
            const double fb[4] = {rb[q][0].x, rb[q][0].y, rb[q][1].x, rb[q][1].y};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // acc[t][u][e] (MFMA output element e at row kq + 4e, column r of its
    // 16 x 16 block) is C row 4 (kq + 4e) + t, column 4 r + u of the wave
    // block: for fixed (u, e) the four t are consecutive rows.
    const Opnd &Co = a.C;
    const int gr0 = Co.r0 + m0, gc0 = Co.c0 + n0;
    double *cb;
    if (Co.ts > 0) {
        const int ti = gr0 / Co.ts, tk = gc0 / Co.ts;
        cb = Co.p + tix(ti, tk, Co.ym, Co.tri) * Co.te + (gr0 - ti * Co.ts) +
             (long long)(gc0 - tk * Co.ts) * Co.ld;
    } else {
        cb = Co.p + gr0 + (long long)gc0 * Co.ld;
    }
    const bool diag = a.lower && gr0 < gc0 + 128;
    const int ldc = (int)Co.ld;
    const double alpha = a.alpha, beta = a.beta;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int cl = wn * 64 + 4 * r + u;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int rl = wm * 64 + 4 * (kq + 4 * e);
            double *c = cb + (rl + cl * ldc);
            double v[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] = alpha * acc[t][u][e];
            if (!diag) {
                d2 *c2 = reinterpret_cast<d2 *>(c);
                d2 w0 = {v[0], v[1]}, w1 = {v[2], v[3]};
                if (beta != 0.0) {
                    const d2 o0 = c2[0], o1 = c2[1];
                    w0 += beta * o0;
                    w1 += beta * o1;
                }
                c2[0] = w0;
                c2[1] = w1;
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (gr0 + rl + t < gc0 + cl) continue;
                    c[t] = beta != 0.0 ? v[t] + beta * c[t] : v[t];
                }
            }
        }
    }
}

// ---- few right-hand sides -------------------------------------------------
// C (M x N) = beta*C + alpha * A * B^T for the solves' right-hand-side
// panels: A and C plain, B a packed (or plain) operand streamed once.  One
// wave owns 16 rows m x 16 (or 32) columns n; A and B go from global memory
// straight into MFMA operand registers (A lane (m = l%16, k = l/16), B lane
// (n = l%16, k = l/16)), no LDS and no barriers, and the K loop is unrolled
// so each wave keeps many loads in flight: the kernel streams B (the
// factor's tile column or tile row) at memory speed.
//
// Element type T of A, B and C: double (v_mfma_f64_16x16x4_f64, result
// element r of a lane at row lane/16 + 4r) or float for the correction
// solves of a low-precision factor (v_mfma_f32_16x16x4_f32, same operand
// layout, result rows as probed on the device: rowmap 0: 4*(lane/16) + r,
// 1: lane/16 + 4r).  FP32 halves the bytes of the factor and of the
// right-hand-side operand that every wave re-reads.
struct RArgs {
    Opnd A, B, C;
    int M, N, K;
    double alpha, beta;
    int rowmap;
};

__device__ __forceinline__ long long elem_off(const Opnd &o, int gr, int gc) {
    if (o.ts <= 0) return gr + (long long)gc * o.ld;
    const int ti = gr / o.ts, tk = gc / o.ts;
    return tix(ti, tk, o.ym, o.tri) * o.te + (gr - ti * o.ts) + (long long)(gc - tk * o.ts) * o.ld;
}

typedef float f4r __attribute__((ext_vector_type(4)));
typedef float f2r __attribute__((ext_vector_type(2)));

template <typename T> struct Rk;
template <> struct Rk<double> {
    typedef d4 acc;
    typedef d2 pair;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r, int) { return (lane >> 4) + 4 * r; }
    // four consecutive elements (16-byte aligned) as two pairs
    static __device__ __forceinline__ void ld4(const double *q, pair &lo, pair &hi) {
        lo = *reinterpret_cast<const d2 *>(q);
        hi = *reinterpret_cast<const d2 *>(q + 2);
    }
};
template <> struct Rk<float> {
    typedef f4r acc;
    typedef f2r pair;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r, int rowmap) {
        return rowmap ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
    }
    static __device__ __forceinline__ void ld4(const float *q, pair &lo, pair &hi) {
        const f4r v = *reinterpret_cast<const f4r *>(q);
        lo = f2r{v.x, v.y};
        hi = f2r{v.z, v.w};
    }
};

template <typename T> __device__ __forceinline__ T *optr(const Opnd &o, long long off) {
    return reinterpret_cast<T *>(o.p) + off;
}

// C[rows m0 + row(r), column n] = alpha * acc[r] + beta * C (rows < M)
template <typename T>
__device__ __forceinline__ void rhs_store(const RArgs &a, int m0, int n, int lane,
                                          const typename Rk<T>::acc &acc) {
    const T alpha = (T)a.alpha, beta = (T)a.beta;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = m0 + Rk<T>::row(lane, r, a.rowmap);
        if (m >= a.M) continue;
        T *c = optr<T>(a.C, elem_off(a.C, a.C.r0 + m, a.C.c0 + n));
        T v = alpha * acc[r];
        if (a.beta != 0.0) v += beta * *c;
        *c = v;
    }
}

// K may span several tiles of a packed B (the solves' bulk updates take
// 4096/NB tile columns or rows at once): the kernels walk K in chunks that
// end at tile edges, with the operand pointers recomputed per chunk.
__device__ __forceinline__ int rhs_chunk(const Opnd &B, int gk, int left) {
    return B.ts > 0 ? min(left, B.ts - gk % B.ts) : left;
}

// acc += A(m, 4t + kl) * B(n, 4t + kl) over t < nk: lane (il, kl) of A at
// pa + t*astep and of B at pb + t*bstep; two batches of U loads in flight.
template <int U, typename T>
__device__ __forceinline__ void rhs_run(const T *pa, long long astep, const T *pb,
                                        long long bstep, int nk, typename Rk<T>::acc &acc) {
    int t = 0;
    if (nk >= U) {
        T av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            av[u] = pa[u * astep];
            bv[u] = pb[u * bstep];
        }
        for (; t + U <= nk; t += U) {
            T an[U], bq[U];
            const bool more = t + 2 * U <= nk;
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    an[u] = pa[(t + U + u) * astep];
                    bq[u] = pb[(t + U + u) * bstep];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) acc = Rk<T>::mma(av[u], bv[u], acc);
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    av[u] = an[u];
                    bv[u] = bq[u];
                }
            }
        }
    }
    for (; t < nk; ++t) acc = Rk<T>::mma(pa[t * astep], pb[t * bstep], acc);
}

// Transposed B with K a multiple of 16: K runs down a stored column, so a
// lane loads 4 consecutive K values of its column (rows 4kq..4kq+3 of each
// 16-wide K group, 16-byte loads) and MFMA j of the group takes element j:
// MFMA j covers K = 16g + 4kq + j over the 4 lane groups kq.  A is loaded to
// match (qa = A[m][4kq], element j of group g at qa + (16g + j)*lda).  Two
// batches of 4 groups in flight.
template <typename T>
__device__ __forceinline__ void rhs_run_tb16(const T *qa, long long lda, const T *qb, int ng,
                                             typename Rk<T>::acc &acc) {
    typedef typename Rk<T>::pair P2;
    constexpr int G = 4;
    T av[G][4];
    P2 bv[G][2];
    auto load = [&](int g0, T (&x)[G][4], P2 (&y)[G][2]) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const long long k = 16LL * (g0 + g);
#pragma unroll
            for (int j = 0; j < 4; ++j) x[g][j] = qa[(k + j) * lda];
            Rk<T>::ld4(qb + k, y[g][0], y[g][1]);
        }
    };
    int g0 = 0;
    if (ng >= G) {
        load(0, av, bv);
        for (; g0 + G <= ng; g0 += G) {
            T an[G][4];
            P2 bn[G][2];
            const bool more = g0 + 2 * G <= ng;
            if (more) load(g0 + G, an, bn);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                acc = Rk<T>::mma(av[g][0], bv[g][0].x, acc);
                acc = Rk<T>::mma(av[g][1], bv[g][0].y, acc);
                acc = Rk<T>::mma(av[g][2], bv[g][1].x, acc);
                acc = Rk<T>::mma(av[g][3], bv[g][1].y, acc);
            }
            if (more) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) av[g][j] = an[g][j];
                    bv[g][0] = bn[g][0];
                    bv[g][1] = bn[g][1];
                }
            }
        }
    }
    // remaining groups (fewer than G), one at a time
    for (; g0 < ng; ++g0) {
        const long long k = 16LL * g0;
        P2 b0, b1;
        Rk<T>::ld4(qb + k, b0, b1);
        acc = Rk<T>::mma(qa[k * lda], b0.x, acc);
        acc = Rk<T>::mma(qa[(k + 1) * lda], b0.y, acc);
        acc = Rk<T>::mma(qa[(k + 2) * lda], b1.x, acc);
        acc = Rk<T>::mma(qa[(k + 3) * lda], b1.y, acc);
    }
}

// B not transposed, N a multiple of 32: one wave owns 16 rows x 32 columns.
// Lane (il, kl) loads B rows n0 + 2il and n0 + 2il + 1 of K column kl (one
// load of two contiguous elements) and feeds them to two MFMAs, one over the
// wave's even output columns and one over its odd ones, so one A load serves
// two MFMAs.  One-wave workgroups: the dispatcher hands out work in units of
// one wave, so a grid a little over a multiple of the CU count does not
// leave a few CUs a whole extra round.
template <typename T>
__global__ __launch_bounds__(64) void k_rhs_w2(RArgs a) {
    typedef typename Rk<T>::pair P2;
    typedef typename Rk<T>::acc Acc;
    constexpr int U = 8;   // K steps per batch; two batches in flight
    const int lane = threadIdx.x;
    const int n0 = blockIdx.x * 32;
    const int m0 = blockIdx.y * 16;
    if (n0 >= a.N) return;   // whole wave
    const int il = lane & 15, kl = lane >> 4;
    const int mi = min(m0 + il, a.M - 1);
    const long long astep = 4 * a.A.ld, bstep = 4 * a.B.ld;
    Acc ev = {0, 0, 0, 0}, od = {0, 0, 0, 0};
    for (int kd = 0; kd < a.K;) {
        const int len = rhs_chunk(a.B, a.B.c0 + kd, a.K - kd);
        const T *pa = optr<T>(a.A, elem_off(a.A, a.A.r0 + mi, a.A.c0 + kd + kl));
        const T *pb = optr<T>(a.B, elem_off(a.B, a.B.r0 + n0 + 2 * il, a.B.c0 + kd + kl));
        const int nk = len / 4;
        int t = 0;
        if (nk >= U) {
            T av[U];
            P2 bv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                av[u] = pa[u * astep];
                bv[u] = *reinterpret_cast<const P2 *>(pb + u * bstep);
            }
            for (; t + U <= nk; t += U) {
                T an[U];
                P2 bn[U];
                const bool more = t + 2 * U <= nk;
                if (more) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        an[u] = pa[(t + U + u) * astep];
                        bn[u] = *reinterpret_cast<const P2 *>(pb + (t + U + u) * bstep);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ev = Rk<T>::mma(av[u], bv[u].x, ev);
                    od = Rk<T>::mma(av[u], bv[u].y, od);
                }
                if (more) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        av[u] = an[u];
                        bv[u] = bn[u];
                    }
                }
            }
        }
        for (; t < nk; ++t) {
            const T x = pa[t * astep];
            const P2 y = *reinterpret_cast<const P2 *>(pb + t * bstep);
            ev = Rk<T>::mma(x, y.x, ev);
            od = Rk<T>::mma(x, y.y, od);
        }
        kd += len;
    }
    rhs_store<T>(a, m0, n0 + 2 * il, lane, ev);
    rhs_store<T>(a, m0, n0 + 2 * il + 1, lane, od);
}

// B transposed (K down a stored column), N a multiple of 32, K chunks of
// whole 16-wide groups: one wave owns 16 rows x 32 columns, the two 16-column
// halves taking the same A registers (rhs_run_tb16's K order, 4 consecutive
// K per lane), so A costs one load per two MFMAs.  One-wave workgroups.
template <typename T>
__global__ __launch_bounds__(64) void k_rhs_tw2(RArgs a) {
    typedef typename Rk<T>::pair P2;
    typedef typename Rk<T>::acc Acc;
    constexpr int G = 2;   // 16-wide K groups per batch; two batches in flight
    const int lane = threadIdx.x;
    const int n0 = blockIdx.x * 32;
    const int m0 = blockIdx.y * 16;
    if (n0 >= a.N) return;
    const int il = lane & 15, kl = lane >> 4;
    const int mi = min(m0 + il, a.M - 1);
    const long long lda = a.A.ld;
    Acc c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
    for (int kd = 0; kd < a.K;) {
        const int len = rhs_chunk(a.B, a.B.r0 + kd, a.K - kd);
        const T *qa = optr<T>(a.A, elem_off(a.A, a.A.r0 + mi, a.A.c0 + kd + 4 * kl));
        const T *qb0 = optr<T>(a.B, elem_off(a.B, a.B.r0 + kd + 4 * kl, a.B.c0 + n0 + il));
        const T *qb1 = optr<T>(a.B, elem_off(a.B, a.B.r0 + kd + 4 * kl, a.B.c0 + n0 + 16 + il));
        const int ng = len / 16;
        T av[G][4];
        P2 bv[G][4];
        auto load = [&](int g0, T (&x)[G][4], P2 (&y)[G][4]) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const long long k = 16LL * (g0 + g);
#pragma unroll
                for (int j = 0; j < 4; ++j) x[g][j] = qa[(k + j) * lda];
                Rk<T>::ld4(qb0 + k, y[g][0], y[g][1]);
                Rk<T>::ld4(qb1 + k, y[g][2], y[g][3]);
            }
        };
        auto mma = [&](const T (&x)[4], const P2 (&y)[4]) {
            c0 = Rk<T>::mma(x[0], y[0].x, c0);
            c1 = Rk<T>::mma(x[0], y[2].x, c1);
            c0 = Rk<T>::mma(x[1], y[0].y, c0);
            c1 = Rk<T>::mma(x[1], y[2].y, c1);
            c0 = Rk<T>::mma(x[2], y[1].x, c0);
            c1 = Rk<T>::mma(x[2], y[3].x, c1);
            c0 = Rk<T>::mma(x[3], y[1].y, c0);
            c1 = Rk<T>::mma(x[3], y[3].y, c1);
        };
        int g0 = 0;
        if (ng >= G) {
            load(0, av, bv);
            for (; g0 + G <= ng; g0 += G) {
                T an[G][4];
                P2 bn[G][4];
                const bool more = g0 + 2 * G <= ng;
                if (more) load(g0 + G, an, bn);
#pragma unroll
                for (int g = 0; g < G; ++g) mma(av[g], bv[g]);
                if (more) {
#pragma unroll
                    for (int g = 0; g < G; ++g) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            av[g][j] = an[g][j];
                            bv[g][j] = bn[g][j];
                        }
                    }
                }
            }
        }
        for (; g0 < ng; ++g0) {
            T x[1][4];
            P2 y[1][4];
            const long long k = 16LL * g0;
#pragma unroll
            for (int j = 0; j < 4; ++j) x[0][j] = qa[(k + j) * lda];
            Rk<T>::ld4(qb0 + k, y[0][0], y[0][1]);
            Rk<T>::ld4(qb1 + k, y[0][2], y[0][3]);
            mma(x[0], y[0]);
        }
        kd += len;
    }
    rhs_store<T>(a, m0, n0 + il, lane, c0);
    rhs_store<T>(a, m0, n0 + 16 + il, lane, c1);
}

// The general form: one wave per 16 x 16 output.  KS = 4 (split K): the
// block's four waves share one 16 x 16 output and take a quarter of K each,
// reduced through LDS in a fixed order.  Each wave's serial MFMA and load
// chain is a quarter as long, which is what bounds the small launches (a
// diagonal-tile product: 8 blocks).
template <int TB, int KS, typename T>
__global__ __launch_bounds__(256) void k_rhs(RArgs a) {
    typedef typename Rk<T>::acc Acc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n0 = (KS == 1 ? blockIdx.x * 4 + wave : blockIdx.x) * 16;
    const int m0 = blockIdx.y * 16;
    if (n0 >= a.N) return;   // whole wave (whole block for KS = 4)
    if (KS > 1) {
        a.K /= KS;
        const int kb = wave * a.K;
        a.A.c0 += kb;
        if (TB == 0) a.B.c0 += kb; else a.B.r0 += kb;
    }
    const int il = lane & 15, kl = lane >> 4;
    // lanes past M or N read a valid row/column instead: they only feed
    // output rows/columns that are never stored
    const int mi = min(m0 + il, a.M - 1), ni = min(n0 + il, a.N - 1);
    Acc acc = {0, 0, 0, 0};
    for (int kd = 0; kd < a.K;) {
        const int len = rhs_chunk(a.B, (TB == 0 ? a.B.c0 : a.B.r0) + kd, a.K - kd);
        if (TB == 1 && (len & 15) == 0 && len >= 128) {
            const T *qa = optr<T>(a.A, elem_off(a.A, a.A.r0 + mi, a.A.c0 + kd + 4 * kl));
            const T *qb = optr<T>(a.B, elem_off(a.B, a.B.r0 + kd + 4 * kl, a.B.c0 + ni));
            rhs_run_tb16<T>(qa, a.A.ld, qb, len / 16, acc);
        } else {
            const T *pa = optr<T>(a.A, elem_off(a.A, a.A.r0 + mi, a.A.c0 + kd + kl));
            const T *pb;
            long long bstep;
            if (TB == 0) {
                pb = optr<T>(a.B, elem_off(a.B, a.B.r0 + ni, a.B.c0 + kd + kl));
                bstep = 4 * a.B.ld;
            } else {
                pb = optr<T>(a.B, elem_off(a.B, a.B.r0 + kd + kl, a.B.c0 + ni));
                bstep = 4;
            }
            rhs_run<16, T>(pa, 4 * a.A.ld, pb, bstep, len / 4, acc);
        }
        kd += len;
    }
    if constexpr (KS > 1) {
        __shared__ Acc part[KS - 1][64];
        if (wave > 0) part[wave - 1][lane] = acc;
        __syncthreads();
        if (wave > 0) return;
#pragma unroll
        for (int w = 0; w < KS - 1; ++w) {
            const Acc q = part[w][lane];
            acc[0] += q[0];
            acc[1] += q[1];
            acc[2] += q[2];
            acc[3] += q[3];
        }
    }
    if (n0 + il >= a.N) return;
    rhs_store<T>(a, m0, n0 + il, lane, acc);
}

// The block's rows of o (starting at a multiple of X past o.r0) stay in one
// tile: tiles and the start row are multiples of the block height.
static bool block_aligned(const Opnd &o, int X) {
    return o.ts <= 0 || ((o.ts % X) == 0 && (o.r0 % X) == 0);
}

template <int BN, int MODE, int KIND, int TB = 0, int BM = 128>
hipError_t launch_one(const Gemm &g, hipStream_t s) {
    KArgs a;
    a.A = g.A;
    a.B = g.B;
    a.C = g.C;
    a.M = g.M;
    a.N = g.N;
    a.K = g.K;
    a.alpha = g.alpha;
    a.beta = g.beta;
    a.lower = g.lower;
    a.mb = (g.M + BM - 1) / BM;
    a.nb = (g.N + BN - 1) / BN;
    if (MODE == 1) {
        // lower trapezoid: column n has mb - n blocks
        long long nbk = 0;
        for (int n = 0; n < a.nb; ++n) nbk += (a.mb - n);
        a.nblk = nbk;
    } else {
        a.nblk = (long long)a.mb * a.nb;
    }
    if (a.nblk <= 0) return hipSuccess;
    if (a.nblk > 0x7fffffffLL) return hipErrorInvalidValue;
    const bool full = !TB && (g.M % BM) == 0 && (g.N % BN) == 0 && (g.K % BK) == 0 &&
                      block_aligned(g.A, BM) && block_aligned(g.B, BN) && block_aligned(g.C, BM) &&
                      (g.C.ts <= 0 || ((g.C.ts % BN) == 0 && (g.C.c0 % BN) == 0)) &&
                      g.A.ld * BK < 0x7fffffffLL && g.B.ld * BK < 0x7fffffffLL &&
                      g.C.ld * BN < 0x7fffffffLL;
    if constexpr (BM == 128 && BN == 128 && !TB) {
        if (full) {
            auto flat = [&](const Opnd &o) { return o.ts <= 0 || (o.c0 % o.ts) + g.K <= o.ts; };
            if (flat(g.A) && flat(g.B))
                hipLaunchKernelGGL((k_dgemm_dtv<MODE, KIND, true>), dim3((unsigned)a.nblk), dim3(NT),
                                   0, s, a);
            else
                hipLaunchKernelGGL((k_dgemm_dtv<MODE, KIND, false>), dim3((unsigned)a.nblk), dim3(NT),
                                   0, s, a);
            return hipGetLastError();
        }
    }
    if (full)
        hipLaunchKernelGGL((k_dgemm_nt<BM, BN, MODE, KIND, TB, !TB>), dim3((unsigned)a.nblk),
                           dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL((k_dgemm_nt<BM, BN, MODE, KIND, TB, 0>), dim3((unsigned)a.nblk),
                           dim3(NT), 0, s, a);
    return hipGetLastError();
}

template <typename T>
void launch_rhs_kernels(const Gemm &g, const RArgs &a, int tb, hipStream_t s) {
    const Opnd &B = g.B;
    // split K when the unsplit grid would not give every CU a block (a
    // large grid streams B faster unsplit) and each quarter keeps whole
    // 16-wide K groups, at least 64
    const long long blocks1 = (long long)((g.N + 63) / 64) * ((g.M + 15) / 16);
    if (blocks1 < 256 && g.K % 64 == 0 && g.K >= 256) {
        dim3 grid((unsigned)((g.N + 15) / 16), (unsigned)((g.M + 15) / 16));
        if (tb)
            hipLaunchKernelGGL((k_rhs<1, 4, T>), grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_rhs<0, 4, T>), grid, dim3(256), 0, s, a);
    } else if (!tb && g.N % 32 == 0 && (B.ts > 0 ? B.ts % 32 == 0 && B.r0 % 32 == 0
                                                  : B.ld % 2 == 0 && B.r0 % 2 == 0) &&
               ((uintptr_t)B.p & 15) == 0) {
        // 32-column groups inside one tile, aligned pairs of B rows
        dim3 grid((unsigned)(g.N / 32), (unsigned)((g.M + 15) / 16));
        hipLaunchKernelGGL(k_rhs_w2<T>, grid, dim3(64), 0, s, a);
    } else if (tb && g.N % 32 == 0 && g.K % 16 == 0 &&
               (B.ts > 0 ? B.ts % 32 == 0 && B.c0 % 32 == 0 && B.r0 % 16 == 0
                         : B.ld % 4 == 0 && B.r0 % 4 == 0) &&
               ((uintptr_t)B.p & 15) == 0) {
        // 32-column groups inside one tile, K chunks of whole 16-wide groups
        // starting on 16-byte boundaries
        dim3 grid((unsigned)(g.N / 32), (unsigned)((g.M + 15) / 16));
        hipLaunchKernelGGL(k_rhs_tw2<T>, grid, dim3(64), 0, s, a);
    } else {
        dim3 grid((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 15) / 16));
        if (tb)
            hipLaunchKernelGGL((k_rhs<1, 1, T>), grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_rhs<0, 1, T>), grid, dim3(256), 0, s, a);
    }
}

hipError_t launch_rhs(const Gemm &g, int tb, hipStream_t s) {
    if (g.A.ts > 0 || g.C.ts > 0 || (g.K & 3)) return hipErrorInvalidValue;
    const Opnd &B = g.B;
    if (B.ts > 0) {
        // each 16-column group inside one tile; K chunks (split at tile
        // edges) of whole 4-wide steps
        if (B.ts % 16) return hipErrorInvalidValue;
        if (!tb && ((B.r0 % 16) || (B.c0 % 4))) return hipErrorInvalidValue;
        if (tb && ((B.c0 % 16) || (B.r0 % 4))) return hipErrorInvalidValue;
    }
    // the transposed loads take 4 consecutive elements of a stored column
    // on 16-byte boundaries (8 bytes for the pairs of the untransposed
    // form), which for float needs a plain B's ld and r0 multiples of 4
    if (g.f32 && ((B.ts <= 0 && (B.ld % 4 || B.r0 % 4)) || (g.rowmap & ~1)))
        return hipErrorInvalidValue;
    RArgs a;
    a.A = g.A;
    a.B = g.B;
    a.C = g.C;
    a.M = g.M;
    a.N = g.N;
    a.K = g.K;
    a.alpha = g.alpha;
    a.beta = g.beta;
    a.rowmap = g.rowmap;
    if (g.f32)
        launch_rhs_kernels<float>(g, a, tb, s);
    else
        launch_rhs_kernels<double>(g, a, tb, s);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm(const Gemm &g, GemmKind kind, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    if (g.f32 && kind != GEMM_RHS && kind != GEMM_RHS_TB) return hipErrorInvalidValue;
    // Operand constraints of the streaming loads (16-byte row-pair loads,
    // 16-wide K steps that never straddle a tile).
    const Opnd *ops[3] = {&g.A, &g.B, &g.C};
    for (const Opnd *o : ops) {
        if ((o->ld & 1) || (o->r0 & 1)) return hipErrorInvalidValue;
        if (o->ts > 0 && ((o->ts % BK) != 0)) return hipErrorInvalidValue;
        if (((uintptr_t)o->p) & 15) return hipErrorInvalidValue;
    }
    const bool tb = kind == GEMM_GENERAL_TB || kind == GEMM_NARROW_TB || kind == GEMM_APPLY_TB ||
                    kind == GEMM_SKINNY_TB || kind == GEMM_RHS_TB;
    if ((g.A.ts > 0 && (g.A.c0 % BK))) return hipErrorInvalidValue;
    if (!tb && g.B.ts > 0 && (g.B.c0 % BK)) return hipErrorInvalidValue;
    if (tb && g.B.ts > 0 && (g.B.r0 % BK)) return hipErrorInvalidValue;
    if (g.C.ts > 0 && (g.C.r0 % 16)) return hipErrorInvalidValue;
    switch (kind) {
    case GEMM_UPDATE:
        // the triangular block enumeration needs the diagonal on block
        // boundaries; otherwise fall back to the rectangular grid + skipping
        if (g.lower && g.C.r0 == g.C.c0 && g.M >= g.N) return launch_one<128, 1, 0>(g, s);
        return launch_one<128, 0, 0>(g, s);
    case GEMM_NARROW:
        return launch_one<64, 0, 1>(g, s);
    case GEMM_APPLY:
        return launch_one<64, 0, 2>(g, s);
    case GEMM_GENERAL_TB:
        return launch_one<128, 0, 4, 1>(g, s);
    case GEMM_NARROW_TB:
        return launch_one<64, 0, 5, 1>(g, s);
    case GEMM_APPLY_TB:
        return launch_one<64, 0, 6, 1>(g, s);
    case GEMM_SKINNY:
        return launch_one<128, 0, 7, 0, 32>(g, s);
    case GEMM_SKINNY_TB:
        return launch_one<128, 0, 8, 1, 32>(g, s);
    case GEMM_RHS:
        return launch_rhs(g, 0, s);
    case GEMM_RHS_TB:
        return launch_rhs(g, 1, s);
    case GEMM_GENERAL:
    default:
        return launch_one<128, 0, 3>(g, s);
    }
}

}  // namespace ooc
