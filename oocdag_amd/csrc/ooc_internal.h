// Internal interfaces of liboocdag: operand descriptors for the packed tile
// storage and the kernel launchers used by the scheduler.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <functional>

namespace ooc {

// One matrix operand.  Element (r, c) of the operand lives at
//   p + T(ti, tk) * te + (gr - ti*ts) + (gc - tk*ts) * ld
// with gr = r0 + r, gc = c0 + c, ti = gr/ts, tk = gc/ts and
//   T(ti, tk) = ti + tk*ym - tri*tk*(tk+1)/2.
// tri = 1 is the packed lower-trapezoid panel of the reference's device
// buffer Y (oocquark.cpp:256: tile (i,j) at (i + j*Ym - j(j+1)/2)*NB^2);
// tri = 0 is a rectangular grid of tile columns, each ym tiles tall (the X
// staging buffer, oocquark.cpp:304).  ts = 0 means plain column-major with
// leading dimension ld.
struct Opnd {
    double *p;
    long long ld;
    long long te;
    int ts;
    int ym;
    int tri;
    int r0, c0;
};

inline Opnd plain(double *p, long long ld, int r0 = 0, int c0 = 0) {
    Opnd o;
    o.p = p; o.ld = ld; o.te = 0; o.ts = 0; o.ym = 0; o.tri = 0; o.r0 = r0; o.c0 = c0;
    return o;
}

inline Opnd panel(double *p, int nb, int ym, int tri, int r0 = 0, int c0 = 0) {
    Opnd o;
    o.p = p; o.ld = nb; o.te = (long long)nb * nb; o.ts = nb; o.ym = ym; o.tri = tri;
    o.r0 = r0; o.c0 = c0;
    return o;
}

// Storage edge of a tile of logical edge nb: tile (i, j) of the matrix
// (rows i*nb .. i*nb+nb-1) is kept in a ts x ts slot, ts = nb rounded up to a
// multiple of 64 (the diagonal-block edge of the factorization kernels); rows
// and columns nb .. ts-1 of a slot are padding (0, with 1 on the diagonal of
// a diagonal tile), so the device factors a padded SPD matrix whose factor
// holds L at the valid positions.  ts == nb when nb % 64 == 0.
inline int tile_edge(int nb) { return (nb + 63) / 64 * 64; }

inline long long tile_index(int ti, int tk, int ym, int tri) {
    return ti + (long long)tk * ym - (tri ? ((long long)tk * (tk + 1)) / 2 : 0);
}

// C(M x N) = beta*C + alpha * A(M x K) * B(N x K)^T.
// lower != 0: only elements with (C.r0 + m) >= (C.c0 + n) are written, and
// blocks strictly above that diagonal are not computed.
struct Gemm {
    Opnd A, B, C;
    int M, N, K;
    double alpha, beta;
    int lower;
    // GEMM_RHS / GEMM_RHS_TB only: A, B and C point to float elements (same
    // element offsets) and the products run on FP32 MFMAs, whose result
    // rows follow rowmap (mfma_rowmap(0, ...))
    int f32 = 0;
    int rowmap = 0;
};

// Kernel families (distinct kernel symbols, so profiles separate them).
enum GemmKind {
    GEMM_UPDATE = 0,   // 128x128 tiles: trailing / left-looking panel updates
    GEMM_NARROW = 1,   // 128x64 tiles: in-column blocked factorization updates
    GEMM_APPLY = 2,    // 128x64 tiles: B <- B * Dinv^T (triangular solve apply)
    GEMM_GENERAL = 3,  // 128x128 tiles, rectangular: C-ABI dgemm entry point
    // B given transposed: operand element (n, k) is element (B.r0 + k,
    // B.c0 + n) of the stored matrix B (C = A * B_stored, "NN").  Used by
    // the triangular solves with L (not L^T) on the right.
    GEMM_GENERAL_TB = 4,
    GEMM_NARROW_TB = 5,
    GEMM_APPLY_TB = 6,
    // 32 x 128 blocks for M <= 32 (a few right-hand sides in the solves)
    GEMM_SKINNY = 7,
    GEMM_SKINNY_TB = 8,
    // right-hand-side panels: A and C plain, any M (16-row groups), one
    // wave per 16 or 32 columns, operands streamed straight into MFMA
    // registers; a tiled B keeps each 16-column group in one tile, K may
    // span tiles
    GEMM_RHS = 9,
    GEMM_RHS_TB = 10
};

hipError_t launch_gemm(const Gemm &g, GemmKind kind, hipStream_t s);

// Row order of the FP32-output MFMA results (kind 0: f32 16x16x4, 1: bf16
// 16x16x32), read once from a probe: 0 row = 4*(lane/16) + r, 1 row =
// lane/16 + 4*r; -1 if the layouts differ from both (gemm_lp.hip).
int mfma_rowmap(int kind, hipStream_t s);
// Mixed-precision variant bits (oocdag_mp_set_update_variant).
int mp_variant();

// Largest block edge of the GEMM kernels: operands must provide this many
// addressable rows past any block start they cover.
constexpr int GEMM_BM = 128;

// Factor the 64x64 diagonal block whose (0,0) is operand element (0,0) of P
// (lower triangle only; the strict upper triangle is neither read nor
// written) and write its inverse, zero-filled above the diagonal, to
// dinv (64 x 64, ld 64).  A non-positive pivot at block column k stores
// min(*info, gcol0 + k + 1) into *info.  info == nullptr: the block already
// holds a factor; only its inverse is written.
hipError_t launch_potrf64(const Opnd &P, int gcol0, double *dinv, int *info, hipStream_t s);

// Blocked factorization of one tile column of a packed panel:
// rows [row0, row0 + mrows) x columns [col0, col0 + ncols) of P (global
// panel coordinates, row0 == col0 is the diagonal), ncols % 64 == 0.
// Writes L in place (lower triangle of the diagonal part).
hipError_t factor_column(const Opnd &P, int row0, int mrows, int ncols, int gcol_base,
                         double *dinv_ws, int *info, hipStream_t s);

// Trailing-update policy of a panel factorization.
struct PanelUpdater {
    virtual ~PanelUpdater() {}
    // called on the critical-path stream once tile column j is final
    virtual hipError_t prepare(int j, hipStream_t s) { (void)j; (void)s; return hipSuccess; }
    // apply the update from tile column j to panel tile columns [c0, c0 + nc)
    // (rows c0*NB and below)
    virtual hipError_t update(int j, int c0, int nc, hipStream_t s) = 0;
};

// Factors tile columns [0, Yn) of the packed panel P (Ym tile rows) with
// look-ahead: column j's factorization and the update of column j+1 run on
// s0 while the bulk update of columns j+2.. from column j runs on s1.
// gcol0: global column of panel column 0 (for info); dinv + j*dinv_stride
// receives column j's 64-block inverses (stride 0: reuse one buffer).
// on_column(j) runs right after column j's kernels were queued on s0 (the
// column is final once s0 reaches that point); a nonzero return aborts.
// after_updates(j), when given, runs once every kernel that reads tile column
// j has been queued (the update of column j+1 on s0 and the bulk update from
// column j on s1): work queued on s0 and s1 after that point never reads it.
// Returns an OOCDAG-style code (0 ok, -4 HIP error, other: a callback's).
int factor_panel_la(const Opnd &P, int NB, int Ym, int Yn, int gcol0, double *dinv,
                    size_t dinv_stride, int *dinfo, hipStream_t s0, hipStream_t s1,
                    PanelUpdater &up, const std::function<int(int)> &on_column,
                    const std::function<int(int)> *after_updates = nullptr);

// Make a staged tile whose valid part is vr x vc behave like an NB x NB
// tile of the zero-padded problem: entries outside the valid part become 0,
// and 1 on the diagonal when the tile is a diagonal tile (so the padded
// matrix stays SPD and its factor's valid part is unchanged).
hipError_t launch_pad_tile(double *tile, int nb, int vr, int vc, int diag, hipStream_t s);

// PDMATGEN('S'/'H' or 'F', DIAG, ISEED) values written straight into the
// packed lower tiles (tri = 1, ym = bb) of tile columns [jt0, jt1): tiles of
// logical edge nb in slots of edge ts = tile_edge(nb), padded beyond N and
// beyond nb inside each slot (0, 1 on a diagonal tile's diagonal).  Y points
// at the tile with packed index tbase (0: the whole layout; else a window
// holding those columns only).
hipError_t launch_gen_tiles(double *Y, int N, int nb, int bb, int jt0, int jt1, char aform,
                            char diag, int iseed, hipStream_t s,
                            long long tbase = 0);

// ---- solves and residual (solve.hip) ----
// 64x64 inverses of the diagonal blocks of the n x n lower triangle L
// (n % 64 == 0) into dinv (n/64 blocks of 64*64, ld 64).
hipError_t diag_inverses(const Opnd &L, int n, double *dinv, hipStream_t s);
// B (m x n) <- B * L^-T and B <- B * L^-1 with L's block inverses in dinv.
hipError_t trsm_right_lower_t(const Opnd &B, int m, int n, const Opnd &L, const double *dinv,
                              hipStream_t s);
hipError_t trsm_right_lower_n(const Opnd &B, int m, int n, const Opnd &L, const double *dinv,
                              hipStream_t s);
// X^T (nrhs x Npad, ld ldx) <- B^T A^-1 with A = L L^T in packed tiles dY;
// dinv_all: the 64-block inverses of every diagonal tile, tile after tile.
hipError_t potrs_tiles_t(const double *dY, int N, int NB, const double *dinv_all, double *Xt,
                         int nrhs, long long ldx, hipStream_t s);
// out(c, r) = in(r, c).  ts > 0: the tiled dimension holds tiles of nbl
// rows in slots of ts; mode 1 maps in's row index, mode 2 in's column index
// from logical to slot positions (the other side stays logical).
hipError_t launch_transpose(const double *in, long long ld_in, double *out, long long ld_out,
                            int rows, int cols, hipStream_t s, int nbl = 0, int ts = 0,
                            int mode = 0);
// Row sums of |A - L L^T| (rowsum) and of |A| (anorm) for the packed factor
// dY of the PDMATGEN(aform, diag, iseed) matrix.  R and S: scratch of bb
// tiles each.
hipError_t llt_resid_tiles(const double *dY, int N, int NB, char aform, char diag, int iseed,
                           double *R, double *S, double *rowsum, double *anorm, hipStream_t s);
// Adds |A_host - A| row sums of the strict upper tiles above diagonal tile jt
// (U: rows [0, jt*NB) of columns [jt*NB, jt*NB+NB), ld ldu).
hipError_t resid_upper_block(const double *U, long long ldu, int N, int NB, int jt, char aform,
                             char diag, int iseed, double *rowsum, hipStream_t s);

// ---- mixed precision (mixed.hip) ----
// A resident factorization of one matrix: W = factor (packed FP64 tiles),
// Z = inverses of its diagonal tiles, dinv = its 64-block inverses; A is the
// matrix itself (not owned) for the residuals.
struct MpFactor {
    const double *A = nullptr;
    double *W = nullptr, *Z = nullptr, *dinv = nullptr;
    // FP32 copies of W and Z for the correction solves of a low-precision
    // factor (precision > 0)
    float *W32 = nullptr, *Z32 = nullptr;
    int N = 0, NB = 0, precision = 0, info = 0;
    double Anorm = 0, factor_seconds = 0;
    size_t bytes = 0;
};
size_t mp_factor_bytes(int N, int NB);
int mp_factor(MpFactor &f, const double *dA, int N, int NB, int precision, hipStream_t s0,
              hipStream_t s1);
void mp_release(MpFactor &f, hipStream_t s);
// refinement solve: 0 converged, 1 stalled, < 0 error
int mp_solve(const MpFactor &f, const double *dB, int nrhs, int ldb, double *dX, int ldx,
             double tol, int max_iter, hipStream_t s, int *iters, double *eta_out,
             double *lres_out = nullptr);
// PDLASCHK solve residual max_k ||r_k|| / (||x_k|| ||A|| eps N) from norms.
double pdlaschk_resid(const double *rnorm, const double *xnorm, int nrhs, double anorm, int N);
hipError_t gram_tiles(double *dY, int N, int NB, double mu, unsigned long long seed,
                      hipStream_t s);
// Tile column j (tiles j..bb-1, ld NB, diagonal tile full) of the Gram
// matrix gram_tiles writes, with G streamed through Gp ((bb-j)*NB x kw).
hipError_t gram_tile_column(double *col, int N, int NB, int j, double mu, unsigned long long seed,
                            double *Gp, int kw, hipStream_t s);
hipError_t normal_block(double *dB, int rows, int cols, long long ld, unsigned long long seed,
                        unsigned long long stream, hipStream_t s);
hipError_t symm_t(const double *dA, int N, int NB, const double *Xt, double *Yt, long long ld,
                  int nrhs, double alpha, double beta, hipStream_t s);

// Dispatch order -> logical order for a grid of G workgroups: XCD x
// (= blockIdx % 8, round-robin dispatch) gets a contiguous slice of logical
// ids, so neighbouring work shares that XCD's L2.  A bijection on [0, G).
__device__ __forceinline__ long long xcd_remap(long long bid, long long G) {
    const int NX = 8;
    long long per = G / NX, rem = G % NX;
    long long x = bid % NX, i = bid / NX;
    return (x < rem) ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

}  // namespace ooc
