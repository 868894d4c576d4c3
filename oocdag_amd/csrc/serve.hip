// Continuous-batching SPD solve server (no reference counterpart).
//
// Requests (a Gram matrix description, a right-hand-side block, a target
// backward error) are submitted from any thread and served by `slots`
// worker threads, each with its own HIP streams, so the work of different
// requests overlaps on the device.  Admission is by an HBM byte budget: a
// request starts when its working set fits next to what running requests and
// the factor cache hold (least recently used factors are evicted to make
// room).  Routes:
//   small  (n <= 1024): every pending small request that fits is taken
//          together and solved by one launch of the persistent batched kernel
//          (continuous batching: whatever arrived while others ran joins the
//          next launch);
//   HBM    : mixed-precision factor + FP64 refinement (mp_factor / mp_solve);
//          the factor stays resident in the cache keyed by matrix id, so a
//          later request with the same matrix only runs the refinement solve;
//   out-of-core: a request whose working set exceeds the whole budget is
//          factored in FP64 from pinned host memory under that budget
//          (oocdag_cholesky) and solved by streaming the factor's tile
//          columns through the device.
// Every result carries its normwise backward error in FP64.
#include "ooc_internal.h"
#include "../../include/oocdag.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace ooc {
hipError_t gen_small_gram(double *A, int n, int ld, double mu, unsigned long long seed,
                          unsigned long long stream_id, double *Gtmp, hipStream_t s);
int mp_factor_solve(const double *dA, int N, int NB, const double *dB, int nrhs, int ldb,
                    double *dX, int ldx, int precision, double tol, int max_iter, hipStream_t s,
                    hipStream_t s1, oocdag_mpir_stats *st, MpFactor *keep);
}  // namespace ooc

using namespace ooc;

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr int SMALL_N = 1024;
constexpr int MAX_BATCH = 512;

int tile_nb(int n) { return n < 4096 ? 256 : 512; }

// the batched kernel holds systems up to 1024 with up to 64 right-hand sides
bool is_small(const oocdag_request &r) { return r.n <= SMALL_N && r.nrhs <= 64; }

// Below this order a low-precision factor does not pay for its refinement
// steps on this GPU (measured): such requests are factored in FP64.
constexpr int MP_MIN_N = 8192;

// row sums of |A| for a column block (rows < n, w columns, ld)
__global__ void k_rowabs_block(const double *A, long long ld, int n, int w, double *rowsum) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < w; ++c) s += fabs(A[r + (long long)c * ld]);
        rowsum[r] += s;
    }
}

size_t tiles_bytes(int n, int nb) {
    size_t bb = (size_t)((n + nb - 1) / nb);
    return bb * (bb + 1) / 2 * nb * nb * 8;
}

struct Job {
    long long ticket;
    oocdag_request rq;
    double t_submit;
};

struct Done {
    oocdag_result r;
    std::vector<double> x;
};

struct CacheEntry {
    long long id;
    double *A = nullptr;  // packed tiles of A
    MpFactor f;
    size_t bytes = 0;
    int users = 0;
    bool ready = false;
};

__global__ void k_sum(const double *x, long long n, double *out) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long i = blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

}  // namespace

struct oocdag_server {
    size_t budget = 0, used = 0;
    int nslots = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;
    std::unordered_map<long long, Done> done;
    std::list<CacheEntry> cache;  // front = most recently used
    std::vector<std::thread> workers;
    std::atomic<long long> next_ticket{1};
    bool stop = false;
    int running = 0;

    // ---- memory estimates (bytes) ------------------------------------------
    static size_t rhs_bytes(int n, int nrhs) { return (size_t)n * (nrhs + 1) * 8 * 2; }
    static size_t small_bytes(const oocdag_request &r) {
        size_t ld = r.n + (r.n & 1);
        // A, G temp, B, X, W, inverses
        return (ld * r.n * 4 + ld * r.nrhs * 2) * 8 + 65536;
    }
    static size_t large_bytes(const oocdag_request &r, bool hit) {
        int nb = tile_nb(r.n);
        size_t bb = (size_t)((r.n + nb - 1) / nb), np = bb * nb;
        size_t solve = 3 * (size_t)(r.nrhs + 1) * np * 8 + 2 * rhs_bytes(r.n, r.nrhs);
        if (hit) return solve;
        // A tiles + factor + low-precision columns + Gram temporary (n^2)
        return tiles_bytes(r.n, nb) + mp_factor_bytes(r.n, nb) + np * np * 8 + solve;
    }
    CacheEntry *find(long long id) {
        for (auto it = cache.begin(); it != cache.end(); ++it)
            if (it->id == id) {
                cache.splice(cache.begin(), cache, it);
                return &cache.front();
            }
        return nullptr;
    }
    // evict unused cache entries (LRU first) until `need` more bytes fit
    bool make_room(size_t need, hipStream_t s) {
        while (used + need > budget) {
            auto victim = cache.end();
            for (auto it = cache.end(); it != cache.begin();) {
                --it;
                if (it->users == 0 && it->ready) {
                    victim = it;
                    break;
                }
            }
            if (victim == cache.end()) return false;
            mp_release(victim->f, s);
            (void)hipFreeAsync(victim->A, s);
            used -= victim->bytes;
            cache.erase(victim);
        }
        return true;
    }

    void finish(long long ticket, Done &&d) {
        std::lock_guard<std::mutex> lk(mu);
        done[ticket] = std::move(d);
        cv_done.notify_all();
    }

    void worker(int slot);
    void run_small(std::vector<Job> &jobs, hipStream_t s, double t_start);
    void run_large(Job &j, CacheEntry *ce, bool hit, hipStream_t s, hipStream_t s1, double t_start);
    void run_ooc(Job &j, hipStream_t s, double t_start);
};

namespace {
void fill_common(oocdag_result &r, const Job &j, double t_start, double t_end) {
    r.queue_seconds = t_start - j.t_submit;
    r.service_seconds = t_end - t_start;
    r.latency_seconds = t_end - j.t_submit;
}
}  // namespace

void oocdag_server::worker(int slot) {
    hipStream_t s = nullptr, s1 = nullptr;
    (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    (void)hipStreamCreateWithFlags(&s1, hipStreamNonBlocking);
    (void)slot;
    for (;;) {
        std::vector<Job> small;
        Job big;
        bool have_big = false, hit = false, ooc_route = false;
        CacheEntry *ce = nullptr;
        size_t reserved = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stop && pending.empty()) break;
                // first admissible request in arrival order
                for (auto it = pending.begin(); it != pending.end(); ++it) {
                    const oocdag_request &r = it->rq;
                    if (is_small(r)) {
                        // continuous batching: take every pending small request that fits
                        size_t need = 0;
                        for (auto jt = pending.begin(); jt != pending.end() && (int)small.size() < MAX_BATCH;) {
                            if (is_small(jt->rq) && make_room(need + small_bytes(jt->rq), s)) {
                                need += small_bytes(jt->rq);
                                small.push_back(*jt);
                                jt = pending.erase(jt);
                            } else {
                                ++jt;
                            }
                        }
                        if (!small.empty()) {
                            used += need;
                            reserved = need;
                        }
                        break;
                    }
                    CacheEntry *e = find(r.matrix_id);
                    if (e && !e->ready) continue;  // being factored by another slot
                    bool is_hit = e != nullptr;
                    size_t need = large_bytes(r, is_hit);
                    if (!is_hit && need > budget) {
                        big = *it;
                        pending.erase(it);
                        have_big = true;
                        ooc_route = true;
                        break;
                    }
                    if (!make_room(need, s)) continue;
                    used += need;
                    reserved = need;
                    big = *it;
                    pending.erase(it);
                    have_big = true;
                    hit = is_hit;
                    if (is_hit) {
                        ce = e;
                        ce->users++;
                    } else {
                        cache.emplace_front();
                        ce = &cache.front();
                        ce->id = r.matrix_id;
                        ce->users = 1;
                        ce->ready = false;
                    }
                    break;
                }
                if (!small.empty() || have_big) break;
                cv_work.wait(lk);
            }
            if (small.empty() && !have_big) break;  // stopping
            ++running;
        }
        const double t_start = now();
        if (!small.empty())
            run_small(small, s, t_start);
        else if (ooc_route)
            run_ooc(big, s, t_start);
        else
            run_large(big, ce, hit, s, s1, t_start);
        {
            std::lock_guard<std::mutex> lk(mu);
            used -= reserved;
            if (ce) {
                ce->users--;
                if (!hit) {
                    if (ce->A && ce->f.W) {
                        ce->ready = true;
                        ce->bytes = tiles_bytes(ce->f.N, ce->f.NB) + mp_factor_bytes(ce->f.N, ce->f.NB);
                        used += ce->bytes;  // the resident factor now counts against the budget
                    } else {
                        for (auto it = cache.begin(); it != cache.end(); ++it)
                            if (&*it == ce) {
                                cache.erase(it);
                                break;
                            }
                    }
                }
            }
            --running;
            cv_work.notify_all();
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s1);
    (void)hipStreamDestroy(s);
}

void oocdag_server::run_small(std::vector<Job> &jobs, hipStream_t s, double t_start) {
    const int k = (int)jobs.size();
    std::vector<int> n(k), r(k);
    std::vector<long long> a_off(k), b_off(k + 1);
    for (int i = 0; i < k; ++i) {
        n[i] = jobs[i].rq.n;
        r[i] = jobs[i].rq.nrhs;
    }
    long long ta = oocdag_batch_layout(k, n.data(), r.data(), a_off.data(), b_off.data());
    double *A = nullptr, *B = nullptr, *X = nullptr, *G = nullptr;
    std::vector<Done> out(k);
    int rc = OOCDAG_OK;
    if (hipMallocAsync((void **)&A, ta * 8, s) != hipSuccess ||
        hipMallocAsync((void **)&B, b_off[k] * 8, s) != hipSuccess ||
        hipMallocAsync((void **)&X, b_off[k] * 8, s) != hipSuccess ||
        hipMallocAsync((void **)&G, (size_t)(SMALL_N + 1) * SMALL_N * 8, s) != hipSuccess)
        rc = OOCDAG_ENOMEM;
    for (int i = 0; i < k && rc == OOCDAG_OK; ++i) {
        const oocdag_request &q = jobs[i].rq;
        const int ld = q.n + (q.n & 1);
        if (gen_small_gram(A + a_off[i], q.n, ld, q.mu, q.seed, 1, G, s) != hipSuccess ||
            normal_block(B + b_off[i], q.n, q.nrhs, ld, q.rhs_seed, 2, s) != hipSuccess)
            rc = OOCDAG_EHIP;
    }
    if (rc == OOCDAG_OK && hipStreamSynchronize(s) != hipSuccess) rc = OOCDAG_EHIP;
    const double t_input = now() - t_start;
    // small systems: FP64 factor (FP32 + refinement does not pay at these
    // orders); the batched kernel still refines when the error is above tol
    int prec = 0;
    double tol = 1.0;
    for (auto &j : jobs) {
        if (j.rq.precision == 0) prec = 0;
        tol = std::min(tol, j.rq.tol);
    }
    std::vector<double> berr(k);
    std::vector<int> iters(k), status(k);
    oocdag_batch_stats st;
    if (rc == OOCDAG_OK)
        rc = oocdag_batched_solve(k, n.data(), r.data(), a_off.data(), b_off.data(), A, B, X, prec,
                                  tol, 10, 0, berr.data(), iters.data(), status.data(), &st);
    // systems whose FP32 refinement did not reach tol: redo in FP64
    if (rc == OOCDAG_OK && prec == 1) {
        std::vector<int> redo;
        for (int i = 0; i < k; ++i)
            if (status[i] != 0) redo.push_back(i);
        if (!redo.empty()) {
            rc = oocdag_batched_solve(k, n.data(), r.data(), a_off.data(), b_off.data(), A, B, X, 0, tol,
                                      10, 0, berr.data(), iters.data(), status.data(), &st);
            for (int i : redo) out[i].r.fell_back = 1;
            if (!redo.empty()) prec = 0;
        }
    }
    std::vector<double> hx(b_off[k]);
    if (rc == OOCDAG_OK &&
        (hipMemcpyAsync(hx.data(), X, b_off[k] * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = OOCDAG_EHIP;
    for (double *p : {A, B, X, G})
        if (p) (void)hipFreeAsync(p, s);
    (void)hipStreamSynchronize(s);
    const double t_end = now();
    for (int i = 0; i < k; ++i) {
        Done &d = out[i];
        const oocdag_request &q = jobs[i].rq;
        const int ld = q.n + (q.n & 1);
        d.r.status = rc != OOCDAG_OK ? rc : status[i];
        d.r.route = 0;
        d.r.batch_size = k;
        d.r.precision_used = prec;
        d.r.iterations = iters[i];
        d.r.backward_error = berr[i];
        d.x.resize((size_t)q.n * q.nrhs);
        double cs = 0;
        for (int c = 0; c < q.nrhs; ++c)
            for (int row = 0; row < q.n; ++row) {
                double v = rc == OOCDAG_OK ? hx[b_off[i] + row + (size_t)c * ld] : 0.0;
                d.x[row + (size_t)c * q.n] = v;
                cs += v;
            }
        d.r.x_checksum = cs;
        fill_common(d.r, jobs[i], t_start, t_end);
        d.r.input_seconds = t_input;
        finish(jobs[i].ticket, std::move(d));
    }
}

void oocdag_server::run_large(Job &j, CacheEntry *ce, bool hit, hipStream_t s, hipStream_t s1,
                              double t_start) {
    const oocdag_request &q = j.rq;
    const int n = q.n, nb = tile_nb(n), ldb = n + (n & 1);
    Done d;
    d.r.route = 1;
    d.r.cache_hit = hit;
    d.r.batch_size = 1;
    double *B = nullptr, *X = nullptr;
    int rc = OOCDAG_OK;
    if (hipMallocAsync((void **)&B, (size_t)ldb * q.nrhs * 8, s) != hipSuccess ||
        hipMallocAsync((void **)&X, (size_t)ldb * q.nrhs * 8, s) != hipSuccess)
        rc = OOCDAG_ENOMEM;
    if (rc == OOCDAG_OK && normal_block(B, n, q.nrhs, ldb, q.rhs_seed, 2, s) != hipSuccess) rc = OOCDAG_EHIP;
    oocdag_mpir_stats st{};
    double t_input = 0;
    if (rc == OOCDAG_OK && !hit) {
        // materialize A, factor (mixed precision, FP64 fallback), refine, keep the factor
        if (hipMallocAsync((void **)&ce->A, tiles_bytes(n, nb), s) != hipSuccess)
            rc = OOCDAG_ENOMEM;
        else if (gram_tiles(ce->A, n, nb, q.mu, q.seed, s) != hipSuccess ||
                 hipStreamSynchronize(s) != hipSuccess)
            rc = OOCDAG_EHIP;
        t_input = now() - t_start;
        if (rc == OOCDAG_OK) {
            const int prec = n >= MP_MIN_N ? q.precision : 0;
            rc = mp_factor_solve(ce->A, n, nb, B, q.nrhs, ldb, X, ldb, prec, q.tol, 30, s, s1, &st,
                                 &ce->f);
            if (rc == OOCDAG_OK && st.info) rc = st.info;
        }
        if (rc != OOCDAG_OK && ce->A) {
            (void)hipFreeAsync(ce->A, s);
            ce->A = nullptr;
        }
    } else if (rc == OOCDAG_OK) {
        // cached factor: refinement solve only; a stall refactors in FP64
        const double ts = now();
        int it = 0;
        double eta = 0;
        int r = mp_solve(ce->f, B, q.nrhs, ldb, X, ldb, q.tol, 30, s, &it, &eta);
        if (r == 1 && ce->f.precision != 0) {
            MpFactor f64 = ce->f;
            f64.W = f64.Z = f64.dinv = nullptr;
            f64.W32 = f64.Z32 = nullptr;
            rc = mp_factor(f64, ce->A, n, nb, 0, s, s1);
            if (rc == OOCDAG_OK) {
                mp_release(ce->f, s);
                ce->f = f64;
                r = mp_solve(ce->f, B, q.nrhs, ldb, X, ldb, q.tol, 30, s, &it, &eta);
                st.fell_back = 1;
            }
        }
        if (r < 0) rc = r;
        st.iterations = it;
        st.backward_error = eta;
        st.precision_used = ce->f.precision;
        st.solve_seconds = now() - ts;
    }
    if (rc == OOCDAG_OK) {
        d.x.resize((size_t)n * q.nrhs);
        if (hipMemcpy2DAsync(d.x.data(), (size_t)n * 8, X, (size_t)ldb * 8, (size_t)n * 8, q.nrhs,
                             hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = OOCDAG_EHIP;
    }
    if (B) (void)hipFreeAsync(B, s);
    if (X) (void)hipFreeAsync(X, s);
    (void)hipStreamSynchronize(s);
    d.r.status = rc;
    d.r.precision_used = st.precision_used;
    d.r.fell_back = st.fell_back;
    d.r.iterations = st.iterations;
    d.r.backward_error = st.backward_error;
    double cs = 0;
    for (double v : d.x) cs += v;
    d.r.x_checksum = cs;
    fill_common(d.r, j, t_start, now());
    d.r.input_seconds = t_input;
    finish(j.ticket, std::move(d));
}

// ---- out-of-core route ----------------------------------------------------

namespace {

__global__ void k_unpack_tiles_full(const double *Y, int n, int nb, int bb, int tj, double *col,
                                    long long ldc) {
    // tile column tj (tiles tj..bb-1) of the packed lower tiles -> rows
    // [tj*nb, n) x cols [tj*nb, tj*nb + nb) of a column block (ld ldc)
    const long long base = ((long long)tj * bb - ((long long)tj * (tj - 1)) / 2) * nb * nb;
    const long long rows = (long long)(bb - tj) * nb;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < rows * nb;
         idx += (long long)gridDim.x * blockDim.x) {
        long long t = idx / ((long long)nb * nb), w = idx - t * nb * nb;
        int r = (int)(w % nb), c = (int)(w / nb);
        long long gr = (tj + t) * nb + r;
        if (gr < n && tj * nb + c < n) col[gr - (long long)tj * nb + (long long)c * ldc] = Y[base + idx];
    }
}

}  // namespace

void oocdag_server::run_ooc(Job &j, hipStream_t s, double t_start) {
    const oocdag_request &q = j.rq;
    const int n = q.n, nb = tile_nb(n), bb = (n + nb - 1) / nb, ldb = n + (n & 1);
    Done d;
    d.r.route = 2;
    d.r.batch_size = 1;
    d.r.precision_used = 0;
    int rc = OOCDAG_OK;
    // 1. materialize A in pinned host memory (column-major, full symmetric)
    double *Ah = nullptr, *Y = nullptr, *col = nullptr;
    std::vector<double> diag(n);
    if (hipHostMalloc((void **)&Ah, (size_t)n * n * 8, hipHostMallocDefault) != hipSuccess) rc = OOCDAG_ENOMEM;
    if (rc == OOCDAG_OK && (hipMallocAsync((void **)&Y, tiles_bytes(n, nb), s) != hipSuccess ||
                            hipMallocAsync((void **)&col, (size_t)n * nb * 8, s) != hipSuccess))
        rc = OOCDAG_ENOMEM;
    if (rc == OOCDAG_OK && gram_tiles(Y, n, nb, q.mu, q.seed, s) != hipSuccess) rc = OOCDAG_EHIP;
    for (int tj = 0; tj < bb && rc == OOCDAG_OK; ++tj) {
        const int c0 = tj * nb, w = std::min(nb, n - c0);
        hipLaunchKernelGGL(k_unpack_tiles_full, dim3(4096), dim3(256), 0, s, Y, n, nb, bb, tj, col, (long long)n);
        if (hipMemcpy2DAsync(Ah + (size_t)c0 * n + c0, (size_t)n * 8, col, (size_t)n * 8,
                             (size_t)(n - c0) * 8, w, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = OOCDAG_EHIP;
    }
    if (Y) (void)hipFreeAsync(Y, s);
    if (col) (void)hipFreeAsync(col, s);
    (void)hipStreamSynchronize(s);
    if (rc == OOCDAG_OK) {
        for (int c = 0; c < n; ++c) {
            diag[c] = Ah[(size_t)c * n + c];
            for (int r = c + 1; r < n; ++r) Ah[(size_t)r * n + c] = Ah[(size_t)c * n + r];
        }
    }
    const double t_input = now() - t_start;
    // 2. FP64 out-of-core factorization under this server's budget
    if (rc == OOCDAG_OK) {
        int info = 0;
        double t = oocdag_cholesky(Ah, n, nb, n, std::max(budget, (size_t)4 * bb * nb * nb * 8), &info);
        if (t < 0 || info) rc = info ? info : OOCDAG_EHIP;
    }
    // 3. solve by streaming the factor (oocdag_potrs stages it tile column
    //    by tile column) and 4. FP64 residual against A (upper triangle +
    //    saved diagonal), both on the device
    std::vector<double> hb((size_t)n * q.nrhs), hx;
    double *Bd = nullptr;
    if (rc == OOCDAG_OK) {
        if (hipMallocAsync((void **)&Bd, (size_t)ldb * q.nrhs * 8, s) != hipSuccess ||
            normal_block(Bd, n, q.nrhs, ldb, q.rhs_seed, 2, s) != hipSuccess ||
            hipMemcpy2DAsync(hb.data(), (size_t)n * 8, Bd, (size_t)ldb * 8, (size_t)n * 8, q.nrhs,
                             hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = OOCDAG_EHIP;
        if (Bd) (void)hipFreeAsync(Bd, s);
    }
    if (rc == OOCDAG_OK) {
        hx = hb;
        rc = oocdag_potrs(Ah, n, n, hx.data(), q.nrhs, n);
    }
    if (rc == OOCDAG_OK) {
        // A x from the symmetric copy (upper triangle + diagonal) with the
        // device GEMM, one column block at a time
        double *Ablk = nullptr, *Xd = nullptr, *R = nullptr, *rsum = nullptr;
        const int ldr = n + (n & 1);
        if (hipMallocAsync((void **)&rsum, (size_t)n * 8, s) != hipSuccess ||
            hipMemsetAsync(rsum, 0, (size_t)n * 8, s) != hipSuccess ||
            hipMallocAsync((void **)&Ablk, (size_t)ldr * nb * 8, s) != hipSuccess ||
            hipMallocAsync((void **)&Xd, (size_t)ldr * q.nrhs * 8, s) != hipSuccess ||
            hipMallocAsync((void **)&R, (size_t)ldr * q.nrhs * 8, s) != hipSuccess)
            rc = OOCDAG_ENOMEM;
        std::vector<double> blk((size_t)ldr * nb);
        if (rc == OOCDAG_OK &&
            (hipMemcpy2DAsync(Xd, (size_t)ldr * 8, hx.data(), (size_t)n * 8, (size_t)n * 8, q.nrhs,
                              hipMemcpyHostToDevice, s) != hipSuccess ||
             hipMemcpy2DAsync(R, (size_t)ldr * 8, hb.data(), (size_t)n * 8, (size_t)n * 8, q.nrhs,
                              hipMemcpyHostToDevice, s) != hipSuccess))
            rc = OOCDAG_EHIP;
        for (int c0 = 0; c0 < n && rc == OOCDAG_OK; c0 += nb) {
            const int w = std::min(nb, n - c0);
            // A[:, c0:c0+w] rebuilt from the strict upper triangle and diagonal
            for (int c = 0; c < w; ++c) {
                const int gc = c0 + c;
                double *dst = blk.data() + (size_t)c * ldr;
                for (int r = 0; r < gc; ++r) dst[r] = Ah[(size_t)gc * n + r];
                dst[gc] = diag[gc];
                for (int r = gc + 1; r < n; ++r) dst[r] = Ah[(size_t)r * n + gc];
                for (int r = n; r < ldr; ++r) dst[r] = 0.0;
            }
            if (hipMemcpyAsync(Ablk, blk.data(), (size_t)ldr * w * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
                rc = OOCDAG_EHIP;
                break;
            }
            // R -= A[:, c0:c0+w] X[c0:c0+w, :]   (C = A B^T with B = X^T rows)
            // R -= A[:, c0:c0+w] X[c0:c0+w, :]: operand B(rhs, k) = X(c0 + k, rhs),
            // i.e. X read transposed
            Gemm g;
            g.A = plain(Ablk, ldr);
            g.B = plain(Xd, ldr, c0, 0);
            g.C = plain(R, ldr);
            g.M = n; g.N = q.nrhs; g.K = w;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if (launch_gemm(g, GEMM_GENERAL_TB, s) != hipSuccess) rc = OOCDAG_EHIP;
            hipLaunchKernelGGL(k_rowabs_block, dim3(64), dim3(256), 0, s, Ablk, (long long)ldr, n, w, rsum);
            if (hipStreamSynchronize(s) != hipSuccess) rc = OOCDAG_EHIP;
        }
        std::vector<double> hr((size_t)n * q.nrhs), hrs(n);
        if (rc == OOCDAG_OK &&
            (hipMemcpy2DAsync(hr.data(), (size_t)n * 8, R, (size_t)ldr * 8, (size_t)n * 8, q.nrhs,
                              hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipMemcpyAsync(hrs.data(), rsum, (size_t)n * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess))
            rc = OOCDAG_EHIP;
        if (rc == OOCDAG_OK) {
            const double an = *std::max_element(hrs.begin(), hrs.end());  // ||A||_inf
            double eta = 0;
            for (int k = 0; k < q.nrhs; ++k) {
                double rn = 0, xn = 0, bn = 0;
                for (int r = 0; r < n; ++r) {
                    rn = std::max(rn, std::fabs(hr[r + (size_t)k * n]));
                    xn = std::max(xn, std::fabs(hx[r + (size_t)k * n]));
                    bn = std::max(bn, std::fabs(hb[r + (size_t)k * n]));
                }
                eta = std::max(eta, rn / (an * xn + bn));
            }
            d.r.backward_error = eta;
        }
        for (double *p : {Ablk, Xd, R, rsum})
            if (p) (void)hipFreeAsync(p, s);
        (void)hipStreamSynchronize(s);
    }
    if (Ah) (void)hipHostFree(Ah);
    d.r.status = rc;
    d.x = std::move(hx);
    double cs = 0;
    for (double v : d.x) cs += v;
    d.r.x_checksum = cs;
    fill_common(d.r, j, t_start, now());
    finish(j.ticket, std::move(d));
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" oocdag_server *oocdag_server_create(size_t hbm_budget_bytes, int slots) {
    if (slots < 1 || hbm_budget_bytes == 0) return nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nullptr;
    oocdag_server *srv = new oocdag_server();
    srv->budget = hbm_budget_bytes;
    srv->nslots = slots;
    for (int i = 0; i < slots; ++i) srv->workers.emplace_back([srv, i] { srv->worker(i); });
    return srv;
}

extern "C" long long oocdag_server_submit(oocdag_server *srv, const oocdag_request *rq) {
    if (!srv || !rq || rq->n < 1 || rq->nrhs < 1 || rq->nrhs > 64 || rq->precision < 0 ||
        rq->precision > 2 || !(rq->tol > 0))
        return OOCDAG_EARG;
    Job j;
    j.ticket = srv->next_ticket++;
    j.rq = *rq;
    j.t_submit = now();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->pending.push_back(j);
    }
    srv->cv_work.notify_one();
    return j.ticket;
}

extern "C" int oocdag_server_wait(oocdag_server *srv, long long ticket, oocdag_result *res,
                                  double *x, size_t x_len) {
    if (!srv || ticket <= 0) return OOCDAG_EARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return srv->done.count(ticket) > 0; });
    Done d = std::move(srv->done[ticket]);
    srv->done.erase(ticket);
    lk.unlock();
    if (res) *res = d.r;
    if (x && x_len) memcpy(x, d.x.data(), std::min(x_len, d.x.size()) * sizeof(double));
    return OOCDAG_OK;
}

extern "C" int oocdag_server_stats(oocdag_server *srv, size_t *used_bytes, int *cached, int *pending,
                                   int *running) {
    if (!srv) return OOCDAG_EARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (used_bytes) *used_bytes = srv->used;
    if (cached) *cached = (int)srv->cache.size();
    if (pending) *pending = (int)srv->pending.size();
    if (running) *running = srv->running;
    return OOCDAG_OK;
}

extern "C" int oocdag_server_destroy(oocdag_server *srv) {
    if (!srv) return OOCDAG_EARG;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers) t.join();
    hipStream_t s = nullptr;
    (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    for (auto &e : srv->cache) {
        mp_release(e.f, s);
        if (e.A) (void)hipFreeAsync(e.A, s);
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    delete srv;
    return OOCDAG_OK;
}
