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
//          generated tile column by tile column into host memory, factored
//          in FP64 out of core (oocdag_cholesky_host_tiles) under the part
//          of the budget it reserved, solved and checked by streaming the
//          factor's and the matrix's tile columns through the device; at no
//          point does it hold more HBM than its reservation.
// Every result carries its normwise backward error and the PDLASCHK solve
// residual (LIN/pdlaschk.f:186,295) in FP64.
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
    oocdag_result r{};
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
    // The out-of-core route's reservation: at least what it cannot run
    // without (two tile columns for the factorization, as the reference's
    // budget check asks, oocdag.cpp:248-250; one tile column, a G panel and
    // the right-hand sides for generation, solve and residual), otherwise
    // half the budget.
    static size_t ooc_min_bytes(const oocdag_request &r) {
        const int nb = tile_nb(r.n);
        const size_t bb = (size_t)((r.n + nb - 1) / nb), np = bb * nb;
        const size_t col = bb * nb * nb * 8;
        const size_t factor = 2 * col + (size_t)nb * 64 * 8 + 4096;
        const size_t route = 2 * col + 4 * (size_t)(r.nrhs + 2) * np * 8 + np * 64 * 8 + 65536;
        return std::max(factor, route);
    }
    size_t ooc_bytes(const oocdag_request &r) const {
        return std::min(budget, std::max(ooc_min_bytes(r), budget / 2));
    }
    // a request no admission can ever satisfy
    bool never_fits(const oocdag_request &r) const {
        if (is_small(r)) return small_bytes(r) > budget;
        return large_bytes(r, false) > budget && ooc_min_bytes(r) > budget;
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

    // Results of the request(s) a worker is serving, kept in the worker
    // thread's outbox and published after the worker has returned the
    // request's reservation to the budget, so a client that sees its result
    // also sees the memory released.
    void finish(long long ticket, Done &&d) { outbox().emplace_back(ticket, std::move(d)); }
    static std::vector<std::pair<long long, Done>> &outbox() {
        thread_local std::vector<std::pair<long long, Done>> box;
        return box;
    }

    void worker(int slot);
    void run_small(std::vector<Job> &jobs, hipStream_t s, double t_start);
    void run_large(Job &j, CacheEntry *ce, bool hit, hipStream_t s, hipStream_t s1, double t_start);
    void run_ooc(Job &j, size_t reserved, hipStream_t s, double t_start);
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
                // requests that cannot fit even in an empty budget fail at once
                for (auto it = pending.begin(); it != pending.end();) {
                    if (never_fits(it->rq)) {
                        Done d;
                        d.r.status = OOCDAG_EBUDGET;
                        d.r.route = is_small(it->rq) ? 0 : 2;
                        fill_common(d.r, *it, now(), now());
                        done[it->ticket] = std::move(d);
                        cv_done.notify_all();
                        it = pending.erase(it);
                    } else {
                        ++it;
                    }
                }
                if (stop && pending.empty()) break;
                // continuous batching: every pending small request that fits
                {
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
                        break;
                    }
                }
                // else the first admissible large request in arrival order
                for (auto it = pending.begin(); it != pending.end(); ++it) {
                    const oocdag_request &r = it->rq;
                    if (is_small(r)) continue;
                    CacheEntry *e = find(r.matrix_id);
                    if (e && !e->ready) continue;  // being factored by another slot
                    bool is_hit = e != nullptr;
                    size_t need = large_bytes(r, is_hit);
                    if (!is_hit && need > budget) {
                        // out of core, inside a reservation of its own
                        need = ooc_bytes(r);
                        if (!make_room(need, s)) continue;
                        used += need;
                        reserved = need;
                        big = *it;
                        pending.erase(it);
                        have_big = true;
                        ooc_route = true;
                        break;
                    }
                    // pin a hit before making room, so the eviction cannot
                    // pick the entry this request is about to read
                    if (is_hit) e->users++;
                    if (!make_room(need, s)) {
                        if (is_hit) e->users--;
                        continue;
                    }
                    used += need;
                    reserved = need;
                    big = *it;
                    pending.erase(it);
                    have_big = true;
                    hit = is_hit;
                    if (is_hit) {
                        ce = e;
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
            run_ooc(big, reserved, s, t_start);
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
            for (auto &r : outbox()) done[r.first] = std::move(r.second);
            outbox().clear();
            cv_work.notify_all();
            cv_done.notify_all();
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
    // small systems: FP64 factor whatever precision was asked for (FP32 +
    // refinement does not pay at these orders); the batched kernel still
    // refines when the error is above tol, the tightest tol of the batch
    const int prec = 0;
    double tol = 1.0;
    for (auto &j : jobs) tol = std::min(tol, j.rq.tol);
    std::vector<double> berr(k), lres(k);
    std::vector<int> iters(k), status(k);
    oocdag_batch_stats st;
    if (rc == OOCDAG_OK)
        rc = oocdag_batched_solve(k, n.data(), r.data(), a_off.data(), b_off.data(), A, B, X, prec,
                                  tol, 10, 0, berr.data(), lres.data(), iters.data(),
                                  status.data(), &st);
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
        d.r.solve_resid = lres[i];
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
        // cached factor: refinement solve only.  A stall refactors in FP64
        // into a factor of this request's own, inside extra budget reserved
        // for it; other requests may be solving with the cached factor at
        // the same time, so the entry takes the FP64 factor (and the
        // low-precision one is freed) only when this request is its sole user.
        const double ts = now();
        int it = 0;
        double eta = 0, lres = 0;
        int r = mp_solve(ce->f, B, q.nrhs, ldb, X, ldb, q.tol, 30, s, &it, &eta, &lres);
        st.precision_used = ce->f.precision;
        if (r == 1 && ce->f.precision != 0) {
            const size_t extra = mp_factor_bytes(n, nb);
            bool got;
            {
                std::lock_guard<std::mutex> lk(mu);
                got = make_room(extra, s);
                if (got) used += extra;
            }
            if (!got) {
                rc = OOCDAG_EBUDGET;
            } else {
                MpFactor f64 = ce->f;
                f64.W = f64.Z = f64.dinv = nullptr;
                f64.W32 = f64.Z32 = nullptr;
                rc = mp_factor(f64, ce->A, n, nb, 0, s, s1);
                if (rc == OOCDAG_OK) {
                    r = mp_solve(f64, B, q.nrhs, ldb, X, ldb, q.tol, 30, s, &it, &eta, &lres);
                    st.fell_back = 1;
                    st.precision_used = 0;
                }
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (rc == OOCDAG_OK && ce->users == 1) std::swap(ce->f, f64);
                    used -= extra;
                }
                mp_release(f64, s);   // the replaced factor, or this request's own
            }
        }
        if (r < 0) rc = r;
        st.iterations = it;
        st.backward_error = eta;
        st.solve_resid = lres;
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
    d.r.solve_resid = st.solve_resid;
    double cs = 0;
    for (double v : d.x) cs += v;
    d.r.x_checksum = cs;
    fill_common(d.r, j, t_start, now());
    d.r.input_seconds = t_input;
    finish(j.ticket, std::move(d));
}

// ---- out-of-core route ----------------------------------------------------

namespace {

// ||A||_inf row sums from one tile column Ac of a symmetric matrix (tiles
// j..bb-1 of edge nb, consecutive, the diagonal tile full), rows offset to
// the column's first row: the row part of every row of the column, then
// the column sums below the diagonal tile, which are the row sums of A's
// rows j*nb.. over the columns right of the tile.  Two launches on one
// stream, one thread per row or column: a fixed summation order.
__global__ void k_tilecol_rowabs(const double *Ac, int nb, int ntile, double *rowabs) {
    const long long rows = (long long)ntile * nb;
    for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < rows;
         r += (long long)gridDim.x * blockDim.x) {
        const long long t = r / nb, rt = r - t * nb;
        const double *p = Ac + t * nb * nb + rt;
        double s = 0.0;
        for (int c = 0; c < nb; ++c) s += fabs(p[(long long)c * nb]);
        rowabs[r] += s;
    }
}

__global__ void k_tilecol_colabs(const double *Ac, int nb, int ntile, double *rowabs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nb) return;
    double s = 0.0;
    for (int t = 1; t < ntile; ++t) {
        const double *p = Ac + (long long)t * nb * nb + (long long)c * nb;
        for (int r = 0; r < nb; ++r) s += fabs(p[r]);
    }
    rowabs[c] += s;
}

__global__ void k_rows_absmax_t(const double *Xt, long long ldx, int nrhs, long long N,
                                double *out) {
    const int k = blockIdx.x;
    double m = 0.0;
    for (long long c = threadIdx.x; c < N; c += blockDim.x) m = fmax(m, fabs(Xt[k + c * ldx]));
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s2]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
}

// Parallel host copy (the matrix kept for the residual).
void host_copy(double *dst, const double *src, size_t n) {
    const size_t chunk = (size_t)1 << 20;
    const long long nch = (long long)((n + chunk - 1) / chunk);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < nch; ++c) {
        const size_t o = (size_t)c * chunk;
        memcpy(dst + o, src + o, std::min(chunk, n - o) * sizeof(double));
    }
}

}  // namespace

// A request too large for the budget: every phase keeps its device memory
// inside `reserved` bytes.
//   1. A's tile columns are generated one at a time on the device (G
//      streamed in column blocks) into host packed tiles T, and copied to
//      Ah for the residual;
//   2. T is factored out of core in FP64 (oocdag_cholesky_host_tiles, the
//      device budget = the reservation);
//   3. L L^T X = B is solved on transposed right-hand sides X^T (nrhs x N)
//      resident on the device, L streamed one tile column at a time:
//      forward over j (X_j <- X_j L_jj^-T, then X_i -= X_j L_ij^T below),
//      backward over j descending (X_j -= sum_i>j X_i L_ij, then
//      X_j <- X_j L_jj^-1);
//   4. R = B - A X and ||A||_inf with A's tile columns streamed from Ah
//      (tile column j gives rows j.. through L-like products and, by
//      symmetry, the rows of tile j through the tiles below it).
void oocdag_server::run_ooc(Job &j, size_t reserved, hipStream_t s, double t_start) {
    const oocdag_request &q = j.rq;
    const int n = q.n, nb = tile_nb(n), bb = (n + nb - 1) / nb;
    const long long Np = (long long)bb * nb, ldt = q.nrhs + (q.nrhs & 1);
    const int ldb = n + (n & 1);
    const size_t tile = (size_t)nb * nb;
    const size_t col_elems = (size_t)bb * tile;
    Done d;
    d.r.route = 2;
    d.r.batch_size = 1;
    d.r.precision_used = 0;
    int rc = OOCDAG_OK;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    std::vector<double> T(oocdag_tiles_elems(n, nb)), Ah;
    double *col = nullptr, *Gp = nullptr;
    // 1. generation
    if (!ok(hipMallocAsync((void **)&col, col_elems * 8, s)) ||
        !ok(hipMallocAsync((void **)&Gp, col_elems * 8, s)))
        rc = OOCDAG_ENOMEM;
    for (int jt = 0; jt < bb && rc == OOCDAG_OK; ++jt) {
        if (!ok(gram_tile_column(col, n, nb, jt, q.mu, q.seed, Gp, nb, s)) ||
            !ok(hipMemcpyAsync(T.data() + (size_t)tile_index(jt, jt, bb, 1) * tile, col,
                               (size_t)(bb - jt) * tile * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipStreamSynchronize(s)))
            rc = OOCDAG_EHIP;
    }
    if (Gp) (void)hipFreeAsync(Gp, s);
    if (col) (void)hipFreeAsync(col, s);
    (void)hipStreamSynchronize(s);
    if (rc == OOCDAG_OK) {
        Ah.resize(T.size());
        host_copy(Ah.data(), T.data(), T.size());
    }
    const double t_input = now() - t_start;
    // 2. FP64 out-of-core factorization inside the reservation
    if (rc == OOCDAG_OK) {
        int info = 0;
        double t = oocdag_cholesky_host_tiles(T.data(), n, nb, reserved, &info);
        if (t < 0 || info) rc = info ? info : OOCDAG_EHIP;
    }
    // 3. streamed solve
    double *Bd = nullptr, *Xt = nullptr, *Rt = nullptr, *dinv = nullptr, *nrm = nullptr,
           *rowabs = nullptr;
    const size_t vb = (size_t)(ldt * Np) * 8;
    if (rc == OOCDAG_OK &&
        (!ok(hipMallocAsync((void **)&Bd, (size_t)ldb * q.nrhs * 8, s)) ||
         !ok(hipMallocAsync((void **)&Xt, vb, s)) || !ok(hipMallocAsync((void **)&Rt, vb, s)) ||
         !ok(hipMallocAsync((void **)&dinv, (size_t)Np * 64 * 8, s)) ||
         !ok(hipMallocAsync((void **)&nrm, (size_t)3 * ldt * 8, s)) ||
         !ok(hipMallocAsync((void **)&rowabs, (size_t)Np * 8, s)) ||
         !ok(hipMallocAsync((void **)&col, col_elems * 8, s))))
        rc = OOCDAG_ENOMEM;
    const GemmKind nt = q.nrhs <= 64 ? GEMM_RHS : GEMM_GENERAL;
    const GemmKind tt = q.nrhs <= 64 ? GEMM_RHS_TB : GEMM_GENERAL_TB;
    auto load_col = [&](const std::vector<double> &H, int jt) -> bool {
        return ok(hipMemcpyAsync(col, H.data() + (size_t)tile_index(jt, jt, bb, 1) * tile,
                                 (size_t)(bb - jt) * tile * 8, hipMemcpyHostToDevice, s));
    };
    if (rc == OOCDAG_OK &&
        (!ok(normal_block(Bd, n, q.nrhs, ldb, q.rhs_seed, 2, s)) ||
         !ok(hipMemsetAsync(Xt, 0, vb, s)) || !ok(hipMemsetAsync(rowabs, 0, (size_t)Np * 8, s)) ||
         !ok(launch_transpose(Bd, ldb, Xt, ldt, n, q.nrhs, s)) ||
         !ok(hipMemcpyAsync(Rt, Xt, vb, hipMemcpyDeviceToDevice, s))))
        rc = OOCDAG_EHIP;
    if (rc == OOCDAG_OK) {   // ||b_k||_inf, before X^T overwrites B^T
        hipLaunchKernelGGL(k_rows_absmax_t, dim3(q.nrhs), dim3(256), 0, s, Xt, ldt, q.nrhs, Np,
                           nrm + 2 * ldt);
        if (!ok(hipGetLastError())) rc = OOCDAG_EHIP;
    }
    for (int jt = 0; jt < bb && rc == OOCDAG_OK; ++jt) {   // forward
        const Opnd Lc = panel(col, nb, bb - jt, 0);
        if (!load_col(T, jt) || !ok(diag_inverses(Lc, nb, dinv + (size_t)jt * nb * 64, s)) ||
            !ok(trsm_right_lower_t(plain(Xt, ldt, 0, jt * nb), q.nrhs, nb, Lc,
                                   dinv + (size_t)jt * nb * 64, s))) {
            rc = OOCDAG_EHIP;
            break;
        }
        if (jt + 1 < bb) {
            Gemm g;
            g.A = plain(Xt, ldt, 0, jt * nb);
            g.B = Lc; g.B.r0 = nb;
            g.C = plain(Xt, ldt, 0, (jt + 1) * nb);
            g.M = q.nrhs; g.N = (bb - jt - 1) * nb; g.K = nb;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if (!ok(launch_gemm(g, nt, s))) rc = OOCDAG_EHIP;
        }
    }
    for (int jt = bb - 1; jt >= 0 && rc == OOCDAG_OK; --jt) {   // backward
        const Opnd Lc = panel(col, nb, bb - jt, 0);
        if (!load_col(T, jt)) {
            rc = OOCDAG_EHIP;
            break;
        }
        if (jt + 1 < bb) {
            Gemm g;
            g.A = plain(Xt, ldt, 0, (jt + 1) * nb);
            g.B = Lc; g.B.r0 = nb;   // L[(jt+1)*nb.., jt], read transposed
            g.C = plain(Xt, ldt, 0, jt * nb);
            g.M = q.nrhs; g.N = nb; g.K = (bb - jt - 1) * nb;
            g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
            if (!ok(launch_gemm(g, tt, s))) rc = OOCDAG_EHIP;
        }
        if (rc == OOCDAG_OK &&
            (!ok(trsm_right_lower_n(plain(Xt, ldt, 0, jt * nb), q.nrhs, nb, Lc,
                                    dinv + (size_t)jt * nb * 64, s)) ||
             !ok(hipStreamSynchronize(s))))
            rc = OOCDAG_EHIP;
    }
    // 4. residual R^T = B^T - X^T A and ||A||_inf, A streamed from Ah
    for (int jt = 0; jt < bb && rc == OOCDAG_OK; ++jt) {
        const Opnd Ac = panel(col, nb, bb - jt, 0);
        if (!load_col(Ah, jt)) {
            rc = OOCDAG_EHIP;
            break;
        }
        Gemm g;   // rows jt.. of A (columns of tile jt): R_i -= A_ij X_j
        g.A = plain(Xt, ldt, 0, jt * nb);
        g.B = Ac;
        g.C = plain(Rt, ldt, 0, jt * nb);
        g.M = q.nrhs; g.N = (bb - jt) * nb; g.K = nb;
        g.alpha = -1.0; g.beta = 1.0; g.lower = 0;
        if (!ok(launch_gemm(g, nt, s))) rc = OOCDAG_EHIP;
        if (rc == OOCDAG_OK && jt + 1 < bb) {   // R_j -= sum_i>j A_ij^T X_i
            Gemm h;
            h.A = plain(Xt, ldt, 0, (jt + 1) * nb);
            h.B = Ac; h.B.r0 = nb;
            h.C = plain(Rt, ldt, 0, jt * nb);
            h.M = q.nrhs; h.N = nb; h.K = (bb - jt - 1) * nb;
            h.alpha = -1.0; h.beta = 1.0; h.lower = 0;
            if (!ok(launch_gemm(h, tt, s))) rc = OOCDAG_EHIP;
        }
        hipLaunchKernelGGL(k_tilecol_rowabs, dim3(256), dim3(256), 0, s, col, nb, bb - jt,
                           rowabs + (size_t)jt * nb);
        hipLaunchKernelGGL(k_tilecol_colabs, dim3((nb + 255) / 256), dim3(256), 0, s, col, nb,
                           bb - jt, rowabs + (size_t)jt * nb);
        if (rc == OOCDAG_OK && (!ok(hipGetLastError()) || !ok(hipStreamSynchronize(s))))
            rc = OOCDAG_EHIP;
    }
    std::vector<double> hx, hxn(q.nrhs), hrn(q.nrhs), hbn(q.nrhs), hra(Np);
    if (rc == OOCDAG_OK) {
        hipLaunchKernelGGL(k_rows_absmax_t, dim3(q.nrhs), dim3(256), 0, s, Xt, ldt, q.nrhs, Np, nrm);
        hipLaunchKernelGGL(k_rows_absmax_t, dim3(q.nrhs), dim3(256), 0, s, Rt, ldt, q.nrhs, Np,
                           nrm + ldt);
        hx.resize((size_t)n * q.nrhs);
        if (!ok(launch_transpose(Xt, ldt, Bd, ldb, q.nrhs, n, s)) ||
            !ok(hipMemcpy2DAsync(hx.data(), (size_t)n * 8, Bd, (size_t)ldb * 8, (size_t)n * 8,
                                 q.nrhs, hipMemcpyDeviceToHost, s)) ||
            !ok(hipMemcpyAsync(hxn.data(), nrm, q.nrhs * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipMemcpyAsync(hrn.data(), nrm + ldt, q.nrhs * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipMemcpyAsync(hbn.data(), nrm + 2 * ldt, q.nrhs * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipMemcpyAsync(hra.data(), rowabs, (size_t)Np * 8, hipMemcpyDeviceToHost, s)) ||
            !ok(hipStreamSynchronize(s)))
            rc = OOCDAG_EHIP;
    }
    if (rc == OOCDAG_OK) {
        const double an = *std::max_element(hra.begin(), hra.begin() + n);   // ||A||_inf
        double eta = 0;
        for (int k = 0; k < q.nrhs; ++k) eta = std::max(eta, hrn[k] / (an * hxn[k] + hbn[k]));
        d.r.backward_error = eta;
        d.r.solve_resid = pdlaschk_resid(hrn.data(), hxn.data(), q.nrhs, an, n);
    }
    for (double *p : {Bd, Xt, Rt, dinv, nrm, rowabs, col})
        if (p) (void)hipFreeAsync(p, s);
    (void)hipStreamSynchronize(s);
    d.r.status = rc;
    d.x = std::move(hx);
    double cs = 0;
    for (double v : d.x) cs += v;
    d.r.x_checksum = cs;
    fill_common(d.r, j, t_start, now());
    d.r.input_seconds = t_input;
    finish(j.ticket, std::move(d));
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" oocdag_server *oocdag_server_create(size_t hbm_budget_bytes, int slots) {
    if (slots < 1 || hbm_budget_bytes == 0) return nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nullptr;
    // the FP32/BF16 MFMA row-order probes allocate and synchronize once per
    // process: run them here, not inside the first request that needs them
    {
        hipStream_t ps = nullptr;
        if (hipStreamCreateWithFlags(&ps, hipStreamNonBlocking) != hipSuccess) return nullptr;
        const bool probed = mfma_rowmap(0, ps) >= 0 && mfma_rowmap(1, ps) >= 0;
        (void)hipStreamDestroy(ps);
        if (!probed) return nullptr;
    }
    // Requests allocate stream-ordered (hipMallocAsync) from the device's
    // default pool, which by default hands freed memory back to the driver at
    // every synchronization, so each request's allocations went to the
    // driver again (0.1-0.3 ms each).  The pool now keeps up to 1 GiB (or the
    // budget, if smaller) of freed memory for reuse; larger blocks, such as
    // an out-of-core panel, are still returned.  The threshold is raised,
    // never lowered, and stays for the life of the process.
    {
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = std::min<uint64_t>(hbm_budget_bytes, 1ull << 30), cur = 0;
            if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &cur) == hipSuccess &&
                cur < keep)
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    oocdag_server *srv = new oocdag_server();
    srv->budget = hbm_budget_bytes;
    srv->nslots = slots;
    for (int i = 0; i < slots; ++i) srv->workers.emplace_back([srv, i] { srv->worker(i); });
    return srv;
}

extern "C" int oocdag_server_estimate(const oocdag_request *rq, size_t *hit_bytes,
                                      size_t *miss_bytes, size_t *cached_bytes,
                                      size_t *ooc_min_bytes) {
    if (!rq || rq->n < 1 || rq->nrhs < 1) return OOCDAG_EARG;
    const bool small = is_small(*rq);
    const int nb = tile_nb(rq->n);
    if (hit_bytes) *hit_bytes = small ? oocdag_server::small_bytes(*rq)
                                      : oocdag_server::large_bytes(*rq, true);
    if (miss_bytes) *miss_bytes = small ? oocdag_server::small_bytes(*rq)
                                        : oocdag_server::large_bytes(*rq, false);
    if (cached_bytes)
        *cached_bytes = small ? 0 : tiles_bytes(rq->n, nb) + mp_factor_bytes(rq->n, nb);
    if (ooc_min_bytes) *ooc_min_bytes = small ? 0 : oocdag_server::ooc_min_bytes(*rq);
    return OOCDAG_OK;
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
