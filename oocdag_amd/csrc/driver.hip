// Host side of liboocdag: the C ABI, the panel planner and the out-of-core
// scheduler on HIP streams and events.
//
// The reference runs Cholesky() (oocdag.cpp:18-112) as a QUARK task DAG on
// four host threads with cuBLAS on the legacy NULL stream and pageable
// synchronous copies (oocquark.cpp:16-63, oocdag.cpp:279-289).  Here one host
// thread queues all kernels on four HIP streams:
//   h2d  : pinned-slot -> HBM copies of Y panels and X tile columns,
//   comp : updates and the diagonal factorization (critical path),
//   bulk : the look-ahead bulk update of each panel,
//   d2h  : factored tile columns back to the host,
// ordered by events: an X buffer is refilled only after the update that read
// it (double-buffered X), a factored tile column is copied out as soon as it
// is final while later columns are still being computed, host packing of the
// next X column overlaps the update running on the current one, and the next
// panel is copied into the device tiles of finished columns while the current
// panel is factored.  Two helper threads keep host staging off the queueing
// thread: the loader packs and queues the next panel's copies, the unpacker
// empties device->host slots into the caller's array.
#include "ooc_internal.h"
#include "../../include/oocdag.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

using namespace ooc;

namespace {

thread_local std::string g_err;
thread_local double g_stat[5];

// Optional instrumentation of the dominant kernel (the 128x128 panel-update
// GEMM): an event pair around every launch on its own stream, summed after
// the call.  Process-wide; meant for single-threaded benchmarking.
struct UpdateProfile {
    bool on = false;
    std::vector<hipEvent_t> ev;   // start, end pairs
    std::vector<double> flops;    // algorithmic flops of each launch
    std::vector<double> bytes;    // algorithmic HBM bytes of each launch
    long long launches = 0;
    double seconds = 0, total_flops = 0, total_bytes = 0;
} g_prof;

int g_progress_every = 0;  // print a line every this many factored tile columns

// Useful flops of a lower-trapezoid update: M x N output (diagonal of the
// region on its top-left corner), lower part only, K-deep.
double trap_update_flops(long long M, long long N, long long K) {
    return 2.0 * (double)K * ((double)M * N - 0.5 * (double)N * (N - 1));
}

// Algorithmic HBM bytes of a lower-trapezoid update: the C elements read
// and written once (16 B each) and the K-deep operand strips read once
// (8 B per element; B is the top rows of A when both describe the same
// operand, as in every update of the factorization).
double trap_update_bytes(const Gemm &g) {
    const double M = g.M, N = g.N, K = g.K;
    const double c = M * N - 0.5 * N * (N - 1);
    const bool b_in_a = g.B.p == g.A.p && g.B.r0 == g.A.r0 && g.B.c0 == g.A.c0 &&
                        g.B.ts == g.A.ts && g.B.ym == g.A.ym && g.B.tri == g.A.tri;
    return 16.0 * c + 8.0 * M * K + (b_in_a ? 0.0 : 8.0 * N * K);
}

hipError_t launch_update(const Gemm &g, hipStream_t s) {
    if (!g_prof.on) return launch_gemm(g, GEMM_UPDATE, s);
    hipEvent_t a, b;
    hipError_t e;
    if ((e = hipEventCreate(&a)) != hipSuccess) return e;
    if ((e = hipEventCreate(&b)) != hipSuccess) return e;
    if ((e = hipEventRecord(a, s)) != hipSuccess) return e;
    if ((e = launch_gemm(g, GEMM_UPDATE, s)) != hipSuccess) return e;
    if ((e = hipEventRecord(b, s)) != hipSuccess) return e;
    g_prof.ev.push_back(a);
    g_prof.ev.push_back(b);
    g_prof.flops.push_back(trap_update_flops(g.M, g.N, g.K));
    g_prof.bytes.push_back(trap_update_bytes(g));
    return hipSuccess;
}

// Folds finished event pairs into the totals (call after a synchronize).
void profile_collect() {
    for (size_t i = 0; i + 1 < g_prof.ev.size(); i += 2) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, g_prof.ev[i], g_prof.ev[i + 1]) == hipSuccess) {
            g_prof.seconds += ms * 1e-3;
            g_prof.total_flops += g_prof.flops[i / 2];
            g_prof.total_bytes += g_prof.bytes[i / 2];
            g_prof.launches += 1;
        }
        (void)hipEventDestroy(g_prof.ev[i]);
        (void)hipEventDestroy(g_prof.ev[i + 1]);
    }
    g_prof.ev.clear();
    g_prof.flops.clear();
    g_prof.bytes.clear();
}

void set_err(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define HIPTRY(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_err("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));   \
            return OOCDAG_EHIP;                                                          \
        }                                                                                \
    } while (0)

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int device_ok() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_err("no HIP device");
        return 0;
    }
    return 1;
}

// Tiles of the packed lower trapezoid with Ym tile rows and Yn tile columns.
long long trap_tiles(int Ym, int Yn) { return (long long)Yn * Ym - ((long long)Yn * (Yn - 1)) / 2; }

// find_Yn (oocdag.cpp:478-492) with the reserved X space generalised from one
// tile column (nx = 1, the reference) to nx columns.  The function-static of
// the reference becomes the explicit 'prev'.
int find_width(int bb, long long memBlock, int jb, int nx, int prev) {
    int Ym = bb - jb;
    long long room = memBlock - (long long)nx * Ym;  // tiles left for the panel
    if ((long long)Ym * (Ym + 1) <= room * 2) return Ym;
    int Yn = prev;
    while (room * 2 >= (long long)(2 * Ym - Yn) * (Yn + 1)) {
        if (Yn == Ym) break;
        ++Yn;
    }
    return Yn;
}

struct Plan {
    std::vector<int> width;  // tile columns per panel
    long long tiles_needed = 0;
    int nx = 1;
};

// Builds the panel schedule.  Two X buffers when the budget leaves room for
// at least one panel column next to them, else the reference's single one.
// force_nx > 0 fixes the X buffer count (1 reproduces find_Yn exactly).
bool make_plan(int N, int NB, size_t memsize, Plan &pl, int force_nx = 0) {
    int bb = (N + NB - 1) / NB;
    // w = memsize/sizeof(double)/BB/NB/NB - 1; assert(w > 0)  (oocdag.cpp:248-250)
    long long w = (long long)(memsize / sizeof(double) / (size_t)bb / (size_t)NB / (size_t)NB) - 1;
    if (w <= 0) return false;
    long long memBlock = (long long)(memsize / sizeof(double) / (size_t)NB / (size_t)NB);
    for (int nx = 2; nx >= 1; --nx) {
        if (force_nx > 0 && nx != force_nx) continue;
        pl.width.clear();
        pl.tiles_needed = 0;
        pl.nx = nx;
        bool ok = true;
        int prev = 0;
        for (int jb = 0; jb < bb;) {
            int Yn = find_width(bb, memBlock, jb, nx, prev);
            if (Yn <= 0) {
                ok = false;
                break;
            }
            int Ym = bb - jb;
            long long need = trap_tiles(Ym, Yn) + (jb > 0 ? (long long)nx * Ym : 0);
            pl.tiles_needed = std::max(pl.tiles_needed, need);
            pl.width.push_back(Yn);
            prev = Yn;
            jb += Yn;
        }
        if (ok) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// host packing between column-major A (LDA) and NB x NB tiles
// ---------------------------------------------------------------------------

// Packs tiles (ti0 .. bb-1, tj) of A into dst: consecutive slots of edge
// ts = tile_edge(NB) (ld = ts), each holding an NB x NB tile of A, padded
// past N and past NB inside the slot with 0 and a 1 on the diagonal of tile
// (tj, tj).
void pack_column(const double *A, int N, int NB, int LDA, int tj, int ti0, double *dst,
                 int nth = 0) {
    const int bb = (N + NB - 1) / NB;
    const int ts = tile_edge(NB);
    const int ntile = bb - ti0;
    const size_t ts2 = (size_t)ts * ts;
    if (nth <= 0) nth = omp_get_max_threads();
#pragma omp parallel for schedule(static) collapse(2) num_threads(nth)
    for (int t = 0; t < ntile; ++t) {
        for (int c = 0; c < ts; ++c) {
            const int ti = ti0 + t;
            const int gr0 = ti * NB, gc = tj * NB + c;
            double *d = dst + (size_t)t * ts2 + (size_t)c * ts;
            const int vr = std::max(0, std::min(NB, N - gr0));
            if (c < NB && gc < N) {
                if (vr > 0) memcpy(d, A + (size_t)gc * LDA + gr0, (size_t)vr * sizeof(double));
                for (int r = vr; r < ts; ++r) d[r] = 0.0;
            } else {
                for (int r = 0; r < ts; ++r) d[r] = (ti == tj && r == c) ? 1.0 : 0.0;
            }
        }
    }
}

// Writes tiles (ti0 .. bb-1, tj) from slots src back into A: valid rows and
// columns only, and only the lower triangle of the diagonal tile.
void unpack_column(double *A, int N, int NB, int LDA, int tj, int ti0, const double *src,
                   int nth = 0) {
    const int bb = (N + NB - 1) / NB;
    const int ts = tile_edge(NB);
    const int ntile = bb - ti0;
    const size_t ts2 = (size_t)ts * ts;
    if (nth <= 0) nth = omp_get_max_threads();
#pragma omp parallel for schedule(static) collapse(2) num_threads(nth)
    for (int t = 0; t < ntile; ++t) {
        for (int c = 0; c < NB; ++c) {
            const int ti = ti0 + t;
            const int gr0 = ti * NB, gc = tj * NB + c;
            if (gc >= N) continue;
            const double *s = src + (size_t)t * ts2 + (size_t)c * ts;
            int r0 = (ti == tj) ? c : 0;
            int vr = std::max(0, std::min(NB, N - gr0));
            if (vr > r0)
                memcpy(A + (size_t)gc * LDA + gr0 + r0, s + r0, (size_t)(vr - r0) * sizeof(double));
        }
    }
}

// Host copy split across the OpenMP threads (staging of host tiles that are
// not page-locked).
void par_memcpy(double *dst, const double *src, size_t n, int nth = 0) {
    const size_t chunk = (size_t)1 << 20;   // doubles
    const long long nch = (long long)((n + chunk - 1) / chunk);
    if (nth <= 0) nth = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nth)
    for (long long c = 0; c < nch; ++c) {
        const size_t o = (size_t)c * chunk;
        memcpy(dst + o, src + o, std::min(chunk, n - o) * sizeof(double));
    }
}

// A host thread that runs posted tasks in order (staging work that must not
// block the thread queueing kernels).  post() returns the task's sequence
// number; wait_done(q) returns once tasks 1..q have run; drain() waits for
// all of them and returns the first failure (its message is copied into this
// thread's error string).  After a failure, and for tasks still queued when
// the worker is destroyed, the remaining tasks are skipped.
class Worker {
  public:
    Worker() : th_([this] { loop(); }) {}
    ~Worker() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    long long post(std::function<int()> f) {
        std::lock_guard<std::mutex> lk(mu_);
        q_.push_back(std::move(f));
        cv_.notify_all();
        return ++posted_;
    }
    void wait_done(long long q) {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return finished_ >= q; });
    }
    int drain() {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return finished_ == posted_; });
        if (rc_) g_err = err_;
        return rc_;
    }

  private:
    void loop() {
        for (;;) {
            std::function<int()> f;
            bool skip;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
                skip = rc_ != 0 || stop_;   // after a failure or at teardown
            }
            const int r = skip ? 0 : f();
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (r && !rc_) {
                    rc_ = r;
                    err_ = g_err;   // this thread's message
                }
                ++finished_;
            }
            done_cv_.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::function<int()>> q_;
    long long posted_ = 0, finished_ = 0;
    int rc_ = 0;
    std::string err_;
    bool stop_ = false;
    std::thread th_;   // last: starts after the members above exist
};

// A monotone counter one thread raises and another waits on; abort() releases
// every waiter (wait() then returns false).
struct Gate {
    std::mutex mu;
    std::condition_variable cv;
    int value = -1;
    bool aborted = false;
    void set(int v) {
        {
            std::lock_guard<std::mutex> lk(mu);
            value = v;
        }
        cv.notify_all();
    }
    bool wait(int v) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return aborted || value >= v; });
        return !aborted;
    }
    void abort() {
        {
            std::lock_guard<std::mutex> lk(mu);
            aborted = true;
        }
        cv.notify_all();
    }
};

// True when p is page-locked host memory known to HIP.
bool host_pinned(const void *p) {
    hipPointerAttribute_t at;
    const hipError_t e = hipPointerGetAttributes(&at, p);
    (void)hipGetLastError();   // an unknown pointer is an error here, not a failure
    return e == hipSuccess && at.type == hipMemoryTypeHost;
}

struct Streams {
    hipStream_t h2d = nullptr, comp = nullptr, d2h = nullptr, bulk = nullptr;
    bool own_comp = true;
    ~Streams() {
        if (bulk) (void)hipStreamDestroy(bulk);
        if (h2d) (void)hipStreamDestroy(h2d);
        if (d2h) (void)hipStreamDestroy(d2h);
        if (comp && own_comp) (void)hipStreamDestroy(comp);
    }
};

// FP64 trailing update of a packed panel (ooc_syrk, oocquark.cpp:321-356).
struct Fp64Updater : PanelUpdater {
    Opnd P;
    int NB, Ym;
    hipError_t update(int j, int c0, int nc, hipStream_t s) override {
        Gemm g;
        g.A = P; g.A.r0 = c0 * NB; g.A.c0 = j * NB;
        g.B = g.A;
        g.C = P; g.C.r0 = c0 * NB; g.C.c0 = c0 * NB;
        g.M = (Ym - c0) * NB;
        g.N = nc * NB;
        g.K = NB;
        g.alpha = -1.0;
        g.beta = 1.0;
        g.lower = 1;
        return launch_update(g, s);
    }
};

// Factors tile columns [0, Yn) of a packed panel already resident in HBM
// (ooc_incore, oocquark.cpp:358-421) with look-ahead on two streams.
// gcol0 = global column of panel column 0.  on_column(j) is called after
// column j's factor kernels were queued on s0.
int factor_panel(double *dY, int NB, int Ym, int Yn, int gcol0, double *dinv, int *dinfo,
                 hipStream_t s0, hipStream_t s1, const std::function<int(int)> &on_column,
                 const std::function<int(int)> *after_updates = nullptr) {
    Fp64Updater up;
    up.P = panel(dY, NB, Ym, 1);
    up.NB = NB;
    up.Ym = Ym;
    const double tstart = now();
    std::vector<hipEvent_t> marks;
    auto progress = [&](int j) -> int {
        int rc = on_column(j);
        if (rc) return rc;
        if (g_progress_every > 0 && (j + 1) % g_progress_every == 0) {
            hipEvent_t e;
            HIPTRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            HIPTRY(hipEventRecord(e, s0));
            marks.push_back(e);
            if (marks.size() >= 2) {
                HIPTRY(hipEventSynchronize(marks.front()));
                (void)hipEventDestroy(marks.front());
                marks.erase(marks.begin());
                fprintf(stderr, "[oocdag] %d/%d tile columns factored, %.1f s\n",
                        j + 1 - g_progress_every, Yn, now() - tstart);
                fflush(stderr);
            }
        }
        return OOCDAG_OK;
    };
    int rc = factor_panel_la(up.P, NB, Ym, Yn, gcol0, dinv, 0, dinfo, s0, s1, up, progress,
                             after_updates);
    for (hipEvent_t e : marks) (void)hipEventDestroy(e);
    if (rc == -4) {
        set_err("HIP error in the panel factorization: %s", hipGetErrorString(hipGetLastError()));
        return OOCDAG_EHIP;
    }
    return rc;
}

// The first failed pivot (1-based column of the padded device matrix, or
// INT_MAX) as the LAPACK info of the N x N matrix: slot column h of tile
// column t is matrix column t*NB + h (NB = 0: no slots, the index as is).
int read_info(int *dinfo, hipStream_t s, int N, int *info, int NB = 0) {
    int h = INT_MAX;
    HIPTRY(hipMemcpyAsync(&h, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (h != INT_MAX && NB > 0) {
        const int ts = tile_edge(NB);
        const long long t = (h - 1) / ts, off = (h - 1) % ts;
        h = (int)std::min<long long>(INT_MAX - 1, t * NB + std::min<long long>(off, NB - 1) + 1);
    }
    *info = (h == INT_MAX) ? 0 : std::min(h, N);
    return OOCDAG_OK;
}

// Busy time of a copy stream: an event pair around every copy, summed when
// the call has finished.
struct CopyTimer {
    std::vector<hipEvent_t> ev;
    hipError_t copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
        hipEvent_t a, b;
        hipError_t e;
        if ((e = hipEventCreate(&a)) != hipSuccess) return e;
        if ((e = hipEventCreate(&b)) != hipSuccess) return e;
        if ((e = hipEventRecord(a, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(dst, src, bytes, kind, s)) != hipSuccess) return e;
        if ((e = hipEventRecord(b, s)) != hipSuccess) return e;
        ev.push_back(a);
        ev.push_back(b);
        return hipSuccess;
    }
    double seconds() {
        double t = 0;
        for (size_t i = 0; i + 1 < ev.size(); i += 2) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) t += ms * 1e-3;
        }
        return t;
    }
    ~CopyTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// Device scratch for the diagonal-block inverses and the info flag;
// stream-ordered (init(..., true)) when the owner releases it on its stream.
struct Scratch {
    double *dinv = nullptr;
    int *dinfo = nullptr;
    bool async = false;
    ~Scratch() {
        if (dinv) (void)hipFree(dinv);
        if (dinfo) (void)hipFree(dinfo);
    }
    void release(hipStream_t s) {
        if (dinv) (void)hipFreeAsync(dinv, s);
        if (dinfo) (void)hipFreeAsync(dinfo, s);
        dinv = nullptr;
        dinfo = nullptr;
    }
    int init(int NB, hipStream_t s, bool stream_ordered = false) {
        async = stream_ordered;
        const size_t b = (size_t)(NB / 64) * 64 * 64 * sizeof(double);
        if (async) {
            HIPTRY(hipMallocAsync((void **)&dinv, b, s));
            HIPTRY(hipMallocAsync((void **)&dinfo, sizeof(int), s));
        } else {
            HIPTRY(hipMalloc(&dinv, b));
            HIPTRY(hipMalloc(&dinfo, sizeof(int)));
        }
        int big = INT_MAX;
        HIPTRY(hipMemcpyAsync(dinfo, &big, sizeof(int), hipMemcpyHostToDevice, s));
        return OOCDAG_OK;
    }
};

// Process-wide pool of pinned host staging buffers (the out-of-core slots,
// the host-tiles generator): handed out best-fit, returned after use, so
// repeated calls -- benchmark steps, the server's out-of-core route --
// allocate page-locked memory once instead of per call.  At most
// kPinnedKeep idle buffers are kept; oocdag_release_staging frees them.
constexpr int kPinnedKeep = 10;
std::mutex g_pin_mu;
std::vector<std::pair<size_t, double *>> g_pin_free;   // idle buffers
std::unordered_map<double *, size_t> g_pin_size;        // size of every pool buffer

double *pinned_get(size_t bytes) {
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        int best = -1;
        for (int i = 0; i < (int)g_pin_free.size(); ++i)
            if (g_pin_free[i].first >= bytes &&
                (best < 0 || g_pin_free[i].first < g_pin_free[best].first))
                best = i;
        if (best >= 0) {
            double *p = g_pin_free[best].second;
            g_pin_free.erase(g_pin_free.begin() + best);
            return p;
        }
    }
    double *p = nullptr;
    if (hipHostMalloc((void **)&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_size[p] = bytes;
    return p;
}

void pinned_put(double *p) {
    if (!p) return;
    double *drop = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        g_pin_free.emplace_back(g_pin_size[p], p);
        if ((int)g_pin_free.size() > kPinnedKeep) {
            int k = 0;   // drop the smallest idle buffer
            for (int i = 1; i < (int)g_pin_free.size(); ++i)
                if (g_pin_free[i].first < g_pin_free[k].first) k = i;
            drop = g_pin_free[k].second;
            g_pin_free.erase(g_pin_free.begin() + k);
            g_pin_size.erase(drop);
        }
    }
    if (drop) (void)hipHostFree(drop);
}

// Frees every idle pooled buffer; returns how many.
int pinned_trim() {
    std::vector<double *> v;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (auto &e : g_pin_free) {
            v.push_back(e.second);
            g_pin_size.erase(e.second);
        }
        g_pin_free.clear();
    }
    for (double *p : v) (void)hipHostFree(p);
    return (int)v.size();
}

// Device tiles that tile column c of the next panel (Ym2 tile rows, packed
// from tile 0) occupies overlap tile columns 0..k of the current panel (Ym
// tile rows, Yn columns); returns k (-1: none), so the next panel's column
// may be copied in once those columns of the current panel are free.
int last_overlapping_column(int Ym, int Yn, int Ym2, int c) {
    const long long end = tile_index(c, c, Ym2, 1) + (Ym2 - c);
    int k = -1;
    for (int q = 0; q < Yn && tile_index(q, q, Ym, 1) < end; ++q) k = q;
    return k;
}

// Out-of-core factorization.  Host side: the column-major array A (LDA) of
// the reference entry point, packed through pinned slots, or -- tiles_host --
// A already holds the packed lower tiles (the device layout; pinned memory
// for full-speed DMA), so every tile-column copy is one contiguous DMA.
//
// Per panel (widths from the planner):
//   1. its tile columns are in HBM: prefetched during the previous panel
//      (step 4), or, for the first panel, copied by this thread -- its first
//      w0 tile columns are factored while the others arrive, each of which is
//      updated from those w0 columns on arrival;
//   2. left-looking update from every prior tile column, each re-read from
//      the host into one of the X buffers (A2X + ooc_syrk, oocquark.cpp:
//      301-356), the copy of the next X column overlapping the current update;
//   3. in-core factorization with look-ahead (ooc_incore, oocquark.cpp:
//      358-421); every finished tile column is copied back to the host at once
//      (into a pinned slot that the unpack thread empties into A, or straight
//      into pinned host tiles), so this thread never waits on a copy;
//   4. meanwhile the loader thread copies the next panel's tile columns into
//      the device tiles this panel no longer needs: a column's copy waits (on
//      the GPU) for the events that mark the overlapped columns of this panel
//      final, read by every update and copied out.
int cholesky_ooc(double *A, int N, int NB, int LDA, size_t memsize, int *info, double *secs,
                 bool tiles_host = false, int force_nx = 0) {
    Plan pl;
    if (!make_plan(N, NB, memsize, pl, force_nx)) {
        set_err("budget %zu bytes below two tile columns for N=%d NB=%d", memsize, N, NB);
        return OOCDAG_EBUDGET;
    }
    // the reference's tile grid and memBlock (make_plan); on the device and
    // in the staging slots each tile sits in an NBs x NBs slot
    const int bb = (N + NB - 1) / NB;
    const int NBs = tile_edge(NB);
    const size_t nb2 = (size_t)NBs * NBs;
    const size_t col_bytes = (size_t)bb * nb2 * sizeof(double);
    constexpr int ND = 3;   // device->host slots

    Streams st;
    HIPTRY(hipStreamCreateWithFlags(&st.h2d, hipStreamNonBlocking));
    HIPTRY(hipStreamCreateWithFlags(&st.comp, hipStreamNonBlocking));
    HIPTRY(hipStreamCreateWithFlags(&st.d2h, hipStreamNonBlocking));
    HIPTRY(hipStreamCreateWithFlags(&st.bulk, hipStreamNonBlocking));
    // host tiles in page-locked memory are copied directly; otherwise (and
    // for the column-major array) through pinned slots
    const bool direct = tiles_host && host_pinned(A);

    // Everything this call allocates is released by one guard after its own
    // four streams have drained (on an early return too): device memory is
    // stream-ordered (hipMallocAsync / hipFreeAsync, no device-wide
    // synchronization, so concurrent callers on other streams -- the
    // server's slots -- are not stalled) and the pinned staging slots go
    // back to a process-wide pool instead of hipHostFree.
    double *dY = nullptr;
    double *hslot_buf[2] = {nullptr, nullptr}, *dslot_buf[ND] = {};
    std::vector<hipEvent_t> events;   // every event made below, destroyed on return
    Scratch sc;
    struct Release {
        Streams *st;
        double **dY;
        Scratch *sc;
        double **h, **d;
        std::vector<hipEvent_t> *ev;
        ~Release() {
            for (hipStream_t q : {st->h2d, st->bulk, st->d2h, st->comp})
                if (q) (void)hipStreamSynchronize(q);
            if (*dY) (void)hipFreeAsync(*dY, st->comp);
            sc->release(st->comp);
            if (st->comp) (void)hipStreamSynchronize(st->comp);
            for (int i = 0; i < 2; ++i) pinned_put(h[i]);
            for (int i = 0; i < ND; ++i) pinned_put(d[i]);
            for (hipEvent_t e : *ev) (void)hipEventDestroy(e);
        }
    } release{&st, &dY, &sc, hslot_buf, dslot_buf, &events};
    if (hipMallocAsync((void **)&dY, (size_t)pl.tiles_needed * nb2 * sizeof(double), st.comp) !=
        hipSuccess) {
        (void)hipGetLastError();
        dY = nullptr;
        set_err("device allocation of %lld tiles failed", pl.tiles_needed);
        return OOCDAG_ENOMEM;
    }
    int rc = sc.init(NBs, st.comp, true);
    if (rc) return rc;

    // pinned slots: 2 for host->device, ND for device->host
    for (int i = 0; i < 2 + ND && !direct; ++i) {
        double **p = i < 2 ? &hslot_buf[i] : &dslot_buf[i - 2];
        if (!(*p = pinned_get(col_bytes))) {
            set_err("pinned host allocation of %zu bytes failed", col_bytes);
            return OOCDAG_ENOMEM;
        }
    }
    auto new_event = [&](hipEvent_t *e) -> hipError_t {
        hipError_t r = hipEventCreateWithFlags(e, hipEventDisableTiming);
        if (r == hipSuccess) events.push_back(*e);
        return r;
    };
    hipEvent_t ev_hslot[2], ev_xready[2], ev_xfree[2], ev_y, ev_xdone;
    for (auto &e : ev_hslot) HIPTRY(new_event(&e));
    for (auto &e : ev_xready) HIPTRY(new_event(&e));
    for (auto &e : ev_xfree) HIPTRY(new_event(&e));
    HIPTRY(new_event(&ev_y));
    HIPTRY(new_event(&ev_xdone));
    bool hslot_used[2] = {false, false};
    int hslot = 0;
    long long dslot_task[ND] = {};   // unpack task that last used each slot
    HIPTRY(hipStreamSynchronize(st.comp));

    const int nth = std::max(1, omp_get_max_threads());
    const int nth_half = std::max(1, nth / 2);
    double h2d_bytes = 0, d2h_bytes = 0;
    CopyTimer t_h2d, t_h2d_pre, t_d2h;
    Gate freed;   // highest tile column of the current panel whose free-events exist

    // One tile column (rows ti0.. of tile column tj) host -> device through
    // H2D slot hslot (packed first, with `threads` threads) or straight from
    // pinned host tiles.  before_copy() runs just before the copy is queued
    // (it adds the stream waits the copy must follow).
    auto h2d_column = [&](double *dst, int tj, int ti0, int threads, CopyTimer &timer,
                          const std::function<int()> &before_copy) -> int {
        const size_t bytes = (size_t)(bb - ti0) * nb2 * sizeof(double);
        const double *src = tiles_host ? A + (size_t)tile_index(ti0, tj, bb, 1) * nb2 : nullptr;
        if (!direct) {
            if (hslot_used[hslot]) HIPTRY(hipEventSynchronize(ev_hslot[hslot]));
            hslot_used[hslot] = true;
            if (tiles_host)
                par_memcpy(hslot_buf[hslot], src, bytes / sizeof(double), threads);
            else
                pack_column(A, N, NB, LDA, tj, ti0, hslot_buf[hslot], threads);
            src = hslot_buf[hslot];
        }
        int r = before_copy();
        if (r) return r;
        HIPTRY(timer.copy(dst, src, bytes, hipMemcpyHostToDevice, st.h2d));
        if (!direct) {
            HIPTRY(hipEventRecord(ev_hslot[hslot], st.h2d));
            hslot ^= 1;
        }
        h2d_bytes += bytes;
        return OOCDAG_OK;
    };
    const std::function<int()> no_wait = [] { return OOCDAG_OK; };
    // declared after everything their tasks use, so they are joined first;
    // the gate is released before that (an early return leaves the loader
    // waiting for columns that will not be factored)
    Worker unpacker, loader;
    struct AbortGate {
        Gate *g;
        ~AbortGate() { g->abort(); }
    } abort_gate{&freed};

    const double t0 = now();
    for (size_t p = 0, jb = 0; p < pl.width.size(); jb += pl.width[p], ++p) {
        const int Yn = pl.width[p];
        const int Ym = bb - (int)jb;
        // the first panel is split: its first w0 tile columns are factored
        // while the others are still being copied in (nothing to overlap the
        // copy of a whole first panel with otherwise)
        const int w0 = (p == 0 && Yn >= 4) ? std::min(8, Yn / 4) : 0;
        // --- A2Y (oocquark.cpp:253-275): the panel's tile columns -> Y (a
        // split first panel loads them one at a time below)
        if (p == 0) {
            for (int jj = 0; jj < (w0 > 0 ? 0 : Yn); ++jj)
                if ((rc = h2d_column(dY + (size_t)tile_index(jj, jj, Ym, 1) * nb2, (int)jb + jj,
                                     (int)jb + jj, nth, t_h2d, no_wait)))
                    return rc;
        } else if ((rc = loader.drain())) {   // prefetched while the previous panel ran
            return rc;
        }
        HIPTRY(hipEventRecord(ev_y, st.h2d));
        HIPTRY(hipStreamWaitEvent(st.comp, ev_y, 0));

        // --- left-looking updates from every prior tile column (oocdag.cpp:73-77)
        const long long ytiles = trap_tiles(Ym, Yn);
        for (int jjb = 0; jjb < (int)jb; ++jjb) {
            const int b = jjb % pl.nx;
            double *X = dY + (size_t)(ytiles + (long long)b * Ym) * nb2;
            if (jjb >= pl.nx) HIPTRY(hipStreamWaitEvent(st.h2d, ev_xfree[b], 0));
            // A2X (oocquark.cpp:301-319): rows jb.. of tile column jjb
            if ((rc = h2d_column(X, jjb, (int)jb, nth, t_h2d, no_wait))) return rc;
            HIPTRY(hipEventRecord(ev_xready[b], st.h2d));
            HIPTRY(hipStreamWaitEvent(st.comp, ev_xready[b], 0));
            // ooc_syrk(X, Y, YM, YN, NB) (oocquark.cpp:321-356) as one launch
            Gemm g;
            g.A = panel(X, NBs, Ym, 0);
            g.B = g.A;
            g.C = panel(dY, NBs, Ym, 1);
            g.M = Ym * NBs;
            g.N = Yn * NBs;
            g.K = NBs;
            g.alpha = -1.0;
            g.beta = 1.0;
            g.lower = 1;
            HIPTRY(launch_update(g, st.comp));
            HIPTRY(hipEventRecord(ev_xfree[b], st.comp));
        }
        HIPTRY(hipEventRecord(ev_xdone, st.comp));

        // per tile column of this panel: copied out (evd), no longer read on
        // s0 / s1 (ev0 / ev1)
        std::vector<hipEvent_t> evc(Yn), evd(Yn), ev0(Yn), ev1(Yn);
        for (int jj = 0; jj < Yn; ++jj) {
            HIPTRY(new_event(&evc[jj]));
            HIPTRY(new_event(&evd[jj]));
            HIPTRY(new_event(&ev0[jj]));
            HIPTRY(new_event(&ev1[jj]));
        }
        freed.set(-1);

        // --- the next panel's tile columns, queued to the loader thread
        // (which shares the host->device slots: only while this thread queues
        // no copies of its own)
        auto post_prefetch = [&]() {
            if (p + 1 >= pl.width.size()) return;
            const int jb2 = (int)jb + Yn, Ym2 = bb - jb2, Yn2 = pl.width[p + 1];
            for (int c = 0; c < Yn2; ++c) {
                const int need = last_overlapping_column(Ym, Yn, Ym2, c);
                double *dst = dY + (size_t)tile_index(c, c, Ym2, 1) * nb2;
                const hipEvent_t e0 = need >= 0 ? ev0[need] : nullptr;
                const hipEvent_t e1 = need >= 0 ? ev1[need] : nullptr;
                const hipEvent_t ed = need >= 0 ? evd[need] : nullptr;
                loader.post([&, dst, c, need, jb2, e0, e1, ed]() -> int {
                    return h2d_column(dst, jb2 + c, jb2 + c, nth_half, t_h2d_pre, [&]() -> int {
                        if (need >= 0) {
                            if (!freed.wait(need)) {
                                set_err("out-of-core factorization aborted");
                                return OOCDAG_EHIP;
                            }
                            HIPTRY(hipStreamWaitEvent(st.h2d, e0, 0));
                            HIPTRY(hipStreamWaitEvent(st.h2d, e1, 0));
                            HIPTRY(hipStreamWaitEvent(st.h2d, ed, 0));
                        }
                        HIPTRY(hipStreamWaitEvent(st.h2d, ev_xdone, 0));
                        return OOCDAG_OK;
                    });
                });
            }
        };

        // --- in-core factorization of the panel, each finished tile column
        // copied back while the next ones are computed (oocquark.cpp:389-419)
        const std::function<int(int)> on_column = [&](int jj) -> int {
            const int tj = (int)jb + jj;
            const size_t bytes = (size_t)(Ym - jj) * nb2 * sizeof(double);
            const double *src = dY + (size_t)tile_index(jj, jj, Ym, 1) * nb2;
            const int s = jj % ND;
            double *dst;
            if (direct) {
                dst = A + (size_t)tile_index(tj, tj, bb, 1) * nb2;
            } else {
                if (dslot_task[s]) unpacker.wait_done(dslot_task[s]);
                dst = dslot_buf[s];
            }
            HIPTRY(hipEventRecord(evc[jj], st.comp));
            HIPTRY(hipStreamWaitEvent(st.d2h, evc[jj], 0));
            HIPTRY(t_d2h.copy(dst, src, bytes, hipMemcpyDeviceToHost, st.d2h));
            HIPTRY(hipEventRecord(evd[jj], st.d2h));
            d2h_bytes += bytes;
            if (!direct) {
                const hipEvent_t done = evd[jj];
                const double *from = dslot_buf[s];
                const size_t elems = (size_t)(Ym - jj) * nb2;
                dslot_task[s] = unpacker.post([&, done, from, tj, elems]() -> int {
                    HIPTRY(hipEventSynchronize(done));
                    if (tiles_host)
                        par_memcpy(A + (size_t)tile_index(tj, tj, bb, 1) * nb2, from, elems,
                                   nth_half);
                    else
                        unpack_column(A, N, NB, LDA, tj, tj, from, nth_half);
                    return OOCDAG_OK;
                });
            }
            return OOCDAG_OK;
        };
        const std::function<int(int)> after_updates = [&](int jj) -> int {
            HIPTRY(hipEventRecord(ev0[jj], st.comp));
            HIPTRY(hipEventRecord(ev1[jj], st.bulk));
            freed.set(jj);
            return OOCDAG_OK;
        };
        if (w0 == 0) {
            post_prefetch();
            rc = factor_panel(dY, NBs, Ym, Yn, (int)jb * NBs, sc.dinv, sc.dinfo, st.comp, st.bulk,
                              on_column, &after_updates);
            if (rc) return rc;
        } else {
            // columns [0, w0) left-looking, each as soon as it is in HBM:
            // its update from the columns before it (one K = c*NBs launch),
            // then its blocked factorization; then each later column's
            // update from columns [0, w0) (one K = w0*NBs launch) on arrival;
            // then columns [w0, Yn): the trailing part of a packed panel is
            // itself a packed panel of Ym - w0 tile rows at
            // tile_index(w0, w0, Ym)
            const Opnd P0 = panel(dY, NBs, Ym, 1);
            for (int c = 0; c < w0; ++c) {
                if ((rc = h2d_column(dY + (size_t)tile_index(c, c, Ym, 1) * nb2, (int)jb + c,
                                     (int)jb + c, nth, t_h2d, no_wait)))
                    return rc;
                HIPTRY(hipEventRecord(ev_y, st.h2d));
                HIPTRY(hipStreamWaitEvent(st.comp, ev_y, 0));
                if (c > 0) {
                    Gemm g;
                    g.A = P0; g.A.r0 = c * NBs; g.A.c0 = 0;
                    g.B = g.A;
                    g.C = P0; g.C.r0 = c * NBs; g.C.c0 = c * NBs;
                    g.M = (Ym - c) * NBs;
                    g.N = NBs;
                    g.K = c * NBs;
                    g.alpha = -1.0;
                    g.beta = 1.0;
                    g.lower = 1;
                    HIPTRY(launch_update(g, st.comp));
                }
                HIPTRY(factor_column(P0, c * NBs, (Ym - c) * NBs, NBs, ((int)jb + c) * NBs, sc.dinv,
                                     sc.dinfo, st.comp));
                if ((rc = on_column(c))) return rc;
            }
            for (int c = w0; c < Yn; ++c) {
                if ((rc = h2d_column(dY + (size_t)tile_index(c, c, Ym, 1) * nb2, (int)jb + c,
                                     (int)jb + c, nth, t_h2d, no_wait)))
                    return rc;
                HIPTRY(hipEventRecord(ev_y, st.h2d));
                HIPTRY(hipStreamWaitEvent(st.comp, ev_y, 0));
                const Opnd P = panel(dY, NBs, Ym, 1);
                Gemm g;
                g.A = P; g.A.r0 = c * NBs; g.A.c0 = 0;
                g.B = g.A;
                g.C = P; g.C.r0 = c * NBs; g.C.c0 = c * NBs;
                g.M = (Ym - c) * NBs;
                g.N = NBs;
                g.K = w0 * NBs;
                g.alpha = -1.0;
                g.beta = 1.0;
                g.lower = 1;
                HIPTRY(launch_update(g, st.comp));
            }
            // those updates were the last readers of columns [0, w0)
            for (int jj = 0; jj < w0; ++jj) {
                HIPTRY(hipEventRecord(ev0[jj], st.comp));
                HIPTRY(hipEventRecord(ev1[jj], st.bulk));
            }
            freed.set(w0 - 1);
            post_prefetch();
            const std::function<int(int)> on_column_b = [&](int jj) { return on_column(jj + w0); };
            const std::function<int(int)> after_updates_b = [&](int jj) {
                return after_updates(jj + w0);
            };
            rc = factor_panel(dY + (size_t)tile_index(w0, w0, Ym, 1) * nb2, NBs, Ym - w0, Yn - w0,
                              ((int)jb + w0) * NBs, sc.dinv, sc.dinfo, st.comp, st.bulk,
                              on_column_b, &after_updates_b);
            if (rc) return rc;
        }
        // the next panel's X buffers may overwrite this panel's tiles and its
        // X columns are read from A: this panel must be finished and unpacked
        if ((rc = unpacker.drain())) return rc;
        HIPTRY(hipStreamSynchronize(st.comp));
        HIPTRY(hipStreamSynchronize(st.d2h));
    }
    if ((rc = loader.drain())) return rc;
    HIPTRY(hipStreamSynchronize(st.h2d));
    rc = read_info(sc.dinfo, st.comp, N, info, NB);
    *secs = now() - t0;
    profile_collect();
    g_stat[0] = *secs;
    g_stat[1] = h2d_bytes;
    g_stat[2] = d2h_bytes;
    g_stat[3] = t_h2d.seconds() + t_h2d_pre.seconds();
    g_stat[4] = t_d2h.seconds();
    return rc;
}

// Any NB >= 1, as Cholesky() takes (oocdag.cpp:241-247); a tile whose edge
// is not a multiple of 64 is stored in a padded slot (tile_edge).
int check_nb(int N, int NB) {
    if (N < 0 || NB < 1 || NB > (1 << 16)) {
        set_err("N=%d NB=%d: NB must be in [1, 65536]", N, NB);
        return OOCDAG_EARG;
    }
    return OOCDAG_OK;
}

// ---- general-matrix generator for the host-array entry point -------------

struct GenColArgs {
    unsigned pow2[40][2];
    unsigned x1;
    double *D;       // column-major M x ncols, ld M
    int M, N, col0, ncols;
    int mode;        // 0 general (j*M+i), 1 transposed (i*N+j), 2 symmetric, 3 'F'
};

__global__ void k_gen_cols(GenColArgs g) {
    long long total = (long long)g.M * g.ncols;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        int i = (int)(idx % g.M), jl = (int)(idx / g.M);
        int j = g.col0 + jl;
        double v;
        if (g.mode == 3) {
            v = (i == j) ? 10.0 : (double)0.1f;
        } else {
            unsigned long long p;
            if (g.mode == 0)
                p = (unsigned long long)j * g.M + i;
            else if (g.mode == 1)
                p = (unsigned long long)i * g.N + j;
            else
                p = (i >= j) ? (unsigned long long)j * g.M + i : (unsigned long long)i * g.N + j;
            unsigned x = g.x1;
            for (int k = 0; p; ++k, p >>= 1)
                if (p & 1ull) x = (g.pow2[k][0] * x + g.pow2[k][1]) & 0x7fffffffu;
            v = 1.0 - 2.0 * ((double)x / 2147483648.0);
        }
        g.D[idx] = v;
    }
}

void lcg_pow_host(unsigned long long n, unsigned &a, unsigned &c) {
    unsigned ra = 1, rc = 0, ba = 1103515245u, bc = 12345u;
    while (n) {
        if (n & 1ull) {
            unsigned na = (ba * ra) & 0x7fffffffu;
            unsigned nc = (ba * rc + bc) & 0x7fffffffu;
            ra = na;
            rc = nc;
        }
        unsigned na = (ba * ba) & 0x7fffffffu;
        unsigned nc = (ba * bc + bc) & 0x7fffffffu;
        ba = na;
        bc = nc;
        n >>= 1;
    }
    a = ra;
    c = rc;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *oocdag_last_error(void) { return g_err.c_str(); }

int oocdag_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int oocdag_last_stats(double *seconds, double *h2d_bytes, double *d2h_bytes, double *h2d_seconds,
                      double *d2h_seconds) {
    if (seconds) *seconds = g_stat[0];
    if (h2d_bytes) *h2d_bytes = g_stat[1];
    if (d2h_bytes) *d2h_bytes = g_stat[2];
    if (h2d_seconds) *h2d_seconds = g_stat[3];
    if (d2h_seconds) *d2h_seconds = g_stat[4];
    return OOCDAG_OK;
}

int oocdag_panel_plan(int N, int NB, size_t memsize, int *widths, int maxp) {
    return oocdag_panel_plan_nx(N, NB, memsize, 0, widths, maxp);
}

int oocdag_panel_plan_nx(int N, int NB, size_t memsize, int nx, int *widths, int maxp) {
    if (N <= 0 || NB <= 0 || nx < 0 || nx > 2) return OOCDAG_EARG;
    Plan pl;
    if (!make_plan(N, NB, memsize, pl, nx)) return OOCDAG_EBUDGET;
    for (size_t i = 0; i < pl.width.size() && (int)i < maxp; ++i) widths[i] = pl.width[i];
    return (int)pl.width.size();
}

double oocdag_cholesky_nx(double *A, int N, int NB, int LDA, size_t memsize, int nx, int *info) {
    g_err.clear();
    for (double &s : g_stat) s = 0;
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (!A || N < 0 || LDA < std::max(1, N) || nx < 0 || nx > 2 || check_nb(N, NB)) {
        if (g_err.empty()) set_err("invalid argument");
        *info = OOCDAG_EARG;
        return -1.0;
    }
    if (N == 0) return 0.0;
    if (!device_ok()) {
        *info = OOCDAG_ENODEV;
        return -1.0;
    }
    double secs = 0;
    int rc = cholesky_ooc(A, N, NB, LDA, memsize, info, &secs, false, nx);
    if (rc) {
        *info = rc;
        return -1.0;
    }
    return secs;
}

double oocdag_cholesky(double *A, int N, int NB, int LDA, size_t memsize, int *info) {
    return oocdag_cholesky_nx(A, N, NB, LDA, memsize, 0, info);
}

double oocdag_cholesky_host_tiles_nx(double *T, int N, int NB, size_t memsize, int nx,
                                     int *info) {
    g_err.clear();
    for (double &s : g_stat) s = 0;
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (!T || N < 0 || nx < 0 || nx > 2 || check_nb(N, NB)) {
        if (g_err.empty()) set_err("invalid argument");
        *info = OOCDAG_EARG;
        return -1.0;
    }
    if (N == 0) return 0.0;
    if (!device_ok()) {
        *info = OOCDAG_ENODEV;
        return -1.0;
    }
    double secs = 0;
    int rc = cholesky_ooc(T, N, NB, N, memsize, info, &secs, true, nx);
    if (rc) {
        *info = rc;
        return -1.0;
    }
    return secs;
}

double oocdag_cholesky_host_tiles(double *T, int N, int NB, size_t memsize, int *info) {
    return oocdag_cholesky_host_tiles_nx(T, N, NB, memsize, 0, info);
}

int oocdag_pdmatgen_host_tiles(double *T, int N, int NB, char aform, char diag, int iseed) {
    g_err.clear();
    if (!T || N <= 0 || NB <= 0) return OOCDAG_EARG;
    char af = (char)(aform & ~0x20), dg = (char)(diag & ~0x20);
    if (!(af == 'S' || af == 'H' || af == 'F') || !(dg == 'D' || dg == 'N')) {
        set_err("tile generator supports AFORM S/H/F and DIAG D/N");
        return OOCDAG_EARG;
    }
    if (!device_ok()) return OOCDAG_ENODEV;
    const int bb = (N + NB - 1) / NB;
    const size_t nb2 = (size_t)tile_edge(NB) * tile_edge(NB);   // one slot
    // groups of tile columns up to ~2 GB, generated on the device into two
    // alternating windows and copied down while the next group is generated.
    // Pinned T takes the copies directly; pageable T through two pinned
    // staging buffers that the host empties with a parallel memcpy (the
    // first touch of a large pageable array is then spread over all cores
    // instead of the runtime's own staging thread).
    const bool direct = host_pinned(T);
    const size_t cap = std::max<size_t>((size_t)bb * nb2, ((size_t)2 << 30) / sizeof(double));
    double *win[2] = {nullptr, nullptr}, *pin[2] = {nullptr, nullptr};
    hipStream_t s[2] = {nullptr, nullptr};
    struct Free {
        double **w, **p;
        hipStream_t *s;
        ~Free() {
            for (int i = 0; i < 2; ++i) {
                if (s[i]) (void)hipStreamSynchronize(s[i]);
                if (w[i]) (void)hipFree(w[i]);
                pinned_put(p[i]);
                if (s[i]) (void)hipStreamDestroy(s[i]);
            }
        }
    } fr{win, pin, s};
    for (int i = 0; i < 2; ++i) {
        HIPTRY(hipStreamCreateWithFlags(&s[i], hipStreamNonBlocking));
        if (hipMalloc(&win[i], cap * sizeof(double)) != hipSuccess) {
            set_err("device allocation of %zu bytes failed", cap * sizeof(double));
            return OOCDAG_ENOMEM;
        }
        if (!direct && !(pin[i] = pinned_get(cap * sizeof(double)))) {
            set_err("pinned host allocation of %zu bytes failed", cap * sizeof(double));
            return OOCDAG_ENOMEM;
        }
    }
    // group g: tiles [first, first + count) of the packed layout
    std::vector<long long> first, count;
    for (int j0 = 0; j0 < bb;) {
        int j1 = j0;
        size_t tiles = 0;
        while (j1 < bb && (tiles + (size_t)(bb - j1)) * nb2 <= cap && j1 - j0 < 16) {
            tiles += (size_t)(bb - j1);
            ++j1;
        }
        const long long t0 = tile_index(j0, j0, bb, 1);
        const int k = (int)first.size();
        first.push_back(t0);
        count.push_back((long long)tiles);
        hipStream_t sk = s[k & 1];
        HIPTRY(launch_gen_tiles(win[k & 1], N, NB, bb, j0, j1, aform, diag, iseed, sk, t0));
        double *dst = direct ? T + (size_t)t0 * nb2 : pin[k & 1];
        HIPTRY(hipMemcpyAsync(dst, win[k & 1], tiles * nb2 * sizeof(double),
                              hipMemcpyDeviceToHost, sk));
        if (!direct && k >= 1) {
            // group k-1 has landed in its pinned buffer: move it into T while
            // group k is generated; group k+1 reuses that buffer only after
            // this copy (it is queued by this thread afterwards)
            HIPTRY(hipStreamSynchronize(s[(k - 1) & 1]));
            par_memcpy(T + (size_t)first[k - 1] * nb2, pin[(k - 1) & 1],
                       (size_t)count[k - 1] * nb2);
        }
        j0 = j1;
    }
    const int G = (int)first.size();
    if (!direct && G >= 1) {
        HIPTRY(hipStreamSynchronize(s[(G - 1) & 1]));
        par_memcpy(T + (size_t)first[G - 1] * nb2, pin[(G - 1) & 1], (size_t)count[G - 1] * nb2);
    }
    HIPTRY(hipStreamSynchronize(s[0]));
    HIPTRY(hipStreamSynchronize(s[1]));
    return OOCDAG_OK;
}

void *oocdag_host_alloc(size_t bytes) {
    g_err.clear();
    void *p = nullptr;
    if (bytes == 0) return nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();   // a refused request must not fail the next launch check
        set_err("pinned host allocation of %zu bytes failed", bytes);
        return nullptr;
    }
    return p;
}

int oocdag_host_free(void *p) {
    if (p) HIPTRY(hipHostFree(p));
    return OOCDAG_OK;
}

size_t oocdag_tiles_elems(int N, int NB) {
    if (N <= 0 || NB <= 0) return 0;
    size_t bb = (size_t)((N + NB - 1) / NB);
    return bb * (bb + 1) / 2 * (size_t)tile_edge(NB) * tile_edge(NB);
}

int oocdag_tile_edge(int NB) { return NB > 0 ? tile_edge(NB) : OOCDAG_EARG; }

int oocdag_potrf_tiles(double *dY, int N, int NB, void *stream, int *info) {
    g_err.clear();
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (!dY || check_nb(N, NB)) return OOCDAG_EARG;
    if (N == 0) return OOCDAG_OK;
    if (!device_ok()) return OOCDAG_ENODEV;
    const int bb = (N + NB - 1) / NB;
    const int NBs = tile_edge(NB);
    Streams st;
    if (stream) {
        st.comp = (hipStream_t)stream;
        st.own_comp = false;
    } else {
        HIPTRY(hipStreamCreateWithFlags(&st.comp, hipStreamNonBlocking));
    }
    Scratch sc;
    int rc = sc.init(NBs, st.comp);
    if (rc) return rc;
    const double t0 = now();
    HIPTRY(hipStreamCreateWithFlags(&st.bulk, hipStreamNonBlocking));
    rc = factor_panel(dY, NBs, bb, bb, 0, sc.dinv, sc.dinfo, st.comp, st.bulk,
                      [](int) { return 0; });
    if (rc) return rc;
    rc = read_info(sc.dinfo, st.comp, N, info, NB);
    g_stat[0] = now() - t0;
    profile_collect();
    g_stat[1] = g_stat[2] = g_stat[3] = g_stat[4] = 0;
    return rc;
}

int oocdag_pdmatgen_tiles(double *dY, int N, int NB, char aform, char diag, int iseed,
                          void *stream) {
    g_err.clear();
    if (!dY || N <= 0 || NB <= 0) return OOCDAG_EARG;
    char af = (char)(aform & ~0x20), dg = (char)(diag & ~0x20);
    if (!(af == 'S' || af == 'H' || af == 'F') || !(dg == 'D' || dg == 'N')) {
        set_err("tile generator supports AFORM S/H/F and DIAG D/N");
        return OOCDAG_EARG;
    }
    if (!device_ok()) return OOCDAG_ENODEV;
    hipStream_t s = (hipStream_t)stream;
    const int bb = (N + NB - 1) / NB;
    // one launch per group of tile columns keeps the grid within 2^31 blocks
    for (int j0 = 0; j0 < bb; j0 += 16) {
        int j1 = std::min(bb, j0 + 16);
        HIPTRY(launch_gen_tiles(dY, N, NB, bb, j0, j1, aform, diag, iseed, s));
    }
    HIPTRY(hipStreamSynchronize(s));
    return OOCDAG_OK;
}

int oocdag_tiles_from_host(double *dY, const double *A, int N, int NB, int LDA) {
    g_err.clear();
    if (!dY || !A || N <= 0 || NB <= 0 || LDA < N) return OOCDAG_EARG;
    const int bb = (N + NB - 1) / NB;
    const size_t nb2 = (size_t)tile_edge(NB) * tile_edge(NB);
    std::vector<double> buf((size_t)bb * nb2);
    for (int tj = 0; tj < bb; ++tj) {
        pack_column(A, N, NB, LDA, tj, tj, buf.data());
        HIPTRY(hipMemcpy(dY + (size_t)tile_index(tj, tj, bb, 1) * nb2, buf.data(),
                         (size_t)(bb - tj) * nb2 * sizeof(double), hipMemcpyHostToDevice));
    }
    return OOCDAG_OK;
}

int oocdag_tiles_to_host(const double *dY, double *A, int N, int NB, int LDA) {
    g_err.clear();
    if (!dY || !A || N <= 0 || NB <= 0 || LDA < N) return OOCDAG_EARG;
    const int bb = (N + NB - 1) / NB;
    const size_t nb2 = (size_t)tile_edge(NB) * tile_edge(NB);
    std::vector<double> buf((size_t)bb * nb2);
    for (int tj = 0; tj < bb; ++tj) {
        HIPTRY(hipMemcpy(buf.data(), dY + (size_t)tile_index(tj, tj, bb, 1) * nb2,
                         (size_t)(bb - tj) * nb2 * sizeof(double), hipMemcpyDeviceToHost));
        unpack_column(A, N, NB, LDA, tj, tj, buf.data());
    }
    return OOCDAG_OK;
}

int oocdag_pdmatgen(char aform, char diag, int M, int N, int MB, int NB, double *A, int LDA,
                    int iseed) {
    g_err.clear();
    char af = (char)(aform & ~0x20), dg = (char)(diag & ~0x20);
    int symm = (af == 'S' || af == 'H');
    int tran = (af == 'T' || af == 'C');
    // argument checks of LIN/pdmatgen.f:156-195 (1x1 grid)
    int inf = 0;
    if (dg != 'D' && dg != 'N')
        inf = 3;
    else if (symm) {
        if (M != N)
            inf = 5;
        else if (MB != NB)
            inf = 7;
    } else if (M < 0)
        inf = 4;
    else if (N < 0)
        inf = 5;
    else if (MB < 1)
        inf = 6;
    else if (NB < 1)
        inf = 7;
    else if (LDA < 0)
        inf = 9;
    if (inf) return inf;
    if (M == 0 || N == 0) return 0;
    if (!A || LDA < M) return OOCDAG_EARG;
    if (!device_ok()) return OOCDAG_ENODEV;
    GenColArgs g;
    for (int k = 0; k < 40; ++k) lcg_pow_host(1ull << k, g.pow2[k][0], g.pow2[k][1]);
    g.x1 = (1103515245u * ((unsigned)iseed & 0x7fffffffu) + 12345u) & 0x7fffffffu;
    g.M = M;
    g.N = N;
    g.mode = (af == 'F') ? 3 : symm ? 2 : tran ? 1 : 0;
    // column chunks of ~1 GiB generated on the device into two alternating
    // windows and copied down while the next chunk is generated; pinned A
    // takes the copies directly, pageable A goes through two pinned staging
    // buffers that the host threads empty column by column
    const int chunk = (int)std::min<long long>(
        N, std::max<long long>(1, (1ll << 30) / 8 / std::max(1, M)));
    const size_t cap = (size_t)M * chunk;
    const bool direct = host_pinned(A);
    double *win[2] = {nullptr, nullptr}, *pin[2] = {nullptr, nullptr};
    hipStream_t s[2] = {nullptr, nullptr};
    struct Free {
        double **w, **p;
        hipStream_t *s;
        ~Free() {
            for (int i = 0; i < 2; ++i) {
                if (s[i]) (void)hipStreamSynchronize(s[i]);
                if (w[i]) (void)hipFree(w[i]);
                pinned_put(p[i]);
                if (s[i]) (void)hipStreamDestroy(s[i]);
            }
        }
    } fr{win, pin, s};
    for (int i = 0; i < 2; ++i) {
        HIPTRY(hipStreamCreateWithFlags(&s[i], hipStreamNonBlocking));
        HIPTRY(hipMalloc(&win[i], cap * sizeof(double)));
        if (!direct && !(pin[i] = pinned_get(cap * sizeof(double)))) {
            set_err("pinned host allocation of %zu bytes failed", cap * sizeof(double));
            return OOCDAG_ENOMEM;
        }
    }
    // columns [c0, c0 + nc) of pinned buffer b into A
    auto drain = [&](int b, int c0, int nc) {
        if (LDA == M) {
            par_memcpy(A + (size_t)c0 * LDA, pin[b], (size_t)M * nc);
            return;
        }
#pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; ++c)
            memcpy(A + (size_t)(c0 + c) * LDA, pin[b] + (size_t)c * M, (size_t)M * sizeof(double));
    };
    int k = 0, prev_c0 = 0, prev_nc = 0;
    for (int c0 = 0; c0 < N; c0 += chunk, ++k) {
        const int b = k & 1;
        g.col0 = c0;
        g.ncols = std::min(chunk, N - c0);
        g.D = win[b];
        long long total = (long long)M * g.ncols;
        int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(k_gen_cols, dim3(blocks), dim3(256), 0, s[b], g);
        HIPTRY(hipGetLastError());
        if (direct)
            HIPTRY(hipMemcpy2DAsync(A + (size_t)c0 * LDA, (size_t)LDA * sizeof(double), win[b],
                                    (size_t)M * sizeof(double), (size_t)M * sizeof(double),
                                    g.ncols, hipMemcpyDeviceToHost, s[b]));
        else
            HIPTRY(hipMemcpyAsync(pin[b], win[b], (size_t)M * g.ncols * sizeof(double),
                                  hipMemcpyDeviceToHost, s[b]));
        if (!direct && k >= 1) {
            // chunk k-1 has landed in its pinned buffer: move it into A while
            // chunk k is generated; chunk k+1 reuses that buffer only after
            // this copy (it is queued by this thread afterwards)
            HIPTRY(hipStreamSynchronize(s[b ^ 1]));
            drain(b ^ 1, prev_c0, prev_nc);
        }
        prev_c0 = c0;
        prev_nc = g.ncols;
    }
    if (k >= 1) {
        HIPTRY(hipStreamSynchronize(s[(k - 1) & 1]));
        if (!direct) drain((k - 1) & 1, prev_c0, prev_nc);
    }
    HIPTRY(hipStreamSynchronize(s[k & 1]));
    // LIN/pdmatgen.f:516-543; AFORM 'F' overwrites the diagonal afterwards
    // (:546-626), so DIAG has no effect on it
    if (dg == 'D' && MB == NB && af != 'F') {
        double maxmn = (double)std::max(M, N);
        for (int i = 0; i < std::min(M, N); ++i) {
            double *a = A + i + (size_t)i * LDA;
            *a = std::fabs(*a) + maxmn;
        }
    }
    return 0;
}

int oocdag_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                    int ldb, double beta, double *C, int ldc, void *stream) {
    g_err.clear();
    if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, n) ||
        ldc < std::max(1, m))
        return OOCDAG_EARG;
    Gemm g;
    g.A = plain((double *)A, lda);
    g.B = plain((double *)B, ldb);
    g.C = plain(C, ldc);
    g.M = m;
    g.N = n;
    g.K = k;
    g.alpha = alpha;
    g.beta = beta;
    g.lower = 0;
    hipError_t e = launch_gemm(g, GEMM_GENERAL, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) {
        set_err("dgemm_nt: leading dimensions must be even and pointers 16-byte aligned");
        return OOCDAG_EARG;
    }
    HIPTRY(e);
    return OOCDAG_OK;
}

int oocdag_dsyrk_ln(int n, int k, double alpha, const double *A, int lda, double beta, double *C,
                    int ldc, void *stream) {
    g_err.clear();
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldc < std::max(1, n)) return OOCDAG_EARG;
    Gemm g;
    g.A = plain((double *)A, lda);
    g.B = g.A;
    g.C = plain(C, ldc);
    g.M = n;
    g.N = n;
    g.K = k;
    g.alpha = alpha;
    g.beta = beta;
    g.lower = 1;
    hipError_t e = launch_gemm(g, GEMM_UPDATE, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) {
        set_err("dsyrk_ln: leading dimensions must be even and pointers 16-byte aligned");
        return OOCDAG_EARG;
    }
    HIPTRY(e);
    return OOCDAG_OK;
}

// Device scratch freed on scope exit.
struct DevTmp {
    double *p = nullptr;
    ~DevTmp() {
        if (p) (void)hipFree(p);
    }
};

// ---- the allocation-free variant of oocdag_dpotrf_l ----------------------
// Its scratch -- the 64 x 64 diagonal-block inverses of the left-looking
// sweep and the info word -- lives in one file-scope device array of a fixed
// size instead of hipMalloc'ed memory: kNoAllocMaxN / 64 inverse blocks,
// then the info word.  Each device holds its own copy; its address is read
// once per device with hipGetSymbolAddress and only that address is handed
// to kernels and copies (the array's name is a host-side stand-in).  Every
// call on one device shares the array: calls that overlap on a device race.
static constexpr int kNoAllocMaxN = 16384;
static constexpr size_t kNoAllocInvDoubles = (size_t)(kNoAllocMaxN / 64) * 64 * 64;
static __device__ __align__(256) double scratch[kNoAllocInvDoubles + 32];

static std::mutex g_scratch_mu;
static std::unordered_map<int, double *> g_scratch_ptr;   // device -> address of its copy

static double *scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    double *p = nullptr;
    if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_scratch_ptr[dev] = p;
    return p;
}

static __global__ void k_set_int(int *p, int v) { *p = v; }

static int dpotrf_aligned(int n, double *A, int lda, int *info, hipStream_t s) {
    Scratch sc;
    // the left-looking sweep keeps one 64x64 inverse per block column
    HIPTRY(hipMalloc(&sc.dinv, (size_t)(n / 64) * 64 * 64 * sizeof(double)));
    HIPTRY(hipMalloc(&sc.dinfo, sizeof(int)));
    int big = INT_MAX;
    HIPTRY(hipMemcpyAsync(sc.dinfo, &big, sizeof(int), hipMemcpyHostToDevice, s));
    HIPTRY(factor_column(plain(A, lda), 0, n, n, 0, sc.dinv, sc.dinfo, s));
    return read_info(sc.dinfo, s, n, info);
}

int oocdag_dpotrf_l_noalloc(int n, double *A, int lda, int *info, void *stream) {
    g_err.clear();
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (!A || n < 0 || lda < std::max(1, n)) {
        set_err("dpotrf_l_noalloc: invalid n or lda");
        return OOCDAG_EARG;
    }
    if (n > kNoAllocMaxN || n % 64 || (lda & 1)) {
        set_err("dpotrf_l_noalloc: needs n %% 64 == 0, n <= %d and an even lda (n=%d lda=%d); "
                "use oocdag_dpotrf_l for other shapes", kNoAllocMaxN, n, lda);
        return OOCDAG_EARG;
    }
    if (n == 0) return OOCDAG_OK;
    double *ws = scratch_on_current_device();
    if (!ws) {
        set_err("dpotrf_l_noalloc: the device scratch array is not available on this device");
        return OOCDAG_EHIP;
    }
    hipStream_t s = (hipStream_t)stream;
    int *dinfo = reinterpret_cast<int *>(ws + kNoAllocInvDoubles);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, dinfo, INT_MAX);
    HIPTRY(hipGetLastError());
    HIPTRY(factor_column(plain(A, lda), 0, n, n, 0, ws, dinfo, s));
    return read_info(dinfo, s, n, info);
}

int oocdag_dpotrf_l(int n, double *A, int lda, int *info, void *stream) {
    g_err.clear();
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (!A || n < 0 || lda < std::max(1, n)) {
        set_err("dpotrf_l: invalid n or lda");
        return OOCDAG_EARG;
    }
    if (n == 0) return OOCDAG_OK;
    hipStream_t s = (hipStream_t)stream;
    if (n % 64 == 0 && (lda & 1) == 0) return dpotrf_aligned(n, A, lda, info, s);
    // any other shape: factored in a padded copy (n rounded up to 64, the
    // padding an identity block, which leaves the n x n factor unchanged)
    const int np = tile_edge(n);
    DevTmp w;
    HIPTRY(hipMalloc(&w.p, (size_t)np * np * sizeof(double)));
    HIPTRY(hipMemcpy2DAsync(w.p, (size_t)np * 8, A, (size_t)lda * 8, (size_t)n * 8, n,
                            hipMemcpyDeviceToDevice, s));
    HIPTRY(launch_pad_tile(w.p, np, n, n, 1, s));
    int rc = dpotrf_aligned(np, w.p, np, info, s);
    if (rc) return rc;
    if (*info > n) *info = 0;   // the padding's pivots are exactly 1
    HIPTRY(hipMemcpy2DAsync(A, (size_t)lda * 8, w.p, (size_t)np * 8, (size_t)n * 8, n,
                            hipMemcpyDeviceToDevice, s));
    HIPTRY(hipStreamSynchronize(s));
    return OOCDAG_OK;
}

static int dtrsm_aligned(int m, int n, const double *L, int ldl, double *B, int ldb,
                         hipStream_t s) {
    // Invert L's 64x64 diagonal blocks in LDS (L itself is not modified),
    // then sweep block columns left to right:
    //   B[:, c] -= B[:, :c] L[c, :c]^T ;  B[:, c] <- B[:, c] D_c^T
    const int nblk = n / 64;
    DevTmp dinv;
    HIPTRY(hipMalloc(&dinv.p, (size_t)nblk * 64 * 64 * sizeof(double)));
    for (int c = 0; c < nblk; ++c)
        HIPTRY(launch_potrf64(plain((double *)L, ldl, c * 64, c * 64), 0,
                              dinv.p + (size_t)c * 4096, nullptr, s));
    for (int c = 0; c < nblk; ++c) {
        if (c > 0) {
            Gemm g;
            g.A = plain(B, ldb, 0, 0);
            g.B = plain((double *)L, ldl, c * 64, 0);
            g.C = plain(B, ldb, 0, c * 64);
            g.M = m;
            g.N = 64;
            g.K = c * 64;
            g.alpha = -1.0;
            g.beta = 1.0;
            g.lower = 0;
            HIPTRY(launch_gemm(g, GEMM_NARROW, s));
        }
        Gemm g;
        g.A = plain(B, ldb, 0, c * 64);
        g.C = g.A;
        g.B = plain(dinv.p + (size_t)c * 4096, 64);
        g.M = m;
        g.N = 64;
        g.K = 64;
        g.alpha = 1.0;
        g.beta = 0.0;
        g.lower = 0;
        HIPTRY(launch_gemm(g, GEMM_APPLY, s));
    }
    HIPTRY(hipStreamSynchronize(s));
    return OOCDAG_OK;
}

int oocdag_dtrsm_rltn(int m, int n, const double *L, int ldl, double *B, int ldb, void *stream) {
    g_err.clear();
    if (!L || !B || m < 0 || n < 0 || ldl < std::max(1, n) || ldb < std::max(1, m)) {
        set_err("dtrsm_rltn: invalid sizes or leading dimensions");
        return OOCDAG_EARG;
    }
    if (m == 0 || n == 0) return OOCDAG_OK;
    hipStream_t s = (hipStream_t)stream;
    if (n % 64 == 0 && (ldl & 1) == 0 && (ldb & 1) == 0)
        return dtrsm_aligned(m, n, L, ldl, B, ldb, s);
    // any other shape: L padded to np x np with an identity block and B to
    // m x np with zero columns ([B 0] [L 0; 0 I]^-T = [B L^-T 0])
    const int np = tile_edge(n);
    const int ldm = m + (m & 1);
    DevTmp lp, bp;
    HIPTRY(hipMalloc(&lp.p, (size_t)np * np * sizeof(double)));
    HIPTRY(hipMalloc(&bp.p, (size_t)ldm * np * sizeof(double)));
    HIPTRY(hipMemcpy2DAsync(lp.p, (size_t)np * 8, L, (size_t)ldl * 8, (size_t)n * 8, n,
                            hipMemcpyDeviceToDevice, s));
    HIPTRY(launch_pad_tile(lp.p, np, n, n, 1, s));
    HIPTRY(hipMemsetAsync(bp.p, 0, (size_t)ldm * np * sizeof(double), s));
    HIPTRY(hipMemcpy2DAsync(bp.p, (size_t)ldm * 8, B, (size_t)ldb * 8, (size_t)m * 8, n,
                            hipMemcpyDeviceToDevice, s));
    int rc = dtrsm_aligned(m, np, lp.p, np, bp.p, ldm, s);
    if (rc) return rc;
    HIPTRY(hipMemcpy2DAsync(B, (size_t)ldb * 8, bp.p, (size_t)ldm * 8, (size_t)m * 8, n,
                            hipMemcpyDeviceToDevice, s));
    HIPTRY(hipStreamSynchronize(s));
    return OOCDAG_OK;
}

int oocdag_potrs_tiles(const double *dY, int N, int NB, double *dB, int nrhs, int ldb,
                       void *stream) {
    g_err.clear();
    if (!dY || !dB || N <= 0 || nrhs < 0 || ldb < N || check_nb(N, NB)) return OOCDAG_EARG;
    if (nrhs == 0) return OOCDAG_OK;
    hipStream_t s = (hipStream_t)stream;
    // the solve runs on the padded matrix of the slots (padding rows of the
    // right-hand sides zero, so their solution entries stay zero)
    const int bb = (N + NB - 1) / NB;
    const int NBs = tile_edge(NB);
    const long long Np = (long long)bb * NBs;
    const long long ldx = nrhs + (nrhs & 1);
    double *dinv = nullptr, *Xt = nullptr;
    struct F {
        double **a, **b;
        ~F() {
            if (*a) (void)hipFree(*a);
            if (*b) (void)hipFree(*b);
        }
    } f{&dinv, &Xt};
    HIPTRY(hipMalloc(&dinv, (size_t)Np * 64 * sizeof(double)));
    HIPTRY(hipMalloc(&Xt, (size_t)(ldx * Np) * sizeof(double)));
    HIPTRY(hipMemsetAsync(Xt, 0, (size_t)(ldx * Np) * sizeof(double), s));
    const Opnd Lp = panel((double *)dY, NBs, bb, 1);
    for (int k = 0; k < bb; ++k) {
        Opnd Lk = Lp;
        Lk.r0 = k * NBs;
        Lk.c0 = k * NBs;
        HIPTRY(diag_inverses(Lk, NBs, dinv + (size_t)k * NBs * 64, s));
    }
    HIPTRY(launch_transpose(dB, ldb, Xt, ldx, N, nrhs, s, NB, NBs, 1));
    HIPTRY(potrs_tiles_t(dY, (int)Np, NBs, dinv, Xt, nrhs, ldx, s));
    HIPTRY(launch_transpose(Xt, ldx, dB, ldb, nrhs, N, s, NB, NBs, 2));
    HIPTRY(hipStreamSynchronize(s));
    return OOCDAG_OK;
}

int oocdag_potrs(const double *L, int N, int LDL, double *B, int nrhs, int LDB) {
    g_err.clear();
    if (!L || !B || N <= 0 || nrhs < 0 || LDL < N || LDB < N) return OOCDAG_EARG;
    if (!device_ok()) return OOCDAG_ENODEV;
    const int NB = N >= 4096 ? 512 : 128;
    double *dY = nullptr, *dB = nullptr;
    struct F {
        double **a, **b;
        ~F() {
            if (*a) (void)hipFree(*a);
            if (*b) (void)hipFree(*b);
        }
    } f{&dY, &dB};
    if (hipMalloc(&dY, oocdag_tiles_elems(N, NB) * sizeof(double)) != hipSuccess) {
        set_err("factor does not fit in device memory");
        return OOCDAG_ENOMEM;
    }
    int rc = oocdag_tiles_from_host(dY, L, N, NB, LDL);
    if (rc) return rc;
    const int ldb = N + (N & 1);
    HIPTRY(hipMalloc(&dB, (size_t)ldb * nrhs * sizeof(double)));
    HIPTRY(hipMemcpy2D(dB, (size_t)ldb * sizeof(double), B, (size_t)LDB * sizeof(double),
                       (size_t)N * sizeof(double), nrhs, hipMemcpyHostToDevice));
    rc = oocdag_potrs_tiles(dY, N, NB, dB, nrhs, ldb, nullptr);
    if (rc) return rc;
    HIPTRY(hipMemcpy2D(B, (size_t)LDB * sizeof(double), dB, (size_t)ldb * sizeof(double),
                       (size_t)N * sizeof(double), nrhs, hipMemcpyDeviceToHost));
    return OOCDAG_OK;
}

static int resid_finish(const double *rowsum, const double *anorm, int N, double *resid) {
    std::vector<double> rs(N), an(N);
    HIPTRY(hipMemcpy(rs.data(), rowsum, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(an.data(), anorm, N * sizeof(double), hipMemcpyDeviceToHost));
    double worst = 0, amax = 0;
    for (int i = 0; i < N; ++i) {
        worst = std::max(worst, rs[i]);
        amax = std::max(amax, an[i]);
    }
    *resid = worst / ((double)N * std::ldexp(1.0, -53) * amax);
    return OOCDAG_OK;
}

struct ResidScratch {
    double *R = nullptr, *S = nullptr, *rowsum = nullptr, *anorm = nullptr;
    ~ResidScratch() {
        for (double *p : {R, S, rowsum, anorm})
            if (p) (void)hipFree(p);
    }
    int init(int N, int NB) {
        const int bb = (N + NB - 1) / NB;
        const size_t col = (size_t)bb * tile_edge(NB) * tile_edge(NB) * sizeof(double);
        if (hipMalloc(&R, col) != hipSuccess || hipMalloc(&S, col) != hipSuccess) {
            set_err("residual scratch allocation failed");
            return OOCDAG_ENOMEM;
        }
        HIPTRY(hipMalloc(&rowsum, (size_t)N * sizeof(double)));
        HIPTRY(hipMalloc(&anorm, (size_t)N * sizeof(double)));
        HIPTRY(hipMemset(rowsum, 0, (size_t)N * sizeof(double)));
        HIPTRY(hipMemset(anorm, 0, (size_t)N * sizeof(double)));
        return OOCDAG_OK;
    }
};

int oocdag_llt_resid_tiles(const double *dY, int N, int NB, char aform, char diag, int iseed,
                           double *resid) {
    g_err.clear();
    if (!dY || !resid || N <= 0 || check_nb(N, NB)) return OOCDAG_EARG;
    if (!device_ok()) return OOCDAG_ENODEV;
    ResidScratch sc;
    int rc = sc.init(N, NB);
    if (rc) return rc;
    HIPTRY(llt_resid_tiles(dY, N, NB, aform, diag, iseed, sc.R, sc.S, sc.rowsum, sc.anorm, 0));
    HIPTRY(hipDeviceSynchronize());
    return resid_finish(sc.rowsum, sc.anorm, N, resid);
}

int oocdag_llt_resid(const double *A, int N, int LDA, char aform, char diag, int iseed,
                     double *resid) {
    g_err.clear();
    if (!A || !resid || N <= 0 || LDA < N) return OOCDAG_EARG;
    if (!device_ok()) return OOCDAG_ENODEV;
    const int NB = N >= 4096 ? 512 : 128;
    const int bb = (N + NB - 1) / NB;
    double *dY = nullptr, *U = nullptr;
    struct F {
        double **a, **b;
        ~F() {
            if (*a) (void)hipFree(*a);
            if (*b) (void)hipFree(*b);
        }
    } f{&dY, &U};
    if (hipMalloc(&dY, oocdag_tiles_elems(N, NB) * sizeof(double)) != hipSuccess) {
        set_err("factor does not fit in device memory");
        return OOCDAG_ENOMEM;
    }
    int rc = oocdag_tiles_from_host(dY, A, N, NB, LDA);
    if (rc) return rc;
    ResidScratch sc;
    if ((rc = sc.init(N, NB))) return rc;
    HIPTRY(llt_resid_tiles(dY, N, NB, aform, diag, iseed, sc.R, sc.S, sc.rowsum, sc.anorm, 0));
    // strict upper tiles: PDPOTRRV keeps A's upper triangle, so every change
    // the factorization made there counts (LIN/pdlafchk.f:204,219)
    HIPTRY(hipMalloc(&U, (size_t)N * NB * sizeof(double)));
    for (int jt = 1; jt < bb; ++jt) {
        int cols = std::min(NB, N - jt * NB);
        HIPTRY(hipMemcpy2D(U, (size_t)N * sizeof(double), A + (size_t)jt * NB * LDA,
                           (size_t)LDA * sizeof(double), (size_t)jt * NB * sizeof(double), cols,
                           hipMemcpyHostToDevice));
        HIPTRY(resid_upper_block(U, N, N, NB, jt, aform, diag, iseed, sc.rowsum, 0));
    }
    HIPTRY(hipDeviceSynchronize());
    return resid_finish(sc.rowsum, sc.anorm, N, resid);
}

void *oocdag_dev_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_err("hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

int oocdag_dev_free(void *p) {
    HIPTRY(hipFree(p));
    return OOCDAG_OK;
}

int oocdag_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIPTRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return OOCDAG_OK;
}

int oocdag_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIPTRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return OOCDAG_OK;
}

int oocdag_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    HIPTRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return OOCDAG_OK;
}

int oocdag_memset_d(void *dst, int value, size_t bytes) {
    HIPTRY(hipMemset(dst, value, bytes));
    return OOCDAG_OK;
}

int oocdag_profile_enable(int on) {
    profile_collect();
    g_prof.on = on != 0;
    g_prof.launches = 0;
    g_prof.seconds = 0;
    g_prof.total_flops = 0;
    g_prof.total_bytes = 0;
    return OOCDAG_OK;
}

int oocdag_profile_read(long long *launches, double *seconds, double *flops) {
    profile_collect();
    if (launches) *launches = g_prof.launches;
    if (seconds) *seconds = g_prof.seconds;
    if (flops) *flops = g_prof.total_flops;
    return OOCDAG_OK;
}

int oocdag_profile_bytes(double *bytes) {
    profile_collect();
    if (bytes) *bytes = g_prof.total_bytes;
    return OOCDAG_OK;
}

int oocdag_release_staging(void) { return pinned_trim(); }

int oocdag_set_progress(int every_columns) {
    g_progress_every = every_columns > 0 ? every_columns : 0;
    return OOCDAG_OK;
}

int oocdag_device_sync(void) {
    HIPTRY(hipDeviceSynchronize());
    return OOCDAG_OK;
}

}  // extern "C"
