// Heterogeneous batches, serving and the kernel-variant autotuner
// (C ABI in include/fdmgpu.h).  No reference counterpart: the reference solves
// one problem per LaplCube::solve call; this layer decides how many solves
// share a launch and which kernel variant a shape uses.
//
//  * Problems of one solver (same constructor arguments) form one batch: every
//    stage is one launch over the batch, with per-problem pointers in a
//    device table (problems stay in their own buffers).
//  * fdmgpu_group: a fixed set of problems; batches are spread over internal
//    streams (fork/join through events) and the whole sequence can be
//    captured once into a hipGraph and replayed.
//  * fdmgpu_server: a worker thread admits requests in arrival order while
//    their scratch fits an arena of the HBM budget, launches a batch per
//    solver as requests arrive (continuous batching across streams) and
//    completes them when the batch's event fires.  Solver plans (twiddle and
//    eigenvalue tables) stay resident.  An FP32+refine batch launches as many
//    inner solves as that solver's requests have needed so far; a request
//    still above the tolerance continues from its iterate in a follow-up
//    batch (up to the plan's max_iter), so no launches are spent on
//    converged problems.
//  * fdmgpu_tune: times every kernel variant of a shape and keeps the fastest
//    (process-wide table, optionally a JSON file read through
//    FDMGPU_TUNE_CACHE).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "plan.hpp"

using namespace fdmgpu;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

#define SRV_HIPCHK(x)                                   \
    do {                                                \
        hipError_t _e = (x);                            \
        if (_e != hipSuccess) {                         \
            set_error_text(hipGetErrorString(_e));      \
            return _e == hipErrorOutOfMemory ? FDMGPU_ERR_OUT_OF_MEMORY : FDMGPU_ERR_DEVICE; \
        }                                               \
    } while (0)

// ---- solver keys and plans -------------------------------------------------
struct Key {
    int kind, precision, periodic, nx, ny, nz;
    double dx, dy, dz, lx, ly, lz;
    bool operator<(const Key& o) const {
        return std::tie(kind, precision, periodic, nx, ny, nz, dx, dy, dz, lx, ly, lz) <
               std::tie(o.kind, o.precision, o.periodic, o.nx, o.ny, o.nz, o.dx, o.dy, o.dz, o.lx, o.ly, o.lz);
    }
};

Key key_of(const fdmgpu_problem& p) {
    if (p.kind == FDMGPU_KIND_RECT) return Key{2, p.precision, p.periodic, p.nx, p.ny, 1, p.dx, p.dy, 0, p.lx, p.ly, 0};
    return Key{3, p.precision, p.periodic, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, p.lx, p.ly, p.lz};
}

int make_plan(const Key& k, fdmgpu_plan* out) {
    if (k.kind == 2) return fdmgpu_plan_rect(out, k.precision, k.periodic, k.dx, k.dy, k.lx, k.ly, k.nx, k.ny);
    if (k.kind == 3)
        return fdmgpu_plan_cube(out, k.precision, k.periodic, k.dx, k.dy, k.dz, k.lx, k.ly, k.lz, k.nx, k.ny, k.nz);
    return FDMGPU_ERR_INVALID_ARG;
}

// allow_alias: the autotuner, which solves on copies, accepts d_ans == d_rhs
// for every precision.
int check_problem(const fdmgpu_problem& p, bool allow_alias = false) {
    if (p.kind != FDMGPU_KIND_CUBE && p.kind != FDMGPU_KIND_RECT) return FDMGPU_ERR_INVALID_ARG;
    if (!p.d_ans || !p.d_rhs) return FDMGPU_ERR_INVALID_ARG;
    if (!allow_alias && p.precision == FDMGPU_F32_REFINE64 && p.d_ans == p.d_rhs) return FDMGPU_ERR_INVALID_ARG;
    return FDMGPU_OK;
}

// Scratch bytes of a batch: the refinement Work plus the pointer tables.
size_t batch_bytes(const Key& k, int count) {
    return work_bytes_shape(k.kind, k.precision, k.nx, k.ny, k.nz, count) +
           ((2 * sizeof(void*) * count + 255) & ~(size_t)255);
}

long problem_elems(const Key& k) { return (long)k.nx * k.ny * k.nz; }

// ---- autotuner table -------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::string, Variant> g_tune;
bool g_tune_env_read = false;

std::string tune_key(int kind, int precision, int periodic, int nx, int ny, int nz) {
    char b[96];
    std::snprintf(b, sizeof b, "%s:p%d:m%d:%dx%dx%d", kind == 2 ? "rect" : "cube", precision, periodic, nx, ny,
                  kind == 2 ? 1 : nz);
    return b;
}

// Cache file format (written by fdmgpu_tune): one entry per line,
//   "<key>": {"plane": P, "tile_x": X, "tile_y": Y, "zcr_gmul": G, "fuse": F, "ms": T},
void tune_read_file(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char key[128];
        Variant v;
        double ms = 0;
        if (std::sscanf(line,
                        " \"%127[^\"]\": {\"plane\": %d, \"tile_x\": %d, \"tile_y\": %d, \"zcr_gmul\": %d, "
                        "\"fuse\": %d, \"ms\": %lf",
                        key, &v.plane, &v.tile_x, &v.tile_y, &v.zcr_gmul, &v.fuse, &ms) == 7)
            g_tune[key] = v;
    }
    std::fclose(f);
}

void tune_env_locked() {
    if (g_tune_env_read) return;
    g_tune_env_read = true;
    if (const char* p = std::getenv("FDMGPU_TUNE_CACHE")) tune_read_file(p);
}

}  // namespace

void fdmgpu::apply_tuned(fdmgpu_plan_t* P) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_env_locked();
    auto it = g_tune.find(tune_key(P->kind, P->precision, P->periodic, P->nx, P->ny, P->nz));
    if (it != g_tune.end() && variant_valid(P, it->second)) P->var = it->second;
}

// ---- group -----------------------------------------------------------------
namespace {
struct Bucket {
    Key key;
    fdmgpu_plan plan = nullptr;
    std::vector<int> idx;        // problem indices
    Work W;                      // owned (work_reserve)
    void** d_tab = nullptr;      // [ans pointers | rhs pointers]
    int stream = 0;
};
}  // namespace

struct fdmgpu_group_t {
    int device = 0;
    int flags = 0;
    std::vector<fdmgpu_problem> probs;
    std::vector<Bucket> buckets;
    std::vector<hipStream_t> streams;   // [0] = origin stream
    hipEvent_t ev_in = nullptr, ev_out = nullptr, fork = nullptr;
    std::vector<hipEvent_t> joins;
    int runs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void group_free(fdmgpu_group_t* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    for (hipStream_t s : g->streams) (void)hipStreamSynchronize(s);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (Bucket& b : g->buckets) {
        work_free(b.W);
        (void)hipFree(b.d_tab);
        if (b.plan) fdmgpu_destroy(b.plan);
    }
    for (hipEvent_t e : g->joins) (void)hipEventDestroy(e);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->ev_in) (void)hipEventDestroy(g->ev_in);
    if (g->ev_out) (void)hipEventDestroy(g->ev_out);
    for (hipStream_t s : g->streams) (void)hipStreamDestroy(s);
    delete g;
}

// Enqueues every batch on the group's streams, forked from and joined back
// into the origin stream (capturable: events and kernels only).
int group_enqueue(fdmgpu_group_t* g) {
    hipStream_t s0 = g->streams[0];
    SRV_HIPCHK(hipEventRecord(g->fork, s0));
    for (size_t i = 1; i < g->streams.size(); i++) SRV_HIPCHK(hipStreamWaitEvent(g->streams[i], g->fork, 0));
    for (Bucket& b : g->buckets) {
        const int n = (int)b.idx.size();
        int rc = plan_solve(b.plan, b.W, nullptr, (Tab)b.d_tab, nullptr, (Tab)(b.d_tab + n), n, g->streams[b.stream]);
        if (rc) return rc;
    }
    for (size_t i = 1; i < g->streams.size(); i++) {
        SRV_HIPCHK(hipEventRecord(g->joins[i], g->streams[i]));
        SRV_HIPCHK(hipStreamWaitEvent(s0, g->joins[i], 0));
    }
    return FDMGPU_OK;
}

}  // namespace

extern "C" {

int fdmgpu_group_create(fdmgpu_group* group, int count, const fdmgpu_problem* problems, int nstreams, int flags) {
    if (!group || count < 1 || !problems || nstreams < 0 || nstreams > 16) return FDMGPU_ERR_INVALID_ARG;
    *group = nullptr;
    for (int i = 0; i < count; i++) {
        int rc = check_problem(problems[i]);
        if (rc) return rc;
    }
    fdmgpu_group_t* g = new fdmgpu_group_t();
    g->flags = flags;
    g->probs.assign(problems, problems + count);
    if (hipGetDevice(&g->device) != hipSuccess) {
        delete g;
        return FDMGPU_ERR_DEVICE;
    }
    std::map<Key, int> where;
    for (int i = 0; i < count; i++) {
        const Key k = key_of(problems[i]);
        auto it = where.find(k);
        if (it == where.end()) {
            where[k] = (int)g->buckets.size();
            g->buckets.push_back(Bucket{});
            g->buckets.back().key = k;
            g->buckets.back().idx.push_back(i);
        } else {
            g->buckets[it->second].idx.push_back(i);
        }
    }
    int rc = FDMGPU_OK;
    const int ns = 1 + std::min<int>(nstreams ? nstreams : 4, (int)g->buckets.size());
    for (int i = 0; i < ns && !rc; i++) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) rc = FDMGPU_ERR_DEVICE;
        else g->streams.push_back(s);
    }
    if (!rc && (hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&g->ev_in, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&g->ev_out, hipEventDisableTiming) != hipSuccess))
        rc = FDMGPU_ERR_DEVICE;
    g->joins.assign(g->streams.size(), nullptr);
    for (size_t i = 1; i < g->streams.size() && !rc; i++)
        if (hipEventCreateWithFlags(&g->joins[i], hipEventDisableTiming) != hipSuccess) rc = FDMGPU_ERR_DEVICE;
    // largest batches first, each to the least-loaded worker stream
    std::vector<int> order(g->buckets.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    auto cost = [&](const Bucket& b) {
        return (double)problem_elems(b.key) * b.idx.size() * (b.key.precision == FDMGPU_F32_REFINE64 ? 2.0 : 1.0);
    };
    std::sort(order.begin(), order.end(), [&](int a, int b) { return cost(g->buckets[a]) > cost(g->buckets[b]); });
    std::vector<double> load(g->streams.size(), 0.0);
    for (int bi : order) {
        if (rc) break;
        Bucket& b = g->buckets[bi];
        size_t best = 1;
        for (size_t i = 2; i < load.size(); i++)
            if (load[i] < load[best]) best = i;
        b.stream = (int)best;
        load[best] += cost(b);
        rc = make_plan(b.key, &b.plan);
        if (rc) break;
        const int n = (int)b.idx.size();
        if (b.key.precision == FDMGPU_F32_REFINE64 && (rc = work_reserve(b.plan, b.W, n))) break;
        std::vector<const void*> tab(2 * n);
        for (int j = 0; j < n; j++) {
            tab[j] = problems[b.idx[j]].d_ans;
            tab[n + j] = problems[b.idx[j]].d_rhs;
        }
        if (hipMalloc(&b.d_tab, sizeof(void*) * 2 * n) != hipSuccess ||
            hipMemcpy(b.d_tab, tab.data(), sizeof(void*) * 2 * n, hipMemcpyHostToDevice) != hipSuccess)
            rc = FDMGPU_ERR_OUT_OF_MEMORY;
    }
    if (rc) {
        group_free(g);
        return rc;
    }
    *group = g;
    return FDMGPU_OK;
}

int fdmgpu_group_solve(fdmgpu_group g, void* stream) {
    if (!g) return FDMGPU_ERR_INVALID_ARG;
    SRV_HIPCHK(hipSetDevice(g->device));
    hipStream_t user = (hipStream_t)stream;
    hipStream_t s0 = g->streams[0];
    SRV_HIPCHK(hipEventRecord(g->ev_in, user));
    SRV_HIPCHK(hipStreamWaitEvent(s0, g->ev_in, 0));
    int rc = FDMGPU_OK;
    if (!(g->flags & FDMGPU_GROUP_GRAPH) || g->runs == 0) {
        rc = group_enqueue(g);
    } else {
        if (!g->exec) {
            SRV_HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
            rc = group_enqueue(g);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(s0, &graph);
            if (!rc && e != hipSuccess) rc = FDMGPU_ERR_DEVICE;
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            g->graph = graph;
            SRV_HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        }
        SRV_HIPCHK(hipGraphLaunch(g->exec, s0));
    }
    if (rc) return rc;
    SRV_HIPCHK(hipEventRecord(g->ev_out, s0));
    SRV_HIPCHK(hipStreamWaitEvent(user, g->ev_out, 0));
    g->runs++;
    return FDMGPU_OK;
}

int fdmgpu_group_set_refine(fdmgpu_group g, double rel_tol, int max_iter) {
    if (!g || !(rel_tol > 0) || max_iter < 1 || max_iter > 64) return FDMGPU_ERR_INVALID_ARG;
    SRV_HIPCHK(hipSetDevice(g->device));
    for (hipStream_t s : g->streams) SRV_HIPCHK(hipStreamSynchronize(s));
    for (Bucket& b : g->buckets)
        if (b.key.precision == FDMGPU_F32_REFINE64) {
            const int rc = fdmgpu_set_refine(b.plan, rel_tol, max_iter);
            if (rc) return rc;
        }
    // the captured sequence holds the old launches: capture again
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
    return FDMGPU_OK;
}

int fdmgpu_group_result(fdmgpu_group g, int i, fdmgpu_result* r) {
    if (!g || !r || i < 0 || i >= (int)g->probs.size()) return FDMGPU_ERR_INVALID_ARG;
    SRV_HIPCHK(hipSetDevice(g->device));
    SRV_HIPCHK(hipStreamSynchronize(g->streams[0]));
    *r = fdmgpu_result{};
    r->state = g->runs ? 2 : 0;
    r->status = FDMGPU_OK;
    r->iterations = 1;
    r->rel_residual = -1.0;
    for (Bucket& b : g->buckets) {
        auto it = std::find(b.idx.begin(), b.idx.end(), i);
        if (it == b.idx.end()) continue;
        if (b.key.precision == FDMGPU_F32_REFINE64 && g->runs) {
            const int j = (int)(it - b.idx.begin());
            SRV_HIPCHK(hipMemcpy(&r->iterations, b.W.iters + j, sizeof(int), hipMemcpyDeviceToHost));
            SRV_HIPCHK(hipMemcpy(&r->rel_residual, b.W.resid + j, sizeof(double), hipMemcpyDeviceToHost));
            if (!(r->rel_residual <= b.plan->rtol)) r->status = FDMGPU_ERR_NOT_CONVERGED;
        }
        break;
    }
    return FDMGPU_OK;
}

int fdmgpu_group_destroy(fdmgpu_group g) {
    if (!g) return FDMGPU_ERR_INVALID_ARG;
    group_free(g);
    return FDMGPU_OK;
}

}  // extern "C"

// ---- server ----------------------------------------------------------------
namespace {

// First-fit allocator over one device allocation (the HBM budget).
class Arena {
  public:
    void init(char* base, size_t size) {
        base_ = base;
        size_ = size;
        free_.clear();
        free_[0] = size;
    }
    char* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            return base_ + off;
        }
        return nullptr;
    }
    void release(char* p, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        size_t off = (size_t)(p - base_);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

  private:
    char* base_ = nullptr;
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Req {
    unsigned long long ticket = 0;
    fdmgpu_problem p{};
    Key key{};
    int its_done = 0;          // inner solves already applied (continuation when > 0)
    Clock::time_point t_submit, t_launch, t_done;
    fdmgpu_result res{};
};

struct Batch {
    std::vector<Req*> reqs;
    fdmgpu_plan plan = nullptr;
    Key key{};
    int first = 1, last = 1;   // inner solves of this launch (refine)
    Work W;
    char* mem = nullptr;
    size_t bytes = 0;
    int stream = 0;
    hipEvent_t ev = nullptr;
    int cap = 0;
    void** h_tab = nullptr;    // pinned [ans | rhs]
    int* h_iters = nullptr;    // pinned
    double* h_resid = nullptr; // pinned
};

}  // namespace

struct fdmgpu_server_t {
    int device = 0;
    int max_batch = 64;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_load;
    int max_inflight = 2;         // batches in flight per stream
    unsigned long long waiting_head = 0;   // ticket of the queue's head when it last had to wait
    int bypassed = 0;             // batches launched before that head since
    char* d_arena = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<unsigned long long, Req*> reqs;
    std::map<Key, fdmgpu_plan> plans;
    std::map<Key, int> typical;   // inner solves the solver's requests needed so far
    double rtol = 1e-10;          // refinement settings of the server's plans
    int max_iter = 4;
    std::vector<Batch*> running, spare;
    std::vector<std::thread> workers;
    bool stop = false;
    int worker_error = FDMGPU_OK;
    unsigned long long next_ticket = 1;
    size_t outstanding = 0;   // submitted and not completed
};

namespace {

Batch* get_batch(fdmgpu_server_t* S) {
    Batch* b = nullptr;
    if (!S->spare.empty()) {
        b = S->spare.back();
        S->spare.pop_back();
        return b;
    }
    b = new Batch();
    b->cap = S->max_batch;
    if (hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&b->h_tab, sizeof(void*) * 2 * b->cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_iters, sizeof(int) * b->cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_resid, sizeof(double) * b->cap, hipHostMallocDefault) != hipSuccess) {
        return nullptr;
    }
    return b;
}

void free_batch(Batch* b) {
    if (!b) return;
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->h_iters) (void)hipHostFree(b->h_iters);
    if (b->h_resid) (void)hipHostFree(b->h_resid);
    delete b;
}

// Launches one batch (the worker thread; S->mu not held).
int launch_batch(fdmgpu_server_t* S, Batch* b) {
    const int n = (int)b->reqs.size();
    const size_t wb = work_bytes(b->plan, n);
    b->W = work_carve(b->plan, n, b->mem);
    void** d_tab = (void**)(b->mem + wb);
    for (int j = 0; j < n; j++) {
        b->h_tab[j] = b->reqs[j]->p.d_ans;
        b->h_tab[n + j] = const_cast<void*>(b->reqs[j]->p.d_rhs);
    }
    hipStream_t s = S->streams[b->stream];
    SRV_HIPCHK(hipMemcpyAsync(d_tab, b->h_tab, sizeof(void*) * 2 * n, hipMemcpyHostToDevice, s));
    int rc = plan_solve_iters(b->plan, b->W, nullptr, (Tab)d_tab, nullptr, (Tab)(d_tab + n), n, s, b->first,
                              b->last);
    if (rc) return rc;
    if (b->plan->precision == FDMGPU_F32_REFINE64) {
        SRV_HIPCHK(hipMemcpyAsync(b->h_iters, b->W.iters, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        SRV_HIPCHK(hipMemcpyAsync(b->h_resid, b->W.resid, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    }
    SRV_HIPCHK(hipEventRecord(b->ev, s));
    return FDMGPU_OK;
}

// Worker threads.  Admission, batching and completion run under the server
// lock; a batch's launch (a dozen or so kernel launches and copies) runs
// outside it, so several workers can issue batches to their streams at once.
// Two measured the same as one on the mixed serving stream (the host's
// issue rate is not what bounds it), so one it is.
constexpr int kWorkers = 1;

// Compute units the large-problem stream leaves to the others (0: none).
constexpr int kReservedCUs = 32;

// Problems of at least this many points are "large" (a 255^3 problem has
// 16.6M; 127^3 has 2.0M): they get stream 0 to themselves.
constexpr long kLargeElems = 1L << 22;
// Batches that may launch before a waiting head of the queue.
constexpr int kMaxBypass = 8;

// Least-loaded stream of the size class of a problem of `elems` points that
// still takes a batch (stream_load < max_inflight), or -1.
int pick_stream(const fdmgpu_server_t* S, long elems) {
    const int ns = (int)S->streams.size();
    int lo = 0, hi = ns;                   // one stream: every class shares it
    if (ns > 1) {
        if (elems >= kLargeElems) hi = 1;  // stream 0
        else lo = 1;                       // streams 1 .. ns-1
    }
    int best = -1;
    for (int i = lo; i < hi; i++)
        if (S->stream_load[i] < S->max_inflight && (best < 0 || S->stream_load[i] < S->stream_load[best])) best = i;
    return best;
}

void worker_main(fdmgpu_server_t* S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (S->stop && S->queue.empty() && S->running.empty()) break;
        bool progressed = false;

        // 1. admission + batching.  Requests are taken in arrival order; a
        //    batch takes the queued requests of one solver (up to max_batch)
        //    whose scratch fits the arena.  Back-pressure: a stream takes at
        //    most max_inflight batches, so under load arrivals wait on the host,
        //    where the next batch of their solver takes all of them (fewer,
        //    larger launches instead of a stream queue of tiny batches).  Size
        //    classes: problems of at least kLargeElems points run on stream 0,
        //    the others on the remaining streams (with one stream, all share
        //    it), so small requests never queue behind a large solve on the
        //    GPU.  Bypass: when the head must wait (no stream of its class
        //    free, or its scratch does not fit), later requests of other
        //    solvers may launch before it; while it waits for scratch at most
        //    kMaxBypass batches do, then admission waits for the head, which
        //    takes the next scratch that frees up (it cannot starve).
        for (;;) {
            if (S->queue.empty()) break;
            Req* head = S->queue.front();
            if (head->ticket != S->waiting_head) {
                S->waiting_head = head->ticket;
                S->bypassed = 0;
            }
            bool launched = false;
            bool head_needs_scratch = false;   // the head waits for scratch (not for a stream)
            std::vector<const Req*> tried;     // first request of each solver already tried this pass
            for (size_t qi = 0; qi < S->queue.size() && !launched; qi++) {
                Req* cand = S->queue[qi];
                // requests behind a head that waits for scratch compete with it
                // for the arena: at most kMaxBypass batches pass it.  A head
                // that waits for a stream of its size class is not delayed by
                // requests of the other class.
                if (qi > 0 && head_needs_scratch && S->bypassed >= kMaxBypass) break;
                bool seen = false;
                for (const Req* t : tried)
                    if (!(t->key < cand->key) && !(cand->key < t->key) && t->its_done == cand->its_done) seen = true;
                if (seen) continue;
                tried.push_back(cand);
                const Key k = cand->key;
                const int its = cand->its_done;
                const int best = pick_stream(S, problem_elems(k));
                if (best < 0) continue;   // every stream of its class is busy
                std::vector<Req*> take;
                for (Req* r : S->queue) {
                    if ((int)take.size() >= S->max_batch) break;
                    if (!(r->key < k) && !(k < r->key) && r->its_done == its) take.push_back(r);
                }
                char* mem = nullptr;
                size_t bytes = 0;
                while (!take.empty()) {
                    bytes = batch_bytes(k, (int)take.size());
                    mem = S->arena.alloc(bytes);
                    if (mem) break;
                    take.resize(take.size() / 2);
                }
                if (!mem) {   // its scratch does not fit yet
                    if (qi == 0) head_needs_scratch = true;
                    continue;
                }
                if (qi > 0 && head_needs_scratch) S->bypassed++;
                Batch* b = get_batch(S);
                if (!b) {
                    S->arena.release(mem, bytes);
                    S->worker_error = FDMGPU_ERR_OUT_OF_MEMORY;
                    break;
                }
                b->reqs = take;
                b->key = k;
                b->plan = S->plans[k];
                {
                    const int maxit = b->plan->max_iter;
                    const int typ = S->typical.count(k) ? S->typical[k] : maxit;
                    b->first = its + 1;
                    b->last = std::min(maxit, std::max(b->first, typ));
                }
                b->mem = mem;
                b->bytes = bytes;
                b->stream = best;
                S->stream_load[best]++;
                for (Req* r : take) {
                    S->queue.erase(std::find(S->queue.begin(), S->queue.end(), r));
                    if (r->its_done == 0) r->t_launch = Clock::now();
                    r->res.state = 1;
                }
                lk.unlock();
                const int rc = launch_batch(S, b);
                // a launch that failed part way may have queued copies and kernels
                // that still use the batch's scratch and pinned tables: they must
                // finish before either is handed to another batch
                if (rc) (void)hipStreamSynchronize(S->streams[b->stream]);
                lk.lock();
                if (rc) {
                    const Clock::time_point now = Clock::now();
                    for (Req* r : b->reqs) {
                        r->res.state = 2;
                        r->res.status = rc;
                        r->t_done = now;
                        S->outstanding--;
                    }
                    S->arena.release(b->mem, b->bytes);
                    S->stream_load[b->stream]--;
                    S->spare.push_back(b);
                    S->cv_done.notify_all();
                } else {
                    S->running.push_back(b);
                }
                progressed = true;
                launched = true;
            }
            if (!launched) break;   // nothing can launch until completions free streams or scratch
        }
        // 2. completions
        for (size_t i = 0; i < S->running.size();) {
            Batch* b = S->running[i];
            const hipError_t q = hipEventQuery(b->ev);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            const Clock::time_point now = Clock::now();
            const bool refine = b->plan->precision == FDMGPU_F32_REFINE64;
            std::vector<Req*> cont;
            for (size_t j = 0; j < b->reqs.size(); j++) {
                Req* r = b->reqs[j];
                const bool ok = q == hipSuccess;
                const bool converged = !refine || b->h_resid[j] <= b->plan->rtol;
                if (ok && refine) {
                    int& typ = S->typical[b->key];
                    typ = std::max(typ, converged ? b->h_iters[j] : b->last + 1);
                }
                if (ok && !converged && b->last < b->plan->max_iter) {   // continue from the iterate
                    r->its_done = b->last;
                    cont.push_back(r);
                    continue;
                }
                r->t_done = now;
                r->res.state = 2;
                r->res.status = ok ? FDMGPU_OK : FDMGPU_ERR_DEVICE;
                r->res.iterations = refine ? b->h_iters[j] : 1;
                r->res.rel_residual = refine ? b->h_resid[j] : -1.0;
                if (refine && r->res.status == FDMGPU_OK && !converged) r->res.status = FDMGPU_ERR_NOT_CONVERGED;
                if (r->its_done == 0) r->res.queue_ms = ms_between(r->t_submit, r->t_launch);
                r->res.latency_ms = ms_between(r->t_submit, r->t_done);
                S->outstanding--;
            }
            for (auto it = cont.rbegin(); it != cont.rend(); ++it) S->queue.push_front(*it);   // ahead of new work
            S->arena.release(b->mem, b->bytes);
            S->stream_load[b->stream]--;
            b->reqs.clear();
            S->spare.push_back(b);
            S->running.erase(S->running.begin() + i);
            progressed = true;
        }
        if (progressed) S->cv_done.notify_all();
        if (!progressed) {
            // idle: wake on a submission, or poll the running batches again soon
            S->cv_work.wait_for(lk, std::chrono::microseconds(S->running.empty() ? 2000 : 20));
        }
    }
}

}  // namespace

extern "C" {

int fdmgpu_server_create(fdmgpu_server* server, unsigned long long hbm_budget, int max_batch, int nstreams) {
    if (!server || hbm_budget < 4096 || max_batch < 1 || max_batch > 4096 || nstreams < 0 || nstreams > 16)
        return FDMGPU_ERR_INVALID_ARG;
    *server = nullptr;
    fdmgpu_server_t* S = new fdmgpu_server_t();
    S->max_batch = max_batch;
    int rc = FDMGPU_OK;
    if (hipGetDevice(&S->device) != hipSuccess) rc = FDMGPU_ERR_DEVICE;
    for (int i = 0; i < (nstreams ? nstreams : 4) && !rc; i++) {
        hipStream_t s;
        hipError_t e;
        int cus = 0;
        if (i == 0 && nstreams != 1 && kReservedCUs > 0 &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, S->device) == hipSuccess &&
            cus > 2 * kReservedCUs) {
            // the large-problem stream leaves kReservedCUs compute units to the
            // other streams, so a small request does not wait for a large
            // solve's blocks to drain.  CU c is reserved when
            // (c mod 32) mod q == (c / 32) mod q, q = cus / kReservedCUs: the
            // same number in every 32 consecutive CUs and, on 256 CUs, in
            // every XCD whether CUs are numbered per XCD or round-robin
            const int q = std::max(1, cus / kReservedCUs);
            std::vector<uint32_t> mask((cus + 31) / 32, 0u);
            for (int c = 0; c < cus; c++)
                if ((c % 32) % q != (c / 32) % q) mask[c / 32] |= 1u << (c % 32);
            e = hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data());
        } else {
            e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        }
        if (e != hipSuccess) rc = FDMGPU_ERR_DEVICE;
        else S->streams.push_back(s);
    }
    S->stream_load.assign(S->streams.size(), 0);
    if (!rc && hipMalloc(&S->d_arena, hbm_budget) != hipSuccess) rc = FDMGPU_ERR_OUT_OF_MEMORY;
    if (rc) {
        for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
        delete S;
        return rc;
    }
    S->arena.init(S->d_arena, hbm_budget);
    for (int i = 0; i < kWorkers; i++) S->workers.emplace_back(worker_main, S);
    *server = S;
    return FDMGPU_OK;
}

int fdmgpu_server_submit(fdmgpu_server S, const fdmgpu_problem* p, unsigned long long* ticket) {
    if (!S || !p || !ticket) return FDMGPU_ERR_INVALID_ARG;
    int rc = check_problem(*p);
    if (rc) return rc;
    const Key k = key_of(*p);
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop) return FDMGPU_ERR_INVALID_ARG;
    if (batch_bytes(k, 1) > S->arena.size()) return FDMGPU_ERR_ADMISSION;
    if (!S->plans.count(k)) {
        // plan creation validates the problem (sizes, precision) and uploads
        // its tables once; later requests of this solver reuse them
        fdmgpu_plan plan = nullptr;
        (void)hipSetDevice(S->device);
        rc = make_plan(k, &plan);
        if (rc) return rc;
        if (k.precision == FDMGPU_F32_REFINE64) (void)fdmgpu_set_refine(plan, S->rtol, S->max_iter);
        S->plans[k] = plan;
    }
    Req* r = new Req();
    r->ticket = S->next_ticket++;
    r->p = *p;
    r->key = k;
    r->t_submit = Clock::now();
    r->res.state = 0;
    r->res.status = FDMGPU_OK;
    S->reqs[r->ticket] = r;
    S->queue.push_back(r);
    S->outstanding++;
    *ticket = r->ticket;
    lk.unlock();
    S->cv_work.notify_one();
    return FDMGPU_OK;
}

int fdmgpu_server_set_refine(fdmgpu_server S, double rel_tol, int max_iter) {
    if (!S || !(rel_tol > 0) || max_iter < 1 || max_iter > 64) return FDMGPU_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    S->rtol = rel_tol;
    S->max_iter = max_iter;
    for (auto& kv : S->plans)
        if (kv.first.precision == FDMGPU_F32_REFINE64) (void)fdmgpu_set_refine(kv.second, rel_tol, max_iter);
    return FDMGPU_OK;
}

int fdmgpu_server_poll(fdmgpu_server S, unsigned long long ticket, fdmgpu_result* res) {
    if (!S || !res) return FDMGPU_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return FDMGPU_ERR_INVALID_ARG;
    *res = it->second->res;
    return FDMGPU_OK;
}

int fdmgpu_server_wait(fdmgpu_server S, unsigned long long ticket, fdmgpu_result* res) {
    if (!S) return FDMGPU_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return FDMGPU_ERR_INVALID_ARG;
    Req* r = it->second;
    S->cv_done.wait(lk, [&] { return r->res.state == 2; });
    if (res) *res = r->res;
    return FDMGPU_OK;
}

int fdmgpu_server_drain(fdmgpu_server S) {
    if (!S) return FDMGPU_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->outstanding == 0; });
    return S->worker_error;
}

int fdmgpu_server_forget(fdmgpu_server S, unsigned long long ticket) {
    if (!S) return FDMGPU_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end() || it->second->res.state != 2) return FDMGPU_ERR_INVALID_ARG;
    delete it->second;
    S->reqs.erase(it);
    return FDMGPU_OK;
}

int fdmgpu_server_destroy(fdmgpu_server S) {
    if (!S) return FDMGPU_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (std::thread& w : S->workers)
        if (w.joinable()) w.join();
    (void)hipSetDevice(S->device);
    for (hipStream_t s : S->streams) (void)hipStreamSynchronize(s);
    for (Batch* b : S->spare) free_batch(b);
    for (auto& kv : S->plans) fdmgpu_destroy(kv.second);
    for (auto& kv : S->reqs) delete kv.second;
    (void)hipFree(S->d_arena);
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    delete S;
    return FDMGPU_OK;
}

// ---- autotuner ---------------------------------------------------------------
int fdmgpu_tune_set(const fdmgpu_problem* problem, const fdmgpu_variant* v) {
    if (!problem || !v) return FDMGPU_ERR_INVALID_ARG;
    int rc = check_problem(*problem);
    if (rc) return rc;
    fdmgpu_plan plan = nullptr;
    if ((rc = make_plan(key_of(*problem), &plan))) return rc;
    const Variant w{v->plane, v->tile_x, v->tile_y, v->zcr_gmul, v->fuse};
    const bool ok = variant_valid(plan, w);
    fdmgpu_destroy(plan);
    if (!ok) return FDMGPU_ERR_INVALID_ARG;
    const fdmgpu_problem& p = *problem;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_env_locked();
    g_tune[tune_key(p.kind, p.precision, p.periodic, p.nx, p.ny, p.kind == FDMGPU_KIND_RECT ? 1 : p.nz)] = w;
    return FDMGPU_OK;
}

int fdmgpu_tune(int count, const fdmgpu_problem* problems, int reps, const char* cache_path) {
    if (count < 1 || !problems || reps < 1 || reps > 1000) return FDMGPU_ERR_INVALID_ARG;
    for (int i = 0; i < count; i++) {
        int rc = check_problem(problems[i], true);
        if (rc) return rc;
    }
    // problems grouped by shape key; each shape tuned on its own batch
    std::map<std::string, std::vector<int>> shapes;
    for (int i = 0; i < count; i++) {
        const fdmgpu_problem& p = problems[i];
        shapes[tune_key(p.kind, p.precision, p.periodic, p.nx, p.ny, p.kind == FDMGPU_KIND_RECT ? 1 : p.nz)].push_back(i);
    }
    hipStream_t s;
    SRV_HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    SRV_HIPCHK(hipEventCreate(&e0));
    SRV_HIPCHK(hipEventCreate(&e1));
    int rc = FDMGPU_OK;
    std::map<std::string, std::pair<Variant, double>> found;
    for (auto& kv : shapes) {
        const std::vector<int>& ids = kv.second;
        const int n = (int)ids.size();
        fdmgpu_plan plan = nullptr;
        if ((rc = make_plan(key_of(problems[ids[0]]), &plan))) break;
        // every variant runs on copies of the shape's right-hand sides and
        // writes into scratch answers: the caller's arrays are only read
        // (a problem may even alias ans and rhs)
        const size_t pb = (size_t)plan_elems(plan) * plan->elem;
        const size_t pba = (pb + 255) & ~(size_t)255;
        void** d_tab = nullptr;
        char* d_copy = nullptr;
        std::vector<const void*> tab(2 * n);
        bool ok = hipMalloc(&d_copy, 2 * pba * n) == hipSuccess && hipMalloc(&d_tab, sizeof(void*) * 2 * n) == hipSuccess;
        for (int j = 0; j < n && ok; j++) {
            tab[j] = d_copy + pba * j;
            tab[n + j] = d_copy + pba * (n + j);
            ok = hipMemcpyAsync(const_cast<void*>(tab[n + j]), problems[ids[j]].d_rhs, pb, hipMemcpyDeviceToDevice,
                                s) == hipSuccess;
        }
        if (ok) ok = hipMemcpyAsync(d_tab, tab.data(), sizeof(void*) * 2 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
                     hipStreamSynchronize(s) == hipSuccess;
        if (!ok) {
            rc = FDMGPU_ERR_OUT_OF_MEMORY;
            fdmgpu_destroy(plan);
            (void)hipFree(d_tab);
            (void)hipFree(d_copy);
            break;
        }
        double best_ms = 1e300;
        Variant best;
        for (const Variant& v : plan_variants(plan)) {
            plan->var = v;
            // warm-up (also grows the plan's scratch), then `reps` timed solves
            rc = plan_solve(plan, plan->w, nullptr, (Tab)d_tab, nullptr, (Tab)(d_tab + n), n, s);
            for (int r = 0; r < reps && !rc; r++) {
                if (r == 0 && hipEventRecord(e0, s) != hipSuccess) rc = FDMGPU_ERR_DEVICE;
                if (!rc) rc = plan_solve(plan, plan->w, nullptr, (Tab)d_tab, nullptr, (Tab)(d_tab + n), n, s);
            }
            float ms = 0;
            if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = FDMGPU_ERR_DEVICE;
            if (rc) break;
            if (ms / reps < best_ms) {
                best_ms = ms / reps;
                best = v;
            }
        }
        (void)hipStreamSynchronize(s);
        (void)hipFree(d_tab);
        (void)hipFree(d_copy);
        fdmgpu_destroy(plan);
        if (rc) break;
        found[kv.first] = {best, best_ms};
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_env_locked();
    for (auto& kv : found) g_tune[kv.first] = kv.second.first;
    if (cache_path) {
        // keep entries of other shapes already in the file
        std::map<std::string, Variant> keep;
        std::swap(keep, g_tune);
        tune_read_file(cache_path);
        std::map<std::string, Variant> onfile;
        std::swap(onfile, g_tune);
        g_tune = keep;
        FILE* f = std::fopen(cache_path, "w");
        if (!f) return FDMGPU_ERR_INVALID_ARG;
        std::fprintf(f, "{\n");
        std::map<std::string, std::pair<Variant, double>> all;
        for (auto& kv : onfile) all[kv.first] = {kv.second, -1.0};
        for (auto& kv : found) all[kv.first] = kv.second;
        size_t i = 0;
        for (auto& kv : all) {
            const Variant& v = kv.second.first;
            std::fprintf(f,
                         "  \"%s\": {\"plane\": %d, \"tile_x\": %d, \"tile_y\": %d, \"zcr_gmul\": %d, "
                         "\"fuse\": %d, \"ms\": %.6f}%s\n",
                         kv.first.c_str(), v.plane, v.tile_x, v.tile_y, v.zcr_gmul, v.fuse, kv.second.second,
                         ++i < all.size() ? "," : "");
        }
        std::fprintf(f, "}\n");
        std::fclose(f);
    }
    return FDMGPU_OK;
}

}  // extern "C"
