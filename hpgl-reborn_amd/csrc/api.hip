// api.hip -- C ABI of libhpgl.so: the reference's entry points (include/hpgl.h)
// over host buffers and the device-pointer extensions (include/hpglx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/cvariogram.h"
#include "../../include/hpglx.h"
#include "inc_io.h"
#include "introsort.h"
#include "krige_kernels.h"
#include "krige_launch.h"
#include "loo.h"
#include "plan.h"
#include "sgs.h"
#include "variogram_plan.h"

namespace hpglx {

int vario_grid(const VarioTemplate& t, const float* d_data, const unsigned char* d_mask, const int shape[3],
               int lag_count, int percent, float* d_result, Workspace& ws, hipStream_t st, std::string& err);
int vario_points(const VarioTemplate& t, const float* d_pts4, int n, int lag_count, float* d_result, Workspace& ws,
                 hipStream_t st, std::string& err);

// Model evaluation on the device for scattered coordinates (cov_model.h:63-77,
// 146-183); the transform comes from the host.
struct DevModel {
    int type;
    double sill, nugget, r0;
    double t[9];
};

__device__ inline double dev_cov(const DevModel& m, double v0, double v1, double v2) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double d = v0 * m.t[3 * i + 0];
        d += v1 * m.t[3 * i + 1];
        d += v2 * m.t[3 * i + 2];
        acc += d * d;
    }
    const double h = sqrt(acc);
    if (h < 0.0001) return m.sill;
    const double x = h / m.r0;
    if (m.type == 2) return (m.sill - m.nugget) * exp(-3 * (x * x));
    if (m.type == 1) return (m.sill - m.nugget) * exp(-3 * h / m.r0);
    if (h > m.r0) return 0.0;
    return (m.sill - m.nugget) * (1 - 1.5 * x + 0.5 * (x * x * x));
}

// SK weights at float coordinates (simple_kriging.cpp:18-38, the path of
// hpgl_simple_kriging_weights): one block, FP64, any n.  The matrix lives in
// a device buffer of n*n doubles; right-looking Cholesky, one column per step.
__global__ __launch_bounds__(256) void k_sk_weights(DevModel m, const float* center, const float* xs,
                                                    const float* ys, const float* zs, int n, double* A, double* x,
                                                    int* status) {
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) bad = 0;
    for (long long i = tid; i < (long long)n * n; i += nt) {
        const int r = (int)(i / n), c = (int)(i % n);
        A[i] = dev_cov(m, (double)xs[r] - (double)xs[c], (double)ys[r] - (double)ys[c], (double)zs[r] - (double)zs[c]);
    }
    for (int r = tid; r < n; r += nt)
        x[r] = dev_cov(m, (double)xs[r] - (double)center[0], (double)ys[r] - (double)center[1],
                       (double)zs[r] - (double)center[2]);
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            const double d = A[(long long)j * n + j];
            if (!(d > 0.0))
                bad = 1;
            else
                A[(long long)j * n + j] = sqrt(d);
        }
        __syncthreads();
        if (bad) break;
        const double d = A[(long long)j * n + j];
        for (int i = j + 1 + tid; i < n; i += nt) A[(long long)i * n + j] /= d;
        __syncthreads();
        const int t = n - j - 1;
        for (long long idx = tid; idx < (long long)t * t; idx += nt) {
            const int i = j + 1 + (int)(idx / t), k = j + 1 + (int)(idx % t);
            if (k <= i) A[(long long)i * n + k] -= A[(long long)i * n + j] * A[(long long)k * n + j];
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (bad) {
            *status = 1;
            return;
        }
        for (int i = 0; i < n; ++i) {
            double s = x[i];
            for (int k = 0; k < i; ++k) s -= A[(long long)i * n + k] * x[k];
            x[i] = s / A[(long long)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (int k = i + 1; k < n; ++k) s -= A[(long long)k * n + i] * x[k];
            x[i] = s / A[(long long)i * n + i];
        }
        *status = 0;
    }
}

// ------------------------------------------- scratch of hpglx_krige_noalloc
// Every device buffer of an allocation-free call beyond the caller's inputs
// and outputs (plan tables, padded mask, neighbour lists, cluster lists,
// counters) is carved from this array.  Each GPU the library runs on holds
// its own copy until the process ends.  In host code the name `scratch` is
// not a device address: the address is read once per device with
// hipGetSymbolAddress (NoAllocDevice::base) and used everywhere instead.
constexpr size_t kNoAllocScratchBytes = (size_t)HPGLX_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
// Lanes of the full-table cluster search in that mode (its per-lane candidate
// buffers are the largest scratch it would otherwise take).
constexpr int kNoAllocClusterLanes = 512;
constexpr size_t kNoAllocGuard = 64 << 10;

namespace {

// State of the allocation-free entry point on one device.
struct NoAllocDevice {
    char* base = nullptr;       // this device's copy of `scratch` (hipGetSymbolAddress)
    Workspace ws;               // arena over the scratch after the resident plan's tables
    std::string resident;       // key of the plan whose tables start the scratch ("" = none)
    size_t plan_bytes = 0;
};

struct Library {
    std::mutex mu;
    std::string last_error;
    std::map<std::string, std::unique_ptr<DevPlan>> plans;
    std::map<std::string, unsigned long long> plan_use;  // last use, for eviction
    unsigned long long plan_clock = 0;
    Workspace ws;
    int (*out_handler)(char*, void*) = nullptr;
    void* out_param = nullptr;
    int (*progress_handler)(char*, int, void*) = nullptr;
    void* progress_param = nullptr;
    int thread_num = 1;
    int default_precision = PREC_FP64;
    bool profiling = false;
    // Event sets of the profiled calls since profiling was switched on; the
    // newest kTimingPool calls are kept.
    std::vector<CallTiming> tpool;
    size_t tcalls = 0;
    hipEvent_t done = nullptr;   // completion of the last call's work
    const int* last_counters = nullptr;  // device counters of the last grid call (fallbacks, cluster cells)
    // hpglx_krige_noalloc: per-device scratch state and host-built plans
    // (placed into the scratch on demand; keys include the device).
    std::map<int, NoAllocDevice> noalloc;
    std::map<std::string, std::unique_ptr<DevPlan>> na_plans;
    std::map<std::string, unsigned long long> na_use;
    hipStream_t io_stream = nullptr;
    char* pin = nullptr;         // page-locked staging of the host-array entry points (HostIO)
    size_t pin_cap = 0;
    IncText text;                // pinned host copy of the last property file read or written
    unsigned long long plan_serial = 0;
    // Graphs of repeated grid calls: a call equal to the previous one (plan,
    // geometry, buffers, parameters) is captured into a hipGraph once and
    // then replayed while the workspace and the state the capture relied on
    // (the padded mask's cleared halo, the plan's template deltas) stay as
    // they were.  Off unless HPGLX_GRAPH=1 or hpglx_set_graph_mode(1): on
    // MI355X a replay measured no faster than the same launches from the
    // stream.
    struct CallGraph {
        hipGraphExec_t exec = nullptr;
        unsigned long long ws_gen = 0;
        std::string pm_tag;
        int tkey[4] = {0, 0, 0, 0};
        unsigned long long use = 0;
    };
    std::map<std::string, CallGraph> graphs;
    std::string last_key;
    unsigned long long graph_clock = 0;
    hipStream_t cap_stream = nullptr;
    int graph_mode = -1;  // -1: not read yet
    // Launch settings chosen by hpglx_autotune per (parameter set, grid,
    // precision); HPGLX_TUNE_FILE keeps them across processes.
    struct Tuning {
        int search_head = 0, blocks_per_cu = 0, dyn_sched = 0;
    };
    std::map<std::string, Tuning> tuned;
    bool tune_file_read = false;
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
};

Library& lib() {
    static Library* L = new Library();
    return *L;
}

void write_msg(Library& L, const std::string& s) {
    if (L.out_handler) {
        std::vector<char> buf(s.begin(), s.end());
        buf.push_back(0);
        L.out_handler(buf.data(), L.out_param);
    } else {
        std::fputs(("[LOG2]" + s).c_str(), stdout);
        std::fflush(stdout);
    }
}

void progress(Library& L, const char* stage, int pct) {
    if (L.progress_handler) {
        std::vector<char> buf(stage, stage + std::strlen(stage) + 1);
        L.progress_handler(buf.data(), pct, L.progress_param);
    }
}

std::string plan_key(int kind, const CovParams& c, const int r[3], int kmax) {
    char b[512];
    std::snprintf(b, sizeof b, "%d|%d|%.17g,%.17g,%.17g|%.17g,%.17g,%.17g|%.17g|%.17g|%d,%d,%d|%d", kind, c.type,
                  c.ranges[0], c.ranges[1], c.ranges[2], c.angles[0], c.angles[1], c.angles[2], c.sill, c.nugget, r[0],
                  r[1], r[2], kmax);
    int dev = 0;
    (void)hipGetDevice(&dev);
    return std::string(b) + "|dev" + std::to_string(dev);
}

CovParams to_cov(const hpgl_cov_params_t& p) {
    CovParams c;
    c.type = p.m_covariance_type;
    for (int i = 0; i < 3; ++i) {
        c.ranges[i] = p.m_ranges[i];
        c.angles[i] = p.m_angles[i];
    }
    c.sill = p.m_sill;
    c.nugget = p.m_nugget;
    return c;
}

// Plans (host tables + their device copies) are cached per parameter set;
// the least recently used one is dropped beyond kMaxPlans, after the last
// call's work has finished with it.
constexpr size_t kMaxPlans = 8;
constexpr size_t kMaxGraphs = 8;

DevPlan* get_plan(Library& L, int kind, const CovParams& c, const int r[3], int kmax, std::string& err) {
    const std::string key = plan_key(kind, c, r, kmax);
    ++L.plan_clock;
    auto it = L.plans.find(key);
    if (it != L.plans.end()) {
        L.plan_use[key] = L.plan_clock;
        return it->second.get();
    }
    std::unique_ptr<DevPlan> p(new DevPlan());
    if (!build_plan(p->h, (Kind)kind, c, r, kmax, err)) return nullptr;
    p->serial = ++L.plan_serial;
    if (L.plans.size() >= kMaxPlans) {
        auto lru = L.plan_use.begin();
        for (auto u = L.plan_use.begin(); u != L.plan_use.end(); ++u)
            if (u->second < lru->second) lru = u;
        if (L.done && hipEventSynchronize(L.done) != hipSuccess) {
            err = "hipEventSynchronize failed";
            return nullptr;
        }
        L.plans[lru->first]->release();
        L.plans.erase(lru->first);
        L.plan_use.erase(lru);
    }
    if (!p->upload(err)) return nullptr;
    DevPlan* raw = p.get();
    L.plans[key] = std::move(p);
    L.plan_use[key] = L.plan_clock;
    return raw;
}

bool ensure_events(Library& L, std::string& err) {
    if (!L.done) {
        if (hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            err = "hipEventCreate failed";
            return false;
        }
    }
    return true;
}

constexpr size_t kTimingPool = 4096;

// The event set for the next profiled call.
CallTiming* next_timing(Library& L, std::string& err) {
    const size_t slot = L.tcalls % kTimingPool;
    if (slot >= L.tpool.size()) {
        CallTiming t;
        for (int i = 0; i < CallTiming::N; ++i)
            if (hipEventCreate(&t.ev[i]) != hipSuccess) {
                err = "hipEventCreate failed";
                return nullptr;
            }
        t.valid = true;
        L.tpool.push_back(t);
    }
    ++L.tcalls;
    return &L.tpool[slot];
}

// Milliseconds {search kernel, main solve kernel} of one call.
bool read_timing(const CallTiming& t, float ms[2]) {
    if (hipEventSynchronize(t.ev[CallTiming::N - 1]) != hipSuccess) return false;
    return hipEventElapsedTime(&ms[0], t.ev[0], t.ev[1]) == hipSuccess &&
           hipEventElapsedTime(&ms[1], t.ev[2], t.ev[3]) == hipSuccess;
}

// The launcher call a request describes (plan and workspace aside).
bool request_call(const hpglx_grid_request_t* q, GridCall& call, const int* cats_host, std::string& err) {
    call.kind = q->kind;
    call.precision = q->precision;
    call.nx = q->dims[0];
    call.ny = q->dims[1];
    call.nz = q->dims[2];
    call.values = q->values;
    call.mask = q->mask;
    call.ind = q->ind;
    call.zfield = q->zfield;
    call.thresholds = q->thresholds;
    call.nind = q->nind;
    call.marg_host = q->marginal;
    call.auto_mean = q->auto_mean;
    call.mean = q->mean;
    call.out_val = q->out_val;
    call.out_mask = q->out_mask;
    call.out_class = q->out_class;
    call.out_prob = q->out_prob;
    call.out_var = q->out_var;
    call.out_est64 = q->out_est64;
    call.out_status = q->out_status;
    call.order_relations = q->order_relations;
    call.out_count = q->out_count;
    call.out_nboff = q->out_nboff;
    call.pivot_tol = q->pivot_tol > 0 ? q->pivot_tol : 1e-3;
    call.refine_steps = q->refine_steps >= 0 ? q->refine_steps : 1;
    call.lvm_mean = q->lvm_mean;
    call.cats_host = cats_host;
    if (const char* sv = std::getenv("HPGLX_SOLVE")) call.solve_variant = std::atoi(sv);
    if (const char* sv = std::getenv("HPGLX_SEARCH")) call.search_variant = std::atoi(sv);
    if (const char* sv = std::getenv("HPGLX_SEARCH_HEAD")) call.search_head = std::atoi(sv);
    if (const char* sv = std::getenv("HPGLX_SEARCH_SLABS")) call.search_slabs = std::atoi(sv);
    if (const char* sv = std::getenv("HPGLX_CLUSTER_CAP")) call.cluster_cap = std::atoi(sv);
    if (const char* sv = std::getenv("HPGLX_CLUSTER_CAND")) call.cluster_cand = std::atoi(sv);
    if (const char* sv = std::getenv("HPGLX_OVERLAP")) call.overlap = std::atoi(sv);
    if (!q->mask || (q->kind != KIND_MIK && !q->values)) {
        err = "values and mask are required";
        return false;
    }
    if (q->kind == KIND_MIK) {
        if (!q->marginal || (!q->ind && !(q->zfield && q->thresholds))) {
            err = "median IK needs indicator data (or a field and thresholds) and marginal probabilities";
            return false;
        }
        if (q->zfield) call.values = q->zfield;
    }
    return true;
}

// ------------------------------------------------------------ call graphs
template <class T>
void key_put(std::string& k, const T& v) {
    k.append((const char*)&v, sizeof v);
}

// Everything a grid call's launches take: two calls with equal keys enqueue
// the same kernels with the same arguments (given the same workspace).
std::string call_key(const GridCall& c, const DevPlan& plan) {
    std::string k;
    k.reserve(320);
    key_put(k, plan.serial);
    const int ints[] = {c.kind, c.precision, c.nx, c.ny, c.nz, c.nind, c.auto_mean, c.order_relations,
                        c.refine_steps, c.search_variant, c.solve_variant, c.search_head, c.cluster_cap,
                        c.cluster_lanes, c.blocks_per_cu, c.cats_host ? 1 : 0, c.search_slabs, c.cluster_cand,
                        c.dyn_sched, c.overlap};
    key_put(k, ints);
    const void* ptrs[] = {c.values, c.mask, c.ind, c.zfield, c.thresholds, c.lvm_mean, c.out_val, c.out_mask,
                          c.out_class, c.out_prob, c.out_var, c.out_est64, c.out_status, c.out_count, c.out_nboff};
    key_put(k, ptrs);
    key_put(k, c.mean);
    key_put(k, c.pivot_tol);
    for (int i = 0; i < c.nind && i < 16; ++i) {
        if (c.marg_host) key_put(k, c.marg_host[i]);
        if (c.cats_host) key_put(k, c.cats_host[i]);
    }
    return k;
}

bool graphs_enabled(Library& L) {
    if (L.graph_mode < 0) {
        const char* v = std::getenv("HPGLX_GRAPH");
        L.graph_mode = (v && v[0] == '1') ? 1 : 0;
    }
    return L.graph_mode == 1;
}

// Captures the call into a graph (on a library stream, so a caller's legacy
// default stream can still replay it).  False when the capture did not give
// a graph that can stand for the call; then the call runs as usual.
bool capture_call(Library& L, DevPlan& plan, const GridCall& call, const std::string& key) {
    if (!L.cap_stream && hipStreamCreateWithFlags(&L.cap_stream, hipStreamNonBlocking) != hipSuccess) {
        L.cap_stream = nullptr;
        return false;
    }
    const unsigned long long gen0 = L.ws.gen;
    const std::string tag0 = L.ws.tags["pmask"];
    int tk0[4];
    std::copy(plan.tkey, plan.tkey + 4, tk0);
    if (hipStreamBeginCapture(L.cap_stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    std::string cerr;
    const int rc = launch_grid(plan, L.ws, call, L.cap_stream, nullptr, cerr);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(L.cap_stream, &graph);
    const bool steady = L.ws.gen == gen0 && L.ws.tags["pmask"] == tag0 && std::equal(tk0, tk0 + 4, plan.tkey);
    if (rc != 0 || e != hipSuccess || !graph || !steady) {
        if (graph) (void)hipGraphDestroy(graph);
        // set-up work the capture recorded never ran: make the next call redo it
        if (L.ws.tags["pmask"] != tag0) L.ws.tags.erase("pmask");
        if (!std::equal(tk0, tk0 + 4, plan.tkey)) std::fill(plan.tkey, plan.tkey + 4, -1);
        (void)hipGetLastError();
        return false;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess || !exec) {
        (void)hipGetLastError();
        return false;
    }
    if (L.graphs.size() >= kMaxGraphs) {
        auto lru = L.graphs.begin();
        for (auto g = L.graphs.begin(); g != L.graphs.end(); ++g)
            if (g->second.use < lru->second.use) lru = g;
        (void)hipGraphExecDestroy(lru->second.exec);
        L.graphs.erase(lru);
    }
    Library::CallGraph& g = L.graphs[key];
    g.exec = exec;
    g.ws_gen = gen0;
    g.pm_tag = tag0;
    std::copy(tk0, tk0 + 4, g.tkey);
    g.use = ++L.graph_clock;
    return true;
}

// One grid call through the graph cache: replay when a valid graph exists,
// capture when the call repeats the previous one, else launch as usual.
int launch_call(Library& L, DevPlan& plan, const GridCall& call, hipStream_t st, std::string& err) {
    const std::string key = call_key(call, plan);
    auto it = L.graphs.find(key);
    if (it == L.graphs.end() && key == L.last_key) {
        if (capture_call(L, plan, call, key)) it = L.graphs.find(key);
    }
    if (it != L.graphs.end()) {
        Library::CallGraph& g = it->second;
        if (g.ws_gen == L.ws.gen && L.ws.tags["pmask"] == g.pm_tag && std::equal(g.tkey, g.tkey + 4, plan.tkey)) {
            g.use = ++L.graph_clock;
            L.last_key = key;
            if (hipGraphLaunch(g.exec, st) != hipSuccess) {
                err = "hipGraphLaunch failed";
                return -1;
            }
            return 0;
        }
        (void)hipGraphExecDestroy(g.exec);
        L.graphs.erase(it);
    }
    L.last_key = key;
    return launch_grid(plan, L.ws, call, st, nullptr, err);
}

// ---------------------------------------------------------------- tuning
std::string tune_key(const hpglx_grid_request_t* q) {
    const CovParams c = to_cov(q->cov);
    return plan_key(q->kind, c, q->radiuses, q->max_neighbours) + "|" + std::to_string(q->dims[0]) + "x" +
           std::to_string(q->dims[1]) + "x" + std::to_string(q->dims[2]) + "|p" + std::to_string(q->precision);
}

void read_tune_file(Library& L) {
    L.tune_file_read = true;
    const char* path = std::getenv("HPGLX_TUNE_FILE");
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        char* tab = std::strchr(line, '\t');
        Library::Tuning t;
        if (!tab || std::sscanf(tab + 1, "%d %d %d", &t.search_head, &t.blocks_per_cu, &t.dyn_sched) != 3) continue;
        L.tuned[std::string(line, tab)] = t;
    }
    std::fclose(f);
}

void apply_tuning(Library& L, const hpglx_grid_request_t* q, GridCall& call) {
    if (!L.tune_file_read) read_tune_file(L);
    if (L.tuned.empty()) return;
    auto it = L.tuned.find(tune_key(q));
    if (it == L.tuned.end()) return;
    if (!std::getenv("HPGLX_SEARCH_HEAD")) call.search_head = it->second.search_head;
    call.blocks_per_cu = it->second.blocks_per_cu;
    call.dyn_sched = it->second.dyn_sched;
}

int run_request(Library& L, const hpglx_grid_request_t* q, hipStream_t st, std::string& err,
                const int* cats_host = nullptr) {
    if (!q) {
        err = "null request";
        return -1;
    }
    if (!ensure_events(L, err)) return -1;
    GridCall call;
    if (!request_call(q, call, cats_host, err)) return -1;
    apply_tuning(L, q, call);
    const CovParams c = to_cov(q->cov);
    DevPlan* plan = get_plan(L, q->kind, c, q->radiuses, q->max_neighbours, err);
    if (!plan) return -1;
    // The previous call's work must be finished before its workspace is reused.
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess) {
        err = "hipStreamWaitEvent failed";
        return -1;
    }
    CallTiming* t = nullptr;
    if (L.profiling && !(t = next_timing(L, err))) return -1;
    // profiled calls record events on their dispatches, so they run eagerly
    int rc = (!t && graphs_enabled(L)) ? launch_call(L, *plan, call, st, err)
                                       : launch_grid(*plan, L.ws, call, st, t, err);
    if (rc == 0 && hipEventRecord(L.done, st) != hipSuccess) {
        err = "hipEventRecord failed";
        return -1;
    }
    if (rc == 0) L.last_counters = (const int*)L.ws.bufs["counters"].first;
    return rc;
}

bool same_ik_model(const hpgl_ik_params_t& a, const hpgl_ik_params_t& b) {
    bool same = a.m_covariance_type == b.m_covariance_type && a.m_sill == b.m_sill && a.m_nugget == b.m_nugget &&
                a.m_max_neighbours == b.m_max_neighbours;
    for (int i = 0; i < 3; ++i)
        same = same && a.m_ranges[i] == b.m_ranges[i] && a.m_angles[i] == b.m_angles[i] &&
               a.m_radiuses[i] == b.m_radiuses[i];
    return same;
}

// Indicator kriging over categories (indicator_kriging.h:88-161): one median-IK
// style pass per group of categories sharing model and neighbourhood, each
// writing its categories' probability rows, then the per-cell correction and
// most probable category.
int run_indicator(Library& L, const hpgl_ik_params_t* P, int ncat, const int dims[3], const unsigned char* data,
                  const unsigned char* mask, unsigned char* out_cat, unsigned char* out_mask, float* out_prob,
                  int precision, hipStream_t st, std::string& err) {
    if (!P || !dims || !data || !mask) {
        err = "indicator kriging: parameters, dims, data and mask are required";
        return -1;
    }
    if (ncat < 1 || ncat > 256) {
        err = "indicator kriging: the category count must be 1..256";
        return -1;
    }
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || n > 2147483647LL) {
        err = "indicator kriging: grid dimensions must be positive with a volume within INT_MAX";
        return -1;
    }
    float* prob = out_prob;
    if (!prob) {
        prob = (float*)L.ws.get("ik_prob", (size_t)ncat * (size_t)n * 4, err);
        if (!prob) return -1;
    }
    std::vector<char> taken(ncat, 0);
    for (int c = 0; c < ncat; ++c) {
        if (taken[c]) continue;
        std::vector<int> cats;
        std::vector<double> marg;
        for (int d = c; d < ncat && cats.size() < 16; ++d)
            if (!taken[d] && same_ik_model(P[c], P[d])) {
                taken[d] = 1;
                cats.push_back(d);
                marg.push_back(P[d].m_marginal_prob);
            }
        hpglx_grid_request_t q;
        std::memset(&q, 0, sizeof q);
        q.kind = KIND_MIK;
        q.precision = precision;
        for (int i = 0; i < 3; ++i) {
            q.dims[i] = dims[i];
            q.radiuses[i] = P[c].m_radiuses[i];
            q.cov.m_ranges[i] = P[c].m_ranges[i];
            q.cov.m_angles[i] = P[c].m_angles[i];
        }
        q.max_neighbours = P[c].m_max_neighbours;
        q.cov.m_covariance_type = P[c].m_covariance_type;
        q.cov.m_sill = P[c].m_sill;
        q.cov.m_nugget = P[c].m_nugget;
        q.ind = data;
        q.nind = (int)cats.size();
        q.marginal = marg.data();
        q.mask = mask;
        q.out_prob = prob;
        q.refine_steps = -1;
        if (run_request(L, &q, st, err, cats.data()) != 0) return -1;
    }
    if (ik_finalize(n, ncat, prob, out_cat, out_mask, st, err) != 0) return -1;
    if (hipEventRecord(L.done, st) != hipSuccess) {
        err = "hipEventRecord failed";
        return -1;
    }
    return 0;
}

bool ensure_io_stream(Library& L, std::string& err) {
    if (!L.io_stream && hipStreamCreateWithFlags(&L.io_stream, hipStreamNonBlocking) != hipSuccess) {
        err = "hipStreamCreate failed";
        return false;
    }
    if (!ensure_events(L, err)) return false;
    if (hipStreamWaitEvent(L.io_stream, L.done, 0) != hipSuccess) {
        err = "hipStreamWaitEvent failed";
        return false;
    }
    return true;
}

// read_inc_file_float / read_inc_file_byte (read_inc_file.cpp:137-193) into
// device buffers: the file is read on the host, its data section parsed on
// the GPU.
int read_property(Library& L, const char* filename, int as_int, float undef_f, int undef_i, int size, float* d_f,
                  unsigned char* d_b, unsigned char* d_mask, const unsigned char* values, int values_count,
                  hipStream_t st, std::string& err) {
    if (!filename) {
        err = "null file name";
        return -1;
    }
    const long long n = inc_load_file(L.text, filename, err);
    if (n < 0) return -1;
    std::string name;
    const long long off = inc_data_offset(L.text.host, n, name, err);
    if (off < 0) return -1;
    if (inc_parse(L.text.host + off, n - off, size, as_int, undef_f, undef_i, d_f, d_b, d_mask, L.ws, st, err) != 0)
        return -1;
    if (as_int && values && values_count > 0 &&
        inc_remap_bytes(d_b, d_mask, size, values, values_count, L.ws, st, err) != 0)
        return -1;
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "hipStreamSynchronize failed";
        return -1;
    }
    return 0;
}

// property_writer.cpp:44-139: INC (name line, values, "/") or GSLIB
// ("HPGL saved GSLIB file", 1, name line, values).
int write_property(Library& L, const char* filename, int gslib, const float* d_f, const unsigned char* d_b,
                   const unsigned char* d_mask, int size, float undef_f, unsigned char undef_b,
                   const unsigned char* remap, int remap_count, const char* name, hipStream_t st, std::string& err) {
    if (!filename || !name) {
        err = "null file or property name";
        return -1;
    }
    const long long len =
        inc_format(d_f, d_b, d_mask, size, undef_f, undef_b, remap, remap_count, L.text, L.ws, st, err);
    if (len < 0) return -1;
    FILE* f = std::fopen(filename, "w+");
    if (!f) {
        err = std::string("Can't open file '") + filename + "'.";
        return -1;
    }
    bool ok = true;
    if (gslib)
        ok = std::fprintf(f, "HPGL saved GSLIB file\n%d\n%s\n", 1, name) >= 0;
    else
        ok = std::fprintf(f, "%s\n", name) >= 0;
    if (ok && len > 0) ok = std::fwrite(L.text.host, 1, (size_t)len, f) == (size_t)len;
    if (ok && !gslib) ok = std::fputs("/\n", f) >= 0;
    if (std::fclose(f) != 0) ok = false;
    if (!ok) {
        err = std::string("Error writing file '") + filename + "'.";
        return -1;
    }
    return 0;
}

// api.cpp:143-167: the values table for byte properties.
std::vector<unsigned char> remap_table(Library& L, const unsigned char* values, int values_count, int indicator_count) {
    std::vector<unsigned char> t;
    if (values_count == indicator_count) {
        t.assign(values, values + values_count);
    } else if (values_count > indicator_count) {
        write_msg(L, "Given " + std::to_string(values_count) + " values for " + std::to_string(indicator_count) +
                         " indicators. Ignoring extra values.\n");
        t.assign(values, values + indicator_count);
    } else {
        if (values_count > 0)
            write_msg(L, "Warning: Given " + std::to_string(values_count) + " values for " +
                             std::to_string(indicator_count) + " indicators. Ignoring values. Using [0, 1, 2 ..]\n");
        for (int i = 0; i < indicator_count; ++i) t.push_back((unsigned char)i);
    }
    return t;
}

VarioTemplate vario_template(const variogram_search_template_t& t) {
    VarioTemplate v;
    v.lag_width = t.m_lag_width;
    v.lag_separation = t.m_lag_separation;
    v.tol_distance = t.m_tol_distance;
    v.first_lag_distance = t.m_first_lag_distance;
    v.num_lags = t.m_num_lags;
    const vector_t* d[3] = {&t.m_ellipsoid.m_direction1, &t.m_ellipsoid.m_direction2, &t.m_ellipsoid.m_direction3};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v.dir[i][j] = d[i]->m_data[j];
    v.R[0] = t.m_ellipsoid.m_R1;
    v.R[1] = t.m_ellipsoid.m_R2;
    v.R[2] = t.m_ellipsoid.m_R3;
    return v;
}

int fail(Library& L, const std::string& msg) {
    L.last_error = msg;
    return -1;
}

int shape_volume(const hpgl_shape_t* s, std::string& err) {
    long long v = 1;
    for (int i = 0; i < 3; ++i) v *= s->m_data[i];
    if (v > 2147483647LL || v <= 0) {
        err = "get_shape_volume: grid volume exceeds INT_MAX";
        return -1;
    }
    return (int)v;
}

// Host-buffer driver shared by the reference entry points.
//
// Host arrays travel through a page-locked buffer of the library's: the
// runtime copies pageable memory through staging of its own at about half
// the DMA rate (C2's 2.5 MB of inputs: 0.14 ms, against 0.06 ms from
// page-locked memory and 0.04 ms for the CPU to copy them there).  Inputs go
// in chunks, each chunk's DMA running while the CPU copies the next, every
// array into a region of its own (a queued DMA still reads it); outputs come
// back into the buffer's start (queued after every input DMA and kernel of
// the call) and are copied out once the stream is done.  Calls that would
// need more than kPinnedMax bytes of it (or when page-locking fails) copy
// straight from and to the caller's arrays.
constexpr size_t kPinnedMax = (size_t)256 << 20;
struct HostIO {
    int n = 0;
    float* d_val = nullptr;
    unsigned char* d_mask = nullptr;
    unsigned char* d_ind = nullptr;
    float* d_out = nullptr;
    unsigned char* d_out_mask = nullptr;
    unsigned char* d_out_class = nullptr;
    char* stage = nullptr;  // the page-locked buffer and its bytes in use by this call's inputs
    size_t staged = 0, stage_cap = 0;
};

// The page-locked buffer, at least `bytes` long.  Every host-array entry
// point waits for its outputs, so no copy through the buffer is in flight
// here unless a call failed midway; a buffer is freed only once its stream
// has drained.
char* pinned_stage(Library& L, size_t bytes, std::string& err) {
    if (bytes <= L.pin_cap) return L.pin;
    if (L.pin) {
        if (L.io_stream) (void)hipStreamSynchronize(L.io_stream);
        (void)hipHostFree(L.pin);
    }
    L.pin = nullptr;
    L.pin_cap = 0;
    const size_t cap = std::max(bytes, (size_t)1 << 20);
    if (hipHostMalloc((void**)&L.pin, cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();  // not left for a later check to report
        L.pin = nullptr;
        err = "hipHostMalloc failed";
        return nullptr;
    }
    L.pin_cap = cap;
    return L.pin;
}

// Queues the copy of `bytes` from host memory to the device through the next
// region of the call's staging buffer (without one, straight).
bool upload(HostIO& io, void* dev, const void* host, size_t bytes, hipStream_t st, std::string& err) {
    if (!io.stage) {
        if (hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
            err = "host to device copy failed";
            return false;
        }
        return true;
    }
    constexpr size_t kChunk = (size_t)512 << 10;
    const size_t at = (io.staged + 255) / 256 * 256;
    if (at + bytes > io.stage_cap) {
        err = "host staging buffer too small";
        return false;
    }
    for (size_t off = 0; off < bytes; off += kChunk) {
        const size_t len = std::min(kChunk, bytes - off);
        std::memcpy(io.stage + at + off, (const char*)host + off, len);
        if (hipMemcpyAsync((char*)dev + off, io.stage + at + off, len, hipMemcpyHostToDevice, st) != hipSuccess) {
            err = "host to device copy failed";
            return false;
        }
    }
    io.staged = at + bytes;
    return true;
}

// Copies outputs back: device to the staging buffer, wait for the stream,
// then to the host arrays.  Each entry is (host destination, device source,
// bytes).
struct Download {
    void* host;
    const void* dev;
    size_t bytes;
};
bool download(HostIO& io, std::initializer_list<Download> outs, hipStream_t st, std::string& err) {
    if (!io.stage) {
        for (const Download& d : outs)
            if (hipMemcpyAsync(d.host, d.dev, d.bytes, hipMemcpyDeviceToHost, st) != hipSuccess) {
                err = "device to host copy failed";
                return false;
            }
        if (hipStreamSynchronize(st) != hipSuccess) {
            err = "device to host copy failed";
            return false;
        }
        return true;
    }
    size_t at = 0;
    for (const Download& d : outs) {
        if (at + d.bytes > io.stage_cap ||
            hipMemcpyAsync(io.stage + at, d.dev, d.bytes, hipMemcpyDeviceToHost, st) != hipSuccess) {
            err = "device to host copy failed";
            return false;
        }
        at = (at + d.bytes + 255) / 256 * 256;
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "device to host copy failed";
        return false;
    }
    at = 0;
    for (const Download& d : outs) {
        std::memcpy(d.host, io.stage + at, d.bytes);
        at = (at + d.bytes + 255) / 256 * 256;
    }
    return true;
}

bool stage_in(Library& L, HostIO& io, int n, const float* val, const unsigned char* mask, const unsigned char* ind,
              std::string& err) {
    io.n = n;
    io.d_mask = (unsigned char*)L.ws.get("io_mask", n, err);
    io.d_out_mask = (unsigned char*)L.ws.get("io_out_mask", n, err);
    if (val) {
        io.d_val = (float*)L.ws.get("io_val", (size_t)n * 4, err);
        io.d_out = (float*)L.ws.get("io_out", (size_t)n * 4, err);
    }
    if (ind) {
        io.d_ind = (unsigned char*)L.ws.get("io_ind", n, err);
        io.d_out_class = (unsigned char*)L.ws.get("io_out_class", n, err);
    }
    if (!err.empty()) return false;
    if (!L.io_stream && hipStreamCreateWithFlags(&L.io_stream, hipStreamNonBlocking) != hipSuccess) {
        err = "hipStreamCreate failed";
        return false;
    }
    if (!ensure_events(L, err)) return false;
    if (hipStreamWaitEvent(L.io_stream, L.done, 0) != hipSuccess) {
        err = "hipStreamWaitEvent failed";
        return false;
    }
    // room for every input (mask, values or indicators, and a second value
    // array for the LVM mean) and, from the start, the outputs
    io.stage_cap = (size_t)n * 10 + 4 * 256;
    io.stage = nullptr;
    io.staged = 0;
    const char* sv = std::getenv("HPGLX_HOST_STAGING");  // "0": straight copies
    if (io.stage_cap <= kPinnedMax && !(sv && std::atoi(sv) == 0)) {
        std::string perr;
        io.stage = pinned_stage(L, io.stage_cap, perr);  // without it the copies go straight
    }
    return upload(io, io.d_mask, mask, n, L.io_stream, err) &&
           (!val || upload(io, io.d_val, val, (size_t)n * 4, L.io_stream, err)) &&
           (!ind || upload(io, io.d_ind, ind, n, L.io_stream, err));
}

bool sgs_transforms(const hpgl_non_parametric_cdf_t* cdf) {
    return cdf && cdf->m_size > 0 && cdf->m_values && cdf->m_probs;
}

// Sequential Gaussian simulation (sequential_gaussian_simulation.cpp:18-78):
// the normal-score transform and each realization's path, draws and levels
// on the host, the levels on the GPU.  v and m (host, N) are the data; v is
// transformed in place when the CDF is given.  Realization r (seed seeds[r])
// is written, still in normal scores, to d_v + r N and d_m + r N (device);
// all of them advance together, one launch per level index.  simmask (host,
// N) and lvm_host (host, N) may be null.
int sgs_core(Library& L, const hpgl_sgs_params_t* p, const int dims[3], std::vector<float>& v,
             std::vector<unsigned char>& m, const unsigned char* simmask, const hpgl_non_parametric_cdf_t* cdf,
             const float* lvm_host, const long* seeds, int nreal, float* d_v, unsigned char* d_m, hipStream_t st,
             std::string& err) {
    // Every check comes before the schedules are built and any copy is queued.
    if (p->m_kriging_kind != KIND_OK && p->m_kriging_kind != KIND_SK) {
        err = "sequential simulation takes simple (1) or ordinary (0) kriging";
        return -1;
    }
    if (p->m_max_neighbours > 64) {
        err = "max_neighbours above 64 is not supported by the GPU simulation";
        return -1;
    }
    if (lvm_host && p->m_kriging_kind != KIND_SK) {
        err = "a locally varying mean applies to simple kriging only";
        return -1;
    }
    CovParams c;
    c.type = p->m_covariance_type;
    for (int i = 0; i < 3; ++i) {
        c.ranges[i] = p->m_ranges[i];
        c.angles[i] = p->m_angles[i];
    }
    c.sill = p->m_sill;
    c.nugget = p->m_nugget;
    // Both kinds krige over the precalculated table of the radius box, zero
    // outside it (sequential_simulation.h:72): the simple-kriging plan.
    DevPlan* plan = get_plan(L, KIND_SK, c, p->m_radiuses, p->m_max_neighbours, err);
    if (!plan) return -1;
    const long long n = (long long)v.size();
    const bool transform = sgs_transforms(cdf);
    if (transform) sgs_to_normal_scores(v.data(), m.data(), n, cdf->m_values, cdf->m_probs, cdf->m_size);
    std::vector<float> lvm;
    float* d_lvm = nullptr;
    if (lvm_host) {
        lvm.assign(lvm_host, lvm_host + n);
        if (transform) sgs_means_to_normal_scores(lvm.data(), n, cdf->m_values, cdf->m_probs, cdf->m_size);
        d_lvm = (float*)L.ws.get("sgs_io_lvm", (size_t)n * 4, err);
        if (!err.empty()) return -1;
    }
    // The schedules are independent host work: one thread per realization,
    // at most 16 at a time.  A cell's search reaches only the template's
    // offsets, so the levels separate cells within the template's reach (not
    // the whole radius box).
    const int reach[3] = {plan->h.H[0], plan->h.H[1], plan->h.H[2]};
    std::vector<SgsSchedule> sched((size_t)nreal);
    std::vector<std::string> errs((size_t)nreal);
    std::vector<char> ok((size_t)nreal, 0);
    const int workers = std::max(1, std::min(nreal, 16));
    auto work = [&](int w) {
        for (int r = w; r < nreal; r += workers) {
            try {
                ok[(size_t)r] = sgs_schedule(dims, reach, m.data(), simmask, seeds[r], sched[(size_t)r],
                                             errs[(size_t)r], nreal == 1);
            } catch (const std::exception& e) {  // out of host memory: reported, not fatal
                errs[(size_t)r] = std::string("sequential simulation schedule: ") + e.what();
            }
        }
    };
    if (workers == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int w = 0; w < workers; ++w) pool.emplace_back(work, w);
        for (auto& t : pool) t.join();
    }
    for (int r = 0; r < nreal; ++r)
        if (!ok[(size_t)r]) {
            err = errs[(size_t)r];
            return -1;
        }
    if (!ensure_events(L, err)) return -1;
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess ||
        (d_lvm && hipMemcpyAsync(d_lvm, lvm.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess)) {
        err = "host to device copy failed";
        (void)hipStreamSynchronize(st);  // a queued copy must not outlive the host arrays
        return -1;
    }
    // the data once from the host, then on the device into every other slice
    for (int r = 0; r < nreal; ++r) {
        const void* sv = r == 0 ? (const void*)v.data() : (const void*)d_v;
        const void* sm = r == 0 ? (const void*)m.data() : (const void*)d_m;
        const hipMemcpyKind kind = r == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        if (hipMemcpyAsync(d_v + (long long)r * n, sv, (size_t)n * 4, kind, st) != hipSuccess ||
            hipMemcpyAsync(d_m + (long long)r * n, sm, (size_t)n, kind, st) != hipSuccess) {
            err = "host to device copy failed";
            (void)hipStreamSynchronize(st);
            return -1;
        }
    }
    // sgs_run synchronises st before returning on success; on an error the
    // stream is synchronised here, so the host arrays above outlive their
    // copies either way.
    const int rc = sgs_run(*plan, L.ws, p->m_kriging_kind, dims, d_v, d_m, d_lvm, sched, st, err);
    if (rc != 0) (void)hipStreamSynchronize(st);
    return rc;
}

// One realization over host copies of the property: v and m (N) are updated
// in place, back in the data's units when the CDF is given.
int run_sgs(Library& L, const hpgl_sgs_params_t* p, const int dims[3], std::vector<float>& v,
            std::vector<unsigned char>& m, const unsigned char* simmask, const hpgl_non_parametric_cdf_t* cdf,
            const float* lvm_host, hipStream_t st, std::string& err) {
    const long long n = (long long)v.size();
    auto* dv = (float*)L.ws.get("sgs_io_val", (size_t)n * 4, err);
    auto* dm = (unsigned char*)L.ws.get("sgs_io_mask", (size_t)n, err);
    if (!err.empty()) return -1;
    if (sgs_core(L, p, dims, v, m, simmask, cdf, lvm_host, &p->m_seed, 1, dv, dm, st, err) != 0) return -1;
    if (hipMemcpyAsync(v.data(), dv, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(m.data(), dm, (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(L.done, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        err = "device to host copy failed";
        return -1;
    }
    if (sgs_transforms(cdf)) sgs_from_normal_scores(v.data(), m.data(), n, cdf->m_values, cdf->m_probs, cdf->m_size);
    return 0;
}

template <typename F>
void timed_call(Library& L, const char* name, long long cells, F&& f) {
    auto t0 = std::chrono::steady_clock::now();
    progress(L, name, 0);
    std::string err;
    int rc = f(err);
    if (rc != 0) {
        L.last_error = err;
        write_msg(L, std::string(name) + " failed: " + err + "\n");
        return;
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    progress(L, name, -1);
    char b[256];
    std::snprintf(b, sizeof b, "Done. Average speed: %.0f point/sec.\n", sec > 0 ? cells / sec : 0.0);
    write_msg(L, b);
}

}  // namespace
}  // namespace hpglx

using namespace hpglx;

extern "C" {

// ------------------------------------------------------- reference ABI
char* hpgl_get_last_exception_message(void) {
    return (char*)lib().last_error.c_str();
}

void hpgl_set_output_handler(int (*handler)(char*, void*), void* param) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.out_handler = handler;
    L.out_param = param;
}

void hpgl_set_progress_handler(int (*handler)(char*, int, void*), void* param) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.progress_handler = handler;
    L.progress_param = param;
}

void hpgl_set_thread_num(int n_threads) {
    lib().thread_num = n_threads;
}

int hpgl_get_thread_num(void) {
    return lib().thread_num;
}

void hpgl_ordinary_kriging(hpgl_cont_masked_array_t* in, hpgl_ok_params_t* p, hpgl_cont_masked_array_t* out) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(&in->m_shape, err);
    if (n < 0 || shape_volume(&out->m_shape, err) != n) {
        L.last_error = err.empty() ? "input and output sizes differ" : err;
        return;
    }
    timed_call(L, "Ordinary Kriging", n, [&](std::string& e) -> int {
        HostIO io;
        if (!stage_in(L, io, n, in->m_data, in->m_mask, nullptr, e)) return -1;
        hpglx_grid_request_t q;
        std::memset(&q, 0, sizeof q);
        q.kind = KIND_OK;
        q.precision = L.default_precision;
        for (int i = 0; i < 3; ++i) {
            q.dims[i] = in->m_shape.m_data[i];
            q.radiuses[i] = p->m_radiuses[i];
        }
        q.max_neighbours = p->m_max_neighbours;
        q.cov.m_covariance_type = p->m_covariance_type;
        for (int i = 0; i < 3; ++i) {
            q.cov.m_ranges[i] = p->m_ranges[i];
            q.cov.m_angles[i] = p->m_angles[i];
        }
        q.cov.m_sill = p->m_sill;
        q.cov.m_nugget = p->m_nugget;
        q.values = io.d_val;
        q.mask = io.d_mask;
        q.out_val = io.d_out;
        q.out_mask = io.d_out_mask;
        q.refine_steps = -1;
        if (run_request(L, &q, L.io_stream, e) != 0) return -1;
        if (!download(io, {{out->m_data, io.d_out, (size_t)n * 4}, {out->m_mask, io.d_out_mask, (size_t)n}},
                      L.io_stream, e))
            return -1;
        return 0;
    });
}

void hpgl_simple_kriging(float* input_data, unsigned char* input_mask, hpgl_shape_t* in_shape, hpgl_sk_params_t* p,
                         float* output_data, unsigned char* output_mask, hpgl_shape_t* out_shape) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(in_shape, err);
    if (n < 0 || shape_volume(out_shape, err) != n) {
        L.last_error = err.empty() ? "input and output sizes differ" : err;
        return;
    }
    timed_call(L, "Simple Kriging", n, [&](std::string& e) -> int {
        HostIO io;
        if (!stage_in(L, io, n, input_data, input_mask, nullptr, e)) return -1;
        hpglx_grid_request_t q;
        std::memset(&q, 0, sizeof q);
        q.kind = KIND_SK;
        q.precision = L.default_precision;
        for (int i = 0; i < 3; ++i) {
            q.dims[i] = in_shape->m_data[i];
            q.radiuses[i] = p->m_radiuses[i];
        }
        q.max_neighbours = p->m_max_neighbours;
        q.cov.m_covariance_type = p->m_covariance_type;
        for (int i = 0; i < 3; ++i) {
            q.cov.m_ranges[i] = p->m_ranges[i];
            q.cov.m_angles[i] = p->m_angles[i];
        }
        q.cov.m_sill = p->m_sill;
        q.cov.m_nugget = p->m_nugget;
        q.auto_mean = p->m_automatic_mean ? 1 : 0;
        q.mean = p->m_mean;
        q.values = io.d_val;
        q.mask = io.d_mask;
        q.out_val = io.d_out;
        q.out_mask = io.d_out_mask;
        q.refine_steps = -1;
        if (run_request(L, &q, L.io_stream, e) != 0) return -1;
        if (!download(io, {{output_data, io.d_out, (size_t)n * 4}, {output_mask, io.d_out_mask, (size_t)n}},
                      L.io_stream, e))
            return -1;
        return 0;
    });
}

// api.cpp:485-516.  The stationary mean is accepted and, as in the
// reference, not used (simple kriging runs with mean 0 in normal-score space).
void hpgl_sgs_simulation(hpgl_cont_masked_array_t* data, hpgl_sgs_params_t* params, hpgl_non_parametric_cdf_t* cdf,
                         double* mean, hpgl_ubyte_array_t* simulation_mask) {
    (void)mean;
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(&data->m_shape, err);
    if (n < 0) {
        L.last_error = err;
        return;
    }
    if (simulation_mask && shape_volume(&simulation_mask->m_shape, err) != n) {
        L.last_error = err.empty() ? "simulation mask and property sizes differ" : err;
        return;
    }
    timed_call(L, "Sequential Gaussian Simulation", n, [&](std::string& e) -> int {
        if (!L.io_stream && hipStreamCreateWithFlags(&L.io_stream, hipStreamNonBlocking) != hipSuccess) {
            e = "hipStreamCreate failed";
            return -1;
        }
        std::vector<float> v(data->m_data, data->m_data + n);
        std::vector<unsigned char> m(data->m_mask, data->m_mask + n);
        const int dims[3] = {data->m_shape.m_data[0], data->m_shape.m_data[1], data->m_shape.m_data[2]};
        if (run_sgs(L, params, dims, v, m, simulation_mask ? simulation_mask->m_data : nullptr, cdf, nullptr,
                    L.io_stream, e) != 0)
            return -1;
        std::memcpy(data->m_data, v.data(), (size_t)n * 4);
        std::memcpy(data->m_mask, m.data(), (size_t)n);
        return 0;
    });
}

// api.cpp:519-548: simple kriging with a locally varying mean (one per cell,
// taken through the normal-score transform with the data).
void hpgl_sgs_lvm_simulation(hpgl_cont_masked_array_t* data, hpgl_sgs_params_t* params,
                             hpgl_non_parametric_cdf_t* cdf, hpgl_float_array_t* means,
                             hpgl_ubyte_array_t* simulation_mask) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(&data->m_shape, err);
    if (n < 0) {
        L.last_error = err;
        return;
    }
    if (!means || !means->m_data || shape_volume(&means->m_shape, err) != n) {
        L.last_error = err.empty() ? "the mean array and the property sizes differ" : err;
        return;
    }
    if (simulation_mask && shape_volume(&simulation_mask->m_shape, err) != n) {
        L.last_error = err.empty() ? "simulation mask and property sizes differ" : err;
        return;
    }
    timed_call(L, "Sequential Gaussian Simulation with Local Varying Mean", n, [&](std::string& e) -> int {
        if (!L.io_stream && hipStreamCreateWithFlags(&L.io_stream, hipStreamNonBlocking) != hipSuccess) {
            e = "hipStreamCreate failed";
            return -1;
        }
        std::vector<float> v(data->m_data, data->m_data + n);
        std::vector<unsigned char> m(data->m_mask, data->m_mask + n);
        const int dims[3] = {data->m_shape.m_data[0], data->m_shape.m_data[1], data->m_shape.m_data[2]};
        if (run_sgs(L, params, dims, v, m, simulation_mask ? simulation_mask->m_data : nullptr, cdf, means->m_data,
                    L.io_stream, e) != 0)
            return -1;
        std::memcpy(data->m_data, v.data(), (size_t)n * 4);
        std::memcpy(data->m_mask, m.data(), (size_t)n);
        return 0;
    });
}

void hpgl_median_ik(hpgl_ind_masked_array_t* in, hpgl_median_ik_params_t* p, hpgl_ind_masked_array_t* out) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(&in->m_shape, err);
    if (n < 0 || shape_volume(&out->m_shape, err) != n) {
        L.last_error = err.empty() ? "input and output sizes differ" : err;
        return;
    }
    timed_call(L, "Median IK", n, [&](std::string& e) -> int {
        HostIO io;
        if (!stage_in(L, io, n, nullptr, in->m_mask, in->m_data, e)) return -1;
        hpglx_grid_request_t q;
        std::memset(&q, 0, sizeof q);
        q.kind = KIND_MIK;
        q.precision = L.default_precision;
        for (int i = 0; i < 3; ++i) {
            q.dims[i] = in->m_shape.m_data[i];
            q.radiuses[i] = p->m_radiuses[i];
        }
        q.max_neighbours = p->m_max_neighbours;
        q.cov.m_covariance_type = p->m_covariance_type;
        for (int i = 0; i < 3; ++i) {
            q.cov.m_ranges[i] = p->m_ranges[i];
            q.cov.m_angles[i] = p->m_angles[i];
        }
        q.cov.m_sill = p->m_sill;
        q.cov.m_nugget = p->m_nugget;
        double marg = p->m_marginal_probs[1];
        q.ind = io.d_ind;
        q.nind = 1;
        q.marginal = &marg;
        q.mask = io.d_mask;
        q.out_class = io.d_out_class;
        q.out_mask = io.d_out_mask;
        q.refine_steps = -1;
        if (run_request(L, &q, L.io_stream, e) != 0) return -1;
        if (!download(io, {{out->m_data, io.d_out_class, (size_t)n}, {out->m_mask, io.d_out_mask, (size_t)n}},
                      L.io_stream, e))
            return -1;
        return 0;
    });
}

void hpgl_lvm_kriging(float* input_data, unsigned char* input_mask, hpgl_shape_t* input_data_shape, float* mean_data,
                      hpgl_shape_t* mean_data_shape, hpgl_ok_params_t* p, float* output_data, unsigned char* output_mask,
                      hpgl_shape_t* output_data_shape) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(input_data_shape, err);
    if (n < 0 || shape_volume(output_data_shape, err) != n || shape_volume(mean_data_shape, err) != n) {
        L.last_error = err.empty() ? "input, mean and output sizes differ" : err;
        return;
    }
    timed_call(L, "LVM Kriging", n, [&](std::string& e) -> int {
        HostIO io;
        if (!stage_in(L, io, n, input_data, input_mask, nullptr, e)) return -1;
        float* d_mean = (float*)L.ws.get("io_lvm_mean", (size_t)n * 4, e);
        if (!d_mean || !upload(io, d_mean, mean_data, (size_t)n * 4, L.io_stream, e)) return -1;
        hpglx_grid_request_t q;
        std::memset(&q, 0, sizeof q);
        q.kind = KIND_SK;
        q.precision = L.default_precision;
        for (int i = 0; i < 3; ++i) {
            q.dims[i] = input_data_shape->m_data[i];
            q.radiuses[i] = p->m_radiuses[i];
            q.cov.m_ranges[i] = p->m_ranges[i];
            q.cov.m_angles[i] = p->m_angles[i];
        }
        q.max_neighbours = p->m_max_neighbours;
        q.cov.m_covariance_type = p->m_covariance_type;
        q.cov.m_sill = p->m_sill;
        q.cov.m_nugget = p->m_nugget;
        q.values = io.d_val;
        q.mask = io.d_mask;
        q.lvm_mean = d_mean;
        q.out_val = io.d_out;
        q.out_mask = io.d_out_mask;
        q.refine_steps = -1;
        if (run_request(L, &q, L.io_stream, e) != 0) return -1;
        if (!download(io, {{output_data, io.d_out, (size_t)n * 4}, {output_mask, io.d_out_mask, (size_t)n}},
                      L.io_stream, e))
            return -1;
        return 0;
    });
}

void hpgl_indicator_kriging(hpgl_ind_masked_array_t* in, hpgl_ind_masked_array_t* out, hpgl_ik_params_t* params,
                            int indicator_count) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int n = shape_volume(&in->m_shape, err);
    if (n < 0 || shape_volume(&out->m_shape, err) != n) {
        L.last_error = err.empty() ? "hpgl_indicator_kriging: input and output size mismatch" : err;
        return;
    }
    timed_call(L, "Indicator Kriging", n, [&](std::string& e) -> int {
        HostIO io;
        if (!stage_in(L, io, n, nullptr, in->m_mask, in->m_data, e)) return -1;
        if (run_indicator(L, params, indicator_count, in->m_shape.m_data, io.d_ind, io.d_mask, io.d_out_class,
                          io.d_out_mask, nullptr, L.default_precision, L.io_stream, e) != 0)
            return -1;
        if (!download(io, {{out->m_data, io.d_out_class, (size_t)n}, {out->m_mask, io.d_out_mask, (size_t)n}},
                      L.io_stream, e))
            return -1;
        return 0;
    });
}

// ------------------------------------------------ property files (INC/GSLIB)
int hpgl_read_inc_file_float(char* filename, float undefined_value, int size, float* data, unsigned char* mask) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (size > 0 && (!data || !mask)) return fail(L, "null output buffers");
    if (!ensure_io_stream(L, err)) return fail(L, err);
    const size_t n = size > 0 ? (size_t)size : 1;
    float* d_f = (float*)L.ws.get("io_read_f", n * 4, err);
    unsigned char* d_m = (unsigned char*)L.ws.get("io_read_mask", n, err);
    if (!err.empty()) return fail(L, err);
    if (read_property(L, filename, 0, undefined_value, 0, size, d_f, nullptr, d_m, nullptr, 0, L.io_stream, err) != 0)
        return fail(L, err);
    if (size > 0 && (hipMemcpy(data, d_f, (size_t)size * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(mask, d_m, (size_t)size, hipMemcpyDeviceToHost) != hipSuccess))
        return fail(L, "device to host copy failed");
    return 0;
}

int hpgl_read_inc_file_byte(char* filename, int undefined_value, int size, unsigned char* data, unsigned char* mask,
                            unsigned char* values, int values_count) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (size > 0 && (!data || !mask)) return fail(L, "null output buffers");
    if (!ensure_io_stream(L, err)) return fail(L, err);
    const size_t n = size > 0 ? (size_t)size : 1;
    unsigned char* d_b = (unsigned char*)L.ws.get("io_read_b", n, err);
    unsigned char* d_m = (unsigned char*)L.ws.get("io_read_mask", n, err);
    if (!err.empty()) return fail(L, err);
    if (read_property(L, filename, 1, 0.0f, undefined_value, size, nullptr, d_b, d_m, values, values_count,
                      L.io_stream, err) != 0)
        return fail(L, err);
    if (size > 0 && (hipMemcpy(data, d_b, (size_t)size, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(mask, d_m, (size_t)size, hipMemcpyDeviceToHost) != hipSuccess))
        return fail(L, "device to host copy failed");
    return 0;
}

namespace {
int write_host_property(Library& L, const char* filename, int gslib, const float* f, const unsigned char* b,
                        const unsigned char* mask, int size, float undef_f, unsigned char undef_b,
                        const std::vector<unsigned char>& remap, const char* name) {
    std::string err;
    if (size > 0 && (!mask || (!f && !b))) return fail(L, "null property buffers");
    if (!ensure_io_stream(L, err)) return fail(L, err);
    const size_t n = size > 0 ? (size_t)size : 1;
    float* d_f = f ? (float*)L.ws.get("io_write_f", n * 4, err) : nullptr;
    unsigned char* d_b = b ? (unsigned char*)L.ws.get("io_write_b", n, err) : nullptr;
    unsigned char* d_m = (unsigned char*)L.ws.get("io_write_mask", n, err);
    if (!err.empty()) return fail(L, err);
    if (size > 0 && ((f && hipMemcpy(d_f, f, (size_t)size * 4, hipMemcpyHostToDevice) != hipSuccess) ||
                     (b && hipMemcpy(d_b, b, (size_t)size, hipMemcpyHostToDevice) != hipSuccess) ||
                     hipMemcpy(d_m, mask, (size_t)size, hipMemcpyHostToDevice) != hipSuccess))
        return fail(L, "host to device copy failed");
    if (write_property(L, filename, gslib, d_f, d_b, d_m, size, undef_f, undef_b, remap.data(), (int)remap.size(),
                       name, L.io_stream, err) != 0)
        return fail(L, err);
    return 0;
}

int shape_volume_or_fail(Library& L, const hpgl_shape_t* s, int& n) {
    std::string err;
    n = shape_volume(s, err);
    if (n < 0) return fail(L, err);
    return 0;
}
}  // namespace

int hpgl_write_inc_file_float(char* filename, hpgl_cont_masked_array_t* arr, float undefined_value, char* name) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    int n = 0;
    if (!arr || shape_volume_or_fail(L, &arr->m_shape, n) != 0) return -1;
    return write_host_property(L, filename, 0, arr->m_data, nullptr, arr->m_mask, n, undefined_value, 0, {}, name);
}

int hpgl_write_inc_file_byte(char* filename, hpgl_ind_masked_array_t* arr, int undefined_value, char* name,
                             unsigned char* values, int values_count) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    int n = 0;
    if (!arr || shape_volume_or_fail(L, &arr->m_shape, n) != 0) return -1;
    const auto remap = remap_table(L, values, values_count, arr->m_indicator_count);
    return write_host_property(L, filename, 0, nullptr, arr->m_data, arr->m_mask, n, 0.0f,
                               (unsigned char)undefined_value, remap, name);
}

int hpgl_write_gslib_cont_property(hpgl_cont_masked_array_t* data, const char* filename, const char* name,
                                   double undefined_value) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    int n = 0;
    if (!data || shape_volume_or_fail(L, &data->m_shape, n) != 0) return -1;
    return write_host_property(L, filename, 1, data->m_data, nullptr, data->m_mask, n, (float)undefined_value, 0, {},
                               name);
}

int hpgl_write_gslib_byte_property(hpgl_ind_masked_array_t* data, const char* filename, const char* name,
                                   double undefined_value, unsigned char* values, int values_count) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    int n = 0;
    if (!data || shape_volume_or_fail(L, &data->m_shape, n) != 0) return -1;
    const auto remap = remap_table(L, values, values_count, data->m_indicator_count);
    return write_host_property(L, filename, 1, nullptr, data->m_data, data->m_mask, n, 0.0f,
                               (unsigned char)(int)undefined_value, remap, name);
}

int hpgl_simple_kriging_weights(float* center, float* xs, float* ys, float* zs, int n, hpgl_cov_params_t* params,
                                float* weights) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (n < 0) return fail(L, "hpgl_simple_kriging_weights: negative neighbour count");
    CovParams c = to_cov(*params);
    DevModel m;
    if (c.type < 0 || c.type > 2) return fail(L, "Unknown covariance type");
    if (!make_transform(c, m.t)) return fail(L, "All covariance ranges should be non-zero");
    // simple_kriging.cpp:31-35 empties the weights when the system cannot be
    // solved (n == 0 included), and api.cpp:399-402 then fails on .at(i).
    if (n == 0) return 0;
    m.type = c.type;
    m.sill = c.sill;
    m.nugget = c.nugget;
    m.r0 = c.ranges[0];
    std::vector<float> host(3 + 3 * (size_t)n);
    for (int i = 0; i < 3; ++i) host[i] = center[i];
    for (int i = 0; i < n; ++i) {
        host[3 + i] = xs[i];
        host[3 + n + i] = ys[i];
        host[3 + 2 * (size_t)n + i] = zs[i];
    }
    float* d_in = (float*)L.ws.get("skw_in", host.size() * 4, err);
    double* d_a = (double*)L.ws.get("skw_a", (size_t)n * n * 8, err);
    double* d_x = (double*)L.ws.get("skw_x", (size_t)n * 8, err);
    int* d_st = (int*)L.ws.get("skw_st", 4, err);
    if (!err.empty()) return fail(L, err);
    std::vector<double> w(n, 0.0);
    int st = 0;
    if (hipMemcpy(d_in, host.data(), host.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return fail(L, "copy failed");
    hipLaunchKernelGGL(k_sk_weights, dim3(1), dim3(256), 0, 0, m, (const float*)d_in, (const float*)(d_in + 3),
                       (const float*)(d_in + 3 + n), (const float*)(d_in + 3 + 2 * (size_t)n), n, d_a, d_x, d_st);
    if (hipMemcpy(&st, d_st, 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(w.data(), d_x, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(L, "kernel or copy failed");
    if (st != 0) return fail(L, "vector::_M_range_check: the kriging system is singular");
    for (int i = 0; i < n; ++i) weights[i] = (float)w[i];
    return 0;
}

// ------------------------------------------------------ device entry points
int hpglx_krige(const hpglx_grid_request_t* req, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (run_request(L, req, (hipStream_t)stream, err) != 0) return fail(L, err);
    return 0;
}

int hpglx_autotune(const hpglx_grid_request_t* req, void* stream, int reps, int* chosen, float* best_ms) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!req) return fail(L, "null request");
    if (reps < 1) reps = 1;
    const hipStream_t st = (hipStream_t)stream;
    if (!ensure_events(L, err)) return fail(L, err);
    for (auto& e : L.tune_ev)
        if (!e && hipEventCreate(&e) != hipSuccess) return fail(L, "hipEventCreate failed");
    GridCall base;
    if (!request_call(req, base, nullptr, err)) return fail(L, err);
    const CovParams c = to_cov(req->cov);
    DevPlan* plan = get_plan(L, req->kind, c, req->radiuses, req->max_neighbours, err);
    if (!plan) return fail(L, err);
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess) return fail(L, "hipStreamWaitEvent failed");
    // The candidates: template entries staged by the search, persistent
    // blocks per CU of the solve, fixed or dynamic dealing of its chunks.
    const int heads[] = {0, 512, 256};
    const int bpcs[] = {0, 1};
    Library::Tuning best;
    float best_t = 3.0e38f;
    for (int head : heads)
        for (int bpc : bpcs)
            for (int dyn = 0; dyn < 2; ++dyn) {
                GridCall call = base;
                call.search_head = head;
                call.blocks_per_cu = bpc;
                call.dyn_sched = dyn;
                if (launch_grid(*plan, L.ws, call, st, nullptr, err) != 0) return fail(L, err);  // warm
                if (hipEventRecord(L.tune_ev[0], st) != hipSuccess) return fail(L, "hipEventRecord failed");
                for (int r = 0; r < reps; ++r)
                    if (launch_grid(*plan, L.ws, call, st, nullptr, err) != 0) return fail(L, err);
                float ms = 0.f;
                if (hipEventRecord(L.tune_ev[1], st) != hipSuccess || hipEventSynchronize(L.tune_ev[1]) != hipSuccess ||
                    hipEventElapsedTime(&ms, L.tune_ev[0], L.tune_ev[1]) != hipSuccess)
                    return fail(L, "timing a candidate failed");
                ms /= (float)reps;
                if (ms < best_t) {
                    best_t = ms;
                    best.search_head = head;
                    best.blocks_per_cu = bpc;
                    best.dyn_sched = dyn;
                }
            }
    if (hipEventRecord(L.done, st) != hipSuccess) return fail(L, "hipEventRecord failed");
    if (!L.tune_file_read) read_tune_file(L);
    const std::string key = tune_key(req);
    L.tuned[key] = best;
    if (const char* path = std::getenv("HPGLX_TUNE_FILE")) {
        if (FILE* f = std::fopen(path, "a")) {
            std::fprintf(f, "%s\t%d %d %d\n", key.c_str(), best.search_head, best.blocks_per_cu, best.dyn_sched);
            std::fclose(f);
        }
    }
    if (chosen) {
        chosen[0] = best.search_head;
        chosen[1] = best.blocks_per_cu;
        chosen[2] = best.dyn_sched;
    }
    if (best_ms) *best_ms = best_t;
    return 0;
}

int hpglx_krige_noalloc(const hpglx_grid_request_t* req, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!req) return fail(L, "null request");
    const hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(L, "hipGetDevice failed");
    NoAllocDevice& na = L.noalloc[dev];
    if (!na.base) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(L, std::string("hipGetSymbolAddress of the allocation-free scratch failed: ") +
                               hipGetErrorString(e));
        na.base = (char*)p;
    }
    if (!ensure_events(L, err)) return fail(L, err);
    GridCall call;
    if (!request_call(req, call, nullptr, err)) return fail(L, err);
    call.cluster_lanes = kNoAllocClusterLanes;
    // The plan's host tables: built once per parameter set and device, kept
    // (at most kMaxPlans) for the copies into the scratch.
    const CovParams c = to_cov(req->cov);
    const std::string key = plan_key(req->kind, c, req->radiuses, req->max_neighbours);
    ++L.plan_clock;
    DevPlan* plan = nullptr;
    auto it = L.na_plans.find(key);
    if (it != L.na_plans.end()) {
        plan = it->second.get();
    } else {
        std::unique_ptr<DevPlan> p(new DevPlan());
        if (!build_plan(p->h, (Kind)req->kind, c, req->radiuses, req->max_neighbours, err)) return fail(L, err);
        if (L.na_plans.size() >= kMaxPlans) {
            auto lru = L.na_use.begin();
            for (auto u = L.na_use.begin(); u != L.na_use.end(); ++u)
                if (u->second < lru->second) lru = u;
            // its host tables may still feed a pending copy
            if (L.done && hipEventSynchronize(L.done) != hipSuccess) return fail(L, "hipEventSynchronize failed");
            for (auto& d : L.noalloc)
                if (d.second.resident == lru->first) d.second.resident.clear();
            L.na_plans.erase(lru->first);
            L.na_use.erase(lru);
        }
        plan = p.get();
        L.na_plans[key] = std::move(p);
    }
    L.na_use[key] = L.plan_clock;
    // After the previous call's work (of any entry point): the scratch, its
    // resident tables and its cleared halo are shared by every call.
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess) return fail(L, "hipStreamWaitEvent failed");
    int rc = 0;
    if (na.resident != key) {
        na.resident.clear();
        size_t used = 0;
        if (!plan->place(na.base, kNoAllocScratchBytes - kNoAllocGuard, used, st, err)) rc = -1;
        if (rc == 0) {
            na.resident = key;
            na.plan_bytes = used;
        }
    }
    if (rc == 0) {
        // The last 64 KiB stay unused: a margin past the last buffer, like
        // the page rounding of an allocation.
        na.ws.arena = na.base + na.plan_bytes;
        na.ws.arena_cap = kNoAllocScratchBytes - na.plan_bytes - kNoAllocGuard;
        na.ws.reset_arena();
        rc = launch_grid(*plan, na.ws, call, st, nullptr, err);
    }
    // Whatever was enqueued (table copies, kernels) orders the next call.
    if (hipEventRecord(L.done, st) != hipSuccess) return fail(L, "hipEventRecord failed");
    if (rc != 0) return fail(L, err);
    L.last_counters = (const int*)na.ws.bufs["counters"].first;
    return 0;
}

int hpglx_ordinary_kriging(const hpgl_cov_params_t* cov, const int radiuses[3], int max_nb, const int dims[3],
                           const float* values, const unsigned char* mask, float* out_val, unsigned char* out_mask,
                           double* out_var, int precision, void* stream) {
    hpglx_grid_request_t q;
    std::memset(&q, 0, sizeof q);
    q.kind = KIND_OK;
    q.precision = precision;
    for (int i = 0; i < 3; ++i) {
        q.dims[i] = dims[i];
        q.radiuses[i] = radiuses[i];
    }
    q.max_neighbours = max_nb;
    q.cov = *cov;
    q.values = values;
    q.mask = mask;
    q.out_val = out_val;
    q.out_mask = out_mask;
    q.out_var = out_var;
    q.refine_steps = -1;
    return hpglx_krige(&q, stream);
}

int hpglx_simple_kriging(const hpgl_cov_params_t* cov, const int radiuses[3], int max_nb, const int dims[3],
                         const float* values, const unsigned char* mask, int auto_mean, double mean, float* out_val,
                         unsigned char* out_mask, double* out_var, int precision, void* stream) {
    hpglx_grid_request_t q;
    std::memset(&q, 0, sizeof q);
    q.kind = KIND_SK;
    q.precision = precision;
    for (int i = 0; i < 3; ++i) {
        q.dims[i] = dims[i];
        q.radiuses[i] = radiuses[i];
    }
    q.max_neighbours = max_nb;
    q.cov = *cov;
    q.values = values;
    q.mask = mask;
    q.auto_mean = auto_mean;
    q.mean = mean;
    q.out_val = out_val;
    q.out_mask = out_mask;
    q.out_var = out_var;
    q.refine_steps = -1;
    return hpglx_krige(&q, stream);
}

int hpglx_indicator_cdf(const hpgl_cov_params_t* cov, const int radiuses[3], int max_nb, const int dims[3],
                        const float* zfield, const unsigned char* mask, int nthr, const float* thresholds,
                        const double* marginal, float* out_cdf, int precision, void* stream) {
    hpglx_grid_request_t q;
    std::memset(&q, 0, sizeof q);
    q.kind = KIND_MIK;
    q.precision = precision;
    for (int i = 0; i < 3; ++i) {
        q.dims[i] = dims[i];
        q.radiuses[i] = radiuses[i];
    }
    q.max_neighbours = max_nb;
    q.cov = *cov;
    q.zfield = zfield;
    q.thresholds = thresholds;
    q.nind = nthr;
    q.marginal = marginal;
    q.mask = mask;
    q.out_prob = out_cdf;
    q.order_relations = 1;
    q.refine_steps = -1;
    return hpglx_krige(&q, stream);
}

int hpglx_lvm_kriging(const hpgl_cov_params_t* cov, const int radiuses[3], int max_nb, const int dims[3],
                      const float* values, const unsigned char* mask, const float* mean_field, float* out_val,
                      unsigned char* out_mask, double* out_var, int precision, void* stream) {
    if (!mean_field) {
        lib().last_error = "hpglx_lvm_kriging: the mean field is required";
        return -1;
    }
    hpglx_grid_request_t q;
    std::memset(&q, 0, sizeof q);
    q.kind = KIND_SK;
    q.precision = precision;
    for (int i = 0; i < 3; ++i) {
        q.dims[i] = dims[i];
        q.radiuses[i] = radiuses[i];
    }
    q.max_neighbours = max_nb;
    q.cov = *cov;
    q.values = values;
    q.mask = mask;
    q.lvm_mean = mean_field;
    q.out_val = out_val;
    q.out_mask = out_mask;
    q.out_var = out_var;
    q.refine_steps = -1;
    return hpglx_krige(&q, stream);
}

int hpglx_indicator_kriging(const hpgl_ik_params_t* params, int ncat, const int dims[3], const unsigned char* data,
                            const unsigned char* mask, unsigned char* out_cat, unsigned char* out_mask,
                            float* out_prob, int precision, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (run_indicator(L, params, ncat, dims, data, mask, out_cat, out_mask, out_prob, precision, (hipStream_t)stream,
                      err) != 0)
        return fail(L, err);
    return 0;
}

int hpglx_sgs_simulation(float* values, unsigned char* mask, const int dims[3], const hpgl_sgs_params_t* params,
                         const hpgl_non_parametric_cdf_t* cdf, const unsigned char* simulation_mask,
                         const float* mean_field, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!values || !mask || !dims || !params) return fail(L, "hpglx_sgs_simulation: null argument");
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || n > 2147483647LL)
        return fail(L, "hpglx_sgs_simulation: invalid grid dimensions");
    const hipStream_t st = (hipStream_t)stream;
    std::vector<float> v((size_t)n);
    std::vector<unsigned char> m((size_t)n), sm(simulation_mask ? (size_t)n : 0);
    std::vector<float> lvm(mean_field ? (size_t)n : 0);
    if ((mean_field && hipMemcpyAsync(lvm.data(), mean_field, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipMemcpyAsync(v.data(), values, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(m.data(), mask, (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (simulation_mask &&
         hipMemcpyAsync(sm.data(), simulation_mask, (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(L, "hpglx_sgs_simulation: device to host copy failed");
    if (run_sgs(L, params, dims, v, m, simulation_mask ? sm.data() : nullptr, cdf, mean_field ? lvm.data() : nullptr, st,
                err) != 0)
        return fail(L, err);
    if (hipMemcpyAsync(values, v.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(mask, m.data(), (size_t)n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(L, "hpglx_sgs_simulation: host to device copy failed");
    return 0;
}

int hpglx_sgs_realizations(const float* values, const unsigned char* mask, const int dims[3],
                           const hpgl_sgs_params_t* params, const long* seeds, int nreal,
                           const hpgl_non_parametric_cdf_t* cdf, const unsigned char* simulation_mask,
                           const float* mean_field, float* out_values, unsigned char* out_mask, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!values || !mask || !dims || !params || !seeds || !out_values || !out_mask)
        return fail(L, "hpglx_sgs_realizations: null argument");
    if (nreal < 1 || nreal > 4096) return fail(L, "hpglx_sgs_realizations: nreal must be 1..4096");
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || n > 2147483647LL || n * nreal > 2147483647LL)
        return fail(L, "hpglx_sgs_realizations: invalid grid dimensions (nreal x cells must stay below 2^31)");
    const hipStream_t st = (hipStream_t)stream;
    std::vector<float> v((size_t)n);
    std::vector<unsigned char> m((size_t)n), sm(simulation_mask ? (size_t)n : 0);
    std::vector<float> lvm(mean_field ? (size_t)n : 0);
    if ((mean_field && hipMemcpyAsync(lvm.data(), mean_field, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipMemcpyAsync(v.data(), values, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(m.data(), mask, (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (simulation_mask &&
         hipMemcpyAsync(sm.data(), simulation_mask, (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(L, "hpglx_sgs_realizations: device to host copy failed");
    if (sgs_core(L, params, dims, v, m, simulation_mask ? sm.data() : nullptr, cdf, mean_field ? lvm.data() : nullptr,
                 seeds, nreal, out_values, out_mask, st, err) != 0)
        return fail(L, err);
    if (sgs_transforms(cdf)) {
        // back transform on the host, one realization at a time
        std::vector<unsigned char> rm((size_t)n);
        for (int r = 0; r < nreal; ++r) {
            float* dv = out_values + (long long)r * n;
            const unsigned char* dm = out_mask + (long long)r * n;
            if (hipMemcpyAsync(v.data(), dv, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(rm.data(), dm, (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail(L, "hpglx_sgs_realizations: device to host copy failed");
            sgs_from_normal_scores(v.data(), rm.data(), n, cdf->m_values, cdf->m_probs, cdf->m_size);
            if (hipMemcpyAsync(dv, v.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail(L, "hpglx_sgs_realizations: host to device copy failed");
        }
    }
    return 0;
}

int hpglx_sgs_schedule(const int dims[3], const int radiuses[3], const unsigned char* informed,
                       const unsigned char* simulation_mask, long seed, int* out_cells, double* out_deviates,
                       int* out_level, int capacity) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    if (!dims || !radiuses || !informed || capacity < 0 || (capacity > 0 && (!out_cells || !out_deviates || !out_level)))
        return fail(L, "hpglx_sgs_schedule: null argument");
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || (long long)dims[0] * dims[1] * dims[2] > 2147483647LL)
        return fail(L, "hpglx_sgs_schedule: invalid grid dimensions");
    if (radiuses[0] < 0 || radiuses[1] < 0 || radiuses[2] < 0)
        return fail(L, "hpglx_sgs_schedule: negative radius");
    SgsSchedule s;
    std::string err;
    if (!sgs_schedule(dims, radiuses, informed, simulation_mask, seed, s, err, true)) return fail(L, err);
    const int m = (int)s.cells.size();
    if (m > capacity) return m;
    for (size_t l = 0; l + 1 < s.level_off.size(); ++l)
        for (int i = s.level_off[l]; i < s.level_off[l + 1]; ++i) {
            out_cells[i] = s.cells[(size_t)i];
            out_deviates[i] = s.z[(size_t)i];
            out_level[i] = (int)l;
        }
    return m;
}

int hpglx_read_inc_float(const char* filename, float undefined_value, int size, float* d_data, unsigned char* d_mask,
                         void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!ensure_events(L, err)) return fail(L, err);
    if (read_property(L, filename, 0, undefined_value, 0, size, d_data, nullptr, d_mask, nullptr, 0,
                      (hipStream_t)stream, err) != 0)
        return fail(L, err);
    return 0;
}

int hpglx_read_inc_byte(const char* filename, int undefined_value, int size, unsigned char* d_data,
                        unsigned char* d_mask, const unsigned char* values, int values_count, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!ensure_events(L, err)) return fail(L, err);
    if (read_property(L, filename, 1, 0.0f, undefined_value, size, nullptr, d_data, d_mask, values, values_count,
                      (hipStream_t)stream, err) != 0)
        return fail(L, err);
    return 0;
}

int hpglx_write_inc_float(const char* filename, const float* d_data, const unsigned char* d_mask, int size,
                          float undefined_value, const char* name, int gslib, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (write_property(L, filename, gslib, d_data, nullptr, d_mask, size, undefined_value, 0, nullptr, 0, name,
                       (hipStream_t)stream, err) != 0)
        return fail(L, err);
    return 0;
}

int hpglx_write_inc_byte(const char* filename, const unsigned char* d_data, const unsigned char* d_mask, int size,
                         int undefined_value, const char* name, const unsigned char* values, int values_count,
                         int gslib, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (write_property(L, filename, gslib, nullptr, d_data, d_mask, size, 0.0f, (unsigned char)undefined_value,
                       values, values_count, name, (hipStream_t)stream, err) != 0)
        return fail(L, err);
    return 0;
}

// ---------------------------------------------------- experimental variogram
void fill_ellipsoid_directions(ellipsoid_t* ell, double azimuth, double dip, double rotation) {
    double dir[3][3];
    vario_fill_directions(azimuth, dip, rotation, dir);
    vector_t* d[3] = {&ell->m_direction1, &ell->m_direction2, &ell->m_direction3};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d[i]->m_data[j] = dir[i][j];
}

int hpglx_calc_variograms(const variogram_search_template_t* templ, const float* d_data, const unsigned char* d_mask,
                          const int dims[3], int lag_count, int percent, float* d_result, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!templ || !d_data || !d_mask || !dims || !d_result) return fail(L, "null argument");
    if (lag_count <= 0) return fail(L, "calc_variograms: num_lags must be positive");
    if (!ensure_events(L, err)) return fail(L, err);
    hipStream_t st = (hipStream_t)stream;
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess) return fail(L, "hipStreamWaitEvent failed");
    if (vario_grid(vario_template(*templ), d_data, d_mask, dims, lag_count, percent, d_result, L.ws, st, err) != 0 ||
        hipEventRecord(L.done, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return fail(L, err.empty() ? "variogram failed" : err);
    return 0;
}

int hpglx_calc_variograms_points(const variogram_search_template_t* templ, const float* d_xyzv, int n, int lag_count,
                                 float* d_result, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!templ || (n > 0 && !d_xyzv) || !d_result) return fail(L, "null argument");
    if (lag_count <= 0) return fail(L, "calc_variograms_from_point_set: num_lags must be positive");
    if (!ensure_events(L, err)) return fail(L, err);
    hipStream_t st = (hipStream_t)stream;
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess) return fail(L, "hipStreamWaitEvent failed");
    if (vario_points(vario_template(*templ), d_xyzv, n, lag_count, d_result, L.ws, st, err) != 0 ||
        hipEventRecord(L.done, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return fail(L, err.empty() ? "variogram failed" : err);
    return 0;
}

void calc_variograms(variogram_search_template_t* templ, hard_data_t* data, float* result, int result_length,
                     int percentToUse) {
    Library& L = lib();
    std::string err;
    {
        std::lock_guard<std::mutex> g(L.mu);
        L.last_error.clear();
        const int lag_count = std::min(templ->m_num_lags, result_length);
        if (lag_count <= 0) return;
        for (int i = 0; i < 3; ++i)
            if (data->m_data_shape[i] != data->m_mask_shape[i]) {
                L.last_error = "calc_variograms: data and mask shapes differ";
                return;
            }
        const int* sh = data->m_data_shape;
        const long long n = (long long)sh[0] * sh[1] * sh[2];
        if (n <= 0) {
            for (int l = 0; l < lag_count; ++l) result[l] = 0.0f;
            return;
        }
        // the strided host arrays, gathered x fastest
        std::vector<float> hv((size_t)n);
        std::vector<unsigned char> hm((size_t)n);
        size_t k = 0;
        for (int z = 0; z < sh[2]; ++z)
            for (int y = 0; y < sh[1]; ++y)
                for (int x = 0; x < sh[0]; ++x, ++k) {
                    hv[k] = data->m_data[(long long)data->m_data_strides[0] * x +
                                         (long long)data->m_data_strides[1] * y +
                                         (long long)data->m_data_strides[2] * z];
                    hm[k] = data->m_mask[(long long)data->m_mask_strides[0] * x +
                                         (long long)data->m_mask_strides[1] * y +
                                         (long long)data->m_mask_strides[2] * z];
                }
        float* d_v = (float*)L.ws.get("vario_in_v", (size_t)n * 4, err);
        unsigned char* d_m = (unsigned char*)L.ws.get("vario_in_m", (size_t)n, err);
        float* d_r = (float*)L.ws.get("vario_out", (size_t)lag_count * 4, err);
        if (!err.empty() || hipMemcpy(d_v, hv.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_m, hm.data(), (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
            L.last_error = err.empty() ? "host to device copy failed" : err;
            return;
        }
        if (!ensure_events(L, err) ||
            vario_grid(vario_template(*templ), d_v, d_m, sh, lag_count, percentToUse, d_r, L.ws, nullptr, err) != 0 ||
            hipMemcpy(result, d_r, (size_t)lag_count * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            L.last_error = err.empty() ? "variogram failed" : err;
            return;
        }
    }
}

void calc_variograms_from_point_set(variogram_search_template_t* templ, cont_point_set_t* ps, float* result,
                                    int result_length) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    const int lag_count = std::min(templ->m_num_lags, result_length);
    if (lag_count <= 0) return;
    const int n = ps->size > 0 ? ps->size : 0;
    std::vector<float> pts((size_t)std::max(1, n) * 4);
    for (int i = 0; i < n; ++i) {
        pts[4 * (size_t)i + 0] = ps->xs[i];
        pts[4 * (size_t)i + 1] = ps->ys[i];
        pts[4 * (size_t)i + 2] = ps->zs[i];
        pts[4 * (size_t)i + 3] = ps->values[i];
    }
    float* d_p = (float*)L.ws.get("vario_pts", pts.size() * 4, err);
    float* d_r = (float*)L.ws.get("vario_out", (size_t)lag_count * 4, err);
    if (!err.empty() || hipMemcpy(d_p, pts.data(), pts.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        L.last_error = err.empty() ? "host to device copy failed" : err;
        return;
    }
    if (!ensure_events(L, err) ||
        vario_points(vario_template(*templ), d_p, n, lag_count, d_r, L.ws, nullptr, err) != 0 ||
        hipMemcpy(result, d_r, (size_t)lag_count * 4, hipMemcpyDeviceToHost) != hipSuccess)
        L.last_error = err.empty() ? "variogram failed" : err;
}

void hpglx_unpack_offset(int packed, int* dx, int* dy, int* dz) {
    unpack_off(packed, *dx, *dy, *dz);
}

int hpglx_template(const hpgl_cov_params_t* cov, const int radiuses[3], int* out_xyz, int cap) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    PlanHost h;
    std::string err;
    if (!build_plan(h, KIND_OK, to_cov(*cov), radiuses, 1, err)) return fail(L, err);
    const int n = (int)h.tdx.size();
    for (int i = 0; i < n && i < cap && out_xyz; ++i) {
        out_xyz[3 * i + 0] = h.tdx[i];
        out_xyz[3 * i + 1] = h.tdy[i];
        out_xyz[3 * i + 2] = h.tdz[i];
    }
    return n;
}

void hpglx_set_profiling(int on) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.profiling = on != 0;
    if (L.profiling) L.tcalls = 0;
}

int hpglx_last_timing(float* ms2) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    if (L.tcalls == 0) return fail(L, "no profiled call");
    float ms[2];
    if (!read_timing(L.tpool[(L.tcalls - 1) % kTimingPool], ms)) return fail(L, "reading the call's events failed");
    ms2[0] = ms[0];
    ms2[1] = ms[1];
    return 0;
}

int hpglx_profile_summary(float* ms2, int* ncalls) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    const size_t n = L.tcalls < kTimingPool ? L.tcalls : kTimingPool;
    if (n == 0) return fail(L, "no profiled call");
    double sum[2] = {0, 0};
    for (size_t i = 0; i < n; ++i) {
        float ms[2];
        if (!read_timing(L.tpool[i], ms)) return fail(L, "reading the call's events failed");
        for (int k = 0; k < 2; ++k) sum[k] += ms[k];
    }
    for (int k = 0; k < 2; ++k) ms2[k] = (float)(sum[k] / (double)n);
    if (ncalls) *ncalls = (int)n;
    return 0;
}

int hpglx_last_call_stats(unsigned int* fallbacks, int* cluster_cells) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    int h[4] = {0, 0, 0, 0};
    if (L.done && hipEventSynchronize(L.done) != hipSuccess) return fail(L, "hipEventSynchronize failed");
    if (L.last_counters && hipMemcpy(h, L.last_counters, sizeof h, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(L, "copy failed");
    if (fallbacks) *fallbacks = (unsigned)h[3];
    if (cluster_cells) *cluster_cells = h[0];
    return 0;
}

void hpglx_set_graph_mode(int on) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.graph_mode = on ? 1 : 0;
}

void hpglx_set_default_precision(int precision) {
    lib().default_precision = precision == PREC_MIXED ? PREC_MIXED : PREC_FP64;
}

void hpglx_host_introsort_perm(const double* keys, int n, int* perm) {
    hpglx_host_introsort_first(keys, n, n, perm);
}

void hpglx_host_introsort_first(const double* keys, int n, int kneed, int* perm) {
    struct HostStore {
        struct Item {
            double key;
            int off;
        };
        std::vector<Item> v;
        double key(int i) const { return v[i].key; }
        Item get(int i) const { return v[i]; }
        void put(int i, const Item& it) { v[i] = it; }
        void swap(int i, int j) { std::swap(v[i], v[j]); }
    } st;
    st.v.resize(n > 0 ? n : 0);
    for (int i = 0; i < n; ++i) st.v[i] = HostStore::Item{keys[i], i};
    Introsort<HostStore> s(st);
    s.sort(n, kneed);
    for (int i = 0; i < n; ++i) perm[i] = st.v[i].off;
}

const char* hpglx_last_error(void) {
    return lib().last_error.c_str();
}

// ------------------------------------------------ leave-one-out kriging (loo.hip)
int hpglx_knn(const float* xyz, int n, int k, int* nbr, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!xyz || !nbr) return fail(L, "knn: coordinates and output are required");
    if (knn(xyz, n, k, nbr, (hipStream_t)stream, err) != 0) return fail(L, err);
    return 0;
}

size_t hpglx_loo_state_size(int n, int k, int nsets) {
    if (n < 2 || k < 1 || k > 32 || nsets < 1) return 0;
    return loo_state_doubles(n, k, nsets) * sizeof(double);
}

int hpglx_loo_forward(const float* xyz, const float* values, int n, const int* nbr, int k, const double* params,
                      int nsets, int model, double* est, int* status, void* state, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    if (!xyz || !values || !nbr || !params || !est) return fail(L, "loo: inputs and est are required");
    LooArgs a;
    std::memset(&a, 0, sizeof a);
    a.n = n;
    a.k = k;
    a.np = nsets;
    a.model = model;
    a.xyz = xyz;
    a.val = values;
    a.nbr = nbr;
    a.prm = params;
    a.est = est;
    a.status = status;
    a.state = (double*)state;
    std::string err;
    if (loo_forward(a, (hipStream_t)stream, err) != 0) return fail(L, err);
    return 0;
}

int hpglx_loo_backward(const float* xyz, const float* values, int n, const int* nbr, int k, const double* params,
                       int nsets, int model, const void* state, const double* grad_est, double* grad_params,
                       double* jac, void* stream) {
    Library& L = lib();
    std::lock_guard<std::mutex> g(L.mu);
    L.last_error.clear();
    std::string err;
    if (!xyz || !values || !nbr || !params || !state) return fail(L, "loo backward: inputs and state are required");
    if (!ensure_events(L, err)) return fail(L, err);
    hipStream_t st = (hipStream_t)stream;
    // the partial-sum buffer is shared by calls: order after the previous one
    if (hipStreamWaitEvent(st, L.done, 0) != hipSuccess) return fail(L, "hipStreamWaitEvent failed");
    double* part = (double*)L.ws.get("loo_part", loo_part_doubles(n, k, nsets) * sizeof(double), err);
    if (!part) return fail(L, err);
    LooArgs a;
    std::memset(&a, 0, sizeof a);
    a.n = n;
    a.k = k;
    a.np = nsets;
    a.model = model;
    a.xyz = xyz;
    a.val = values;
    a.nbr = nbr;
    a.prm = params;
    a.state = (double*)state;
    a.gest = grad_est;
    a.jac = jac;
    a.part = part;
    if (loo_backward(a, grad_params, st, err) != 0) return fail(L, err);
    if (hipEventRecord(L.done, st) != hipSuccess) return fail(L, "hipEventRecord failed");
    return 0;
}

const char* hpglx_version(void) {
    return "hpgl-mi355x 0.1 (gfx950)";
}

}  // extern "C"
