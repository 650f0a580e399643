// C ABI of the gfx950 hot path: device context, layout conversion and the
// per-batch launch sequences.  See include/grotor_amd.h for the reference
// entry point each function replaces.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grotor_amd.h"
#include "gr_kernels.hpp"
#include "gr_status.hpp"

using namespace grotor;

namespace {

void hip_check(hipError_t status, const char* op) {
    if (status != hipSuccess) throw GrError(GR_E_HIP, std::string(op) + ": " + hipGetErrorString(status));
}

struct DeviceArena {
    std::vector<void*> blocks;
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
        blocks.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T>
    void release(T*& p) {
        if (!p) return;
        auto it = std::find(blocks.begin(), blocks.end(), (void*)p);
        if (it != blocks.end()) blocks.erase(it);
        (void)hipFree(p);
        p = nullptr;
    }
    ~DeviceArena() {
        for (void* p : blocks) (void)hipFree(p);
    }
};

// Sequential aligned sub-buffers of a base pointer (sizes only when base is null).
struct Carver {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += std::max<size_t>(count, 1) * sizeof(T);
        return p;
    }
};

// Grow-only device buffer, carved into aligned sub-buffers by a layout pass
// that runs twice: once without a base to size it, once to assign pointers.
struct Workspace {
    char* base = nullptr;
    size_t cap = 0;
    Workspace() = default;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() {
        if (base) (void)hipFree(base);
    }
    template <class Layout>
    void carve(Layout&& layout) {
        Carver dry{nullptr};
        layout(dry);
        if (dry.off > cap) {
            if (base) hip_check(hipFree(base), "hipFree workspace");
            base = nullptr;
            cap = 0;
            const size_t want = dry.off + dry.off / 4;
            hip_check(hipMalloc(reinterpret_cast<void**>(&base), want), "hipMalloc workspace");
            cap = want;
        }
        Carver real{base};
        layout(real);
    }
};

// Scratch of the allocation-free DP entry point (gr_lshape_batch_noalloc):
// the persistent DP's work-queue counters, then the global DP tables of trees
// too large for LDS.  One copy per GPU, held until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (the array's name is a host-side stand-in).
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocQueueBytes = size_t(kDpQueues) * kDpQueueStride * sizeof(int);
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
double* noalloc_scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw GrError(GR_E_HIP, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

struct gr_ctx {
    DeviceArena mem;
    hipStream_t stream = nullptr;
    // second stream (default priority) for the large-tree DP launch of a
    // batch, which runs beside the small-tree launch; fork/join events
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // third stream (default priority) for the persistent form's smallest
    // trees, beside the other two classes
    hipStream_t side2 = nullptr;
    hipEvent_t join2_ev = nullptr;
    int cus = 0;
    DeviceGrid g{};
    int tracks = 0, maxlen = 0;
    size_t cells = 0;
    float *capacity = nullptr, *demand = nullptr, *vcost = nullptr, *wcost = nullptr;
    double* presum = nullptr;
    int* timestamp = nullptr;
    unsigned* track_dirty = nullptr;
    Stats* stats = nullptr;
    int* error = nullptr;
    CostParams cp;
    // route storage
    int route_per_pin = kRoutePerPin;
    int next_route_per_pin = kRoutePerPin;  // gr_ctx_set_route_stride, applied by gr_ctx_set_routes
    int net_count = 0;
    std::vector<int> pin_acc_num_h;
    int* pin_acc_num = nullptr;
    int* routes = nullptr;
    long long route_words = 0;
    // L-shape stage topology
    int stage_count = 0;
    std::vector<int> node_off_h;
    int *net_ids = nullptr, *node_off = nullptr, *nodes = nullptr, *parents = nullptr;
    int *lvl = nullptr, *link = nullptr;  // L-shape DP level order and child lists, per stage node
    int* dp_queue = nullptr;              // persistent DP: kDpQueues work-queue counters per batch
    int dp_queue_cap = 0;
    std::vector<int> height_h;            // per stage slot: height of the tree's root
    std::vector<int> net_ids_h;
    std::vector<int> slots_h;      // stage slot per execution position
    int4* desc = nullptr;          // per execution position: node begin, nodes, route word, route capacity
    std::vector<int4> desc_h;
    double* gscratch = nullptr;
    size_t gscratch_bytes = 0;
    double* noalloc_scratch = nullptr;    // this device's `scratch` (gr_lshape_batch_noalloc)
    std::vector<int> noalloc_slots;       // sized at gr_lshape_upload
    // rip-up
    int* of_sum = nullptr;
    unsigned char* of_flags = nullptr;
    // batching: resident geometry (gr_batch_geometry) and grow-only scratch
    Workspace geom_ws, batch_ws1, batch_ws2, pm_ws;
    BatchGeom geom{};
    bool has_geom = false;
    long long geom_points = 0, geom_hsegs = 0, geom_vsegs = 0;
    std::vector<int> geom_pt_off;  // host copy of the point offsets
    // soft DP: cost arrays, per-cell gradients, checkpointed tables
    Workspace soft_ws, soft_tables;
    SoftCostArgs soft_cost{};
    SoftArgs soft{};
    int soft_precision = 0;
    int soft_kind = 0;        // 0: L-shape edges, 1: the Z-shape route family
    bool soft_given = false;  // the last forward ran on caller-given costs
    bool soft_ready = false;  // a forward pass whose tables are still held
    double* soft_params = nullptr;
    // Z-shape stage: plan on the device, shared DP map, raw words
    Workspace z_ws;
    int* z_map = nullptr;   // 6 planes
    int* z_prev = nullptr;  // 16 planes
    long long z_plane = 0;
    int z_batches = 0;
    std::vector<int> z_batch_off;
    std::vector<long long> z_rec_off;
    int *z_net_ids = nullptr, *z_records = nullptr, *z_raw_cap = nullptr, *z_raw = nullptr;
    // FLT stage: phase-two route words, resident DAG records, batch tables
    int route2_per_pin = 100;
    int* pin_off2 = nullptr;
    int* routes2 = nullptr;
    long long route2_words = 0;
    std::vector<int> pin_off2_h;
    // nets whose current route is their phase-two words (an FLT stage has
    // routed them): rip-up reads those words, a later stage removes them
    unsigned char* live2 = nullptr;
    std::vector<unsigned char> live2_h;
    Workspace f_ws, f_tables;
    FltArgs f{};
    int f_batches = 0;
    std::vector<int> f_batch_off;
    std::vector<int> f_ids_h;  // subnet ids of the uploaded FLT stage, execution order
    std::vector<int4> f_base;  // per batch: stage-global first node, edge, slot; edges
    std::vector<FltLdsShape> f_batch_max;  // per batch: the largest child slots, nodes, edges, depth width, depths
    // per batch: DAGs whose LDS share exceeds f_lds_split bytes first (f_large
    // of them), then the rest, each class with its own largest shape
    std::vector<int> f_large;
    std::vector<int> f_far_off;  // per batch: offsets of its edges longer than 4 in the stage's far-edge list
    std::vector<FltLdsShape> f_class_max;  // [batch * kFltClasses + class], class 0 the large DAGs
    int f_lds_split = 20 * 1024;
    int* f_queue = nullptr;         // persistent FLT DP: kFltClasses x kFltQueues counters per batch
    char* f_stage = nullptr;        // pinned host staging of the FLT stage records
    size_t f_stage_cap = 0;
    int f_queue_cap = 0;
    int lds_limit = 65536;          // LDS bytes a workgroup may use
    std::vector<double> f_edge_bytes, f_route_bytes;  // algorithmic bytes per batch launch
    int f_per_wave[2] = {1, 1};  // FLT DAGs per wavefront: batches of fewer / at least kFltWideBatch DAGs
    // congestion view (2D)
    unsigned char* cview = nullptr;
    float *cxsum = nullptr, *cysum = nullptr;
    int per_wave[kTreeClasses] = {1, 1};  // L-shape DP nets per wavefront by tree size class (gr_lshape_tune)
    int tiny_bound = 4;  // persistent form: trees of at most this many nodes form its first group (gr_lshape_tune_groups)
    int tiny_per_wave = 0;  // nets per wavefront of that group, 0: the small class's (gr_lshape_tune_tiny)
    int batch_mask_words = 2;  // prefix-mask words of the last batching call
    int color_waves_per_cu = 32;
    // kernel profiling (HIP events on the context stream)
    struct ProfRec {
        int kind;
        hipEvent_t a, b;
        double bytes;
    };
    bool profiling = false;
    unsigned prof_mask = ~0u;  // kernel classes that get events
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> events;
    size_t next_event = 0;
    unsigned long long* refreshed = nullptr;
    std::vector<double> slot_dp_bytes;  // algorithmic bytes of the DP per stage slot
    Stats prof_stats0{};
    unsigned long long prof_wire_edges = 0, prof_vias = 0;  // counts of profiled steps before a reset
    unsigned long long refreshed0 = 0;

    // batch loop captured as a HIP graph (gr_lshape_run mode 1), reused
    // while the stage topology, batch offsets and first stamp are unchanged
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // the last synchronized batch-loop interval (L-shape, FLT or Z-shape run)
    std::chrono::steady_clock::time_point last_t0{}, last_t1{};
    std::vector<int> graph_key;
    int stage_gen = 0;
    // FLT batch loop captured as a HIP graph (gr_flt_run mode 1), reused
    // while the uploaded stage, its storage, first stamp and settings stay
    hipGraph_t fgraph = nullptr;
    hipGraphExec_t fgraph_exec = nullptr;
    std::vector<int> fgraph_key;
    int flt_gen = 0;

    void drop_fgraph() {
        if (fgraph_exec) (void)hipGraphExecDestroy(fgraph_exec);
        if (fgraph) (void)hipGraphDestroy(fgraph);
        fgraph_exec = nullptr;
        fgraph = nullptr;
        fgraph_key.clear();
    }
    // both captured loops: every buffer the L-shape loop holds is one the
    // FLT loop may hold too
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_key.clear();
        drop_fgraph();
    }
    void release_phase2() {
        drop_fgraph();  // a captured FLT loop holds the phase-two addresses
        mem.release(routes2);
        mem.release(pin_off2);
        mem.release(live2);
        pin_off2_h.clear();
        live2_h.clear();
        route2_words = 0;
    }

    ~gr_ctx() {
        drop_graph();
        if (f_stage) (void)hipHostFree(f_stage);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (side) (void)hipStreamDestroy(side);
        if (join2_ev) (void)hipEventDestroy(join2_ev);
        if (side2) (void)hipStreamDestroy(side2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr int kSmallTree = 16;
constexpr int kMaxTreeNodes = 32767;  // node indices and heights are packed in 16 bits

hipEvent_t take_event(gr_ctx* c) {
    if (c->next_event == c->events.size()) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        c->events.push_back(e);
    }
    return c->events[c->next_event++];
}

// Run `launch` between two events on `stream` (the context stream unless
// given) when profiling.
template <class F>
void timed(gr_ctx* c, int kind, double bytes, F&& launch, hipStream_t stream = nullptr) {
    if (!c->profiling || !(c->prof_mask >> kind & 1u)) {
        launch();
        return;
    }
    if (!stream) stream = c->stream;
    hipEvent_t a = take_event(c), b = take_event(c);
    hip_check(hipEventRecord(a, stream), "hipEventRecord");
    launch();
    hip_check(hipEventRecord(b, stream), "hipEventRecord");
    c->prof.push_back({kind, a, b, bytes});
}

void upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}
void download(void* dst, const void* src, size_t bytes, hipStream_t s) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// IDX order <-> track-major order (host side, one pass)
template <class T>
void to_track_major(const DeviceGrid& g, const T* src, std::vector<T>& dst) {
    dst.resize(size_t(g.L) * g.XY);
    for (int l = 0; l < g.L; ++l)
        for (int x = 0; x < g.X; ++x)
            for (int y = 0; y < g.Y; ++y) dst[cell_of(g, l, x, y)] = src[(size_t(l) * g.X + x) * g.Y + y];
}
template <class T>
void from_track_major(const DeviceGrid& g, const std::vector<T>& src, T* dst) {
    for (int l = 0; l < g.L; ++l)
        for (int x = 0; x < g.X; ++x)
            for (int y = 0; y < g.Y; ++y) dst[(size_t(l) * g.X + x) * g.Y + y] = src[cell_of(g, l, x, y)];
}

Stats read_stats(gr_ctx* c) {
    Stats part[kStatStripes];
    download(part, c->stats, sizeof(part), c->stream);
    Stats s{};
    for (const Stats& p : part) {
        s.total_wirelength += p.total_wirelength;
        s.total_vias += p.total_vias;
        s.wire_edges += p.wire_edges;
    }
    return s;
}

unsigned long long read_refreshed(gr_ctx* c) {
    std::vector<unsigned long long> part(size_t(kStatStripes) * kCounterStride);
    download(part.data(), c->refreshed, sizeof(unsigned long long) * part.size(), c->stream);
    unsigned long long sum = 0;
    for (int i = 0; i < kStatStripes; ++i) sum += part[size_t(i) * kCounterStride];
    return sum;
}

void check_device_error(gr_ctx* c, const char* op) {
    int err = 0;
    download(&err, c->error, sizeof(int), c->stream);
    if (!err) return;
    const int zero = 0;
    upload(c->error, &zero, sizeof(int), c->stream);
    switch (err) {
        case kDevRouteOverflow: throw GrError(GR_E_CAPACITY, std::string(op) + ": route storage exceeded");
        case kDevNoSolution: throw GrError(GR_E_DEVICE, std::string(op) + ": no finite layer assignment");
        case kDevSpinLimit: throw GrError(GR_E_DEVICE, std::string(op) + ": batching wait exceeded its spin limit");
        case kDevColorOverflow: throw GrError(GR_E_CAPACITY, std::string(op) + ": more than 2016 batches");
        case kDevBadCoordinate: throw GrError(GR_E_DEVICE, std::string(op) + ": route coordinate outside the grid");
        default: throw GrError(GR_E_DEVICE, std::string(op) + ": device error " + std::to_string(err));
    }
}

void launch_refresh(gr_ctx* c, int force) {
    timed(c, GR_K_REFRESH, 0.0, [&] {
        hipLaunchKernelGGL(refresh_tracks_kernel, dim3(c->tracks), dim3(kRefreshThreads),
                           size_t(c->maxlen) * sizeof(double), c->stream, c->g, c->capacity, c->demand, c->vcost,
                           c->presum, (float*)nullptr, c->track_dirty, force, c->cp, c->refreshed);
    });
    hip_check(hipGetLastError(), "refresh_tracks_kernel");
}

CommitArgs commit_args(gr_ctx* c, int begin, int count, int stamp, int K) {
    CommitArgs a;
    a.g = c->g;
    a.demand = c->demand;
    a.timestamp = c->timestamp;
    a.track_dirty = c->track_dirty;
    a.stats = c->stats;
    a.rs = {c->routes, c->pin_acc_num, c->route_per_pin};
    a.net_ids = c->net_ids;
    a.slots = nullptr;
    a.begin = begin;
    a.count = count;
    a.stamp = stamp;
    a.K = K;
    a.rs2 = {c->routes2, c->pin_off2, c->route2_per_pin};
    a.live2 = nullptr;
    return a;
}

constexpr int kCommitPackedNets = 4096;  // batches of at least this many nets commit four nets per wavefront

// Commit (K = 1) or removal (K = -1) of stage slots [begin, begin + count):
// routes from `rs` (phase-one storage when null), or, with `current`, each
// net's current route (phase-two words for nets marked live).
void launch_commit(gr_ctx* c, int begin, int count, int stamp, int K, const int* net_ids = nullptr,
                   const RouteStore* rs = nullptr, bool current = false) {
    if (!count) return;
    CommitArgs a = commit_args(c, begin, count, stamp, K);
    if (net_ids) a.net_ids = net_ids;
    if (rs) a.rs = *rs;
    if (current) a.live2 = c->live2;
    // many nets (a large batch): four per wavefront, so the batch's nets are
    // resident at once; otherwise a wavefront per net
    if (count >= kCommitPackedNets) {
        const int blocks = std::min((count + 15) / 16, 4096);
        timed(c, GR_K_COMMIT_WIRES, 0.0,
              [&] { hipLaunchKernelGGL(commit_wires_kernel<16>, dim3(blocks), dim3(256), 0, c->stream, a); });
        timed(c, GR_K_COMMIT_VIAS, 0.0,
              [&] { hipLaunchKernelGGL(commit_vias_kernel<16>, dim3(blocks), dim3(256), 0, c->stream, a); });
    } else {
        const int blocks = std::min((count + 3) / 4, 4096);
        timed(c, GR_K_COMMIT_WIRES, 0.0,
              [&] { hipLaunchKernelGGL(commit_wires_kernel<64>, dim3(blocks), dim3(256), 0, c->stream, a); });
        timed(c, GR_K_COMMIT_VIAS, 0.0,
              [&] { hipLaunchKernelGGL(commit_vias_kernel<64>, dim3(blocks), dim3(256), 0, c->stream, a); });
    }
    hip_check(hipGetLastError(), "commit kernels");
}

// Per batch: stage slots grouped by tree size class (up to kSmallTree nodes,
// larger), in stage order inside a class, and the node capacity each class
// needs.  Each class is one DP launch with its own nets per wavefront.
// (Finer classes were measured slower: every extra launch per batch costs
// more in launch and tail time than the smaller tables save.)
constexpr int kClassNodes[kTreeClasses - 1] = {kSmallTree};

struct BatchLaunch {
    int begin;
    int count[kTreeClasses];
    int nmax[kTreeClasses];
    // plan_batch: the last `tiny` slots of class 0 are its trees of at most
    // tiny_bound nodes (the persistent form drains them from a queue of their
    // own, with tables sized for them)
    int tiny = 0;
    int tiny_nmax = 1;
};

int tree_class(int n) {
    int k = 0;
    while (k < kTreeClasses - 1 && n > kClassNodes[k]) ++k;
    return k;
}

BatchLaunch plan_batch(gr_ctx* c, int begin, int count, std::vector<int>& slots) {
    BatchLaunch b{begin, {0, 0}, {1, 1}};
    for (int s = begin; s < begin + count; ++s) {
        const int n = c->node_off_h[s + 1] - c->node_off_h[s], k = tree_class(n);
        ++b.count[k];
        b.nmax[k] = std::max(b.nmax[k], n);
    }
    int at[kTreeClasses];
    at[0] = begin;
    for (int k = 1; k < kTreeClasses; ++k) at[k] = at[k - 1] + b.count[k - 1];
    for (int s = begin; s < begin + count; ++s) slots[at[tree_class(c->node_off_h[s + 1] - c->node_off_h[s])]++] = s;
    // tallest trees first inside a class: the nets with the longest level
    // chains start first and the launch's tail is short; class 0's trees of
    // at most c->tiny_bound nodes after its others
    // (slots enter in ascending order, so ordering ties by slot is the
    // stable order, without the temporary buffer of a stable sort)
    auto tiny = [&](int u) { return c->node_off_h[u + 1] - c->node_off_h[u] <= c->tiny_bound; };
    for (int k = 0, lo = begin; k < kTreeClasses; lo += b.count[k++])
        std::sort(slots.begin() + lo, slots.begin() + lo + b.count[k], [&](int u, int v) {
            if (k == 0 && tiny(u) != tiny(v)) return tiny(v);
            return c->height_h[u] != c->height_h[v] ? c->height_h[u] > c->height_h[v] : u < v;
        });
    for (int j = begin; j < begin + b.count[0]; ++j)
        if (tiny(slots[j])) {
            ++b.tiny;
            b.tiny_nmax = std::max(b.tiny_nmax, c->node_off_h[slots[j] + 1] - c->node_off_h[slots[j]]);
        }
    return b;
}

// f(k) for k in [0, n) on up to 16 host threads (chunks of 256 items);
// the first exception thrown by any item is rethrown once all have stopped.
template <class F>
void parallel_items(int n, F&& f) {
    const int threads = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
    if (n < 4096 || threads == 1) {
        for (int k = 0; k < n; ++k) f(k);
        return;
    }
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr error;
    std::mutex m;
    auto work = [&] {
        try {
            while (!failed.load(std::memory_order_relaxed)) {
                const int k0 = next.fetch_add(256);
                if (k0 >= n) return;
                for (int k = k0; k < std::min(n, k0 + 256); ++k) f(k);
            }
        } catch (...) {
            std::lock_guard<std::mutex> lock(m);
            if (!error) error = std::current_exception();
            failed = true;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (error) std::rethrow_exception(error);
}

// f(0) .. f(n - 1) for a few heavy items, one at a time per thread.
template <class F>
void parallel_tasks(int n, F&& f) {
    const int threads = std::max(1, std::min<int>({16, n, int(std::thread::hardware_concurrency())}));
    if (threads <= 1) {
        for (int k = 0; k < n; ++k) f(k);
        return;
    }
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr error;
    std::mutex m;
    auto work = [&] {
        try {
            for (int k = next++; k < n && !failed.load(std::memory_order_relaxed); k = next++) f(k);
        } catch (...) {
            std::lock_guard<std::mutex> lock(m);
            if (!error) error = std::current_exception();
            failed = true;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (error) std::rethrow_exception(error);
}

constexpr size_t kLdsLimit = 64 * 1024;

void ensure_gscratch(gr_ctx* c, size_t bytes) {
    if (bytes <= c->gscratch_bytes) return;
    c->drop_graph();  // a captured batch loop holds the old scratch address
    c->mem.release(c->gscratch);
    c->gscratch = c->mem.alloc<double>((bytes + 7) / 8);
    c->gscratch_bytes = bytes;
}

// Nets per wavefront for a class launch: the tuned value, halved until the
// workgroup's tables fit in LDS; 0 when one net's tables do not fit (the
// global-scratch kernel then runs with one net per wavefront).
int per_wave_for(gr_ctx* c, int k, int nmax) {
    const size_t bytes = (lshape_scratch_bytes(c->g.L, nmax) + 7) / 8 * 8;
    if (bytes > kLdsLimit) return 0;
    int p = c->per_wave[k];
    while (p > 1 && bytes * p > kLdsLimit) p /= 2;
    return p;
}

// Global DP tables of the trees that do not fit in LDS: the context's
// grow-only scratch, or a caller-given buffer of `bytes`.
struct DpTables {
    double* base = nullptr;
    size_t bytes = 0;
};

void launch_lshape_class(gr_ctx* c, const BatchLaunch& b, int k, int per_wave_override = 0,
                         hipStream_t stream = nullptr, DpTables tables = {}) {
    if (!b.count[k]) return;
    if (!stream) stream = c->stream;
    LshapeArgs a;
    a.g = c->g;
    a.vcost = c->vcost;
    a.presum = c->presum;
    a.rs = {c->routes, c->pin_acc_num, c->route_per_pin};
    a.desc = c->desc;
    a.nodes = c->nodes;
    a.parents = c->parents;
    a.lvl = c->lvl;
    a.link = c->link;
    if (!tables.base) tables = {c->gscratch, c->gscratch_bytes};
    a.gscratch = tables.base;
    a.error = c->error;
    a.queue = nullptr;
    a.cus = c->cus;
    a.begin = b.begin;
    for (int j = 0; j < k; ++j) a.begin += b.count[j];
    a.count = b.count[k];
    a.nmax = b.nmax[k];
    double alg = 0;
    if (c->profiling)
        for (int i = a.begin; i < a.begin + a.count; ++i) alg += c->slot_dp_bytes[c->slots_h[i]];
    const int p = per_wave_override ? per_wave_override : per_wave_for(c, k, a.nmax);
    if (p > 0) {
        a.per_wave = p;
        timed(c, GR_K_LSHAPE, alg, [&] { launch_lshape_dp(a, (a.count + p - 1) / p, true, stream); }, stream);
    } else {
        const size_t bytes = lshape_scratch_bytes(c->g.L, a.nmax);
        if (((bytes + 7) / 8) * 8 * size_t(a.count) > tables.bytes) throw GrError(GR_E_STATE, "DP scratch not reserved");
        a.per_wave = 1;
        timed(c, GR_K_LSHAPE, alg, [&] { launch_lshape_dp(a, a.count, false, stream); }, stream);
    }
    hip_check(hipGetLastError(), "lshape_dp kernels");
}

// The two size classes of a batch are independent: the large trees (the
// batch's critical path) run on the side stream beside the small ones.
void launch_lshape(gr_ctx* c, const BatchLaunch& b, DpTables tables = {}) {
    if (!b.count[0] || !b.count[1] || c->profiling) {  // profiled runs time each launch alone
        for (int k = 0; k < kTreeClasses; ++k) launch_lshape_class(c, b, k, 0, nullptr, tables);
        return;
    }
    hip_check(hipEventRecord(c->fork_ev, c->stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->side, c->fork_ev, 0), "hipStreamWaitEvent");
    launch_lshape_class(c, b, 1, 0, c->side, tables);
    launch_lshape_class(c, b, 0, 0, c->stream, tables);
    hip_check(hipEventRecord(c->join_ev, c->side), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->stream, c->join_ev, 0), "hipStreamWaitEvent");
}

// Persistent form of one batch: its slots in one work queue, tallest trees
// first; BatchLaunch class 0 carries the whole batch (its node capacity is
// the batch's largest tree) and class 1 is empty.
BatchLaunch plan_batch_queue(gr_ctx* c, int begin, int count, std::vector<int>& slots) {
    BatchLaunch b{begin, {count, 0}, {1, 1}};
    for (int s = begin; s < begin + count; ++s) {
        slots[s] = s;
        b.nmax[0] = std::max(b.nmax[0], c->node_off_h[s + 1] - c->node_off_h[s]);
    }
    std::sort(slots.begin() + begin, slots.begin() + begin + count, [&](int u, int v) {
        return c->height_h[u] != c->height_h[v] ? c->height_h[u] > c->height_h[v] : u < v;
    });
    return b;
}

// Plain form of one batch (the reference's execution: one net per
// wavefront in batch order, every net's DP tables in global memory, one
// launch): class 0 carries the batch, its node capacity the largest tree.
BatchLaunch plan_batch_plain(gr_ctx* c, int begin, int count, std::vector<int>& slots) {
    BatchLaunch b{begin, {count, 0}, {1, 1}};
    for (int s = begin; s < begin + count; ++s) {
        slots[s] = s;
        b.nmax[0] = std::max(b.nmax[0], c->node_off_h[s + 1] - c->node_off_h[s]);
    }
    return b;
}

size_t plain_tables_needed(const gr_ctx* c, const BatchLaunch& b) {
    return b.count[0] ? ((lshape_scratch_bytes(c->g.L, b.nmax[0]) + 7) / 8) * 8 * size_t(b.count[0]) : 0;
}

void launch_lshape_plain(gr_ctx* c, const BatchLaunch& b) {
    if (!b.count[0]) return;
    LshapeArgs a;
    a.g = c->g;
    a.vcost = c->vcost;
    a.presum = c->presum;
    a.rs = {c->routes, c->pin_acc_num, c->route_per_pin};
    a.desc = c->desc;
    a.nodes = c->nodes;
    a.parents = c->parents;
    a.lvl = c->lvl;
    a.link = c->link;
    a.gscratch = c->gscratch;
    a.error = c->error;
    a.queue = nullptr;
    a.cus = c->cus;
    a.begin = b.begin;
    a.count = b.count[0];
    a.nmax = b.nmax[0];
    a.per_wave = 1;
    if (plain_tables_needed(c, b) > c->gscratch_bytes) throw GrError(GR_E_STATE, "DP scratch not reserved");
    double alg = 0;
    if (c->profiling)
        for (int i = a.begin; i < a.begin + a.count; ++i) alg += c->slot_dp_bytes[c->slots_h[i]];
    timed(c, GR_K_LSHAPE, alg, [&] { launch_lshape_dp(a, a.count, false, c->stream); });
    hip_check(hipGetLastError(), "lshape_dp plain kernel");
}

// Slots [begin, begin + count) of class k (trees of at most nmax nodes)
// drained by the persistent DP kernel from its queue counters (zero
// beforehand), on `stream`; trees too large for LDS take the launch form.
void launch_lshape_queue_range(gr_ctx* c, int begin, int count, int nmax, int k, int* queue, hipStream_t stream,
                               int per_wave = 0) {
    if (!count) return;
    int p = per_wave_for(c, k, nmax);
    if (p > 0 && per_wave > 0) {  // an override, halved until the workgroup's tables fit in LDS
        const size_t bytes = (lshape_scratch_bytes(c->g.L, nmax) + 7) / 8 * 8;
        for (p = per_wave; p > 1 && bytes * p > kLdsLimit;) p /= 2;
    }
    if (p <= 0) {  // tables beyond LDS: the global-table launch over the range
        const BatchLaunch one{begin, {count, 0}, {nmax, 1}};
        launch_lshape_class(c, one, 0, 0, stream);
        return;
    }
    LshapeArgs a;
    a.g = c->g;
    a.vcost = c->vcost;
    a.presum = c->presum;
    a.rs = {c->routes, c->pin_acc_num, c->route_per_pin};
    a.desc = c->desc;
    a.nodes = c->nodes;
    a.parents = c->parents;
    a.lvl = c->lvl;
    a.link = c->link;
    a.gscratch = c->gscratch;
    a.error = c->error;
    a.queue = queue;
    a.cus = c->cus;
    a.begin = begin;
    a.count = count;
    a.nmax = nmax;
    a.per_wave = p;
    double alg = 0;
    if (c->profiling)
        for (int i = a.begin; i < a.begin + a.count; ++i) alg += c->slot_dp_bytes[c->slots_h[i]];
    timed(c, GR_K_LSHAPE, alg, [&] { launch_lshape_dp(a, (a.count + p - 1) / p, true, stream); }, stream);
    hip_check(hipGetLastError(), "lshape_dp queue kernel");
}

// Class 0 of the batch (or all of it) drained from one queue on the main
// stream: the allocation-free entry point's form.
void launch_lshape_queue(gr_ctx* c, const BatchLaunch& b, int* queue) {
    launch_lshape_queue_range(c, b.begin, b.count[0], b.nmax[0], 0, queue, c->stream);
}

constexpr int kPersistentGroups = 3;  // queues per batch in the persistent form

// Persistent form of one batch (plan_batch), in three groups with a queue
// each (kDpQueues counters at queue + g * kDpQueues * kDpQueueStride): the
// trees of at most tiny_bound nodes (most nets), the rest of the small
// class, the large trees.  Each group's tables are sized for its own
// largest tree, so the many tiny trees keep a small LDS footprint and a
// high occupancy; the groups run on three streams, side by side.  The tiny
// group takes its own nets per wavefront (tiny_per_wave), the other two one
// net per wavefront whatever the launch form was tuned to: their trees
// are solved fastest with a whole wavefront each.
void launch_lshape_persistent(gr_ctx* c, const BatchLaunch& b, int* queue) {
    const size_t qs = size_t(kDpQueues) * kDpQueueStride;
    const int mid = b.count[0] - b.tiny;
    struct Group {
        int begin, count, nmax, k;
    } grp[kPersistentGroups] = {{b.begin + mid, b.tiny, b.tiny_nmax, 0},
                                {b.begin, mid, b.nmax[0], 0},
                                {b.begin + b.count[0], b.count[1], b.nmax[1], 1}};
    int live = 0;
    for (const Group& g : grp) live += g.count > 0;
    // profiled runs time each launch alone: every group on the main stream
    if (live < 2 || c->profiling) {
        for (int j = 0; j < kPersistentGroups; ++j)
            launch_lshape_queue_range(c, grp[j].begin, grp[j].count, grp[j].nmax, grp[j].k, queue + j * qs, c->stream,
                                      j == 0 ? c->tiny_per_wave : 1);
        return;
    }
    hip_check(hipEventRecord(c->fork_ev, c->stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->side, c->fork_ev, 0), "hipStreamWaitEvent");
    hip_check(hipStreamWaitEvent(c->side2, c->fork_ev, 0), "hipStreamWaitEvent");
    launch_lshape_queue_range(c, grp[2].begin, grp[2].count, grp[2].nmax, 1, queue + 2 * qs, c->side, 1);
    launch_lshape_queue_range(c, grp[1].begin, grp[1].count, grp[1].nmax, 0, queue + qs, c->side2, 1);
    launch_lshape_queue_range(c, grp[0].begin, grp[0].count, grp[0].nmax, 0, queue, c->stream, c->tiny_per_wave);
    hip_check(hipEventRecord(c->join_ev, c->side), "hipEventRecord");
    hip_check(hipEventRecord(c->join2_ev, c->side2), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->stream, c->join_ev, 0), "hipStreamWaitEvent");
    hip_check(hipStreamWaitEvent(c->stream, c->join2_ev, 0), "hipStreamWaitEvent");
}

// Reserve the global DP scratch the batches will need (trees whose tables
// do not fit in LDS) before anything is launched or captured.
size_t dp_tables_needed(const gr_ctx* c, const BatchLaunch& b) {
    size_t need = 0;
    for (int k = 0; k < kTreeClasses; ++k) {
        if (!b.count[k]) continue;
        const size_t bytes = lshape_scratch_bytes(c->g.L, b.nmax[k]);
        if (bytes > kLdsLimit) need = std::max(need, ((bytes + 7) / 8) * 8 * size_t(b.count[k]));
    }
    return need;
}

void reserve_dp_scratch(gr_ctx* c, const std::vector<BatchLaunch>& plan) {
    size_t need = 0;
    for (const BatchLaunch& b : plan) need = std::max(need, dp_tables_needed(c, b));
    ensure_gscratch(c, need);
}

// Execution-order descriptors of the stage slots in slots[first, last):
// each DP workgroup reads one int4 instead of chasing slot -> net -> offsets.
void upload_desc(gr_ctx* c, const std::vector<int>& slots, int first, int last) {
    if (c->desc_h.size() != size_t(c->stage_count)) c->desc_h.assign(c->stage_count, int4{-1, -1, -1, -1});
    bool changed = false;
    for (int j = first; j < last; ++j) {
        const int s = slots[j], id = c->net_ids_h[s];
        // nodes (at most kMaxTreeNodes) | the root's height << 16
        const int4 d{c->node_off_h[s], (c->node_off_h[s + 1] - c->node_off_h[s]) | c->height_h[s] << 16,
                     c->route_per_pin * c->pin_acc_num_h[id],
                     c->route_per_pin * (c->pin_acc_num_h[id + 1] - c->pin_acc_num_h[id])};
        int4& h = c->desc_h[j];
        if (h.x != d.x || h.y != d.y || h.z != d.z || h.w != d.w) changed = true;
        h = d;
    }
    if (changed) upload(c->desc + first, c->desc_h.data() + first, sizeof(int4) * size_t(last - first), c->stream);
}

// Launch arguments of FLT batch b of the uploaded stage.
FltArgs flt_batch_args(gr_ctx* c, int b) {
    const int begin = c->f_batch_off[b], count = c->f_batch_off[b + 1] - begin;
    FltArgs a = c->f;
    a.presum = c->presum;
    a.vcost = c->vcost;
    a.desc = c->f.desc + begin;
    a.sizes = c->f.sizes + begin;
    a.net_ids = c->f.net_ids + begin;
    a.count = count;
    a.node_base = c->f_base[b].x;
    a.edge_base = c->f_base[b].y;
    a.slot_base = c->f_base[b].z;
    a.edges = c->f_base[b].w;
    a.queue = nullptr;
    a.lds_dag_bytes = 0;
    a.lds_shape = c->f_batch_max[b];
    a.far_edges = c->f.far_edges + c->f_far_off[b];
    a.far_count = c->f_far_off[b + 1] - c->f_far_off[b];
    return a;
}

// One DAG class of batch b: DAGs [off, off + count) of the batch with
// `per_wave` DAGs per wavefront, each DAG's tables in LDS when the class's
// largest DAG fits (per_wave halved until the workgroup's share fits), the
// persistent form draining `queue` when given; otherwise the batch tables in
// global memory.
void launch_flt_class(gr_ctx* c, FltArgs a, const FltLdsShape& mx, int off, int count, int per_wave, int* queue,
                      hipStream_t stream) {
    a.desc += off;
    a.sizes += off;
    a.net_ids += off;
    a.count = count;
    const size_t bytes = flt_lds_bytes(c->g.L, mx.slots, mx.nodes, mx.edges, mx.dnodes, mx.depths);
    const size_t avail = size_t(c->lds_limit) - 256;  // the kernel's static LDS stays beside the tables
    int p = per_wave;
    while (p > 1 && bytes * size_t(p) > avail) p /= 2;
    a.per_wave = p;
    if (bytes <= avail) {
        a.lds_dag_bytes = int(bytes);
        a.lds_shape = mx;
        a.queue = queue;
        launch_flt_route(a, true, c->cus, stream);
    } else {
        a.queue = nullptr;
        launch_flt_route(a, false, c->cus, stream);
    }
}

// The DP of batch b: its large-DAG class on the side stream beside the
// rest (one stream when profiling, so each launch is timed alone), each
// class with LDS shares sized by its own largest DAG.
void launch_flt_dp(gr_ctx* c, const FltArgs& a, int b, int per_wave, int* queue) {
    const int count = c->f_batch_off[b + 1] - c->f_batch_off[b], large = c->f_large[b];
    const FltLdsShape* mx = &c->f_class_max[size_t(b) * kFltClasses];
    const size_t qs = size_t(kFltQueues) * kFltQueueStride;
    int* q1 = queue ? queue + qs : nullptr;
    if (!large || large == count || c->profiling) {
        if (large) launch_flt_class(c, a, mx[0], 0, large, per_wave, queue, c->stream);
        if (count > large) launch_flt_class(c, a, mx[1], large, count - large, per_wave, q1, c->stream);
        return;
    }
    hip_check(hipEventRecord(c->fork_ev, c->stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->side, c->fork_ev, 0), "hipStreamWaitEvent");
    launch_flt_class(c, a, mx[0], 0, large, per_wave, queue, c->side);
    launch_flt_class(c, a, mx[1], large, count - large, per_wave, q1, c->stream);
    hip_check(hipEventRecord(c->join_ev, c->side), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->stream, c->join_ev, 0), "hipStreamWaitEvent");
}

// Route storage may have been replaced since the FLT upload: bind the
// current buffers and check the stage's nets against them.
void bind_flt_storage(gr_ctx* c) {
    if (!c->routes2) throw GrError(GR_E_STATE, "no phase-two route storage");
    for (int id : c->f_ids_h)
        if (id >= c->net_count || c->pin_off2_h[id] < 0)
            throw GrError(GR_E_STATE, "FLT stage net without phase-two route storage");
    c->f.routes = RouteStore{c->routes2, c->pin_off2, c->route2_per_pin};
    c->f.pin_acc_num = c->pin_acc_num;
}

// The per-net line inputs of the evaluator score and connectivity check,
// checked and uploaded into `ar`: every index the kernels derive from them
// is validated here.
ScoreArgs net_lines(gr_ctx* c, DeviceArena& ar, int nets, const int* pin_off, const int* pins, const int* extra_off,
                    const int* extra, const int* sub_off, const int* subs, const int* layer_dir) {
    if (nets < 0 || !pin_off || !extra_off || !sub_off || !layer_dir) throw GrError(GR_E_ARG, "null argument");
    if (!c->routes) throw GrError(GR_E_STATE, "routes not allocated (gr_ctx_set_routes)");
    const int X = c->g.X, Y = c->g.Y, M = c->g.L + 1;
    // lines pack x | y << 16 into a signed int: coordinates must stay below 32768
    if (X > 32768 || Y > 32768 || M > kScoreLayers)
        throw GrError(GR_E_ARG, "grid too large for the score's line packing and layer counters");
    const long long XY = (long long)X * Y, cells = XY * M;
    if (cells >= (1LL << 31)) throw GrError(GR_E_ARG, "metal-layer grid too large for the score's cell keys");
    const int np = pin_off[nets], ne = extra_off[nets], ns = sub_off[nets];
    if (np < 0 || ne < 0 || (ne & 1) || ns < 0 || (np && !pins) || (ne && !extra) || (ns && !subs))
        throw GrError(GR_E_ARG, "invalid net line offsets");
    for (int k = 0; k < nets; ++k)
        if (pin_off[k + 1] < pin_off[k] || extra_off[k + 1] < extra_off[k] || sub_off[k + 1] < sub_off[k])
            throw GrError(GR_E_ARG, "net offsets must be non-decreasing");
    for (int i = 0; i < np; ++i)
        if (pins[i] < 0 || pins[i] >= cells) throw GrError(GR_E_ARG, "pin cell outside the metal-layer grid");
    for (int i = 0; i < ne; ++i)
        if (extra[i] < 0 || extra[i] >= cells) throw GrError(GR_E_ARG, "extra route cell outside the grid");
    const int nsub = int(c->pin_acc_num_h.size()) - 1;
    for (int i = 0; i < ns; ++i)
        if (subs[i] < 0 || subs[i] >= nsub) throw GrError(GR_E_ARG, "subnet id outside the route store");
    for (int z = 0; z < M; ++z)
        if (layer_dir[z] != 0 && layer_dir[z] != 1) throw GrError(GR_E_ARG, "layer direction must be 0 or 1");
    ScoreArgs a{};
    a.nets = nets, a.X = X, a.Y = Y, a.M = M;
    int* d_pin_off = ar.alloc<int>(nets + 1);
    int* d_pins = ar.alloc<int>(np);
    int* d_extra_off = ar.alloc<int>(nets + 1);
    int* d_extra = ar.alloc<int>(ne);
    int* d_sub_off = ar.alloc<int>(nets + 1);
    int* d_subs = ar.alloc<int>(ns);
    int* d_dir = ar.alloc<int>(M);
    hipStream_t st = c->stream;
    upload(d_pin_off, pin_off, sizeof(int) * (nets + 1), st);
    if (np) upload(d_pins, pins, sizeof(int) * np, st);
    upload(d_extra_off, extra_off, sizeof(int) * (nets + 1), st);
    if (ne) upload(d_extra, extra, sizeof(int) * ne, st);
    upload(d_sub_off, sub_off, sizeof(int) * (nets + 1), st);
    if (ns) upload(d_subs, subs, sizeof(int) * ns, st);
    upload(d_dir, layer_dir, sizeof(int) * M, st);
    a.pin_off = d_pin_off, a.pins = d_pins, a.extra_off = d_extra_off, a.extra = d_extra;
    a.sub_off = d_sub_off, a.subs = d_subs, a.dir = d_dir;
    a.rs = {c->routes, c->pin_acc_num, c->route_per_pin};
    a.routes2 = c->routes2, a.pin_off2 = c->routes2 ? c->pin_off2 : nullptr, a.per_pin2 = c->route2_per_pin;
    return a;
}

void require_stage(gr_ctx* c, int begin, int count) {
    if (!c->net_ids) throw GrError(GR_E_STATE, "no L-shape stage topology uploaded");
    if (!c->routes) throw GrError(GR_E_STATE, "no route storage");
    if (begin < 0 || count < 0 || begin + count > c->stage_count) throw GrError(GR_E_ARG, "slot range outside the stage");
}

// Soft DP forward over slots [begin, begin + count): the smooth cost model
// at (slope, zero_slope, via_weight), or, with wcost and vcost (IDX order,
// host), those costs.
void soft_forward(gr_ctx* c, int begin, int count, double slope, double zero_slope, double via_weight,
                  const double* wcost, const double* vcost, double tau, int precision, double* total, double* per_net,
                  double* elapsed_ms) {
    if (!c || !total) throw GrError(GR_E_ARG, "null argument");
    if (precision != 32 && precision != 64) throw GrError(GR_E_ARG, "precision must be 32 or 64");
    if (!(tau > 0.0)) throw GrError(GR_E_ARG, "tau must be positive");
    if (c->g.L > kSoftLayers) throw GrError(GR_E_ARG, "the soft DP supports at most 10 routing layers");
    require_stage(c, begin, count);
    c->soft_ready = false;
    const size_t cells = c->cells;
    SoftCostArgs& sc = c->soft_cost;
    SoftArgs& a = c->soft;
    c->soft_ws.carve([&](auto& cv) {
        sc.wcost = cv.template take<double>(cells);
        sc.presum = cv.template take<double>(cells);
        sc.vcost = cv.template take<double>(cells);
        sc.dw_ds = cv.template take<double>(cells);
        sc.dw_dz = cv.template take<double>(cells);
        sc.dv_ds = cv.template take<double>(cells);
        sc.dv_dz = cv.template take<double>(cells);
        a.gw = cv.template take<double>(cells);
        a.gv = cv.template take<double>(cells);
        c->soft_params = cv.template take<double>(size_t(kStatStripes) * kCounterStride);
        a.value = cv.template take<double>(std::max(count, 1));
    });
    const long long nodes = (long long)c->node_off_h[begin + count] - c->node_off_h[begin];
    void* tables = nullptr;
    c->soft_tables.carve(
        [&](auto& cv) { tables = cv.template take<char>(soft_scratch_bytes(c->g.L, nodes, precision)); });
    sc.g = c->g;
    sc.capacity = c->capacity;
    sc.demand = c->demand;
    sc.slope = slope;
    sc.zero_slope = zero_slope;
    sc.via_weight = via_weight;
    const bool given = wcost != nullptr;
    if (given) {
        std::vector<double> h;
        to_track_major(c->g, wcost, h);
        upload(sc.wcost, h.data(), sizeof(double) * cells, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "soft cost upload");
        to_track_major(c->g, vcost, h);
        upload(sc.vcost, h.data(), sizeof(double) * cells, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "soft cost upload");
    }
    a.g = c->g;
    a.presum = sc.presum;
    a.vcost = sc.vcost;
    a.node_off = c->node_off;
    a.nodes = c->nodes;
    a.parents = c->parents;
    a.begin = begin;
    a.count = count;
    a.tau = tau;
    a.scratch = tables;
    hipEvent_t e0 = take_event(c), e1 = take_event(c);
    hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
    if (given)
        hipLaunchKernelGGL(soft_prefix_kernel, dim3(c->tracks), dim3(256), 0, c->stream, c->g, sc.wcost, sc.presum);
    else
        hipLaunchKernelGGL(soft_costs_kernel, dim3(c->tracks), dim3(256), 0, c->stream, sc);
    if (count) launch_soft_forward(a, precision, c->soft_kind, c->cus, c->stream);
    hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
    hip_check(hipGetLastError(), "soft forward kernels");
    std::vector<double> v(std::max(count, 1));
    download(v.data(), a.value, sizeof(double) * count, c->stream);
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (elapsed_ms) *elapsed_ms = ms;
    c->next_event -= 2;
    double sum = 0.0;
    for (int k = 0; k < count; ++k) sum += v[k];
    *total = sum;
    if (per_net) std::memcpy(per_net, v.data(), sizeof(double) * count);
    c->soft_precision = precision;
    c->soft_given = given;
    c->soft_ready = true;
}

}  // namespace

extern "C" {

int gr_device_name(char* buffer, int size) {
    return guarded([&] {
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (!count) throw GrError(GR_E_HIP, "no HIP device");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
        snprintf(buffer, size_t(size), "%s %s", prop.name, prop.gcnArchName);
    });
}

int gr_ctx_create(const gr_grid_desc* grid, gr_ctx** out) {
    return guarded([&] {
        if (!grid || !out) throw GrError(GR_E_ARG, "null argument");
        if (grid->L < 1 || grid->L >= kMaxLayers) throw GrError(GR_E_ARG, "routing layers must be 1..15");
        if (grid->X < 2 || grid->Y < 2) throw GrError(GR_E_ARG, "grid must be at least 2 x 2");
        // a track's Fenwick values live in LDS during the refresh: 64 KiB
        if (grid->X > 8192 || grid->Y > 8192) throw GrError(GR_E_ARG, "tracks longer than 8192 cells");
        if (size_t(grid->L) * grid->X * grid->Y > size_t(INT_MAX)) throw GrError(GR_E_ARG, "grid exceeds 2^31 cells");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (!count) throw GrError(GR_E_HIP, "no HIP device");
        auto c = std::make_unique<gr_ctx>();
        hip_check(hipSetDevice(0), "hipSetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
        c->cus = prop.multiProcessorCount;
        c->noalloc_scratch = noalloc_scratch_address();  // with device 0 current
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) == hipSuccess && lds > 0)
            c->lds_limit = std::min(lds, 160 * 1024);
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&c->join2_ev, hipEventDisableTiming), "hipEventCreate");
        DeviceGrid& g = c->g;
        g.L = grid->L, g.X = grid->X, g.Y = grid->Y, g.DIR = grid->DIR & 1, g.XY = grid->X * grid->Y;
        g.unit_wire = grid->unit_length_wire_cost;
        g.unit_via = grid->unit_via_cost;
        for (int l = 0; l < kMaxLayers; ++l) g.short_cost[l] = l < g.L ? grid->unit_length_short_costs[l] : 0.0;
        int t = 0;
        for (int l = 0; l < g.L; ++l) {
            g.track_begin[l] = t;
            t += layer_dir(g, l) ? g.X : g.Y;
        }
        for (int l = g.L; l <= kMaxLayers; ++l) g.track_begin[l] = t;
        c->tracks = t;
        c->maxlen = std::max(g.X, g.Y);
        c->cells = size_t(g.L) * g.XY;
        int* xl = c->mem.alloc<int>(g.X - 1);
        int* yl = c->mem.alloc<int>(g.Y - 1);
        upload(xl, grid->x_edge_len, sizeof(int) * (g.X - 1), c->stream);
        upload(yl, grid->y_edge_len, sizeof(int) * (g.Y - 1), c->stream);
        g.x_edge_len = xl, g.y_edge_len = yl;
        c->capacity = c->mem.alloc<float>(c->cells);
        c->demand = c->mem.alloc<float>(c->cells);
        c->vcost = c->mem.alloc<float>(c->cells);
        c->presum = c->mem.alloc<double>(c->cells);
        c->timestamp = c->mem.alloc<int>(c->cells);
        c->track_dirty = c->mem.alloc<unsigned>(c->tracks);
        c->stats = c->mem.alloc<Stats>(kStatStripes);
        c->error = c->mem.alloc<int>(1);
        c->refreshed = c->mem.alloc<unsigned long long>(kStatStripes * kCounterStride);
        hip_check(hipMemsetAsync(c->refreshed, 0, sizeof(unsigned long long) * kStatStripes * kCounterStride, c->stream),
                  "clear counter");
        std::vector<float> cap;
        to_track_major(g, grid->capacity, cap);
        upload(c->capacity, cap.data(), sizeof(float) * c->cells, c->stream);
        hip_check(hipMemsetAsync(c->demand, 0, sizeof(float) * c->cells, c->stream), "clear demand");
        hip_check(hipMemsetAsync(c->timestamp, 0, sizeof(int) * c->cells, c->stream), "clear timestamps");
        hip_check(hipMemsetAsync(c->stats, 0, sizeof(Stats) * kStatStripes, c->stream), "clear statistics");
        hip_check(hipMemsetAsync(c->error, 0, sizeof(int), c->stream), "clear error");
        std::vector<unsigned> ones(c->tracks, 1u);
        upload(c->track_dirty, ones.data(), sizeof(unsigned) * c->tracks, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "context setup");
        *out = c.release();
    });
}

int gr_ctx_destroy(gr_ctx* ctx) {
    return guarded([&] {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int gr_ctx_set_routes(gr_ctx* c, int net_count, const int* pin_acc_num) {
    return guarded([&] {
        if (!c || net_count < 0 || !pin_acc_num) throw GrError(GR_E_ARG, "invalid route storage arguments");
        c->drop_graph();  // a captured batch loop holds the old route storage addresses
        c->route_per_pin = c->next_route_per_pin;
        c->release_phase2();  // phase-two offsets index the old subnets
        c->mem.release(c->routes);
        c->mem.release(c->pin_acc_num);
        c->mem.release(c->of_flags);
        c->net_count = net_count;
        c->pin_acc_num_h.assign(pin_acc_num, pin_acc_num + net_count + 1);
        c->route_words = (long long)c->route_per_pin * pin_acc_num[net_count] + 1;
        c->pin_acc_num = c->mem.alloc<int>(net_count + 1);
        upload(c->pin_acc_num, pin_acc_num, sizeof(int) * (net_count + 1), c->stream);
        c->routes = c->mem.alloc<int>(size_t(c->route_words));
        hip_check(hipMemsetAsync(c->routes, 0, sizeof(int) * size_t(c->route_words), c->stream), "clear routes");
        hip_check(hipStreamSynchronize(c->stream), "route storage");
    });
}

int gr_ctx_set_route_stride(gr_ctx* c, int words_per_pin) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (words_per_pin < 4) throw GrError(GR_E_ARG, "route words per pin must be at least 4");
        c->next_route_per_pin = words_per_pin;  // the current storage keeps its stride
    });
}

int gr_ctx_set_cost_params(gr_ctx* c, double slope, double zero_slope, double via_weight, int integer_wire) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        c->cp.slope = slope;
        c->cp.zero_slope = zero_slope;
        c->cp.via_weight = via_weight;
        c->cp.integer_wire = integer_wire ? 1 : 0;
        std::vector<unsigned> ones(c->tracks, 1u);
        upload(c->track_dirty, ones.data(), sizeof(unsigned) * c->tracks, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "cost parameters");
    });
}

int gr_cost_refresh(gr_ctx* c) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        launch_refresh(c, 0);
        hip_check(hipStreamSynchronize(c->stream), "gr_cost_refresh");
    });
}

int gr_cost_refresh_full(gr_ctx* c) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        launch_refresh(c, 1);
        hip_check(hipStreamSynchronize(c->stream), "gr_cost_refresh_full");
    });
}

int gr_cost_download(gr_ctx* c, float* wcost, float* vcost, double* presum) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (wcost) {
            // wcost is not kept resident: recompute it from the current demand
            // with the refresh kernel into temporaries, leaving the resident
            // costs and dirty flags untouched
            DeviceArena scratch;
            float* tmp = scratch.alloc<float>(c->cells);
            float* tv = scratch.alloc<float>(c->cells);
            double* tp = scratch.alloc<double>(c->cells);
            unsigned* td = scratch.alloc<unsigned>(c->tracks);
            hipLaunchKernelGGL(refresh_tracks_kernel, dim3(c->tracks), dim3(kRefreshThreads),
                               size_t(c->maxlen) * sizeof(double), c->stream, c->g, c->capacity, c->demand, tv, tp,
                               tmp, td, 1, c->cp, (unsigned long long*)nullptr);
            hip_check(hipGetLastError(), "refresh for wcost");
            std::vector<float> h(c->cells);
            download(h.data(), tmp, sizeof(float) * c->cells, c->stream);
            from_track_major(c->g, h, wcost);
        }
        if (vcost) {
            std::vector<float> h(c->cells);
            download(h.data(), c->vcost, sizeof(float) * c->cells, c->stream);
            from_track_major(c->g, h, vcost);
        }
        if (presum) {
            std::vector<double> h(c->cells);
            download(h.data(), c->presum, sizeof(double) * c->cells, c->stream);
            from_track_major(c->g, h, presum);
        }
    });
}

int gr_demand_upload(gr_ctx* c, const float* demand) {
    return guarded([&] {
        if (!c || !demand) throw GrError(GR_E_ARG, "null argument");
        std::vector<float> h;
        to_track_major(c->g, demand, h);
        upload(c->demand, h.data(), sizeof(float) * c->cells, c->stream);
        std::vector<unsigned> ones(c->tracks, 1u);
        upload(c->track_dirty, ones.data(), sizeof(unsigned) * c->tracks, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "demand upload");
    });
}

int gr_demand_download(gr_ctx* c, float* demand) {
    return guarded([&] {
        if (!c || !demand) throw GrError(GR_E_ARG, "null argument");
        std::vector<float> h(c->cells);
        download(h.data(), c->demand, sizeof(float) * c->cells, c->stream);
        from_track_major(c->g, h, demand);
    });
}

int gr_timestamp_download(gr_ctx* c, int* ts) {
    return guarded([&] {
        if (!c || !ts) throw GrError(GR_E_ARG, "null argument");
        std::vector<int> h(c->cells);
        download(h.data(), c->timestamp, sizeof(int) * c->cells, c->stream);
        from_track_major(c->g, h, ts);
    });
}

int gr_stats_download(gr_ctx* c, double* wl, long long* vias) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        const Stats s = read_stats(c);
        if (wl) *wl = s.total_wirelength;
        if (vias) *vias = (long long)s.total_vias;
    });
}

int gr_lshape_upload(gr_ctx* c, int count, const int* net_ids, const int* node_off, const int* nodes,
                     const int* parents) {
    return guarded([&] {
        if (!c || count < 0 || !net_ids || !node_off || !nodes || !parents) throw GrError(GR_E_ARG, "null argument");
        for (int k = 0; k < count; ++k) {
            if (net_ids[k] < 0 || net_ids[k] >= c->net_count) throw GrError(GR_E_ARG, "net id outside route storage");
            const int n = node_off[k + 1] - node_off[k];
            if (n < 1) throw GrError(GR_E_ARG, "empty tree");
            if (n > kMaxTreeNodes) throw GrError(GR_E_ARG, "tree has more than 32767 nodes");
            if (parents[node_off[k]] != -1) throw GrError(GR_E_ARG, "tree root must have parent -1");
            for (int i = 1; i < n; ++i)
                if (parents[node_off[k] + i] < 0 || parents[node_off[k] + i] >= i)
                    throw GrError(GR_E_ARG, "parents must precede children (DFS preorder)");
            for (int i = 0; i < n; ++i) {
                const int v = nodes[node_off[k] + i];
                if (v < 0 || v >= (c->g.L + 1) * c->g.XY) throw GrError(GR_E_ARG, "node index outside the grid");
            }
        }
        for (int* p : {c->net_ids, c->node_off, c->nodes, c->parents, c->lvl, c->link}) {
            int* q = p;
            c->mem.release(q);
        }
        c->mem.release(c->desc);
        c->net_ids = c->node_off = c->nodes = c->parents = c->lvl = c->link = nullptr;
        c->slots_h.clear();
        c->desc_h.clear();  // the new descriptor buffer holds nothing yet
        c->soft_ready = false;
        c->drop_graph();
        ++c->stage_gen;
        const int total = node_off[count];
        c->stage_count = count;
        c->node_off_h.assign(node_off, node_off + count + 1);
        c->net_ids = c->mem.alloc<int>(count);
        c->node_off = c->mem.alloc<int>(count + 1);
        c->nodes = c->mem.alloc<int>(total);
        c->parents = c->mem.alloc<int>(total);
        c->desc = c->mem.alloc<int4>(count);
        c->net_ids_h.assign(net_ids, net_ids + count);
        c->noalloc_slots.assign(size_t(count), 0);
        c->slots_h.assign(size_t(count), 0);
        c->desc_h.assign(size_t(count), int4{-1, -1, -1, -1});
        upload(c->net_ids, net_ids, sizeof(int) * count, c->stream);
        upload(c->node_off, node_off, sizeof(int) * (count + 1), c->stream);
        upload(c->nodes, nodes, sizeof(int) * total, c->stream);
        upload(c->parents, parents, sizeof(int) * total, c->stream);
        {
            // per net: node heights (leaves 0), the nodes ordered by (height,
            // index), and each node's children in descending index as a list
            std::vector<int> lvl(total), link(total), height, cnt, first, last;
            c->height_h.assign(count, 0);
            for (int k = 0; k < count; ++k) {
                const int b = node_off[k], n = node_off[k + 1] - b;
                height.assign(n, 0);
                first.assign(n, -1);
                last.assign(n, -1);
                std::vector<int> next(n, -1);
                for (int i = n - 1; i >= 1; --i) {
                    const int p = parents[b + i];
                    height[p] = std::max(height[p], height[i] + 1);
                    if (first[p] < 0) first[p] = i;
                    else next[last[p]] = i;
                    last[p] = i;
                }
                cnt.assign(n + 1, 0);
                for (int i = 0; i < n; ++i) ++cnt[height[i] + 1];
                for (int h = 0; h < n; ++h) cnt[h + 1] += cnt[h];
                for (int i = 0; i < n; ++i) lvl[b + cnt[height[i]]++] = i | height[i] << 16;
                for (int i = 0; i < n; ++i) link[b + i] = (first[i] + 1) | (next[i] + 1) << 16;
                c->height_h[k] = height[0];
            }
            c->lvl = c->mem.alloc<int>(total);
            c->link = c->mem.alloc<int>(total);
            upload(c->lvl, lvl.data(), sizeof(int) * total, c->stream);
            upload(c->link, link.data(), sizeof(int) * total, c->stream);
            hip_check(hipStreamSynchronize(c->stream), "stage upload");
        }
        // algorithmic bytes of the DP per net: tree input 8n, via stacks
        // 4*L*n, straight edges 16*L, bent edges 40*L, route words 32n
        const int L = c->g.L, X = c->g.X, Y = c->g.Y;
        c->slot_dp_bytes.assign(count, 0.0);
        for (int k = 0; k < count; ++k) {
            const int b = node_off[k], n = node_off[k + 1] - b;
            double bytes = 8.0 * n + 4.0 * L * n + 32.0 * n;
            for (int i = 1; i < n; ++i) {
                const int v = nodes[b + i], q = nodes[b + parents[b + i]];
                const bool st = (v / Y % X == q / Y % X) || (v % Y == q % Y);
                bytes += (st ? 16.0 : 40.0) * L;
            }
            c->slot_dp_bytes[k] = bytes;
        }
        hip_check(hipStreamSynchronize(c->stream), "stage upload");
    });
}

int gr_lshape_batch(gr_ctx* c, int begin, int count, int stamp) {
    (void)stamp;
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        require_stage(c, begin, count);
        std::vector<int> slots(c->stage_count);
        const BatchLaunch b = plan_batch(c, begin, count, slots);
        reserve_dp_scratch(c, {b});
        if (c->slots_h.size() != slots.size()) c->slots_h.assign(slots.size(), 0);
        std::copy(slots.begin() + begin, slots.begin() + begin + count, c->slots_h.begin() + begin);
        upload_desc(c, slots, begin, begin + count);
        launch_lshape(c, b);
        hip_check(hipStreamSynchronize(c->stream), "gr_lshape_batch");
        check_device_error(c, "gr_lshape_batch");
    });
}

int gr_lshape_batch_noalloc(gr_ctx* c, int begin, int count, int stamp) {
    (void)stamp;
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        require_stage(c, begin, count);
        if (!c->noalloc_scratch) throw GrError(GR_E_STATE, "no scratch address for this device");
        if (c->noalloc_slots.size() != size_t(c->stage_count) || c->desc_h.size() != size_t(c->stage_count))
            throw GrError(GR_E_STATE, "stage topology not uploaded");
        if (!count) return;
        std::vector<int>& slots = c->noalloc_slots;
        int* queue = reinterpret_cast<int*>(c->noalloc_scratch);
        const DpTables tables{c->noalloc_scratch + kNoAllocQueueBytes / sizeof(double),
                              kNoAllocScratchBytes - kNoAllocQueueBytes};
        BatchLaunch b = plan_batch_queue(c, begin, count, slots);
        const bool persistent = per_wave_for(c, 0, b.nmax[0]) > 0;
        if (!persistent) {
            // a tree too large for LDS: the size-class launches, its tables in scratch
            b = plan_batch(c, begin, count, slots);
            const size_t need = dp_tables_needed(c, b);
            if (need > tables.bytes)
                throw GrError(GR_E_CAPACITY, "batch needs " + std::to_string(need) + " bytes of DP tables; the "
                              "allocation-free entry point holds " + std::to_string(tables.bytes));
        }
        std::copy(slots.begin() + begin, slots.begin() + begin + count, c->slots_h.begin() + begin);
        upload_desc(c, slots, begin, begin + count);
        if (persistent) {
            // the counters keep the last call's values: zero them (a kernel, not a memset)
            hipLaunchKernelGGL(fill_int_kernel, dim3((kDpQueues * kDpQueueStride + 255) / 256), dim3(256), 0, c->stream,
                               queue, 0, (long long)kDpQueues * kDpQueueStride);
            hip_check(hipGetLastError(), "clear work-queue counters");
            launch_lshape_queue(c, b, queue);
        } else {
            launch_lshape(c, b, tables);
        }
        hip_check(hipStreamSynchronize(c->stream), "gr_lshape_batch_noalloc");
        check_device_error(c, "gr_lshape_batch_noalloc");
    });
}

int gr_commit_batch(gr_ctx* c, int begin, int count, int stamp, int K) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        require_stage(c, begin, count);
        launch_commit(c, begin, count, stamp, K);
        hip_check(hipStreamSynchronize(c->stream), "gr_commit_batch");
    });
}

int gr_lshape_run(gr_ctx* c, int nbatch, const int* batch_off, int first_stamp, int mode, double* elapsed_ms) {
    return guarded([&] {
        if (!c || nbatch < 0 || !batch_off) throw GrError(GR_E_ARG, "invalid arguments");
        if (mode < 0 || mode > 2)
            throw GrError(GR_E_ARG, "execution mode not available (0 = plain, 1 = launches, 2 = persistent + graph)");
        require_stage(c, 0, batch_off[nbatch]);
        if (batch_off[0] != 0) throw GrError(GR_E_ARG, "batch offsets must start at 0");
        std::vector<int> slots(c->stage_count);
        std::vector<BatchLaunch> plan(nbatch);
        size_t plain_need = 0;
        for (int b = 0; b < nbatch; ++b) {
            if (batch_off[b + 1] < batch_off[b]) throw GrError(GR_E_ARG, "batch offsets must be non-decreasing");
            const int begin = batch_off[b], count = batch_off[b + 1] - batch_off[b];
            plan[b] = mode == 0 ? plan_batch_plain(c, begin, count, slots) : plan_batch(c, begin, count, slots);
            if (mode == 0) plain_need = std::max(plain_need, plain_tables_needed(c, plan[b]));
        }
        if (mode == 0) ensure_gscratch(c, plain_need);
        else reserve_dp_scratch(c, plan);
        const long long qwords = (long long)kPersistentGroups * kDpQueues * kDpQueueStride;  // per batch
        if (mode == 2 && c->dp_queue_cap < nbatch) {
            c->drop_graph();
            c->mem.release(c->dp_queue);
            c->dp_queue = c->mem.alloc<int>(size_t(std::max(nbatch, 1)) * qwords);
            c->dp_queue_cap = nbatch;
        }
        auto enqueue = [&] {
            if (mode == 2 && nbatch)  // a kernel, not a memset: the loop may be captured
                hipLaunchKernelGGL(fill_int_kernel, dim3(unsigned((nbatch * qwords + 255) / 256)), dim3(256), 0,
                                   c->stream, c->dp_queue, 0, nbatch * qwords);
            for (int b = 0; b < nbatch; ++b) {
                launch_refresh(c, 0);
                if (mode == 2) launch_lshape_persistent(c, plan[b], c->dp_queue + size_t(b) * qwords);
                else if (mode == 1) launch_lshape(c, plan[b]);
                else launch_lshape_plain(c, plan[b]);
                launch_commit(c, batch_off[b], batch_off[b + 1] - batch_off[b], first_stamp + b, 1);
            }
        };
        // a captured graph cannot carry profiling events: profiled runs of
        // mode 2 issue the same launches directly
        const bool graph = mode == 2 && !c->profiling;
        if (graph) {
            std::vector<int> key{c->stage_gen, first_stamp, nbatch, c->per_wave[0], c->per_wave[1], mode, c->tiny_bound,
                                 c->tiny_per_wave};
            key.insert(key.end(), batch_off, batch_off + nbatch + 1);
            // the graph reads the descriptors when it runs
            c->slots_h = slots;
            upload_desc(c, slots, 0, c->stage_count);
            if (!c->graph_exec || key != c->graph_key) {
                c->drop_graph();
                hip_check(hipStreamSynchronize(c->stream), "before graph capture");
                const bool profiling = c->profiling;
                c->profiling = false;
                hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try {
                    enqueue();
                } catch (...) {
                    hipGraph_t broken = nullptr;
                    (void)hipStreamEndCapture(c->stream, &broken);
                    if (broken) (void)hipGraphDestroy(broken);
                    c->profiling = profiling;
                    throw;
                }
                c->profiling = profiling;
                hip_check(hipStreamEndCapture(c->stream, &c->graph), "hipStreamEndCapture");
                hip_check(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                c->graph_key = key;
            }
        } else {
            c->slots_h = slots;
            upload_desc(c, slots, 0, c->stage_count);
        }
        hip_check(hipStreamSynchronize(c->stream), "stage plan");
        hip_check(hipDeviceSynchronize(), "before DAG interval");
        const auto t0 = std::chrono::steady_clock::now();
        if (graph) hip_check(hipGraphLaunch(c->graph_exec, c->stream), "hipGraphLaunch");
        else enqueue();
        hip_check(hipStreamSynchronize(c->stream), "DAG interval");
        const auto t1 = std::chrono::steady_clock::now();
        c->last_t0 = t0, c->last_t1 = t1;
        if (elapsed_ms) *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        check_device_error(c, "gr_lshape_run");
    });
}

int gr_ctx_synchronize(gr_ctx* c) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int gr_ctx_last_interval(gr_ctx* c, long long* start_ns, long long* end_ns) {
    return guarded([&] {
        if (!c || !start_ns || !end_ns) throw GrError(GR_E_ARG, "null argument");
        using std::chrono::duration_cast;
        using std::chrono::nanoseconds;
        *start_ns = (long long)duration_cast<nanoseconds>(c->last_t0.time_since_epoch()).count();
        *end_ns = (long long)duration_cast<nanoseconds>(c->last_t1.time_since_epoch()).count();
    });
}

int gr_ctx_set_nets_per_wave(gr_ctx* c, const int* per_wave) {
    return guarded([&] {
        if (!c || !per_wave) throw GrError(GR_E_ARG, "null argument");
        for (int k = 0; k < kTreeClasses; ++k)
            if (per_wave[k] != 1 && per_wave[k] != 2 && per_wave[k] != 4 && per_wave[k] != 8)
                throw GrError(GR_E_ARG, "nets per wavefront must be 1, 2, 4 or 8");
        std::copy(per_wave, per_wave + kTreeClasses, c->per_wave);
    });
}

int gr_lshape_tune(gr_ctx* c, int nbatch, const int* batch_off, int* per_wave, double* ms) {
    return guarded([&] {
        if (!c || nbatch < 0 || !batch_off) throw GrError(GR_E_ARG, "invalid arguments");
        require_stage(c, 0, batch_off[nbatch]);
        if (batch_off[0] != 0) throw GrError(GR_E_ARG, "batch offsets must start at 0");
        std::vector<int> slots(c->stage_count);
        std::vector<BatchLaunch> plan(nbatch);
        for (int b = 0; b < nbatch; ++b) {
            if (batch_off[b + 1] < batch_off[b]) throw GrError(GR_E_ARG, "batch offsets must be non-decreasing");
            plan[b] = plan_batch(c, batch_off[b], batch_off[b + 1] - batch_off[b], slots);
        }
        reserve_dp_scratch(c, plan);
        c->slots_h = slots;
        upload_desc(c, slots, 0, c->stage_count);
        launch_refresh(c, 0);
        const bool profiling = c->profiling;
        c->profiling = false;
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        const int candidates[4] = {1, 2, 4, 8};
        for (int k = 0; k < kTreeClasses; ++k) {
            int nmax = 1;
            bool any = false;
            for (const BatchLaunch& b : plan)
                if (b.count[k]) any = true, nmax = std::max(nmax, b.nmax[k]);
            double best = 1e30;
            int choice = c->per_wave[k];
            for (int ci = 0; ci < 4; ++ci) {
                const int p = candidates[ci];
                const size_t bytes = (lshape_scratch_bytes(c->g.L, nmax) + 7) / 8 * 8;
                double t = -1;
                if (any && bytes * p <= kLdsLimit) {
                    // every batch's launch of this class at the stage's current
                    // costs; the routes it writes are rewritten by the stage
                    t = 1e30;
                    for (int rep = 0; rep < 2; ++rep) {
                        hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
                        for (const BatchLaunch& b : plan) launch_lshape_class(c, b, k, p);
                        hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
                        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                        float f = 0;
                        hip_check(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
                        t = std::min(t, double(f));
                    }
                    if (t < best) best = t, choice = p;
                }
                if (ms) ms[k * 4 + ci] = t;
            }
            c->per_wave[k] = choice;
            if (per_wave) per_wave[k] = choice;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->profiling = profiling;
        c->drop_graph();
        check_device_error(c, "gr_lshape_tune");
    });
}

// Times the persistent DP of the uploaded stage's batches (mode 2's launches,
// the groups on their streams, at the stage's current costs; the routes it
// writes are rewritten by the stage) for each candidate setting applied by
// `set`: the best of two runs per candidate, into ms[i]; keeps the fastest
// (applied again before returning) and returns its index.
extern "C++" {
template <class Set>
int tune_persistent(gr_ctx* c, int nbatch, const int* batch_off, int ncand, Set&& set, double* ms) {
    if (!c || nbatch < 0 || !batch_off) throw GrError(GR_E_ARG, "invalid arguments");
    require_stage(c, 0, batch_off[nbatch]);
    if (batch_off[0] != 0) throw GrError(GR_E_ARG, "batch offsets must start at 0");
    for (int b = 0; b < nbatch; ++b)
        if (batch_off[b + 1] < batch_off[b]) throw GrError(GR_E_ARG, "batch offsets must be non-decreasing");
    const long long qwords = (long long)kPersistentGroups * kDpQueues * kDpQueueStride;
    int* queue = c->mem.alloc<int>(size_t(std::max(nbatch, 1)) * qwords);
    const bool profiling = c->profiling;
    c->profiling = false;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    double best = 1e30;
    int choice = 0;
    std::vector<int> slots(c->stage_count);
    launch_refresh(c, 0);
    for (int ci = 0; ci < ncand; ++ci) {
        set(ci);
        std::vector<BatchLaunch> plan(nbatch);
        for (int b = 0; b < nbatch; ++b) plan[b] = plan_batch(c, batch_off[b], batch_off[b + 1] - batch_off[b], slots);
        reserve_dp_scratch(c, plan);
        c->slots_h = slots;
        upload_desc(c, slots, 0, c->stage_count);
        double t = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            hipLaunchKernelGGL(fill_int_kernel, dim3(unsigned((nbatch * qwords + 255) / 256)), dim3(256), 0, c->stream,
                               queue, 0, nbatch * qwords);
            hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
            for (int b = 0; b < nbatch; ++b) launch_lshape_persistent(c, plan[b], queue + size_t(b) * qwords);
            hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float f = 0;
            hip_check(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
            t = std::min(t, double(f));
        }
        if (ms) ms[ci] = t;
        if (t < best) best = t, choice = ci;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->mem.release(queue);
    set(choice);
    c->profiling = profiling;
    c->drop_graph();
    check_device_error(c, "persistent-form tuning");
    return choice;
}
}  // extern "C++"

int gr_lshape_tune_groups(gr_ctx* c, int nbatch, const int* batch_off, int* tiny_nodes, double* ms) {
    return guarded([&] {
        const int candidates[4] = {1, 2, 4, 8};  // 1: no group of its own (every routed tree has 2 nodes or more)
        const int ci = tune_persistent(c, nbatch, batch_off, 4, [&](int i) { c->tiny_bound = candidates[i]; }, ms);
        if (tiny_nodes) *tiny_nodes = candidates[ci];
    });
}

int gr_lshape_tune_tiny(gr_ctx* c, int nbatch, const int* batch_off, int* per_wave, double* ms) {
    return guarded([&] {
        const int candidates[4] = {1, 2, 4, 8};
        const int ci = tune_persistent(c, nbatch, batch_off, 4, [&](int i) { c->tiny_per_wave = candidates[i]; }, ms);
        if (per_wave) *per_wave = candidates[ci];
    });
}

int gr_ctx_set_tiny_per_wave(gr_ctx* c, int per_wave) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (per_wave != 0 && per_wave != 1 && per_wave != 2 && per_wave != 4 && per_wave != 8)
            throw GrError(GR_E_ARG, "nets per wavefront must be 0 (the small class's), 1, 2, 4 or 8");
        c->tiny_per_wave = per_wave;
    });
}

int gr_ctx_set_tiny_tree(gr_ctx* c, int nodes) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (nodes < 1 || nodes > kSmallTree) throw GrError(GR_E_ARG, "tiny-tree bound must lie in [1, 16]");
        c->tiny_bound = nodes;
    });
}

int gr_ctx_set_flt_split(gr_ctx* c, int bytes) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (bytes < 0) throw GrError(GR_E_ARG, "FLT class split must be at least 0 bytes");
        c->f_lds_split = bytes;
    });
}

int gr_ctx_reset(gr_ctx* c) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (c->profiling) {
            // keep the committed edge and via counts of the profiled interval
            const Stats now = read_stats(c);
            c->prof_wire_edges += now.wire_edges - c->prof_stats0.wire_edges;
            c->prof_vias += now.total_vias - c->prof_stats0.total_vias;
            c->prof_stats0 = Stats{};
        }
        hip_check(hipMemsetAsync(c->demand, 0, sizeof(float) * c->cells, c->stream), "clear demand");
        hip_check(hipMemsetAsync(c->timestamp, 0, sizeof(int) * c->cells, c->stream), "clear timestamps");
        hip_check(hipMemsetAsync(c->stats, 0, sizeof(Stats) * kStatStripes, c->stream), "clear statistics");
        if (c->routes)
            hip_check(hipMemsetAsync(c->routes, 0, sizeof(int) * size_t(c->route_words), c->stream), "clear routes");
        // a fresh context has no phase-two storage (it belongs to an FLT stage)
        c->release_phase2();
        std::vector<unsigned> ones(c->tracks, 1u);
        upload(c->track_dirty, ones.data(), sizeof(unsigned) * c->tracks, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "gr_ctx_reset");
    });
}

int gr_profile_begin(gr_ctx* c) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        hip_check(hipStreamSynchronize(c->stream), "gr_profile_begin");
        c->prof.clear();
        c->next_event = 0;
        c->prof_stats0 = read_stats(c);
        c->prof_wire_edges = c->prof_vias = 0;
        c->refreshed0 = read_refreshed(c);
        c->profiling = true;
    });
}

int gr_profile_select(gr_ctx* c, unsigned kind_mask) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        c->prof_mask = kind_mask;
    });
}

int gr_profile_end(gr_ctx* c, double* ms, long long* launches, double* bytes, int kinds) {
    return guarded([&] {
        if (!c || kinds < 0) throw GrError(GR_E_ARG, "invalid arguments");
        hip_check(hipStreamSynchronize(c->stream), "gr_profile_end");
        c->profiling = false;
        std::vector<double> t(GR_K_COUNT, 0.0), by(GR_K_COUNT, 0.0);
        std::vector<long long> n(GR_K_COUNT, 0);
        for (const auto& r : c->prof) {
            float e = 0.f;
            hip_check(hipEventElapsedTime(&e, r.a, r.b), "hipEventElapsedTime");
            t[r.kind] += e;
            by[r.kind] += r.bytes;
            n[r.kind] += 1;
        }
        const Stats s1 = read_stats(c);
        const unsigned long long refreshed1 = read_refreshed(c);
        // refresh, per cell of a refreshed track: read capacity + demand, write
        // wcost + vcost (4 B each) and presum (8 B); wcost stays in LDS here
        // but is counted, as for the reference's full refresh
        const auto selected = [&](int kind) { return (c->prof_mask >> kind & 1u) != 0; };
        if (selected(GR_K_REFRESH)) by[GR_K_REFRESH] += 24.0 * double(refreshed1 - c->refreshed0);
        // wires: 4 B atomic read-modify-write + 4 B timestamp + 4 B dirty mark per unit edge
        if (selected(GR_K_COMMIT_WIRES))
            by[GR_K_COMMIT_WIRES] += 12.0 * double(c->prof_wire_edges + (s1.wire_edges - c->prof_stats0.wire_edges));
        // vias: 4 B timestamp + 4 B demand per via layer step
        if (selected(GR_K_COMMIT_VIAS))
            by[GR_K_COMMIT_VIAS] += 8.0 * double(c->prof_vias + (s1.total_vias - c->prof_stats0.total_vias));
        for (int k = 0; k < std::min(kinds, (int)GR_K_COUNT); ++k) {
            if (ms) ms[k] = t[k];
            if (launches) launches[k] = n[k];
            if (bytes) bytes[k] = by[k];
        }
        c->prof.clear();
        c->next_event = 0;
    });
}

int gr_routes_download(gr_ctx* c, int* routes, long long words) {
    return guarded([&] {
        if (!c || !routes || words < 0) throw GrError(GR_E_ARG, "invalid arguments");
        if (words > c->route_words) throw GrError(GR_E_ARG, "more words than the route storage holds");
        download(routes, c->routes, sizeof(int) * size_t(words), c->stream);
    });
}

int gr_ripup(gr_ctx* c, int threshold, unsigned char* is_of_net) {
    return guarded([&] {
        if (!c || !is_of_net) throw GrError(GR_E_ARG, "null argument");
        if (!c->routes) throw GrError(GR_E_STATE, "no route storage");
        if (!c->of_sum) c->of_sum = c->mem.alloc<int>(c->cells);
        if (!c->of_flags) c->of_flags = c->mem.alloc<unsigned char>(c->net_count);
        RipupArgs a{c->g, c->capacity, c->demand, c->of_sum, threshold};
        hipLaunchKernelGGL(overflow_prefix_kernel, dim3(c->tracks), dim3(256), 0, c->stream, a);
        if (c->net_count)
            hipLaunchKernelGGL(overflow_nets_kernel, dim3((c->net_count + 255) / 256), dim3(256), 0, c->stream, c->g,
                               c->of_sum, RouteStore{c->routes, c->pin_acc_num, c->route_per_pin},
                               RouteStore{c->routes2, c->pin_off2, c->route2_per_pin}, c->live2, c->net_count,
                               c->of_flags);
        hip_check(hipGetLastError(), "ripup kernels");
        download(is_of_net, c->of_flags, size_t(c->net_count), c->stream);
    });
}

int gr_zshape_upload(gr_ctx* c, int nbatch, const int* batch_off, const int* net_ids, const long long* rec_off,
                     const int* records, const int* raw_cap) {
    return guarded([&] {
        if (!c || nbatch < 0 || !batch_off || !rec_off) throw GrError(GR_E_ARG, "null argument");
        if (!c->routes) throw GrError(GR_E_STATE, "no route storage");
        const DeviceGrid& g = c->g;
        const int n = std::max(g.X, g.Y), nn = n * n, L = g.L;
        if (batch_off[0] != 0 || rec_off[0] != 0) throw GrError(GR_E_ARG, "offsets must start at 0");
        const int nets = batch_off[nbatch];
        const long long words = rec_off[nbatch];
        if ((nets && (!net_ids || !raw_cap)) || (words && !records)) throw GrError(GR_E_ARG, "null argument");
        long long max_raw = 1;
        // every index the kernels derive from the plan is checked here
        for (int b = 0; b < nbatch; ++b) {
            const int begin = batch_off[b], count = batch_off[b + 1] - begin;
            const long long r0 = rec_off[b], r1 = rec_off[b + 1];
            if (count < 0 || r1 < r0 + count) throw GrError(GR_E_ARG, "batch offsets out of order");
            long long raw = 0;
            for (int k = 0; k < count; ++k) raw += raw_cap[begin + k];
            max_raw = std::max(max_raw, raw);
            for (int k = 0; k < count; ++k) {
                const int id = net_ids[begin + k];
                if (id < 0 || id >= c->net_count) throw GrError(GR_E_ARG, "net id outside route storage");
                if (raw_cap[begin + k] < 2) throw GrError(GR_E_ARG, "raw capacity too small");
                const long long base = r0 + records[r0 + k];
                if (records[r0 + k] < count || base + 2 > r1) throw GrError(GR_E_ARG, "record offset outside the batch");
                const int* rec = records + base;
                const int nodes = rec[1];
                if (nodes < 1 || base + 2 + 6ll * nodes > r1) throw GrError(GR_E_ARG, "record length outside the batch");
                if (rec[0] < 0 || rec[0] + raw_cap[begin + k] > raw) throw GrError(GR_E_ARG, "raw offset outside the batch");
                for (int i = 0; i < nodes; ++i) {
                    const int* r = rec + 2 + 6 * i;
                    if (r[0] < 0 || r[0] >= nn || r[0] / n >= g.X || r[0] % n >= g.Y)
                        throw GrError(GR_E_ARG, "node position outside the grid");
                    if (!((r[1] == -1 && r[2] == -1) || (r[1] >= 1 && r[1] <= L && r[2] >= r[1] && r[2] <= L)))
                        throw GrError(GR_E_ARG, "node layer outside the routing layers");
                    for (int j = 3; j < 6; ++j)
                        if (r[j] != -1 && (r[j] < 6 * (i + 1) || r[j] >= 6 * nodes || r[j] % 6))
                            throw GrError(GR_E_ARG, "child record must follow its parent in preorder");
                }
            }
        }
        c->z_ws.carve([&](auto& cv) {
            c->z_net_ids = cv.template take<int>(nets);
            c->z_records = cv.template take<int>(words);
            c->z_raw_cap = cv.template take<int>(nets);
            c->z_raw = cv.template take<int>(max_raw);
        });
        if (nets) {
            upload(c->z_net_ids, net_ids, sizeof(int) * nets, c->stream);
            upload(c->z_raw_cap, raw_cap, sizeof(int) * nets, c->stream);
        }
        if (words) upload(c->z_records, records, sizeof(int) * words, c->stream);
        const long long plane = (long long)(L + 1) * nn;
        if (plane != c->z_plane) {
            c->mem.release(c->z_map);
            c->mem.release(c->z_prev);
            c->z_map = c->mem.alloc<int>(6 * plane);
            c->z_prev = c->mem.alloc<int>(16 * plane);
            hip_check(hipMemsetAsync(c->z_map, 0, sizeof(int) * 6 * plane, c->stream), "clear Z-shape map");
            hip_check(hipMemsetAsync(c->z_prev, 0, sizeof(int) * 16 * plane, c->stream), "clear Z-shape map");
            c->z_plane = plane;
        }
        c->z_batches = nbatch;
        c->z_batch_off.assign(batch_off, batch_off + nbatch + 1);
        c->z_rec_off.assign(rec_off, rec_off + nbatch + 1);
        hip_check(hipStreamSynchronize(c->stream), "Z-shape upload");
    });
}

int gr_zshape_run(gr_ctx* c, int first_stamp, double* elapsed_ms) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (!c->z_map) throw GrError(GR_E_STATE, "no Z-shape plan uploaded");
        const DeviceGrid& g = c->g;
        ZshapeArgs a;
        a.g = g;
        a.presum = c->presum;
        a.vcost = c->vcost;
        a.map = c->z_map;
        a.prev = c->z_prev;
        a.raw = c->z_raw;
        a.n = std::max(g.X, g.Y);
        a.layers = g.L + 1;
        a.dirv = g.DIR ^ 1;
        a.error = c->error;
        const RouteStore rs{c->routes, c->pin_acc_num, c->route_per_pin};
        hip_check(hipDeviceSynchronize(), "before DAG interval");
        const auto t0 = std::chrono::steady_clock::now();
        for (int b = 0; b < c->z_batches; ++b) {
            const int begin = c->z_batch_off[b], count = c->z_batch_off[b + 1] - begin;
            launch_refresh(c, 0);
            if (!count) continue;
            a.points = c->z_records + c->z_rec_off[b];
            a.count = count;
            a.raw_cap = c->z_raw_cap + begin;
            timed(c, GR_K_ZSHAPE, 0.0, [&] {
                hipLaunchKernelGGL(zshape_reset_kernel, dim3(count), dim3(256), 0, c->stream, a);
                hipLaunchKernelGGL(zshape_solve_kernel, dim3(count), dim3(kWave), 0, c->stream, a);
                hipLaunchKernelGGL(zshape_convert_kernel, dim3((count + 255) / 256), dim3(256), 0, c->stream, a,
                                   (const int*)(c->z_net_ids + begin), rs);
            });
            launch_commit(c, begin, count, first_stamp + b, 1, c->z_net_ids);
            hip_check(hipGetLastError(), "Z-shape kernels");
        }
        hip_check(hipStreamSynchronize(c->stream), "DAG interval");
        const auto t1 = std::chrono::steady_clock::now();
        c->last_t0 = t0, c->last_t1 = t1;
        if (elapsed_ms) *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        check_device_error(c, "gr_zshape_run");
    });
}

int gr_ctx_set_routes_phase2(gr_ctx* c, const int* pin_off2, long long pins2, int words_per_pin) {
    return guarded([&] {
        if (!c || !pin_off2 || pins2 < 0) throw GrError(GR_E_ARG, "invalid phase-two route arguments");
        if (!c->routes) throw GrError(GR_E_STATE, "no route storage");
        if (words_per_pin < 8) throw GrError(GR_E_ARG, "phase-two route words per pin must be at least 8");
        for (int i = 0; i < c->net_count; ++i) {
            const int off = pin_off2[i];
            const int pins = c->pin_acc_num_h[i + 1] - c->pin_acc_num_h[i];
            if (off < -1 || (off >= 0 && (long long)off + pins > pins2))
                throw GrError(GR_E_ARG, "phase-two pin offset outside the storage");
        }
        c->release_phase2();
        c->route2_per_pin = words_per_pin;
        c->pin_off2_h.assign(pin_off2, pin_off2 + c->net_count);
        c->live2_h.assign(c->net_count, 0);
        c->route2_words = (long long)words_per_pin * pins2 + 1;
        c->pin_off2 = c->mem.alloc<int>(std::max(c->net_count, 1));
        if (c->net_count) upload(c->pin_off2, pin_off2, sizeof(int) * c->net_count, c->stream);
        c->live2 = c->mem.alloc<unsigned char>(std::max(c->net_count, 1));
        hip_check(hipMemsetAsync(c->live2, 0, std::max(c->net_count, 1), c->stream), "clear live flags");
        c->routes2 = c->mem.alloc<int>(size_t(c->route2_words));
        hip_check(hipMemsetAsync(c->routes2, 0, sizeof(int) * size_t(c->route2_words), c->stream), "clear routes");
        hip_check(hipStreamSynchronize(c->stream), "phase-two route storage");
    });
}

int gr_ctx_reserve_routes_phase2(gr_ctx* c, int count, const int* net_ids, int words_per_pin, int* pin_off2,
                                 long long* pins2) {
    return guarded([&] {
        if (!c || count < 0 || (count && !net_ids)) throw GrError(GR_E_ARG, "invalid phase-two reservation arguments");
        if (!c->routes) throw GrError(GR_E_STATE, "no route storage");
        if (words_per_pin < 8) throw GrError(GR_E_ARG, "phase-two route words per pin must be at least 8");
        const int n = c->net_count;
        const bool had = c->routes2 != nullptr;
        const int old_stride = had ? c->route2_per_pin : words_per_pin;
        const int stride = std::max(old_stride, words_per_pin);
        const long long old_pins = had ? (c->route2_words - 1) / old_stride : 0;
        std::vector<int> off = had ? c->pin_off2_h : std::vector<int>(size_t(n), -1);
        long long pins = old_pins;
        for (int k = 0; k < count; ++k) {
            const int id = net_ids[k];
            if (id < 0 || id >= n) throw GrError(GR_E_ARG, "net id outside route storage");
            if (off[id] >= 0) continue;
            off[id] = int(pins);
            pins += c->pin_acc_num_h[id + 1] - c->pin_acc_num_h[id];
            if (pins > INT32_MAX) throw GrError(GR_E_ARG, "too many phase-two pins");
        }
        if (!had || pins != old_pins || stride != old_stride) {
            // new storage at the wider stride; kept nets keep their offsets
            // in pins, their words (and live flags) carried over
            std::vector<int> kept;
            if (had)
                for (int id = 0; id < n; ++id)
                    if (c->pin_off2_h[id] >= 0) kept.push_back(id);
            const long long words = (long long)stride * pins + 1;
            int* routes = c->mem.alloc<int>(size_t(words));
            int* offsets = c->mem.alloc<int>(std::max(n, 1));
            hip_check(hipMemsetAsync(routes, 0, sizeof(int) * size_t(words), c->stream), "clear routes");
            if (n) upload(offsets, off.data(), sizeof(int) * n, c->stream);
            if (had && stride == old_stride) {
                hip_check(hipMemcpyAsync(routes, c->routes2, sizeof(int) * size_t(c->route2_words),
                                         hipMemcpyDeviceToDevice, c->stream),
                          "carry phase-two routes");
            } else if (!kept.empty()) {
                int* ids = c->mem.alloc<int>(kept.size());
                upload(ids, kept.data(), sizeof(int) * kept.size(), c->stream);
                hipLaunchKernelGGL(restride_routes_kernel, dim3(unsigned(kept.size())), dim3(256), 0, c->stream,
                                   c->routes2, routes, ids, offsets, c->pin_acc_num, old_stride, stride);
                hip_check(hipGetLastError(), "restride kernel");
                hip_check(hipStreamSynchronize(c->stream), "restride phase-two routes");
                c->mem.release(ids);
            }
            hip_check(hipStreamSynchronize(c->stream), "phase-two route storage");
            std::vector<unsigned char> live = had ? c->live2_h : std::vector<unsigned char>(size_t(n), 0);
            c->release_phase2();
            c->routes2 = routes;
            c->pin_off2 = offsets;
            c->route2_per_pin = stride;
            c->route2_words = words;
            c->pin_off2_h = std::move(off);
            c->live2_h = std::move(live);
            c->live2 = c->mem.alloc<unsigned char>(std::max(n, 1));
            if (n) upload(c->live2, c->live2_h.data(), size_t(n), c->stream);
            hip_check(hipStreamSynchronize(c->stream), "phase-two live flags");
        }
        if (pin_off2 && n) std::copy(c->pin_off2_h.begin(), c->pin_off2_h.end(), pin_off2);
        if (pins2) *pins2 = (c->route2_words - 1) / c->route2_per_pin;
    });
}

int gr_routes2_download(gr_ctx* c, int* routes, long long words) {
    return guarded([&] {
        if (!c || !routes || words < 0) throw GrError(GR_E_ARG, "invalid arguments");
        if (words > c->route2_words) throw GrError(GR_E_ARG, "more words than the phase-two storage holds");
        download(routes, c->routes2, sizeof(int) * size_t(words), c->stream);
    });
}

int gr_congestion_view(gr_ctx* c, unsigned char* congested, float* xsum, float* ysum) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        const DeviceGrid& g = c->g;
        if (!c->cview) {
            c->cview = c->mem.alloc<unsigned char>(g.XY);
            c->cxsum = c->mem.alloc<float>(g.XY);
            c->cysum = c->mem.alloc<float>(g.XY);
        }
        hipLaunchKernelGGL(congestion_view_kernel, dim3((g.XY + 255) / 256), dim3(256), 0, c->stream, g, c->capacity,
                           c->demand, c->cview, c->cxsum, c->cysum);
        hipLaunchKernelGGL(congestion_prefix_kernel, dim3(g.Y), dim3(256), sizeof(float) * g.X, c->stream, g.Y, g.X, 1,
                           g.Y, c->cxsum);
        hipLaunchKernelGGL(congestion_prefix_kernel, dim3(g.X), dim3(256), sizeof(float) * g.Y, c->stream, g.X, g.Y, g.Y,
                           1, c->cysum);
        hip_check(hipGetLastError(), "congestion view kernels");
        if (congested) download(congested, c->cview, g.XY, c->stream);
        if (xsum) download(xsum, c->cxsum, sizeof(float) * g.XY, c->stream);
        if (ysum) download(ysum, c->cysum, sizeof(float) * g.XY, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "congestion view");
    });
}

int gr_flt_upload(gr_ctx* c, int nbatch, const int* batch_off, const int* net_ids, const long long* node_off,
                  const int* nodes, const long long* edge_off, const int* edges) {
    return guarded([&] {
        if (!c || nbatch < 0 || !batch_off || !node_off || !edge_off) throw GrError(GR_E_ARG, "null argument");
        if (!c->routes2) throw GrError(GR_E_STATE, "no phase-two route storage");
        const DeviceGrid& g = c->g;
        if (g.L > 9) throw GrError(GR_E_ARG, "FLT supports at most 9 routing layers");
        if (g.X > 65535 || g.Y > 65535) throw GrError(GR_E_ARG, "FLT packs node positions in 16 bits per axis");
        if (batch_off[0] != 0 || node_off[0] != 0 || edge_off[0] != 0) throw GrError(GR_E_ARG, "offsets must start at 0");
        for (int b = 0; b < nbatch; ++b)
            if (batch_off[b + 1] < batch_off[b]) throw GrError(GR_E_ARG, "batch offsets out of order");
        const int nets = batch_off[nbatch];
        if (nets && !net_ids) throw GrError(GR_E_ARG, "null net ids");
        for (int k = 0; k < nets; ++k)
            if (node_off[k + 1] < node_off[k] || edge_off[k + 1] < edge_off[k])
                throw GrError(GR_E_ARG, "FLT record offsets out of order");
        const long long total_nodes = node_off[nets], total_edges = edge_off[nets];
        if ((total_nodes && !nodes) || (total_edges && !edges)) throw GrError(GR_E_ARG, "null records");
        if (total_nodes > INT32_MAX / 16 || total_edges > INT32_MAX / 16) throw GrError(GR_E_ARG, "FLT stage too large");
        const int max_coord = (g.L + 1) * g.XY;
        const int L = g.L;

        // pass 1, per net (threads): node records validated, depth and slot counts
        std::vector<int> net_depths(nets), net_slots(nets);
        parallel_items(nets, [&](int k) {
            const int id = net_ids[k];
            if (id < 0 || id >= c->net_count) throw GrError(GR_E_ARG, "net id outside route storage");
            if (c->pin_off2_h[id] < 0) throw GrError(GR_E_ARG, "net without phase-two route storage");
            const long long n0 = node_off[k], nn = node_off[k + 1] - n0, ne = edge_off[k + 1] - edge_off[k];
            if (nn < 1) throw GrError(GR_E_ARG, "FLT net without nodes");
            int slot = 0, roots = 0, prev_end = 0, maxdepth = 0;
            for (long long i = 0; i < nn; ++i) {
                const int* r = nodes + 4 * (n0 + i);
                if (r[0] < 0 || r[0] >= max_coord) throw GrError(GR_E_ARG, "FLT node outside the grid");
                if (r[1] < 0 || r[1] >= 5000) throw GrError(GR_E_ARG, "FLT depth outside [0, 5000)");
                if (r[2] < 0 || r[2] > 6) throw GrError(GR_E_ARG, "FLT node with more than 6 child slots");
                if (r[3] < prev_end || r[3] > ne) throw GrError(GR_E_ARG, "FLT parent_end out of order");
                if (r[3] == prev_end) {
                    ++roots;
                    if (r[1] != 0) throw GrError(GR_E_ARG, "FLT node without parents below depth 0");
                }
                maxdepth = std::max(maxdepth, r[1]);
                slot += r[2];
                prev_end = r[3];
            }
            if (prev_end != ne) throw GrError(GR_E_ARG, "FLT edge count differs from the last parent_end");
            if (roots != 1) throw GrError(GR_E_ARG, "FLT DAG needs exactly one root");
            net_depths[k] = maxdepth + 1;
            net_slots[k] = slot;
        });
        std::vector<long long> dbase(nets + 1, 0), sbase(nets + 1, 0);
        for (int k = 0; k < nets; ++k) {
            dbase[k + 1] = dbase[k] + net_depths[k] + 1;
            sbase[k + 1] = sbase[k] + net_slots[k];
        }
        if (sbase[nets] > INT32_MAX / 16 || dbase[nets] > INT32_MAX) throw GrError(GR_E_ARG, "FLT stage too large");

        // the device workspace layout, staged in pinned host memory so the
        // whole stage crosses in one copy
        const size_t ndoff = size_t(dbase[nets]);
        struct Stage {
            int4 *nrec, *erec, *desc, *sizes;
            int *order, *doff, *eorder, *edoff, *ids, *far;
        };
        auto layout = [&](auto& cv, Stage& st) {
            st.nrec = cv.template take<int4>(size_t(total_nodes));
            st.erec = cv.template take<int4>(size_t(total_edges));
            st.order = cv.template take<int>(size_t(total_nodes));
            st.doff = cv.template take<int>(ndoff);
            st.eorder = cv.template take<int>(size_t(total_edges));
            st.edoff = cv.template take<int>(ndoff);
            st.desc = cv.template take<int4>(nets);
            st.sizes = cv.template take<int4>(nets);
            st.ids = cv.template take<int>(nets);
            st.far = cv.template take<int>(size_t(total_edges));  // used up to the far-edge count
        };
        Carver dry{nullptr};
        {
            Stage unused{};
            layout(dry, unused);
        }
        const size_t stage_bytes = dry.off;
        if (stage_bytes > c->f_stage_cap) {
            if (c->f_stage) (void)hipHostFree(c->f_stage);
            c->f_stage = nullptr;
            c->f_stage_cap = 0;
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->f_stage), stage_bytes + stage_bytes / 4), "hipHostMalloc");
            c->f_stage_cap = stage_bytes + stage_bytes / 4;
        }
        Carver host{c->f_stage};
        Stage hs{};
        layout(host, hs);
        int4 *nrec = hs.nrec, *erec = hs.erec, *desc = hs.desc, *sizes = hs.sizes;
        int *order = hs.order, *doff = hs.doff, *ids = hs.ids;

        // pass 2, per net (threads): records, edges validated, nodes by
        // depth (stable, the reference's radix sort on (net, depth)), bytes
        std::vector<double> net_eb(nets), net_rb(nets);
        std::vector<int> net_dnodes(nets);
        std::vector<char> big(nets);  // the DAG's LDS share exceeds f_lds_split
        std::vector<int> net_far(nets);  // edges longer than 4
        std::vector<char> edge_far(static_cast<size_t>(total_edges));
        parallel_items(nets, [&](int k) {
            const long long n0 = node_off[k], e0 = edge_off[k];
            const long long nn = node_off[k + 1] - n0, ne = edge_off[k + 1] - e0;
            int slot = 0, prev_end = 0;
            for (long long i = 0; i < nn; ++i) {
                const int* r = nodes + 4 * (n0 + i);
                nrec[n0 + i] = make_int4(r[0], prev_end, r[3], slot * 8 + r[2]);
                slot += r[2];
                prev_end = r[3];
            }
            double eb = 0;
            for (long long i = 0; i < nn; ++i) {
                const int4 nr = nrec[n0 + i];
                const int di = nodes[4 * (n0 + i) + 1];
                for (int e = nr.y; e < nr.z; ++e) {
                    const int* r = edges + 2 * (e0 + e);
                    const int par = r[0], s2 = r[1] % 10;
                    if (par < 0 || par >= nn || r[1] / 10 != i) throw GrError(GR_E_ARG, "FLT edge record inconsistent");
                    if (nodes[4 * (n0 + par) + 1] >= di) throw GrError(GR_E_ARG, "FLT parent not shallower than its child");
                    if (s2 >= nodes[4 * (n0 + par) + 2]) throw GrError(GR_E_ARG, "FLT edge slot outside its parent");
                    const int pc = nrec[n0 + par].x, cc = nr.x;
                    const int pcx = pc / g.Y % g.X, pcy = pc % g.Y, ccx = cc / g.Y % g.X, ccy = cc % g.Y;
                    if (pcx != ccx && pcy != ccy) throw GrError(GR_E_ARG, "FLT edge not straight");
                    erec[e0 + e] = make_int4(par, s2, pc, cc);
                    const int len = std::abs(pcx - ccx) + std::abs(pcy - ccy);
                    edge_far[e0 + e] = len > 4;
                    if (len > 4) eb += 12.0 * L * (len + 1), ++net_far[k];  // shorter edges' change tables are never read
                }
            }
            const int depths = net_depths[k];
            int* dk = doff + dbase[k];
            for (int d = 0; d <= depths; ++d) dk[d] = 0;
            for (long long i = 0; i < nn; ++i) ++dk[nodes[4 * (n0 + i) + 1] + 1];
            for (int d = 0; d < depths; ++d) dk[d + 1] += dk[d];
            // dk[d] is the first position of depth d; place, then restore
            for (long long i = 0; i < nn; ++i) order[n0 + dk[nodes[4 * (n0 + i) + 1]]++] = int(i);
            for (int d = depths; d > 0; --d) dk[d] = dk[d - 1];
            dk[0] = 0;
            // edges by the depth of their child, in node order within a depth
            int* ek = hs.edoff + dbase[k];
            int dn_max = 0, ne_at = 0;
            for (int d = 0; d < depths; ++d) {
                ek[d] = ne_at;
                dn_max = std::max(dn_max, dk[d + 1] - dk[d]);
                for (int q = dk[d]; q < dk[d + 1]; ++q) {
                    const int4 nr = nrec[n0 + order[n0 + q]];
                    for (int e = nr.y; e < nr.z; ++e) hs.eorder[e0 + ne_at++] = e;
                }
            }
            ek[depths] = ne_at;
            net_dnodes[k] = dn_max;
            big[k] = flt_lds_bytes(L, slot, int(nn), int(ne), dn_max, depths) > size_t(c->f_lds_split);
            desc[k] = make_int4(int(n0), int(e0), int(sbase[k]), int(dbase[k]));
            sizes[k] = make_int4(int(nn), int(ne), slot, depths);
            ids[k] = net_ids[k];
            net_eb[k] = eb;
            // the FLT unit cost of one DAG in the DP: per node 16 + 8 * L
            // (record, layer costs), per edge 8 (its record)
            net_rb[k] = nn * (16.0 + 8.0 * L) + 8.0 * ne;
        });
        // the records cross to the device while the host orders the batches
        // below; the per-net descriptors, last in the layout, follow them
        Stage dev{};
        c->f_ws.carve([&](auto& cv) { layout(cv, dev); });
        const size_t bulk = size_t(reinterpret_cast<char*>(hs.desc) - c->f_stage);
        if (bulk) upload(dev.nrec, c->f_stage, bulk, c->stream);

        // per batch: stage-global bases, the table sizes and the algorithmic
        // bytes per launch (the layer-change search, 12 * L * (len + 1) per
        // edge longer than 4 for the presum and via cost gathers along it;
        // the per-edge change tables are scratch and not counted)
        long long max_nodes = 1, max_edges = 1, max_slots = 1;
        c->f_base.assign(nbatch, make_int4(0, 0, 0, 0));
        c->f_batch_max.assign(nbatch, FltLdsShape{0, 0, 0, 0, 0});
        c->f_edge_bytes.assign(nbatch, 0.0);
        c->f_route_bytes.assign(nbatch, 0.0);
        for (int b = 0; b < nbatch; ++b) {
            const int k0 = batch_off[b], k1 = batch_off[b + 1];
            if (k0 == k1) continue;
            const long long nodes_b = node_off[k1] - node_off[k0], edges_b = edge_off[k1] - edge_off[k0];
            const long long slots_b = sbase[k1] - sbase[k0];
            max_nodes = std::max(max_nodes, nodes_b);
            max_edges = std::max(max_edges, edges_b);
            max_slots = std::max(max_slots, slots_b);
            c->f_base[b] = make_int4(desc[k0].x, desc[k0].y, desc[k0].z, int(edges_b));
            for (int k = k0; k < k1; ++k) {
                FltLdsShape& m = c->f_batch_max[b];
                m.slots = std::max(m.slots, sizes[k].z);
                m.nodes = std::max(m.nodes, sizes[k].x);
                m.edges = std::max(m.edges, sizes[k].y);
                m.dnodes = std::max(m.dnodes, net_dnodes[k]);
                m.depths = std::max(m.depths, sizes[k].w);
                c->f_edge_bytes[b] += net_eb[k];
                c->f_route_bytes[b] += net_rb[k];
            }
        }
        // DAG classes by LDS share: the DAGs above f_lds_split bytes first,
        // each class launched with shares sized by its own largest DAG, so
        // the many small DAGs of a batch are not held to its largest one's
        // share.  Inside a class the DAGs of a wavefront step through their
        // depths together: decreasing depth, then size (results do not
        // depend on the order of nets inside a batch).
        // the far-edge list: per batch its edges longer than 4, in edge order
        c->f_far_off.assign(nbatch + 1, 0);
        for (int b = 0; b < nbatch; ++b) {
            long long nf = 0;
            for (int k = batch_off[b]; k < batch_off[b + 1]; ++k) nf += net_far[k];
            c->f_far_off[b + 1] = c->f_far_off[b] + int(nf);
        }
        parallel_tasks(nbatch, [&](int b) {
            int* out = hs.far + c->f_far_off[b];
            for (int k = batch_off[b]; k < batch_off[b + 1]; ++k) {
                if (!net_far[k]) continue;
                for (long long e = edge_off[k]; e < edge_off[k + 1]; ++e)
                    if (edge_far[e]) *out++ = int(e);
            }
        });
        c->f_large.assign(nbatch, 0);
        c->f_class_max.assign(size_t(nbatch) * kFltClasses, FltLdsShape{0, 0, 0, 0, 0});
        parallel_tasks(nbatch, [&](int b) {
            std::vector<int> perm;
            std::vector<int4> dp, sp;
            std::vector<int> ip;
            {
                const int k0 = batch_off[b], k1 = batch_off[b + 1];
                perm.resize(k1 - k0);
                for (int i = 0; i < k1 - k0; ++i) perm[i] = k0 + i;
                std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
                    if (big[x] != big[y]) return big[x] > big[y];
                    return sizes[x].w != sizes[y].w ? sizes[x].w > sizes[y].w : sizes[x].x > sizes[y].x;
                });
                dp.resize(k1 - k0), sp.resize(k1 - k0), ip.resize(k1 - k0);
                for (int i = 0; i < k1 - k0; ++i) {
                    const int k = perm[i];
                    dp[i] = desc[k], sp[i] = sizes[k], ip[i] = ids[k];
                    c->f_large[b] += big[k];
                    FltLdsShape& m = c->f_class_max[size_t(b) * kFltClasses + (big[k] ? 0 : 1)];
                    m.slots = std::max(m.slots, sizes[k].z);
                    m.nodes = std::max(m.nodes, sizes[k].x);
                    m.edges = std::max(m.edges, sizes[k].y);
                    m.dnodes = std::max(m.dnodes, net_dnodes[k]);
                    m.depths = std::max(m.depths, sizes[k].w);
                }
                for (int i = 0; i < k1 - k0; ++i) desc[k0 + i] = dp[i], sizes[k0 + i] = sp[i], ids[k0 + i] = ip[i];
            }
        });
        if (c->f_queue_cap < nbatch) {
            c->drop_graph();
            c->mem.release(c->f_queue);
            c->f_queue = c->mem.alloc<int>(size_t(std::max(nbatch, 1)) * kFltClasses * kFltQueues * kFltQueueStride);
            c->f_queue_cap = nbatch;
        }
        FltArgs& f = c->f;
        c->f_tables.carve([&](auto& cv) {
            f.costs = cv.template take<double>(size_t(max_nodes) * kFltStride);
            f.packed = cv.template take<unsigned long long>(size_t(max_slots) * kFltStride);
            f.best_path = cv.template take<unsigned long long>(size_t(max_slots) * kFltStride);
            f.layer_output = cv.template take<int>(size_t(max_nodes));
            f.cost_edges = cv.template take<double>(size_t(max_edges) * 9 * kFltChangeWidth);
            f.best_change = cv.template take<int>(size_t(max_edges) * 9 * kFltChangeWidth);
        });
        // the descriptors and the used part of the far-edge list
        const size_t tail = size_t(reinterpret_cast<char*>(hs.far + c->f_far_off[nbatch]) - c->f_stage);
        if (tail > bulk) upload(reinterpret_cast<char*>(dev.nrec) + bulk, c->f_stage + bulk, tail - bulk, c->stream);
        f.g = g;
        f.node_rec = dev.nrec;
        f.edge_rec = dev.erec;
        f.order = dev.order;
        f.depth_off = dev.doff;
        f.eorder = dev.eorder;
        f.edepth_off = dev.edoff;
        f.desc = dev.desc;
        f.sizes = dev.sizes;
        f.net_ids = dev.ids;
        f.far_edges = dev.far;
        f.routes = RouteStore{c->routes2, c->pin_off2, c->route2_per_pin};
        f.pin_acc_num = c->pin_acc_num;
        f.error = c->error;
        c->drop_fgraph();  // the workspaces may have moved
        ++c->flt_gen;
        c->f_batches = nbatch;
        c->f_batch_off.assign(batch_off, batch_off + nbatch + 1);
        c->f_ids_h.assign(net_ids, net_ids + nets);
        hip_check(hipStreamSynchronize(c->stream), "FLT upload");
    });
}

int gr_flt_run(gr_ctx* c, int first_stamp, int remove, int mode, double* elapsed_ms) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (mode < 0 || mode > 2)
            throw GrError(GR_E_ARG, "FLT execution mode not available (0 = plain, 1 = launches, 2 = persistent + graph)");
        if (remove < 0 || remove > 2)
            throw GrError(GR_E_ARG, "FLT removal must be 0 (none), 1 (phase-one routes) or 2 (current routes)");
        if (!c->f.node_rec && c->f_batches) throw GrError(GR_E_STATE, "no FLT stage uploaded");
        bind_flt_storage(c);
        const RouteStore rs2{c->routes2, c->pin_off2, c->route2_per_pin};
        auto enqueue = [&] {
            int stamp = first_stamp;
            const long long qwords = (long long)c->f_batches * kFltClasses * kFltQueues * kFltQueueStride;
            if (mode == 2 && qwords)  // a kernel, not a memset: the loop may be captured
                hipLaunchKernelGGL(fill_int_kernel, dim3(unsigned((qwords + 255) / 256)), dim3(256), 0, c->stream,
                                   c->f_queue, 0, qwords);
            for (int b = 0; b < c->f_batches; ++b) {
                const int begin = c->f_batch_off[b], count = c->f_batch_off[b + 1] - begin;
                const int* ids = c->f.net_ids + begin;
                if (remove) launch_commit(c, 0, count, stamp++, -1, ids, nullptr, remove == 2);
                const int s = stamp++;
                launch_refresh(c, 0);
                if (!count) continue;
                const FltArgs a = flt_batch_args(c, b);
                timed(c, GR_K_FLT_EDGE, c->f_edge_bytes[b], [&] { launch_flt_edges(a, c->stream); });
                int* queue = mode == 2 ? c->f_queue + size_t(b) * kFltClasses * kFltQueues * kFltQueueStride : nullptr;
                timed(c, GR_K_FLT, c->f_route_bytes[b], [&] {
                    if (mode == 0) {
                        // plain: one DAG per wavefront, the batch tables in global memory
                        FltArgs p = a;
                        p.per_wave = 1;
                        launch_flt_route(p, false, c->cus, c->stream);
                    } else {
                        launch_flt_dp(c, a, b, c->f_per_wave[count >= kFltWideBatch], queue);
                    }
                });
                hip_check(hipGetLastError(), "FLT kernels");
                launch_commit(c, 0, count, s, 1, ids, &rs2);
            }
        };
        // a captured graph cannot carry profiling events: profiled runs of
        // mode 2 issue the same launches directly
        const bool graph = mode == 2 && !c->profiling;
        if (graph) {
            const std::vector<int> key{c->flt_gen, first_stamp, remove, c->f_per_wave[0], c->f_per_wave[1]};
            if (!c->fgraph_exec || key != c->fgraph_key) {
                c->drop_fgraph();
                hip_check(hipStreamSynchronize(c->stream), "before graph capture");
                hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try {
                    enqueue();
                } catch (...) {
                    hipGraph_t broken = nullptr;
                    (void)hipStreamEndCapture(c->stream, &broken);
                    if (broken) (void)hipGraphDestroy(broken);
                    throw;
                }
                hip_check(hipStreamEndCapture(c->stream, &c->fgraph), "hipStreamEndCapture");
                hip_check(hipGraphInstantiate(&c->fgraph_exec, c->fgraph, nullptr, nullptr, 0), "hipGraphInstantiate");
                c->fgraph_key = key;
            }
        }
        hip_check(hipDeviceSynchronize(), "before DAG interval");
        const auto t0 = std::chrono::steady_clock::now();
        if (graph) hip_check(hipGraphLaunch(c->fgraph_exec, c->stream), "hipGraphLaunch");
        else enqueue();
        hip_check(hipStreamSynchronize(c->stream), "DAG interval");
        const auto t1 = std::chrono::steady_clock::now();
        c->last_t0 = t0, c->last_t1 = t1;
        if (elapsed_ms) *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        check_device_error(c, "gr_flt_run");
        // the stage's nets now hold their phase-two routes
        bool changed = false;
        for (int id : c->f_ids_h) changed |= !c->live2_h[id], c->live2_h[id] = 1;
        if (changed) {
            upload(c->live2, c->live2_h.data(), size_t(c->net_count), c->stream);
            hip_check(hipStreamSynchronize(c->stream), "phase-two live flags");
        }
    });
}

int gr_flt_tune(gr_ctx* c, int* per_wave, double* ms) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (!c->f.node_rec && c->f_batches) throw GrError(GR_E_STATE, "no FLT stage uploaded");
        bind_flt_storage(c);
        int* saved = c->mem.alloc<int>(size_t(c->route2_words));
        hip_check(hipMemcpyAsync(saved, c->routes2, sizeof(int) * size_t(c->route2_words), hipMemcpyDeviceToDevice,
                                 c->stream),
                  "save phase-two routes");
        launch_refresh(c, 0);
        const bool profiling = c->profiling;
        c->profiling = false;
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        const int candidates[4] = {1, 2, 4, 8};
        double t[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        bool seen[2] = {false, false};
        for (int b = 0; b < c->f_batches; ++b) {
            const int begin = c->f_batch_off[b], count = c->f_batch_off[b + 1] - begin;
            if (!count) continue;
            const FltArgs a = flt_batch_args(c, b);
            launch_flt_edges(a, c->stream);
            const int cls = count >= kFltWideBatch;
            seen[cls] = true;
            for (int ci = 0; ci < 4; ++ci) {
                // the DP of this batch at the current costs, in the launch
                // form; its words are rewritten when the stage runs
                hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
                launch_flt_dp(c, a, b, candidates[ci], nullptr);
                hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float f = 0;
                hip_check(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
                t[cls][ci] += f;
            }
        }
        hip_check(hipGetLastError(), "FLT tuning launches");
        // the tuning runs wrote phase-two words: put back what the stage
        // starts from (earlier stages' routes included)
        hip_check(hipMemcpyAsync(c->routes2, saved, sizeof(int) * size_t(c->route2_words), hipMemcpyDeviceToDevice,
                                 c->stream),
                  "restore phase-two routes");
        hip_check(hipStreamSynchronize(c->stream), "FLT tuning");
        c->mem.release(saved);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->profiling = profiling;
        for (int cls = 0; cls < 2; ++cls) {
            int best = 0;
            for (int ci = 1; ci < 4; ++ci)
                if (t[cls][ci] < t[cls][best]) best = ci;
            if (seen[cls]) c->f_per_wave[cls] = candidates[best];
            if (per_wave) per_wave[cls] = c->f_per_wave[cls];
            if (ms)
                for (int ci = 0; ci < 4; ++ci) ms[cls * 4 + ci] = seen[cls] ? t[cls][ci] : -1.0;
        }
        check_device_error(c, "gr_flt_tune");
    });
}

int gr_ctx_set_flt_per_wave(gr_ctx* c, const int* per_wave) {
    return guarded([&] {
        if (!c || !per_wave) throw GrError(GR_E_ARG, "null argument");
        for (int k = 0; k < 2; ++k)
            if (per_wave[k] != 1 && per_wave[k] != 2 && per_wave[k] != 4 && per_wave[k] != 8)
                throw GrError(GR_E_ARG, "DAGs per wavefront must be 1, 2, 4 or 8");
        c->f_per_wave[0] = per_wave[0];
        c->f_per_wave[1] = per_wave[1];
    });
}

int gr_soft_forward(gr_ctx* c, int begin, int count, double slope, double zero_slope, double via_weight, double tau,
                    int precision, double* total, double* per_net, double* elapsed_ms) {
    return guarded([&] {
        soft_forward(c, begin, count, slope, zero_slope, via_weight, nullptr, nullptr, tau, precision, total, per_net,
                     elapsed_ms);
    });
}

int gr_soft_forward_costs(gr_ctx* c, int begin, int count, const double* wcost, const double* vcost, double tau,
                          int precision, double* total, double* per_net, double* elapsed_ms) {
    return guarded([&] {
        if (!wcost || !vcost) throw GrError(GR_E_ARG, "null cost arrays");
        soft_forward(c, begin, count, 0.0, 0.0, 0.0, wcost, vcost, tau, precision, total, per_net, elapsed_ms);
    });
}

int gr_ctx_set_soft_kind(gr_ctx* c, int kind) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (kind != 0 && kind != 1) throw GrError(GR_E_ARG, "soft DP kind must be 0 (L-shape) or 1 (Z-shape)");
        c->soft_kind = kind;
        c->soft_ready = false;
    });
}

int gr_soft_backward(gr_ctx* c, double* grad_params, double* elapsed_ms) {
    return guarded([&] {
        if (!c || !grad_params) throw GrError(GR_E_ARG, "null argument");
        if (!c->soft_ready) throw GrError(GR_E_STATE, "no soft forward pass to differentiate");
        const size_t cells = c->cells;
        SoftArgs& a = c->soft;
        hip_check(hipMemsetAsync(a.gw, 0, sizeof(double) * cells, c->stream), "clear gradients");
        hip_check(hipMemsetAsync(a.gv, 0, sizeof(double) * cells, c->stream), "clear gradients");
        hip_check(hipMemsetAsync(c->soft_params, 0, sizeof(double) * kStatStripes * kCounterStride, c->stream),
                  "clear parameter gradients");
        hipEvent_t e0 = take_event(c), e1 = take_event(c);
        hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
        if (a.count) launch_soft_backward(a, c->soft_precision, c->soft_kind, c->cus, c->stream);
        const int blocks = int(std::min<size_t>(1024, (cells + 255) / 256));
        if (!c->soft_given)  // the cost-model parameters exist only for model costs
            hipLaunchKernelGGL(soft_params_kernel, dim3(blocks), dim3(256), 0, c->stream, c->soft_cost, a.gw, a.gv,
                               (long long)cells, c->soft_params);
        hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
        hip_check(hipGetLastError(), "soft backward kernels");
        std::vector<double> part(size_t(kStatStripes) * kCounterStride);
        download(part.data(), c->soft_params, sizeof(double) * part.size(), c->stream);
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (elapsed_ms) *elapsed_ms = ms;
        c->next_event -= 2;
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int i = 0; i < kStatStripes; ++i) s += part[size_t(i) * kCounterStride + j];
            grad_params[j] = s;
        }
    });
}

int gr_soft_edge_grads(gr_ctx* c, double* grad_wcost, double* grad_vcost) {
    return guarded([&] {
        if (!c) throw GrError(GR_E_ARG, "null context");
        if (!c->soft_ready) throw GrError(GR_E_STATE, "no soft pass");
        for (auto [dst, src] : {std::pair<double*, double*>{grad_wcost, c->soft.gw}, {grad_vcost, c->soft.gv}}) {
            if (!dst) continue;
            std::vector<double> h(c->cells);
            download(h.data(), src, sizeof(double) * c->cells, c->stream);
            from_track_major(c->g, h, dst);
        }
    });
}

int gr_batch_geometry(gr_ctx* c, int n, const int* pt_off, const int* pts, const int* hs_off, const int* hsegs,
                      const int* vs_off, const int* vsegs) {
    return guarded([&] {
        if (!c || n < 0 || !pt_off || !hs_off || !vs_off) throw GrError(GR_E_ARG, "null argument");
        c->has_geom = false;
        const int XY = c->g.XY;
        auto check_off = [&](const int* off, const char* what) {
            if (off[0] != 0) throw GrError(GR_E_ARG, std::string(what) + " offsets must start at 0");
            for (int k = 0; k < n; ++k)
                if (off[k + 1] < off[k]) throw GrError(GR_E_ARG, std::string(what) + " offsets must not decrease");
        };
        check_off(pt_off, "point");
        check_off(hs_off, "horizontal segment");
        check_off(vs_off, "vertical segment");
        const long long np = pt_off[n], nh = hs_off[n], nv = vs_off[n];
        if ((np && !pts) || (nh && !hsegs) || (nv && !vsegs)) throw GrError(GR_E_ARG, "null geometry array");
        for (long long q = 0; q < np; ++q)
            if (pts[q] < 0 || pts[q] >= XY) throw GrError(GR_E_ARG, "batching point outside the grid");
        for (long long s = 0; s < nh; ++s)
            if (hsegs[2 * s] < 0 || hsegs[2 * s + 1] >= XY) throw GrError(GR_E_ARG, "segment outside the grid");
        for (long long s = 0; s < nv; ++s)
            if (vsegs[2 * s] < 0 || vsegs[2 * s + 1] >= XY) throw GrError(GR_E_ARG, "segment outside the grid");
        BatchGeom& b = c->geom;
        c->geom_ws.carve([&](auto& cv) {
            b.pt_off = cv.template take<int>(n + 1);
            b.pts = cv.template take<int>(np);
            b.hs_off = cv.template take<int>(n + 1);
            b.hsegs = cv.template take<int>(2 * nh);
            b.vs_off = cv.template take<int>(n + 1);
            b.vsegs = cv.template take<int>(2 * nv);
        });
        hipStream_t st = c->stream;
        upload((void*)b.pt_off, pt_off, sizeof(int) * (n + 1), st);
        if (np) upload((void*)b.pts, pts, sizeof(int) * np, st);
        upload((void*)b.hs_off, hs_off, sizeof(int) * (n + 1), st);
        if (nh) upload((void*)b.hsegs, hsegs, sizeof(int) * 2 * nh, st);
        upload((void*)b.vs_off, vs_off, sizeof(int) * (n + 1), st);
        if (nv) upload((void*)b.vsegs, vsegs, sizeof(int) * 2 * nv, st);
        hip_check(hipStreamSynchronize(st), "batching geometry upload");
        b.n = n, b.X = c->g.X, b.Y = c->g.Y, b.sparsity = 1;
        c->geom_points = np, c->geom_hsegs = nh, c->geom_vsegs = nv;
        c->geom_pt_off.assign(pt_off, pt_off + n + 1);
        c->has_geom = true;
    });
}

int gr_batch_colors(gr_ctx* c, int limit, int sparsity, int* color, int* num_batches) {
    return guarded([&] {
        if (!c || !color || !num_batches) throw GrError(GR_E_ARG, "null argument");
        if (!c->has_geom) throw GrError(GR_E_STATE, "no batching geometry (gr_batch_geometry)");
        if (sparsity < 1) throw GrError(GR_E_ARG, "sparsity must be positive");
        if (limit < 1) throw GrError(GR_E_ARG, "batch limit must be positive");
        *num_batches = 0;
        BatchGeom b = c->geom;
        b.sparsity = sparsity;
        // run-head flags of the sorted marks as a scan input, without storing them
        auto run_heads = [](const unsigned* cell, const unsigned* tag, int m) {
            return hipcub::TransformInputIterator<int, RunHead, hipcub::CountingInputIterator<int>>(
                hipcub::CountingInputIterator<int>(0), RunHead{cell, tag, m});
        };
        const int n = b.n;
        if (!n) return;
        const int XY = c->g.XY;
        const long long npts = c->geom_points;
        hipStream_t st = c->stream;

        // marks per net and their offsets
        int *mark_count = nullptr, *mark_off = nullptr, *cand_count = nullptr;
        void* scan_tmp = nullptr;
        size_t scan_bytes = 0;
        hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, mark_count, mark_off, n + 1, st), "scan size");
        c->batch_ws1.carve([&](auto& cv) {
            mark_count = cv.template take<int>(n + 1);
            mark_off = cv.template take<int>(n + 1);
            cand_count = cv.template take<int>(n);
            scan_tmp = cv.template take<char>(scan_bytes);
        });
        hip_check(hipMemsetAsync(mark_count + n, 0, sizeof(int), st), "clear mark count tail");
        const double geom_bytes = 4.0 * npts + 8.0 * (c->geom_hsegs + c->geom_vsegs) + 4.0 * n;
        // the algorithmic bytes charge every candidate mark (repeats included)
        // and are only needed when the colouring launch is being timed
        const bool color_profiled = c->profiling && (c->prof_mask >> GR_K_BATCH_COLOR & 1u);
        timed(c, GR_K_BATCH_MARKS, geom_bytes,
              [&] {
                  hipLaunchKernelGGL(count_marks_kernel, dim3((n + 15) / 16), dim3(256), 0, st, b, mark_count,
                                     color_profiled ? cand_count : nullptr);
              });
        hip_check(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, mark_count, mark_off, n + 1, st), "scan marks");
        int total_marks = 0;
        download(&total_marks, mark_off + n, sizeof(int), st);
        if (total_marks < 0) throw GrError(GR_E_CAPACITY, "more than 2^31 batching marks");
        const int M = total_marks;
        double cand_marks = M;
        if (color_profiled) {
            std::vector<int> cand(n);
            download(cand.data(), cand_count, sizeof(int) * n, st);
            cand_marks = 0;
            for (int v : cand) cand_marks += v;
        }

        // per-cell lists, reader records
        int cell_bits = 1;
        while ((1ll << cell_bits) < XY) ++cell_bits;
        unsigned *cell = nullptr, *tag = nullptr, *cell2 = nullptr, *tag2 = nullptr;
        int *eidx = nullptr, *list_net = nullptr, *einfo = nullptr, *marker = nullptr;
        int *prev_marker = nullptr, *slot_of = nullptr, *rcount = nullptr, *d_color = nullptr, *d_count = nullptr;
        int *d_misc = nullptr, *d_closed = nullptr;
        int4* rec = nullptr;
        void* tmp = nullptr;
        size_t sort_bytes = 0, sum_bytes = 0, max_bytes = 0;
        {
            hipcub::DoubleBuffer<unsigned> k0(cell, cell2), v0(tag, tag2);
            hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, v0, std::max(M, 1), 0, cell_bits, st),
                      "sort size");
            hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, sum_bytes, run_heads(nullptr, nullptr, M), eidx, M + 1, st),
                      "scan size");
            hip_check(hipcub::DeviceScan::ExclusiveScan(nullptr, max_bytes, marker, prev_marker, hipcub::Max(), -1,
                                                        std::max(M, 1), st),
                      "scan size");
        }
        constexpr int kMaxColors = kColorsPerWord * kMaxMaskWords;
        c->batch_ws2.carve([&](auto& cv) {
            cell = cv.template take<unsigned>(M);
            tag = cv.template take<unsigned>(M);
            cell2 = cv.template take<unsigned>(M);
            tag2 = cv.template take<unsigned>(M);
            eidx = cv.template take<int>(M + 1);
            list_net = cv.template take<int>(M);
            einfo = cv.template take<int>(M);
            marker = cv.template take<int>(M);
            prev_marker = cv.template take<int>(M);
            slot_of = cv.template take<int>(M);
            rec = cv.template take<int4>(npts);
            rcount = cv.template take<int>(n);
            d_color = cv.template take<int>(n);
            d_count = cv.template take<int>(kMaxColors);
            d_misc = cv.template take<int>(kColorQueues * kColorQueueStride);  // dequeue counters
            d_closed = cv.template take<int>(2 * kMaxColors);
            tmp = cv.template take<char>(std::max(sort_bytes, std::max(sum_bytes, max_bytes)));
        });
        const int grid_m = M / 256 + 1;
        hip_check(hipMemsetAsync(rcount, 0, sizeof(int) * n, st), "clear reader counts");
        if (M) {
            timed(c, GR_K_BATCH_MARKS, 8.0 * M, [&] {
                hipLaunchKernelGGL(emit_marks_kernel, dim3((n + 15) / 16), dim3(256), 0, st, b, mark_off, cell, tag);
            });
            hipcub::DoubleBuffer<unsigned> keys(cell, cell2), vals(tag, tag2);
            timed(c, GR_K_BATCH_MARKS, 32.0 * M, [&] {
                hip_check(hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, vals, M, 0, cell_bits, st),
                          "sort marks");
            });
            const unsigned* sc = keys.Current();
            const unsigned* sv = vals.Current();
            timed(c, GR_K_BATCH_MARKS, 12.0 * M, [&] {
                hip_check(hipcub::DeviceScan::ExclusiveSum(tmp, sum_bytes, run_heads(sc, sv, M), eidx, M + 1, st),
                          "scan entries");
                hipLaunchKernelGGL(build_entries_kernel, dim3(grid_m), dim3(256), 0, st, sc, sv, eidx, M, list_net,
                                   einfo, marker);
            });
            // entries: eidx[M] on the device; the scans and kernels below run over
            // all M positions and ignore those past it
            timed(c, GR_K_BATCH_MARKS, 20.0 * M, [&] {
                hip_check(hipcub::DeviceScan::ExclusiveScan(tmp, max_bytes, marker, prev_marker, hipcub::Max(), -1, M, st),
                          "scan markers");
                hipLaunchKernelGGL(reader_slots_kernel, dim3(grid_m), dim3(256), 0, st, eidx + M, list_net, einfo,
                                   b.pt_off, rcount, slot_of);
                hipLaunchKernelGGL(reader_records_kernel, dim3(grid_m), dim3(256), 0, st, eidx + M, einfo, prev_marker,
                                   slot_of, rec);
            });
            hip_check(hipGetLastError(), "batching list kernels");
        }

        // coloring, widened on color overflow, re-run from the first net that
        // would overfill a batch
        std::vector<int> closed_color, closed_from;
        int first = 0;
        int words = std::max(1, std::min(kMaxMaskWords, c->batch_mask_words));
        std::vector<int> counts;
        std::vector<int> host_color;
        bool have_colors = false;
        while (true) {
            unsigned long long* pm = nullptr;
            c->pm_ws.carve([&](auto& cv) { pm = cv.template take<unsigned long long>(size_t(npts) * words); });
            const long long slot0 = c->geom_pt_off[first];
            hip_check(hipMemsetAsync(pm + slot0 * words, 0, sizeof(unsigned long long) * size_t(npts - slot0) * words, st),
                      "clear prefix masks");
            hipLaunchKernelGGL(fill_int_kernel, dim3((n - first + 255) / 256), dim3(256), 0, st, d_color + first, -1,
                               (long long)(n - first));
            hip_check(hipMemsetAsync(d_count, 0, sizeof(int) * kMaxColors, st), "clear color counts");
            hip_check(hipMemsetAsync(d_misc, 0, sizeof(int) * kColorQueues * kColorQueueStride, st),
                      "clear dequeue counters");
            const int nclosed = int(closed_color.size());
            if (nclosed) {
                upload(d_closed, closed_color.data(), sizeof(int) * nclosed, st);
                upload(d_closed + kMaxColors, closed_from.data(), sizeof(int) * nclosed, st);
            }
            ColorArgs ca;
            ca.n = n;
            ca.pt_off = b.pt_off;
            ca.rcount = rcount;
            ca.rec = rec;
            ca.list_net = list_net;
            ca.pm = pm;
            ca.mask_words = words;
            ca.color = d_color;
            ca.next = d_misc;
            ca.first = first;
            ca.error = c->error;
            ca.spin_limit = 1ll << 22;
            ca.nclosed = nclosed;
            ca.closed_color = d_closed;
            ca.closed_from = d_closed + kMaxColors;
            const int waves = std::max(1, std::min(c->cus * c->color_waves_per_cu, n - first));
            const size_t lds = sizeof(unsigned long long) * size_t(kWave + 1) * words;
            // algorithmic bytes of the whole batching, charged to this launch:
            // per net 4 B per point, 8 B per segment and 4 B of color, plus one
            // 4-byte read-modify-write (8 B) per marked cell
            const double color_bytes =
                4.0 * npts + 8.0 * (c->geom_hsegs + c->geom_vsegs) + 4.0 * n + 8.0 * cand_marks;
            const size_t color_rec = c->prof.size();
            timed(c, GR_K_BATCH_COLOR, color_bytes,
                  [&] { hipLaunchKernelGGL(color_kernel, dim3(waves), dim3(kWave), lds, st, ca); });
            hipLaunchKernelGGL(color_hist_kernel, dim3(std::min(1024, (n - first + 255) / 256)), dim3(256), 0, st,
                               d_color, first, n, d_count);
            hip_check(hipGetLastError(), "color kernels");
            int err = 0;
            download(&err, c->error, sizeof(int), st);
            if (err == kDevColorOverflow && words < kMaxMaskWords) {
                const int zero = 0;
                upload(c->error, &zero, sizeof(int), st);
                if (c->profiling) c->prof.resize(color_rec);
                words = std::min(kMaxMaskWords, words * 2);
                first = 0;
                closed_color.clear();
                closed_from.clear();
                continue;
            }
            if (err) check_device_error(c, "gr_batch_colors");  // reads the word again, throws, clears it
            counts.assign(kMaxColors, 0);
            download(counts.data(), d_count, sizeof(int) * kMaxColors, st);
            while (!counts.empty() && !counts.back()) counts.pop_back();
            bool over = false;
            for (int v : counts) over |= v > limit;
            if (!over && !first) break;
            // exact limit: find the first net that would overfill a batch
            host_color.resize(n);
            download(host_color.data(), d_color, sizeof(int) * n, st);
            have_colors = true;
            std::vector<int> fill(kMaxColors, 0);
            int kstar = -1;
            for (int k = 0; k < n; ++k) {
                if (++fill[host_color[k]] > limit) {
                    kstar = k;
                    break;
                }
            }
            if (kstar < 0) {
                counts.assign(kMaxColors, 0);
                int top = -1;
                for (int k = 0; k < n; ++k) ++counts[host_color[k]], top = std::max(top, host_color[k]);
                counts.resize(top + 1);
                break;
            }
            if (int(closed_color.size()) >= kMaxColors) throw GrError(GR_E_CAPACITY, "too many full batches");
            closed_color.push_back(host_color[kstar]);
            closed_from.push_back(kstar);
            first = kstar;
        }
        c->batch_mask_words = words;
        if (have_colors) std::memcpy(color, host_color.data(), sizeof(int) * n);
        else download(color, d_color, sizeof(int) * n, st);
        *num_batches = int(counts.size());
    });
}

int gr_batch_plan(gr_ctx* c, int n, const int* pt_off, const int* pts, const int* hs_off, const int* hsegs,
                  const int* vs_off, const int* vsegs, int limit, int sparsity, int* color, int* num_batches) {
    const int rc = gr_batch_geometry(c, n, pt_off, pts, hs_off, hsegs, vs_off, vsegs);
    if (rc) return rc;
    return gr_batch_colors(c, limit, sparsity, color, num_batches);
}

int gr_score(gr_ctx* c, int nets, const int* pin_off, const int* pins, const int* extra_off, const int* extra,
             const int* sub_off, const int* subs, const int* layer_dir, const double* capacity,
             const double* overflow_cost, double* out, unsigned short* demand) {
    return guarded([&] {
        if (!c || !capacity || !overflow_cost || !out) throw GrError(GR_E_ARG, "null argument");
        DeviceArena ar;
        ScoreArgs a = net_lines(c, ar, nets, pin_off, pins, extra_off, extra, sub_off, subs, layer_dir);
        const int X = c->g.X, Y = c->g.Y, M = c->g.L + 1;
        const long long XY = (long long)X * Y, cells = XY * M;
        const int blocks_per_layer = int((XY + kScoreCells - 1) / kScoreCells);
        double* d_cap = ar.alloc<double>(cells);
        int* d_wire = ar.alloc<int>(cells);
        int* d_stacked = ar.alloc<int>(cells);
        const int ncount = kStatStripes * kScoreLayers + 1;  // striped via counters, rejected nets
        unsigned long long* d_counts = ar.alloc<unsigned long long>(ncount);
        double* d_ploss = ar.alloc<double>(size_t(M) * blocks_per_layer);
        unsigned long long* d_pwl = ar.alloc<unsigned long long>(size_t(M) * blocks_per_layer);
        unsigned short* d_demand = demand ? ar.alloc<unsigned short>(cells) : nullptr;
        hipStream_t st = c->stream;
        upload(d_cap, capacity, sizeof(double) * cells, st);
        a.wire = d_wire, a.stacked = d_stacked, a.vias = d_counts, a.bad_nets = d_counts + ncount - 1;
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, st), "hipEventRecord");
        hipLaunchKernelGGL(fill_int_kernel, dim3(int(std::min<long long>(4096, (cells + 255) / 256))), dim3(256), 0, st,
                           d_wire, 0, cells);
        hipLaunchKernelGGL(fill_int_kernel, dim3(int(std::min<long long>(4096, (cells + 255) / 256))), dim3(256), 0, st,
                           d_stacked, 0, cells);
        hipLaunchKernelGGL(fill_int_kernel, dim3(16), dim3(256), 0, st, (int*)d_counts, 0, 2LL * ncount);
        if (nets)
            hipLaunchKernelGGL(score_nets_kernel, dim3(std::min(nets, std::max(1, c->cus) * 16)), dim3(kWave), 0, st, a);
        hipLaunchKernelGGL(score_cells_kernel, dim3(blocks_per_layer, M), dim3(256), 0, st, a, d_cap, c->g.x_edge_len,
                           c->g.y_edge_len, d_ploss, d_pwl, d_demand);
        hip_check(hipEventRecord(e1, st), "hipEventRecord");
        hip_check(hipGetLastError(), "score kernels");
        std::vector<unsigned long long> counts(ncount), pwl(size_t(M) * blocks_per_layer);
        std::vector<double> ploss(size_t(M) * blocks_per_layer);
        download(counts.data(), d_counts, sizeof(unsigned long long) * ncount, st);
        download(pwl.data(), d_pwl, sizeof(unsigned long long) * pwl.size(), st);
        download(ploss.data(), d_ploss, sizeof(double) * ploss.size(), st);
        if (demand) download(demand, d_demand, sizeof(unsigned short) * cells, st);
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // resources/evaluator.cpp:575-623, every layer routable (CAP
        // directions are 0 or 1)
        double wl_cost = 0, via_cost = 0, overflow = 0;
        for (int z = 0; z < M; ++z) {
            unsigned long long wl = 0;
            double layer_loss = 0;
            for (int b = 0; b < blocks_per_layer; ++b) {
                wl += pwl[size_t(z) * blocks_per_layer + b];
                layer_loss += ploss[size_t(z) * blocks_per_layer + b];
            }
            unsigned long long vias = 0;
            for (int t = 0; t < kStatStripes; ++t) vias += counts[size_t(t) * kScoreLayers + z];
            overflow += layer_loss * overflow_cost[z];
            via_cost += double(vias) * c->g.unit_via;
            wl_cost += double(wl) * c->g.unit_wire;
        }
        out[0] = wl_cost;
        out[1] = via_cost;
        out[2] = overflow;
        out[3] = overflow + via_cost + wl_cost;
        out[4] = double(counts[ncount - 1]);
        out[5] = double(ms);
    });
}

int gr_check_connectivity(gr_ctx* c, int nets, const int* pin_off, const int* pins, const int* extra_off,
                          const int* extra, const int* sub_off, const int* subs, const int* layer_dir,
                          const int* access_pin_off, const int* access_off, const int* access, unsigned char* status,
                          long long* counts) {
    return guarded([&] {
        if (!c || !access_pin_off || !access_off || !counts) throw GrError(GR_E_ARG, "null argument");
        DeviceArena ar;
        ScoreArgs a = net_lines(c, ar, nets, pin_off, pins, extra_off, extra, sub_off, subs, layer_dir);
        const long long cells = (long long)c->g.X * c->g.Y * (c->g.L + 1);
        if (access_pin_off[0] != 0) throw GrError(GR_E_ARG, "pin offsets must start at 0");
        for (int k = 0; k < nets; ++k)
            if (access_pin_off[k + 1] < access_pin_off[k]) throw GrError(GR_E_ARG, "pin offsets must be non-decreasing");
        const int npins = access_pin_off[nets];
        if (access_off[0] != 0) throw GrError(GR_E_ARG, "access offsets must start at 0");
        for (int p = 0; p < npins; ++p)
            if (access_off[p + 1] < access_off[p]) throw GrError(GR_E_ARG, "access offsets must be non-decreasing");
        const int nacc = access_off[npins];
        if (nacc && !access) throw GrError(GR_E_ARG, "null access points");
        for (int t = 0; t < nacc; ++t)
            if (access[t] < 0 || access[t] >= cells) throw GrError(GR_E_ARG, "access point outside the metal-layer grid");
        // staging of the nets whose lines can exceed kScoreLines: at most
        // their pin entries, extra pairs and the pairs their subnets' route
        // words can hold
        std::vector<long long> big_off(size_t(std::max(nets, 1)), -1);
        long long big = 0;
        for (int k = 0; k < nets; ++k) {
            long long bound = (pin_off[k + 1] - pin_off[k]) + (extra_off[k + 1] - extra_off[k]) / 2;
            for (int t = sub_off[k]; t < sub_off[k + 1]; ++t) {
                const int id = subs[t];
                const long long np = c->pin_acc_num_h[id + 1] - c->pin_acc_num_h[id];
                const bool two = c->routes2 && c->pin_off2_h[id] >= 0;
                bound += std::max(0LL, (np * (two ? c->route2_per_pin : c->route_per_pin) - 1) / 2);
            }
            if (bound > kScoreLines) big_off[k] = big, big += bound;
        }
        ConnArgs q{};
        int* d_apin = ar.alloc<int>(nets + 1);
        int* d_aoff = ar.alloc<int>(npins + 1);
        int* d_acc = ar.alloc<int>(std::max(nacc, 1));
        long long* d_big_off = ar.alloc<long long>(big_off.size());
        q.big_lines = ar.alloc<int4>(size_t(std::max(big, 1LL)));
        q.big_par = ar.alloc<int>(size_t(std::max(big, 1LL)));
        q.status = ar.alloc<unsigned char>(std::max(nets, 1));
        q.counts = ar.alloc<unsigned long long>(2);
        hipStream_t st = c->stream;
        upload(d_apin, access_pin_off, sizeof(int) * (nets + 1), st);
        upload(d_aoff, access_off, sizeof(int) * (npins + 1), st);
        if (nacc) upload(d_acc, access, sizeof(int) * nacc, st);
        upload(d_big_off, big_off.data(), sizeof(long long) * big_off.size(), st);
        hip_check(hipMemsetAsync(q.counts, 0, 2 * sizeof(unsigned long long), st), "clear counts");
        q.apin_off = d_apin, q.acc_off = d_aoff, q.acc = d_acc, q.big_off = d_big_off;
        if (nets)
            hipLaunchKernelGGL(connectivity_kernel, dim3(std::min(nets, std::max(1, c->cus) * 16)), dim3(kWave), 0, st,
                               a, q);
        hip_check(hipGetLastError(), "connectivity kernel");
        unsigned long long n2[2];
        download(n2, q.counts, sizeof(n2), st);
        if (status && nets) download(status, q.status, size_t(nets), st);
        counts[0] = (long long)n2[0];
        counts[1] = (long long)n2[1];
        counts[2] = counts[0] + counts[1];
    });
}

}  // extern "C"
