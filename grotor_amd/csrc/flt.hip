// FLT routing-DAG DP on gfx950.
//
// Two launches per batch: flt_edge_kernel searches, for every DAG edge and
// from-layer, the best layer-change position towards every same-parity
// to-layer; flt_route_kernel gives each DAG one wavefront, which initializes
// the DAG's tables, solves its depths bottom-up (children push packed
// (cost, node, choice) keys into their parents' child slots with 64-bit
// atomic min) and traces them top-down.  Between depths the wavefront waits
// for its own stores and atomics, and lanes read what other lanes wrote with
// loads that bypass L1.
//
// Segments are appended in node order within a depth (a wavefront prefix
// sum), so the route words are deterministic.
//
// Reference: precompute_cost_edge (src/routing/flt_cost.cuh:50-130),
// solve_flt_node (src/routing/flt_block.cuh:6-142), trace_flt_node
// (:144-331), route_flt_block (:334-355).
#include "gr_kernels.hpp"

namespace grotor {
namespace {

constexpr int kS = kFltStride;  // layer index 1..L of costs and slot tables, 0 unused
constexpr int kStraight = 82;   // choice code of a same-layer edge
// float 1.01e22 in the upper word, code 200000000 (src/routing/flt_cost.cuh:48)
constexpr unsigned long long kEmpty = 7208259049189261824ULL;

__device__ __forceinline__ unsigned long long pack(float cost, int code) {
    const int i = __float_as_int(cost);
    return ((unsigned long long)(unsigned)(i >= 0 ? i : i ^ 0x7FFFFFFF) << 32) | (unsigned)code;
}
__device__ __forceinline__ double unpack_cost(unsigned long long v) {
    const int i = (int)(unsigned)(v >> 32);
    return (double)__int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF);
}

__device__ __forceinline__ double wire(const FltArgs& a, int l, int xmin, int xmax, int ymin, int ymax) {
    return a.presum[cell_of(a.g, l, xmax, ymax)] - a.presum[cell_of(a.g, l, xmin, ymin)];
}

// Between depths: every store and atomic of this wavefront has reached L2.
// Values another lane wrote (child slots lowered by atomics, best paths,
// layer outputs) are then read with L1-bypassing loads (l2_load), so no
// cache invalidation is needed.
__device__ __forceinline__ void depth_barrier() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

template <class T>
__device__ __forceinline__ T l2_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void raise(int* error, int code) { atomicCAS(error, 0, code); }

// Layer-change tables: per edge and from-layer f the five to-layers of f's
// parity (the only ones the DP reads), dense, instead of the reference's
// 9 x 9 rows (src/routing/flt_cost.cuh:71, 127).
__host__ __device__ __forceinline__ size_t change_slot(int e, int f, int t) {
    return ((size_t)e * 9 + f) * kFltChangeWidth + (t - (f & 1)) / 2;
}

// One DAG's records and layer-change tables (global memory, read-only
// during the DP), net-local.
struct Dag {
    const int4* nodes;  // {coordinate, first edge, end edge, first slot * 8 + slots}
    const int4* edges;  // {parent, slot, parent coordinate, child coordinate}
    const double* cost_edges;  // [change_slot(edge, from, to)]
    const int* best_change;
};

// Where a DAG's DP tables live.  Rows are child slots; a row holds the
// layers 1..L.  GlobalTables: the batch tables in global memory (row stride
// kS), values other lanes wrote read past L1 (l2_load).  LdsTables: the
// DAG's rows in LDS (row stride L) with its nodes' via costs staged there
// once; the slot keys, best paths and layer outputs never leave the
// compute unit.  Both give the same values: the DP below is one code for
// both.
struct GlobalTables {
    static constexpr bool kReload = false;  // keep slot keys in registers rather than reload them
    double* costs;  // [node * kS + layer]; only the root's row is read
    unsigned long long* packed;
    unsigned long long* best;
    int* layer;
    const float* vcost;  // the grid's via costs
    DeviceGrid g;
    __device__ __forceinline__ unsigned long long key(int row, int li) const { return l2_load(packed + (size_t)row * kS + li); }
    __device__ __forceinline__ void lower(int row, int li, unsigned long long v) const { atomicMin(packed + (size_t)row * kS + li, v); }
    __device__ __forceinline__ unsigned long long path(int row, int li) const { return l2_load(best + (size_t)row * kS + li); }
    __device__ __forceinline__ void set_path(int row, int li, unsigned long long v) const { best[(size_t)row * kS + li] = v; }
    __device__ __forceinline__ int out(int v) const { return l2_load(layer + v); }
    __device__ __forceinline__ void set_out(int v, int x) const { layer[v] = x; }
    __device__ __forceinline__ double root_cost(int v, int li) const { return costs[(size_t)v * kS + li]; }
    __device__ __forceinline__ void set_root_cost(int v, int li, double c) const { costs[(size_t)v * kS + li] = c; }
    __device__ __forceinline__ float via(int /*v*/, int j, int x, int y) const { return vcost[cell_of(g, j, x, y)]; }
    __device__ __forceinline__ void clear(const Dag&, int slots, int /*nodes*/, int lane, int W) const {
        for (int i = lane; i < slots * kS; i += W) packed[i] = kEmpty;
    }
    __device__ __forceinline__ static void barrier() { depth_barrier(); }
};

template <int L>
struct LdsTables {
    static constexpr bool kReload = true;  // slot keys reloaded from LDS instead of held in registers
    unsigned long long* packed;  // [row * L + li - 1]
    unsigned* best;              // best paths, code * kS + layer (below 2^32: codes are node * 100 + choice)
    double* rootc;               // [kS], the root's layer costs
    int4* nrec;                  // the DAG's node records
    int4* erec;                  // edge records {parent, slot, parent coordinate, child node}
    int* eorder;                 // edges by the depth of their child (a.eorder)
    int* order;                  // nodes by depth (a.order)
    int* doff;                   // [depths + 1] offsets of the depths in order
    int* edoff;                  // [depths + 1] offsets of the depths in eorder
    double* cvs;                 // [position in depth * L + li - 1]: layer costs of a depth's solved nodes
    int* layer;                  // [node]
    int* posof;                  // [node]: position in its depth
    unsigned* xy;                // [node]: x << 16 | y, decoded once at staging
    float* vc;                   // [node * (L - 1) + j]: via cost of layer j at the node's cell
    __device__ __forceinline__ unsigned long long key(int row, int li) const { return packed[row * L + li - 1]; }
    // the float cost of a slot key, from its upper word alone
    __device__ __forceinline__ float cost(int row, int li) const {
        const int i = reinterpret_cast<const int*>(packed)[2 * (row * L + li - 1) + 1];
        return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF);
    }
    __device__ __forceinline__ void lower(int row, int li, unsigned long long v) const { atomicMin(packed + row * L + li - 1, v); }
    __device__ __forceinline__ unsigned long long path(int row, int li) const { return best[row * L + li - 1]; }
    __device__ __forceinline__ void set_path(int row, int li, unsigned long long v) const { best[row * L + li - 1] = (unsigned)v; }
    __device__ __forceinline__ int out(int v) const { return layer[v]; }
    __device__ __forceinline__ void set_out(int v, int x) const { layer[v] = x; }
    __device__ __forceinline__ double root_cost(int, int li) const { return rootc[li]; }
    __device__ __forceinline__ void set_root_cost(int, int li, double c) const { rootc[li] = c; }
    __device__ __forceinline__ float via(int v, int j, int, int) const { return vc[v * (L - 1) + j]; }
    // every LDS access of the wavefront has completed
    __device__ __forceinline__ static void barrier() {
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_wave_barrier();
    }
};

using LdsShape = FltLdsShape;

// Lays out LdsTables t over base (or only sizes them with base = nullptr);
// returns the bytes, 16-byte aligned.  (t is a reference: a null test on the
// tables' address would keep them in private memory.)
template <int L>
__host__ __device__ inline size_t lds_layout(char* base, const LdsShape& s, LdsTables<L>& t) {
    size_t off = 0;
    auto take = [&](size_t bytes, size_t align) {
        off = (off + align - 1) & ~(align - 1);
        char* p = base ? base + off : nullptr;
        off += bytes;
        return p;
    };
    char* packed = take((size_t)s.slots * L * 8, 16);
    char* best = take((size_t)s.slots * L * 4, 8);
    char* rootc = take((size_t)kS * 8, 8);
    char* nrec = take((size_t)s.nodes * 16, 16);
    char* erec = take((size_t)s.edges * 16, 16);
    char* cvs = take((size_t)s.dnodes * L * 8, 8);
    char* eorder = take((size_t)s.edges * 4, 4);
    char* order = take((size_t)s.nodes * 4, 4);
    char* doff = take((size_t)(s.depths + 1) * 4, 4);
    char* edoff = take((size_t)(s.depths + 1) * 4, 4);
    char* layer = take((size_t)s.nodes * 4, 4);
    char* posof = take((size_t)s.nodes * 4, 4);
    char* xy = take((size_t)s.nodes * 4, 4);
    char* vc = take((size_t)s.nodes * (L > 1 ? L - 1 : 0) * 4, 4);
    t.packed = reinterpret_cast<unsigned long long*>(packed);
    t.best = reinterpret_cast<unsigned*>(best);
    t.rootc = reinterpret_cast<double*>(rootc);
    t.nrec = reinterpret_cast<int4*>(nrec);
    t.erec = reinterpret_cast<int4*>(erec);
    t.cvs = reinterpret_cast<double*>(cvs);
    t.eorder = reinterpret_cast<int*>(eorder);
    t.order = reinterpret_cast<int*>(order);
    t.doff = reinterpret_cast<int*>(doff);
    t.edoff = reinterpret_cast<int*>(edoff);
    t.layer = reinterpret_cast<int*>(layer);
    t.posof = reinterpret_cast<int*>(posof);
    t.xy = reinterpret_cast<unsigned*>(xy);
    t.vc = reinterpret_cast<float*>(vc);
    return (off + 15) & ~size_t(15);
}

// solve_flt_node (src/routing/flt_block.cuh:6-142) for any number of child
// slots, the node's layer costs in a private array.
template <bool kPush, class Tb>
__device__ void solve_node(const FltArgs& a, const Dag& d, const Tb& tb, int v, int pos) {
    const DeviceGrid& g = a.g;
    const int L = g.L;
    const int4 r = d.nodes[v];
    const int low_fixed = 1 + r.x / g.XY;
    const int x = r.x / g.Y % g.X, y = r.x % g.Y;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const int kids = r.w & 7;
    const int row0 = r.w >> 3;
    double cv[kS];
    for (int li = 0; li < kS; ++li) cv[li] = kInf;
    double mins[6];
    unsigned long long paths[6];
    for (int low = 1; low <= low_fixed; ++low) {
        for (int k = 0; k < kids; ++k) mins[k] = kInf;
        double via = 0.0;
        for (int li = low; li < L + 1; ++li) {
            if (li > low) via += (double)tb.via(v, li - 2, x, y);
            for (int k = 0; k < kids; ++k) {
                const unsigned long long p = tb.key(row0 + k, li);
                const double c = unpack_cost(p);
                if (c < mins[k]) {
                    mins[k] = c;
                    paths[k] = (p & 0xFFFFFFFFULL) * kS + (unsigned long long)li;
                }
            }
            if (li >= high_fixed) {
                double cost = via;
                for (int k = 0; k < kids; ++k) cost += mins[k];
                if (cost < kInf && cost < cv[li]) {
                    cv[li] = cost;
                    for (int k = 0; k < kids; ++k) tb.set_path(row0 + k, li, paths[k]);
                }
            }
        }
        for (int li = L - 1; li >= low; --li)
            if (cv[li + 1] < cv[li]) {
                cv[li] = cv[li + 1];
                for (int k = 0; k < kids; ++k) tb.set_path(row0 + k, li, tb.path(row0 + k, li + 1));
            }
    }
    if (r.y == r.z)  // the root's trace reads its layer costs
        for (int li = 0; li < kS; ++li) tb.set_root_cost(v, li, li <= L ? cv[li] : kInf);
    if constexpr (!kPush) {
        // the depth's push items read the layer costs from LDS
        tb.posof[v] = pos;
        for (int li = 1; li <= L; ++li) tb.cvs[pos * L + li - 1] = cv[li];
        return;
    }
    unsigned long long cache[kS];
    for (int e = r.y; e < r.z; ++e) {
        const int4 er = d.edges[e];
        const int prow = (d.nodes[er.x].w >> 3) + er.y;
        const int px = er.z / g.Y % g.X, py = er.z % g.Y;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        for (int t = 1; t < L + 1; ++t) cache[t] = kEmpty;
        for (int layer = 1; layer < L + 1; ++layer) {
            if (((layer & 1) ^ g.DIR) == 1 && y != py) continue;
            if (((layer & 1) ^ g.DIR) == 0 && x != px) continue;
            const double base = cv[layer];
            if (!(base < kInf)) {  // the reference asserts here
                raise(a.error, kDevNoSolution);
                continue;
            }
            for (int t = 1 + (1 - layer % 2); t < L + 1; t += 2) {
                double cost2;
                int choice = kStraight;
                if (t == layer) {
                    cost2 = base + wire(a, layer - 1, minx, maxx, miny, maxy);
                } else if (abs(x - px) + abs(y - py) > 4) {
                    const int seq = (layer - 1) * 9 + (t - 1);
                    const double ce = d.cost_edges[change_slot(e, layer - 1, t - 1)];
                    if (!(ce < kInf)) continue;
                    cost2 = base + ce;
                    choice = seq;
                } else {
                    continue;
                }
                const unsigned long long nv = pack((float)cost2, v * 100 + choice);
                if (nv < cache[t]) cache[t] = nv;
            }
        }
        for (int t = 1; t < L + 1; ++t)
            if (cache[t] < kEmpty) tb.lower(prow, t, cache[t]);
    }
}

// solve_flt_node for nodes of at most kFastKids child slots with the
// node's tables in registers: the slot values and via costs are loaded
// once (one memory latency instead of one per inner step) and the layer
// costs live in registers.  Instead of copying best-path rows, the passes
// keep, per layer, the slot positions (4 bits per child slot) of the
// running first minima that produced the layer's cost; the best-path rows
// are written once at the end from those positions, which is what the
// reference's rows hold after its copies.
constexpr int kFastKids = 4;

// The candidates of a solved node towards its parents (the second half of
// solve_flt_node, :81-141) with the layer costs cv in registers.  A layer
// can run along an edge only if its direction matches, which for an edge of
// non-zero length selects one layer parity, so each lane walks only that
// parity's layers (and to-layers); an edge between two nodes on the same
// cell takes every layer.
template <int L, class Tb>
__device__ __forceinline__ void push_parents(const FltArgs& a, const Dag& d, const Tb& tb, int v, int x, int y,
                                             const int4 r, const double* cv) {
    const DeviceGrid& g = a.g;
    constexpr int H = (L + 1) / 2;  // layers of one parity, at most
    for (int e = r.y; e < r.z; ++e) {
        const int4 er = d.edges[e];
        const int prow = (d.nodes[er.x].w >> 3) + er.y;
        const int px = er.z / g.Y % g.X, py = er.z % g.Y;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const bool far = abs(x - px) + abs(y - py) > 4;
        if (x == px && y == py) {
            // zero length: every layer, straight only (far is false)
#pragma unroll
            for (int layer = 1; layer <= L; ++layer) {
                const double base = cv[layer];
                if (!(base < kInf)) {  // the reference asserts here
                    raise(a.error, kDevNoSolution);
                    continue;
                }
                const unsigned long long nv =
                    pack((float)(base + wire(a, layer - 1, minx, maxx, miny, maxy)), v * 100 + kStraight);
                tb.lower(prow, layer, nv);
            }
            continue;
        }
        // a layer runs along y when ((layer & 1) ^ DIR) == 0 (the reference's
        // direction tests, :101-102): odd layers do when DIR == 1
        const bool odd_runs_y = g.DIR == 1;
        const int par = (x == px) == odd_runs_y ? 1 : 2;  // first layer of the usable parity
        unsigned long long cache[H];
#pragma unroll
        for (int m = 0; m < H; ++m) cache[m] = kEmpty;
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const int layer = par + 2 * j;
            if (layer > L) continue;
            const double base = par == 1 ? cv[min(2 * j + 1, L)] : cv[min(2 * j + 2, L)];
            if (!(base < kInf)) {  // the reference asserts here
                raise(a.error, kDevNoSolution);
                continue;
            }
#pragma unroll
            for (int m = 0; m < H; ++m) {
                const int t = par + 2 * m;
                if (t > L) continue;
                double cost2;
                int choice = kStraight;
                if (m == j) {
                    cost2 = base + wire(a, layer - 1, minx, maxx, miny, maxy);
                } else if (far) {
                    const double ce = d.cost_edges[change_slot(e, layer - 1, t - 1)];
                    if (!(ce < kInf)) continue;
                    cost2 = base + ce;
                    choice = (layer - 1) * 9 + (t - 1);
                } else {
                    continue;
                }
                const unsigned long long nv = pack((float)cost2, v * 100 + choice);
                if (nv < cache[m]) cache[m] = nv;
            }
        }
#pragma unroll
        for (int m = 0; m < H; ++m)
            if (cache[m] < kEmpty && par + 2 * m <= L) tb.lower(prow, par + 2 * m, cache[m]);
    }
}

template <int L, int K, bool kPush, class Tb>
__device__ __forceinline__ void solve_node_fast(const FltArgs& a, const Dag& d, const Tb& tb, int v, const int4 r,
                                                int pos) {
    const DeviceGrid& g = a.g;
    const int kids = r.w & 7;
    const int low_fixed = 1 + r.x / g.XY;
    const int x = r.x / g.Y % g.X, y = r.x % g.Y;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const int row0 = r.w >> 3;
    // slot keys, loaded once: the costs for the pass below, the codes for
    // the best-path rows
    constexpr int KK = Tb::kReload ? 1 : K;  // keys held in registers (LDS tables reload them)
    constexpr int LK = Tb::kReload ? 1 : L + 1;
    unsigned long long pkv[KK][LK];
    float pv[K][L + 1];  // slot costs (the packed key's float)
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int li = 1; li <= L; ++li) {
            const unsigned long long key = k < kids ? tb.key(row0 + k, li) : 0ull;
            if constexpr (!Tb::kReload) pkv[k][li] = key;
            pv[k][li] = k < kids ? (float)unpack_cost(key) : 0.f;
        }
    float vc[L];
#pragma unroll
    for (int j = 0; j + 1 < L; ++j) vc[j] = tb.via(v, j, x, y);
    // The reference's passes over the stack's low end (low = 1..lf, each an
    // upward sweep with strict updates, then a downward sweep copying only
    // strictly smaller costs) leave at every layer li the first minimum, in
    // the order (low, then li' >= li), of the costs C(low, li') of the via
    // stacks [low, li']: a pass never touches layers below its low, keeps an
    // earlier pass's cost on a tie and copies down only what is strictly
    // smaller.  The C(low, li') are independent of each other, so they are
    // computed side by side (each pass's prefix sum of via costs and running
    // first minima of the slot costs in their own registers) rather than pass
    // after pass; the selection then keeps, per low, the first minimum over
    // li' >= li and takes the first minimum over the lows.  Every cost is the
    // same sum in the same order as in the reference's passes.
    const int lf = min(L, low_fixed);
    double C[L][L];    // C[low - 1][li - 1], kInf where the stack has no finite cost
    unsigned R[L][L];  // slot positions of C's running minima, 4 bits per child slot
    {
        double V[L];   // per low: via costs of layers low+1..li, summed upward
        float M[L][K];  // per low: running first minima of the slot costs over low..li
        unsigned P[L];  // their positions
#pragma unroll
        for (int li = 1; li <= L; ++li) {
#pragma unroll
            for (int low = 1; low <= li; ++low) {
                const int i = low - 1;
                if (li == low) {
                    V[i] = 0.0;
                    P[i] = 0;
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        if (k < kids) {
                            M[i][k] = pv[k][li];
                            P[i] |= (unsigned)li << (4 * k);
                        }
                } else {
                    V[i] += (double)vc[li - 2];
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        if (k < kids && pv[k][li] < M[i][k]) {
                            M[i][k] = pv[k][li];
                            P[i] = (P[i] & ~(15u << (4 * k))) | ((unsigned)li << (4 * k));
                        }
                }
                double cost = V[i];
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (k < kids) cost += (double)M[i][k];
                C[i][li - 1] = low <= lf && li >= high_fixed && cost < kInf ? cost : kInf;
                R[i][li - 1] = P[i];
            }
        }
    }
    double cv[L + 1];
    unsigned at[L + 1];  // per layer: the slot positions of its cost, 4 bits per child slot
    {
        double S[L];     // per low: first minimum of C(low, li') over li' >= li
        unsigned SR[L];
#pragma unroll
        for (int i = 0; i < L; ++i) S[i] = kInf, SR[i] = 0;
        cv[0] = kInf, at[0] = 0;
#pragma unroll
        for (int li = L; li >= 1; --li) {
#pragma unroll
            for (int low = 1; low <= li; ++low) {
                const int i = low - 1;
                const double c = C[i][li - 1];
                if (c < kInf && c <= S[i]) {  // going down, a tie goes to the lower li'
                    S[i] = c;
                    SR[i] = R[i][li - 1];
                }
            }
            double best = kInf;
            unsigned pos = 0;
#pragma unroll
            for (int low = 1; low <= li; ++low)
                if (S[low - 1] < best) {  // a tie goes to the lower low
                    best = S[low - 1];
                    pos = SR[low - 1];
                }
            cv[li] = best;
            at[li] = pos;
        }
    }
#pragma unroll
    for (int li = 1; li <= L; ++li) {
        if (!(cv[li] < kInf)) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k >= kids) continue;
            const int j = (at[li] >> (4 * k)) & 15;
            unsigned long long key = 0;  // slot k's key at layer j, j a run-time layer
            if constexpr (Tb::kReload) {
                key = tb.key(row0 + k, j);
            } else {
#pragma unroll
                for (int lj = 1; lj <= L; ++lj)
                    if (lj == j) key = pkv[k][lj];
            }
            tb.set_path(row0 + k, li, (key & 0xFFFFFFFFULL) * kS + (unsigned long long)j);
        }
    }
    if (r.y == r.z) {  // the root's trace reads its layer costs
        tb.set_root_cost(v, 0, kInf);
#pragma unroll
        for (int li = 1; li <= L; ++li) tb.set_root_cost(v, li, cv[li]);
    }
    if constexpr (kPush) {
        push_parents<L>(a, d, tb, v, x, y, r, cv);
    } else {
        // the depth's push items read the layer costs from LDS
        tb.posof[v] = pos;
#pragma unroll
        for (int li = 1; li <= L; ++li) tb.cvs[pos * L + li - 1] = cv[li];
    }
}

// A node's route words: its edge's segments e0..e5 (ne of them) and its via
// segment v0, v1 (nv of them), in named registers.
struct Words {
    int e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0, v0 = 0, v1 = 0;
    int ne = 0, nv = 0;
    __device__ __forceinline__ int count() const { return ne + nv; }
    // every word is a cell of the grid
    __device__ __forceinline__ bool in_grid(int cells) const {
        auto in = [&](int w) { return w >= 0 && w < cells; };
        const bool edge2 = in(e0) && in(e1), edge6 = edge2 && in(e2) && in(e3) && in(e4) && in(e5);
        return (ne == 0 || (ne == 2 ? edge2 : edge6)) && (nv == 0 || (in(v0) && in(v1)));
    }
    __device__ __forceinline__ void put(int* routes, int at) const {
        int* p = routes + at;
        if (ne) {
            p[0] = e0;
            p[1] = e1;
        }
        if (ne == 6) {
            p[2] = e2;
            p[3] = e3;
            p[4] = e4;
            p[5] = e5;
        }
        if (nv) {
            p[ne] = v0;
            p[ne + 1] = v1;
        }
    }
};

// The segments of edge e (record er) from node (cx, cy) to its parent for
// the parent's best path `path` and the node's layer lv (trace_flt_node,
// src/routing/flt_block.cuh:144-331): a straight wire on lv's routing layer,
// or wire, via stack and wire through the edge's best layer-change position.
// A zero-length edge has none.  A straight wire is the first two words of a
// layer change on lv at the parent's end.
__device__ __forceinline__ void edge_words(const FltArgs& a, const Dag& d, int e, const int4 er,
                                           unsigned long long path, int lv, int cx, int cy, Words& wd) {
    const DeviceGrid& g = a.g;
    const int choice = (int)((path / kS) % 100);
    const int px = er.z / g.Y % g.X, py = er.z % g.Y;
    const bool vert = px == cx && cy != py, horz = py == cy && cx != px;
    if (!vert && !horz) {
        wd.ne = 0;
        return;
    }
    const bool straight = choice == kStraight;
    const int an = vert ? cy : cx, ap = vert ? py : px;  // coordinates along the edge
    const int fl = straight ? lv - 1 : choice / 9, tl = straight ? lv - 1 : choice % 9;
    const int pos = straight ? ap : d.best_change[change_slot(e, fl, tl)];
    auto idx = [&](int ll, int along) { return vert ? (ll * g.X + cx) * g.Y + along : (ll * g.X + along) * g.Y + cy; };
    wd.e0 = idx(fl, min(an, pos));
    wd.e1 = idx(fl, max(an, pos));
    wd.e2 = idx(min(fl, tl), pos);
    wd.e3 = idx(max(fl, tl), pos);
    wd.e4 = idx(tl, min(ap, pos));
    wd.e5 = idx(tl, max(ap, pos));
    wd.ne = straight ? 2 : 6;
}

// The node's via stack: from its own layer (pin nodes) or lv over the
// layers its children's best paths enter at lv, as one segment.
template <class Tb>
__device__ __forceinline__ void via_words(const FltArgs& a, const Tb& tb, const int4 r, int lv, Words& wd) {
    const DeviceGrid& g = a.g;
    const int L = g.L;
    const int l = r.x / g.XY, cx = r.x / g.Y % g.X, cy = r.x % g.Y;
    const int row0 = r.w >> 3;
    const int kids = r.w & 7;
    int minl = l + 1;
    int maxl = (l + 1) == L + 1 ? 1 : minl;
    minl = min(minl, lv);
    maxl = max(maxl, lv);
    for (int k = 0; k < kids; ++k) {
        const int lc = (int)(tb.path(row0 + k, lv) % kS);
        minl = min(minl, lc);
        maxl = max(maxl, lc);
    }
    wd.v0 = ((minl - 1) * g.X + cx) * g.Y + cy;
    wd.v1 = ((maxl - 1) * g.X + cx) * g.Y + cy;
    wd.nv = minl < maxl ? 2 : 0;
}

// The root's layer (the first of its least-cost layers, 0 when none is
// finite), as trace_flt_node picks it.
template <class Tb>
__device__ __forceinline__ int root_layer(const FltArgs& a, const Tb& tb, int v) {
    int best = 0;
    double bc = tb.root_cost(v, 0);
    for (int layer = 1; layer < a.g.L + 1; ++layer) {
        const double c = tb.root_cost(v, layer);
        if (c < bc) {
            bc = c;
            best = layer;
        }
    }
    return best;
}

// trace_flt_node (src/routing/flt_block.cuh:144-331): the node's layer, the
// segments of the edge its chosen parent uses and its via stack, into wd.
template <class Tb>
__device__ __forceinline__ void trace_node(const FltArgs& a, const Dag& d, const Tb& tb, int v, Words& wd) {
    const DeviceGrid& g = a.g;
    const int4 r = d.nodes[v];
    const int cx = r.x / g.Y % g.X, cy = r.x % g.Y;
    int lv;
    if (r.y == r.z) {
        lv = root_layer(a, tb, v);
        tb.set_out(v, lv);
        if (!lv) {  // no finite cost at any layer: nothing to trace
            raise(a.error, kDevNoSolution);
            return;
        }
    } else {
        int par_layer = -1, par = -1, par_edge = -1;
        lv = -1;
        for (int e = r.y; e < r.z; ++e) {
            const int4 er = d.edges[e];
            const int pl = tb.out(er.x);
            if (pl <= 0) continue;
            par = er.x;
            par_edge = e;
            par_layer = pl;
            const unsigned long long path = tb.path((d.nodes[par].w >> 3) + er.y, par_layer);
            const unsigned long long child = path / kS / 100, choice = (path / kS) % 100;
            if (child != (unsigned long long)v) {
                lv = -1;
                continue;
            }
            lv = choice == kStraight ? (int)(path % kS) : (int)(choice / 9 + 1);
            edge_words(a, d, e, er, path, lv, cx, cy, wd);
            break;
        }
        if (par_layer == -1) {
            tb.set_out(v, -1);
            wd.ne = 0;
            return;
        }
        const int4 er = d.edges[par_edge];
        const unsigned long long path = tb.path((d.nodes[par].w >> 3) + er.y, par_layer);
        if (path / kS / 100 != (unsigned long long)v) {
            tb.set_out(v, -1);
            wd.ne = 0;
            return;
        }
        tb.set_out(v, lv);
    }
    via_words(a, tb, r, lv, wd);
}

// The trace in two passes (LDS tables).  Pass one, depth by depth from the
// root: each node's layer as trace_node sets it, and its chosen parent edge
// into posof[] (free once the DP is done).  Pass two, every node at once:
// trace_node's words from the layers and edges pass one left.  Only pass
// one waits on the depth before it; pass two's global loads (layer-change
// positions) are issued together.
template <int L>
__device__ __forceinline__ void trace_layer(const FltArgs& a, const Dag& d, const LdsTables<L>& tb, int v) {
    const int4 r = d.nodes[v];
    if (r.y == r.z) {
        const int lv = root_layer(a, tb, v);
        tb.set_out(v, lv);
        if (!lv) raise(a.error, kDevNoSolution);  // no finite cost at any layer: nothing to trace
        return;
    }
    int par_edge = -1, lv = -1;
    unsigned long long path = 0;
    for (int e = r.y; e < r.z; ++e) {
        const int4 er = d.edges[e];
        const int pl = tb.out(er.x);
        if (pl <= 0) continue;
        par_edge = e;
        path = tb.path((d.nodes[er.x].w >> 3) + er.y, pl);
        if (path / kS / 100 != (unsigned long long)v) {
            lv = -1;
            continue;
        }
        const unsigned long long choice = (path / kS) % 100;
        lv = choice == kStraight ? (int)(path % kS) : (int)(choice / 9 + 1);
        break;
    }
    // trace_node's second read of the last parent path is this `path`
    if (par_edge == -1 || path / kS / 100 != (unsigned long long)v) {
        tb.set_out(v, -1);
        return;
    }
    tb.set_out(v, lv);
    tb.posof[v] = par_edge;
}

template <int L>
__device__ __forceinline__ void trace_words(const FltArgs& a, const Dag& d, const LdsTables<L>& tb, int v, Words& wd) {
    const DeviceGrid& g = a.g;
    const int lv = tb.out(v);
    if (lv <= 0) return;
    const int4 r = d.nodes[v];
    if (r.y != r.z) {
        const int e = tb.posof[v];
        const int4 er = d.edges[e];
        const unsigned long long path = tb.path((d.nodes[er.x].w >> 3) + er.y, tb.out(er.x));
        edge_words(a, d, e, er, path, lv, r.x / g.Y % g.X, r.x % g.Y, wd);
    }
    via_words(a, tb, r, lv, wd);
}

// One DAG of the batch on W lanes (lane = this lane's index among them),
// stepping with the other DAGs of its wavefront through `depths` (the
// wavefront's largest depth count): clear the tables, solve the depths
// bottom-up, trace them top-down and write the route words.
template <int L, class Tb>
__device__ __forceinline__ void route_dag(const FltArgs& a, int k, bool live, int lane, int W, const Dag& d,
                                          const Tb& tb, const int4 dn, const int* order, const int* doff, int depths) {
    tb.clear(d, dn.z, dn.x, lane, W);
    Tb::barrier();
    for (int dep = depths - 1; dep >= 0; --dep) {
        if (dep < dn.w) {
            const int end = doff[dep + 1];
            for (int i = doff[dep] + lane; i < end; i += W) {
                const int v = order[i];
                const int4 r = d.nodes[v];
                if ((r.w & 7) <= 2)
                    solve_node_fast<L, 2, true>(a, d, tb, v, r, 0);
                else if ((r.w & 7) <= kFastKids)
                    solve_node_fast<L, kFastKids, true>(a, d, tb, v, r, 0);
                else
                    solve_node<true>(a, d, tb, v, 0);
            }
        }
        Tb::barrier();
    }
    const int id = live ? a.net_ids[k] : 0;
    int* routes = a.routes.routes + (live ? (size_t)a.routes.pin_acc_num[id] * a.routes.per_pin : 0);
    const long long cap = live ? (long long)(a.pin_acc_num[id + 1] - a.pin_acc_num[id]) * a.routes.per_pin : 0;
    const int g_cells = a.g.L * a.g.XY;
    int count = 1;  // the root's trace sets the word count to 1
    for (int dep = 0; dep < depths; ++dep) {
        const int first = dep < dn.w ? doff[dep] : 0, end = dep < dn.w ? doff[dep + 1] : 0;
        const int chunks = wave_max((end - first + W - 1) / W);
        for (int ch = 0; ch < chunks; ++ch) {
            Words wd;
            const int i = first + ch * W + lane;
            if (i < end) trace_node(a, d, tb, order[i], wd);
            if (!wd.in_grid(g_cells)) {
                raise(a.error, kDevBadCoordinate);
                wd.ne = wd.nv = 0;
            }
            const int n = wd.count();
            // inclusive prefix of the word counts within this DAG's lanes
            int incl = n;
            for (int s = 1; s < W; s <<= 1) {
                const int o = __shfl_up(incl, s, W);
                if (lane >= s) incl += o;
            }
            const int at = count + incl - n;
            if (n) {
                if (at + n - 1 >= cap) {
                    raise(a.error, kDevRouteOverflow);
                } else {
                    wd.put(routes, at);
                }
            }
            count += __shfl(incl, W - 1, W);
        }
        Tb::barrier();
    }
    if (live && lane == 0) routes[0] = count <= cap ? count : 1;
}

// solve_flt_node (src/routing/flt_block.cuh:6-142) for nodes of three to six
// child slots, by the reference's passes themselves: per low end of the via
// stack an upward sweep (running first minima of each slot's costs, strict
// updates of the layer costs) and a downward sweep copying strictly smaller
// costs.  The slot costs, layer costs and their slot positions (4 bits per
// child slot) stay in registers; every loop over layers is unrolled, the
// run-time low end masking the layers below it.  The best-path rows are
// written once at the end from the positions.
template <int L>
__device__ __forceinline__ void solve_node_wide(const FltArgs& a, const Dag& d, const LdsTables<L>& tb, int v,
                                                const int4 r, int pos) {
    constexpr int K = 6;
    const DeviceGrid& g = a.g;
    const int low_fixed = 1 + r.x / g.XY;
    const int x = r.x / g.Y % g.X, y = r.x % g.Y;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const int kids = r.w & 7;
    const int row0 = r.w >> 3;
    float pv[K][L];  // slot costs (the packed keys' floats), layer li at [li - 1]
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int li = 1; li <= L; ++li) pv[k][li - 1] = k < kids ? (float)unpack_cost(tb.key(row0 + k, li)) : 0.f;
    float vc[L];
#pragma unroll
    for (int j = 0; j + 1 < L; ++j) vc[j] = tb.via(v, j, x, y);
    double cv[L];    // layer li at [li - 1]
    unsigned at[L];  // slot positions of cv, 4 bits per child slot
#pragma unroll
    for (int i = 0; i < L; ++i) cv[i] = kInf, at[i] = 0;
    const int lf = min(L, low_fixed);
    for (int low = 1; low <= lf; ++low) {
        float mins[K];
#pragma unroll
        for (int k = 0; k < K; ++k) mins[k] = (float)kInf;  // an empty key (1.01e22) never lowers it
        unsigned run = 0;
        double via = 0.0;
#pragma unroll
        for (int li = 1; li <= L; ++li) {
            if (li < low) continue;
            if (li > low) via += (double)vc[li - 2];
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < kids && pv[k][li - 1] < mins[k]) {
                    mins[k] = pv[k][li - 1];
                    run = (run & ~(15u << (4 * k))) | ((unsigned)li << (4 * k));
                }
            if (li >= high_fixed) {
                double cost = via;
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (k < kids) cost += (double)mins[k];
                if (cost < kInf && cost < cv[li - 1]) {
                    cv[li - 1] = cost;
                    at[li - 1] = run;
                }
            }
        }
#pragma unroll
        for (int li = L - 1; li >= 1; --li)
            if (li >= low && cv[li] < cv[li - 1]) {
                cv[li - 1] = cv[li];
                at[li - 1] = at[li];
            }
    }
#pragma unroll
    for (int li = 1; li <= L; ++li) {
        if (!(cv[li - 1] < kInf)) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k >= kids) continue;
            const int j = (at[li - 1] >> (4 * k)) & 15;
            const unsigned long long key = tb.key(row0 + k, j);
            tb.set_path(row0 + k, li, (key & 0xFFFFFFFFULL) * kS + (unsigned long long)j);
        }
    }
    if (r.y == r.z) {  // the root's trace reads its layer costs
        tb.set_root_cost(v, 0, kInf);
#pragma unroll
        for (int li = 1; li <= L; ++li) tb.set_root_cost(v, li, cv[li - 1]);
    }
    tb.posof[v] = pos;
#pragma unroll
    for (int li = 1; li <= L; ++li) tb.cvs[pos * L + li - 1] = cv[li - 1];
}

// Lanes per node of the lane-parallel solve: a power of two, at least L,
// inside one 16-lane DPP row.
template <int L>
constexpr int kNodeLanes = L > 8 ? 16 : 8;

// The first minimum over the lows of T[m] of lane g - m for m = M..0 (lows
// ascending), a tie going to the lower low: lane g's layer cost and slot
// positions.
template <int M>
__device__ __forceinline__ void merge_lows(const double* T, const unsigned* TR, int g, double& best, unsigned& at) {
    if constexpr (M >= 0) {
        const double t = row_shr<M>(T[M]);
        const unsigned tr = (unsigned)row_shr<M>((int)TR[M]);
        const bool take = g >= M && t < best;
        best = take ? t : best;
        at = take ? tr : at;
        merge_lows<M - 1>(T, TR, g, best, at);
    }
}

// solve_flt_node (src/routing/flt_block.cuh:6-142) for one node on G lanes
// (g = this lane's index among them), for depths with few nodes.  Lane g
// takes the via stacks with low end low = g + 1: it computes the costs
// C(low, li') for li' = low..L (the via costs of layers low+1..li' summed
// upward, then each slot's running first minimum over low..li' added, in the
// order of the single-lane solves) and keeps T[m], the first minimum of
// C(low, li') over li' >= low + m, a tie going to the lower li'.  Lane g then
// reads, for its layer li = g + 1, T[li - low] of every lane with low <= li,
// ascending, keeping the first strictly smaller one: the first minimum in
// the order (low, then li' >= li) that the reference's passes leave at li.
// Every cost is the same sum in the same order as in solve_node_fast and
// solve_node_wide.
template <int L, int K>
__device__ __forceinline__ void solve_node_lanes(const FltArgs& a, const LdsTables<L>& tb, int v, const int4 r,
                                                 int pos, int g) {
    constexpr int G = kNodeLanes<L>;
    const int kids = r.w & 7;
    const int low_fixed = 1 + r.x / a.g.XY;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const int lf = min(L, low_fixed);
    const int row0 = r.w >> 3;
    const int low = g + 1;
    // Straight-line code: lane-varying conditions are selects, and reads
    // past a lane's layers or child slots read an in-range entry whose value
    // is not used.
    double T[L];
    unsigned TR[L];  // slot positions of T, 4 bits per child slot
    {
        float M[K];
        unsigned run = 0;
        double via = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) M[k] = 0.f;
#pragma unroll
        for (int m = 0; m < L; ++m) {
            const int li = low + m;
            const bool in = li <= L;
            const int lc = in ? li : L;
            if (m) via += in ? (double)tb.vc[v * (L - 1) + lc - 2] : 0.0;
            double cost = via;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const bool on = k < kids;
                const float p = tb.cost(on ? row0 + k : row0, lc);
                const bool take = on && in && (m == 0 || p < M[k]);
                M[k] = take ? p : M[k];
                run = take ? (run & ~(15u << (4 * k))) | ((unsigned)li << (4 * k)) : run;
                cost = on ? cost + (double)M[k] : cost;
            }
            const bool ok = in && low <= lf && li >= high_fixed && cost < kInf;
            T[m] = ok ? cost : kInf;
            TR[m] = ok ? run : 0u;
        }
        double s = kInf;
        unsigned sr = 0;
#pragma unroll
        for (int m = L - 1; m >= 0; --m) {
            const bool take = T[m] < kInf && T[m] <= s;  // going down, a tie goes to the lower li'
            s = take ? T[m] : s;
            sr = take ? TR[m] : sr;
            T[m] = s;
            TR[m] = sr;
        }
    }
    static_assert(G <= 16, "a node's lanes lie in one DPP row");
    double best = kInf;
    unsigned at = 0;
    merge_lows<L - 1>(T, TR, g, best, at);  // lane g - m has low = li - m
    if (g >= L) return;
    const int li = g + 1;
    if (best < kInf) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k >= kids) break;
            const int j = (at >> (4 * k)) & 15;
            const unsigned long long key = tb.key(row0 + k, j);
            tb.set_path(row0 + k, li, (key & 0xFFFFFFFFULL) * kS + (unsigned long long)j);
        }
    }
    if (r.y == r.z) {  // the root's trace reads its layer costs
        if (!g) tb.set_root_cost(v, 0, kInf);
        tb.set_root_cost(v, li, best);
    }
    if (!g) tb.posof[v] = pos;
    tb.cvs[pos * L + li - 1] = best;
}

// One push item of a depth: edge e of a node at that depth towards target
// layer t of its parent's slot, with the additive costs of every from-layer
// of the parity that runs along the edge (the second half of
// solve_flt_node, :81-141, split by (edge, target) across lanes).  Its
// global loads (prefix sums, layer-change costs) depend only on the edge, so
// they are issued before the depth's nodes are solved.
// An edge has H = ceil(L / 2) items: item mi of an edge that changes its
// coordinate targets layer par + 2 mi (the only layers of the usable
// parity); of a zero-length edge, layers mi + 1 and mi + 1 + H.
template <int L>
struct PushItem {
    static constexpr int H = (L + 1) / 2;
    int t = 0;           // target layer, 0: no item
    int t2 = 0;          // zero-length edge: second target layer, 0: none
    int v = 0, row = 0;  // child node, parent slot row
    int par = 1;         // first from-layer of the usable parity
    bool zero = false;   // zero-length edge: straight on every layer
    bool far = false;    // long enough to change layers (> 4)
    double add[H];       // per from-layer par + 2i: wire cost (its layer == t) or layer-change cost
    double add2 = 0.0;   // zero-length edge: the wire cost on layer t2
};

template <int L>
__device__ __forceinline__ PushItem<L> load_item(const FltArgs& a, const Dag& d, const LdsTables<L>& tb, int e0,
                                                 int q) {
    constexpr int H = PushItem<L>::H;
    const DeviceGrid& g = a.g;
    PushItem<L> it;
    const int e = tb.eorder[e0 + q / H], mi = q % H;
    const int4 er = tb.erec[e];  // {parent, slot, parent coordinate, child node}
    const int v = er.w;
    const unsigned cxy = tb.xy[v], pxy = tb.xy[er.x];  // the parent's cell is er.z
    const int x = int(cxy >> 16), y = int(cxy & 0xFFFFu), px = int(pxy >> 16), py = int(pxy & 0xFFFFu);
    const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
    it.v = v;
    it.row = (tb.nrec[er.x].w >> 3) + er.y;
    it.zero = x == px && y == py;
    it.far = abs(x - px) + abs(y - py) > 4;
    if (it.zero) {
        it.t = mi + 1;
        it.add[0] = wire(a, mi, minx, maxx, miny, maxy);
        if (mi + 1 + H <= L) {
            it.t2 = mi + 1 + H;
            it.add2 = wire(a, mi + H, minx, maxx, miny, maxy);
        }
        return it;
    }
    // a layer runs along y when ((layer & 1) ^ DIR) == 0 (:101-102)
    it.par = (x == px) == (g.DIR == 1) ? 1 : 2;
    const int t = it.par + 2 * mi;
    if (t > L) return it;  // it.t stays 0
    it.t = t;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const int j = it.par + 2 * i;
        if (j > L) continue;
        if (j == t) it.add[i] = wire(a, j - 1, minx, maxx, miny, maxy);
        else if (it.far) it.add[i] = d.cost_edges[change_slot(e, j - 1, t - 1)];
    }
    return it;
}

template <int L>
__device__ __forceinline__ void finish_item(const FltArgs& a, const LdsTables<L>& tb, const PushItem<L>& it) {
    constexpr int H = PushItem<L>::H;
    if (!it.t) return;
    const double* cv = tb.cvs + tb.posof[it.v] * L - 1;  // cv[layer], layer 1..L
    if (it.zero) {
        const double base = cv[it.t];
        if (!(base < kInf))  // the reference asserts here
            raise(a.error, kDevNoSolution);
        else
            tb.lower(it.row, it.t, pack((float)(base + it.add[0]), it.v * 100 + kStraight));
        if (it.t2) {
            const double base2 = cv[it.t2];
            if (!(base2 < kInf))
                raise(a.error, kDevNoSolution);
            else
                tb.lower(it.row, it.t2, pack((float)(base2 + it.add2), it.v * 100 + kStraight));
        }
        return;
    }
    unsigned long long best = kEmpty;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const int j = it.par + 2 * i;
        if (j > L) continue;
        const double base = cv[j];
        if (!(base < kInf)) {  // the reference asserts here
            raise(a.error, kDevNoSolution);
            continue;
        }
        double cost2;
        int choice = kStraight;
        if (j == it.t) {
            cost2 = base + it.add[i];
        } else if (it.far) {
            if (!(it.add[i] < kInf)) continue;
            cost2 = base + it.add[i];
            choice = (j - 1) * 9 + (it.t - 1);
        } else {
            continue;
        }
        const unsigned long long nv = pack((float)cost2, it.v * 100 + choice);
        if (nv < best) best = nv;
    }
    if (best < kEmpty) tb.lower(it.row, it.t, best);
}

// One DAG on W lanes with its tables and records in LDS.  Per depth: the
// push items' global loads are issued first, the depth's nodes are solved
// into LDS, then the items lower their parents' slot keys; every other
// access of the depth loop stays in LDS.
template <int L>
__device__ __forceinline__ void route_dag_lds(const FltArgs& a, int k, bool live, int lane_in, int W, const Dag& dg,
                                              const LdsTables<L>& tb, const int4 ds, const int4 dn, int depths) {
    const int lane = lane_in;
    // stage the DAG: child slots to the empty key, records, orders, via costs
    for (int i = lane; i < dn.z * L; i += W) tb.packed[i] = kEmpty;
    for (int v = lane; v < dn.x; v += W) {
        const int4 r = dg.nodes[v];
        tb.nrec[v] = r;
        tb.order[v] = a.order[ds.x + v];
        const int x = r.x / a.g.Y % a.g.X, y = r.x % a.g.Y;
        tb.xy[v] = (unsigned)x << 16 | (unsigned)y;
        for (int j = 0; j + 1 < L; ++j) tb.vc[v * (L - 1) + j] = a.vcost[cell_of(a.g, j, x, y)];
        for (int e = r.y; e < r.z; ++e) {
            const int4 er = dg.edges[e];
            tb.erec[e] = make_int4(er.x, er.y, er.z, v);
        }
    }
    for (int i = lane; i < dn.y; i += W) tb.eorder[i] = a.eorder[ds.y + i];
    for (int i = lane; i <= dn.w; i += W) {
        tb.doff[i] = a.depth_off[ds.w + i];
        tb.edoff[i] = a.edepth_off[ds.w + i];
    }
    LdsTables<L>::barrier();
    Dag d = dg;
    d.nodes = tb.nrec;
    d.edges = tb.erec;  // {parent, slot, parent coordinate, child node}: the trace reads the first three
    for (int dep = depths - 1; dep >= 0; --dep) {
        const int lane = opaque(lane_in);  // lane-derived addresses formed per depth
        const bool here = dep < dn.w;
        const int e0 = here ? tb.edoff[dep] : 0, items = here ? (tb.edoff[dep + 1] - e0) * PushItem<L>::H : 0;
        PushItem<L> first;
        if (lane < items) first = load_item<L>(a, d, tb, e0, lane);
        if (here) {
            const int n0 = tb.doff[dep], end = tb.doff[dep + 1];
            // A node on kNodeLanes lanes costs about a quarter of a node on
            // one lane: the lane-parallel solve takes depths whose nodes it
            // covers in at most four times the single-lane passes.
            constexpr int G = kNodeLanes<L>;
            const int per = W / G, width = end - n0;
            if (per && (width + per - 1) / per <= 4 * ((width + W - 1) / W)) {
                for (int i = n0 + lane / G; i < end; i += per) {  // uniform over each node's G lanes
                    const int v = tb.order[i];
                    const int4 r = tb.nrec[v];
                    const int kids = r.w & 7;
                    if (kids <= 2)
                        solve_node_lanes<L, 2>(a, tb, v, r, i - n0, lane % G);
                    else if (kids <= 4)
                        solve_node_lanes<L, 4>(a, tb, v, r, i - n0, lane % G);
                    else
                        solve_node_lanes<L, 6>(a, tb, v, r, i - n0, lane % G);
                }
            } else {
                for (int i = n0 + lane; i < end; i += W) {
                    const int v = tb.order[i];
                    const int4 r = tb.nrec[v];
                    if ((r.w & 7) <= 2)
                        solve_node_fast<L, 2, false>(a, d, tb, v, r, i - n0);
                    else
                        solve_node_wide<L>(a, d, tb, v, r, i - n0);
                }
            }
        }
        LdsTables<L>::barrier();
        for (int q = lane; q < items; q += W) finish_item<L>(a, tb, q == lane ? first : load_item<L>(a, d, tb, e0, q));
        LdsTables<L>::barrier();
    }
    const int id = live ? a.net_ids[k] : 0;
    int* routes = a.routes.routes + (live ? (size_t)a.routes.pin_acc_num[id] * a.routes.per_pin : 0);
    const long long cap = live ? (long long)(a.pin_acc_num[id + 1] - a.pin_acc_num[id]) * a.routes.per_pin : 0;
    const int g_cells = a.g.L * a.g.XY;
    // trace pass one: the layers, depth by depth from the root
    for (int dep = 0; dep < dn.w; ++dep) {
        const int end = tb.doff[dep + 1];
        for (int i = tb.doff[dep] + lane; i < end; i += W) trace_layer<L>(a, d, tb, tb.order[i]);
        LdsTables<L>::barrier();
    }
    // pass two: the words of every node in depth order, appended through a
    // prefix sum over the DAG's lanes
    const int n0 = dn.w ? tb.doff[0] : 0, total = dn.w ? tb.doff[dn.w] - n0 : 0;
    const int chunks = wave_max((total + W - 1) / W);
    int count = 1;  // the root's trace sets the word count to 1
    for (int ch = 0; ch < chunks; ++ch) {
        Words wd;
        const int q = ch * W + lane;
        if (q < total) trace_words<L>(a, d, tb, tb.order[n0 + q], wd);
        if (!wd.in_grid(g_cells)) {
            raise(a.error, kDevBadCoordinate);
            wd.ne = wd.nv = 0;
        }
        const int n = wd.count();
        int incl = n;
        for (int s = 1; s < W; s <<= 1) {
            const int o = __shfl_up(incl, s, W);
            if (lane >= s) incl += o;
        }
        const int at = count + incl - n;
        if (n) {
            if (at + n - 1 >= cap) {
                raise(a.error, kDevRouteOverflow);
            } else {
                wd.put(routes, at);
            }
        }
        count += __shfl(incl, W - 1, W);
    }
    if (live && lane == 0) routes[0] = count <= cap ? count : 1;
}

// The records of DAG k of the batch and its layer-change tables.
__device__ __forceinline__ Dag dag_of(const FltArgs& a, const int4 ds) {
    Dag d;
    d.nodes = a.node_rec + ds.x;
    d.edges = a.edge_rec + ds.y;
    d.cost_edges = a.cost_edges + change_slot(ds.y - a.edge_base, 0, 0);
    d.best_change = a.best_change + change_slot(ds.y - a.edge_base, 0, 0);
    return d;
}

}  // namespace

// precompute_cost_edge (src/routing/flt_cost.cuh:50-130) for the batch's
// edges longer than 4 (the DP and the trace read no other change tables; a
// list built at upload).  A workgroup takes edge_group(L) of them: it first
// gathers, at every sample
// position of each edge, the via costs of all layers and the prefix sums of
// the layers that run along the edge into LDS (each value once, instead of
// once per layer pair), then one thread per (edge, from-layer, same-parity
// to-layer) scans the positions from LDS, summing exactly what the
// reference sums in the same order.
constexpr int kEdgeThreads = 256;
constexpr int kEdgePositions = 40;  // sample positions of an edge: at most 38 (length 39)
__host__ __device__ constexpr int edge_group(int L) { return kEdgeThreads / (L * kFltChangeWidth); }

__global__ __launch_bounds__(kEdgeThreads) void flt_edge_kernel(FltArgs a) {
    const DeviceGrid& g = a.g;
    const int L = g.L;
    const int E = edge_group(L);
    struct EdgeInfo {
        int fromx, fromy, tox, toy, start, dd, count, vertical, near;
    };
    __shared__ EdgeInfo info[kEdgeThreads / kFltChangeWidth];
    extern __shared__ __align__(16) char smem[];
    // per edge of the group: vc[k * (L - 1) + l] (float), ps[k * L + l] (double), k = position - 1
    float* vcs = reinterpret_cast<float*>(smem);
    double* pss = reinterpret_cast<double*>(smem + size_t(E) * kEdgePositions * (L - 1) * 4 + 8);
    const int e0 = blockIdx.x * E;  // position in the batch's far-edge list
    if (threadIdx.x < E) {
        EdgeInfo& in = info[threadIdx.x];
        in.count = 0;
        in.near = 1;
        if (e0 + int(threadIdx.x) < a.far_count) {
            const int4 er = a.edge_rec[a.far_edges[e0 + threadIdx.x]];
            in.tox = er.z / g.Y % g.X, in.toy = er.z % g.Y;
            in.fromx = er.w / g.Y % g.X, in.fromy = er.w % g.Y;
            const int distance = abs(in.tox - in.fromx) + abs(in.toy - in.fromy);
            in.vertical = in.fromx == in.tox;
            in.start = in.vertical ? min(in.fromy, in.toy) : min(in.fromx, in.tox);
            in.dd = distance ? distance / min(distance, 20) : 1;
            // positions start + dd * k < start + distance, k >= 1
            in.count = distance ? (distance - 1) / in.dd : 0;
            // the DP and the trace read the change tables of edges longer
            // than 4 only (solve_flt_node's distance test): shorter edges
            // are skipped
            in.near = distance <= 4;
        }
    }
    __syncthreads();
    // gather: every (edge, position, value) of the group, the positions of
    // an edge being its cells at start + dd * k
    const int per_pos = 2 * L - 1;  // L - 1 via costs, then L prefix sums
    for (int q = threadIdx.x; q < E * kEdgePositions * per_pos; q += kEdgeThreads) {
        const int j = q / (kEdgePositions * per_pos), rem = q % (kEdgePositions * per_pos);
        const int k = rem / per_pos, v = rem % per_pos;
        const EdgeInfo& in = info[j];
        if (in.near || k >= in.count || k >= kEdgePositions) continue;
        const int sp = in.start + in.dd * (k + 1);
        const int sx = in.vertical ? in.fromx : sp, sy = in.vertical ? sp : in.fromy;
        if (v < L - 1) {
            vcs[(j * kEdgePositions + k) * (L - 1) + v] = a.vcost[cell_of(g, v, sx, sy)];
        } else {
            const int l = v - (L - 1);
            // only the layers along the edge are read
            if (layer_dir(g, l) == (in.vertical ? 1 : 0)) pss[(j * kEdgePositions + k) * L + l] = a.presum[cell_of(g, l, sx, sy)];
        }
    }
    __syncthreads();
    const int j = threadIdx.x / (L * kFltChangeWidth), r = threadIdx.x % (L * kFltChangeWidth);
    if (j >= E || e0 + j >= a.far_count) return;
    const int e = a.far_edges[e0 + j] - a.edge_base;  // the batch's edge index
    const int f = r / kFltChangeWidth, to = (f & 1) + 2 * (r % kFltChangeWidth);
    // to == f is the straight edge, which the DP prices itself
    if (to >= L || to == f) return;
    const EdgeInfo& in = info[j];
    // a from-layer across the edge: the reference fills its row with
    // infinity; the DP never reads it (its direction test comes first)
    if ((layer_dir(g, f) == 0 && in.fromy != in.toy) || (layer_dir(g, f) == 1 && in.fromx != in.tox)) return;
    if (in.near) return;  // zero length or not read
    double best = kInf;
    int pos = -1;
    const int lo = min(f, to), hi = max(f, to);
    if (in.count <= kEdgePositions) {
        // the wires from each end to the position: prefix sums at the far
        // corner minus the near corner, as wire() forms them
        const double pf = a.presum[cell_of(g, f, in.fromx, in.fromy)], pt = a.presum[cell_of(g, to, in.tox, in.toy)];
        const int fa = in.vertical ? in.fromy : in.fromx, ta = in.vertical ? in.toy : in.tox;
        const float* vc = vcs + size_t(j) * kEdgePositions * (L - 1);
        const double* ps = pss + size_t(j) * kEdgePositions * L;
        for (int k = 0; k < in.count; ++k) {
            const int sp = in.start + in.dd * (k + 1);
            double cost = 0.0;
            for (int l = lo; l < hi; ++l) cost += (double)vc[k * (L - 1) + l];
            const double sf = ps[k * L + f], st = ps[k * L + to];
            const double wf = fa < sp ? sf - pf : pf - sf;
            const double wt = ta < sp ? st - pt : pt - st;
            cost += wf + wt;
            if (cost < best) {
                best = cost;
                pos = sp;
            }
        }
    } else {  // longer than any edge the sampling rule gives: the direct search
        for (int sp = in.start + in.dd; sp < in.start + in.dd * (in.count + 1); sp += in.dd) {
            const int sx = in.vertical ? in.fromx : sp, sy = in.vertical ? sp : in.fromy;
            double cost = 0.0;
            for (int l = lo; l < hi; ++l) cost += (double)a.vcost[cell_of(g, l, sx, sy)];
            if (in.vertical)
                cost += wire(a, f, in.fromx, in.fromx, min(in.fromy, sy), max(in.fromy, sy)) +
                        wire(a, to, in.tox, in.tox, min(in.toy, sy), max(in.toy, sy));
            else
                cost += wire(a, f, min(in.fromx, sx), max(in.fromx, sx), in.fromy, in.fromy) +
                        wire(a, to, min(in.tox, sx), max(in.tox, sx), in.toy, in.toy);
            if (cost < best) {
                best = cost;
                pos = sp;
            }
        }
    }
    a.cost_edges[change_slot(e, f, to)] = best;
    a.best_change[change_slot(e, f, to)] = pos;
}

void launch_flt_edges(const FltArgs& a, hipStream_t stream) {
    const int L = a.g.L, E = edge_group(L);
    if (!a.far_count || L < 2) return;
    const size_t lds = size_t(E) * kEdgePositions * ((L - 1) * 4 + L * 8) + 8;
    hipLaunchKernelGGL(flt_edge_kernel, dim3(unsigned((a.far_count + E - 1) / E)), dim3(kEdgeThreads), lds, stream, a);
}

namespace {

// route_flt_block (src/routing/flt_block.cuh:334-355), tables in global
// memory (DAGs of any size): a.per_wave DAGs per wavefront (1, 2, 4 or 8;
// each gets 64 / per_wave lanes), four wavefronts per workgroup.  The DAGs
// of a wavefront step through their depths together, so a batch is
// uploaded with its DAGs in decreasing depth order.
template <int L>
__global__ __launch_bounds__(256) void flt_route_kernel(FltArgs a) {
    const int P = a.per_wave, W = kWave / P;
    const int wl = threadIdx.x & (kWave - 1);
    const int sub = wl / W, lane = wl % W;  // lane within this DAG's group
    const int first_dag = (blockIdx.x * (256 / kWave) + threadIdx.x / kWave) * P;
    if (first_dag >= a.count) return;  // the whole wavefront
    const int k = first_dag + sub;
    const bool live = k < a.count;
    const int4 ds = live ? a.desc[k] : make_int4(0, 0, 0, 0);  // stage-global first node, edge, slot, depth offset
    const int4 dn = live ? a.sizes[k] : make_int4(0, 0, 0, 0);  // nodes, edges, slots, depths
    const Dag d = dag_of(a, ds);
    GlobalTables tb;
    tb.costs = a.costs + (size_t)(ds.x - a.node_base) * kS;
    tb.packed = a.packed + (size_t)(ds.z - a.slot_base) * kS;
    tb.best = a.best_path + (size_t)(ds.z - a.slot_base) * kS;
    tb.layer = a.layer_output + (ds.x - a.node_base);
    tb.vcost = a.vcost;
    tb.g = a.g;
    const int depths = wave_max(dn.w);  // the wavefront's deepest DAG
    route_dag<L>(a, k, live, lane, W, d, tb, dn, a.order + ds.x, a.depth_off + ds.w, depths);
}

// The same DP with each DAG's tables in LDS: one wavefront per workgroup,
// a.per_wave DAGs per wavefront, each with a.lds_dag_bytes of LDS (sized by
// the batch's largest DAG).  Without a queue, workgroup w takes DAG group w;
// with one (the persistent form), a grid no larger than what stays resident
// drains the batch's DAG groups in order (deepest DAGs first) from
// kFltQueues striped counters, zero at launch: workgroup w dequeues from
// counter w % kFltQueues, which hands out groups q, q + kFltQueues, ...
template <int L>
// two wavefronts per SIMD (256 registers) hold without spills
__global__ __launch_bounds__(kWave, 2) void flt_route_lds_kernel(FltArgs a) {
    extern __shared__ __align__(16) char smem[];
    __shared__ int s_group;
    const int P = a.per_wave, W = kWave / P;
    const int sub = threadIdx.x / W, lane = threadIdx.x % W;
    const int q = blockIdx.x % kFltQueues;
    int* counter = a.queue ? a.queue + q * kFltQueueStride : nullptr;
    for (int round = 0;; ++round) {
        int group;
        if (counter) {
            if (threadIdx.x == 0) s_group = atomicAdd(counter, 1) * kFltQueues + q;
            __syncthreads();
            group = s_group;
            __syncthreads();  // every lane has read s_group before the next dequeue
        } else {
            if (round) break;
            group = blockIdx.x;
        }
        const int first_dag = group * P;
        if (first_dag >= a.count) break;
        const int k = first_dag + sub;
        const bool live = k < a.count;
        const int4 ds = live ? a.desc[k] : make_int4(0, 0, 0, 0);
        const int4 dn = live ? a.sizes[k] : make_int4(0, 0, 0, 0);
        const Dag d = dag_of(a, ds);
        LdsTables<L> tb;
        lds_layout<L>(smem + (size_t)sub * a.lds_dag_bytes, a.lds_shape, tb);
        const int depths = wave_max(dn.w);
        route_dag_lds<L>(a, k, live, lane, W, d, tb, ds, dn, depths);
        __syncthreads();  // the tables are reused by the next group
    }
}

template <int L>
void launch_for(const FltArgs& a, int mode_lds, int cus, hipStream_t stream) {
    if (!mode_lds) {
        const int blocks = (a.count + 4 * a.per_wave - 1) / (4 * a.per_wave);
        hipLaunchKernelGGL(flt_route_kernel<L>, dim3(blocks), dim3(256), 0, stream, a);
        return;
    }
    const size_t lds = (size_t)a.lds_dag_bytes * a.per_wave;
    const int groups = (a.count + a.per_wave - 1) / a.per_wave;
    int grid = groups;
    if (a.queue) {
        int resident = 0;  // workgroups per compute unit; any value keeps the queue correct
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, flt_route_lds_kernel<L>, kWave, lds) != hipSuccess ||
            resident < 1)
            resident = 1;
        grid = std::max(1, std::min(groups, resident * cus));
    }
    hipLaunchKernelGGL(flt_route_lds_kernel<L>, dim3(grid), dim3(kWave), lds, stream, a);
}

}  // namespace

size_t flt_lds_bytes(int L, int slots, int nodes, int edges, int dnodes, int depths) {
    const LdsShape shape{slots, nodes, edges, dnodes, depths};
    switch (L) {
#define GR_FLT_BYTES(n) \
    case n: {                                                    \
        LdsTables<n> t;                                          \
        return lds_layout<n>(nullptr, shape, t);                 \
    }
        GR_FLT_BYTES(1)
        GR_FLT_BYTES(2)
        GR_FLT_BYTES(3)
        GR_FLT_BYTES(4)
        GR_FLT_BYTES(5)
        GR_FLT_BYTES(6)
        GR_FLT_BYTES(7)
        GR_FLT_BYTES(8)
        GR_FLT_BYTES(9)
#undef GR_FLT_BYTES
        default: return 0;
    }
}

void launch_flt_route(const FltArgs& a, bool lds, int cus, hipStream_t stream) {
    switch (a.g.L) {
#define GR_FLT_CASE(n) \
    case n: launch_for<n>(a, lds, cus, stream); break;
        GR_FLT_CASE(1)
        GR_FLT_CASE(2)
        GR_FLT_CASE(3)
        GR_FLT_CASE(4)
        GR_FLT_CASE(5)
        GR_FLT_CASE(6)
        GR_FLT_CASE(7)
        GR_FLT_CASE(8)
        GR_FLT_CASE(9)
#undef GR_FLT_CASE
        default: break;
    }
}

}  // namespace grotor
