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

// One DAG's tables, net-local.
struct Dag {
    const int4* nodes;  // {coordinate, first edge, end edge, first slot * 8 + slots}
    const int4* edges;  // {parent, slot, parent coordinate, child coordinate}
    double* costs;      // [node * kS + layer]
    unsigned long long* packed;
    unsigned long long* best;
    int* layer;
    const double* cost_edges;  // [change_slot(edge, from, to)]
    const int* best_change;
};

// solve_flt_node (src/routing/flt_block.cuh:6-142), tables in global memory
// (any number of child slots).
__device__ void solve_node(const FltArgs& a, const Dag& d, int v) {
    const DeviceGrid& g = a.g;
    const int L = g.L;
    const int4 r = d.nodes[v];
    const int low_fixed = 1 + r.x / g.XY;
    const int x = r.x / g.Y % g.X, y = r.x % g.Y;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const int kids = r.w & 7;
    unsigned long long* bp = d.best + (size_t)(r.w >> 3) * kS;
    const unsigned long long* pk = d.packed + (size_t)(r.w >> 3) * kS;
    double* cv = d.costs + (size_t)v * kS;
    for (int li = 0; li < kS; ++li) cv[li] = kInf;
    double mins[6];
    unsigned long long paths[6];
    for (int low = 1; low <= low_fixed; ++low) {
        for (int k = 0; k < kids; ++k) mins[k] = kInf;
        double via = 0.0;
        for (int li = low; li < L + 1; ++li) {
            if (li > low) via += (double)a.vcost[cell_of(g, li - 2, x, y)];
            for (int k = 0; k < kids; ++k) {
                const unsigned long long p = l2_load(pk + k * kS + li);
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
                    for (int k = 0; k < kids; ++k) bp[k * kS + li] = paths[k];
                }
            }
        }
        for (int li = L - 1; li >= low; --li)
            if (cv[li + 1] < cv[li]) {
                cv[li] = cv[li + 1];
                for (int k = 0; k < kids; ++k) bp[k * kS + li] = bp[k * kS + li + 1];
            }
    }
    unsigned long long cache[kS];
    for (int e = r.y; e < r.z; ++e) {
        const int4 er = d.edges[e];
        unsigned long long* ppk = d.packed + (size_t)((d.nodes[er.x].w >> 3) + er.y) * kS;
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
            if (cache[t] < kEmpty) atomicMin(ppk + t, cache[t]);
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
template <int L>
__device__ __forceinline__ void push_parents(const FltArgs& a, const Dag& d, int v, int x, int y, const int4 r, const double* cv) {
    const DeviceGrid& g = a.g;
    constexpr int H = (L + 1) / 2;  // layers of one parity, at most
    for (int e = r.y; e < r.z; ++e) {
        const int4 er = d.edges[e];
        unsigned long long* ppk = d.packed + (size_t)((d.nodes[er.x].w >> 3) + er.y) * kS;
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
                atomicMin(ppk + layer, nv);
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
            if (cache[m] < kEmpty && par + 2 * m <= L) atomicMin(ppk + par + 2 * m, cache[m]);
    }
}

template <int L, int K>
__device__ __forceinline__ void solve_node_fast(const FltArgs& a, const Dag& d, int v, const int4 r) {
    const DeviceGrid& g = a.g;
    const int kids = r.w & 7;
    const int low_fixed = 1 + r.x / g.XY;
    const int x = r.x / g.Y % g.X, y = r.x % g.Y;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const unsigned long long* pk = d.packed + (size_t)(r.w >> 3) * kS;
    // slot keys, loaded once: the costs for the pass below, the codes for
    // the best-path rows
    unsigned long long pkv[K][L + 1];
    float pv[K][L + 1];  // slot costs (the packed key's float)
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int li = 1; li <= L; ++li) {
            pkv[k][li] = k < kids ? l2_load(pk + k * kS + li) : 0ull;
            pv[k][li] = k < kids ? (float)unpack_cost(pkv[k][li]) : 0.f;
        }
    float vc[L];
#pragma unroll
    for (int j = 0; j + 1 < L; ++j) vc[j] = a.vcost[cell_of(g, j, x, y)];
    double cv[L + 1];
    unsigned at[L + 1];  // per layer: the slot positions of its cost, 4 bits per child slot
#pragma unroll
    for (int li = 0; li <= L; ++li) cv[li] = kInf, at[li] = 0;
    // passes over the stack's low end stay a loop; every layer index inside
    // is a compile-time constant, so the tables stay in registers
#pragma unroll 1
    for (int low = 1; low <= min(L, low_fixed); ++low) {
        {
            double mins[K];
#pragma unroll
            for (int k = 0; k < K; ++k) mins[k] = kInf;
            unsigned run = 0;
            double via = 0.0;
#pragma unroll
            for (int li = 1; li <= L; ++li) {
                if (li < low) continue;
                if (li > low) via += (double)vc[li - 2];
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (k < kids) {
                        const double c = (double)pv[k][li];
                        if (c < mins[k]) {
                            mins[k] = c;
                            run = (run & ~(15u << (4 * k))) | ((unsigned)li << (4 * k));
                        }
                    }
                if (li >= high_fixed) {
                    double cost = via;
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        if (k < kids) cost += mins[k];
                    if (cost < kInf && cost < cv[li]) {
                        cv[li] = cost;
                        at[li] = run;
                    }
                }
            }
#pragma unroll
            for (int li = L - 1; li >= 1; --li)
                if (li >= low && cv[li + 1] < cv[li]) {
                    cv[li] = cv[li + 1];
                    at[li] = at[li + 1];
                }
        }
    }
    unsigned long long* bp = d.best + (size_t)(r.w >> 3) * kS;
#pragma unroll
    for (int li = 1; li <= L; ++li) {
        if (!(cv[li] < kInf)) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k >= kids) continue;
            const int j = (at[li] >> (4 * k)) & 15;
            unsigned long long key = 0;  // pkv[k][j], j a run-time layer
#pragma unroll
            for (int lj = 1; lj <= L; ++lj)
                if (lj == j) key = pkv[k][lj];
            bp[k * kS + li] = (key & 0xFFFFFFFFULL) * kS + (unsigned long long)j;
        }
    }
    if (r.y == r.z) {  // the root's trace reads its layer costs
        double* cg = d.costs + (size_t)v * kS;
        cg[0] = kInf;
#pragma unroll
        for (int li = 1; li <= L; ++li) cg[li] = cv[li];
    }
    push_parents<L>(a, d, v, x, y, r, cv);
}

// trace_flt_node (src/routing/flt_block.cuh:144-331): the node's layer, the
// segments of the edge its chosen parent uses and its via stack, into w[].
__device__ int trace_node(const FltArgs& a, const Dag& d, int v, int* w) {
    const DeviceGrid& g = a.g;
    const int L = g.L;
    const int4 r = d.nodes[v];
    const int l = r.x / g.XY, cx = r.x / g.Y % g.X, cy = r.x % g.Y;
    const unsigned long long* bp = d.best + (size_t)(r.w >> 3) * kS;
    int n = 0;
    int lv;
    if (r.y == r.z) {
        int best = 0;
        double bc = d.costs[(size_t)v * kS];
        for (int layer = 1; layer < L + 1; ++layer) {
            const double c = d.costs[(size_t)v * kS + layer];
            if (c < bc) {
                bc = c;
                best = layer;
            }
        }
        lv = best;
        d.layer[v] = lv;
        if (!lv) {  // no finite cost at any layer: nothing to trace
            raise(a.error, kDevNoSolution);
            return 0;
        }
    } else {
        int par_layer = -1, par = -1, par_edge = -1;
        lv = -1;
        for (int e = r.y; e < r.z; ++e) {
            const int4 er = d.edges[e];
            const int pl = l2_load(d.layer + er.x);
            if (pl <= 0) continue;
            par = er.x;
            par_edge = e;
            par_layer = pl;
            const unsigned long long path = l2_load(d.best + (size_t)((d.nodes[par].w >> 3) + er.y) * kS + par_layer);
            const unsigned long long child = path / kS / 100, choice = (path / kS) % 100;
            if (child != (unsigned long long)v) {
                lv = -1;
                continue;
            }
            lv = choice == kStraight ? (int)(path % kS) : (int)(choice / 9 + 1);
            const int px = er.z / g.Y % g.X, py = er.z % g.Y;
            const int rl = lv - 1;
            auto idx = [&](int ll, int xx, int yy) { return (ll * g.X + xx) * g.Y + yy; };
            if (px == cx && cy != py) {
                if (choice == kStraight) {
                    w[n++] = idx(rl, px, min(cy, py));
                    w[n++] = idx(rl, px, max(cy, py));
                } else {
                    const int fl = (int)choice / 9, tl = (int)choice % 9;
                    const int pos = d.best_change[change_slot(e, fl, tl)];
                    w[n++] = idx(fl, cx, min(cy, pos));
                    w[n++] = idx(fl, cx, max(cy, pos));
                    w[n++] = idx(min(fl, tl), cx, pos);
                    w[n++] = idx(max(fl, tl), cx, pos);
                    w[n++] = idx(tl, cx, min(py, pos));
                    w[n++] = idx(tl, cx, max(py, pos));
                }
            } else if (py == cy && cx != px) {
                if (choice == kStraight) {
                    w[n++] = idx(rl, min(cx, px), py);
                    w[n++] = idx(rl, max(cx, px), py);
                } else {
                    const int fl = (int)choice / 9, tl = (int)choice % 9;
                    const int pos = d.best_change[change_slot(e, fl, tl)];
                    w[n++] = idx(fl, min(cx, pos), cy);
                    w[n++] = idx(fl, max(cx, pos), cy);
                    w[n++] = idx(min(fl, tl), pos, cy);
                    w[n++] = idx(max(fl, tl), pos, cy);
                    w[n++] = idx(tl, min(px, pos), cy);
                    w[n++] = idx(tl, max(px, pos), cy);
                }
            }
            break;
        }
        if (par_layer == -1) {
            d.layer[v] = -1;
            return 0;
        }
        const int4 er = d.edges[par_edge];
        const unsigned long long path = l2_load(d.best + (size_t)((d.nodes[par].w >> 3) + er.y) * kS + par_layer);
        if (path / kS / 100 != (unsigned long long)v) {
            d.layer[v] = -1;
            return 0;
        }
        d.layer[v] = lv;
    }
    const int kids = r.w & 7;
    int minl = l + 1;
    int maxl = (l + 1) == L + 1 ? 1 : minl;
    minl = min(minl, lv);
    maxl = max(maxl, lv);
    for (int k = 0; k < kids; ++k) {
        const int lc = (int)(l2_load(bp + k * kS + lv) % kS);
        minl = min(minl, lc);
        maxl = max(maxl, lc);
    }
    if (minl < maxl) {
        w[n++] = ((minl - 1) * g.X + cx) * g.Y + cy;
        w[n++] = ((maxl - 1) * g.X + cx) * g.Y + cy;
    }
    return n;
}


}  // namespace

// precompute_cost_edge (src/routing/flt_cost.cuh:50-130); one thread per
// (edge, from-layer, to-layer of the same parity) of the batch.
__global__ void flt_edge_kernel(FltArgs a) {
    const DeviceGrid& g = a.g;
    const int L = g.L;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.edges * L * kFltChangeWidth) return;
    const int e = (int)(t / (L * kFltChangeWidth));
    const int r = (int)(t % (L * kFltChangeWidth));
    const int f = r / kFltChangeWidth, to = (f & 1) + 2 * (r % kFltChangeWidth);
    // to == f is the straight edge, which the DP prices itself
    if (to >= L || to == f) return;
    const int4 er = a.edge_rec[a.edge_base + e];
    const int tox = er.z / g.Y % g.X, toy = er.z % g.Y;
    const int fromx = er.w / g.Y % g.X, fromy = er.w % g.Y;
    // a from-layer across the edge: the reference fills its row with
    // infinity; the DP never reads it (its direction test comes first)
    if ((layer_dir(g, f) == 0 && fromy != toy) || (layer_dir(g, f) == 1 && fromx != tox)) return;
    const int distance = abs(tox - fromx) + abs(toy - fromy);
    if (!distance) return;
    const int dd = distance / min(distance, 20);
    const bool vertical = fromx == tox;
    const int start = vertical ? min(fromy, toy) : min(fromx, tox);
    const int end = vertical ? max(fromy, toy) : max(fromx, tox);
    double best = kInf;
    int pos = -1;
    for (int s = start + dd; s < end; s += dd) {
        const int sx = vertical ? fromx : s, sy = vertical ? s : fromy;
        double cost = 0.0;
        for (int l = min(f, to); l < max(f, to); ++l) cost += (double)a.vcost[cell_of(g, l, sx, sy)];
        if (vertical)
            cost += wire(a, f, fromx, fromx, min(fromy, sy), max(fromy, sy)) +
                    wire(a, to, tox, tox, min(toy, sy), max(toy, sy));
        else
            cost += wire(a, f, min(fromx, sx), max(fromx, sx), fromy, fromy) +
                    wire(a, to, min(tox, sx), max(tox, sx), toy, toy);
        if (cost < best) {
            best = cost;
            pos = s;
        }
    }
    a.cost_edges[change_slot(e, f, to)] = best;
    a.best_change[change_slot(e, f, to)] = pos;
}

// route_flt_block (src/routing/flt_block.cuh:334-355): a.per_wave DAGs per
// wavefront (1, 2, 4 or 8; each gets 64 / per_wave lanes), four wavefronts
// per workgroup.  The DAGs of a wavefront step through their depths
// together, so a batch is uploaded with its DAGs in decreasing depth order.
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
    Dag d;
    d.nodes = a.node_rec + ds.x;
    d.edges = a.edge_rec + ds.y;
    d.costs = a.costs + (size_t)(ds.x - a.node_base) * kS;
    d.packed = a.packed + (size_t)(ds.z - a.slot_base) * kS;
    d.best = a.best_path + (size_t)(ds.z - a.slot_base) * kS;
    d.layer = a.layer_output + (ds.x - a.node_base);
    d.cost_edges = a.cost_edges + change_slot(ds.y - a.edge_base, 0, 0);
    d.best_change = a.best_change + change_slot(ds.y - a.edge_base, 0, 0);
    const int* order = a.order + ds.x;
    const int* doff = a.depth_off + ds.w;
    int depths = dn.w;  // the wavefront's deepest DAG
    for (int off = kWave / 2; off; off >>= 1) depths = max(depths, __shfl_xor(depths, off, kWave));

    // initialize_batch (src/routing/flt_workspace.cuh:60-69) for this DAG:
    // only the child slots need it here (a node's layer costs start at
    // infinity in its solve, and every node's trace sets its layer output
    // before any child reads it)
    for (int i = lane; i < dn.z * kS; i += W) d.packed[i] = kEmpty;
    depth_barrier();
    for (int dep = depths - 1; dep >= 0; --dep) {
        if (dep < dn.w) {
            const int end = doff[dep + 1];
            for (int i = doff[dep] + lane; i < end; i += W) {
                const int v = order[i];
                const int4 r = d.nodes[v];
                if ((r.w & 7) <= 2)
                    solve_node_fast<L, 2>(a, d, v, r);
                else if ((r.w & 7) <= kFastKids)
                    solve_node_fast<L, kFastKids>(a, d, v, r);
                else
                    solve_node(a, d, v);
            }
        }
        depth_barrier();
    }
    const int id = live ? a.net_ids[k] : 0;
    int* routes = a.routes.routes + (live ? (size_t)a.routes.pin_acc_num[id] * a.routes.per_pin : 0);
    const long long cap = live ? (long long)(a.pin_acc_num[id + 1] - a.pin_acc_num[id]) * a.routes.per_pin : 0;
    const int g_cells = a.g.L * a.g.XY;
    int count = 1;  // the root's trace sets the word count to 1
    for (int dep = 0; dep < depths; ++dep) {
        const int first = dep < dn.w ? doff[dep] : 0, end = dep < dn.w ? doff[dep + 1] : 0;
        int chunks = (end - first + W - 1) / W;
        for (int off = kWave / 2; off; off >>= 1) chunks = max(chunks, __shfl_xor(chunks, off, kWave));
        for (int ch = 0; ch < chunks; ++ch) {
            int w[8];
            const int i = first + ch * W + lane;
            int n = i < end ? trace_node(a, d, order[i], w) : 0;
            for (int j = 0; j < n; ++j)
                if (w[j] < 0 || w[j] >= g_cells) {
                    raise(a.error, kDevBadCoordinate);
                    n = 0;
                }
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
                    for (int j = 0; j < n; ++j) routes[at + j] = w[j];
                }
            }
            count += __shfl(incl, W - 1, W);
        }
        depth_barrier();
    }
    if (live && lane == 0) routes[0] = count <= cap ? count : 1;
}

void launch_flt_route(const FltArgs& a, int blocks, hipStream_t stream) {
    switch (a.g.L) {
#define GR_FLT_CASE(n) \
    case n: hipLaunchKernelGGL(flt_route_kernel<n>, dim3(blocks), dim3(256), 0, stream, a); break;
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
