// L-shape routing-DAG dynamic program for gfx950.
//
// Reference: Lshape_route_warp with node_diagonal / emit_lshape_node
// (src/routing/lshape_warp.cuh:8-247) and its scalar twin Lshape_route_cuda
// (src/routing/lshape_thread.cuh:18-171).
//
// One 64-lane wavefront (one workgroup) owns 1, 2, 4 or 8 nets (nets per
// wavefront, a launch parameter chosen by the autotuner): each net gets
// 64/P lanes and its own tables.  Within a net, lanes run across layers and
// across tree nodes: the net's lanes form G = max(1, (64/P) / L) groups and
// every group solves one node, lane c of a group taking layer c.  Nodes are
// solved level by level, a node's level being its height (leaves 0, a parent
// one more than its highest child), so all nodes of a level are independent.
// Per node (lshape_warp.cuh:103-189):
//   G  lane lo: the node's interval table row lo receives its children's
//      closures (each child's best arrival per layer, closed over [lo, hi]
//      with the first minimum), children in descending preorder index, which
//      is the scalar kernel's order of accumulation, so every double is formed
//      by the same additions in the same order;
//   A  lane c: best interval containing layer c (first minimum in
//      (low, high) row-major order);
//   B  lane c: best arrival on parent layer c (straight edge, or the first
//      minimum over the opposite-parity child layers of a bent edge, with the
//      corner via column summed left to right as the reference does).
// Three wavefront barriers per level instead of four per node.  The
// traceback walks the levels top down, one lane per node, and the route
// words are placed by a wavefront prefix sum.  Heights, the level order and
// the child lists are computed on the host when the stage is uploaded.
// All costs the sweep reads (prefix-sum differences and via costs) are
// gathered for every node before the sweep starts, so the sequential part
// of the net touches LDS only.  The layer count is a template parameter.
// The interval tables live in LDS (or, for trees too large for LDS, in a
// per-workgroup slice of global scratch); traceback recomputes the closure
// argmin from the stored per-node best arrivals, so no per-interval
// traceback table is kept.  Results equal the reference's bit for bit.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

// Per-net tables for trees of up to nmax nodes; L is a compile-time
// constant so that every per-layer loop unrolls.
template <int L>
struct Tables {
    double* dist;   // nmax * T      interval tables, upper triangle (low <= high) row by row
    double* diag;   // nmax * L      best arrival on each parent layer
    double* w1;     // nmax * L      first-leg wire (child side; the whole edge when straight)
    double* w2;     // nmax * L      second-leg wire (parent side, bent edges)
    double* gbest;  // lanes         per lane group: best interval containing each layer
    float* vs;      // nmax * L      via costs on the node's column
    float* va;      // nmax * L      via costs at corner (x, py)
    float* vb;      // nmax * L      via costs at corner (px, y)
    int* dprev;     // nmax * L
    int* gchoice;   // lanes         per lane group: that interval, low * L + high
    int* nd;        // nmax
    int* par;       // nmax
    int* rng;       // nmax
    int* path;      // nmax
    int* xy;        // nmax          node position x << 16 | y
    int* lay;       // nmax          node layer (L for Steiner nodes)
    int* order;     // nmax          nodes by (height, index)
    int* link;      // nmax          (first child + 1) | (next smaller sibling + 1) << 16
    int* woff;      // nmax + 1
    int* lstart;    // nmax + 1      first order position of each level
};

// entries of an interval table: intervals low <= high of L layers
template <int L>
constexpr int kTri = L * (L + 1) / 2;

// position of interval (lo, hi), lo <= hi, in a triangular table
template <int L>
__host__ __device__ constexpr int tri(int lo, int hi) {
    return lo * (2 * L - lo + 1) / 2 + (hi - lo);
}

// Entries of the per-group arrays (gbest, gchoice) of a net with 64 /
// per_wave lanes: its lane groups use at most that many, and at least L
// (one group of L layers).
template <int L>
__host__ __device__ constexpr int net_lanes(int per_wave) {
    return kWave / per_wave > L ? kWave / per_wave : L;
}

// Bytes of one net's tables (lanes: net_lanes; the host sizes LDS fits and
// global scratch with kWave, an upper bound).  The per-group arrays come
// last, so the other arrays' offsets do not depend on the lane count.
template <int L>
__host__ __device__ constexpr size_t table_head_bytes(int nmax) {
    return 8 * (size_t(nmax) * (kTri<L> + 3 * L)) + 4 * (size_t(nmax) * 3 * L) +
           4 * (size_t(nmax) * L + 8 * size_t(nmax) + 2 * (size_t(nmax) + 1));
}
template <int L>
__host__ __device__ constexpr size_t table_bytes(int nmax, int lanes = kWave) {
    return (table_head_bytes<L>(nmax) + 7) / 8 * 8 + 12 * size_t(lanes);
}

template <int L>
__device__ inline Tables<L> carve(double* base, int nmax, int lanes) {
    Tables<L> s;
    s.dist = base;
    s.diag = s.dist + size_t(nmax) * kTri<L>;
    s.w1 = s.diag + size_t(nmax) * L;
    s.w2 = s.w1 + size_t(nmax) * L;
    s.vs = reinterpret_cast<float*>(s.w2 + size_t(nmax) * L);
    s.va = s.vs + size_t(nmax) * L;
    s.vb = s.va + size_t(nmax) * L;
    s.dprev = reinterpret_cast<int*>(s.vb + size_t(nmax) * L);
    s.nd = s.dprev + size_t(nmax) * L;
    s.par = s.nd + nmax;
    s.rng = s.par + nmax;
    s.path = s.rng + nmax;
    s.xy = s.path + nmax;
    s.lay = s.xy + nmax;
    s.order = s.lay + nmax;
    s.link = s.order + nmax;
    s.woff = s.link + nmax;
    s.lstart = s.woff + nmax + 1;
    s.gbest = base + (table_head_bytes<L>(nmax) + 7) / 8;
    s.gchoice = reinterpret_cast<int*>(s.gbest + lanes);
    return s;
}

__device__ __forceinline__ int idx3(const DeviceGrid& g, int l, int x, int y) { return (l * g.X + x) * g.Y + y; }

// src/cost/prefix_scan.cuh:38
__device__ __forceinline__ double wire_cost(const DeviceGrid& g, const double* presum, int l, int xmin, int xmax,
                                            int ymin, int ymax) {
    return presum[cell_of(g, l, xmax, ymax)] - presum[cell_of(g, l, xmin, ymin)];
}

// Inclusive prefix sum over each group of W consecutive lanes (W a power of
// two up to 64) and the group's total; every lane must be active.  Inside a
// row of 16 lanes the sum doubles its reach through DPP row shifts (a source
// outside the row reads 0, one outside the group is not added); the row
// totals in lanes 15, 31, 47 and 63 carry across the rows of a group.
__device__ __forceinline__ int group_scan(int v, int W, int& total) {
    const int l = threadIdx.x & (kWave - 1), gi = l & (W - 1), row = l >> 4;
    int t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);
    v += gi >= 1 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);
    v += gi >= 2 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);
    v += gi >= 4 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);
    v += gi >= 8 ? t : 0;
    const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31);
    const int r2 = __builtin_amdgcn_readlane(v, 47), r3 = __builtin_amdgcn_readlane(v, 63);
    if (W == 64) {
        v += (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
        total = r0 + r1 + r2 + r3;
    } else if (W == 32) {
        v += row == 1 ? r0 : row == 3 ? r2 : 0;
        total = row < 2 ? r0 + r1 : r2 + r3;
    } else if (W == 16) {
        total = row == 0 ? r0 : row == 1 ? r1 : row == 2 ? r2 : r3;
    } else {
        total = __shfl(v, l | (W - 1), kWave);
    }
    return v;
}

// Row lo of node i's table receives the closures of i's children, in
// descending child index (lshape_warp.cuh:165-189).  The loop runs over
// every high end with selects instead of branches (row[h - lo] for h < lo
// still lies inside node i's table), so the row stays in registers.
template <int L>
__device__ __forceinline__ void gather_children(const Tables<L>& s, int i, int lo) {
    double* row = s.dist + size_t(i) * kTri<L> + tri<L>(lo, lo);  // row[h - lo] = interval (lo, h)
    double acc[L];
#pragma unroll
    for (int hi = 0; hi < L; ++hi) acc[hi] = row[hi - lo];
    for (int k = (s.link[i] & 0xFFFF) - 1; k >= 0; k = (s.link[k] >> 16) - 1) {
        const double* dg = s.diag + k * L;
        double c = dg[lo];
#pragma unroll
        for (int hi = 0; hi < L; ++hi) {
            const double d = dg[hi];
            if (hi > lo && !(c <= d)) c = d;
            const double sum = acc[hi] + c;
            acc[hi] = hi >= lo ? sum : acc[hi];
        }
    }
#pragma unroll
    for (int hi = 0; hi < L; ++hi)
        if (hi >= lo) row[hi - lo] = acc[hi];
}

// Phase A on a node's L lanes (gl = layer c, wl = the lane's index in the
// wavefront).  Lane gl takes row low = gl of the node's interval table and
// walks its high ends downward, keeping the first minimum over high >= low
// + m (a tie going to the lower high); at step m lane c reads that running
// minimum from lane c - m (low = c - m, so the lows ascend as m falls) and
// keeps the first strictly smaller one.  The result is the first minimum
// over the intervals containing c in (low, high) row-major order, the
// order the reference scans them in.
template <int L>
__device__ __forceinline__ void phase_a_lanes(const double* nd, int gl_in, int wl_in, double* gbest, int* gchoice) {
    const int gl = opaque(gl_in), wl = opaque(wl_in);  // lane-derived addresses formed here
    const int lo = gl;
    double sm = kInf, best = kInf;
    int sh = 0, chs = 0;
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
        const int hi = lo + m;
        const bool in = hi < L;  // lanes gl >= L: no row
        const double v = nd[in ? tri<L>(lo, hi) : 0];
        const bool take = in && v < kInf && v <= sm;  // going down, a tie goes to the lower high
        sm = take ? v : sm;
        sh = take ? hi : sh;
        const int src = gl >= m ? wl - m : wl;
        const double t = __shfl(sm, src, kWave);
        const int th = __shfl(sh, src, kWave);
        const bool pick = gl >= m && t < best;  // a tie goes to the lower low
        best = pick ? t : best;
        chs = pick ? (gl - m) * L + th : chs;
        if (m % 2 == 0) asm volatile("" ::: "memory");  // bounds the loads in flight
    }
    if (gl < L) {
        gbest[gl] = best;
        gchoice[gl] = chs;
    }
}

// Phase B (src/routing/lshape_warp.cuh:18-48) on a node's L lanes, lane gl
// taking parent layer pl = gl.  A bent edge's candidates are the child
// layers cl of the opposite parity; the corner via column between cl and pl
// is summed from the lower layer up, starting from 0.0, exactly as the
// reference folds it.  Lane a keeps those sums F(a, a + d) as d grows, over
// both corner columns: the one its own layers' direction selects (lane
// a + d reads it when a is that lane's candidate) and the other (lane a's
// own candidates above it).  The reference keeps the first strictly smaller
// candidate in ascending cl; taking the least (cost, cl) pair is the same
// choice in any visiting order.
template <int L>
__device__ __forceinline__ void phase_b_lanes(const DeviceGrid& g, const Tables<L>& s, int i, int gl_in, int wl_in,
                                              const double* gb, const int* gc) {
    constexpr int LL = L * L;
    const int gl = opaque(gl_in), wl = opaque(wl_in);
    const int pl = gl;
    const int vi = s.xy[i], vp = s.xy[s.par[i]];
    const int x = vi >> 16, y = vi & 0xFFFF, px = vp >> 16, py = vp & 0xFFFF;
    double best = kInf;
    int bp = 0;
    if (x == px || y == py) {
        if (pl < L) {
            const int d = layer_dir(g, pl);
            if (!((d == 0 && y != py) || (d == 1 && x != px))) {
                best = gb[pl] + s.w1[i * L + pl];
                bp = gc[pl] * LL + pl * L + pl;
            }
        }
    } else {
        const int a = pl < L ? pl : L - 1;  // lanes gl >= L compute on layer L - 1, unused
        // corner columns: (x, py) for a vertical first leg, (px, y) otherwise
        const float* own = (layer_dir(g, a) ? s.va : s.vb) + i * L;
        const float* other = (layer_dir(g, a ^ 1) ? s.va : s.vb) + i * L;
        const double second = s.w2[i * L + a];
        int bcl = L;
        auto consider = [&](int cl, double fold) {
            const double cost = gb[cl] + s.w1[i * L + cl];
            const double cost2 = fold + second;
            const double c = cost + cost2;
            if (c < kInf && (c < best || (c == best && cl < bcl))) {
                best = c;
                bcl = cl;
                bp = gc[cl] * LL + cl * L + pl;
            }
        };
        double ua = 0.0, ub = 0.0;  // F(a, a + d) over the own and the other column
#pragma unroll
        for (int d = 1; d < L; ++d) {
            if (d % 2) {
                const int src = gl >= d ? wl - d : wl;
                const double fold = __shfl(ua, src, kWave);  // F(pl - d, pl), lane pl - d's own column
                if (pl < L && gl >= d) consider(pl - d, fold);
                if (pl + d < L) consider(pl + d, ub);
            }
            const int j = a + d < L ? a + d : L - 1;
            ua += (double)own[j];
            ub += (double)other[j];
            asm volatile("" ::: "memory");  // bounds the loads in flight
        }
    }
    if (pl < L) {
        s.diag[i * L + pl] = best;
        s.dprev[i * L + pl] = bp;
    }
}

// The root's first minimum in row-major order on nets of fewer than 16
// lanes: per row the first minimum over high (lanes), then lane 0 takes the
// first minimum over the rows.
template <int L>
__device__ __forceinline__ void root_first_min(const LshapeArgs& a, const Tables<L>& s, int lane, int W, bool live) {
    for (int lo = live ? lane : L; lo < L; lo += W) {
        double m = s.dist[tri<L>(lo, lo)];
        int mh = lo;
#pragma unroll
        for (int hi = 1; hi < L; ++hi) {
            if (hi <= lo) continue;
            const double v = s.dist[tri<L>(lo, hi)];
            if (v < m) {
                m = v;
                mh = hi;
            }
        }
        s.gbest[lo] = m;
        s.gchoice[lo] = mh;
    }
    __syncthreads();
    if (lane == 0 && live) {
        int ch = s.gchoice[0];
        double cv = s.gbest[0];
        for (int lo = 1; lo < L; ++lo)
            if (s.gbest[lo] < cv) {
                cv = s.gbest[lo];
                ch = lo * L + s.gchoice[lo];
            }
        s.rng[0] = ch;
        s.path[0] = 0;
        if (!(cv < kInf)) atomicExch(a.error, (int)kDevNoSolution);
    }
    __syncthreads();
}

// One DPP step of the root's first minimum: lane g takes the value of lane
// g - M of its row (a lower row of the root's table) when it is not larger.
template <int M>
__device__ __forceinline__ void root_merge(double& m, int& ch) {
    const double om = row_shr<M>(m);
    const int oc = row_shr<M>(ch);
    const bool take = om <= m;
    m = take ? om : m;
    ch = take ? oc : ch;
}

// rows (node, layer) per lane and pass of the cost gathers
constexpr int kGatherRows = 2;

// With a queue counter, lane 0 dequeues the workgroup's next chunk before
// the traceback and parks it in *stash once the route words are placed, so
// the atomic's round trip overlaps the last phases of this net's solve.
template <int L>
__device__ void solve_net(const LshapeArgs& a, double* base, size_t words, int first_slot, int* counter = nullptr,
                          int* stash = nullptr) {
    constexpr int LL = L * L, T = kTri<L>;
    const int P = a.per_wave, W = kWave / P;
    const int sub = threadIdx.x / W, lane = threadIdx.x % W;  // lane within the net's lanes
    const int G = max(1, W / L), GL = W / G;                   // lane groups of the net, lanes per group
    const int grp = lane / GL, gl = lane % GL;                 // grp >= G: an idle lane
    const int slot = first_slot + sub;
    const bool live = slot < a.count;
    const Tables<L> s = carve<L>(base + words * sub, a.nmax, net_lanes<L>(P));
    const DeviceGrid& g = a.g;
    const int X = g.X, Y = g.Y;
    const int4 desc = live ? a.desc[a.begin + slot] : make_int4(0, 0, 0, 0);
    // nodes | height of the root (the number of levels below it) << 16
    const int n0 = desc.x, n = desc.y & 0xFFFF, H = desc.y >> 16;
    const int hmax = wave_max(H);
    for (int i = lane; i < n; i += W) {
        const int v = a.nodes[n0 + i];
        const int q = a.lvl[n0 + i], qp = a.lvl[n0 + i - (i > 0)], h = q >> 16;
        s.nd[i] = v;
        s.par[i] = a.parents[n0 + i];
        s.xy[i] = (v / Y % X) << 16 | (v % Y);  // decoded once: no divisions below
        s.lay[i] = v / g.XY;
        s.link[i] = a.link[n0 + i];
        s.order[i] = q & 0xFFFF;
        if (i == 0 || (qp >> 16) != h) s.lstart[h] = i;
    }
    if (live && lane == 0) s.lstart[H + 1] = n;
    __syncthreads();
    // Every cost the DP reads, gathered up front for all nodes so that the
    // bottom-up sweep below touches only LDS: the via stack at each node
    // (src/routing/lshape_warp.cuh:137-145), the edge's wire costs per layer
    // and the via columns at both corners of a bent edge (:18-48).
    // A lane takes kGatherRows rows per pass with no branch between their
    // loads (rows past the end read the root's cells), so all of them are in
    // flight at once; the root row keeps only its via stack.
    constexpr int U = kGatherRows;
    for (int r0 = lane; r0 < n * L; r0 += U * W) {
        double pe0[U], pe1[U], pf0[U], pf1[U];
        float vsv[U], vav[U], vbv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = r0 + u * W < n * L ? r0 + u * W : 0;
            const int i = r / L, c = r % L;
            const int vq = s.xy[i], pq = s.xy[i ? s.par[i] : 0];
            const int x = vq >> 16, y = vq & 0xFFFF, px = pq >> 16, py = pq & 0xFFFF;
            const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
            const bool straight = x == px || y == py, d = layer_dir(g, c);
            // wire costs as prefix-sum differences P[e1] - P[e0] (the first
            // leg, or the whole straight edge) and P[f1] - P[f0] (second leg)
            const int ex0 = straight ? minx : d ? x : minx, ex1 = straight ? maxx : d ? x : maxx;
            const int ey0 = straight ? miny : d ? miny : y, ey1 = straight ? maxy : d ? maxy : y;
            const int fx0 = straight ? ex0 : d ? px : minx, fx1 = straight ? ex1 : d ? px : maxx;
            const int fy0 = straight ? ey0 : d ? miny : py, fy1 = straight ? ey1 : d ? maxy : py;
            pe0[u] = a.presum[cell_of(g, c, ex0, ey0)];
            pe1[u] = a.presum[cell_of(g, c, ex1, ey1)];
            pf0[u] = a.presum[cell_of(g, c, fx0, fy0)];
            pf1[u] = a.presum[cell_of(g, c, fx1, fy1)];
            vsv[u] = a.vcost[cell_of(g, c, x, y)];
            vav[u] = a.vcost[cell_of(g, c, x, straight ? y : py)];  // corner (x, py)
            vbv[u] = a.vcost[cell_of(g, c, straight ? x : px, y)];  // corner (px, y)
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = r0 + u * W;
            if (r >= n * L) continue;
            const int i = r / L, c = r % L;
            s.vs[r] = c < L - 1 ? vsv[u] : 0.f;
            if (!i) continue;
            const int vq = s.xy[i], pq = s.xy[s.par[i]];
            const int x = vq >> 16, y = vq & 0xFFFF, px = pq >> 16, py = pq & 0xFFFF;
            const bool straight = x == px || y == py;
            const int d = layer_dir(g, c);
            const bool along = !((d == 0 && y != py) || (d == 1 && x != px));
            s.w1[r] = straight && !along ? 0.0 : pe1[u] - pe0[u];
            s.w2[r] = straight ? 0.0 : pf1[u] - pf0[u];
            s.va[r] = straight ? 0.f : vav[u];
            s.vb[r] = straight ? 0.f : vbv[u];
        }
    }
    __syncthreads();
    for (int r = lane; r < n * L; r += W) {
        const int i = r / L, lo = r % L;
        const int fixed = s.lay[i];
        const bool pin = fixed < L;  // a pin's intervals must contain its layer
        double* row = s.dist + size_t(i) * T + tri<L>(lo, lo);  // row[h - lo] = interval (lo, h)
        const float* vs = s.vs + i * L;
        double v = 0.0;
        row[0] = pin && lo != fixed ? kInf : 0.0;
        // every high end with selects, so the via loads issue together
#pragma unroll
        for (int hi = 1; hi < L; ++hi) {
            const float f = vs[hi - 1];
            if (hi > lo) {
                v = v + (double)f;
                row[hi - lo] = pin && (lo > fixed || hi < fixed) ? kInf : v;
            }
        }
    }
    __syncthreads();
    for (int lv = 0; lv < hmax; ++lv) {
        const bool lact = live && lv < H;
        const int q0 = lact ? s.lstart[lv] : 0, cnt = lact ? s.lstart[lv + 1] - q0 : 0;
        const int chunks = wave_max((cnt + G - 1) / G);
        for (int ch = 0; ch < chunks; ++ch) {
            const int k = ch * G + grp;
            const bool act = lact && grp < G && k < cnt;
            const int i = act ? s.order[q0 + k] : 0;
            if (lv > 0) {  // level-0 nodes are leaves
                if (act)
                    for (int lo = gl; lo < L; lo += GL) gather_children<L>(s, i, lo);
                __syncthreads();
            }
            // phase A: the best interval containing each layer c, first
            // minimum in (low, high) row-major order
            if (act && GL >= L) {
                phase_a_lanes<L>(s.dist + size_t(i) * T, gl, int(threadIdx.x), s.gbest + grp * L, s.gchoice + grp * L);
            } else if (act) {
                const double* nd = s.dist + size_t(i) * T;
                for (int c = gl; c < L; c += GL) {
                    double best = kInf;
                    int chs = 0;
#pragma unroll
                    for (int lo = 0; lo < L; ++lo) {
#pragma unroll
                        for (int hi = lo; hi < L; ++hi) {
                            const double v = nd[tri<L>(lo, hi)];
                            if (lo <= c && c <= hi && v < best) {
                                best = v;
                                chs = lo * L + hi;
                            }
                        }
                        if (lo % 2 == 1) asm volatile("" ::: "memory");  // bounds the loads in flight
                    }
                    s.gbest[grp * L + c] = best;
                    s.gchoice[grp * L + c] = chs;
                }
            }
            __syncthreads();
            // phase B (src/routing/lshape_warp.cuh:18-48)
            if (act && GL >= L) {
                phase_b_lanes<L>(g, s, i, gl, int(threadIdx.x), s.gbest + grp * L, s.gchoice + grp * L);
            } else if (act) {
                const int vi = s.xy[i], vp = s.xy[s.par[i]];
                const int x = vi >> 16, y = vi & 0xFFFF, px = vp >> 16, py = vp & 0xFFFF;
                const double* gb = s.gbest + grp * L;
                const int* gc = s.gchoice + grp * L;
                for (int pl = gl; pl < L; pl += GL) {
                    double best = kInf;
                    int bp = 0;
                    if (x == px || y == py) {
                        const int d = layer_dir(g, pl);
                        if (!((d == 0 && y != py) || (d == 1 && x != px))) {
                            best = gb[pl] + s.w1[i * L + pl];
                            bp = gc[pl] * LL + pl * L + pl;
                        }
                    } else {
                        const double second = s.w2[i * L + pl];
                        // every candidate layer has the opposite parity, hence
                        // one direction and one corner column: (x, py) for a
                        // vertical first leg, (px, y) otherwise
                        const float* corner = (layer_dir(g, pl ^ 1) ? s.va : s.vb) + i * L;
                        float cf[L];
#pragma unroll
                        for (int j = 0; j < L; ++j) cf[j] = corner[j];
                        double up = 0.0;  // corner[pl+1 .. cl-1], left to right
#pragma unroll
                        for (int cl = 0; cl < L; ++cl) {
                            if (cl >= pl + 2) up += (double)cf[cl - 1];
                            double fold = up;
                            if (cl < pl) {  // corner[cl+1 .. pl-1], left to right
                                double v = 0.0;
#pragma unroll
                                for (int j = cl + 1; j < L; ++j)
                                    if (j < pl) v += (double)cf[j];
                                fold = v;
                            }
                            if (((cl ^ pl) & 1) == 0) continue;
                            const double cost = gb[cl] + s.w1[i * L + cl];
                            const double cost2 = fold + second;
                            if (cost + cost2 < best) {
                                best = cost + cost2;
                                bp = gc[cl] * LL + cl * L + pl;
                            }
                        }
                    }
                    s.diag[i * L + pl] = best;
                    s.dprev[i * L + pl] = bp;
                }
            }
            __syncthreads();
        }
    }
    // root: its children's closures, then its first minimum in row-major
    // order (src/routing/lshape_warp.cuh:193-199): per row the first minimum
    // over high (lanes), then the first minimum over the rows
    if (live && hmax > 0)
        for (int lo = lane; lo < L; lo += W) gather_children<L>(s, 0, lo);
    if (hmax > 0) __syncthreads();
    if (W >= 16) {
        // rows lo < L on lanes lo of the net's first 16-lane row, then the
        // first minimum over the rows through DPP row shifts, a lower row
        // winning ties; other lanes carry +infinity, which every table value
        // undercuts (the first minimum lands in lane 15)
        const int lo = opaque(lane < L ? lane : 0);
        double m = s.dist[tri<L>(lo, lo)];
        int mh = lo;
#pragma unroll
        for (int hi = 1; hi < L; ++hi) {
            const double v = s.dist[tri<L>(lo, hi > lo ? hi : lo)];
            const bool take = hi > lo && v < m;
            m = take ? v : m;
            mh = take ? hi : mh;
        }
        int ch = lo * L + mh;
        if (!live || lane >= L) m = __builtin_huge_val();
        root_merge<1>(m, ch);
        root_merge<2>(m, ch);
        root_merge<4>(m, ch);
        root_merge<8>(m, ch);
        if (live && lane == 15) {
            s.rng[0] = ch;
            s.path[0] = 0;
            if (!(m < kInf)) atomicExch(a.error, (int)kDevNoSolution);
        }
        __syncthreads();
    } else {
        root_first_min<L>(a, s, lane, W, live);
    }
    int next = 0;
    if (counter && threadIdx.x == 0) next = atomicAdd(counter, 1);
    // traceback (src/routing/lshape_warp.cuh:200-213), top level first: a
    // node's interval follows from its parent's
    for (int lv = hmax - 1; lv >= 0; --lv) {
        if (live && lv < H)
            for (int q = s.lstart[lv] + lane; q < s.lstart[lv + 1]; q += W) {
                const int i = s.order[q];
                const int pr = s.rng[s.par[i]], plo = pr / L, phi = pr % L;
                const double* dg = s.diag + i * L;
                const int* dp = s.dprev + i * L;
                double c = dg[plo];
                int pth = dp[plo];
                for (int h = plo + 1; h <= phi; ++h)
                    if (!(c <= dg[h])) {
                        c = dg[h];
                        pth = dp[h];
                    }
                s.path[i] = pth;
                const int r = pth / LL;
                s.rng[i] = r;
                if (!(s.dist[size_t(i) * T + tri<L>(r / L, r % L)] < kInf)) atomicExch(a.error, (int)kDevNoSolution);
            }
        __syncthreads();
    }
    // route words per node, placed by a prefix sum over the net's lanes
    const int nchunks = wave_max(live ? (n + W - 1) / W : 0);
    int carry = 1;
    for (int b = 0; b < nchunks; ++b) {
        const int i = b * W + lane;
        int w = 0;
        if (live && i < n) {
            w = s.rng[i] / L < s.rng[i] % L ? 2 : 0;
            if (i) {
                const int v = s.xy[i], q = s.xy[s.par[i]];
                const bool st = (v >> 16) == (q >> 16) || (v & 0xFFFF) == (q & 0xFFFF);
                w += st ? 2 : 6;
            }
        }
        int total;
        const int incl = group_scan(w, W, total);
        if (live && i < n) s.woff[i] = carry + incl - w;
        carry += total;
    }
    if (live && lane == 0) {
        s.woff[n] = carry;
        if (carry > desc.w) atomicExch(a.error, (int)kDevRouteOverflow);
    }
    if (counter && threadIdx.x == 0) *stash = next;
    __syncthreads();
    const int base_word = desc.z, cap = desc.w;
    if (!live || s.woff[n] > cap) return;
    int* routes = a.rs.routes + base_word;
    // emission (src/routing/lshape_warp.cuh:51-100)
    for (int i = lane; i < n; i += W) {
        int* out = routes + s.woff[i];
        int c = 0;
        const int lo = s.rng[i] / L, hi = s.rng[i] % L;
        const int x = s.xy[i] >> 16, y = s.xy[i] & 0xFFFF;
        if (lo < hi) {
            out[c++] = idx3(g, lo, x, y);
            out[c++] = idx3(g, hi, x, y);
        }
        if (i) {
            const int q = s.xy[s.par[i]];
            const int px = q >> 16, py = q & 0xFFFF;
            const int cl = s.path[i] / L % L, pl = s.path[i] % L;
            if (px == x || py == y) {
                out[c++] = idx3(g, cl, min(x, px), min(y, py));
                out[c++] = idx3(g, cl, max(x, px), max(y, py));
            } else if (layer_dir(g, cl)) {
                out[c++] = idx3(g, cl, x, y);
                out[c++] = idx3(g, cl, x, py);
                out[c++] = idx3(g, cl, x, py);
                out[c++] = idx3(g, pl, x, py);
                out[c++] = idx3(g, pl, x, py);
                out[c++] = idx3(g, pl, px, py);
            } else {
                out[c++] = idx3(g, cl, x, y);
                out[c++] = idx3(g, cl, px, y);
                out[c++] = idx3(g, cl, px, y);
                out[c++] = idx3(g, pl, px, y);
                out[c++] = idx3(g, pl, px, y);
                out[c++] = idx3(g, pl, px, py);
            }
        }
    }
    if (lane == 0) routes[0] = s.woff[n];
}

template <int L>
__global__ __launch_bounds__(kWave) void lshape_dp_lds_kernel(LshapeArgs a) {
    extern __shared__ double smem[];
    solve_net<L>(a, smem, (table_bytes<L>(a.nmax, net_lanes<L>(a.per_wave)) + 7) / 8, blockIdx.x * a.per_wave);
}

template <int L>
__global__ __launch_bounds__(kWave) void lshape_dp_global_kernel(LshapeArgs a) {
    const size_t words = (table_bytes<L>(a.nmax, net_lanes<L>(a.per_wave)) + 7) / 8;
    solve_net<L>(a, a.gscratch + words * a.per_wave * blockIdx.x, words, blockIdx.x * a.per_wave);
}

// Persistent form: a grid no larger than what stays resident drains the
// batch's nets from a work queue, per_wave nets per dequeue, in execution
// order (tallest trees first).  The queue is kDpQueues counters, one
// 128-byte line apart and zero at launch: workgroup w dequeues from counter
// w % kDpQueues, which hands out chunks q, q + kDpQueues, ... in order (one
// counter for all workgroups serializes the dequeues).
template <int L>
__global__ __launch_bounds__(kWave) void lshape_dp_queue_kernel(LshapeArgs a) {
    extern __shared__ double smem[];
    __shared__ int s_next;
    const size_t words = (table_bytes<L>(a.nmax, net_lanes<L>(a.per_wave)) + 7) / 8;
    const int q = blockIdx.x % kDpQueues;
    int* counter = a.queue + q * kDpQueueStride;
    // Lane 0 dequeues; the workgroup is one wavefront, which reads the chunk
    // from lane 0.  Each solve dequeues the next chunk near its end (late
    // enough that chunks still go to the workgroups that free up first).
    int got = 0;
    if (threadIdx.x == 0) got = atomicAdd(counter, 1);
    int first = (__builtin_amdgcn_readfirstlane(got) * kDpQueues + q) * a.per_wave;
    while (first < a.count) {
        solve_net<L>(a, smem, words, first, counter, &s_next);
        __syncthreads();  // the tables and s_next are reused by the next nets
        first = (__builtin_amdgcn_readfirstlane(s_next) * kDpQueues + q) * a.per_wave;
        __syncthreads();  // every lane has read s_next before the next solve writes it
    }
}

template <int L>
void launch_for(const LshapeArgs& a, int blocks, bool lds, hipStream_t st) {
    const size_t lds_bytes = ((table_bytes<L>(a.nmax, net_lanes<L>(a.per_wave)) + 7) / 8) * 8 * a.per_wave;
    if (a.queue) {
        int resident = 0;  // workgroups per compute unit; any value keeps the queue correct
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, lshape_dp_queue_kernel<L>, kWave, lds_bytes) !=
                hipSuccess ||
            resident < 1)
            resident = 1;
        const int grid = std::max(1, std::min(blocks, resident * a.cus));
        hipLaunchKernelGGL(lshape_dp_queue_kernel<L>, dim3(grid), dim3(kWave), lds_bytes, st, a);
    } else if (lds)
        hipLaunchKernelGGL(lshape_dp_lds_kernel<L>, dim3(blocks), dim3(kWave), lds_bytes, st, a);
    else
        hipLaunchKernelGGL(lshape_dp_global_kernel<L>, dim3(blocks), dim3(kWave), 0, st, a);
}

}  // namespace

size_t lshape_scratch_bytes(int L, int nmax) {
    switch (L) {
#define GR_L_CASE(n) \
    case n: return table_bytes<n>(nmax);
        GR_L_CASE(1) GR_L_CASE(2) GR_L_CASE(3) GR_L_CASE(4) GR_L_CASE(5) GR_L_CASE(6) GR_L_CASE(7) GR_L_CASE(8)
        GR_L_CASE(9) GR_L_CASE(10) GR_L_CASE(11) GR_L_CASE(12) GR_L_CASE(13) GR_L_CASE(14) GR_L_CASE(15)
#undef GR_L_CASE
        default: return 0;
    }
}

bool launch_lshape_dp(const LshapeArgs& a, int blocks, bool lds, hipStream_t st) {
    switch (a.g.L) {
#define GR_L_CASE(n)                     \
    case n:                              \
        launch_for<n>(a, blocks, lds, st); \
        return true;
        GR_L_CASE(1) GR_L_CASE(2) GR_L_CASE(3) GR_L_CASE(4) GR_L_CASE(5) GR_L_CASE(6) GR_L_CASE(7) GR_L_CASE(8)
        GR_L_CASE(9) GR_L_CASE(10) GR_L_CASE(11) GR_L_CASE(12) GR_L_CASE(13) GR_L_CASE(14) GR_L_CASE(15)
#undef GR_L_CASE
        default: return false;
    }
}

}  // namespace grotor
