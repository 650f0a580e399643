// L-shape routing-DAG dynamic program for gfx950.
//
// Reference: Lshape_route_warp with node_diagonal / emit_lshape_node
// (src/routing/lshape_warp.cuh:8-247) and its scalar twin Lshape_route_cuda
// (src/routing/lshape_thread.cuh:18-171).
//
// One 64-lane wavefront (one workgroup) owns 1, 2, 4 or 8 nets (nets per
// wavefront, a launch parameter chosen by the autotuner): each net gets a
// group of 64/P lanes and its own tables.  Lanes run across layers instead
// of across tree nodes: nodes are solved one at a time in descending
// preorder index, which is the scalar kernel's order and makes every parent
// accumulate its children's contributions in descending child order without
// floating-point atomics.  Per node the work splits into three lane-parallel
// phases separated by workgroup barriers (the nets of a wavefront step
// through their nodes together; a net with fewer nodes idles):
//   A  lane c: best interval containing layer c (first minimum in
//      (low, high) row-major order); lanes L..3L-1 fold the corner via
//      columns of a bent edge into per-(low, high) partial sums, in the
//      reference's left-to-right order;
//   B  lane c: best arrival on parent layer c (straight edge, or the first
//      minimum over the opposite-parity child layers of a bent edge);
//   C  lane c: interval closure of row low = c and accumulation into the
//      parent's table.
// All costs the sweep reads (prefix-sum differences and via costs) are
// gathered for every node before the sweep starts, so the sequential part
// of the net touches LDS only.  The layer count is a template parameter.
// The interval tables live in LDS (or, for trees too large for LDS, in a
// per-workgroup slice of global scratch); traceback recomputes the closure
// argmin from the stored per-node diagonal, so no per-interval traceback
// table is kept.  Results equal the reference's bit for bit: every double
// is formed by the same operations in the same order.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

// Per-workgroup tables for trees of up to nmax nodes; L is a compile-time
// constant so that every per-layer loop unrolls.
template <int L>
struct Tables {
    double* dist;   // nmax * T      interval tables, upper triangle (low <= high) row by row
    double* diag;   // nmax * L      best arrival on each parent layer
    double* w1;     // nmax * L      first-leg wire (child side; the whole edge when straight)
    double* w2;     // nmax * L      second-leg wire (parent side, bent edges)
    double* best;   // L             best interval containing each layer (current node)
    double* fold;   // 2 * L * L     corner via sums (va, vb) of the current node
    float* vs;      // nmax * L      via costs on the node's column
    float* va;      // nmax * L      via costs at corner (x, py)
    float* vb;      // nmax * L      via costs at corner (px, y)
    int* dprev;     // nmax * L
    int* nd;        // nmax
    int* par;       // nmax
    int* rng;       // nmax
    int* path;      // nmax
    int* choice;    // L
    int* woff;      // nmax + 1
    int* xy;        // nmax          node position x << 16 | y
    int* lay;       // nmax          node layer (L for Steiner nodes)
    double* rmin;   // L * L         row suffix minima of the current node's table
    int* rarg;      // L * L         their high ends
};

// entries of an interval table: intervals low <= high of L layers
template <int L>
constexpr int kTri = L * (L + 1) / 2;

// position of interval (lo, hi), lo <= hi, in a triangular table
template <int L>
__host__ __device__ constexpr int tri(int lo, int hi) {
    return lo * (2 * L - lo + 1) / 2 + (hi - lo);
}

template <int L>
__host__ __device__ constexpr size_t table_bytes(int nmax) {
    return 8 * (size_t(nmax) * (kTri<L> + 3 * L) + L + 3 * L * L) + 4 * (size_t(nmax) * 3 * L) +
           4 * (size_t(nmax) * L + 4 * size_t(nmax) + L + nmax + 1) + 4 * (2 * size_t(nmax) + L * L) + 8;
}

template <int L>
__device__ inline Tables<L> carve(double* base, int nmax) {
    Tables<L> s;
    s.dist = base;
    s.diag = s.dist + size_t(nmax) * kTri<L>;
    s.w1 = s.diag + size_t(nmax) * L;
    s.w2 = s.w1 + size_t(nmax) * L;
    s.best = s.w2 + size_t(nmax) * L;
    s.fold = s.best + L;
    s.vs = reinterpret_cast<float*>(s.fold + 2 * L * L);
    s.va = s.vs + size_t(nmax) * L;
    s.vb = s.va + size_t(nmax) * L;
    s.dprev = reinterpret_cast<int*>(s.vb + size_t(nmax) * L);
    s.nd = s.dprev + size_t(nmax) * L;
    s.par = s.nd + nmax;
    s.rng = s.par + nmax;
    s.path = s.rng + nmax;
    s.choice = s.path + nmax;
    s.woff = s.choice + L;
    s.xy = s.woff + nmax + 1;
    s.lay = s.xy + nmax;
    s.rarg = s.lay + nmax;
    // doubles after the ints, 8-byte aligned
    const size_t ints = size_t(s.rarg + L * L - reinterpret_cast<int*>(base));
    s.rmin = base + (ints + 1) / 2;
    return s;
}

__device__ __forceinline__ int idx3(const DeviceGrid& g, int l, int x, int y) { return (l * g.X + x) * g.Y + y; }

// src/cost/prefix_scan.cuh:38
__device__ __forceinline__ double wire_cost(const DeviceGrid& g, const double* presum, int l, int xmin, int xmax,
                                            int ymin, int ymax) {
    return presum[cell_of(g, l, xmax, ymax)] - presum[cell_of(g, l, xmin, ymin)];
}

template <int L>
__device__ void solve_net(const LshapeArgs& a, double* base, size_t words) {
    constexpr int LL = L * L, T = kTri<L>;
    const int P = a.per_wave, W = kWave / P;
    const int sub = threadIdx.x / W, lane = threadIdx.x % W;  // lane within the net's group
    const int slot = blockIdx.x * P + sub;
    const bool live = slot < a.count;
    const Tables<L> s = carve<L>(base + words * sub, a.nmax);
    const DeviceGrid& g = a.g;
    const int X = g.X, Y = g.Y;
    const int4 desc = live ? a.desc[a.begin + slot] : make_int4(0, 0, 0, 0);
    const int n0 = desc.x, n = desc.y;
    int steps = n;  // nodes of the wavefront's largest net
    for (int off = kWave / 2; off; off >>= 1) steps = max(steps, __shfl_xor(steps, off, kWave));
    for (int i = lane; i < n; i += W) {
        const int v = a.nodes[n0 + i];
        s.nd[i] = v;
        s.par[i] = a.parents[n0 + i];
        s.xy[i] = (v / Y % X) << 16 | (v % Y);  // decoded once: no divisions below
        s.lay[i] = v / g.XY;
    }
    __syncthreads();
    // Every cost the DP reads, gathered up front for all nodes so that the
    // bottom-up sweep below touches only LDS: the via stack at each node
    // (src/routing/lshape_warp.cuh:137-145), the edge's wire costs per layer
    // and the via columns at both corners of a bent edge (:18-48).
    for (int r = lane; r < n * L; r += W) {
        const int i = r / L, c = r % L;
        const int x = s.xy[i] >> 16, y = s.xy[i] & 0xFFFF;
        s.vs[r] = c < L - 1 ? a.vcost[cell_of(g, c, x, y)] : 0.f;
        if (!i) continue;
        const int q = s.xy[s.par[i]];
        const int px = q >> 16, py = q & 0xFFFF;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const int d = layer_dir(g, c);
        double w1 = 0.0, w2 = 0.0;
        float va = 0.f, vb = 0.f;
        if (x == px || y == py) {
            if (!((d == 0 && y != py) || (d == 1 && x != px))) w1 = wire_cost(g, a.presum, c, minx, maxx, miny, maxy);
        } else {
            w1 = d ? wire_cost(g, a.presum, c, x, x, miny, maxy) : wire_cost(g, a.presum, c, minx, maxx, y, y);
            w2 = d ? wire_cost(g, a.presum, c, px, px, miny, maxy) : wire_cost(g, a.presum, c, minx, maxx, py, py);
            va = a.vcost[cell_of(g, c, x, py)];
            vb = a.vcost[cell_of(g, c, px, y)];
        }
        s.w1[r] = w1;
        s.w2[r] = w2;
        s.va[r] = va;
        s.vb[r] = vb;
    }
    __syncthreads();
    for (int r = lane; r < n * L; r += W) {
        const int i = r / L, lo = r % L;
        const int fixed = s.lay[i];
        double* row = s.dist + size_t(i) * T + tri<L>(lo, lo);  // row[h - lo] = interval (lo, h)
        double v = 0.0;
        row[0] = 0.0;
        for (int hi = lo + 1; hi < L; ++hi) {
            v = v + (double)s.vs[i * L + hi - 1];
            row[hi - lo] = v;
        }
        if (fixed < L)
            for (int hi = lo; hi < L; ++hi)
                if (lo > fixed || hi < fixed) row[hi - lo] = kInf;
    }
    __syncthreads();
    for (int t = 1; t < steps; ++t) {
        const int i = n - t;  // this net's node at this step (none once i < 1)
        const bool act = i >= 1;
        const int p = act ? s.par[i] : 0;
        const int vi = act ? s.xy[i] : 0, vp = act ? s.xy[p] : 0;
        const int x = vi >> 16, y = vi & 0xFFFF, px = vp >> 16, py = vp & 0xFFFF;
        const bool straight = x == px || y == py;
        // phase A: the best interval containing each layer c, first minimum
        // in (low, high) row-major order, in two steps.  A1: lane lo < L
        // takes the suffix minima of row lo (the smallest high end among
        // equal values); lanes L..3L-1 fold the corner via columns of a bent
        // edge.  A2: lane c takes the first minimum over lo <= c of row lo's
        // suffix minimum at c.
        const double* nd = s.dist + size_t(act ? i : 0) * T;
        for (int q = act ? lane : 3 * L; q < 3 * L; q += W) {
          if (q < L) {
            const int lo = q;
            double m = nd[tri<L>(lo, L - 1)];
            int mh = L - 1;
            s.rmin[lo * L + L - 1] = m;
            s.rarg[lo * L + L - 1] = mh;
#pragma unroll
            for (int hi = L - 2; hi >= 0; --hi) {
                if (hi < lo) continue;
                const double v = nd[tri<L>(lo, hi)];
                if (v <= m) {
                    m = v;
                    mh = hi;
                }
                s.rmin[lo * L + hi] = m;
                s.rarg[lo * L + hi] = mh;
            }
          } else if (!straight) {
            // fold[k][lo][hi] = sum of corner[lo+1 .. hi-1], left to right
            const int k = (q - L) / L, lo = (q - L) % L;
            const float* corner = (k ? s.vb : s.va) + i * L;
            double* f = s.fold + k * LL + lo * L;
            double v = 0.0;
#pragma unroll
            for (int hi = 0; hi < L; ++hi) {
                if (hi <= lo) continue;
                if (hi > lo + 1) v += (double)corner[hi - 1];
                f[hi] = v;
            }
          }
        }
        __syncthreads();
        for (int c = act ? lane : L; c < L; c += W) {
            double best = kInf;
            int ch = 0;
#pragma unroll
            for (int lo = 0; lo < L; ++lo) {
                if (lo > c) continue;
                const double v = s.rmin[lo * L + c];
                if (v < best) {
                    best = v;
                    ch = lo * L + s.rarg[lo * L + c];
                }
            }
            s.best[c] = best;
            s.choice[c] = ch;
        }
        __syncthreads();
        // phase B (src/routing/lshape_warp.cuh:18-48)
        for (int pl = act ? lane : L; pl < L; pl += W) {
            double best = kInf;
            int bp = 0;
            if (straight) {
                const int d = layer_dir(g, pl);
                if (!((d == 0 && y != py) || (d == 1 && x != px))) {
                    best = s.best[pl] + s.w1[i * L + pl];
                    bp = s.choice[pl] * LL + pl * L + pl;
                }
            } else {
                const double second = s.w2[i * L + pl];
#pragma unroll
                for (int cl = 0; cl < L; ++cl) {
                    if (((cl ^ pl) & 1) == 0) continue;
                    const double cost = s.best[cl] + s.w1[i * L + cl];
                    // vertical first leg: corner at (x, py), the va column
                    const double* f = s.fold + (layer_dir(g, cl) ? 0 : LL);
                    const int lo = min(cl, pl), hi = max(cl, pl);
                    const double cost2 = f[lo * L + hi] + second;
                    if (cost + cost2 < best) {
                        best = cost + cost2;
                        bp = s.choice[cl] * LL + cl * L + pl;
                    }
                }
            }
            s.diag[i * L + pl] = best;
            s.dprev[i * L + pl] = bp;
        }
        __syncthreads();
        // phase C (src/routing/lshape_warp.cuh:165-189)
        for (int lo = act ? lane : L; lo < L; lo += W) {
            const double* dg = s.diag + i * L;
            double* prow = s.dist + size_t(p) * T + tri<L>(lo, lo);
            double c = dg[lo];
#pragma unroll
            for (int hi = 0; hi < L; ++hi) {
                if (hi < lo) continue;
                if (hi > lo && !(c <= dg[hi])) c = dg[hi];
                prow[hi - lo] += c;
            }
        }
        __syncthreads();
    }
    // root choice and traceback (src/routing/lshape_warp.cuh:193-213)
    const int base_word = desc.z, cap = desc.w;
    // the root's first minimum in row-major order: per row its first
    // minimum over high (lanes), then the first minimum over the rows
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
        s.rmin[lo] = m;
        s.rarg[lo] = mh;
    }
    __syncthreads();
    if (lane == 0 && live) {
        int ch = s.rarg[0];
        double cv = s.rmin[0];
        for (int lo = 1; lo < L; ++lo)
            if (s.rmin[lo] < cv) {
                cv = s.rmin[lo];
                ch = lo * L + s.rarg[lo];
            }
        s.rng[0] = ch;
        s.path[0] = 0;
        bool ok = cv < kInf;
        for (int i = 1; i < n; ++i) {
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
            s.rng[i] = pth / LL;
            if (!(s.dist[size_t(i) * T + tri<L>(s.rng[i] / L, s.rng[i] % L)] < kInf)) ok = false;
        }
        int total = 1;
        for (int i = 0; i < n; ++i) {
            s.woff[i] = total;
            int w = s.rng[i] / L < s.rng[i] % L ? 2 : 0;
            if (i) {
                const int v = s.xy[i], q = s.xy[s.par[i]];
                const bool st = (v >> 16) == (q >> 16) || (v & 0xFFFF) == (q & 0xFFFF);
                w += st ? 2 : 6;
            }
            total += w;
        }
        s.woff[n] = total;
        if (!ok) atomicExch(a.error, (int)kDevNoSolution);
        if (total > cap) atomicExch(a.error, (int)kDevRouteOverflow);
    }
    __syncthreads();
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
    solve_net<L>(a, smem, (table_bytes<L>(a.nmax) + 7) / 8);
}

template <int L>
__global__ __launch_bounds__(kWave) void lshape_dp_global_kernel(LshapeArgs a) {
    const size_t words = (table_bytes<L>(a.nmax) + 7) / 8;
    solve_net<L>(a, a.gscratch + words * a.per_wave * blockIdx.x, words);
}

template <int L>
void launch_for(const LshapeArgs& a, int blocks, bool lds, hipStream_t st) {
    if (lds)
        hipLaunchKernelGGL(lshape_dp_lds_kernel<L>, dim3(blocks), dim3(kWave),
                           ((table_bytes<L>(a.nmax) + 7) / 8) * 8 * a.per_wave, st, a);
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
