// L-shape routing-DAG dynamic program for gfx950.
//
// Reference: Lshape_route_warp with node_diagonal / emit_lshape_node
// (src/routing/lshape_warp.cuh:8-247) and its scalar twin Lshape_route_cuda
// (src/routing/lshape_thread.cuh:18-171).
//
// One 64-lane wavefront (one workgroup) owns one net.  Lanes run across
// layers instead of across tree nodes: nodes are solved one at a time in
// descending preorder index, which is the scalar kernel's order and makes
// every parent accumulate its children's contributions in descending child
// order without floating-point atomics.  Per node the work splits into three
// lane-parallel phases separated by workgroup barriers:
//   A  lane c: best interval containing layer c (first minimum in
//      (low, high) row-major order); lanes L..4L-1 issue the prefix-sum and
//      corner via gathers for the same node;
//   B  lane c: best arrival on parent layer c (straight edge, or the first
//      minimum over the opposite-parity child layers of a bent edge);
//   C  lane c: interval closure of row low = c and accumulation into the
//      parent's table.
// The interval tables live in LDS (or, for trees too large for LDS, in a
// per-workgroup slice of global scratch); traceback recomputes the closure
// argmin from the stored per-node diagonal, so no per-interval traceback
// table is kept.  Results equal the reference's bit for bit: every double
// is formed by the same operations in the same order.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

struct Scratch {
    double* dist;   // nmax * L * L
    double* diag;   // nmax * L
    double* tmpd;   // 3 * L: best, first-leg wire, second-leg wire
    float* vs;      // nmax * L: via costs on the node's column
    float* va;      // L: via costs at corner (x, py)
    float* vb;      // L: via costs at corner (px, y)
    int* dprev;     // nmax * L
    int* nd;        // nmax
    int* par;       // nmax
    int* rng;       // nmax
    int* path;      // nmax
    int* choice;    // L
    int* woff;      // nmax + 1
};

__host__ __device__ inline size_t scratch_bytes(int L, int nmax) {
    const size_t LL = size_t(L) * L;
    return 8 * (nmax * LL + size_t(nmax) * L + 3 * L) + 4 * (size_t(nmax) * L + 2 * L) +
           4 * (size_t(nmax) * L + 4 * size_t(nmax) + L + nmax + 1);
}

__device__ inline Scratch carve(double* base, int L, int nmax) {
    Scratch s;
    s.dist = base;
    s.diag = s.dist + size_t(nmax) * L * L;
    s.tmpd = s.diag + size_t(nmax) * L;
    s.vs = reinterpret_cast<float*>(s.tmpd + 3 * L);
    s.va = s.vs + size_t(nmax) * L;
    s.vb = s.va + L;
    s.dprev = reinterpret_cast<int*>(s.vb + L);
    s.nd = s.dprev + size_t(nmax) * L;
    s.par = s.nd + nmax;
    s.rng = s.par + nmax;
    s.path = s.rng + nmax;
    s.choice = s.path + nmax;
    s.woff = s.choice + L;
    return s;
}

__device__ __forceinline__ int idx3(const DeviceGrid& g, int l, int x, int y) { return (l * g.X + x) * g.Y + y; }

// src/cost/prefix_scan.cuh:38
__device__ __forceinline__ double wire_cost(const DeviceGrid& g, const double* presum, int l, int xmin, int xmax,
                                            int ymin, int ymax) {
    return presum[cell_of(g, l, xmax, ymax)] - presum[cell_of(g, l, xmin, ymin)];
}

__device__ void solve_net(const LshapeArgs& a, const Scratch& s) {
    const int lane = threadIdx.x;
    const DeviceGrid& g = a.g;
    const int L = g.L, LL = L * L, X = g.X, Y = g.Y;
    const int slot = a.slots ? a.slots[a.begin + blockIdx.x] : a.begin + blockIdx.x;
    const int id = a.net_ids[slot];
    const int n0 = a.node_off[slot], n = a.node_off[slot + 1] - n0;
    for (int i = lane; i < n; i += kWave) {
        s.nd[i] = a.nodes[n0 + i];
        s.par[i] = a.parents[n0 + i];
    }
    __syncthreads();
    // via stacks (src/routing/lshape_warp.cuh:137-145)
    for (int r = lane; r < n * L; r += kWave) {
        const int i = r / L, l = r % L;
        if (l < L - 1) {
            const int v = s.nd[i];
            s.vs[i * L + l] = a.vcost[cell_of(g, l, v / Y % X, v % Y)];
        }
    }
    __syncthreads();
    for (int r = lane; r < n * L; r += kWave) {
        const int i = r / L, lo = r % L;
        const int fixed = s.nd[i] / g.XY;
        double* row = s.dist + size_t(i) * LL + lo * L;
        double v = 0.0;
        row[lo] = 0.0;
        for (int hi = lo + 1; hi < L; ++hi) {
            v = v + (double)s.vs[i * L + hi - 1];
            row[hi] = v;
        }
        if (fixed < L)
            for (int hi = lo; hi < L; ++hi)
                if (lo > fixed || hi < fixed) row[hi] = kInf;
    }
    __syncthreads();
    for (int i = n - 1; i >= 1; --i) {
        const int p = s.par[i];
        const int x = s.nd[i] / Y % X, y = s.nd[i] % Y;
        const int px = s.nd[p] / Y % X, py = s.nd[p] % Y;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const bool straight = x == px || y == py;
        // phase A
        if (lane < L) {
            const int c = lane;
            const double* nd = s.dist + size_t(i) * LL;
            double best = kInf;
            int ch = 0;
            for (int lo = 0; lo <= c; ++lo)
                for (int hi = c; hi < L; ++hi)
                    if (nd[lo * L + hi] < best) {
                        best = nd[lo * L + hi];
                        ch = lo * L + hi;
                    }
            s.tmpd[c] = best;
            s.choice[c] = ch;
        } else if (lane < 2 * L) {
            const int c = lane - L;
            const int d = layer_dir(g, c);
            double w = 0.0;
            if (straight) {
                if (!((d == 0 && y != py) || (d == 1 && x != px))) w = wire_cost(g, a.presum, c, minx, maxx, miny, maxy);
            } else {
                w = d ? wire_cost(g, a.presum, c, x, x, miny, maxy) : wire_cost(g, a.presum, c, minx, maxx, y, y);
            }
            s.tmpd[L + c] = w;
        } else if (lane < 3 * L) {
            const int c = lane - 2 * L;
            if (!straight)
                s.tmpd[2 * L + c] = layer_dir(g, c) ? wire_cost(g, a.presum, c, px, px, miny, maxy)
                                                    : wire_cost(g, a.presum, c, minx, maxx, py, py);
        } else if (lane < 4 * L) {
            const int c = lane - 3 * L;
            if (!straight) {
                s.va[c] = a.vcost[cell_of(g, c, x, py)];
                s.vb[c] = a.vcost[cell_of(g, c, px, y)];
            }
        }
        __syncthreads();
        // phase B (src/routing/lshape_warp.cuh:18-48)
        if (lane < L) {
            const int pl = lane;
            double best = kInf;
            int bp = 0;
            if (straight) {
                const int d = layer_dir(g, pl);
                if (!((d == 0 && y != py) || (d == 1 && x != px))) {
                    best = s.tmpd[pl] + s.tmpd[L + pl];
                    bp = s.choice[pl] * LL + pl * L + pl;
                }
            } else {
                for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                    const double cost = s.tmpd[cl] + s.tmpd[L + cl];
                    const float* corner = layer_dir(g, cl) ? s.va : s.vb;
                    const int lo = min(cl, pl), hi = max(cl, pl);
                    double cost2 = 0.0;
                    for (int l = lo + 1; l < hi; ++l) cost2 += (double)corner[l];
                    cost2 += s.tmpd[2 * L + pl];
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
        if (lane < L) {
            const int lo = lane;
            const double* dg = s.diag + i * L;
            double* prow = s.dist + size_t(p) * LL + lo * L;
            double c = dg[lo];
            for (int hi = lo; hi < L; ++hi) {
                if (hi > lo && !(c <= dg[hi])) c = dg[hi];
                prow[hi] += c;
            }
        }
        __syncthreads();
    }
    // root choice and traceback (src/routing/lshape_warp.cuh:193-213)
    const int base_word = a.rs.pin_acc_num[id] * kRoutePerPin;
    const int cap = kRoutePerPin * (a.rs.pin_acc_num[id + 1] - a.rs.pin_acc_num[id]);
    if (lane == 0) {
        int ch = 0;
        for (int lo = 0; lo < L; ++lo)
            for (int hi = lo; hi < L; ++hi)
                if (s.dist[lo * L + hi] < s.dist[ch]) ch = lo * L + hi;
        s.rng[0] = ch;
        s.path[0] = 0;
        bool ok = s.dist[ch] < kInf;
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
            if (!(s.dist[size_t(i) * LL + s.rng[i]] < kInf)) ok = false;
        }
        int total = 1;
        for (int i = 0; i < n; ++i) {
            s.woff[i] = total;
            int w = s.rng[i] / L < s.rng[i] % L ? 2 : 0;
            if (i) {
                const int v = s.nd[i], q = s.nd[s.par[i]];
                const bool st = (v / Y % X == q / Y % X) || (v % Y == q % Y);
                w += st ? 2 : 6;
            }
            total += w;
        }
        s.woff[n] = total;
        if (!ok) atomicExch(a.error, (int)kDevNoSolution);
        if (total > cap) atomicExch(a.error, (int)kDevRouteOverflow);
    }
    __syncthreads();
    if (s.woff[n] > cap) return;
    int* routes = a.rs.routes + base_word;
    // emission (src/routing/lshape_warp.cuh:51-100)
    for (int i = lane; i < n; i += kWave) {
        int* out = routes + s.woff[i];
        int c = 0;
        const int lo = s.rng[i] / L, hi = s.rng[i] % L;
        const int x = s.nd[i] / Y % X, y = s.nd[i] % Y;
        if (lo < hi) {
            out[c++] = idx3(g, lo, x, y);
            out[c++] = idx3(g, hi, x, y);
        }
        if (i) {
            const int q = s.nd[s.par[i]];
            const int px = q / Y % X, py = q % Y;
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

}  // namespace

size_t lshape_scratch_bytes(int L, int nmax) { return scratch_bytes(L, nmax); }

__global__ __launch_bounds__(kWave) void lshape_dp_lds_kernel(LshapeArgs a) {
    extern __shared__ double smem[];
    solve_net(a, carve(smem, a.g.L, a.nmax));
}

__global__ __launch_bounds__(kWave) void lshape_dp_global_kernel(LshapeArgs a) {
    const size_t words = (scratch_bytes(a.g.L, a.nmax) + 7) / 8;
    solve_net(a, carve(a.gscratch + words * blockIdx.x, a.g.L, a.nmax));
}

}  // namespace grotor
