// Soft (log-sum-exp) L-shape DP and its adjoint for gfx950.
//
// No reference counterpart: a relaxation of the hard L-shape DP
// (src/routing/lshape_warp.cuh:8-247, lshape.hip here) in which every
// minimum becomes smin_tau(x) = -tau * log sum exp(-x / tau) over the finite
// entries.  Per net, with intervals lo <= hi of the L routing layers:
//   V_i[lo][hi]  via stack sum vcost(l), lo <= l < hi, infinite outside the pin layer
//   D_i          V_i + sum over children c of C_c
//   best_c[k]    smin over intervals containing k of D_c
//   diag_c[pl]   straight edge: best_c[pl] + W1_c[pl] on the layers of the edge's direction;
//                bent edge: smin over cl of the other parity of
//                best_c[cl] + W1_c[cl] + corner vias strictly between cl and pl + W2_c[pl]
//   C_c[lo][hi]  smin over lo <= h <= hi of diag_c[h]
//   F            smin over all intervals of D_root
// The cost arrays come from the smooth form of the cost model
// (src/cost/full_update.cuh:9-83 with exact exp), so F is differentiable in
// the overflow slopes and the via weight.  The forward pass checkpoints the
// per-node tables (D, best, diag, leg wire costs); the adjoint pass replays
// the graph from the root down, with d smin / d x_j = exp(-(x_j - smin)/tau),
// and scatters dF/dwcost (expected wire edge use) and dF/dvcost (expected via
// landings) into per-cell FP64 accumulators.
//
// One thread per net: the work per node is O(L^3) scalar exp/log and the
// tables live in a per-net slice of global scratch.  The arithmetic type R
// is float (FP32 tables, exp and log) or double.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

constexpr double kSoftInf = 1e22;
constexpr double kSoftBig = 1e20;

__device__ __forceinline__ double soft_e(double cap, double dem, double s, double z, double& ds, double& dz) {
    ds = dz = 0.0;
    if (cap > 0.001) {
        const double e = exp(s * (dem - cap));
        ds = (dem - cap) * e;
        return e;
    }
    if (dem > 0) {
        const double e = exp(z * dem);
        dz = dem * e;
        return e;
    }
    return 0.0;
}

__device__ __forceinline__ double soft_incr(double cap, double dem, double inc, double sc, double s, double z,
                                            double& ds, double& dz) {
    double s0, z0, s1, z1;
    const double e0 = soft_e(cap, dem, s, z, s0, z0);
    const double e1 = soft_e(cap, dem + inc, s, z, s1, z1);
    const double v = (e1 - e0) * sc;
    if (v >= 1e12) {
        ds = dz = 0.0;
        return 1e12;
    }
    ds = (s1 - s0) * sc;
    dz = (z1 - z0) * sc;
    return v;
}

template <class R>
__device__ __forceinline__ R r_exp(R x);
template <>
__device__ __forceinline__ float r_exp<float>(float x) { return expf(x); }
template <>
__device__ __forceinline__ double r_exp<double>(double x) { return exp(x); }
template <class R>
__device__ __forceinline__ R r_log(R x);
template <>
__device__ __forceinline__ float r_log<float>(float x) { return logf(x); }
template <>
__device__ __forceinline__ double r_log<double>(double x) { return log(x); }

template <class R>
__device__ __forceinline__ R smin(const R* x, int m, R tau) {
    R mn = R(kSoftInf);
    for (int j = 0; j < m; ++j) mn = x[j] < mn ? x[j] : mn;
    if (mn >= R(kSoftBig)) return R(kSoftInf);
    R s = 0;
    for (int j = 0; j < m; ++j)
        if (x[j] < R(kSoftBig)) s += r_exp<R>(-(x[j] - mn) / tau);
    return mn - tau * r_log<R>(s);
}

template <class R>
__device__ __forceinline__ R soft_weight(R x, R y, R tau) {
    if (x >= R(kSoftBig) || y >= R(kSoftBig)) return R(0);
    return r_exp<R>(-(x - y) / tau);
}

__device__ __forceinline__ int tri_index(int L, int lo, int hi) { return lo * (2 * L - lo + 1) / 2 + (hi - lo); }

// leg of layer l through `fixed` over track positions [lo, hi): its wire cost
// from the prefix sums, and the scatter of its gradient
__device__ __forceinline__ double soft_leg(const SoftArgs& a, int l, int x0, int y0, int x1, int y1) {
    return a.presum[cell_of(a.g, l, x1, y1)] - a.presum[cell_of(a.g, l, x0, y0)];
}
__device__ __forceinline__ void soft_leg_grad(const SoftArgs& a, int l, int along_y, int fixed, int lo, int hi,
                                              double gr) {
    for (int p = lo; p < hi; ++p)
        atomicAdd(a.gw + (along_y ? cell_of(a.g, l, fixed, p) : cell_of(a.g, l, p, fixed)), gr);
}

template <class R>
struct NetTables {
    R *D, *best, *diag, *W1, *W2, *gD;
};

template <class R>
__device__ __forceinline__ NetTables<R> net_tables(const SoftArgs& a, int slot) {
    const int L = a.g.L, T = L * (L + 1) / 2;
    const size_t stride = size_t(2 * T + 4 * L);
    R* base = reinterpret_cast<R*>(a.scratch) + size_t(a.node_off[slot] - a.node_off[a.begin]) * stride;
    const int n = a.node_off[slot + 1] - a.node_off[slot];
    NetTables<R> t;
    t.D = base;
    t.gD = t.D + size_t(n) * T;
    t.best = t.gD + size_t(n) * T;
    t.diag = t.best + size_t(n) * L;
    t.W1 = t.diag + size_t(n) * L;
    t.W2 = t.W1 + size_t(n) * L;
    return t;
}

}  // namespace

// Smooth costs per track (one workgroup per track): vcost, the exclusive
// prefix of wcost, and the partial derivatives in slope / zero_slope.
__global__ __launch_bounds__(256) void soft_costs_kernel(SoftCostArgs a) {
    const DeviceGrid& g = a.g;
    const int T = blockIdx.x;
    int l = 0;
#pragma unroll
    for (int i = 1; i < kMaxLayers; ++i)
        if (i < g.L && T >= g.track_begin[i]) l = i;
    const int dir = layer_dir(g, l);
    const int len = dir ? g.Y : g.X;
    const int base = l * g.XY + (T - g.track_begin[l]) * len;
    const int* elen = dir ? g.y_edge_len : g.x_edge_len;
    const double sc = g.short_cost[l];
    __shared__ double part[256];
    const int k = (len + blockDim.x - 1) / blockDim.x;
    const int p0 = threadIdx.x * k, p1 = min(len, p0 + k);
    double local = 0.0;
    for (int p = p0; p < p1; ++p) {
        const int i = base + p, q = p > 0 ? i - 1 : i;
        double ws, wz;
        double w = soft_incr(a.capacity[i], a.demand[i], 1.0, sc, a.slope, a.zero_slope, ws, wz);
        if (p < len - 1) w += g.unit_wire * (double)elen[p];
        double v, vs, vz;
        if (p == 0) {
            v = soft_incr(a.capacity[i], a.demand[i], 1.0, sc, a.slope, a.zero_slope, vs, vz);
        } else if (p == len - 1) {
            v = soft_incr(a.capacity[q], a.demand[q], 1.0, sc, a.slope, a.zero_slope, vs, vz);
        } else {
            double s1, z1, s2, z2;
            v = soft_incr(a.capacity[i], a.demand[i], 0.5, sc, a.slope, a.zero_slope, s1, z1) +
                soft_incr(a.capacity[q], a.demand[q], 0.5, sc, a.slope, a.zero_slope, s2, z2);
            vs = s1 + s2;
            vz = z1 + z2;
        }
        a.vcost[i] = v + a.via_weight * g.unit_via;
        a.dw_ds[i] = ws;
        a.dw_dz[i] = wz;
        a.dv_ds[i] = vs;
        a.dv_dz[i] = vz;
        a.wcost[i] = w;
        local += w;
    }
    part[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const double v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = part[threadIdx.x];
    for (int p = p0; p < p1; ++p) {
        a.presum[base + p] = run;
        run += a.wcost[base + p];
    }
}

template <class R>
__global__ __launch_bounds__(256) void soft_forward_kernel(SoftArgs a) {
    const int slot = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.begin + a.count) return;
    const DeviceGrid& g = a.g;
    const int L = g.L, X = g.X, Y = g.Y, XY = g.XY, T = L * (L + 1) / 2;
    const R tau = R(a.tau);
    const int n0 = a.node_off[slot], n = a.node_off[slot + 1] - n0;
    const int* nodes = a.nodes + n0;
    const int* parents = a.parents + n0;
    const NetTables<R> t = net_tables<R>(a, slot);
    R tmp[64];
    for (int i = 0; i < n; ++i) {
        const int v = nodes[i], fixed = v / XY, x = v / Y % X, y = v % Y;
        for (int lo = 0; lo < L; ++lo) {
            R s = 0;
            for (int hi = lo; hi < L; ++hi) {
                if (hi > lo) s += R(a.vcost[cell_of(g, hi - 1, x, y)]);
                t.D[size_t(i) * T + tri_index(L, lo, hi)] = (fixed < L && (lo > fixed || hi < fixed)) ? R(kSoftInf) : s;
            }
        }
    }
    for (int i = n - 1; i >= 1; --i) {
        const int p = parents[i];
        const int x = nodes[i] / Y % X, y = nodes[i] % Y, px = nodes[p] / Y % X, py = nodes[p] % Y;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const bool straight = x == px || y == py;
        const R* Di = t.D + size_t(i) * T;
        R* best = t.best + size_t(i) * L;
        R* diag = t.diag + size_t(i) * L;
        R* W1 = t.W1 + size_t(i) * L;
        R* W2 = t.W2 + size_t(i) * L;
        for (int k = 0; k < L; ++k) {
            int m = 0;
            for (int lo = 0; lo <= k; ++lo)
                for (int hi = k; hi < L; ++hi) tmp[m++] = Di[tri_index(L, lo, hi)];
            best[k] = smin<R>(tmp, m, tau);
        }
        for (int c = 0; c < L; ++c) {
            const int d = layer_dir(g, c);
            double w1 = 0.0, w2 = 0.0;
            if (straight) {
                if (d ? x == px : y == py)
                    w1 = d ? soft_leg(a, c, x, miny, x, maxy) : soft_leg(a, c, minx, y, maxx, y);
            } else {
                w1 = d ? soft_leg(a, c, x, miny, x, maxy) : soft_leg(a, c, minx, y, maxx, y);
                w2 = d ? soft_leg(a, c, px, miny, px, maxy) : soft_leg(a, c, minx, py, maxx, py);
            }
            W1[c] = R(w1);
            W2[c] = R(w2);
        }
        for (int pl = 0; pl < L; ++pl) {
            const int d = layer_dir(g, pl);
            R val = R(kSoftInf);
            if (straight) {
                if ((d ? x == px : y == py) && best[pl] < R(kSoftBig)) val = best[pl] + W1[pl];
            } else {
                int m = 0;
                for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                    const int lo = min(cl, pl), hi = max(cl, pl);
                    R f = 0;
                    for (int l = lo + 1; l < hi; ++l)
                        f += R(layer_dir(g, cl) ? a.vcost[cell_of(g, l, x, py)] : a.vcost[cell_of(g, l, px, y)]);
                    tmp[m++] = best[cl] >= R(kSoftBig) ? R(kSoftInf) : best[cl] + W1[cl] + f + W2[pl];
                }
                val = smin<R>(tmp, m, tau);
            }
            diag[pl] = val;
        }
        R* Dp = t.D + size_t(p) * T;
        for (int lo = 0; lo < L; ++lo)
            for (int hi = lo; hi < L; ++hi) {
                const R cval = smin<R>(diag + lo, hi - lo + 1, tau);
                R* e = Dp + tri_index(L, lo, hi);
                *e = (*e >= R(kSoftBig) || cval >= R(kSoftBig)) ? R(kSoftInf) : *e + cval;
            }
    }
    a.value[slot - a.begin] = double(smin<R>(t.D, T, tau));
}

template <class R>
__global__ __launch_bounds__(256) void soft_backward_kernel(SoftArgs a) {
    const int slot = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.begin + a.count) return;
    const DeviceGrid& g = a.g;
    const int L = g.L, X = g.X, Y = g.Y, T = L * (L + 1) / 2;
    const R tau = R(a.tau);
    const int n0 = a.node_off[slot], n = a.node_off[slot + 1] - n0;
    const int* nodes = a.nodes + n0;
    const int* parents = a.parents + n0;
    const NetTables<R> t = net_tables<R>(a, slot);
    const R F = R(a.value[slot - a.begin]);
    if (!(F < R(kSoftBig))) return;
    for (int e = 0; e < T; ++e) t.gD[e] = soft_weight<R>(t.D[e], F, tau);
    R gdiag[kMaxLayers], gbest[kMaxLayers];
    for (int c = 1; c < n; ++c) {
        const int p = parents[c];
        const int x = nodes[c] / Y % X, y = nodes[c] % Y, px = nodes[p] / Y % X, py = nodes[p] % Y;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const bool straight = x == px || y == py;
        const R* gC = t.gD + size_t(p) * T;
        const R* dg = t.diag + size_t(c) * L;
        const R* best = t.best + size_t(c) * L;
        const R* W1 = t.W1 + size_t(c) * L;
        const R* W2 = t.W2 + size_t(c) * L;
        for (int h = 0; h < L; ++h) gdiag[h] = gbest[h] = 0;
        for (int lo = 0; lo < L; ++lo)
            for (int hi = lo; hi < L; ++hi) {
                const R gc = gC[tri_index(L, lo, hi)];
                if (gc == R(0)) continue;
                const R cval = smin<R>(dg + lo, hi - lo + 1, tau);
                for (int h = lo; h <= hi; ++h) gdiag[h] += gc * soft_weight<R>(dg[h], cval, tau);
            }
        for (int pl = 0; pl < L; ++pl) {
            if (gdiag[pl] == R(0)) continue;
            const int d = layer_dir(g, pl);
            if (straight) {
                gbest[pl] += gdiag[pl];
                if (d) soft_leg_grad(a, pl, 1, x, miny, maxy, double(gdiag[pl]));
                else soft_leg_grad(a, pl, 0, y, minx, maxx, double(gdiag[pl]));
                continue;
            }
            for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                if (best[cl] >= R(kSoftBig)) continue;
                const int lo = min(cl, pl), hi = max(cl, pl);
                const int dcl = layer_dir(g, cl);
                R f = 0;
                for (int l = lo + 1; l < hi; ++l)
                    f += R(dcl ? a.vcost[cell_of(g, l, x, py)] : a.vcost[cell_of(g, l, px, y)]);
                const R tc = best[cl] + W1[cl] + f + W2[pl];
                const R gr = gdiag[pl] * soft_weight<R>(tc, dg[pl], tau);
                if (gr == R(0)) continue;
                gbest[cl] += gr;
                if (dcl) soft_leg_grad(a, cl, 1, x, miny, maxy, double(gr));
                else soft_leg_grad(a, cl, 0, y, minx, maxx, double(gr));
                if (d) soft_leg_grad(a, pl, 1, px, miny, maxy, double(gr));
                else soft_leg_grad(a, pl, 0, py, minx, maxx, double(gr));
                for (int l = lo + 1; l < hi; ++l)
                    atomicAdd(a.gv + (dcl ? cell_of(g, l, x, py) : cell_of(g, l, px, y)), double(gr));
            }
        }
        const R* Dc = t.D + size_t(c) * T;
        R* gDc = t.gD + size_t(c) * T;
        for (int lo = 0; lo < L; ++lo)
            for (int hi = lo; hi < L; ++hi) {
                R s = 0;
                for (int k = lo; k <= hi; ++k)
                    if (gbest[k] != R(0)) s += gbest[k] * soft_weight<R>(Dc[tri_index(L, lo, hi)], best[k], tau);
                gDc[tri_index(L, lo, hi)] = s;
            }
    }
    // via stacks: D_i[lo][hi] holds vcost(l) for lo <= l < hi
    for (int i = 0; i < n; ++i) {
        const int x = nodes[i] / Y % X, y = nodes[i] % Y;
        for (int l = 0; l + 1 < L; ++l) {
            R s = 0;
            for (int lo = 0; lo <= l; ++lo)
                for (int hi = l + 1; hi < L; ++hi) s += t.gD[size_t(i) * T + tri_index(L, lo, hi)];
            if (s != R(0)) atomicAdd(a.gv + cell_of(g, l, x, y), double(s));
        }
    }
}

// dF/d(slope, zero_slope, via_weight) = sum over cells of the per-cell
// gradients times the cost model's partial derivatives.
__global__ __launch_bounds__(256) void soft_params_kernel(SoftCostArgs c, const double* gw, const double* gv,
                                                          long long cells, double* out) {
    __shared__ double red[3][256];
    double s = 0.0, z = 0.0, u = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < cells; i += (long long)gridDim.x * blockDim.x) {
        s += gw[i] * c.dw_ds[i] + gv[i] * c.dv_ds[i];
        z += gw[i] * c.dw_dz[i] + gv[i] * c.dv_dz[i];
        u += gv[i] * c.g.unit_via;
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = z;
    red[2][threadIdx.x] = u;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        __syncthreads();
    }
    // three sums per 128-byte stripe, stripes summed by the host
    if (threadIdx.x < 3) atomicAdd(out + (blockIdx.x % kStatStripes) * kCounterStride + threadIdx.x, red[threadIdx.x][0]);
}

template __global__ void soft_forward_kernel<float>(SoftArgs);
template __global__ void soft_forward_kernel<double>(SoftArgs);
template __global__ void soft_backward_kernel<float>(SoftArgs);
template __global__ void soft_backward_kernel<double>(SoftArgs);

size_t soft_scratch_bytes(int L, long long nodes, int precision) {
    const int T = L * (L + 1) / 2;
    return size_t(nodes) * size_t(2 * T + 4 * L) * (precision == 64 ? 8 : 4);
}

void launch_soft_forward(const SoftArgs& a, int precision, hipStream_t st) {
    const dim3 grid((a.count + 255) / 256);
    if (precision == 64) hipLaunchKernelGGL(soft_forward_kernel<double>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(soft_forward_kernel<float>, grid, dim3(256), 0, st, a);
}

void launch_soft_backward(const SoftArgs& a, int precision, hipStream_t st) {
    const dim3 grid((a.count + 255) / 256);
    if (precision == 64) hipLaunchKernelGGL(soft_backward_kernel<double>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(soft_backward_kernel<float>, grid, dim3(256), 0, st, a);
}

}  // namespace grotor
