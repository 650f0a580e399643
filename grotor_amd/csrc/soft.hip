// Soft (log-sum-exp) DP over L-shape or Z-shape edges, and its adjoint,
// for gfx950.
//
// No reference counterpart: a relaxation of the hard pattern-routing DPs in
// which every minimum becomes smin_tau(x) = -tau * log sum exp(-x / tau)
// over the finite entries.  Per net, with intervals lo <= hi of the L
// routing layers:
//   V_i[lo][hi]  via stack sum vcost(l), lo <= l < hi, infinite outside the pin layer
//   D_i          V_i + sum over children c of C_c
//   best_c[k]    smin over intervals containing k of D_c
//   diag_c[e]    smin over the edge's routes leaving the child on layer s
//                and arriving on layer e of best_c[s] + the route's cost
//   C_c[lo][hi]  smin over lo <= h <= hi of diag_c[h]
//   F            smin over all intervals of D_root
// Edge routes of kind 0, the L-shape DP (src/routing/lshape_warp.cuh:8-247,
// lshape.hip here): a straight edge is one wire on a layer of its direction
// (s = e); a bent edge a first leg on s, the corner vias strictly between s
// and e, a second leg on e.  Kind 1, the Z-shape route family
// (src/routing/zshape.cuh:62-288): straight as above (a column when the ends
// share both coordinates); bent: the two L-shapes and the Z-shapes whose
// middle leg runs on a row strictly between the ends' rows or a column
// strictly between their columns, each leg on one layer of its direction
// and a via stack vcost(k), lower <= k < upper layer, at each bend.  Nested
// soft minima compose exactly (smin_x(c_x + smin_y f_xy) = smin_xy(c_x +
// f_xy)), so F is the soft minimum over every route of the family.
//
// The costs are the smooth cost model (src/cost/full_update.cuh:9-83 with
// exact exp; F is differentiable in the overflow slopes and the via weight)
// or arrays the caller gives.  The forward pass checkpoints D, best and diag
// per node; the adjoint replays the tree from the root down, recomputing the
// edges' route costs, with d smin / d x_j = exp(-(x_j - smin) / tau), and
// scatters dF/dwcost (expected wire edge use) and dF/dvcost (expected via
// landings) into FP64 accumulators.
//
// One wavefront per net: lanes over a node's intervals and layers, and for
// kind 1 over an edge's bend positions (64 at a time), whose per-lane soft
// minima (running minimum and scaled sum) merge across the wave.  R, the
// type of the tables and of exp / log, is float or double; costs and
// gradients stay FP64.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

constexpr double kSoftInf = 1e22;
constexpr double kSoftBig = 1e20;
constexpr int kSoftMaxT = kSoftLayers * (kSoftLayers + 1) / 2;

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

// A soft minimum in progress: value mn - tau * log(s), s = sum of
// exp(-(x - mn) / tau) over the entries so far (empty: mn infinite).
template <class R>
struct Lse {
    R mn, s;
};

template <class R>
__device__ __forceinline__ Lse<R> lse_empty() {
    return Lse<R>{R(kSoftInf), R(0)};
}

template <class R>
__device__ __forceinline__ void lse_add(Lse<R>& a, R x, R tau) {
    if (!(x < R(kSoftBig))) return;
    if (x < a.mn) {
        a.s = (a.mn < R(kSoftBig) ? a.s * r_exp<R>(-(a.mn - x) / tau) : R(0)) + R(1);
        a.mn = x;
    } else {
        a.s += r_exp<R>(-(x - a.mn) / tau);
    }
}

template <class R>
__device__ __forceinline__ Lse<R> lse_merge(const Lse<R>& a, const Lse<R>& b, R tau) {
    if (!(b.mn < R(kSoftBig))) return a;
    if (!(a.mn < R(kSoftBig))) return b;
    const R m = a.mn < b.mn ? a.mn : b.mn;
    return Lse<R>{m, a.s * r_exp<R>(-(a.mn - m) / tau) + b.s * r_exp<R>(-(b.mn - m) / tau)};
}

// every lane receives the merge of all lanes' states (the butterfly adds
// the same two terms on both partners, so all lanes agree bit for bit)
template <class R>
__device__ __forceinline__ Lse<R> lse_wave(Lse<R> a, R tau) {
    for (int off = kWave / 2; off; off >>= 1) {
        const Lse<R> b{__shfl_xor(a.mn, off, kWave), __shfl_xor(a.s, off, kWave)};
        a = lse_merge(a, b, tau);
    }
    return a;
}

template <class R>
__device__ __forceinline__ R lse_value(const Lse<R>& a, R tau) {
    return a.mn < R(kSoftBig) ? a.mn - tau * r_log<R>(a.s) : R(kSoftInf);
}

template <class R>
__device__ __forceinline__ R soft_weight(R x, R y, R tau) {
    if (!(x < R(kSoftBig)) || !(y < R(kSoftBig))) return R(0);
    return r_exp<R>(-(x - y) / tau);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = kWave / 2; off; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int tri_index(int L, int lo, int hi) { return lo * (2 * L - lo + 1) / 2 + (hi - lo); }

// Wire cost of layer l along its track between positions p0 and p1 of the
// track through `fixed` (x for along_y, else y), from the prefix sums; the
// gradient variant adds gr to each edge cell [min, max) of the leg.
__device__ __forceinline__ double leg(const SoftArgs& a, int l, int along_y, int fixed, int p0, int p1) {
    const int lo = min(p0, p1), hi = max(p0, p1);
    const int c0 = along_y ? cell_of(a.g, l, fixed, lo) : cell_of(a.g, l, lo, fixed);
    const int c1 = along_y ? cell_of(a.g, l, fixed, hi) : cell_of(a.g, l, hi, fixed);
    return a.presum[c1] - a.presum[c0];
}
__device__ __forceinline__ void leg_grad(const SoftArgs& a, int l, int along_y, int fixed, int p0, int p1, double gr) {
    if (gr == 0.0) return;
    const int lo = min(p0, p1), hi = max(p0, p1);
    for (int p = lo; p < hi; ++p) atomicAdd(a.gw + (along_y ? cell_of(a.g, l, fixed, p) : cell_of(a.g, l, p, fixed)), gr);
}
// via stack at (x, y) between layers l0 and l1: vcost(k), min <= k < max
__device__ __forceinline__ double stack(const SoftArgs& a, int x, int y, int l0, int l1) {
    double s = 0.0;
    for (int k = min(l0, l1); k < max(l0, l1); ++k) s += a.vcost[cell_of(a.g, k, x, y)];
    return s;
}
__device__ __forceinline__ void stack_grad(const SoftArgs& a, int x, int y, int l0, int l1, double gr) {
    if (gr == 0.0) return;
    for (int k = min(l0, l1); k < max(l0, l1); ++k) atomicAdd(a.gv + cell_of(a.g, k, x, y), gr);
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

// Interval table of a workgroup: s_lo[e], s_hi[e] for e = tri_index(L, lo, hi).
__device__ __forceinline__ void interval_table(int L, unsigned char* s_lo, unsigned char* s_hi) {
    const int T = L * (L + 1) / 2;
    for (int e = threadIdx.x; e < T; e += blockDim.x) {
        int lo = 0, rem = e;
        while (rem >= L - lo) rem -= L - lo, ++lo;
        s_lo[e] = (unsigned char)lo;
        s_hi[e] = (unsigned char)(lo + rem);
    }
}

// Geometry of a tree edge from child (x, y) to parent (px, py).
struct Edge {
    int x, y, px, py;
};

// ---- kind 1: the Z-shape route family ----
//
// zedge() visits every route of the edge as (leaving layer s, arriving
// layer e, cost): straight and L-shape routes are split over the lanes by
// arriving layer, the Z-shapes by bend position.  Visit(s, e, cost,
// scatter) is called with the route's cost; in the adjoint pass it returns
// the route's weight and then scatter(w) adds w to the route's cells.  The
// Z-shapes are visited through their nested soft minima: per bend, F1 over
// the leaving layer, M over the middle layer, then the arriving layer.

// Forward: per-lane soft minima per arriving layer, then merged over the wave.
template <class R>
__device__ __forceinline__ void zedge_forward(const SoftArgs& a, const Edge& E, const R* best, R tau, R* out) {
    const DeviceGrid& g = a.g;
    const int L = g.L, lane = threadIdx.x;
    const int x = E.x, y = E.y, px = E.px, py = E.py;
    Lse<R> st[kSoftLayers];
#pragma unroll
    for (int e = 0; e < kSoftLayers; ++e) st[e] = lse_empty<R>();
    if (x == px || y == py) {
        const int e = lane;
        const int dv = x == px;  // a column (also when both coordinates agree)
        if (e < L && layer_dir(g, e) == dv) {
            const double c = dv ? leg(a, e, 1, x, y, py) : leg(a, e, 0, y, x, px);
            if (best[e] < R(kSoftBig)) lse_add(st[e], best[e] + R(c), tau);
        }
    } else {
        // L-shapes, lane e = arriving layer
        const int e = lane;
        if (e < L) {
            if (layer_dir(g, e) == 0) {  // vertical, corner (x, py), horizontal on e
                const double last = leg(a, e, 0, py, x, px);
                for (int s = 0; s < L; ++s)
                    if (layer_dir(g, s) == 1 && best[s] < R(kSoftBig))
                        lse_add(st[e], best[s] + R(leg(a, s, 1, x, y, py) + stack(a, x, py, s, e) + last), tau);
            } else {  // horizontal, corner (px, y), vertical on e
                const double last = leg(a, e, 1, px, y, py);
                for (int s = 0; s < L; ++s)
                    if (layer_dir(g, s) == 0 && best[s] < R(kSoftBig))
                        lse_add(st[e], best[s] + R(leg(a, s, 0, y, x, px) + stack(a, px, y, s, e) + last), tau);
            }
        }
        // Z-shapes: rows strictly between y and py (vertical, horizontal, vertical),
        // then columns strictly between x and px (horizontal, vertical, horizontal)
        for (int pass = 0; pass < 2; ++pass) {
            const int lo = pass == 0 ? min(y, py) : min(x, px), hi = pass == 0 ? max(y, py) : max(x, px);
            const int d1 = pass == 0 ? 1 : 0;  // direction of the first and last legs
            for (int b0 = lo + 1; b0 < hi; b0 += kWave) {
                const int b = b0 + lane;
                if (b >= hi) continue;
                // bend cells: (x, b) and (px, b) for rows, (b, y) and (b, py) for columns
                const int ax = pass == 0 ? x : b, ay = pass == 0 ? b : y;
                const int bx = pass == 0 ? px : b, by = pass == 0 ? b : py;
                R f1[kSoftLayers];
                for (int m = 0; m < L; ++m) {
                    Lse<R> s1 = lse_empty<R>();
                    if (layer_dir(g, m) != d1)
                        for (int s = 0; s < L; ++s)
                            if (layer_dir(g, s) == d1 && best[s] < R(kSoftBig)) {
                                const double c = (pass == 0 ? leg(a, s, 1, x, y, b) : leg(a, s, 0, y, x, b)) +
                                                 stack(a, ax, ay, s, m);
                                lse_add(s1, best[s] + R(c), tau);
                            }
                    f1[m] = lse_value(s1, tau);
                }
                for (int e = 0; e < L; ++e) {
                    if (layer_dir(g, e) != d1) continue;
                    Lse<R> s2 = lse_empty<R>();
                    for (int m = 0; m < L; ++m)
                        if (layer_dir(g, m) != d1 && f1[m] < R(kSoftBig)) {
                            const double c = (pass == 0 ? leg(a, m, 0, b, x, px) : leg(a, m, 1, b, y, py)) +
                                             stack(a, bx, by, m, e);
                            lse_add(s2, f1[m] + R(c), tau);
                        }
                    const R mv = lse_value(s2, tau);
                    if (mv < R(kSoftBig)) {
                        const double c = pass == 0 ? leg(a, e, 1, px, b, py) : leg(a, e, 0, py, b, px);
                        lse_add(st[e], mv + R(c), tau);
                    }
                }
            }
        }
    }
    for (int e = 0; e < L; ++e) {
        const Lse<R> m = lse_wave(st[e], tau);
        if (lane == e) out[e] = lse_value(m, tau);
    }
}

// Adjoint of zedge_forward: gdiag[e] = dF/d diag[e]; adds each route's
// weight to gb[s] (this lane's share of dF/d best[s]) and its cells.
template <class R>
__device__ __forceinline__ void zedge_backward(const SoftArgs& a, const Edge& E, const R* best, const R* diag,
                                               const R* gdiag, R tau, double* gb) {
    const DeviceGrid& g = a.g;
    const int L = g.L, lane = threadIdx.x;
    const int x = E.x, y = E.y, px = E.px, py = E.py;
    if (x == px || y == py) {
        const int e = lane;
        const int dv = x == px;
        if (e < L && layer_dir(g, e) == dv && gdiag[e] != R(0) && best[e] < R(kSoftBig)) {
            const double c = dv ? leg(a, e, 1, x, y, py) : leg(a, e, 0, y, x, px);
            const double w = double(gdiag[e] * soft_weight<R>(best[e] + R(c), diag[e], tau));
            gb[e] += w;
            if (dv) leg_grad(a, e, 1, x, y, py, w);
            else leg_grad(a, e, 0, y, x, px, w);
        }
        return;
    }
    {
        const int e = lane;
        if (e < L && gdiag[e] != R(0)) {
            const int h = layer_dir(g, e) == 0;  // arriving horizontal: first leg vertical through (x, py)
            const double last = h ? leg(a, e, 0, py, x, px) : leg(a, e, 1, px, y, py);
            double wl = 0.0;
            for (int s = 0; s < L; ++s) {
                if (layer_dir(g, s) != (h ? 1 : 0) || !(best[s] < R(kSoftBig))) continue;
                const double first = h ? leg(a, s, 1, x, y, py) : leg(a, s, 0, y, x, px);
                const double c = first + (h ? stack(a, x, py, s, e) : stack(a, px, y, s, e)) + last;
                const double w = double(gdiag[e] * soft_weight<R>(best[s] + R(c), diag[e], tau));
                if (w == 0.0) continue;
                gb[s] += w;
                wl += w;
                if (h) {
                    leg_grad(a, s, 1, x, y, py, w);
                    stack_grad(a, x, py, s, e, w);
                } else {
                    leg_grad(a, s, 0, y, x, px, w);
                    stack_grad(a, px, y, s, e, w);
                }
            }
            if (h) leg_grad(a, e, 0, py, x, px, wl);
            else leg_grad(a, e, 1, px, y, py, wl);
        }
    }
    for (int pass = 0; pass < 2; ++pass) {
        const int lo = pass == 0 ? min(y, py) : min(x, px), hi = pass == 0 ? max(y, py) : max(x, px);
        const int d1 = pass == 0 ? 1 : 0;
        for (int b0 = lo + 1; b0 < hi; b0 += kWave) {
            const int b = b0 + lane;
            if (b >= hi) continue;
            const int ax = pass == 0 ? x : b, ay = pass == 0 ? b : y;
            const int bx = pass == 0 ? px : b, by = pass == 0 ? b : py;
            R f1[kSoftLayers], mv[kSoftLayers];
            double gf1[kSoftLayers], gm[kSoftLayers];
            for (int m = 0; m < L; ++m) {
                Lse<R> s1 = lse_empty<R>();
                if (layer_dir(g, m) != d1)
                    for (int s = 0; s < L; ++s)
                        if (layer_dir(g, s) == d1 && best[s] < R(kSoftBig)) {
                            const double c = (pass == 0 ? leg(a, s, 1, x, y, b) : leg(a, s, 0, y, x, b)) +
                                             stack(a, ax, ay, s, m);
                            lse_add(s1, best[s] + R(c), tau);
                        }
                f1[m] = lse_value(s1, tau);
                gf1[m] = 0.0;
            }
            for (int e = 0; e < L; ++e) {
                gm[e] = 0.0;
                mv[e] = R(kSoftInf);
                if (layer_dir(g, e) != d1) continue;
                Lse<R> s2 = lse_empty<R>();
                for (int m = 0; m < L; ++m)
                    if (layer_dir(g, m) != d1 && f1[m] < R(kSoftBig)) {
                        const double c = (pass == 0 ? leg(a, m, 0, b, x, px) : leg(a, m, 1, b, y, py)) +
                                         stack(a, bx, by, m, e);
                        lse_add(s2, f1[m] + R(c), tau);
                    }
                mv[e] = lse_value(s2, tau);
                if (mv[e] < R(kSoftBig) && gdiag[e] != R(0)) {
                    const double c = pass == 0 ? leg(a, e, 1, px, b, py) : leg(a, e, 0, py, b, px);
                    gm[e] = double(gdiag[e] * soft_weight<R>(mv[e] + R(c), diag[e], tau));
                    if (pass == 0) leg_grad(a, e, 1, px, b, py, gm[e]);
                    else leg_grad(a, e, 0, py, b, px, gm[e]);
                }
            }
            // middle legs and the second bend's stacks
            for (int m = 0; m < L; ++m) {
                if (layer_dir(g, m) == d1 || !(f1[m] < R(kSoftBig))) continue;
                const double mid = pass == 0 ? leg(a, m, 0, b, x, px) : leg(a, m, 1, b, y, py);
                for (int e = 0; e < L; ++e) {
                    if (layer_dir(g, e) != d1 || gm[e] == 0.0) continue;
                    const double c = mid + stack(a, bx, by, m, e);
                    const double w = gm[e] * double(soft_weight<R>(f1[m] + R(c), mv[e], tau));
                    gf1[m] += w;
                    stack_grad(a, bx, by, m, e, w);
                }
                if (pass == 0) leg_grad(a, m, 0, b, x, px, gf1[m]);
                else leg_grad(a, m, 1, b, y, py, gf1[m]);
            }
            // first legs and the first bend's stacks
            for (int s = 0; s < L; ++s) {
                if (layer_dir(g, s) != d1 || !(best[s] < R(kSoftBig))) continue;
                const double first = pass == 0 ? leg(a, s, 1, x, y, b) : leg(a, s, 0, y, x, b);
                double ws = 0.0;
                for (int m = 0; m < L; ++m) {
                    if (layer_dir(g, m) == d1 || gf1[m] == 0.0) continue;
                    const double c = first + stack(a, ax, ay, s, m);
                    const double w = gf1[m] * double(soft_weight<R>(best[s] + R(c), f1[m], tau));
                    ws += w;
                    stack_grad(a, ax, ay, s, m, w);
                }
                gb[s] += ws;
                if (pass == 0) leg_grad(a, s, 1, x, y, b, ws);
                else leg_grad(a, s, 0, y, x, b, ws);
            }
        }
    }
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


// One wavefront per net (grid-stride over the slots [begin, begin + count)).
template <class R, int K>
__global__ __launch_bounds__(kWave) void soft_forward_kernel(SoftArgs a) {
    __shared__ unsigned char s_lo[kSoftMaxT], s_hi[kSoftMaxT];
    __shared__ R s_best[kSoftLayers], s_diag[kSoftLayers];
    __shared__ double s_w1[kSoftLayers], s_w2[kSoftLayers];
    const int lane = threadIdx.x;
    const DeviceGrid& g = a.g;
    const int L = g.L, X = g.X, Y = g.Y, XY = g.XY, T = L * (L + 1) / 2;
    const R tau = R(a.tau);
    interval_table(L, s_lo, s_hi);
    __syncthreads();
    for (int slot = a.begin + blockIdx.x; slot < a.begin + a.count; slot += gridDim.x) {
        const int n0 = a.node_off[slot], n = a.node_off[slot + 1] - n0;
        const int* nodes = a.nodes + n0;
        const int* parents = a.parents + n0;
        const NetTables<R> t = net_tables<R>(a, slot);
        for (int q = lane; q < n * T; q += kWave) {
            const int i = q / T, e = q % T;
            const int v = nodes[i], fixed = v / XY, x = v / Y % X, y = v % Y;
            const int lo = s_lo[e], hi = s_hi[e];
            R val = R(kSoftInf);
            if (!(fixed < L && (lo > fixed || hi < fixed))) {
                R s = 0;
                for (int l = lo; l < hi; ++l) s += R(a.vcost[cell_of(g, l, x, y)]);
                val = s;
            }
            t.D[size_t(i) * T + e] = val;
        }
        __syncthreads();
        for (int i = n - 1; i >= 1; --i) {
            const int p = parents[i];
            const Edge E{nodes[i] / Y % X, nodes[i] % Y, nodes[p] / Y % X, nodes[p] % Y};
            const R* Di = t.D + size_t(i) * T;
            if (lane < L) {
                Lse<R> st = lse_empty<R>();
                for (int lo = 0; lo <= lane; ++lo)
                    for (int hi = lane; hi < L; ++hi) lse_add(st, Di[tri_index(L, lo, hi)], tau);
                const R b = lse_value(st, tau);
                s_best[lane] = b;
                t.best[size_t(i) * L + lane] = b;
            }
            __syncthreads();
            if (K == 0) {
                const int x = E.x, y = E.y, px = E.px, py = E.py;
                const bool straight = x == px || y == py;
                if (lane < L) {
                    const int c = lane, d = layer_dir(g, c);
                    double w1 = 0.0, w2 = 0.0;
                    if (!straight || (d ? x == px : y == py)) w1 = d ? leg(a, c, 1, x, y, py) : leg(a, c, 0, y, x, px);
                    if (!straight) w2 = d ? leg(a, c, 1, px, y, py) : leg(a, c, 0, py, x, px);
                    s_w1[c] = w1;
                    s_w2[c] = w2;
                    t.W1[size_t(i) * L + c] = R(w1);
                    t.W2[size_t(i) * L + c] = R(w2);
                }
                __syncthreads();
                if (lane < L) {
                    const int pl = lane, d = layer_dir(g, pl);
                    Lse<R> st = lse_empty<R>();
                    if (straight) {
                        if ((d ? x == px : y == py) && s_best[pl] < R(kSoftBig))
                            lse_add(st, s_best[pl] + R(s_w1[pl]), tau);
                    } else {
                        for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                            if (!(s_best[cl] < R(kSoftBig))) continue;
                            const int lo = min(cl, pl), hi = max(cl, pl);
                            double f = 0.0;
                            for (int l = lo + 1; l < hi; ++l)
                                f += layer_dir(g, cl) ? a.vcost[cell_of(g, l, x, py)] : a.vcost[cell_of(g, l, px, y)];
                            lse_add(st, s_best[cl] + R(s_w1[cl] + f + s_w2[pl]), tau);
                        }
                    }
                    s_diag[pl] = lse_value(st, tau);
                }
            } else {
                zedge_forward<R>(a, E, s_best, tau, s_diag);
            }
            __syncthreads();
            if (lane < L) t.diag[size_t(i) * L + lane] = s_diag[lane];
            R* Dp = t.D + size_t(p) * T;
            for (int e = lane; e < T; e += kWave) {
                Lse<R> st = lse_empty<R>();
                for (int h = s_lo[e]; h <= s_hi[e]; ++h) lse_add(st, s_diag[h], tau);
                const R cval = lse_value(st, tau);
                const R d = Dp[e];
                Dp[e] = (!(d < R(kSoftBig)) || !(cval < R(kSoftBig))) ? R(kSoftInf) : d + cval;
            }
            __syncthreads();
        }
        Lse<R> st = lse_empty<R>();
        for (int e = lane; e < T; e += kWave) lse_add(st, t.D[e], tau);
        st = lse_wave(st, tau);
        if (lane == 0) a.value[slot - a.begin] = double(lse_value(st, tau));
        __syncthreads();
    }
}

template <class R, int K>
__global__ __launch_bounds__(kWave) void soft_backward_kernel(SoftArgs a) {
    __shared__ unsigned char s_lo[kSoftMaxT], s_hi[kSoftMaxT];
    __shared__ R s_C[kSoftMaxT], s_gC[kSoftMaxT];
    __shared__ R s_best[kSoftLayers], s_diag[kSoftLayers], s_gdiag[kSoftLayers], s_gbest[kSoftLayers];
    const int lane = threadIdx.x;
    const DeviceGrid& g = a.g;
    const int L = g.L, X = g.X, Y = g.Y, T = L * (L + 1) / 2;
    const R tau = R(a.tau);
    interval_table(L, s_lo, s_hi);
    __syncthreads();
    for (int slot = a.begin + blockIdx.x; slot < a.begin + a.count; slot += gridDim.x) {
        const int n0 = a.node_off[slot], n = a.node_off[slot + 1] - n0;
        const int* nodes = a.nodes + n0;
        const int* parents = a.parents + n0;
        const NetTables<R> t = net_tables<R>(a, slot);
        const R F = R(a.value[slot - a.begin]);
        if (!(F < R(kSoftBig))) continue;  // uniform over the wave
        for (int e = lane; e < T; e += kWave) t.gD[e] = soft_weight<R>(t.D[e], F, tau);
        __syncthreads();
        for (int c = 1; c < n; ++c) {
            const int p = parents[c];
            const Edge E{nodes[c] / Y % X, nodes[c] % Y, nodes[p] / Y % X, nodes[p] % Y};
            const R* dg = t.diag + size_t(c) * L;
            const R* gC = t.gD + size_t(p) * T;
            for (int e = lane; e < T; e += kWave) {
                Lse<R> st = lse_empty<R>();
                for (int h = s_lo[e]; h <= s_hi[e]; ++h) lse_add(st, dg[h], tau);
                s_C[e] = lse_value(st, tau);
                s_gC[e] = gC[e];
            }
            if (lane < L) {
                s_best[lane] = t.best[size_t(c) * L + lane];
                s_diag[lane] = dg[lane];
            }
            __syncthreads();
            if (lane < L) {
                const int h = lane;
                R s = 0;
                for (int lo = 0; lo <= h; ++lo)
                    for (int hi = h; hi < L; ++hi) {
                        const int e = tri_index(L, lo, hi);
                        if (s_gC[e] != R(0)) s += s_gC[e] * soft_weight<R>(s_diag[h], s_C[e], tau);
                    }
                s_gdiag[h] = s;
            }
            __syncthreads();
            double gb[kSoftLayers];
#pragma unroll
            for (int k = 0; k < kSoftLayers; ++k) gb[k] = 0.0;
            if (K == 0) {
                const int x = E.x, y = E.y, px = E.px, py = E.py;
                const bool straight = x == px || y == py;
                if (lane < L && s_gdiag[lane] != R(0)) {
                    const int pl = lane, d = layer_dir(g, pl);
                    if (straight) {
                        if (d ? x == px : y == py) {
                            gb[pl] += double(s_gdiag[pl]);
                            if (d) leg_grad(a, pl, 1, x, y, py, double(s_gdiag[pl]));
                            else leg_grad(a, pl, 0, y, x, px, double(s_gdiag[pl]));
                        }
                    } else {
                        // the forward's candidate values, recomputed the same way
                        const double wpl = d ? leg(a, pl, 1, px, y, py) : leg(a, pl, 0, py, x, px);
                        double w2 = 0.0;
                        for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                            if (!(s_best[cl] < R(kSoftBig))) continue;
                            const int lo = min(cl, pl), hi = max(cl, pl);
                            const int dcl = layer_dir(g, cl);
                            double f = 0.0;
                            for (int l = lo + 1; l < hi; ++l)
                                f += dcl ? a.vcost[cell_of(g, l, x, py)] : a.vcost[cell_of(g, l, px, y)];
                            const double wcl = dcl ? leg(a, cl, 1, x, y, py) : leg(a, cl, 0, y, x, px);
                            const R tc = s_best[cl] + R(wcl + f + wpl);
                            const double gr = double(s_gdiag[pl] * soft_weight<R>(tc, s_diag[pl], tau));
                            if (gr == 0.0) continue;
                            gb[cl] += gr;
                            w2 += gr;
                            if (dcl) leg_grad(a, cl, 1, x, y, py, gr);
                            else leg_grad(a, cl, 0, y, x, px, gr);
                            for (int l = lo + 1; l < hi; ++l)
                                atomicAdd(a.gv + (dcl ? cell_of(g, l, x, py) : cell_of(g, l, px, y)), gr);
                        }
                        if (d) leg_grad(a, pl, 1, px, y, py, w2);
                        else leg_grad(a, pl, 0, py, x, px, w2);
                    }
                }
            } else {
                zedge_backward<R>(a, E, s_best, s_diag, s_gdiag, tau, gb);
            }
            for (int k = 0; k < L; ++k) {
                const double v = wave_sum(gb[k]);
                if (lane == k) s_gbest[k] = R(v);
            }
            __syncthreads();
            const R* Dc = t.D + size_t(c) * T;
            R* gDc = t.gD + size_t(c) * T;
            for (int e = lane; e < T; e += kWave) {
                R s = 0;
                const R d = Dc[e];
                for (int k = s_lo[e]; k <= s_hi[e]; ++k)
                    if (s_gbest[k] != R(0)) s += s_gbest[k] * soft_weight<R>(d, s_best[k], tau);
                gDc[e] = s;
            }
            __syncthreads();
        }
        // via stacks: D_i[lo][hi] holds vcost(l) for lo <= l < hi
        for (int q = lane; q < n * (L - 1); q += kWave) {
            const int i = q / (L - 1), l = q % (L - 1);
            const int x = nodes[i] / Y % X, y = nodes[i] % Y;
            R s = 0;
            for (int lo = 0; lo <= l; ++lo)
                for (int hi = l + 1; hi < L; ++hi) s += t.gD[size_t(i) * T + tri_index(L, lo, hi)];
            if (s != R(0)) atomicAdd(a.gv + cell_of(g, l, x, y), double(s));
        }
        __syncthreads();
    }
}

// Exclusive prefix of given wire costs along each track (one workgroup per
// track), for soft passes over caller-supplied costs.
__global__ __launch_bounds__(256) void soft_prefix_kernel(DeviceGrid g, const double* wcost, double* presum) {
    const int T = blockIdx.x;
    int l = 0;
#pragma unroll
    for (int i = 1; i < kMaxLayers; ++i)
        if (i < g.L && T >= g.track_begin[i]) l = i;
    const int len = layer_dir(g, l) ? g.Y : g.X;
    const int base = l * g.XY + (T - g.track_begin[l]) * len;
    __shared__ double part[256];
    const int k = (len + blockDim.x - 1) / blockDim.x;
    const int p0 = threadIdx.x * k, p1 = min(len, p0 + k);
    double local = 0.0;
    for (int p = p0; p < p1; ++p) local += wcost[base + p];
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
        presum[base + p] = run;
        run += wcost[base + p];
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


size_t soft_scratch_bytes(int L, long long nodes, int precision) {
    const int T = L * (L + 1) / 2;
    return size_t(nodes) * size_t(2 * T + 4 * L) * (precision == 64 ? 8 : 4);
}

namespace {
int soft_grid(const SoftArgs& a, int cus) { return std::max(1, std::min(a.count, std::max(1, cus) * 32)); }
}  // namespace

void launch_soft_forward(const SoftArgs& a, int precision, int kind, int cus, hipStream_t st) {
    const dim3 grid(soft_grid(a, cus));
    if (precision == 64) {
        if (kind) hipLaunchKernelGGL((soft_forward_kernel<double, 1>), grid, dim3(kWave), 0, st, a);
        else hipLaunchKernelGGL((soft_forward_kernel<double, 0>), grid, dim3(kWave), 0, st, a);
    } else {
        if (kind) hipLaunchKernelGGL((soft_forward_kernel<float, 1>), grid, dim3(kWave), 0, st, a);
        else hipLaunchKernelGGL((soft_forward_kernel<float, 0>), grid, dim3(kWave), 0, st, a);
    }
}

void launch_soft_backward(const SoftArgs& a, int precision, int kind, int cus, hipStream_t st) {
    const dim3 grid(soft_grid(a, cus));
    if (precision == 64) {
        if (kind) hipLaunchKernelGGL((soft_backward_kernel<double, 1>), grid, dim3(kWave), 0, st, a);
        else hipLaunchKernelGGL((soft_backward_kernel<double, 0>), grid, dim3(kWave), 0, st, a);
    } else {
        if (kind) hipLaunchKernelGGL((soft_backward_kernel<float, 1>), grid, dim3(kWave), 0, st, a);
        else hipLaunchKernelGGL((soft_backward_kernel<float, 0>), grid, dim3(kWave), 0, st, a);
    }
}

}  // namespace grotor
