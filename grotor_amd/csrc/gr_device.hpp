// Device-side definitions shared by the gfx950 kernels.
//
// HBM layout ("track-major"): routing layer l occupies cells
// [l * X * Y, (l + 1) * X * Y).  Inside a layer every routing track is one
// contiguous run in its preferred direction:
//   vertical layer   (dir 1): track = x, position = y, cell = l*XY + x*Y + y
//   horizontal layer (dir 0): track = y, position = x, cell = l*XY + y*X + x
// so every per-track scan, commit span and prefix read is unit stride.  Routes
// and the C-ABI keep the reference encoding IDX(l, x, y) = (l*X + x)*Y + y
// (include/router_state.hpp:267).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gr_exp.hpp"

namespace grotor {

constexpr double kInf = 1e22;  // include/router_state.hpp:271
constexpr int kRoutePerPin = 12;  // include/router_state.hpp:31
constexpr int kMaxLayers = 16;

struct DeviceGrid {
    int L, X, Y, DIR, XY;
    double unit_wire, unit_via;
    double short_cost[kMaxLayers];
    const int* x_edge_len;
    const int* y_edge_len;
    int track_begin[kMaxLayers + 1];  // first global track index of each layer
};

__host__ __device__ inline int layer_dir(const DeviceGrid& g, int l) { return (l & 1) ^ g.DIR; }

// v through an opaque copy: values derived from it are formed where it is
// used, not hoisted out of an enclosing loop where they would hold
// registers for the whole kernel.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Lane g of each 16-lane row receives x of lane g - M of its row (DPP
// row_shr); lanes with g < M keep their own x.
template <int M>
__device__ __forceinline__ int row_shr(int x) {
    if constexpr (M == 0) return x;
    else return __builtin_amdgcn_update_dpp(x, x, 0x110 + M, 0xF, 0xF, false);
}
template <int M>
__device__ __forceinline__ double row_shr(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = row_shr<M>((int)(b & 0xFFFFFFFFLL)), hi = row_shr<M>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Maximum over the wavefront's 64 lanes, the same in every lane; every lane
// must be active.  Each row of 16 lanes reduces through DPP row shifts by 1,
// 2, 4 and 8 (a lane whose source lies outside its row keeps its own
// value), so lanes 15, 31, 47 and 63 hold the row maxima, read as scalars.
__device__ __forceinline__ int wave_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xF, 0xF, false));
    return max(max(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
               max(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}

__host__ __device__ inline int cell_of(const DeviceGrid& g, int l, int x, int y) {
    return l * g.XY + (layer_dir(g, l) ? x * g.Y + y : y * g.X + x);
}

// Decode a reference 3D index into a track-major cell.
__host__ __device__ inline int cell_of_idx(const DeviceGrid& g, int idx) {
    const int l = idx / g.XY, x = idx / g.Y % g.X, y = idx % g.Y;
    return cell_of(g, l, x, y);
}

__host__ __device__ inline int track_of_cell(const DeviceGrid& g, int cell) {
    const int l = cell / g.XY, r = cell % g.XY;
    const int len = layer_dir(g, l) ? g.Y : g.X;
    return g.track_begin[l] + r / len;
}

// of_cost_scaled with of_cost_scale = 1 (src/cost/full_update.cuh:9-14) at
// the three demands d, d + 0.5 and d + 1 of the increments (demand_plus),
// each e = gr_expf(x) or 0.  The first exponential goes through the table
// form and the other two step from it (x + 1/4 and x + 1/2 with capacity,
// x + 3/4 and x + 3/2 without, at the default slopes); every value equals
// the reference form's.
__device__ __forceinline__ float demand_plus(float d, double inc) { return (float)((double)d + inc + 0.0); }

__device__ __forceinline__ void of_cost3(float capacity, float demand, double slope, double zero_slope, double& e0,
                                         double& eh, double& e1) {
    float x[3];
    bool need[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float d = demand_plus(demand, 0.5 * k);
        if ((double)capacity > 0.001) {
            const float diff = d - capacity;
            x[k] = (float)fmin(slope * (double)diff, 1.0 * slope * (double)diff);
            need[k] = true;
        } else {
            x[k] = (float)fmin(zero_slope * (double)d, 1.0 * zero_slope * (double)d);
            need[k] = d > 0.0f;
        }
    }
    // the base exponential is taken even where e(d) = 0 (no capacity and no
    // demand: x = 0), so that the other two can step from it
    double base_v, v;
    const float f0 = gr_expf_fast(x[0], &base_v);
    e0 = need[0] ? (double)f0 : 0.0;
    const float fh = gr_expf_step(x[1], x[0], base_v, &v);
    eh = need[1] ? (double)fh : 0.0;
    const float f1 = gr_expf_step(x[2], x[0], base_v, &v);
    e1 = need[2] ? (double)f1 : 0.0;
}

// incremental_of_cost (src/cost/full_update.cuh:23-32) given the three
// overflow costs e(d), e(d + 0.5), e(d + 1) of a cell.
__device__ __forceinline__ double incr_from(double e_lo, double e_hi, double short_cost) {
    return fmin(1e12, (e_hi - e_lo) * short_cost);
}

}  // namespace grotor
