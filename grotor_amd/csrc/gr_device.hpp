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

// Project exponential, binary32 -> binary32.  Cody-Waite reduction, Taylor
// series to degree 10 in binary64 Horner form (no contraction: the whole
// library is compiled with -ffp-contract=off), scaling by 2^n, one rounding
// to binary32, results below FLT_MIN flushed to +0.  Stands in for CUDA's
// __expf in src/cost/full_update.cuh:11-20.
__device__ __forceinline__ float gr_expf(float xf) {
    if (xf != xf) return xf;
    if (xf > 89.0f) return __builtin_huge_valf();
    if (xf < -88.0f) return 0.0f;
    const double x = (double)xf;
    const double n = floor(x * 0x1.71547652b82fep+0 + 0.5);
    double r = x - n * 0x1.62e42fee00000p-1;
    r = r - n * 0x1.a39ef35793c76p-33;
    double p = 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0x1.0000000000000p-1;
    p = p * r + 1.0;
    p = p * r + 1.0;
    const double v = ldexp(p, (int)n);
    if (v < 0x1.0p-126) return 0.0f;
    return (float)v;
}

// of_cost_scaled with of_cost_scale = 1 (src/cost/full_update.cuh:9-14).
__device__ __forceinline__ double of_cost(float capacity, float demand, double slope, double zero_slope) {
    if ((double)capacity > 0.001) {
        const float diff = demand - capacity;
        return (double)gr_expf((float)fmin(slope * (double)diff, 1.0 * slope * (double)diff));
    }
    if (demand > 0.0f) return (double)gr_expf((float)fmin(zero_slope * (double)demand, 1.0 * zero_slope * (double)demand));
    return 0.0;
}

// incremental_of_cost (src/cost/full_update.cuh:23-32) given the three
// overflow costs e(d), e(d + 0.5), e(d + 1) of a cell.
__device__ __forceinline__ double incr_from(double e_lo, double e_hi, double short_cost) {
    return fmin(1e12, (e_hi - e_lo) * short_cost);
}

__device__ __forceinline__ float demand_plus(float d, double inc) { return (float)((double)d + inc + 0.0); }

}  // namespace grotor
