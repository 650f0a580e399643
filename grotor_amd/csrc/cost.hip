// Cost refresh for gfx950: wire/via costs and the Fenwick-ordered prefix.
//
// Reference: update_single_wcost_ispd24 / update_single_vcost_ispd24
// (src/cost/full_update.cuh:34-83), refresh_fenwick_tracks
// (src/cost/fenwick.cuh:129-168) driven by graph::refresh_cost
// (src/cost/incremental_update.cuh:98-132).
//
// All three outputs are pure functions of (capacity, demand) along one
// routing track: wcost and vcost read the cell and its predecessor on the
// track, presum reads the track's wcost.  The reference tracks dirtiness per
// cell plus per Fenwick node; here one flag per track suffices and a dirty
// track is recomputed whole, which yields the same values bit for bit.
//
// One workgroup per track.  With the track-major layout a track is one
// contiguous run, so thread t owns the k consecutive positions
// [t*k, t*k + k) and every global access is a unit-stride vector access.
// Fenwick nodes are built level by level in LDS in the reference's exact
// addition order; wcost itself never leaves the chip unless requested.
#include "gr_kernels.hpp"

namespace grotor {

__global__ __launch_bounds__(kRefreshThreads) void refresh_tracks_kernel(
    DeviceGrid g, const float* __restrict__ capacity, const float* __restrict__ demand,
    float* __restrict__ vcost, double* __restrict__ presum, float* __restrict__ wcost_out,
    unsigned* __restrict__ track_dirty, int force, CostParams cp, unsigned long long* __restrict__ refreshed) {
    const int T = blockIdx.x;
    if (!force && !track_dirty[T]) return;
    int l = 0;
#pragma unroll
    for (int i = 1; i < kMaxLayers; ++i)
        if (i < g.L && T >= g.track_begin[i]) l = i;
    const int dir = layer_dir(g, l);
    const int len = dir ? g.Y : g.X;
    const int base = l * g.XY + (T - g.track_begin[l]) * len;
    const int* elen = dir ? g.y_edge_len : g.x_edge_len;
    const double sc = g.short_cost[l];
    extern __shared__ double fv[];  // Fenwick values, index = BIT position

    // increments of each thread's last cell: the via cost of the next
    // thread's first cell needs them
    __shared__ double last_h[kRefreshThreads], last_1[kRefreshThreads];
    const int k = (len + blockDim.x - 1) / blockDim.x;
    const int p0 = threadIdx.x * k;
    const int p1 = min(len, p0 + k);
    double first_h = 0.0, first_1 = 0.0;  // increments of cell p0
    if (p0 < len) {
        double prev_h = 0.0, prev_1 = 0.0;
        for (int p = p0; p < p1; ++p) {
            const float c = capacity[base + p], d = demand[base + p];
            double e0, eh, e1;
            of_cost3(c, d, cp.slope, cp.zero_slope, e0, eh, e1);
            const double inc_h = incr_from(e0, eh, sc);
            const double inc_1 = incr_from(e0, e1, sc);
            // wcost: src/cost/full_update.cuh:44-46
            float w = (float)inc_1;
            if (p < len - 1) w = (float)((double)w + g.unit_wire * (double)elen[p]);
            if (cp.integer_wire) w = (float)(long long)fminf(fmaxf(w, 0.f), 1.e7f);
            if (wcost_out) wcost_out[base + p] = w;
            if (p < len - 1) fv[p + 1] = (double)w;
            if (p == p0) {
                first_h = inc_h;
                first_1 = inc_1;
            } else {
                // vcost: src/cost/full_update.cuh:67-82
                const double v = p == len - 1 ? prev_1 : inc_h + prev_h;
                float vf = (float)v;
                vf = (float)((double)vf + cp.via_weight * g.unit_via);
                vcost[base + p] = vf;
            }
            prev_h = inc_h;
            prev_1 = inc_1;
        }
        last_h[threadIdx.x] = prev_h;
        last_1[threadIdx.x] = prev_1;
    }
    if (threadIdx.x == 0) fv[0] = 0.0;
    __syncthreads();
    if (p0 < len) {
        double v;
        if (p0 == 0) v = first_1;
        else if (p0 == len - 1) v = last_1[threadIdx.x - 1];
        else v = first_h + last_h[threadIdx.x - 1];
        float vf = (float)v;
        vf = (float)((double)vf + cp.via_weight * g.unit_via);
        vcost[base + p0] = vf;
    }
    // Fenwick nodes: node(i) = leaf(i) + node(i-1) + node(i-2) + ... +
    // node(i - width/2) for i = (2s+1)*width - 1 (src/cost/fenwick.cuh:149-159)
    for (int width = 2; width <= len; width <<= 1) {
        for (int i = (2 * threadIdx.x + 1) * width - 1; i < len; i += 2 * blockDim.x * width) {
            double value = fv[i];
            for (int off = 1; off < width; off <<= 1) value += fv[i - off];
            fv[i] = value;
        }
        __syncthreads();
    }
    // exclusive prefix along the Fenwick path (src/cost/fenwick.cuh:160-166)
    for (int p = p0; p < p1; ++p) {
        double sum = 0.0;
        for (int j = p; j >= 0; j = (j & (j + 1)) - 1) sum += fv[j];
        presum[base + p] = sum;
    }
    if (threadIdx.x == 0) {
        track_dirty[T] = 0;
        // refreshed cells, striped like the stage statistics
        if (refreshed) atomicAdd(refreshed + (T % kStatStripes) * kCounterStride, (unsigned long long)len);
    }
}

}  // namespace grotor
