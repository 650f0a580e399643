// Demand commit and rip-up classification for gfx950.
//
// Reference: commit_batch_wires / commit_batch_vias / commit_batch_demand
// (src/routing/demand.cuh:236-319); mark_overflow_edges,
// compute_presum_general and mark_overflow_nets driven by graph::ripup
// (src/routing/congestion.cuh:8-58, 114-136).
//
// One 64-lane wavefront walks one net's route words.  Demand changes are
// exact in binary32 (multiples of 0.5) and wirelength sums are integers, so
// the float atomics give order-independent results.  Wires and vias stay in
// two launches so that every via sees all wire timestamps of its batch.  A
// changed cell marks its track dirty for the next cost refresh.
#include "gr_kernels.hpp"

namespace grotor {

__device__ __forceinline__ const int* route_of(const RouteStore& rs, int id) {
    return rs.routes + (size_t)rs.pin_acc_num[id] * kRoutePerPin;
}

__device__ __forceinline__ void add_demand(const CommitArgs& a, int cell, float value) {
    atomicAdd(a.demand + cell, value);
    a.track_dirty[track_of_cell(a.g, cell)] = 1u;
}

__global__ __launch_bounds__(256) void commit_wires_kernel(CommitArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int waves = gridDim.x * blockDim.x / kWave;
    const DeviceGrid& g = a.g;
    for (int s = wave; s < a.count; s += waves) {
        const int slot = a.slots ? a.slots[a.begin + s] : a.begin + s;
        const int* r = route_of(a.rs, a.net_ids[slot]);
        const int end = r[0];
        double length = 0.0;
        long long edges = 0;
        for (int i = 1; i < end; i += 2) {
            const int ra = r[i], rb = r[i + 1], layer = ra / g.XY;
            if (layer != rb / g.XY) continue;
            const int x0 = ra / g.Y % g.X, y0 = ra % g.Y, x1 = rb / g.Y % g.X, y1 = rb % g.Y;
            const bool along_y = x0 == x1;
            const int lo = along_y ? min(y0, y1) : min(x0, x1);
            const int hi = along_y ? max(y0, y1) : max(x0, x1);
            edges += hi - lo;
            if (lane == 0) {
                a.timestamp[cell_of(g, layer, x0, y0)] = a.stamp;
                a.timestamp[cell_of(g, layer, x1, y1)] = a.stamp;
            }
            for (int pos = lo + lane; pos < hi; pos += kWave) {
                const int cell = along_y ? cell_of(g, layer, x0, pos) : cell_of(g, layer, pos, y0);
                add_demand(a, cell, (float)a.K);
                a.timestamp[cell] = a.stamp;
                length += along_y ? (double)g.y_edge_len[pos] : (double)g.x_edge_len[pos];
            }
        }
        for (int off = kWave / 2; off; off >>= 1) length += __shfl_xor(length, off, kWave);
        if (lane == 0 && length != 0.0) atomicAdd(&a.stats->total_wirelength, length * a.K);
        if (lane == 0 && edges) atomicAdd(&a.stats->wire_edges, (unsigned long long)edges);
    }
}

__global__ __launch_bounds__(256) void commit_vias_kernel(CommitArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int waves = gridDim.x * blockDim.x / kWave;
    const DeviceGrid& g = a.g;
    for (int s = wave; s < a.count; s += waves) {
        const int slot = a.slots ? a.slots[a.begin + s] : a.begin + s;
        const int* r = route_of(a.rs, a.net_ids[slot]);
        const int end = r[0];
        long long vias = 0;
        for (int i = 1 + 2 * lane; i < end; i += 2 * kWave) {
            const int ra = r[i], rb = r[i + 1];
            const int la = ra / g.XY, lb = rb / g.XY;
            const int low = min(la, lb), high = max(la, lb);
            if (low == high) continue;
            const int x = ra / g.Y % g.X, y = ra % g.Y;
            vias += high - low;
            for (int layer = low; layer < high; ++layer) {
                const int cell = cell_of(g, layer, x, y);
                if (a.timestamp[cell] >= a.stamp) continue;
                const bool along_y = layer_dir(g, layer);
                const int position = along_y ? y : x, length = along_y ? g.Y : g.X;
                // the previous cell on the track is cell - 1 in track-major order
                if (!position) add_demand(a, cell, (float)a.K);
                else if (position == length - 1) add_demand(a, cell - 1, (float)a.K);
                else {
                    add_demand(a, cell - 1, .5f * a.K);
                    add_demand(a, cell, .5f * a.K);
                }
            }
        }
        for (int off = kWave / 2; off; off >>= 1) vias += __shfl_xor(vias, off, kWave);
        if (lane == 0 && vias) atomicAdd(&a.stats->total_vias, (unsigned long long)(vias * a.K));
    }
}

// Exclusive prefix of per-edge overflow flags along every track; the last
// cell of a track has no edge (src/routing/congestion.cuh:8-21, 42-72 of
// src/cost/prefix_scan.cuh).  Integer sums: any order is exact.
__global__ __launch_bounds__(256) void overflow_prefix_kernel(RipupArgs a) {
    const DeviceGrid& g = a.g;
    const int T = blockIdx.x;
    int l = 0;
#pragma unroll
    for (int i = 1; i < kMaxLayers; ++i)
        if (i < g.L && T >= g.track_begin[i]) l = i;
    const int len = layer_dir(g, l) ? g.Y : g.X;
    const int base = l * g.XY + (T - g.track_begin[l]) * len;
    const int k = (len + blockDim.x - 1) / blockDim.x;
    const int p0 = threadIdx.x * k, p1 = min(len, p0 + k);
    __shared__ int part[256];
    int local = 0;
    for (int p = p0; p < p1; ++p)
        local += (p + 1 < len) && (a.capacity[base + p] + a.threshold <= a.demand[base + p]);
    part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int acc = part[threadIdx.x] - local;
    for (int p = p0; p < p1; ++p) {
        a.of_sum[base + p] = acc;
        acc += (p + 1 < len) && (a.capacity[base + p] + a.threshold <= a.demand[base + p]);
    }
}

__global__ __launch_bounds__(256) void overflow_nets_kernel(DeviceGrid g, const int* of_sum, RouteStore rs,
                                                            int net_count, unsigned char* is_of_net) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= net_count) return;
    const int* r = route_of(rs, id);
    int of = 0;
    unsigned char flag = 0;
    for (int i = 1; i < r[0]; i += 2) {
        const int l = r[i] / g.XY, x0 = r[i] / g.Y % g.X, y0 = r[i] % g.Y;
        const int x1 = r[i + 1] / g.Y % g.X, y1 = r[i + 1] % g.Y;
        if (x0 != x1) of += of_sum[cell_of(g, l, max(x0, x1), y0)] - of_sum[cell_of(g, l, min(x0, x1), y0)];
        if (y0 != y1) of += of_sum[cell_of(g, l, x0, max(y0, y1))] - of_sum[cell_of(g, l, x0, min(y0, y1))];
        if (of > 10) {
            flag = 1;
            break;
        }
    }
    is_of_net[id] = flag;
}

__global__ void fill_int_kernel(int* p, int value, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

}  // namespace grotor
