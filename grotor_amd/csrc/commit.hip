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
    return rs.routes + (size_t)rs.pin_acc_num[id] * rs.per_pin;
}

// The route a net holds now: its phase-two words once a rip-up stage has
// routed it there (the ripup_flag choice of commit_wire_demand,
// src/routing/demand.cuh:148-155), else its phase-one words.
__device__ __forceinline__ const int* current_route(const RouteStore& rs, const RouteStore& rs2,
                                                    const unsigned char* live2, int id) {
    return live2 && live2[id] ? route_of(rs2, id) : route_of(rs, id);
}

__device__ __forceinline__ void add_demand(const CommitArgs& a, int cell, float value) {
    atomicAdd(a.demand + cell, value);
    a.track_dirty[track_of_cell(a.g, cell)] = 1u;
}

// Wires: the unit edges of all of a net's segments are spread over the W
// lanes of its group (W = 64, a wavefront, or 16: four nets per wavefront;
// segment geometry per lane, a group prefix sum of edge counts, then edge t
// goes to lane t mod W), so a net costs ceil(edges / W) steps however its
// edges split into segments.
template <int W>
__global__ __launch_bounds__(256) void commit_wires_kernel(CommitArgs a) {
    constexpr int G = 256 / W;  // groups per workgroup
    __shared__ int s_off[G][W], s_cell[G][W], s_step[G][W], s_lo[G][W];
    __shared__ unsigned char s_flags[G][W];
    const int lane = threadIdx.x & (W - 1);
    const int wib = threadIdx.x / W;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;  // this lane's group
    const int waves = gridDim.x * blockDim.x / W;
    const DeviceGrid& g = a.g;
    int* off = s_off[wib];
    int* cell0 = s_cell[wib];
    int* step = s_step[wib];
    int* lo_of = s_lo[wib];
    unsigned char* flags = s_flags[wib];  // bit 0: along y, bit 1: one track
    // every group of a wavefront runs the same number of net steps (the
    // group operations below need all lanes of the wavefront)
    const int wave64 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int steps = (a.count - (wave64 * (kWave / W)) + waves - 1) / waves;
    for (int it = 0, s = wave; it < steps; ++it, s += waves) {
        const bool live = s < a.count;
        const int slot = live ? (a.slots ? a.slots[a.begin + s] : a.begin + s) : 0;
        const int* r = live ? current_route(a.rs, a.rs2, a.live2, a.net_ids[slot]) : nullptr;
        const int nseg = live ? (r[0] - 1) / 2 : 0;
        double length = 0.0;
        long long edges = 0;
        int rounds = (nseg + W - 1) / W;  // the same in every group of the wavefront
        for (int o = W; o < kWave; o <<= 1) rounds = max(rounds, __shfl_xor(rounds, o, kWave));
        for (int base = 0; base < rounds * W; base += W) {
            int cnt = 0;
            if (base + lane < nseg) {
                const int ra = r[1 + 2 * (base + lane)], rb = r[2 + 2 * (base + lane)];
                const int layer = ra / g.XY;
                if (layer == rb / g.XY) {
                    const int x0 = ra / g.Y % g.X, y0 = ra % g.Y, x1 = rb / g.Y % g.X, y1 = rb % g.Y;
                    const bool along_y = x0 == x1;
                    const int lo = along_y ? min(y0, y1) : min(x0, x1);
                    cnt = (along_y ? max(y0, y1) : max(x0, x1)) - lo;
                    // a segment along its layer's direction lies on one track
                    const bool one_track = along_y == (bool)layer_dir(g, layer);
                    const int c0 = along_y ? cell_of(g, layer, x0, lo) : cell_of(g, layer, lo, y0);
                    a.timestamp[cell_of(g, layer, x0, y0)] = a.stamp;
                    a.timestamp[cell_of(g, layer, x1, y1)] = a.stamp;
                    if (one_track && cnt) a.track_dirty[track_of_cell(g, c0)] = 1u;
                    cell0[lane] = c0;
                    step[lane] = one_track ? 1 : (along_y ? g.X : g.Y);
                    lo_of[lane] = lo;
                    flags[lane] = (along_y ? 1 : 0) | (one_track ? 2 : 0);
                }
            }
            int incl = cnt;
            for (int d = 1; d < W; d <<= 1) {
                const int v = __shfl_up(incl, d, W);
                if (lane >= d) incl += v;
            }
            const int total = __shfl(incl, W - 1, W);
            off[lane] = incl - cnt;
            edges += cnt;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            for (int t = lane; t < total; t += W) {
                int lo = 0, hi = W - 1;  // last lane whose edges start at or before t
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (off[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                const int d = t - off[lo];
                const int cell = cell0[lo] + d * step[lo];
                const int f = flags[lo];
                atomicAdd(a.demand + cell, (float)a.K);
                if (!(f & 2)) a.track_dirty[track_of_cell(g, cell)] = 1u;
                a.timestamp[cell] = a.stamp;
                length += (f & 1) ? (double)g.y_edge_len[lo_of[lo] + d] : (double)g.x_edge_len[lo_of[lo] + d];
            }
            __builtin_amdgcn_wave_barrier();
        }
        for (int o = W / 2; o; o >>= 1) {
            length += __shfl_xor(length, o, W);
            edges += __shfl_xor(edges, o, W);
        }
        Stats* st = a.stats + wave % kStatStripes;
        if (lane == 0 && length != 0.0) atomicAdd(&st->total_wirelength, length * a.K);
        if (lane == 0 && edges) atomicAdd(&st->wire_edges, (unsigned long long)edges);
    }
}

// Vias: one lane per route segment (W lanes per net, as for the wires); the
// timestamps of all its layer steps are loaded before any is tested, so a
// stack costs one memory round trip.
template <int W>
__global__ __launch_bounds__(256) void commit_vias_kernel(CommitArgs a) {
    const int lane = threadIdx.x & (W - 1);
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const int waves = gridDim.x * blockDim.x / W;
    const DeviceGrid& g = a.g;
    const int wave64 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int steps = (a.count - (wave64 * (kWave / W)) + waves - 1) / waves;
    for (int it = 0, s = wave; it < steps; ++it, s += waves) {
        const bool live = s < a.count;
        const int slot = live ? (a.slots ? a.slots[a.begin + s] : a.begin + s) : 0;
        const int* r = live ? current_route(a.rs, a.rs2, a.live2, a.net_ids[slot]) : nullptr;
        const int end = live ? r[0] : 0;
        long long vias = 0;
        for (int i = 1 + 2 * lane; i < end; i += 2 * W) {
            const int ra = r[i], rb = r[i + 1];
            const int la = ra / g.XY, lb = rb / g.XY;
            const int low = min(la, lb), high = max(la, lb);
            if (low == high) continue;
            const int x = ra / g.Y % g.X, y = ra % g.Y;
            vias += high - low;
            unsigned stamped = 0;
#pragma unroll
            for (int j = 0; j < kMaxLayers - 1; ++j) {
                const int layer = min(low + j, high - 1);  // in range for every j
                if (a.timestamp[cell_of(g, layer, x, y)] >= a.stamp) stamped |= 1u << j;
            }
            for (int layer = low; layer < high; ++layer) {
                if (stamped >> (layer - low) & 1u) continue;
                const int cell = cell_of(g, layer, x, y);
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
        for (int off = W / 2; off; off >>= 1) vias += __shfl_xor(vias, off, W);
        if (lane == 0 && vias) atomicAdd(&a.stats[wave % kStatStripes].total_vias, (unsigned long long)(vias * a.K));
    }
}

template __global__ void commit_wires_kernel<16>(CommitArgs);
template __global__ void commit_wires_kernel<64>(CommitArgs);
template __global__ void commit_vias_kernel<16>(CommitArgs);
template __global__ void commit_vias_kernel<64>(CommitArgs);

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
                                                            RouteStore rs2, const unsigned char* live2, int net_count,
                                                            unsigned char* is_of_net) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= net_count) return;
    const int* r = current_route(rs, rs2, live2, id);
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

// Phase-two storage re-laid at a wider stride: one workgroup copies one
// kept net's words (its pins at old_per_pin words each), offsets in pins
// unchanged.
__global__ __launch_bounds__(256) void restride_routes_kernel(const int* src, int* dst, const int* ids,
                                                              const int* pin_off2, const int* pin_acc_num,
                                                              int old_per_pin, int per_pin) {
    const int id = ids[blockIdx.x];
    const long long n = (long long)(pin_acc_num[id + 1] - pin_acc_num[id]) * old_per_pin;
    const int* s = src + (size_t)pin_off2[id] * old_per_pin;
    int* d = dst + (size_t)pin_off2[id] * per_pin;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) d[i] = s[i];
}

__global__ void fill_int_kernel(int* p, int value, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

// extract_congestionView (src/routing/congestion.cuh:60-79): per 2D cell the
// worst demand - capacity over the routing layers (>= 0).
__global__ void congestion_view_kernel(DeviceGrid g, const float* capacity, const float* demand,
                                       unsigned char* congested, float* xsum, float* ysum) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.XY) return;
    const int x = idx / g.Y, y = idx % g.Y;
    double worst = 0.0;
    for (int l = 0; l < g.L; ++l) {
        const int c = cell_of(g, l, x, y);
        worst = fmax(worst, (double)(demand[c] - capacity[c]));
    }
    congested[idx] = worst > 0.0;
    xsum[idx] = (float)worst;
    ysum[idx] = (float)worst;
}

// extract_congestionView_xsum / _ysum (src/routing/congestion.cuh:81-112):
// exclusive prefix of one row or column in the reference's in-place order
// (at level d every index with bit d set adds the last element of the block
// below it), so the float sums are bit-identical.
__global__ void congestion_prefix_kernel(int tracks, int len, int track_stride, int pos_stride, float* sum) {
    extern __shared__ float s[];
    float* v = sum + (size_t)blockIdx.x * track_stride;
    if (threadIdx.x == 0) s[0] = 0.f;
    for (int p = threadIdx.x; p < len - 1; p += blockDim.x) s[p + 1] = v[(size_t)p * pos_stride];
    __syncthreads();
    for (int d = 0; (1 << d) < len; ++d) {
        for (int i = threadIdx.x; i < len; i += blockDim.x)
            if (i >> d & 1) s[i] += s[(i >> d << d) - 1];
        __syncthreads();
    }
    for (int p = threadIdx.x; p < len; p += blockDim.x) v[(size_t)p * pos_stride] = s[p];
    (void)tracks;
}

}  // namespace grotor
