// Exact first-fit greedy batching on gfx950.
//
// Reference: cudb::generate_batches_rsmt (src/batching/generate_batches_cpu.hpp:7-73).
// Nets are taken in priority order k; net k joins the first batch none of
// whose earlier members marked any of k's points.  Marks are the cells of
// the net's horizontal / vertical segments (every `sparsity` cells) and the
// stencil {(x-1,y), (x,y-1), (x,y), (x+1,y)} of every point (:35-40).
//
// Equivalently color(k) = mex{ color(j) : j < k, marks(j) contain a point of
// k }.  The GPU form (the design the reference's HIPRT chunk path suggests,
// src/batching/hiprt/grid_chunk.cuh, with the BVH query replaced by direct
// indexing of a uniform G-cell bin grid):
//   1. count_marks / emit_marks: every net's marked cells, scanned into one
//      array, with a per-cell counter;
//   2. fill_lists: per-cell lists of the nets that mark the cell (CSR);
//   3. color_kernel: a persistent kernel dequeues nets in priority order from
//      a device counter; the wavefront owning net k reads the lists of k's
//      points, waits (acquire loads, bounded spin) for every earlier listed
//      net to publish its color, builds the set of their colors in an LDS
//      bitmap and publishes the minimum excluded color with a release store.
//      A net only waits on nets dequeued before it, so the earliest
//      unfinished net always makes progress and every wave exits when the
//      counter passes n.
// The batch size limit of the reference is checked on the host afterwards:
// if no batch exceeds it the unlimited coloring equals the limited one.
#include "gr_kernels.hpp"

namespace grotor {

constexpr int kColorWords = 64;  // 64 x 64-bit words: colors 0..4095

__device__ __forceinline__ int hseg_marks(int first, int second, int step) {
    return second >= first ? (second - first) / step + 1 : 0;
}

__global__ __launch_bounds__(256) void count_marks_kernel(BatchGeom b, int* mark_count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.n) return;
    int m = 0;
    for (int s = b.hs_off[k]; s < b.hs_off[k + 1]; ++s) m += hseg_marks(b.hsegs[2 * s], b.hsegs[2 * s + 1], b.sparsity * b.Y);
    for (int s = b.vs_off[k]; s < b.vs_off[k + 1]; ++s) m += hseg_marks(b.vsegs[2 * s], b.vsegs[2 * s + 1], b.sparsity);
    for (int q = b.pt_off[k]; q < b.pt_off[k + 1]; ++q) {
        const int px = b.pts[q] / b.Y, py = b.pts[q] % b.Y;
        m += 1 + (px > 0) + (px + 1 < b.X) + (py > 0);
    }
    mark_count[k] = m;
}

template <class F>
__device__ __forceinline__ void for_each_mark(const BatchGeom& b, int k, F&& f) {
    for (int s = b.hs_off[k]; s < b.hs_off[k + 1]; ++s)
        for (int p = b.hsegs[2 * s]; p <= b.hsegs[2 * s + 1]; p += b.sparsity * b.Y) f(p);
    for (int s = b.vs_off[k]; s < b.vs_off[k + 1]; ++s)
        for (int p = b.vsegs[2 * s]; p <= b.vsegs[2 * s + 1]; p += b.sparsity) f(p);
    for (int q = b.pt_off[k]; q < b.pt_off[k + 1]; ++q) {
        const int px = b.pts[q] / b.Y, py = b.pts[q] % b.Y;
        if (px > 0) f((px - 1) * b.Y + py);
        if (py > 0) f(px * b.Y + py - 1);
        f(px * b.Y + py);
        if (px + 1 < b.X) f((px + 1) * b.Y + py);
    }
}

__global__ __launch_bounds__(256) void emit_marks_kernel(BatchGeom b, const int* mark_off, int* cell_count,
                                                         int* mark_cell) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.n) return;
    int m = mark_off[k];
    for_each_mark(b, k, [&](int cell) {
        mark_cell[m++] = cell;
        atomicAdd(cell_count + cell, 1);
    });
}

__global__ __launch_bounds__(256) void fill_lists_kernel(BatchGeom b, const int* mark_off, const int* mark_cell,
                                                         const int* cell_off, int* cell_cursor, int* cell_list) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.n) return;
    for (int m = mark_off[k]; m < mark_off[k + 1]; ++m) {
        const int cell = mark_cell[m];
        cell_list[cell_off[cell] + atomicAdd(cell_cursor + cell, 1)] = k;
    }
}

__global__ __launch_bounds__(kWave) void color_kernel(ColorArgs a) {
    __shared__ unsigned long long used[kColorWords];
    __shared__ int s_k;
    const int lane = threadIdx.x;
    while (true) {
        if (lane == 0) s_k = atomicAdd(a.next, 1);
        used[lane] = 0ull;
        __syncthreads();
        const int k = s_k;
        if (k >= a.b.n) break;
        for (int q = a.b.pt_off[k]; q < a.b.pt_off[k + 1]; ++q) {
            const int p = a.b.pts[q];
            const int e1 = a.cell_off[p + 1];
            for (int e = a.cell_off[p] + lane; e < e1; e += kWave) {
                const int j = a.cell_list[e];
                if (j >= k) continue;
                int c = __hip_atomic_load(a.color + j, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                long long spins = 0;
                while (c < 0) {
                    if (++spins > a.spin_limit) {
                        atomicExch(a.error, (int)kDevSpinLimit);
                        c = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    c = __hip_atomic_load(a.color + j, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (c < kColorWords * 64) atomicOr(&used[c >> 6], 1ull << (c & 63));
            }
        }
        __syncthreads();
        if (lane == 0) {
            int mex = -1;
            for (int w = 0; w < kColorWords && mex < 0; ++w)
                if (~used[w]) mex = w * 64 + __builtin_ctzll(~used[w]);
            if (mex < 0) {
                atomicExch(a.error, (int)kDevColorOverflow);
                mex = kColorWords * 64 - 1;
            }
            __hip_atomic_store(a.color + k, mex, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            atomicAdd(a.color_count + mex, 1);
            atomicMax(a.max_color, mex);
        }
        __syncthreads();
    }
}

}  // namespace grotor
