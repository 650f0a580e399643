// Exact first-fit greedy batching on gfx950.
//
// Reference: cudb::generate_batches_rsmt (src/batching/generate_batches_cpu.hpp:7-73).
// Nets are taken in priority order k; net k joins the first batch none of
// whose earlier members marked any of k's points.  Marks are the cells of
// the net's horizontal / vertical segments (every `sparsity` cells) and the
// stencil {(x-1,y), (x,y-1), (x,y), (x+1,y)} of every point (:35-40).
//
// Equivalently color(k) = mex{ color(j) : j < k, marks(j) contain a point of
// k }: net k needs, for each of its points p, the set of colors of the
// earlier nets that mark p.  The GPU form (after the reference's HIPRT chunk
// path, src/batching/hiprt/grid_chunk.cuh, with the BVH query replaced by
// direct indexing of a uniform G-cell bin grid):
//   1. count_marks / emit_keys: every net's marked cells as 64-bit keys
//      (cell << 32 | k); a radix sort and a unique pass give, per cell, the
//      list of nets marking it in priority order (CSR, one entry per net);
//   2. every cell keeps a color bitmask, a count of completed writes and a
//      count of completed reads.  A list entry is a reader when the cell is
//      one of its net's points (every entry writes: the point stencil marks
//      the cell itself).  A reader at list position i needs exactly the
//      colors of the i entries before it; writes commute, so the only order
//      kept is between a reader and the writes around it: the reader waits
//      until i writes are done, and a writer waits until every reader before
//      it has read;
//   3. color_kernel: a persistent kernel dequeues nets in priority order from
//      a device counter.  The wavefront owning net k reads the masks of its
//      points (each once the cell's write count reaches k's position), takes
//      the minimum excluded color, then ORs its color into every cell it
//      marks (each once the cell's read count covers the readers before k's
//      entry) and counts the write (release).  Spins use relaxed loads with
//      one acquire fence once satisfied.
// Every wait is on strictly earlier nets, which were dequeued earlier, so the
// earliest unfinished net always progresses and every wave exits once the
// counter passes n.  Work is linear in points + marks and the wait chains
// follow the true dependency depth of the greedy coloring.
// The batch size limit of the reference is checked on the host afterwards:
// if no batch exceeds it the unlimited coloring equals the limited one.
#include "gr_kernels.hpp"

namespace grotor {


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

__global__ __launch_bounds__(256) void emit_keys_kernel(BatchGeom b, const int* mark_off,
                                                        unsigned long long* keys) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.n) return;
    int m = mark_off[k];
    for_each_mark(b, k, [&](int cell) { keys[m++] = ((unsigned long long)cell << 32) | (unsigned)k; });
}

// cell_off[c] = first list entry of cell c; list_net[e] = net of entry e.
__global__ __launch_bounds__(256) void build_lists_kernel(const unsigned long long* ukeys, const int* unique_count,
                                                          int cells, int* cell_off, int* list_net) {
    const int U = *unique_count;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > U) return;
    const int cell = e < U ? int(ukeys[e] >> 32) : cells;
    const int prev = e > 0 ? int(ukeys[e - 1] >> 32) : -1;
    for (int c = prev + 1; c <= cell; ++c) cell_off[c] = e;
    if (e < U) list_net[e] = int(ukeys[e] & 0xffffffffu);
}

// position of net k in the sorted list [begin, end) (k is present)
__device__ __forceinline__ int list_position(const int* list, int begin, int end, int k) {
    int lo = begin, hi = end;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// entry_info[e] = 1 for the list entry of every (point cell, net) pair
__global__ __launch_bounds__(256) void mark_readers_kernel(BatchGeom b, const int* cell_off, const int* list_net,
                                                           int* is_reader) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.n) return;
    for (int q = b.pt_off[k]; q < b.pt_off[k + 1]; ++q) {
        const int p = b.pts[q];
        is_reader[list_position(list_net, cell_off[p], cell_off[p + 1], k)] = 1;
    }
}

// entry_info[e] = 2 * (readers before e in its cell's list) + is_reader[e]
__global__ __launch_bounds__(256) void entry_info_kernel(const unsigned long long* ukeys, const int* unique_count,
                                                         const int* cell_off, const int* is_reader,
                                                         const int* reader_scan, int* entry_info) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *unique_count) return;
    const int cell = int(ukeys[e] >> 32);
    entry_info[e] = 2 * (reader_scan[e] - reader_scan[cell_off[cell]]) + is_reader[e];
}

// Spin until *counter >= target with relaxed loads, then one acquire fence.
// False if the spin limit was hit or another wave reported an error.
__device__ __forceinline__ bool wait_count(const ColorArgs& a, const int* counter, int target) {
    long long spins = 0;
    int sleep = 1;
    while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > a.spin_limit) {
            atomicExch(a.error, (int)kDevSpinLimit);
            return false;
        }
        if ((spins & 255) == 0 && __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        if (sleep == 1) __builtin_amdgcn_s_sleep(1);
        else __builtin_amdgcn_s_sleep(4);
        if (spins > 64) sleep = 4;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

__global__ __launch_bounds__(kWave) void color_kernel(ColorArgs a) {
    __shared__ unsigned long long forbidden[kMaxMaskWords];
    __shared__ int s_k, s_color;
    const int lane = threadIdx.x;
    const int W = a.mask_words;
    while (true) {
        if (lane == 0) s_k = atomicAdd(a.next, 1);
        if (lane < kMaxMaskWords) forbidden[lane] = 0ull;
        __syncthreads();
        const int k = s_k;
        if (k >= a.b.n) break;
        // reads: colors of the earlier nets that mark k's points
        for (int q = a.b.pt_off[k] + lane; q < a.b.pt_off[k + 1]; q += kWave) {
            const int p = a.b.pts[q];
            const int begin = a.cell_off[p];
            const int e = list_position(a.list_net, begin, a.cell_off[p + 1], k);
            if (!wait_count(a, a.writes + p, e - begin)) continue;
            for (int w = 0; w < W; ++w) {
                const unsigned long long m = __hip_atomic_load(a.mask + (size_t)p * W + w, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
                if (m) atomicOr(&forbidden[w], m);
            }
            // reader rank r done: later writers of this cell may proceed
            __hip_atomic_fetch_max(a.reads + p, (a.entry_info[e] >> 1) + 1, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (lane == 0) {
            int mex = -1;
            for (int w = 0; w < W && mex < 0; ++w)
                if (~forbidden[w]) mex = w * 64 + __builtin_ctzll(~forbidden[w]);
            if (mex < 0) {
                atomicExch(a.error, (int)kDevColorOverflow);
                mex = 64 * W - 1;
            }
            s_color = mex;
            a.color[k] = mex;
            atomicAdd(a.color_count + mex, 1);
            atomicMax(a.max_color, mex);
        }
        __syncthreads();
        // writes: k's color into every cell it marks (each list entry once)
        const int color = s_color;
        const unsigned long long bit = 1ull << (color & 63);
        for (int m = a.mark_off[k] + lane; m < a.mark_off[k + 1]; m += kWave) {
            const int cell = int(a.keys[m] >> 32);
            const int e = list_position(a.list_net, a.cell_off[cell], a.cell_off[cell + 1], k);
            if (atomicExch(a.claimed + e, 1)) continue;  // duplicate mark of the same cell
            const int info = a.entry_info[e];
            if (!wait_count(a, a.reads + cell, (info >> 1) + (info & 1))) continue;
            atomicOr(a.mask + (size_t)cell * W + (color >> 6), bit);
            __hip_atomic_fetch_add(a.writes + cell, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

}  // namespace grotor
