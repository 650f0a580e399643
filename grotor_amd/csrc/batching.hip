// Exact first-fit greedy batching on gfx950.
//
// Reference: cudb::generate_batches_rsmt (src/batching/generate_batches_cpu.hpp:7-73).
// Nets are taken in priority order k; net k joins the first batch that is not
// full and none of whose earlier members marked any of k's points.  Marks are
// the cells of the net's horizontal / vertical segments (every `sparsity`
// cells) and the stencil {(x-1,y), (x,y-1), (x,y), (x+1,y)} of every point
// (:35-40, :59-66).
//
// Equivalently color(k) = mex( {color(j) : j < k, marks(j) contain a point of
// k} + {batches already full when k is placed} ).  The GPU form:
//   1. count_marks / emit_marks: every mark as (cell, net << 1 | self), where
//      `self` flags the stencil centre of one of the net's points, each cell
//      of a net once; a stable radix sort on the cell gives, per cell, the
//      marks in priority order;
//   2. entry_heads / build_entries: one list entry per (cell, net) pair.  An
//      entry is a reader when the cell is one of its net's points.  A reader
//      at list position e needs the colors of every entry before it; the
//      previous reader r of the same cell has already folded the colors of
//      the entries before r into its prefix mask P(r), so
//          P(e) = P(r) | colors of entries r .. e-1
//      (from the list start when e is the cell's first reader).  A max-scan
//      over "reader or first entry of a cell" positions finds r;
//   3. reader_slots / reader_records: per net, a contiguous record per reader
//      (start entry, own entry, slot of the previous reader);
//   4. color_kernel: a persistent kernel dequeues nets in priority order.  The
//      wavefront owning net k waits for the prefix masks of the previous
//      readers and for the colors of the entries in its ranges, publishes its
//      own prefix masks, takes the minimum excluded color and publishes it.
// A wave holds one net at a time: a wave that held several consecutive nets
// would finish them one after the other, and every net waiting on a later one
// of them would wait for the whole run.  Nets are split over kColorQueues
// queues by residue (net - first) % kColorQueues, each handed out in
// increasing order by its own counter, and a wave dequeues its next net while
// it colors the current one, so the dequeue latency is hidden.
// Every value another wavefront waits on is a single word written once: a
// net's color (-1 until set) or one 64-bit prefix-mask word whose top bit
// says it is ready (63 colors per word).  Waits are therefore relaxed atomic
// loads of that one word, with no fences: nothing is published through a
// second location.  Progress: each wait is on a strictly earlier net.  Let
// k be the earliest uncolored net, in queue q.  If a wave holds k as its
// current net, every net k waits on is colored, so k completes; if a wave
// holds it as its next net, that wave's current net is earlier, hence
// colored, and the wave moves on to k.  Otherwise every net that queue q's
// waves hold is earlier than k, so they are colored and the next dequeue of
// queue q returns k.  The first kColorQueues waves of the grid cover every
// queue, and a wave exits only after its queue passes n.  Each list entry is
// read by at most one reader, so the work is linear in points + marks.
// The batch size limit is applied by the host: colors stay exact up to the
// first net that would overfill a batch; from that net on the batch is
// closed (a forbidden color for later nets) and the suffix is colored again.
#include "gr_kernels.hpp"

namespace grotor {

__device__ __forceinline__ int seg_marks(int first, int second, int step) {
    return second >= first ? (second - first) / step + 1 : 0;
}

// A net's marks, enumerated by its wavefront 64 candidates at a time in a
// fixed order: the stencil centres of its points (its reader marks) first,
// so that after the stable sort the first mark of every (cell, net) run is a
// reader mark whenever the run has one; then the segment cells, spread over
// the lanes (a wave prefix sum of per-segment counts, then candidate t goes
// to lane t mod 64); then the other stencil cells.  Repeated cells of one net
// are dropped through a per-wavefront hash set in LDS (the colouring needs
// each (cell, net) pair once; a net with more than kDedupMax candidates keeps
// its repeats, which the entry runs then merge).  count_marks_kernel and
// emit_marks_kernel run the same enumeration, so they keep the same marks.
constexpr int kDedupSlots = 512, kDedupMax = 256;
constexpr unsigned kNoCell = 0xFFFFFFFFu;
// Nets share a wavefront in groups of kMarkLanes lanes: most nets have a few
// dozen candidate marks, which a group covers in a few steps.
constexpr int kMarkLanes = 16, kMarkGroups = kWave / kMarkLanes;

struct MarkScratch {
    int off[kMarkLanes], start[kMarkLanes], step[kMarkLanes];
    unsigned set[kDedupSlots];
};

// true the first time `cell` is offered to a set of 2^bits slots (at least
// twice the candidates, so the probe ends)
__device__ __forceinline__ bool first_visit(unsigned* set, int bits, unsigned cell) {
    unsigned h = (cell * 2654435761u) >> (32 - bits);
    while (true) {
        const unsigned old = atomicCAS(&set[h], kNoCell, cell);
        if (old == kNoCell) return true;
        if (old == cell) return false;
        h = (h + 1) & ((1u << bits) - 1);
    }
}

// The group's bits of a wavefront ballot.
__device__ __forceinline__ unsigned group_bits(bool pred, int group) {
    return unsigned(__ballot(pred) >> (group * kMarkLanes)) & ((1u << kMarkLanes) - 1);
}

// chunk(keep, cell, reader) once per kMarkLanes candidates, uniformly over
// the net's group of lanes (gl = lane within the group); returns the number
// of candidates
template <class Chunk>
__device__ __forceinline__ int for_each_mark(const BatchGeom& b, int k, int gl, MarkScratch& ms, Chunk&& chunk) {
    constexpr int GW = kMarkLanes;
    const int p0 = b.pt_off[k], np = b.pt_off[k + 1] - p0;
    const int h0 = b.hs_off[k], nh = b.hs_off[k + 1] - h0;
    const int v0 = b.vs_off[k], ns = nh + b.vs_off[k + 1] - v0;
    auto seg_of = [&](int sg, int& first, int& st) {
        const int* seg = sg < nh ? b.hsegs + 2 * (h0 + sg) : b.vsegs + 2 * (v0 + sg - nh);
        st = sg < nh ? b.sparsity * b.Y : b.sparsity;
        first = seg[0];
        return seg_marks(seg[0], seg[1], st);
    };
    // candidates, to decide on the dedup
    int cand = 0;
    for (int q = gl; q < np; q += GW) {
        const int p = b.pts[p0 + q], px = p / b.Y, py = p % b.Y;
        cand += 1 + (px > 0) + (px + 1 < b.X) + (py > 0);
    }
    for (int sg = gl; sg < ns; sg += GW) {
        int first, st;
        cand += seg_of(sg, first, st);
    }
    for (int o = GW / 2; o; o >>= 1) cand += __shfl_xor(cand, o, GW);
    const bool dedup = cand <= kDedupMax;
    int bits = 5;  // the set: at least 32 slots, at least twice the candidates
    while ((1 << bits) < 2 * cand) ++bits;
    if (dedup)
        for (int i = gl; i < (1 << bits); i += GW) ms.set[i] = kNoCell;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    auto offer = [&](bool valid, unsigned cell, unsigned reader) {
        const bool keep = valid && (!dedup || first_visit(ms.set, bits, cell));
        chunk(keep, cell, reader);
    };
    for (int base = 0; base < np; base += GW) {
        const int q = base + gl;
        offer(q < np, q < np ? unsigned(b.pts[p0 + q]) : 0u, 1u);
    }
    for (int sbase = 0; sbase < ns; sbase += GW) {
        const int sg = sbase + gl;
        int cnt = 0, first = 0, st = 1;
        if (sg < ns) cnt = seg_of(sg, first, st);
        int incl = cnt;
        for (int d = 1; d < GW; d <<= 1) {
            const int v = __shfl_up(incl, d, GW);
            if (gl >= d) incl += v;
        }
        const int total = __shfl(incl, GW - 1, GW);
        ms.off[gl] = incl - cnt;
        ms.start[gl] = first;
        ms.step[gl] = st;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (int base = 0; base < total; base += GW) {
            const int t = base + gl;
            unsigned cell = 0;
            if (t < total) {
                int lo = 0, hi = GW - 1;  // last lane whose candidates start at or before t
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (ms.off[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                cell = unsigned(ms.start[lo] + (t - ms.off[lo]) * ms.step[lo]);
            }
            offer(t < total, cell, 0u);
        }
        __builtin_amdgcn_wave_barrier();
    }
    for (int base = 0; base < np; base += GW) {
        const int q = base + gl;
        const bool valid = q < np;
        const int p = valid ? b.pts[p0 + q] : 0, px = p / b.Y, py = p % b.Y;
        offer(valid && px > 0, unsigned(p - b.Y), 0u);
        offer(valid && py > 0, unsigned(p - 1), 0u);
        offer(valid && px + 1 < b.X, unsigned(p + b.Y), 0u);
    }
    return cand;
}

// mark_count[k] = marks net k keeps, cand_count[k] (if given) = its
// candidates before the dedup; kMarkGroups nets per wavefront
__global__ __launch_bounds__(256) void count_marks_kernel(BatchGeom b, int* mark_count, int* cand_count) {
    __shared__ MarkScratch s_ms[4 * kMarkGroups];
    const int g = threadIdx.x / kMarkLanes, gl = threadIdx.x % kMarkLanes, grp = g % kMarkGroups;
    const int k = blockIdx.x * 4 * kMarkGroups + g;
    if (k >= b.n) return;  // uniform per group
    int m = 0;
    const int cand = for_each_mark(b, k, gl, s_ms[g],
                                   [&](bool keep, unsigned, unsigned) { m += __popc(group_bits(keep, grp)); });
    if (gl == 0) {
        mark_count[k] = m;
        if (cand_count) cand_count[k] = cand;
    }
}

// The kept marks of net k at mark_off[k], each as (cell, k << 1 | reader).
__global__ __launch_bounds__(256) void emit_marks_kernel(BatchGeom b, const int* mark_off, unsigned* cell,
                                                         unsigned* tag) {
    __shared__ MarkScratch s_ms[4 * kMarkGroups];
    const int g = threadIdx.x / kMarkLanes, gl = threadIdx.x % kMarkLanes, grp = g % kMarkGroups;
    const int k = blockIdx.x * 4 * kMarkGroups + g;
    if (k >= b.n) return;  // uniform per group
    const unsigned plain = unsigned(k) << 1;
    const unsigned below = (1u << gl) - 1;
    int m = mark_off[k];
    (void)for_each_mark(b, k, gl, s_ms[g], [&](bool keep, unsigned c, unsigned reader) {
        const unsigned mask = group_bits(keep, grp);
        if (keep) {
            const int w = m + __popc(mask & below);
            cell[w] = c;
            tag[w] = plain | reader;
        }
        m += __popc(mask);
    });
}

// One list entry per run: its net, kEntryReader / kEntryFirst flags and the
// marker position used to find the previous reader or the list start.  The
// run's first mark carries the reader flag (reader marks are emitted first).
__global__ __launch_bounds__(256) void build_entries_kernel(const unsigned* cell, const unsigned* tag, const int* eidx,
                                                            int m, int* list_net, int* einfo, int* marker) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int e = eidx[i];
    if (eidx[i + 1] == e) return;  // not a run head
    const unsigned c = cell[i], t = tag[i];
    const int first = (i == 0 || cell[i - 1] != c) ? kEntryFirst : 0;
    const int info = int(t & 1u) * kEntryReader | first;
    list_net[e] = int(t >> 1);
    einfo[e] = info;
    marker[e] = info ? e : -1;
}

// slot_of[e] = record slot of reader entry e inside its net's point range;
// *u (on the device) is the number of entries
__global__ __launch_bounds__(256) void reader_slots_kernel(const int* u, const int* list_net, const int* einfo,
                                                           const int* pt_off, int* rcount, int* slot_of) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *u || !(einfo[e] & kEntryReader)) return;
    const int k = list_net[e];
    slot_of[e] = pt_off[k] + atomicAdd(rcount + k, 1);
}

__global__ __launch_bounds__(256) void reader_records_kernel(const int* u, const int* einfo, const int* prev_marker,
                                                             const int* slot_of, int4* rec) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *u) return;
    const int info = einfo[e];
    if (!(info & kEntryReader)) return;
    int start = e, prev = -1;
    if (!(info & kEntryFirst)) {
        start = prev_marker[e];  // previous reader of this cell, or its first entry
        if (einfo[start] & kEntryReader) prev = slot_of[start];
    }
    rec[slot_of[e]] = make_int4(start, e, prev, 0);
}

namespace {

constexpr unsigned long long kReady = 1ull << 63;
constexpr unsigned long long kColorBits = kReady - 1;

__device__ __forceinline__ bool spin_abort(const ColorArgs& a, long long& spins) {
    if (++spins > a.spin_limit) {
        atomicExch(a.error, (int)kDevSpinLimit);
        return true;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return true;
    __builtin_amdgcn_s_sleep(2);
    return false;
}

// Color of an earlier net: a near (possibly stale) read first, then
// coherent reads until it is set.
__device__ __forceinline__ int await_color(const ColorArgs& a, int net) {
    int c = __hip_atomic_load(a.color + net, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    long long spins = 0;
    while (c < 0) {
        c = __hip_atomic_load(a.color + net, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c < 0 && spin_abort(a, spins)) return 0;
    }
    return c;
}

__device__ __forceinline__ unsigned long long await_mask(const ColorArgs& a, const unsigned long long* p) {
    unsigned long long v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    long long spins = 0;
    while (!(v & kReady)) {
        v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(v & kReady) && spin_abort(a, spins)) return 0;
    }
    return v & kColorBits;
}

}  // namespace

// The colouring workgroup is one wavefront: its LDS exchanges need only the
// wavefront's own ordering, not a workgroup barrier (which would also wait
// for every outstanding global load and store).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Dynamic LDS: per lane W prefix-mask words, then W forbidden words.
__global__ __launch_bounds__(kWave) void color_kernel(ColorArgs a) {
    extern __shared__ unsigned long long lds[];
    __shared__ int s_off[kWave], s_start[kWave];
    __shared__ int s_k;
    const int lane = threadIdx.x;
    const int W = a.mask_words;
    unsigned long long* P = lds;
    unsigned long long* forbidden = lds + kWave * W;
    const int q = blockIdx.x % kColorQueues;
    int* counter = a.next + q * kColorQueueStride;
    // net of the c-th dequeue of queue q, a.n once the queue is exhausted
    auto net_of = [&](int c) {
        const long long k = (long long)a.first + q + (long long)kColorQueues * c;
        return k < a.n ? (int)k : a.n;
    };
    if (lane == 0) s_k = atomicAdd(counter, 1);
    wave_sync();
    int k = net_of(s_k);
    while (k < a.n) {
        int next = 0;
        if (lane == 0) next = atomicAdd(counter, 1);  // result used after net k
        for (int w = lane; w < W; w += kWave) forbidden[w] = 0ull;
        wave_sync();
        for (int i = lane; i < a.nclosed; i += kWave) {  // batches already full when k is placed
            if (a.closed_from[i] > k) continue;
            const int c = a.closed_color[i];
            atomicOr(&forbidden[c / kColorsPerWord], 1ull << (c % kColorsPerWord));
        }
        const int r0 = a.pt_off[k];
        const int nr = a.rcount[k];
        for (int base = 0; base < nr; base += kWave) {
            const bool active = base + lane < nr;
            int4 rc = make_int4(0, 0, -1, 0);
            if (active) rc = a.rec[r0 + base + lane];
            // own row: the previous reader's prefix (or nothing)
            for (int w = 0; w < W; ++w)
                P[lane * W + w] = rc.z >= 0 ? await_mask(a, a.pm + (size_t)rc.z * W + w) : 0ull;
            // entries rc.x .. rc.y-1 of every lane, spread over the wave
            const int cnt = active ? rc.y - rc.x : 0;
            int incl = cnt;
            for (int d = 1; d < kWave; d <<= 1) {
                const int v = __shfl_up(incl, d, kWave);
                if (lane >= d) incl += v;
            }
            const int total = __shfl(incl, kWave - 1, kWave);
            s_off[lane] = incl - cnt;
            s_start[lane] = rc.x;
            wave_sync();
            for (int t = lane; t < total; t += kWave) {
                int lo = 0, hi = kWave - 1;  // last lane whose range starts at or before t
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                const int c = await_color(a, a.list_net[s_start[lo] + (t - s_off[lo])]);
                atomicOr(&P[lo * W + c / kColorsPerWord], 1ull << (c % kColorsPerWord));
            }
            wave_sync();
            if (active) {
                for (int w = 0; w < W; ++w) {
                    const unsigned long long v = P[lane * W + w];
                    __hip_atomic_store(a.pm + (size_t)(r0 + base + lane) * W + w, v | kReady, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    if (v) atomicOr(&forbidden[w], v);
                }
            }
            wave_sync();
        }
        wave_sync();
        if (lane == 0) {
            int mex = -1;
            for (int w = 0; w < W && mex < 0; ++w) {
                const unsigned long long free = ~forbidden[w] & kColorBits;
                if (free) mex = w * kColorsPerWord + __builtin_ctzll(free);
            }
            if (mex < 0) {
                atomicExch(a.error, (int)kDevColorOverflow);
                mex = 0;
            }
            __hip_atomic_store(a.color + k, mex, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_k = next;
        }
        wave_sync();
        k = net_of(s_k);
    }
}

// counts[c] += nets of [first, n) with color c; per-block LDS histogram
__global__ __launch_bounds__(256) void color_hist_kernel(const int* color, int first, int n, int* counts) {
    __shared__ int h[kColorsPerWord * kMaxMaskWords];
    for (int c = threadIdx.x; c < kColorsPerWord * kMaxMaskWords; c += blockDim.x) h[c] = 0;
    __syncthreads();
    for (int k = first + blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        atomicAdd(&h[color[k]], 1);
    __syncthreads();
    for (int c = threadIdx.x; c < kColorsPerWord * kMaxMaskWords; c += blockDim.x)
        if (h[c]) atomicAdd(counts + c, h[c]);
}

}  // namespace grotor
