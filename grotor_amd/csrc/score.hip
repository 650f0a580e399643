// Evaluator-equivalent score of a routing solution, computed from the
// resident route words: NVR_DB::read_gr_solution and
// update_stacked_via_counter (resources/evaluator.cpp:458-675).
//
// The evaluator reads the .out text net by net.  Per net it counts every
// unit wire edge into wire_counter and every via layer-step into the layer's
// via total, then adds a stacked-via demand on the via's own layer for each
// distinct via location that none of the net's own wires covers (its flag
// grid marks the net's wire cells, endpoints included, before the vias are
// counted).  Demand per cell is 2 * wires + stacked vias in a 16-bit field;
// the overflow loss is exp(0.5 * (demand - 2 * capacity) / 2) where the
// capacity exceeds 0.001, exp(0.5 * 1.5 * demand) where it is zero and the
// cell has demand.  Metal 0 never gets a direction in the evaluator
// (:383-388): it is no routing layer, its wires are rejected and only its
// vias are costed.
//
// score_nets_kernel: one wavefront per design net, its lines in the .out
// order (metal-0 pin vias, extra routes, then every subnet's route pairs
// shifted up one metal layer).  The stacking test of a via scans the net's
// own lines (decoded once into LDS), so nets run in parallel with the
// evaluator's sequential result.
// score_cells_kernel: per cell the demand, loss and wirelength, summed per
// layer in fixed-shape partial sums that the host adds in order.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

struct Line {
    int xl, yl, zl, xh, yh, zh;
};

__device__ __forceinline__ Line decode_line(int p0, int p1, int X, int Y) {
    if (p0 > p1) {
        const int t = p0;
        p0 = p1;
        p1 = t;
    }
    const int XY = X * Y;
    return Line{p0 / Y % X, p0 % Y, p0 / XY, p1 / Y % X, p1 % Y, p1 / XY};
}

// Route words of subnet id: its phase-two words when an FLT stage rerouted
// it (src/output.hpp:98-106), else its phase-one words; *pairs receives
// the route pairs, at most what those words can hold.
__device__ __forceinline__ const int* subnet_words(const ScoreArgs& a, int id, int* pairs) {
    const int pins = a.rs.pin_acc_num[id + 1] - a.rs.pin_acc_num[id];
    const int* r;
    int cap;
    if (a.pin_off2 && a.pin_off2[id] >= 0) {
        r = a.routes2 + (size_t)a.pin_off2[id] * a.per_pin2;
        cap = pins * a.per_pin2;
    } else {
        r = a.rs.routes + (size_t)a.rs.pin_acc_num[id] * a.rs.per_pin;
        cap = pins * a.rs.per_pin;
    }
    *pairs = max(0, (min(r[0], cap) - 1) / 2);
    return r;
}

// Calls f(ord, p0, p1) for the lines of net k from lane `lane` with stride
// `stride`, in .out order; ord numbers the lines (pin entries not on metal 0
// keep their number without producing a line).
template <class F>
__device__ __forceinline__ void for_each_line(const ScoreArgs& a, int k, int lane, int stride, F&& f) {
    const int XY = a.X * a.Y;
    int ord = 0;
    const int p0 = a.pin_off[k], p1 = a.pin_off[k + 1];
    for (int j = p0 + lane; j < p1; j += stride) {
        const int p = a.pins[j];
        if (p < XY && f(ord + (j - p0), p, p + XY)) return;
    }
    ord += p1 - p0;
    const int e0 = a.extra_off[k], e1 = a.extra_off[k + 1];
    for (int j = e0 + 2 * lane; j + 1 < e1; j += 2 * stride)
        if (f(ord + (j - e0) / 2, a.extra[j], a.extra[j + 1])) return;
    ord += (e1 - e0) / 2;
    for (int s = a.sub_off[k]; s < a.sub_off[k + 1]; ++s) {
        int pairs;
        const int* r = subnet_words(a, a.subs[s], &pairs);
        for (int j = lane; j < pairs; j += stride)
            if (f(ord + j, r[1 + 2 * j] + XY, r[2 + 2 * j] + XY)) return;
        ord += pairs;
    }
}

// Counts of one line (resources/evaluator.cpp:518-569, 635-675); covered(x,
// y, z, ord) tells whether a wire of the net covers (x, y) on layer z or an
// earlier via line of the net passes (x, y, z).
template <class Covered>
__device__ __forceinline__ void count_line(const ScoreArgs& a, const Line& ln, int ord, bool& bad,
                                           unsigned long long* vias, Covered&& covered) {
    const int X = a.X, Y = a.Y;
    if (ln.zh != ln.zl) {
        if (ln.xh != ln.xl || ln.yh != ln.yl) {
            bad = true;  // a via that moves in x or y
            return;
        }
        for (int z = ln.zl; z < ln.zh; ++z) {
            atomicAdd(vias + z, 1ull);
            if (covered(ln.xl, ln.yl, z, ord)) continue;
            int* st = a.stacked + (size_t)z * X * Y;
            if (a.dir[z] == 0) {
                if (ln.xl > 0 && ln.xl < X - 1) {
                    atomicAdd(st + (ln.xl - 1) * Y + ln.yl, 1);
                    atomicAdd(st + ln.xl * Y + ln.yl, 1);
                } else if (ln.xl > 0) {
                    atomicAdd(st + (ln.xl - 1) * Y + ln.yl, 2);
                } else if (ln.xl < X - 1) {
                    atomicAdd(st + ln.xl * Y + ln.yl, 2);
                }
            } else {
                if (ln.yl > 0 && ln.yl < Y - 1) {
                    atomicAdd(st + ln.xl * Y + ln.yl - 1, 1);
                    atomicAdd(st + ln.xl * Y + ln.yl, 1);
                } else if (ln.yl > 0) {
                    atomicAdd(st + ln.xl * Y + ln.yl - 1, 2);
                } else if (ln.yl < Y - 1) {
                    atomicAdd(st + ln.xl * Y + ln.yl, 2);
                }
            }
        }
        return;
    }
    if (ln.zl == 0) {
        bad = true;  // metal 0 is no routing layer (:383-388, :556-559)
        return;
    }
    int* wc = a.wire + (size_t)ln.zl * X * Y;
    if (a.dir[ln.zl] == 0) {
        if (!(ln.xh > ln.xl && ln.yh == ln.yl)) {
            bad = true;
            return;
        }
        for (int x = ln.xl; x < ln.xh; ++x) atomicAdd(wc + x * Y + ln.yl, 1);
    } else {
        if (!(ln.yh > ln.yl && ln.xh == ln.xl)) {
            bad = true;
            return;
        }
        for (int y = ln.yl; y < ln.yh; ++y) atomicAdd(wc + ln.xl * Y + y, 1);
    }
}

// Whether line m (of the same net) makes the via step at (x, y, z) of line
// ord not stacked.
__device__ __forceinline__ bool covers(const ScoreArgs& a, const Line& m, int mord, int x, int y, int z, int ord) {
    if (m.zl == m.zh) {
        if (m.zl != z || z == 0) return false;
        if (a.dir[z] == 0) return m.xh > m.xl && m.yh == m.yl && m.yl == y && m.xl <= x && x <= m.xh;
        return m.yh > m.yl && m.xh == m.xl && m.xl == x && m.yl <= y && y <= m.yh;
    }
    return mord < ord && m.xl == m.xh && m.yl == m.yh && m.xl == x && m.yl == y && m.zl <= z && z < m.zh;
}

__device__ __forceinline__ int pack_xy(int x, int y) { return x | y << 16; }

}  // namespace

// One wavefront (workgroup) per design net.  A net of at most kScoreLines
// lines is decoded once into LDS, one int4 per line (xl | yl << 16,
// xh | yh << 16, zl | zh << 8 or -1 for a metal-1 pin without a line), and
// every via scans the LDS copy; larger nets scan their lines in global memory.  Via counts
// per layer gather in LDS and leave once per workgroup, into stripe
// blockIdx.x % kStatStripes of a.vias (kScoreLayers counters a stripe): one
// global counter per layer would serialize millions of atomics.
__global__ __launch_bounds__(kWave) void score_nets_kernel(ScoreArgs a) {
    __shared__ int4 s_line[kScoreLines];  // (xl | yl << 16, xh | yh << 16, zl | zh << 8 or -1, 0)
    __shared__ unsigned long long s_vias[kScoreLayers];
    const int lane = threadIdx.x;
    const int X = a.X, Y = a.Y;
    if (lane < kScoreLayers) s_vias[lane] = 0;
    __syncthreads();
    for (int k = blockIdx.x; k < a.nets; k += gridDim.x) {
        // line count: pin entries, extra pairs, subnet pairs
        int n = (a.pin_off[k + 1] - a.pin_off[k]) + (a.extra_off[k + 1] - a.extra_off[k]) / 2;
        int subs_pairs = 0;
        for (int s = a.sub_off[k] + lane; s < a.sub_off[k + 1]; s += kWave) {
            int pairs;
            subnet_words(a, a.subs[s], &pairs);
            subs_pairs += pairs;
        }
        for (int off = kWave / 2; off; off >>= 1) subs_pairs += __shfl_xor(subs_pairs, off, kWave);
        n += subs_pairs;
        bool bad = false;
        if (n <= kScoreLines) {
            for_each_line(a, k, lane, kWave, [&](int ord, int q0, int q1) {
                const Line ln = decode_line(q0, q1, X, Y);
                s_line[ord] = make_int4(pack_xy(ln.xl, ln.yl), pack_xy(ln.xh, ln.yh), ln.zl | ln.zh << 8, 0);
                return false;
            });
            // metal-1 pins print no line
            for (int j = a.pin_off[k] + lane; j < a.pin_off[k + 1]; j += kWave)
                if (a.pins[j] >= X * Y) s_line[j - a.pin_off[k]].z = -1;
            __syncthreads();
            for (int i = lane; i < n; i += kWave) {
                const int4 q = s_line[i];
                if (q.z < 0) continue;
                const Line ln{q.x & 0xFFFF, q.x >> 16, q.z & 0xFF, q.y & 0xFFFF, q.y >> 16, q.z >> 8};
                count_line(a, ln, i, bad, s_vias, [&](int x, int y, int z, int ord) {
                    // covers() on the packed lines: a wire of layer z through
                    // (x, y), or an earlier via column at (x, y) spanning z
                    const int xy = pack_xy(x, y);
                    const bool hor = a.dir[z] == 0;
#pragma unroll 4
                    for (int m = 0; m < n; ++m) {
                        const int4 w = s_line[m];
                        if (w.z < 0) continue;
                        const int zl = w.z & 0xFF, zh = w.z >> 8;
                        if (zl == zh) {
                            if (zl != z || z == 0) continue;
                            if (hor) {
                                if ((w.x >> 16) == y && (w.y >> 16) == y && (w.x & 0xFFFF) <= x &&
                                    x <= (w.y & 0xFFFF) && (w.y & 0xFFFF) > (w.x & 0xFFFF))
                                    return true;
                            } else if ((w.x & 0xFFFF) == x && (w.y & 0xFFFF) == x && (w.x >> 16) <= y &&
                                       y <= (w.y >> 16) && (w.y >> 16) > (w.x >> 16)) {
                                return true;
                            }
                        } else if (m < ord && w.x == xy && w.y == xy && zl <= z && z < zh) {
                            return true;
                        }
                    }
                    return false;
                });
            }
            __syncthreads();  // the LDS lines are reused by the next net
        } else {
            for_each_line(a, k, lane, kWave, [&](int ord, int q0, int q1) {
                count_line(a, decode_line(q0, q1, X, Y), ord, bad, s_vias, [&](int x, int y, int z, int o) {
                    bool hit = false;
                    for_each_line(a, k, 0, 1, [&](int mord, int r0, int r1) {
                        hit = covers(a, decode_line(r0, r1, X, Y), mord, x, y, z, o);
                        return hit;
                    });
                    return hit;
                });
                return false;
            });
        }
        if (__any(bad) && lane == 0) atomicAdd(a.bad_nets, 1ull);
    }
    __syncthreads();
    if (lane < a.M && s_vias[lane])
        atomicAdd(a.vias + (blockIdx.x % kStatStripes) * kScoreLayers + lane, s_vias[lane]);
}

namespace {

// A staged line as a box of metal-layer cells: the cells the evaluator
// marks for it (a wire's cells, both ends included; a via column from zl to
// zh).  Lines of an open net are not boxes; entries without a line (metal-1
// pins) have z = -1.
struct Box {
    int x0, y0, z0, x1, y1, z1;
};

__device__ __forceinline__ Box box_of(int4 q) {
    return Box{q.x & 0xFFFF, q.x >> 16, q.z & 0xFF, q.y & 0xFFFF, q.y >> 16, q.z >> 8};
}

// The evaluator rejects the line (resources/evaluator.cpp:518-569): a via
// that moves, a wire on metal 0 or across its layer's direction, an empty
// wire.
__device__ __forceinline__ bool rejected(const Box& b, const int* dir) {
    if (b.z0 != b.z1) return b.x0 != b.x1 || b.y0 != b.y1;
    if (b.z0 == 0) return true;
    if (dir[b.z0] == 0) return !(b.x1 > b.x0 && b.y1 == b.y0);
    return !(b.y1 > b.y0 && b.x1 == b.x0);
}

__device__ __forceinline__ bool meet(int a0, int a1, int b0, int b1, int gap) {
    return a0 <= b1 + gap && b0 <= a1 + gap;
}

// Whether the flood fill of check_connectivity passes between a cell of A
// and a cell of B: a shared cell, the same (x, y) on adjacent layers, or
// neighbours along a common routing layer's direction (metal 0 has none).
__device__ __forceinline__ bool touch(const Box& A, const Box& B, const int* dir) {
    const bool xi = meet(A.x0, A.x1, B.x0, B.x1, 0), yi = meet(A.y0, A.y1, B.y0, B.y1, 0);
    if (xi && yi && meet(A.z0, A.z1, B.z0, B.z1, 1)) return true;
    const bool x1 = meet(A.x0, A.x1, B.x0, B.x1, 1), y1 = meet(A.y0, A.y1, B.y0, B.y1, 1);
    if (!(yi && x1) && !(xi && y1)) return false;
    for (int z = max(max(A.z0, B.z0), 1); z <= min(A.z1, B.z1); ++z)
        if (dir[z] == 0 ? yi && x1 : xi && y1) return true;
    return false;
}

__device__ __forceinline__ bool holds(const Box& b, int x, int y, int z) {
    return b.x0 <= x && x <= b.x1 && b.y0 <= y && y <= b.y1 && b.z0 <= z && z <= b.z1;
}

__device__ __forceinline__ int uf_load(int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Root of i with path halving; roots only ever link to smaller indices.
__device__ __forceinline__ int uf_find(int* par, int i) {
    for (;;) {
        const int p = uf_load(par + i);
        if (p == i) return i;
        const int g = uf_load(par + p);
        if (g != p) __hip_atomic_store(par + i, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        i = p;
    }
}

__device__ __forceinline__ void uf_union(int* par, int a, int b) {
    for (;;) {
        a = uf_find(par, a);
        b = uf_find(par, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(par + a, a, b) == a) return;
    }
}

}  // namespace

// One wavefront per design net.  Its lines are staged as boxes (in LDS, or
// for nets that may exceed kScoreLines lines in the global staging the host
// sized), lines whose cells the flood fill connects are united (union-find,
// lanes over line pairs), the components holding a marked access point of
// the first pin are flagged, and every other pin needs an access point in a
// flagged component: the evaluator's traced cells are exactly the marked
// cells of those components.
__global__ __launch_bounds__(kWave) void connectivity_kernel(ScoreArgs a, ConnArgs c) {
    __shared__ int4 s_line[kScoreLines];
    __shared__ int s_par[kScoreLines];
    __shared__ unsigned long long s_counts[2];
    const int lane = threadIdx.x;
    const int X = a.X, Y = a.Y, XY = X * Y;
    if (lane < 2) s_counts[lane] = 0;
    __syncthreads();
    for (int k = blockIdx.x; k < a.nets; k += gridDim.x) {
        int n = (a.pin_off[k + 1] - a.pin_off[k]) + (a.extra_off[k + 1] - a.extra_off[k]) / 2;
        int subs_pairs = 0;
        for (int s = a.sub_off[k] + lane; s < a.sub_off[k + 1]; s += kWave) {
            int pairs;
            subnet_words(a, a.subs[s], &pairs);
            subs_pairs += pairs;
        }
        for (int off = kWave / 2; off; off >>= 1) subs_pairs += __shfl_xor(subs_pairs, off, kWave);
        n += subs_pairs;
        const bool big = n > kScoreLines;
        // the host sizes every net that can exceed kScoreLines lines
        int4* L = big ? c.big_lines + c.big_off[k] : s_line;
        int* par = big ? c.big_par + c.big_off[k] : s_par;
        for_each_line(a, k, lane, kWave, [&](int ord, int q0, int q1) {
            const Line ln = decode_line(q0, q1, X, Y);
            L[ord] = make_int4(pack_xy(ln.xl, ln.yl), pack_xy(ln.xh, ln.yh), ln.zl | ln.zh << 8, 0);
            return false;
        });
        for (int j = a.pin_off[k] + lane; j < a.pin_off[k + 1]; j += kWave)
            if (a.pins[j] >= XY) L[j - a.pin_off[k]].z = -1;  // metal-1 pins print no line
        __syncthreads();
        bool bad = false;
        for (int i = lane; i < n; i += kWave) {
            const int4 q = L[i];
            par[i] = i;
            if (q.z >= 0) bad |= rejected(box_of(q), a.dir);
        }
        unsigned char st = 0;
        if (__any(bad)) {
            st = 1;  // has_connectivity_violation: counted open, never flood filled (:529-530)
        } else {
            __syncthreads();
            for (int i = lane; i < n; i += kWave) {
                const int4 qi = L[i];
                if (qi.z < 0) continue;
                const Box bi = box_of(qi);
                for (int j = 0; j < i; ++j) {
                    const int4 qj = L[j];
                    if (qj.z >= 0 && touch(bi, box_of(qj), a.dir)) uf_union(par, i, j);
                }
            }
            __syncthreads();
            for (int i = lane; i < n; i += kWave) par[i] = uf_find(par, i);
            __syncthreads();
            // components holding a marked access point of the first pin
            const int p0 = c.apin_off[k], p1 = c.apin_off[k + 1];
            if (p1 > p0) {
                for (int t = c.acc_off[p0]; t < c.acc_off[p0 + 1]; ++t) {
                    const int cell = c.acc[t];
                    const int z = cell / XY, x = cell / Y % X, y = cell % Y;
                    for (int i = lane; i < n; i += kWave) {
                        const int4 q = L[i];
                        if (q.z >= 0 && holds(box_of(q), x, y, z)) L[par[i]].w = 1;
                    }
                }
                __syncthreads();
                bool open = false;
                for (int p = p0 + 1 + lane; p < p1; p += kWave) {
                    bool ok = false;
                    for (int t = c.acc_off[p]; t < c.acc_off[p + 1] && !ok; ++t) {
                        const int cell = c.acc[t];
                        const int z = cell / XY, x = cell / Y % X, y = cell % Y;
                        for (int i = 0; i < n && !ok; ++i) {
                            const int4 q = L[i];
                            ok = q.z >= 0 && L[par[i]].w && holds(box_of(q), x, y, z);
                        }
                    }
                    open |= !ok;
                }
                if (__any(open)) st = 2;
            }
        }
        if (lane == 0) {
            c.status[k] = st;
            if (st) atomicAdd(&s_counts[st - 1], 1ull);
        }
        __syncthreads();  // the staging is reused by the next net
    }
    __syncthreads();
    if (lane < 2 && s_counts[lane]) atomicAdd(c.counts + lane, s_counts[lane]);
}

// Block b of layer blockIdx.y covers cells [b * kScoreCells, (b + 1) *
// kScoreCells) of the layer; its loss and wirelength sums go to
// partial_loss / partial_wl[layer * gridDim.x + b].
__global__ __launch_bounds__(256) void score_cells_kernel(ScoreArgs a, const double* capacity, const int* x_len,
                                                          const int* y_len, double* partial_loss,
                                                          unsigned long long* partial_wl, unsigned short* demand_out) {
    __shared__ double s_loss[256];
    __shared__ unsigned long long s_wl[256];
    const int z = blockIdx.y;
    const int XY = a.X * a.Y;
    if (z == 0) {  // metal 0 is no routing layer: no demand, loss or wirelength
        const int c0 = blockIdx.x * kScoreCells, c1 = min(XY, c0 + kScoreCells);
        if (demand_out)
            for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) demand_out[c] = 0;
        if (threadIdx.x == 0) {
            partial_loss[blockIdx.x] = 0.0;
            partial_wl[blockIdx.x] = 0;
        }
        return;
    }
    const int dir = a.dir[z];
    double loss = 0.0;
    unsigned long long wl = 0;
    const int c0 = blockIdx.x * kScoreCells;
    const int c1 = min(XY, c0 + kScoreCells);
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const size_t cell = (size_t)z * XY + c;
        const int w = a.wire[cell];
        const unsigned short d = (unsigned short)(2 * w + a.stacked[cell]);  // NVR_Gcell::m_demand : 16
        if (demand_out) demand_out[cell] = d;
        const double cap = capacity[cell];
        if (cap > 0.001) {
            const double overflow = (double)d - 2 * cap;
            loss += exp(overflow / 2 * 0.5);
        } else if (cap >= 0 && d > 0) {
            loss += exp(1.5 * (double)d * 0.5);
        }
        const int x = c / a.Y, y = c % a.Y;
        if (dir == 0) {
            if (x < a.X - 1) wl += (unsigned long long)(w * x_len[x]);
        } else if (y < a.Y - 1) {
            wl += (unsigned long long)(w * y_len[y]);
        }
    }
    s_loss[threadIdx.x] = loss;
    s_wl[threadIdx.x] = wl;
    __syncthreads();
    for (int h = blockDim.x / 2; h; h >>= 1) {
        if (threadIdx.x < h) {
            s_loss[threadIdx.x] += s_loss[threadIdx.x + h];
            s_wl[threadIdx.x] += s_wl[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial_loss[z * gridDim.x + blockIdx.x] = s_loss[0];
        partial_wl[z * gridDim.x + blockIdx.x] = s_wl[0];
    }
}

}  // namespace grotor
