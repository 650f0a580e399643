// Z-shape pattern routing (GGR) for gfx950.
//
// Reference: zshape::solve<Warps>, viaSweep and reset_nodes
// (src/routing/zshape.cuh:42-404), convert_routes
// (src/routing/zshape_topology.hpp:233-289).
//
// One 64-lane wavefront routes one tree (the reference uses a 32-lane
// warp).  Nets of a batch have disjoint bounding boxes, so they share one
// map of LAYER = L + 1 planes of n x n cells (n = max(X, Y)) without
// touching each other's cells; map layer i is routing layer i - 1 and its
// second index runs along the layer's tracks.  Per tree edge (child -> node)
// the phases are: clear the box's two rows and two columns on plane 4; first
// leg along the child's row/column; via sweeps at every cell of that leg;
// middle leg across the box; via sweeps on the node's row/column; last leg
// into child slot j of the node.  Lanes take disjoint cells in every phase,
// so the values equal a sequential run; the last leg's per-layer first
// minimum is a wavefront reduction.  Each cell's via sweep, the child's
// value in the first leg and the node's interval combination work on values
// loaded once into registers instead of chains of dependent global accesses.
// The node's interval combination and the traceback run on lane 0, as in
// the reference.  Costs are integers:
// int(min(|presum(a) - presum(b)|, 5e8)) per wire, int(clamp(vcost, 1, 1e6))
// per via landing.  Indices decoded from the traceback tables are checked
// against the map before use; a bad one sets the error word.
#include "gr_kernels.hpp"

namespace grotor {

namespace {

constexpr int kZInf = 1000000000;

struct Z {
    const ZshapeArgs& a;
    long long nn, P;
    __device__ Z(const ZshapeArgs& args) : a(args), nn((long long)args.n * args.n), P(args.layers * nn) {}
    __device__ __forceinline__ bool vert(int i) const { return ((i & 1) ^ a.dirv) != 0; }
    __device__ __forceinline__ long long cell(int i, int x, int y) const {
        return i * nn + (vert(i) ? (long long)x * a.n + y : (long long)y * a.n + x);
    }
    // the same (x, y) in the orientation of the layer next to i
    __device__ __forceinline__ long long cell_swapped(int i_other, int i, int x, int y) const {
        return i_other * nn + (vert(i) ? (long long)y * a.n + x : (long long)x * a.n + y);
    }
    __device__ __forceinline__ double wire_sum(long long index) const {
        const int layer = int(index / nn) - 1;
        if (layer < 0 || layer >= a.g.L) return 0.0;
        const int p = int(index / a.n % a.n), q = int(index % a.n);
        const bool along_y = layer_dir(a.g, layer);
        return a.presum[cell_of(a.g, layer, along_y ? p : q, along_y ? q : p)];
    }
    __device__ __forceinline__ int via(long long index) const {
        const int layer = int(index / nn) - 1;
        if (layer < 0 || layer >= a.g.L) return kZInf;
        const int p = int(index / a.n % a.n), q = int(index % a.n);
        const bool along_y = layer_dir(a.g, layer);
        return (int)fminf(fmaxf(a.vcost[cell_of(a.g, layer, along_y ? p : q, along_y ? q : p)], 1.f), 1.e6f);
    }
    __device__ __forceinline__ int wire(long long p, long long q) const {
        return (int)fmin(fabs(wire_sum(p) - wire_sum(q)), double(kZInf / 2));
    }
    // wire(p, q) with wire_sum(p) already loaded
    __device__ __forceinline__ int wire_from(double wp, long long q) const {
        return (int)fmin(fabs(wp - wire_sum(q)), double(kZInf / 2));
    }
};

// The via sweeps at one (x, y) of plane 4: upwards then downwards through
// the layer stack, each step taking the neighbour's value plus the via cost
// when it is smaller.  The stack is loaded once and swept in registers, in
// the reference's order (zshape.cuh viaSweep); only this lane touches it.
__device__ void via_sweep(const Z& z, int x, int y) {
    int* m = z.a.map + 4 * z.P;
    int* pr = z.a.prev + 4 * z.P;
    const int L1 = z.a.layers;
    long long idx[kMaxLayers];
    int mv[kMaxLayers], pv[kMaxLayers], pl[kMaxLayers], vc[kMaxLayers];
#pragma unroll
    for (int i = 1; i < kMaxLayers; i++) {
        if (i >= L1) continue;
        idx[i] = z.cell(i, x, y);
        mv[i] = m[idx[i]];
        pv[i] = pr[idx[i]];
        pl[i] = (int)idx[i];
        vc[i] = i < L1 - 1 ? z.via(idx[i]) : 0;
    }
#pragma unroll
    for (int i = 2; i < kMaxLayers; i++) {
        if (i >= L1) continue;
        const int v = mv[i - 1] + vc[i - 1];
        if (mv[i] > v) {
            mv[i] = v;
            pv[i] = pv[i - 1];
            pl[i] = (int)idx[i - 1];
        }
    }
#pragma unroll
    for (int i = kMaxLayers - 2; i >= 1; i--) {
        if (i > L1 - 2) continue;
        const int v = mv[i + 1] + vc[i];
        if (mv[i] > v) {
            mv[i] = v;
            pv[i] = pv[i + 1];
            pl[i] = (int)idx[i + 1];
        }
    }
#pragma unroll
    for (int i = 1; i < kMaxLayers; i++) {
        if (i >= L1) continue;
        m[idx[i]] = mv[i];
        pr[idx[i]] = pv[i];
        pr[idx[i] + z.P] = pl[i];
    }
}

__device__ void zedge(const Z& z, int lane, int j, int fromx, int fromy, int tox, int toy) {
    const ZshapeArgs& a = z.a;
    const long long nn = z.nn, P = z.P;
    const int L1 = a.layers, n = a.n;
    int* map = a.map;
    int* prev = a.prev;
    int* prev2 = a.prev + 10 * P;
    int* m4 = map + 4 * P;
    int* p4 = prev + 4 * P;
    const int x0 = min(fromx, tox), x1 = max(fromx, tox), y0 = min(fromy, toy), y1 = max(fromy, toy);
    if (fromx != tox && fromy != toy) {
        for (int q = x0 + lane; q <= x1; q += kWave)
            for (int i = 1; i < L1; i++) {
                if (z.vert(i)) {
                    m4[i * nn + (long long)q * n + toy] = kZInf;
                    m4[i * nn + (long long)q * n + fromy] = kZInf;
                } else {
                    m4[i * nn + (long long)toy * n + q] = kZInf;
                    m4[i * nn + (long long)fromy * n + q] = kZInf;
                }
            }
        __syncthreads();
        for (int q = y0 + lane; q <= y1; q += kWave)
            for (int i = 1; i < L1; i++) {
                if (z.vert(i)) {
                    m4[i * nn + (long long)tox * n + q] = kZInf;
                    m4[i * nn + (long long)fromx * n + q] = kZInf;
                } else {
                    m4[i * nn + (long long)q * n + tox] = kZInf;
                    m4[i * nn + (long long)q * n + fromx] = kZInf;
                }
            }
        __syncthreads();
        // first leg along the child's column (vertical layers) or row
        // (horizontal layers); the child's value and prefix sum per layer are
        // loaded once, and the layers need no barrier between them
        {
            int cv[kMaxLayers];
            double cw[kMaxLayers];
#pragma unroll
            for (int i = 1; i < kMaxLayers; i++) {
                if (i >= L1) continue;
                const long long last = z.vert(i) ? i * nn + (long long)fromx * n + fromy : i * nn + (long long)fromy * n + fromx;
                cv[i] = map[last];
                cw[i] = z.wire_sum(last);
            }
            for (int q = y0 + lane; q <= y1; q += kWave) {
                if (q == fromy) continue;
#pragma unroll
                for (int i = 1; i < kMaxLayers; i++) {
                    if (i >= L1 || !z.vert(i)) continue;
                    const long long cur = i * nn + (long long)fromx * n + q;
                    m4[cur] = cv[i] + z.wire_from(cw[i], cur);
                    p4[cur] = (int)(i * nn + (long long)fromx * n + fromy);
                }
            }
            for (int q = x0 + lane; q <= x1; q += kWave) {
                if (q == fromx) continue;
#pragma unroll
                for (int i = 1; i < kMaxLayers; i++) {
                    if (i >= L1 || z.vert(i)) continue;
                    const long long cur = i * nn + (long long)fromy * n + q;
                    m4[cur] = cv[i] + z.wire_from(cw[i], cur);
                    p4[cur] = (int)(i * nn + (long long)fromy * n + fromx);
                }
            }
            __syncthreads();
        }
        for (int q = y0 + lane; q <= y1; q += kWave)
            if (q != fromy) via_sweep(z, fromx, q);
        __syncthreads();
        for (int q = x0 + lane; q <= x1; q += kWave)
            if (q != fromx) via_sweep(z, q, fromy);
        __syncthreads();
        // middle leg across the box: vertical layers from row fromy to row
        // toy, horizontal layers from column fromx to column tox; cells of
        // different layers and lanes are disjoint, so one barrier suffices
        for (int q = x0 + lane; q <= x1; q += kWave) {
#pragma unroll
            for (int i = 1; i < kMaxLayers; i++) {
                if (i >= L1 || !z.vert(i)) continue;
                const long long last = i * nn + (long long)q * n + fromy, cur = i * nn + (long long)q * n + toy;
                const int cost = z.wire(last, cur), ml = m4[last], mc = m4[cur], pl = p4[last];
                if (ml + cost < mc) {
                    m4[cur] = ml + cost;
                    p4[cur] = pl;
                }
            }
        }
        for (int q = y0 + lane; q <= y1; q += kWave) {
#pragma unroll
            for (int i = 1; i < kMaxLayers; i++) {
                if (i >= L1 || z.vert(i)) continue;
                const long long last = i * nn + (long long)q * n + fromx, cur = i * nn + (long long)q * n + tox;
                const int cost = z.wire(last, cur), ml = m4[last], mc = m4[cur], pl = p4[last];
                if (ml + cost < mc) {
                    m4[cur] = ml + cost;
                    p4[cur] = pl;
                }
            }
        }
        __syncthreads();
        for (int q = y0 + lane; q <= y1; q += kWave)
            if (q != toy) via_sweep(z, tox, q);
        __syncthreads();
        for (int q = x0 + lane; q <= x1; q += kWave)
            if (q != tox) via_sweep(z, q, toy);
        __syncthreads();
        // last leg into child slot j of the node, per layer the first
        // minimum over the leg's cells in ascending order: each lane takes
        // every 64th cell, then a wavefront reduction keeps the smallest
        // value and, among equal values, the smallest cell
#pragma unroll 1
        for (int i = 1; i < L1; i++) {
            const bool v = z.vert(i);
            const long long cur = v ? i * nn + (long long)tox * n + toy : i * nn + (long long)toy * n + tox;
            const int lo = v ? y0 : x0, hi = v ? y1 : x1, skip = v ? toy : tox;
            const double wc = z.wire_sum(cur);
            int best = kZInf * 2, bq = -1;  // above every candidate (at most kZInf + kZInf / 2)
            for (int q = lo + lane; q <= hi; q += kWave) {
                if (q == skip) continue;
                const long long last = v ? i * nn + (long long)tox * n + q : i * nn + (long long)toy * n + q;
                const int cand = m4[last] + (int)fmin(fabs(z.wire_sum(last) - wc), double(kZInf / 2));
                if (cand < best) {
                    best = cand;
                    bq = q;
                }
            }
            for (int o = kWave / 2; o; o >>= 1) {
                const int ob = __shfl_xor(best, o, kWave), oq = __shfl_xor(bq, o, kWave);
                if (ob < best || (ob == best && oq >= 0 && (bq < 0 || oq < bq))) {
                    best = ob;
                    bq = oq;
                }
            }
            if (lane == 0 && bq >= 0 && best < map[cur + j * P]) {
                const long long last = v ? i * nn + (long long)tox * n + bq : i * nn + (long long)toy * n + bq;
                map[cur + j * P] = best;
                prev[cur + j * P] = p4[last];
                prev2[cur + j * P] = prev[5 * P + last];
            }
        }
    } else if (fromx == tox) {
        for (int i = 1 + lane; i < L1; i += kWave) {
            if (z.vert(i)) {
                const long long last = i * nn + (long long)fromx * n + fromy, cur = i * nn + (long long)tox * n + toy;
                const int cost = z.wire(last, cur);
                if (map[cur + j * P] > map[last] + cost) {
                    map[cur + j * P] = map[last] + cost;
                    prev[cur + j * P] = (int)last;
                }
            } else {
                map[i * nn + (long long)toy * n + tox + j * P] = kZInf;
            }
        }
    } else {
        for (int i = 1 + lane; i < L1; i += kWave) {
            if (!z.vert(i)) {
                const long long last = i * nn + (long long)fromy * n + fromx, cur = i * nn + (long long)toy * n + tox;
                const int cost = z.wire(last, cur);
                if (map[cur + j * P] > map[last] + cost) {
                    map[cur + j * P] = map[last] + cost;
                    prev[cur + j * P] = (int)last;
                }
            } else {
                map[i * nn + (long long)tox * n + toy + j * P] = kZInf;
            }
        }
    }
    __syncthreads();
}

// The node's interval combination (zshape.cuh, lane 0 as in the
// reference): for every layer interval [lower, upper] the cheapest entry of
// each child slot inside it plus the vias of the interval, kept where it
// lowers the node's plane-0 value.  The node's cells are loaded once and
// combined in registers in the reference's order.
__device__ void znode(const Z& z, const int* rec) {
    const ZshapeArgs& a = z.a;
    const int tox = rec[0] / a.n, toy = rec[0] % a.n, minlayer = rec[1], maxlayer = rec[2];
    const int L1 = a.layers;
    const bool has[3] = {rec[3] != -1, rec[4] != -1, rec[5] != -1};
    long long idx[kMaxLayers];
    int slot[3][kMaxLayers], mv[kMaxLayers], pv[kMaxLayers], via[kMaxLayers];
#pragma unroll
    for (int i = 0; i < kMaxLayers; i++) {
        if (i >= L1) continue;
        idx[i] = z.cell(i, tox, toy);
#pragma unroll
        for (int j = 0; j < 3; j++) slot[j][i] = has[j] ? a.map[idx[i] + (j + 1) * z.P] : 0;
        mv[i] = a.map[idx[i]];
        pv[i] = a.prev[idx[i]];
        via[i] = z.via(idx[i]);
    }
#pragma unroll 1
    for (int lower = 0; lower < L1; lower++) {
        int via_sum = 0, min_cost = kZInf, bestpos[3] = {0, 0, 0};
        int costs[3] = {has[0] ? kZInf : 0, has[1] ? kZInf : 0, has[2] ? kZInf : 0};
#pragma unroll
        for (int upper = 0; upper < kMaxLayers; upper++) {
            if (upper < lower || upper >= L1) continue;
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (has[j] && slot[j][upper] < costs[j]) {
                    costs[j] = slot[j][upper];
                    bestpos[j] = upper;
                }
            if (1LL * costs[0] + costs[1] + costs[2] < min_cost) min_cost = costs[0] + costs[1] + costs[2];
            if (minlayer == -1 || (lower <= minlayer && maxlayer <= upper)) {
                const int code = (((lower * L1 + upper) * L1 + bestpos[2]) * L1 + bestpos[1]) * L1 + bestpos[0];
#pragma unroll
                for (int i = 0; i < kMaxLayers; i++) {
                    if (i < lower || i > upper) continue;
                    if (mv[i] > min_cost + via_sum) {
                        mv[i] = min_cost + via_sum;
                        pv[i] = code;
                    }
                }
            }
            via_sum += via[upper];
        }
    }
#pragma unroll
    for (int i = 0; i < kMaxLayers; i++) {
        if (i >= L1) continue;
        a.map[idx[i]] = mv[i];
        a.prev[idx[i]] = pv[i];
    }
}

struct Emitter {
    int* routes;
    int cap;
    int* error;
    __device__ __forceinline__ void operator()(long long p, long long q) {
        if (routes[0] + 1 >= cap) {
            atomicExch(error, (int)kDevRouteOverflow);
            return;
        }
        routes[routes[0]++] = (int)p;
        routes[routes[0]++] = (int)q;
    }
};

__device__ __forceinline__ bool in_map(const Z& z, long long idx) { return idx >= 0 && idx < z.P; }

__device__ void ztrace(const Z& z, const int* pts, int node_cnt, int* routes, int cap) {
    const ZshapeArgs& a = z.a;
    const long long nn = z.nn, P = z.P;
    const int L1 = a.layers, n = a.n;
    int* map = a.map;
    const int* prev = a.prev;
    const int* prev2 = a.prev + 10 * P;
    Emitter emit{routes, cap, a.error};
    int min_total = kZInf;
    for (int i = 0; i < L1; i++) {
        const long long idx = z.cell(i, pts[0] / n, pts[0] % n);
        if (map[idx] < min_total) {
            min_total = map[idx];
            map[5 * P + pts[0]] = i;
        }
    }
    if (min_total == kZInf || min_total < 0) {
        routes[0] = -1;
        return;
    }
    for (int t = 0; t < 6 * node_cnt; t += 6) {
        const int x = pts[t] / n, y = pts[t] % n, layer = map[5 * P + pts[t]];
        if (layer < 0 || layer >= L1) {
            atomicExch(a.error, (int)kDevNoSolution);
            return;
        }
        const long long idx = z.cell(layer, x, y);
        const int code = prev[idx];
        const int minl = code / L1 / L1 / L1 / L1, maxl = code / L1 / L1 / L1 % L1;
        for (int i = minl; i < maxl && i < L1; i++) emit(z.cell(i, x, y), -1);
        int bestpos = code;
        for (int j = 0; j < 3; j++) {
            if (pts[t + 3 + j] == -1) break;
            const int child = pts[pts[t + 3 + j]];
            const int entry_layer = bestpos % L1;
            const long long entry = z.cell(entry_layer, x, y);
            bestpos /= L1;
            const long long last = prev[(j + 1) * P + entry];
            if (!in_map(z, last)) {
                atomicExch(a.error, (int)kDevNoSolution);
                return;
            }
            const int last_layer = int(last / nn);
            int lastx = int(last / n % n), lasty = int(last % n);
            if (!z.vert(last_layer)) {
                const int s = lastx;
                lastx = lasty;
                lasty = s;
            }
            if (lastx != x && lasty != y) {
                const long long mid = prev2[(j + 1) * P + entry];
                if (!in_map(z, mid)) {
                    atomicExch(a.error, (int)kDevNoSolution);
                    return;
                }
                const int mid_layer = int(mid / nn);
                int midx = int(mid / n % n), midy = int(mid % n);
                if (!z.vert(mid_layer)) {
                    const int s = midx;
                    midx = midy;
                    midy = s;
                }
                const long long p1 = z.vert(entry_layer) ? entry - y + midy : entry - x + midx;
                const long long p3 = z.vert(mid_layer) ? mid - midy + lasty : mid - midx + lastx;
                const long long p4 = z.vert(last_layer) ? last - lasty + midy : last - lastx + midx;
                if (entry != p1) emit(min(entry, p1), entry < p1 ? p1 - entry : entry - p1);
                if (p1 != mid)
                    for (int i = min(entry_layer, mid_layer); i < max(entry_layer, mid_layer); i++)
                        emit(z.cell(i, midx, midy), -1);
                if (mid != p3) emit(min(mid, p3), mid < p3 ? p3 - mid : mid - p3);
                if (p3 != p4) {
                    int tx = int(p3 / n % n), ty = int(p3 % n);
                    if (!z.vert(int(p3 / nn))) {
                        const int s = tx;
                        tx = ty;
                        ty = s;
                    }
                    for (int i = min(mid_layer, last_layer); i < max(mid_layer, last_layer); i++)
                        emit(z.cell(i, tx, ty), -1);
                }
                if (p4 != last) emit(min(p4, last), p4 < last ? last - p4 : p4 - last);
            } else {
                emit(min(entry, last), entry < last ? last - entry : entry - last);
            }
            map[5 * P + child] = last_layer;
        }
    }
}

}  // namespace

__global__ __launch_bounds__(kWave) void zshape_solve_kernel(ZshapeArgs a) {
    const int k = blockIdx.x;
    if (k >= a.count) return;
    const int lane = threadIdx.x;
    const Z z(a);
    const int* pts = a.points + a.points[k];
    int* routes = a.raw + pts[0];
    const int cap = a.raw_cap[k];
    if (lane == 0) routes[0] = 1;
    __syncthreads();
    const int node_cnt = pts[1];
    pts += 2;
    for (int t = 6 * (node_cnt - 1); t >= 0; t -= 6) {
        const int tox = pts[t] / a.n, toy = pts[t] % a.n;
        for (int j = 1; j <= 3; j++) {
            if (pts[t + 2 + j] == -1) continue;
            const int child = pts[pts[t + 2 + j]];
            zedge(z, lane, j, child / a.n, child % a.n, tox, toy);
        }
        if (lane == 0) znode(z, pts + t);
        __syncthreads();
    }
    if (lane == 0) ztrace(z, pts, node_cnt, routes, cap);
}

__global__ __launch_bounds__(256) void zshape_reset_kernel(ZshapeArgs a) {
    const int k = blockIdx.x;
    if (k >= a.count) return;
    const Z z(a);
    const int* rec = a.points + a.points[k];
    const int nodes = rec[1];
    rec += 2;
    for (int i = threadIdx.x; i < nodes * a.layers; i += blockDim.x) {
        const int layer = i % a.layers, point = rec[(i / a.layers) * 6];
        const long long index = z.cell(layer, point / a.n, point % a.n);
        for (int plane = 0; plane < 4; ++plane) a.map[plane * z.P + index] = kZInf;
    }
}

// Raw words [map start, length or -1 for a via] into route words
// (src/routing/zshape_topology.hpp:233-289).
__global__ __launch_bounds__(256) void zshape_convert_kernel(ZshapeArgs a, const int* net_ids, RouteStore rs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.count) return;
    const DeviceGrid& g = a.g;
    const int n = a.n, X = g.X, Y = g.Y, L = g.L, XY = g.XY;
    const int id = net_ids[k];
    const int* src = a.raw + a.points[a.points[k]];
    int* dst = rs.routes + (size_t)rs.pin_acc_num[id] * rs.per_pin;
    const int cap = rs.per_pin * (rs.pin_acc_num[id + 1] - rs.pin_acc_num[id]);
    int out = 1;
    for (int p = 1; p + 1 < src[0]; p += 2) {
        const int start = src[p], len = src[p + 1];
        const int layer = start / (n * n) - 1;
        const int u = start / n % n, v = start % n;
        const bool vert = layer >= 0 && layer_dir(g, layer);
        const int x = vert ? u : v, y = vert ? v : u;
        if (layer < 0 || layer >= L || x >= X || y >= Y) {
            atomicExch(a.error, (int)kDevBadCoordinate);
            continue;
        }
        const int pos0 = (layer * X + x) * Y + y;
        int pos1 = pos0;
        if (len == -1) {
            if (layer + 1 >= L) {
                atomicExch(a.error, (int)kDevBadCoordinate);
                continue;
            }
            pos1 += XY;
        } else {
            const int end = start + len, eu = end / n % n, ev = end % n;
            const int ex = vert ? eu : ev, ey = vert ? ev : eu;
            if (ex >= X || ey >= Y) {
                atomicExch(a.error, (int)kDevBadCoordinate);
                continue;
            }
            pos1 = (layer * X + ex) * Y + ey;
        }
        if (out + 1 < cap) {
            dst[out++] = pos0;
            dst[out++] = pos1;
        } else {
            atomicExch(a.error, (int)kDevRouteOverflow);
        }
    }
    if (src[0] <= 0) atomicExch(a.error, (int)kDevNoSolution);
    dst[0] = out;
}

}  // namespace grotor
