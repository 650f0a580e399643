/*
 * CPU oracle for the GRotor routing hot path -- TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; never part of the product path.  See oracle.h.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fPIC -shared
 * (oracle/Makefile).  Each function cites the reference code it restates.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_INF 1e22 /* include/router_state.hpp:271 */
#define IDX3(g, l, x, y) (((l) * (g)->X + (x)) * (g)->Y + (y))

/* ------------------------------------------------------------------------ */
/* Project exponential.                                                      */
/* Cody-Waite reduction x = n*ln2 + r (ln2 split into a 32-bit-exact high    */
/* part and a low part), exp(r) by its Taylor series to degree 10 in Horner  */
/* form, scaling by 2^n, all in binary64 with no contraction, then one       */
/* rounding to binary32.  Results below FLT_MIN are flushed to +0 so that no */
/* platform's denormal mode can change a result.                            */
/* ------------------------------------------------------------------------ */
float ora_expf(float xf) {
    if (xf != xf) return xf;
    if (xf > 89.0f) return INFINITY;
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

/* src/cost/full_update.cuh:9-14 (of_cost_scaled, of_cost_scale = 1).  The
 * float parameters mirror the reference signature: the demand argument has
 * already been narrowed to binary32 by the caller. */
double ora_of_cost(float capacity, float demand, double slope, double zero_slope) {
    if ((double)capacity > 0.001) {
        const float diff = demand - capacity;
        const double a = slope * (double)diff;
        const double b = 1.0 * slope * (double)diff;
        return (double)ora_expf((float)fmin(a, b));
    }
    if (demand > 0.0f) {
        const double a = zero_slope * (double)demand;
        const double b = 1.0 * zero_slope * (double)demand;
        return (double)ora_expf((float)fmin(a, b));
    }
    return 0.0;
}

/* src/cost/full_update.cuh:23-32, incre_base = cross_points = 0
 * (src/preprocess.hpp:1757 clears cross_points). */
double ora_incremental_of_cost(float capacity, float demand, double incre_demand,
                               double short_cost, double slope, double zero_slope) {
    const float d1 = (float)((double)demand + incre_demand + 0.0);
    const float d0 = (float)((double)demand + 0.0);
    const double v = (ora_of_cost(capacity, d1, slope, zero_slope) -
                      ora_of_cost(capacity, d0, slope, zero_slope)) * short_cost;
    return fmin(1e12, v);
}

/* Fenwick tree over the shifted leaves [0, w0, w1, ...]: odd indices are
 * stored nodes, even indices read the leaf.  src/cost/fenwick.cuh:129-168. */
static double fen_val(const double *tree, const float *w, int j) {
    return (j & 1) ? tree[j >> 1] : (j ? (double)w[j - 1] : 0.0);
}

void ora_fenwick_track(const float *w, int len, double *presum) {
    double *tree = (double *)calloc((size_t)len / 2 + 1, sizeof(double));
    for (int width = 1; width <= len; width <<= 1) {
        for (int i = width - 1; i < len; i += 2 * width) {
            double value = i ? (double)w[i - 1] : 0.0;
            for (int off = 1; off < width; off <<= 1) value += fen_val(tree, w, i - off);
            if (i & 1) tree[i >> 1] = value;
        }
    }
    for (int i = 0; i < len; ++i) {
        double sum = 0.0;
        for (int j = i; j >= 0; j = (j & (j + 1)) - 1) sum += fen_val(tree, w, j);
        presum[i] = sum;
    }
    free(tree);
}

/* src/cost/prefix_scan.cuh:74-114: Hillis-Steele over sum[0] = 0,
 * sum[i+1] = w[i]; in round d every index with bit d set adds the last
 * element of the preceding 2^d block. */
void ora_dense_prefix_track(const float *w, int len, double *presum) {
    double *s = (double *)malloc(sizeof(double) * (size_t)len);
    s[0] = 0.0;
    for (int i = 0; i + 1 < len; ++i) s[i + 1] = (double)w[i];
    for (int d = 0; (1 << d) < len; ++d)
        for (int i = len - 1; i >= 0; --i)
            if ((i >> d) & 1) s[i] += s[((i >> d) << d) - 1];
    memcpy(presum, s, sizeof(double) * (size_t)len);
    free(s);
}

/* One routing track: wcost/vcost of every cell (src/cost/full_update.cuh:34-83)
 * and the Fenwick prefix of the track (src/cost/fenwick.cuh:129-168).  Tracks
 * are independent: a cell's costs read only itself and its predecessor on
 * the track, and presum reads only the track's wcost. */
/* Z-shape (GGR) builds round each wire cost to an integer in [0, 1e7]
 * (src/cost/full_update.cuh:47-50). */
static int integer_wire = 0;
void ora_set_integer_wire(int on) { integer_wire = on; }

static void refresh_track(const ora_grid *g, const float *capacity, const float *demand,
                          double slope, double zero_slope, double via_weight, int l, int t,
                          float *w, float *vcost, double *presum, float *tw, double *tp) {
    const int X = g->X, Y = g->Y;
    const int dir = (l & 1) ^ g->DIR;
    const int len = dir ? Y : X;
    const double sc = g->unit_length_short_costs[l];
    for (int p = 0; p < len; ++p) {
        const int x = dir ? t : p, y = dir ? p : t;
        const size_t i = (size_t)IDX3(g, l, x, y);
        float wc = (float)ora_incremental_of_cost(capacity[i], demand[i], 1.0, sc, slope, zero_slope);
        if (dir == 0 && x < X - 1)
            wc = (float)((double)wc + g->unit_length_wire_cost * (double)g->x_edge_len[x]);
        if (dir == 1 && y < Y - 1)
            wc = (float)((double)wc + g->unit_length_wire_cost * (double)g->y_edge_len[y]);
        if (integer_wire) wc = (float)(long long)fminf(fmaxf(wc, 0.f), 1.e7f);
        if (w) w[i] = wc;
        tw[p] = wc;
        if (vcost) {
            const size_t q = (size_t)(p > 0 ? (dir ? IDX3(g, l, x, y - 1) : IDX3(g, l, x - 1, y)) : (int)i);
            double v;
            if (p == 0)
                v = ora_incremental_of_cost(capacity[i], demand[i], 1.0, sc, slope, zero_slope);
            else if (p == len - 1)
                v = ora_incremental_of_cost(capacity[q], demand[q], 1.0, sc, slope, zero_slope);
            else
                v = ora_incremental_of_cost(capacity[i], demand[i], 0.5, sc, slope, zero_slope) +
                    ora_incremental_of_cost(capacity[q], demand[q], 0.5, sc, slope, zero_slope);
            float vc = (float)v;
            vc = (float)((double)vc + via_weight * g->unit_via_cost);
            vcost[i] = vc;
        }
    }
    if (presum) {
        ora_fenwick_track(tw, len, tp);
        for (int p = 0; p < len; ++p) presum[dir ? IDX3(g, l, t, p) : IDX3(g, l, p, t)] = tp[p];
    }
}

static int track_count(const ora_grid *g, int l) { return ((l & 1) ^ g->DIR) ? g->X : g->Y; }

/* Refresh the tracks whose dirty flag is set (all when dirty is NULL) and
 * clear the flags: graph::refresh_cost (src/cost/incremental_update.cuh:98-132)
 * at track granularity, which gives the same values as its per-cell form. */
static void refresh_dirty(const ora_grid *g, const float *capacity, const float *demand,
                          double slope, double zero_slope, double via_weight, unsigned char *dirty,
                          float *wcost, float *vcost, double *presum, int threads) {
    int total = 0;
    for (int l = 0; l < g->L; ++l) total += track_count(g, l);
    const int maxlen = g->X > g->Y ? g->X : g->Y;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
    {
        float *tw = (float *)malloc(sizeof(float) * (size_t)maxlen);
        double *tp = (double *)malloc(sizeof(double) * (size_t)maxlen);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int T = 0; T < total; ++T) {
            if (dirty && !dirty[T]) continue;
            int l = 0, t = T;
            while (t >= track_count(g, l)) t -= track_count(g, l++);
            refresh_track(g, capacity, demand, slope, zero_slope, via_weight, l, t, wcost, vcost, presum, tw, tp);
            if (dirty) dirty[T] = 0;
        }
        free(tw);
        free(tp);
    }
}

void ora_cost_refresh_slopes(const ora_grid *g, const float *capacity, const float *demand,
                             double slope, double zero_slope, double via_weight,
                             float *wcost, float *vcost, double *presum) {
    refresh_dirty(g, capacity, demand, slope, zero_slope, via_weight, NULL, wcost, vcost, presum, 1);
}

void ora_cost_refresh(const ora_grid *g, const float *capacity, const float *demand,
                      float *wcost, float *vcost, double *presum) {
    ora_cost_refresh_slopes(g, capacity, demand, 0.5, 1.5, 1.0, wcost, vcost, presum);
}

/* ------------------------------------------------------------------------ */
/* Batching: src/batching/generate_batches_cpu.hpp:7-73                     */
/* ------------------------------------------------------------------------ */
typedef struct bitgrid { uint64_t *bits; } bitgrid;

int ora_batches(int X, int Y, int n,
                const int *pt_off, const int *pts,
                const int *hs_off, const int *hsegs,
                const int *vs_off, const int *vsegs,
                int limit, int sparsity, int *color) {
    const size_t words = ((size_t)X * Y + 63) / 64;
    int nb = 0, cap = 16;
    bitgrid *vis = (bitgrid *)malloc(sizeof(bitgrid) * (size_t)cap);
    int *size = (int *)malloc(sizeof(int) * (size_t)cap);
#define TEST(b, p) ((vis[b].bits[(size_t)(p) >> 6] >> ((p) & 63)) & 1u)
#define SET(b, p) (vis[b].bits[(size_t)(p) >> 6] |= (uint64_t)1 << ((p) & 63))
    for (int k = 0; k < n; ++k) {
        int bid = -1;
        for (int b = 0; b < nb; ++b) {
            if (size[b] >= limit) continue;
            int conflict = 0;
            for (int q = pt_off[k]; q < pt_off[k + 1]; ++q)
                if (TEST(b, pts[q])) { conflict = 1; break; }
            if (!conflict) { bid = b; break; }
        }
        if (bid < 0) {
            if (nb == cap) {
                cap *= 2;
                vis = (bitgrid *)realloc(vis, sizeof(bitgrid) * (size_t)cap);
                size = (int *)realloc(size, sizeof(int) * (size_t)cap);
            }
            vis[nb].bits = (uint64_t *)calloc(words, sizeof(uint64_t));
            size[nb] = 0;
            bid = nb++;
        }
        size[bid]++;
        color[k] = bid;
        for (int s = hs_off[k]; s < hs_off[k + 1]; ++s)
            for (int p = hsegs[2 * s]; p <= hsegs[2 * s + 1]; p += sparsity * Y) SET(bid, p);
        for (int s = vs_off[k]; s < vs_off[k + 1]; ++s)
            for (int p = vsegs[2 * s]; p <= vsegs[2 * s + 1]; p += sparsity) SET(bid, p);
        for (int q = pt_off[k]; q < pt_off[k + 1]; ++q) {
            /* mark_3x3: x in [_x-1, _x+1], y in [_y-1, _y+1) with x == _x || y == _y */
            const int px = pts[q] / Y, py = pts[q] % Y;
            for (int x = px - 1; x <= px + 1; ++x) {
                if (x < 0 || x >= X) continue;
                for (int y = py - 1; y < py + 1; ++y) {
                    if (y < 0 || y >= Y) continue;
                    if (x == px || y == py) SET(bid, x * Y + y);
                }
            }
        }
    }
#undef TEST
#undef SET
    for (int b = 0; b < nb; ++b) free(vis[b].bits);
    free(vis);
    free(size);
    return nb;
}

/* ------------------------------------------------------------------------ */
/* L-shape DP: src/routing/lshape_thread.cuh:18-171 (scalar semantics) with */
/* first-minimum ties and the `choice = 0` default of                       */
/* src/routing/lshape_warp.cuh:8-49, 193-199.                               */
/* ------------------------------------------------------------------------ */
static double wire_cost(const ora_grid *g, const double *presum, int l, int xmin, int xmax, int ymin, int ymax) {
    /* src/cost/prefix_scan.cuh:10-40 (default branch :38) */
    return presum[IDX3(g, l, xmax, ymax)] - presum[IDX3(g, l, xmin, ymin)];
}

int ora_lshape_net_cost(const ora_grid *g, const float *vcost, const double *presum,
                        int n, const int *nodes, const int *parents, int *routes, int cap, double *best);

int ora_lshape_net(const ora_grid *g, const float *vcost, const double *presum,
                   int n, const int *nodes, const int *parents, int *routes, int cap) {
    return ora_lshape_net_cost(g, vcost, presum, n, nodes, parents, routes, cap, NULL);
}

int ora_lshape_net_cost(const ora_grid *g, const float *vcost, const double *presum,
                        int n, const int *nodes, const int *parents, int *routes, int cap, double *best) {
    const int L = g->L, X = g->X, Y = g->Y, DIR = g->DIR, LL = L * L;
    double *dist = (double *)malloc(sizeof(double) * (size_t)n * LL);
    int *from = (int *)calloc((size_t)n * LL, sizeof(int));
    int *range = (int *)malloc(sizeof(int) * (size_t)n);
    double diag[64];
    int diag_prev[64];
    /* via stacks (lshape_thread.cuh:39-49) */
    for (int i = 0; i < n; ++i) {
        const int fixed = nodes[i] / X / Y, x = nodes[i] / Y % X, y = nodes[i] % Y;
        double *nd = dist + (size_t)i * LL;
        for (int lo = 0; lo < L; ++lo) {
            nd[lo * L + lo] = 0.0;
            for (int hi = lo + 1; hi < L; ++hi)
                nd[lo * L + hi] = nd[lo * L + hi - 1] + (double)vcost[IDX3(g, hi - 1, x, y)];
            for (int hi = lo; hi < L; ++hi)
                if (fixed < L && (lo > fixed || hi < fixed)) nd[lo * L + hi] = ORA_INF;
        }
    }
    for (int i = n - 1; i >= 1; --i) {
        const int p = parents[i];
        const int x = nodes[i] / Y % X, y = nodes[i] % Y;
        const int px = nodes[p] / Y % X, py = nodes[p] % Y;
        const int minx = x < px ? x : px, maxx = x < px ? px : x;
        const int miny = y < py ? y : py, maxy = y < py ? py : y;
        const double *nd = dist + (size_t)i * LL;
        double cur[64];
        int choice[64];
        for (int c = 0; c < L; ++c) {
            cur[c] = ORA_INF;
            choice[c] = 0;
            for (int lo = 0; lo <= c; ++lo)
                for (int hi = c; hi < L; ++hi)
                    if (nd[lo * L + hi] < cur[c]) { cur[c] = nd[lo * L + hi]; choice[c] = lo * L + hi; }
            diag[c] = ORA_INF;
            diag_prev[c] = 0;
        }
        if (x == px || y == py) {
            for (int c = 0; c < L; ++c) {
                const int d = (c & 1) ^ DIR;
                if (d == 0 && y != py) continue;
                if (d == 1 && x != px) continue;
                diag[c] = cur[c] + wire_cost(g, presum, c, minx, maxx, miny, maxy);
                diag_prev[c] = choice[c] * LL + c * L + c;
            }
        } else {
            for (int cl = 0; cl < L; ++cl) {
                double cost = cur[cl];
                const int dcl = (cl & 1) ^ DIR;
                if (dcl) cost += wire_cost(g, presum, cl, x, x, miny, maxy);
                else cost += wire_cost(g, presum, cl, minx, maxx, y, y);
                for (int pl = (cl & 1) ^ 1; pl < L; pl += 2) {
                    double cost2 = 0.0;
                    const int lo = cl < pl ? cl : pl, hi = cl < pl ? pl : cl;
                    for (int l = lo + 1; l < hi; ++l)
                        cost2 += dcl ? (double)vcost[IDX3(g, l, x, py)] : (double)vcost[IDX3(g, l, px, y)];
                    if ((pl & 1) ^ DIR) cost2 += wire_cost(g, presum, pl, px, px, miny, maxy);
                    else cost2 += wire_cost(g, presum, pl, minx, maxx, py, py);
                    if (cost + cost2 < diag[pl]) {
                        diag[pl] = cost + cost2;
                        diag_prev[pl] = choice[cl] * LL + cl * L + pl;
                    }
                }
            }
        }
        /* interval closure and parent accumulation (lshape_thread.cuh:299-313,
         * lshape_warp.cuh:165-189): first minimum over [lo..hi] */
        double *pd = dist + (size_t)p * LL;
        int *fr = from + (size_t)i * LL;
        for (int lo = 0; lo < L; ++lo) {
            double c = diag[lo];
            int path = diag_prev[lo];
            for (int hi = lo; hi < L; ++hi) {
                if (hi > lo && !(c <= diag[hi])) { c = diag[hi]; path = diag_prev[hi]; }
                fr[lo * L + hi] = path;
                pd[lo * L + hi] += c;
            }
        }
    }
    int choice = 0;
    for (int lo = 0; lo < L; ++lo)
        for (int hi = lo; hi < L; ++hi)
            if (dist[lo * L + hi] < dist[choice]) choice = lo * L + hi;
    range[0] = choice;
    if (best) *best = dist[choice];
    int status = 0;
    for (int i = 1; i < n; ++i) range[i] = from[(size_t)i * LL + range[parents[i]]] / LL;
    for (int i = 0; i < n; ++i)
        if (!(dist[(size_t)i * LL + range[i]] < ORA_INF)) status = -2;
    /* emission (lshape_thread.cuh:316-369, lshape_warp.cuh:51-100) */
    int cnt = 1;
#define EMIT(v) do { if (cnt >= cap) { status = -1; goto done; } routes[cnt++] = (v); } while (0)
    for (int i = 0; i < n; ++i) {
        const int lo = range[i] / L, hi = range[i] % L;
        const int x = nodes[i] / Y % X, y = nodes[i] % Y;
        if (lo < hi) { EMIT(IDX3(g, lo, x, y)); EMIT(IDX3(g, hi, x, y)); }
        if (i == 0) continue;
        const int p = parents[i];
        const int px = nodes[p] / Y % X, py = nodes[p] % Y;
        const int minx = x < px ? x : px, maxx = x < px ? px : x;
        const int miny = y < py ? y : py, maxy = y < py ? py : y;
        const int path = from[(size_t)i * LL + range[p]];
        const int cl = path / L % L, pl = path % L;
        if (px == x || py == y) {
            EMIT(IDX3(g, cl, minx, miny));
            EMIT(IDX3(g, cl, maxx, maxy));
        } else if ((cl & 1) ^ DIR) {
            EMIT(IDX3(g, cl, x, y)); EMIT(IDX3(g, cl, x, py));
            EMIT(IDX3(g, cl, x, py)); EMIT(IDX3(g, pl, x, py));
            EMIT(IDX3(g, pl, x, py)); EMIT(IDX3(g, pl, px, py));
        } else {
            EMIT(IDX3(g, cl, x, y)); EMIT(IDX3(g, cl, px, y));
            EMIT(IDX3(g, cl, px, y)); EMIT(IDX3(g, pl, px, y));
            EMIT(IDX3(g, pl, px, y)); EMIT(IDX3(g, pl, px, py));
        }
    }
#undef EMIT
    routes[0] = cnt;
done:
    free(dist);
    free(from);
    free(range);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Demand commit: src/routing/demand.cuh:236-319                            */
/* ------------------------------------------------------------------------ */
static int track_of(const ora_grid *g, int l, int x, int y) {
    int T = 0;
    for (int k = 0; k < l; ++k) T += track_count(g, k);
    return T + (((l & 1) ^ g->DIR) ? x : y);
}

void ora_commit_batch(const ora_grid *g, float *demand, int *timestamp,
                      int count, int *const *routes_of, int stamp, int K,
                      double *total_wirelength, long long *total_vias, unsigned char *dirty) {
    const int X = g->X, Y = g->Y, XY = X * Y;
    for (int k = 0; k < count; ++k) { /* commit_batch_wires */
        const int *r = routes_of[k];
        double length = 0.0;
        for (int i = 1; i < r[0]; i += 2) {
            const int a = r[i], b = r[i + 1], layer = a / XY;
            if (layer != b / XY) continue;
            const int x0 = a / Y % X, y0 = a % Y, x1 = b / Y % X, y1 = b % Y;
            const int along_y = x0 == x1;
            const int lo = along_y ? (y0 < y1 ? y0 : y1) : (x0 < x1 ? x0 : x1);
            const int hi = along_y ? (y0 < y1 ? y1 : y0) : (x0 < x1 ? x1 : x0);
            timestamp[a] = stamp;
            timestamp[b] = stamp;
            for (int pos = lo; pos < hi; ++pos) {
                const int idx = layer * XY + (along_y ? x0 * Y + pos : pos * Y + y0);
                demand[idx] += (float)K;
                if (dirty) dirty[track_of(g, layer, idx / Y % X, idx % Y)] = 1;
                timestamp[idx] = stamp;
                length += along_y ? (double)g->y_edge_len[pos] : (double)g->x_edge_len[pos];
            }
        }
        *total_wirelength += length * K;
    }
    for (int k = 0; k < count; ++k) { /* commit_batch_vias */
        const int *r = routes_of[k];
        long long vias = 0;
        for (int i = 1; i < r[0]; i += 2) {
            const int a = r[i], b = r[i + 1];
            const int la = a / XY, lb = b / XY;
            const int low = la < lb ? la : lb, high = la < lb ? lb : la;
            if (low == high) continue;
            const int x = a / Y % X, y = a % Y;
            vias += high - low;
            for (int layer = low; layer < high; ++layer) {
                const int idx = layer * XY + x * Y + y;
                if (timestamp[idx] >= stamp) continue;
                const int along_y = (layer & 1) ^ g->DIR;
                const int position = along_y ? y : x, length = along_y ? Y : X;
                const int stride = along_y ? 1 : Y;
                if (dirty) dirty[track_of(g, layer, x, y)] = 1;
                if (!position) demand[idx] += (float)K;
                else if (position == length - 1) demand[idx - stride] += (float)K;
                else {
                    demand[idx - stride] += .5f * K;
                    demand[idx] += .5f * K;
                }
            }
        }
        *total_vias += vias * K;
    }
}

/* ------------------------------------------------------------------------ */
/* Rip-up classification: src/routing/congestion.cuh:8-58, 114-136          */
/* ------------------------------------------------------------------------ */
void ora_ripup(const ora_grid *g, const float *capacity, const float *demand,
               int threshold, int net_count, int *const *routes_of,
               unsigned char *is_of_net) {
    const int L = g->L, X = g->X, Y = g->Y, XY = X * Y;
    int *sum = (int *)calloc((size_t)L * XY, sizeof(int));
    for (int l = 0; l < L; ++l) {
        const int dir = (l & 1) ^ g->DIR;
        const int len = dir ? Y : X, tracks = dir ? X : Y;
        for (int t = 0; t < tracks; ++t) {
            /* mark_overflow_edges then compute_presum_general: exclusive prefix
             * of the overflow flags of edges 0..len-2 along the track */
            int acc = 0;
            for (int p = 0; p < len; ++p) {
                const int idx = dir ? IDX3(g, l, t, p) : IDX3(g, l, p, t);
                const int flag = (p + 1 < len) && (capacity[idx] + threshold <= demand[idx]);
                sum[idx] = acc;
                acc += flag;
            }
        }
    }
    for (int k = 0; k < net_count; ++k) { /* mark_overflow_nets */
        const int *r = routes_of[k];
        int of = 0;
        is_of_net[k] = 0;
        for (int i = 1; i < r[0]; i += 2) {
            const int l = r[i] / XY, x0 = r[i] / Y % X, y0 = r[i] % Y;
            const int x1 = r[i + 1] / Y % X, y1 = r[i + 1] % Y;
            if (x0 != x1) of += sum[IDX3(g, l, x0 > x1 ? x0 : x1, y0)] - sum[IDX3(g, l, x0 < x1 ? x0 : x1, y0)];
            if (y0 != y1) of += sum[IDX3(g, l, x0, y0 > y1 ? y0 : y1)] - sum[IDX3(g, l, x0, y0 < y1 ? y0 : y1)];
            if (of > 10) { is_of_net[k] = 1; break; }
        }
    }
    free(sum);
}

/* ------------------------------------------------------------------------ */
/* L-shape stage loop: src/routing/lshape.cuh:153-179                       */
/* ------------------------------------------------------------------------ */
int ora_lshape_stage(const ora_grid *g, const float *capacity, float *demand,
                     int *timestamp, float *wcost, float *vcost, double *presum,
                     int nbatch, const int *batch_off, const int *net_ids,
                     const int *node_off, const int *nodes, const int *parents,
                     const long long *route_off, const int *route_cap, int *routes,
                     int *stamp, double *total_wirelength, long long *total_vias,
                     int threads) {
    int status = 0;
    int maxcount = 0;
    for (int b = 0; b < nbatch; ++b)
        if (batch_off[b + 1] - batch_off[b] > maxcount) maxcount = batch_off[b + 1] - batch_off[b];
    int **rof = (int **)malloc(sizeof(int *) * (size_t)(maxcount > 0 ? maxcount : 1));
    int tracks = 0;
    for (int l = 0; l < g->L; ++l) tracks += track_count(g, l);
    unsigned char *dirty = (unsigned char *)malloc((size_t)tracks);
    memset(dirty, 1, (size_t)tracks);
    for (int b = 0; b < nbatch && !status; ++b) {
        const int begin = batch_off[b], count = batch_off[b + 1] - begin;
        const int s = ++*stamp;
        refresh_dirty(g, capacity, demand, 0.5, 1.5, 1.0, dirty, wcost, vcost, presum, threads);
        int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1) reduction(| : bad)
#endif
        for (int k = 0; k < count; ++k) {
            const int id = net_ids[begin + k];
            const int e = begin + k;
            const int rc = ora_lshape_net(g, vcost, presum, node_off[e + 1] - node_off[e],
                                          nodes + node_off[e], parents + node_off[e],
                                          routes + route_off[id], route_cap[id]);
            if (rc) bad |= 1;
        }
        if (bad) { status = -1; break; }
        for (int k = 0; k < count; ++k) rof[k] = routes + route_off[net_ids[begin + k]];
        ora_commit_batch(g, demand, timestamp, count, rof, s, 1, total_wirelength, total_vias, dirty);
    }
    free(dirty);
    free(rof);
    return status;
}

/* One Z-shape stage (src/pipeline_ggr.hpp:114-127): per batch, refresh the
 * tracks whose demand changed, reset the batch's tree-node cells, solve,
 * convert the raw words into route words, commit wires then vias.  records
 * holds each batch's encoded nets from rec_off[b]; raw_cap[e] is the raw
 * word capacity of the net at execution position e.  Returns 0 or the
 * convert error code. */
int ora_zshape_stage(const ora_grid *g, const float *capacity, float *demand, int *timestamp, float *wcost,
                     float *vcost, double *presum, int nbatch, const int *batch_off, const int *net_ids,
                     const long long *rec_off, const int *records, const int *raw_cap,
                     const long long *route_off, const int *route_cap, int *routes, int *stamp,
                     double *total_wirelength, long long *total_vias, int threads) {
    const int n = g->X > g->Y ? g->X : g->Y;
    const size_t plane = (size_t)(g->L + 1) * n * n;
    int *map = (int *)calloc(6 * plane, sizeof(int));
    int *prev = (int *)calloc(16 * plane, sizeof(int));
    long long maxraw = 1;
    int maxcount = 1;
    for (int b = 0; b < nbatch; ++b) {
        long long s = 0;
        for (int e = batch_off[b]; e < batch_off[b + 1]; ++e) s += raw_cap[e];
        if (s > maxraw) maxraw = s;
        if (batch_off[b + 1] - batch_off[b] > maxcount) maxcount = batch_off[b + 1] - batch_off[b];
    }
    int *raw = (int *)calloc((size_t)maxraw, sizeof(int));
    int *raw_off = (int *)malloc(sizeof(int) * (size_t)maxcount);
    int **rof = (int **)malloc(sizeof(int *) * (size_t)maxcount);
    int tracks = 0;
    for (int l = 0; l < g->L; ++l) tracks += track_count(g, l);
    unsigned char *dirty = (unsigned char *)malloc((size_t)tracks);
    memset(dirty, 1, (size_t)tracks);
    int status = 0;
    for (int b = 0; b < nbatch; ++b) {
        const int begin = batch_off[b], count = batch_off[b + 1] - begin;
        const int *pts = records + rec_off[b];
        const int s = ++*stamp;
        refresh_dirty(g, capacity, demand, 0.5, 1.5, 1.0, dirty, wcost, vcost, presum, threads);
        ora_zshape_reset(g, map, pts, count);
        ora_zshape_solve(g, presum, vcost, map, prev, pts, count, raw, NULL);
        for (int k = 0; k < count; ++k) raw_off[k] = pts[pts[k]];
        const int err = ora_zshape_convert(g, raw, raw_off, net_ids + begin, count, route_off, route_cap, routes);
        if (err) status = err;
        for (int k = 0; k < count; ++k) rof[k] = routes + route_off[net_ids[begin + k]];
        ora_commit_batch(g, demand, timestamp, count, rof, s, 1, total_wirelength, total_vias, dirty);
    }
    free(map);
    free(prev);
    free(raw);
    free(raw_off);
    free(rof);
    free(dirty);
    return status;
}

int ora_flt_stage(const ora_grid *g, const float *capacity, float *demand, int *timestamp, float *wcost,
                  float *vcost, double *presum, int nbatch, const int *batch_off, const int *net_ids,
                  const long long *node_off, const int *node_rec, const long long *edge_off, const int *edge_rec,
                  const long long *route_off1, int *routes1, const long long *route_off2, const int *route_cap2,
                  int *routes2, int remove, const unsigned char *live2, int *stamp, double *total_wirelength,
                  long long *total_vias, int threads) {
    int maxcount = 1;
    for (int b = 0; b < nbatch; ++b)
        if (batch_off[b + 1] - batch_off[b] > maxcount) maxcount = batch_off[b + 1] - batch_off[b];
    int **rof = (int **)malloc(sizeof(int *) * (size_t)maxcount);
    int tracks = 0;
    for (int l = 0; l < g->L; ++l) tracks += track_count(g, l);
    unsigned char *dirty = (unsigned char *)malloc((size_t)tracks);
    memset(dirty, 1, (size_t)tracks);
    int status = 0;
    for (int b = 0; b < nbatch; ++b) {
        const int begin = batch_off[b], count = batch_off[b + 1] - begin;
        if (remove) {
            /* remove 1: the phase-one routes (route_prepared_topology's
             * remove_phase_one, src/routing/flt.cuh:145-152); remove 2: each
             * net's current route, its phase-two words where live2 marks them
             * (a later rip-up stage of the same nets) */
            for (int k = 0; k < count; ++k) {
                const int id = net_ids[begin + k];
                rof[k] = remove == 2 && live2 && live2[id] ? routes2 + route_off2[id] : routes1 + route_off1[id];
            }
            const int s1 = ++*stamp;
            ora_commit_batch(g, demand, timestamp, count, rof, s1, -1, total_wirelength, total_vias, dirty);
        }
        const int s = ++*stamp;
        refresh_dirty(g, capacity, demand, 0.5, 1.5, 1.0, dirty, wcost, vcost, presum, threads);
        const int err = ora_flt_batch(g, presum, vcost, count, net_ids + begin, node_off + begin, node_rec,
                                      edge_off + begin, edge_rec, route_off2, route_cap2, routes2, NULL);
        if (err) status = err;
        for (int k = 0; k < count; ++k) rof[k] = routes2 + route_off2[net_ids[begin + k]];
        ora_commit_batch(g, demand, timestamp, count, rof, s, 1, total_wirelength, total_vias, dirty);
    }
    free(rof);
    free(dirty);
    return status;
}

/* extract_congestionView, _xsum, _ysum (src/routing/congestion.cuh:60-112) */
static void view_prefix(float *s, int len) {
    for (int d = 0; (1 << d) < len; ++d)
        for (int i = 0; i < len; ++i)
            if (i >> d & 1) s[i] += s[(i >> d << d) - 1];
}

void ora_congestion_view(const ora_grid *g, const float *capacity, const float *demand, unsigned char *congested,
                         float *xsum, float *ysum) {
    const int X = g->X, Y = g->Y, XY = X * Y;
    for (int idx = 0; idx < XY; ++idx) {
        double worst = 0.0;
        for (int l = 0; l < g->L; ++l) {
            const double d = (double)(demand[(size_t)l * XY + idx] - capacity[(size_t)l * XY + idx]);
            worst = fmax(worst, d);
        }
        if (congested) congested[idx] = worst > 0.0;
        xsum[idx] = (float)worst;
        ysum[idx] = (float)worst;
    }
    float *s = (float *)malloc(sizeof(float) * (size_t)(X > Y ? X : Y));
    for (int x = 0; x < X; ++x) {
        s[0] = 0.f;
        for (int y = 0; y + 1 < Y; ++y) s[y + 1] = ysum[x * Y + y];
        view_prefix(s, Y);
        for (int y = 0; y < Y; ++y) ysum[x * Y + y] = s[y];
    }
    for (int y = 0; y < Y; ++y) {
        s[0] = 0.f;
        for (int x = 0; x + 1 < X; ++x) s[x + 1] = xsum[x * Y + y];
        view_prefix(s, X);
        for (int x = 0; x < X; ++x) xsum[x * Y + y] = s[x];
    }
    free(s);
}
