/*
 * FLT (routing-DAG) DP -- TEST INFRASTRUCTURE ONLY (the checker for the GPU
 * FLT kernels; see oracle.h).
 *
 * Restates the batch assembly of src/routing/flt_workspace.cuh:29-85,202-224
 * (compact node/edge records -> batch arrays, exclusive scan of child slots,
 * stable order by (net, depth)), precompute_cost_edge
 * (src/routing/flt_cost.cuh:50-130), solve_flt_node / trace_flt_node
 * (src/routing/flt_block.cuh:6-331) and the per-DAG depth sweep of
 * route_flt_block (:334-351) as sequential C.  The reference fixes ten metal
 * layers (MAX_LAYER = 10: nine routing layers, 9 x 9 layer-pair tables); the
 * loops here run over the grid's L routing layers (L <= 9) with the same
 * table strides, which is the reference's arithmetic when L = 9.
 *
 * Trace appends segments in node order within a depth, depth by depth; the
 * reference appends with atomicAdd, so its order inside one net varies and
 * only the per-net segment multiset is defined.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define FLT_INF 1e22
#define FLT_ML 10       /* MAX_LAYER: layer index 1..L, 0 unused */
#define FLT_STRAIGHT 82 /* choice code of a same-layer edge */
#define IDX3(g, l, x, y) (((l) * (g)->X + (x)) * (g)->Y + (y))

/* src/routing/flt_cost.cuh:48: float 1.01e22 in the upper word, code 200000000 */
static const unsigned long long kFltEmpty = 7208259049189261824ULL;

/* src/routing/flt_cost.cuh:25-46 */
static int32_t ordered_int(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
static float ordered_float(int32_t i) {
    const int32_t b = i >= 0 ? i : i ^ 0x7FFFFFFF;
    float f;
    memcpy(&f, &b, 4);
    return f;
}
unsigned long long ora_flt_pack(float cost, int code) {
    return ((unsigned long long)(uint32_t)ordered_int(cost) << 32) | (unsigned int)code;
}
static void flt_unpack(unsigned long long v, double *cost, int *code) {
    *cost = (double)ordered_float((int32_t)(uint32_t)(v >> 32));
    *code = (int)(v & 0xFFFFFFFFULL);
}

static double wire(const ora_grid *g, const double *presum, int l, int xmin, int xmax, int ymin, int ymax) {
    return presum[IDX3(g, l, xmax, ymax)] - presum[IDX3(g, l, xmin, ymin)];
}
static int dir_of(const ora_grid *g, int l) { return (l & 1) ^ g->DIR; }

/* Batch arrays in the reference's FltDeviceView form (batch-wide indices). */
typedef struct {
    const ora_grid *g;
    const double *presum;
    const float *vcost;
    int *nodes, *par_num_sum, *par_nodes, *child_idx, *child_num_sum;
    double *cost_edges;
    int *best_change;
    double *costs;                 /* [node * 10 + layer] */
    unsigned long long *best_path; /* [slot * 10 + layer] */
    unsigned long long *packed;    /* [slot * 10 + layer] */
    int *layer_output;
    int error;
} flt_ctx;

/* precompute_cost_edge: per edge and from-layer f, the first best change
 * position (stride d / min(d, 20)) for every same-parity to-layer */
static void flt_precompute(flt_ctx *c, int edges) {
    const ora_grid *g = c->g;
    const int X = g->X, Y = g->Y, L = g->L;
    for (int e = 0; e < edges; ++e) {
        const int to = c->par_nodes[e], from = c->child_idx[e] / 10;
        const int tox = c->nodes[to] / Y % X, toy = c->nodes[to] % Y;
        const int fromx = c->nodes[from] / Y % X, fromy = c->nodes[from] % Y;
        for (int f = 0; f < L; ++f) {
            double *ce = c->cost_edges + (size_t)e * 81 + f * 9;
            int *bc = c->best_change + (size_t)e * 81 + f * 9;
            if ((dir_of(g, f) == 0 && fromy != toy) || (dir_of(g, f) == 1 && fromx != tox)) {
                for (int t = 0; t < 9; ++t) ce[t] = FLT_INF;
                continue;
            }
            const int distance = abs(tox - fromx) + abs(toy - fromy);
            if (!distance) continue;
            const int dd = distance / (distance < 20 ? distance : 20);
            const int vertical = fromx == tox;
            const int start = vertical ? (fromy < toy ? fromy : toy) : (fromx < tox ? fromx : tox);
            const int end = vertical ? (fromy < toy ? toy : fromy) : (fromx < tox ? tox : fromx);
            for (int t = 0; t < L; ++t) {
                if ((t & 1) != (f & 1)) continue;
                double best = FLT_INF;
                int pos = -1;
                for (int s = start + dd; s < end; s += dd) {
                    const int sx = vertical ? fromx : s, sy = vertical ? s : fromy;
                    double cost = 0.0;
                    for (int l = (f < t ? f : t); l < (f < t ? t : f); ++l) cost += (double)c->vcost[IDX3(g, l, sx, sy)];
                    if (vertical)
                        cost += wire(g, c->presum, f, fromx, fromx, fromy < sy ? fromy : sy, fromy < sy ? sy : fromy) +
                                wire(g, c->presum, t, tox, tox, toy < sy ? toy : sy, toy < sy ? sy : toy);
                    else
                        cost += wire(g, c->presum, f, fromx < sx ? fromx : sx, fromx < sx ? sx : fromx, fromy, fromy) +
                                wire(g, c->presum, t, tox < sx ? tox : sx, tox < sx ? sx : tox, toy, toy);
                    if (cost < best) {
                        best = cost;
                        pos = s;
                    }
                }
                ce[t] = best;
                bc[t] = pos;
            }
        }
    }
}

/* solve_flt_node (src/routing/flt_block.cuh:6-142) */
static void flt_solve(flt_ctx *c, int v) {
    const ora_grid *g = c->g;
    const int L = g->L, X = g->X, Y = g->Y;
    const int parents = c->par_num_sum[v + 1] - c->par_num_sum[v];
    const int low_fixed = 1 + c->nodes[v] / X / Y;
    const int x = c->nodes[v] / Y % X, y = c->nodes[v] % Y;
    const int high_fixed = low_fixed == L + 1 ? 0 : low_fixed;
    const int kids = c->child_num_sum[v + 1] - c->child_num_sum[v];
    unsigned long long *bp = c->best_path + (size_t)c->child_num_sum[v] * FLT_ML;
    const unsigned long long *pk = c->packed + (size_t)c->child_num_sum[v] * FLT_ML;
    double *cv = c->costs + (size_t)v * FLT_ML;
    double mins[6];
    unsigned long long paths[6];
    for (int low = 1; low <= low_fixed; ++low) {
        for (int k = 0; k < kids; ++k) mins[k] = FLT_INF;
        double via = 0.0;
        for (int li = low; li < L + 1; ++li) {
            if (li > low) via += (double)c->vcost[IDX3(g, li - 2, x, y)];
            for (int k = 0; k < kids; ++k) {
                double cost;
                int code;
                flt_unpack(pk[k * FLT_ML + li], &cost, &code);
                if (cost < mins[k]) {
                    mins[k] = cost;
                    paths[k] = (unsigned long long)code * FLT_ML + (unsigned long long)li;
                }
            }
            if (li >= high_fixed) {
                double cost = via;
                for (int k = 0; k < kids; ++k) cost += mins[k];
                if (cost < FLT_INF && cost < cv[li]) {
                    cv[li] = cost;
                    for (int k = 0; k < kids; ++k) bp[k * FLT_ML + li] = paths[k];
                }
            }
        }
        for (int li = L - 1; li >= low; --li)
            if (cv[li + 1] < cv[li]) {
                cv[li] = cv[li + 1];
                for (int k = 0; k < kids; ++k) bp[k * FLT_ML + li] = bp[k * FLT_ML + li + 1];
            }
    }
    unsigned long long cache[FLT_ML];
    for (int p = 0; p < parents; ++p) {
        const int e = c->par_num_sum[v] + p;
        const int par = c->par_nodes[e], slot = c->child_idx[e] % 10;
        unsigned long long *ppk = c->packed + (size_t)c->child_num_sum[par] * FLT_ML + (size_t)slot * FLT_ML;
        const int px = c->nodes[par] / Y % X, py = c->nodes[par] % Y;
        const int minx = x < px ? x : px, maxx = x < px ? px : x, miny = y < py ? y : py, maxy = y < py ? py : y;
        for (int t = 1; t < L + 1; ++t) cache[t] = kFltEmpty;
        for (int layer = 1; layer < L + 1; ++layer) {
            if (((layer & 1) ^ g->DIR) == 1 && y != py) continue;
            if (((layer & 1) ^ g->DIR) == 0 && x != px) continue;
            const double base = cv[layer];
            if (!(base < FLT_INF)) { /* the reference asserts here */
                c->error = 2;
                continue;
            }
            for (int t = 1 + (1 - layer % 2); t < L + 1; t += 2) {
                double cost2;
                int choice = FLT_STRAIGHT;
                if (t == layer) {
                    cost2 = base + wire(g, c->presum, layer - 1, minx, maxx, miny, maxy);
                } else if (abs(x - px) + abs(y - py) > 4) {
                    const int seq = (layer - 1) * 9 + (t - 1);
                    const double ce = c->cost_edges[(size_t)e * 81 + seq];
                    if (!(ce < FLT_INF)) continue;
                    cost2 = base + ce;
                    choice = seq;
                } else {
                    continue;
                }
                const unsigned long long nv = ora_flt_pack((float)cost2, v * 100 + choice);
                if (nv < cache[t]) cache[t] = nv;
            }
        }
        for (int t = 1; t < L + 1; ++t)
            if (cache[t] < ppk[t]) ppk[t] = cache[t];
    }
}

static void emit(int *routes, int cap, const int *w, int n, int *err) {
    if (routes[0] + n - 1 >= cap) {
        *err = 1;
        return;
    }
    for (int i = 0; i < n; ++i) routes[routes[0] + i] = w[i];
    routes[0] += n;
}

/* trace_flt_node (src/routing/flt_block.cuh:144-331) */
static void flt_trace(flt_ctx *c, int v, int *routes, int cap) {
    const ora_grid *g = c->g;
    const int L = g->L, X = g->X, Y = g->Y;
    const unsigned long long *bp = c->best_path + (size_t)c->child_num_sum[v] * FLT_ML;
    const int l = c->nodes[v] / Y / X;
    const int cx = c->nodes[v] / Y % X, cy = c->nodes[v] % Y;
    const int parents = c->par_num_sum[v + 1] - c->par_num_sum[v];
    int *lo = c->layer_output;
    if (!parents) {
        int best = 0;
        double bc = c->costs[(size_t)v * FLT_ML];
        for (int layer = 1; layer < L + 1; ++layer)
            if (c->costs[(size_t)v * FLT_ML + layer] < bc) {
                bc = c->costs[(size_t)v * FLT_ML + layer];
                best = layer;
            }
        lo[v] = best;
        routes[0] = 1;
        if (!best) { /* no finite cost at any layer: nothing to trace */
            c->error = 2;
            return;
        }
    } else {
        int par_layer = -1, par_idx = -1, par_seq = -1;
        for (int p = 0; p < parents; ++p) {
            const int e = c->par_num_sum[v] + p;
            const int par = c->par_nodes[e];
            if (lo[par] <= 0) continue;
            par_idx = par;
            par_seq = p;
            par_layer = lo[par];
            const int slot = c->child_idx[e] % 10;
            const unsigned long long path = c->best_path[((size_t)c->child_num_sum[par] + slot) * FLT_ML + par_layer];
            const unsigned long long child = path / FLT_ML / 100, choice = (path / FLT_ML) % 100;
            if (child != (unsigned long long)v) {
                lo[v] = -1;
                continue;
            }
            lo[v] = choice == FLT_STRAIGHT ? (int)(path % FLT_ML) : (int)(choice / 9 + 1);
            const int px = c->nodes[par] / Y % X, py = c->nodes[par] % Y;
            const int r = lo[v] - 1;
            int w[6];
            if (px == cx && cy != py) {
                if (choice == FLT_STRAIGHT) {
                    w[0] = IDX3(g, r, px, cy < py ? cy : py);
                    w[1] = IDX3(g, r, px, cy < py ? py : cy);
                    emit(routes, cap, w, 2, &c->error);
                } else {
                    const int fl = (int)choice / 9, tl = (int)choice % 9;
                    const int pos = c->best_change[(size_t)e * 81 + fl * 9 + tl];
                    w[0] = IDX3(g, fl, cx, cy < pos ? cy : pos);
                    w[1] = IDX3(g, fl, cx, cy < pos ? pos : cy);
                    w[2] = IDX3(g, fl < tl ? fl : tl, cx, pos);
                    w[3] = IDX3(g, fl < tl ? tl : fl, cx, pos);
                    w[4] = IDX3(g, tl, cx, py < pos ? py : pos);
                    w[5] = IDX3(g, tl, cx, py < pos ? pos : py);
                    emit(routes, cap, w, 6, &c->error);
                }
            } else if (py == cy && cx != px) {
                if (choice == FLT_STRAIGHT) {
                    w[0] = IDX3(g, r, cx < px ? cx : px, py);
                    w[1] = IDX3(g, r, cx < px ? px : cx, py);
                    emit(routes, cap, w, 2, &c->error);
                } else {
                    const int fl = (int)choice / 9, tl = (int)choice % 9;
                    const int pos = c->best_change[(size_t)e * 81 + fl * 9 + tl];
                    w[0] = IDX3(g, fl, cx < pos ? cx : pos, cy);
                    w[1] = IDX3(g, fl, cx < pos ? pos : cx, cy);
                    w[2] = IDX3(g, fl < tl ? fl : tl, pos, cy);
                    w[3] = IDX3(g, fl < tl ? tl : fl, pos, cy);
                    w[4] = IDX3(g, tl, px < pos ? px : pos, cy);
                    w[5] = IDX3(g, tl, px < pos ? pos : px, cy);
                    emit(routes, cap, w, 6, &c->error);
                }
            }
            break;
        }
        if (par_layer == -1) {
            lo[v] = -1;
            return;
        }
        const int slot = c->child_idx[c->par_num_sum[v] + par_seq] % 10;
        const unsigned long long path = c->best_path[((size_t)c->child_num_sum[par_idx] + slot) * FLT_ML + par_layer];
        if (path / FLT_ML / 100 != (unsigned long long)v) {
            lo[v] = -1;
            return;
        }
    }
    const int kids = c->child_num_sum[v + 1] - c->child_num_sum[v];
    int minl = l + 1;
    int maxl = (l + 1) == L + 1 ? 1 : minl;
    if (lo[v] < minl) minl = lo[v];
    if (lo[v] > maxl) maxl = lo[v];
    for (int k = 0; k < kids; ++k) {
        const int lc = (int)(bp[k * FLT_ML + lo[v]] % FLT_ML);
        if (lc < minl) minl = lc;
        if (lc > maxl) maxl = lc;
    }
    if (minl < maxl) {
        const int w[2] = {IDX3(g, minl - 1, cx, cy), IDX3(g, maxl - 1, cx, cy)};
        emit(routes, cap, w, 2, &c->error);
    }
}

typedef struct {
    unsigned long long key;
    int index;
} flt_key;

static int key_cmp(const void *a, const void *b) {
    const flt_key *p = (const flt_key *)a, *q = (const flt_key *)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->index - q->index; /* stable */
}

int ora_flt_batch(const ora_grid *g, const double *presum, const float *vcost, int count, const int *net_ids,
                  const long long *node_off, const int *node_rec, const long long *edge_off, const int *edge_rec,
                  const long long *route_off, const int *route_cap, int *routes, int *layer_output) {
    const long long n0 = node_off[0], e0 = edge_off[0];
    const int nodes_n = (int)(node_off[count] - n0), edges_n = (int)(edge_off[count] - e0);
    flt_ctx c;
    memset(&c, 0, sizeof(c));
    c.g = g;
    c.presum = presum;
    c.vcost = vcost;
    c.nodes = (int *)malloc(sizeof(int) * (size_t)(nodes_n + 1));
    c.par_num_sum = (int *)malloc(sizeof(int) * (size_t)(nodes_n + 1));
    c.child_num_sum = (int *)malloc(sizeof(int) * (size_t)(nodes_n + 1));
    c.par_nodes = (int *)malloc(sizeof(int) * (size_t)(edges_n + 1));
    c.child_idx = (int *)malloc(sizeof(int) * (size_t)(edges_n + 1));
    c.cost_edges = (double *)malloc(sizeof(double) * (size_t)(edges_n + 1) * 81);
    c.best_change = (int *)malloc(sizeof(int) * (size_t)(edges_n + 1) * 81);
    c.layer_output = layer_output ? layer_output : (int *)malloc(sizeof(int) * (size_t)(nodes_n + 1));
    flt_key *keys = (flt_key *)malloc(sizeof(flt_key) * (size_t)(nodes_n + 1));
    /* assemble_batch: batch-wide node and edge indices, keys (net << 13) | depth */
    c.par_num_sum[0] = 0;
    for (int k = 0; k < count; ++k) {
        const int tn = (int)(node_off[k] - n0), te = (int)(edge_off[k] - e0);
        const int nn = (int)(node_off[k + 1] - node_off[k]), ne = (int)(edge_off[k + 1] - edge_off[k]);
        for (int i = 0; i < nn; ++i) {
            const int *r = node_rec + 4 * (node_off[k] + i);
            c.nodes[tn + i] = r[0];
            c.par_num_sum[tn + i + 1] = te + r[3];
            c.child_num_sum[tn + i] = r[2];
            keys[tn + i].key = ((unsigned long long)k << 13) | (unsigned)r[1];
            keys[tn + i].index = tn + i;
        }
        for (int i = 0; i < ne; ++i) {
            const int *r = edge_rec + 2 * (edge_off[k] + i);
            c.par_nodes[te + i] = r[0] + tn;
            c.child_idx[te + i] = r[1] + tn * 10;
        }
    }
    /* exclusive scan of child slot counts */
    int acc = 0;
    for (int i = 0; i < nodes_n; ++i) {
        const int k = c.child_num_sum[i];
        c.child_num_sum[i] = acc;
        acc += k;
    }
    c.child_num_sum[nodes_n] = acc;
    qsort(keys, (size_t)nodes_n, sizeof(flt_key), key_cmp);
    /* initialize_batch */
    c.costs = (double *)malloc(sizeof(double) * (size_t)nodes_n * FLT_ML + 8);
    c.packed = (unsigned long long *)malloc(sizeof(unsigned long long) * (size_t)(acc + 1) * FLT_ML);
    /* best_path is not cleared by the reference; every entry trace reads was written by solve */
    c.best_path = (unsigned long long *)calloc((size_t)(acc + 1) * FLT_ML, sizeof(unsigned long long));
    for (size_t i = 0; i < (size_t)nodes_n * FLT_ML; ++i) c.costs[i] = FLT_INF;
    for (size_t i = 0; i < (size_t)(acc + 1) * FLT_ML; ++i) c.packed[i] = kFltEmpty;
    for (int i = 0; i < nodes_n; ++i) c.layer_output[i] = -1;
    flt_precompute(&c, edges_n);
    /* route_flt_block: per DAG, depths deepest first (solve), then from the root (trace) */
    int pos = 0;
    for (int k = 0; k < count; ++k) {
        const int nn = (int)(node_off[k + 1] - node_off[k]);
        const int first = pos;
        pos += nn;
        int *r = routes + route_off[net_ids[k]];
        const int cap = route_cap[net_ids[k]];
        for (int i = pos - 1; i >= first; --i) flt_solve(&c, keys[i].index);
        for (int i = first; i < pos; ++i) flt_trace(&c, keys[i].index, r, cap);
    }
    const int err = c.error;
    free(c.nodes);
    free(c.par_num_sum);
    free(c.child_num_sum);
    free(c.par_nodes);
    free(c.child_idx);
    free(c.cost_edges);
    free(c.best_change);
    if (!layer_output) free(c.layer_output);
    free(c.costs);
    free(c.packed);
    free(c.best_path);
    free(keys);
    return err;
}
