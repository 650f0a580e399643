/*
 * Z-shape DP (GGR pattern routing) -- TEST INFRASTRUCTURE ONLY (the checker
 * for the GPU Z-shape kernel; see oracle.h).
 *
 * Restates zshape::solve<Warps>, viaSweep and reset_nodes
 * (src/routing/zshape.cuh:42-404) and convert_routes
 * (src/routing/zshape_topology.hpp:233-289) as sequential C.  Every phase of
 * the reference kernel writes disjoint cells from its lanes, so running its
 * lanes one after another gives the same values.
 *
 * Map space: LAYER = L + 1 planes of n x n cells, n = max(X, Y).  Map layer i
 * is routing layer i - 1 (map layer 0 carries no wires: wire cost 0, via
 * cost "infinity"); a cell of layer i is i*n*n + (vertical(i) ? x*n + y :
 * y*n + x) with vertical(i) = (i & 1) ^ DIRECTION and DIRECTION = DIR ^ 1,
 * so that the second index runs along the layer's tracks.  Integer costs:
 * a wire from cell a to cell b costs int(min(|presum(a) - presum(b)|, 5e8)),
 * a via landing int(clamp(vcost, 1, 1e6)).
 *
 * Net records (zshape_topology.hpp:59-121): [route word offset, node count,
 * then per node in preorder: x*n + y, lowest layer, highest layer (map
 * layers of a pin, -1 for a Steiner node), and up to three children as
 * record offsets (6 * preorder index, -1 when absent)].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ZINF 1000000000
#define IDX3(g, l, x, y) (((l) * (g)->X + (x)) * (g)->Y + (y))

typedef struct {
    const ora_grid *g;
    const double *presum;
    const float *vcost;
    int n, layers, dirv; /* dirv: DIRECTION */
    long long plane;     /* layers * n * n */
} zctx;

/* Test probe (ora_zshape_probe): when set, the last leg also records in
 * prev planes 7-9 the layer cell where the middle leg really ran (the
 * origin of the via-sweep chain at the second bend), and the traceback can
 * take the middle leg from there instead of from the one recorded via step.
 * Off in every stage run. */
static int probe_origin = 0;

static int zmin(int a, int b) { return a < b ? a : b; }
static int zmax(int a, int b) { return a < b ? b : a; }

/* map cell of layer i at (x, y) */
static long long zcell(const zctx *z, int i, int x, int y) {
    const long long nn = (long long)z->n * z->n;
    return i * nn + (((i & 1) ^ z->dirv) ? (long long)x * z->n + y : (long long)y * z->n + x);
}

/* WireCosts / ViaCosts (src/routing/zshape.cuh:8-31) */
static double zwire_sum(const zctx *z, long long index) {
    const long long nn = (long long)z->n * z->n;
    const int layer = (int)(index / nn) - 1;
    if (layer < 0 || layer >= z->g->L) return 0;
    const int a = (int)(index / z->n % z->n), b = (int)(index % z->n);
    const int along_y = (layer & 1) ^ z->g->DIR;
    const int x = along_y ? a : b, y = along_y ? b : a;
    return z->presum[IDX3(z->g, layer, x, y)];
}
static int zvia(const zctx *z, long long index) {
    const long long nn = (long long)z->n * z->n;
    const int layer = (int)(index / nn) - 1;
    if (layer < 0 || layer >= z->g->L) return ZINF;
    const int a = (int)(index / z->n % z->n), b = (int)(index % z->n);
    const int along_y = (layer & 1) ^ z->g->DIR;
    const int x = along_y ? a : b, y = along_y ? b : a;
    return (int)fminf(fmaxf(z->vcost[IDX3(z->g, layer, x, y)], 1.f), 1.e6f);
}
static int zwire(const zctx *z, long long a, long long b) {
    const double v = fabs(zwire_sum(z, a) - zwire_sum(z, b));
    return (int)fmin(v, (double)(ZINF / 2));
}

/* viaSweep (src/routing/zshape.cuh:39-59) on plane 4 at (x, y) */
static void via_sweep(const zctx *z, int *map, int *prev, int x, int y) {
    const long long P = z->plane, nn = (long long)z->n * z->n;
    int *m = map + 4 * P, *pr = prev + 4 * P;
    for (int i = 1; i < z->layers; i++) {
        const long long cur = zcell(z, i, x, y);
        pr[cur + P] = (int)cur;
    }
    for (int i = 2; i < z->layers; i++) {
        const long long cur = zcell(z, i, x, y);
        /* the layer below, indexed with this layer's orientation swapped */
        const long long last = (i - 1) * nn + (((i & 1) ^ z->dirv) ? (long long)y * z->n + x : (long long)x * z->n + y);
        if (m[cur] > m[last] + zvia(z, last)) {
            m[cur] = m[last] + zvia(z, last);
            pr[cur] = pr[last];
            pr[cur + P] = (int)last;
        }
    }
    for (int i = z->layers - 2; i >= 1; i--) {
        const long long cur = zcell(z, i, x, y);
        const long long last = (i + 1) * nn + (((i & 1) ^ z->dirv) ? (long long)y * z->n + x : (long long)x * z->n + y);
        if (m[cur] > m[last] + zvia(z, cur)) {
            m[cur] = m[last] + zvia(z, cur);
            pr[cur] = pr[last];
            pr[cur + P] = (int)last;
        }
    }
}

/* probe only: follow the via-sweep predecessors (plane 5) from a bend cell
 * down to the layer cell whose value a leg wrote (its own predecessor) */
static int sweep_origin(const int *pr5, int cell) {
    while (pr5[cell] != cell) cell = pr5[cell];
    return cell;
}

/* one child edge (from -> to) into child slot j of the node at `to`
 * (src/routing/zshape.cuh:82-259) */
static void zedge(const zctx *z, int *map, int *prev, int *prev2, int j, int fromx, int fromy, int tox, int toy) {
    const long long P = z->plane, nn = (long long)z->n * z->n;
    const int L1 = z->layers;
    int *m4 = map + 4 * P, *p4 = prev + 4 * P;
    if (fromx != tox && fromy != toy) {
        /* the two rows and two columns of the box start at infinity */
        for (int a = zmin(fromx, tox); a <= zmax(fromx, tox); a++)
            for (int i = 1; i < L1; i++) {
                if ((i & 1) ^ z->dirv) {
                    m4[i * nn + (long long)a * z->n + toy] = ZINF;
                    m4[i * nn + (long long)a * z->n + fromy] = ZINF;
                } else {
                    m4[i * nn + (long long)toy * z->n + a] = ZINF;
                    m4[i * nn + (long long)fromy * z->n + a] = ZINF;
                }
            }
        for (int b = zmin(fromy, toy); b <= zmax(fromy, toy); b++)
            for (int i = 1; i < L1; i++) {
                if ((i & 1) ^ z->dirv) {
                    m4[i * nn + (long long)tox * z->n + b] = ZINF;
                    m4[i * nn + (long long)fromx * z->n + b] = ZINF;
                } else {
                    m4[i * nn + (long long)b * z->n + tox] = ZINF;
                    m4[i * nn + (long long)b * z->n + fromx] = ZINF;
                }
            }
        /* first leg: along the child's column (vertical layers) or row */
        for (int i = 1; i < L1; i++) {
            if ((i & 1) ^ z->dirv) {
                const long long last = i * nn + (long long)fromx * z->n + fromy;
                for (int b = zmin(fromy, toy); b <= zmax(fromy, toy); b++) {
                    if (b == fromy) continue;
                    const long long cur = i * nn + (long long)fromx * z->n + b;
                    m4[cur] = map[last] + zwire(z, last, cur);
                    p4[cur] = (int)last;
                }
            } else {
                const long long last = i * nn + (long long)fromy * z->n + fromx;
                for (int b = zmin(fromx, tox); b <= zmax(fromx, tox); b++) {
                    if (b == fromx) continue;
                    const long long cur = i * nn + (long long)fromy * z->n + b;
                    m4[cur] = map[last] + zwire(z, last, cur);
                    p4[cur] = (int)last;
                }
            }
        }
        for (int b = zmin(fromy, toy); b <= zmax(fromy, toy); b++)
            if (b != fromy) via_sweep(z, map, prev, fromx, b);
        for (int b = zmin(fromx, tox); b <= zmax(fromx, tox); b++)
            if (b != fromx) via_sweep(z, map, prev, b, fromy);
        /* middle leg: across the box, keeping the first leg's start */
        for (int i = 1; i < L1; i++) {
            if ((i & 1) ^ z->dirv) {
                for (int b = zmin(fromx, tox); b <= zmax(fromx, tox); b++) {
                    const long long last = i * nn + (long long)b * z->n + fromy, cur = i * nn + (long long)b * z->n + toy;
                    const int cost = zwire(z, last, cur);
                    if (m4[last] + cost < m4[cur]) {
                        m4[cur] = m4[last] + cost;
                        p4[cur] = p4[last];
                    }
                }
            } else {
                for (int b = zmin(fromy, toy); b <= zmax(fromy, toy); b++) {
                    const long long last = i * nn + (long long)b * z->n + fromx, cur = i * nn + (long long)b * z->n + tox;
                    const int cost = zwire(z, last, cur);
                    if (m4[last] + cost < m4[cur]) {
                        m4[cur] = m4[last] + cost;
                        p4[cur] = p4[last];
                    }
                }
            }
        }
        for (int b = zmin(fromy, toy); b <= zmax(fromy, toy); b++)
            if (b != toy) via_sweep(z, map, prev, tox, b);
        for (int b = zmin(fromx, tox); b <= zmax(fromx, tox); b++)
            if (b != tox) via_sweep(z, map, prev, b, toy);
        /* last leg into the node, remembering the first leg's start (prev)
         * and the middle leg's start (prev2) */
        for (int i = 1; i < L1; i++) {
            if ((i & 1) ^ z->dirv) {
                const long long cur = i * nn + (long long)tox * z->n + toy;
                for (int b = zmin(fromy, toy); b <= zmax(fromy, toy); b++) {
                    if (b == toy) continue;
                    const long long last = i * nn + (long long)tox * z->n + b;
                    const int cost = zwire(z, last, cur);
                    if (m4[last] + cost < map[cur + j * P]) {
                        map[cur + j * P] = m4[last] + cost;
                        prev[cur + j * P] = p4[last];
                        prev2[cur + j * P] = prev[5 * P + last];
                        if (probe_origin) prev[cur + (6 + j) * P] = sweep_origin(prev + 5 * P, prev[5 * P + last]);
                    }
                }
            } else {
                const long long cur = i * nn + (long long)toy * z->n + tox;
                for (int b = zmin(fromx, tox); b <= zmax(fromx, tox); b++) {
                    if (b == tox) continue;
                    const long long last = i * nn + (long long)toy * z->n + b;
                    const int cost = zwire(z, last, cur);
                    if (m4[last] + cost < map[cur + j * P]) {
                        map[cur + j * P] = m4[last] + cost;
                        prev[cur + j * P] = p4[last];
                        prev2[cur + j * P] = prev[5 * P + last];
                        if (probe_origin) prev[cur + (6 + j) * P] = sweep_origin(prev + 5 * P, prev[5 * P + last]);
                    }
                }
            }
        }
    } else if (fromx == tox) {
        for (int i = 1; i < L1; i++) {
            if ((i & 1) ^ z->dirv) {
                const long long last = i * nn + (long long)fromx * z->n + fromy, cur = i * nn + (long long)tox * z->n + toy;
                const int cost = zwire(z, last, cur);
                if (map[cur + j * P] > map[last] + cost) {
                    map[cur + j * P] = map[last] + cost;
                    prev[cur + j * P] = (int)last;
                }
            } else {
                map[i * nn + (long long)toy * z->n + tox + j * P] = ZINF;
            }
        }
    } else {
        for (int i = 1; i < L1; i++) {
            if (!((i & 1) ^ z->dirv)) {
                const long long last = i * nn + (long long)fromy * z->n + fromx, cur = i * nn + (long long)toy * z->n + tox;
                const int cost = zwire(z, last, cur);
                if (map[cur + j * P] > map[last] + cost) {
                    map[cur + j * P] = map[last] + cost;
                    prev[cur + j * P] = (int)last;
                }
            } else {
                map[i * nn + (long long)tox * z->n + toy + j * P] = ZINF;
            }
        }
    }
}

/* zshape.cuh:260-288: combine the children of the node at rec with the via
 * stack of every layer interval [lower, upper] */
static void znode(const zctx *z, int *map, int *prev, const int *rec) {
    const long long P = z->plane;
    const int tox = rec[0] / z->n, toy = rec[0] % z->n, minlayer = rec[1], maxlayer = rec[2];
    const int L1 = z->layers;
    for (int lower = 0; lower < L1; lower++) {
        int via_sum = 0, min_cost = ZINF, bestpos[3] = {0, 0, 0};
        int costs[3] = {rec[3] == -1 ? 0 : ZINF, rec[4] == -1 ? 0 : ZINF, rec[5] == -1 ? 0 : ZINF};
        for (int upper = lower; upper < L1; upper++) {
            const long long idx = zcell(z, upper, tox, toy);
            for (int j = 0; j < 3; j++)
                if (rec[3 + j] != -1 && map[idx + (j + 1) * P] < costs[j]) {
                    costs[j] = map[idx + (j + 1) * P];
                    bestpos[j] = upper;
                }
            if (1LL * costs[0] + costs[1] + costs[2] < min_cost) min_cost = costs[0] + costs[1] + costs[2];
            if (minlayer == -1 || (lower <= minlayer && maxlayer <= upper)) {
                for (int i = lower; i <= upper; i++) {
                    const long long index = zcell(z, i, tox, toy);
                    if (map[index] > min_cost + via_sum) {
                        map[index] = min_cost + via_sum;
                        prev[index] = (((lower * L1 + upper) * L1 + bestpos[2]) * L1 + bestpos[1]) * L1 + bestpos[0];
                    }
                }
            }
            via_sum += zvia(z, idx);
        }
    }
}

static void zemit(int *routes, long long a, long long b) {
    routes[routes[0]++] = (int)a;
    routes[routes[0]++] = (int)b;
}

/* zshape.cuh:290-399: root layer, then per node its via stack and the
 * leg pieces to each child, into raw words [start, length or -1] */
static int ztrace(const zctx *z, int *map, int *prev, int *prev2, const int *pts, int node_cnt, int *routes,
                  int ideal, int *quirk) {
    const long long P = z->plane, nn = (long long)z->n * z->n;
    const int L1 = z->layers, n = z->n;
    int min_total = ZINF;
    for (int i = 0; i < L1; i++) {
        const long long idx = zcell(z, i, pts[0] / n, pts[0] % n);
        if (map[idx] < min_total) {
            min_total = map[idx];
            map[5 * P + pts[0]] = i;
        }
    }
    if (min_total == ZINF || min_total < 0) {
        routes[0] = -1;
        return min_total;
    }
    for (int t = 0; t < 6 * node_cnt; t += 6) {
        const int x = pts[t] / n, y = pts[t] % n, layer = map[5 * P + pts[t]];
        const long long idx = zcell(z, layer, x, y);
        const int minl = prev[idx] / L1 / L1 / L1 / L1, maxl = prev[idx] / L1 / L1 / L1 % L1;
        for (int i = minl; i < maxl; i++) zemit(routes, zcell(z, i, x, y), -1);
        int bestpos = prev[idx];
        for (int j = 0; j < 3; j++) {
            if (pts[t + 3 + j] == -1) break;
            const int child = pts[pts[t + 3 + j]];
            const int entry_layer = bestpos % L1;
            const long long entry = zcell(z, entry_layer, x, y);
            bestpos /= L1;
            const long long last = prev[(j + 1) * P + entry];
            const int last_layer = (int)(last / nn);
            int lastx = (int)(last / n % n), lasty = (int)(last % n);
            if (!((last_layer & 1) ^ z->dirv)) {
                const int s = lastx;
                lastx = lasty;
                lasty = s;
            }
            if (lastx != x && lasty != y) {
                const long long rec_mid = prev2[(j + 1) * P + entry];
                const long long mid = ideal ? prev[(7 + j) * P + entry] : rec_mid;
                if (quirk && probe_origin && prev[(7 + j) * P + entry] != rec_mid) *quirk = 1;
                const int mid_layer = (int)(mid / nn);
                int midx = (int)(mid / n % n), midy = (int)(mid % n);
                if (!((mid_layer & 1) ^ z->dirv)) {
                    const int s = midx;
                    midx = midy;
                    midy = s;
                }
                const long long p1 = ((entry_layer & 1) ^ z->dirv) ? entry - y + midy : entry - x + midx;
                const long long p3 = ((mid_layer & 1) ^ z->dirv) ? mid - midy + lasty : mid - midx + lastx;
                const long long p4 = ((last_layer & 1) ^ z->dirv) ? last - lasty + midy : last - lastx + midx;
                if (entry != p1) zemit(routes, entry < p1 ? entry : p1, entry < p1 ? p1 - entry : entry - p1);
                if (p1 != mid)
                    for (int i = zmin(entry_layer, mid_layer); i < zmax(entry_layer, mid_layer); i++)
                        zemit(routes, zcell(z, i, midx, midy), -1);
                if (mid != p3) zemit(routes, mid < p3 ? mid : p3, mid < p3 ? p3 - mid : mid - p3);
                if (p3 != p4) {
                    int tx = (int)(p3 / n % n), ty = (int)(p3 % n);
                    if (!(((p3 / nn) & 1) ^ z->dirv)) {
                        const int s = tx;
                        tx = ty;
                        ty = s;
                    }
                    for (int i = zmin(mid_layer, last_layer); i < zmax(mid_layer, last_layer); i++)
                        zemit(routes, zcell(z, i, tx, ty), -1);
                }
                if (p4 != last) zemit(routes, p4 < last ? p4 : last, p4 < last ? last - p4 : p4 - last);
            } else {
                zemit(routes, entry < last ? entry : last, entry < last ? last - entry : entry - last);
            }
            map[5 * P + child] = last_layer;
        }
    }
    return min_total;
}

void ora_zshape_reset(const ora_grid *g, int *map, const int *points, int count) {
    const int n = g->X > g->Y ? g->X : g->Y, layers = g->L + 1;
    zctx z = {g, NULL, NULL, n, layers, g->DIR ^ 1, (long long)layers * n * n};
    for (int k = 0; k < count; ++k) {
        const int *rec = points + points[k];
        const int nodes = rec[1];
        rec += 2;
        for (int i = 0; i < nodes * layers; ++i) {
            const int layer = i % layers, point = rec[(i / layers) * 6];
            const long long index = zcell(&z, layer, point / n, point % n);
            for (int plane = 0; plane < 4; ++plane) map[plane * z.plane + index] = ZINF;
        }
    }
}

static void zsolve(const ora_grid *g, const double *presum, const float *vcost, int *map, int *prev,
                   const int *points, int count, int *raw, int *cost, int ideal, int *quirk) {
    const int n = g->X > g->Y ? g->X : g->Y, layers = g->L + 1;
    zctx z = {g, presum, vcost, n, layers, g->DIR ^ 1, (long long)layers * n * n};
    int *prev2 = prev + 10 * z.plane;
    for (int k = 0; k < count; ++k) {
        const int *pts = points + points[k];
        int *routes = raw + pts[0];
        routes[0] = 1;
        const int node_cnt = pts[1];
        pts += 2;
        for (int t = 6 * (node_cnt - 1); t >= 0; t -= 6) {
            const int tox = pts[t] / n, toy = pts[t] % n;
            for (int j = 1; j <= 3; j++) {
                if (pts[t + 2 + j] == -1) continue;
                const int child = pts[pts[t + 2 + j]];
                zedge(&z, map, prev, prev2, j, child / n, child % n, tox, toy);
            }
            znode(&z, map, prev, pts + t);
        }
        if (quirk) quirk[k] = 0;
        const int c = ztrace(&z, map, prev, prev2, pts, node_cnt, routes, ideal, quirk ? quirk + k : NULL);
        if (cost) cost[k] = c;
    }
}

void ora_zshape_solve(const ora_grid *g, const double *presum, const float *vcost, int *map, int *prev,
                      const int *points, int count, int *raw, int *cost) {
    zsolve(g, presum, vcost, map, prev, points, count, raw, cost, 0, NULL);
}

void ora_zshape_probe(const ora_grid *g, const double *presum, const float *vcost, int *map, int *prev,
                      const int *points, int count, int *raw, int *cost, int ideal, int *quirk) {
    probe_origin = 1;
    zsolve(g, presum, vcost, map, prev, points, count, raw, cost, ideal, quirk);
    probe_origin = 0;
}

int ora_zshape_convert(const ora_grid *g, const int *raw, const int *raw_off, const int *net_ids, int count,
                       const long long *route_off, const int *route_cap, int *routes) {
    const int n = g->X > g->Y ? g->X : g->Y, X = g->X, Y = g->Y, L = g->L;
    int err = 0;
    for (int k = 0; k < count; ++k) {
        const int *src = raw + raw_off[k];
        int *dst = routes + route_off[net_ids[k]];
        const int cap = route_cap[net_ids[k]];
        int out = 1;
        for (int p = 1; p + 1 < src[0]; p += 2) {
            const int start = src[p], len = src[p + 1];
            const int layer = start / (n * n) - 1;
            const int a = start / n % n, b = start % n;
            const int vert = (layer & 1) ^ g->DIR;
            const int x = vert ? a : b, y = vert ? b : a;
            if (layer < 0 || layer >= L || x >= X || y >= Y) {
                err = 3;
                continue;
            }
            const int pos0 = IDX3(g, layer, x, y);
            int pos1 = pos0;
            if (len == -1) {
                if (layer + 1 >= L) {
                    err = 3;
                    continue;
                }
                pos1 += X * Y;
            } else {
                const int end = start + len, ea = end / n % n, eb = end % n;
                const int ex = vert ? ea : eb, ey = vert ? eb : ea;
                if (ex >= X || ey >= Y) {
                    err = 3;
                    continue;
                }
                pos1 = IDX3(g, layer, ex, ey);
            }
            if (out + 1 < cap) {
                dst[out++] = pos0;
                dst[out++] = pos1;
            } else {
                err = 1;
            }
        }
        if (src[0] <= 0) err = 2;
        dst[0] = out;
    }
    return err;
}
