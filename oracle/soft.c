/*
 * Soft (log-sum-exp) L-shape DP and its adjoint, FP64 -- TEST
 * INFRASTRUCTURE ONLY (the checker for the GPU soft DP; see oracle.h).
 *
 * The soft DP has no counterpart in the reference.  It relaxes the hard
 * L-shape DP of src/routing/lshape_warp.cuh:8-247 (restated in
 * ora_lshape_net_cost) by replacing every minimum with the soft minimum
 *     smin_tau(x_1..x_m) = -tau * log sum_j exp(-x_j / tau),
 * entries at or above ORA_BIG (masked pin layers, infeasible edges) being
 * left out.  Per net, with T = L(L+1)/2 intervals (lo <= hi) per node:
 *   V_i[lo][hi]  = sum_{l=lo}^{hi-1} vcost(l, node i), INF outside the pin layer
 *   D_i          = V_i + sum over children c of C_c
 *   best_c[k]    = smin over intervals lo <= k <= hi of D_c
 *   diag_c[pl]   = straight edge: best_c[pl] + W1_c[pl] on layers of the edge's direction
 *                  bent edge: smin over cl of the other parity of
 *                             best_c[cl] + W1_c[cl] + corner vias strictly between + W2_c[pl]
 *   C_c[lo][hi]  = smin over lo <= h <= hi of diag_c[h]
 *   F            = smin over all intervals of D_root
 * W1/W2 are the wire costs of the two legs (one leg for a straight edge),
 * i.e. sums of wcost over the track cells the leg covers.  As tau -> 0 each
 * smin tends to the hard minimum and F to the hard DP's optimum.
 *
 * The adjoint runs the same graph backwards: d smin / d x_j is the softmax
 * weight exp(-(x_j - smin) / tau).  Gradients are scattered to the cells:
 * dF/dwcost is the expected use of each wire edge, dF/dvcost the expected
 * use of each via landing.  The cost model is the smooth (FP64, exact exp)
 * form of src/cost/full_update.cuh:9-83, so F is differentiable in the
 * overflow slopes and the via weight.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ORA_INF 1e22
#define ORA_BIG 1e20
#define IDX3(g, l, x, y) (((l) * (g)->X + (x)) * (g)->Y + (y))

static int layer_dir(const ora_grid *g, int l) { return (l & 1) ^ g->DIR; }

/* overflow cost and its derivatives in (slope, zero_slope)
 * (src/cost/full_update.cuh:9-14 with of_cost_scale = 1) */
static double soft_e(double cap, double dem, double s, double z, double *ds, double *dz) {
    *ds = *dz = 0.0;
    if (cap > 0.001) {
        const double e = exp(s * (dem - cap));
        *ds = (dem - cap) * e;
        return e;
    }
    if (dem > 0) {
        const double e = exp(z * dem);
        *dz = dem * e;
        return e;
    }
    return 0.0;
}

/* incremental_of_cost (src/cost/full_update.cuh:23-32) and derivatives */
static double soft_incr(double cap, double dem, double inc, double sc, double s, double z, double *ds, double *dz) {
    double s0, z0, s1, z1;
    const double e0 = soft_e(cap, dem, s, z, &s0, &z0);
    const double e1 = soft_e(cap, dem + inc, s, z, &s1, &z1);
    const double v = (e1 - e0) * sc;
    if (v >= 1e12) {
        *ds = *dz = 0.0;
        return 1e12;
    }
    *ds = (s1 - s0) * sc;
    *dz = (z1 - z0) * sc;
    return v;
}

void ora_soft_costs(const ora_grid *g, const float *capacity, const float *demand, double slope, double zero_slope,
                    double via_weight, double *wcost, double *vcost, double *dw_ds, double *dw_dz, double *dv_ds,
                    double *dv_dz) {
    const int L = g->L, X = g->X, Y = g->Y;
    for (int l = 0; l < L; ++l) {
        const int dir = layer_dir(g, l);
        const double sc = g->unit_length_short_costs[l];
        for (int x = 0; x < X; ++x)
            for (int y = 0; y < Y; ++y) {
                const int i = IDX3(g, l, x, y);
                const int p = dir ? y : x, len = dir ? Y : X;
                const int q = p > 0 ? (dir ? IDX3(g, l, x, y - 1) : IDX3(g, l, x - 1, y)) : i;
                double a, b;
                /* wcost: src/cost/full_update.cuh:34-51 */
                double w = soft_incr(capacity[i], demand[i], 1.0, sc, slope, zero_slope, &a, &b);
                if (p < len - 1) w += g->unit_length_wire_cost * (dir ? g->y_edge_len[y] : g->x_edge_len[x]);
                wcost[i] = w;
                if (dw_ds) dw_ds[i] = a;
                if (dw_dz) dw_dz[i] = b;
                /* vcost: src/cost/full_update.cuh:58-83 */
                double v, vs = 0.0, vz = 0.0;
                if (p == 0) {
                    v = soft_incr(capacity[i], demand[i], 1.0, sc, slope, zero_slope, &vs, &vz);
                } else if (p == len - 1) {
                    v = soft_incr(capacity[q], demand[q], 1.0, sc, slope, zero_slope, &vs, &vz);
                } else {
                    double a2, b2;
                    v = soft_incr(capacity[i], demand[i], 0.5, sc, slope, zero_slope, &a, &b) +
                        soft_incr(capacity[q], demand[q], 0.5, sc, slope, zero_slope, &a2, &b2);
                    vs = a + a2;
                    vz = b + b2;
                }
                vcost[i] = v + via_weight * g->unit_via_cost;
                if (dv_ds) dv_ds[i] = vs;
                if (dv_dz) dv_dz[i] = vz;
            }
    }
}

static double smin(const double *x, int m, double tau) {
    double mn = ORA_INF;
    for (int j = 0; j < m; ++j)
        if (x[j] < mn) mn = x[j];
    if (mn >= ORA_BIG) return ORA_INF;
    double s = 0.0;
    for (int j = 0; j < m; ++j)
        if (x[j] < ORA_BIG) s += exp(-(x[j] - mn) / tau);
    return mn - tau * log(s);
}

static double weight(double x, double y, double tau) {
    if (x >= ORA_BIG || y >= ORA_BIG) return 0.0;
    return exp(-(x - y) / tau);
}

static int tri(int L, int lo, int hi) { return lo * (2 * L - lo + 1) / 2 + (hi - lo); }

/* wire cost of the track cells [lo, hi) of layer l through (x, y) along the
 * layer's direction; the gradient variant adds g to each of those cells */
static double leg(const ora_grid *g, const double *wcost, int l, int along_y, int fixed, int lo, int hi) {
    double s = 0.0;
    for (int p = lo; p < hi; ++p) s += along_y ? wcost[IDX3(g, l, fixed, p)] : wcost[IDX3(g, l, p, fixed)];
    return s;
}
static void leg_grad(const ora_grid *g, double *gw, int l, int along_y, int fixed, int lo, int hi, double gr) {
    for (int p = lo; p < hi; ++p) gw[along_y ? IDX3(g, l, fixed, p) : IDX3(g, l, p, fixed)] += gr;
}

typedef struct {
    int x, y, px, py, straight;
} edge_geom;

/* ---- Z-shape family (kind 1) ----
 * The route family the Z-shape DP prices per tree edge
 * (src/routing/zshape.cuh:62-288): from child (x, y) to parent (px, py) a
 * straight wire on one layer of the edge's direction when they share a
 * column or row (a column when they share both); otherwise the L-shape
 * through (x, py) (first leg vertical) or through (px, y) (first leg
 * horizontal), and the Z-shapes whose middle leg crosses the box on a row
 * strictly between y and py or a column strictly between x and px; each leg
 * on one layer of its direction, a via stack at each bend costing
 * vcost(k) for k from the lower layer up to below the upper one.  The edge
 * leaves the child on its first leg's layer s and arrives on its last leg's
 * layer e.  visit() calls back every candidate (s, e, cost); the adjoint
 * variant scatters a weight to the candidate's wire and via cells. */
typedef struct {
    const ora_grid *g;
    const double *wcost, *vcost;
    double *gw, *gv; /* NULL in the forward pass */
    double w;        /* the adjoint weight of the current candidate */
} zsoft;

static double zleg(zsoft *z, int l, int along_y, int fixed, int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    if (z->gw) leg_grad(z->g, z->gw, l, along_y, fixed, lo, hi, z->w);
    return leg(z->g, z->wcost, l, along_y, fixed, lo, hi);
}

static double zstack(zsoft *z, int x, int y, int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    double s = 0.0;
    for (int k = lo; k < hi; ++k) {
        s += z->vcost[IDX3(z->g, k, x, y)];
        if (z->gv) z->gv[IDX3(z->g, k, x, y)] += z->w;
    }
    return s;
}

/* One pass over the candidates of an edge.  Forward (z->gw == NULL): the
 * online soft minimum per arriving layer into mn[e], sm[e].  Adjoint: each
 * candidate's weight gdiag[e] * exp(-(best[s] + cost - diag[e]) / tau) goes
 * to gbest[s] and to its cells (the cost is evaluated once to get the
 * weight, then again with the weight set to scatter it). */
static void zcandidate(zsoft *z, int s, int e, double cost, const double *best, double tau, double *mn, double *sm,
                       const double *diag, const double *gdiag, double *gbest, int *rescatter) {
    const double b = best[s];
    if (b >= ORA_BIG) return;
    const double x = b + cost;
    if (!z->gw) {
        if (x < mn[e]) {
            sm[e] = (mn[e] >= ORA_BIG ? 0.0 : sm[e] * exp(-(mn[e] - x) / tau)) + 1.0;
            mn[e] = x;
        } else {
            sm[e] += exp(-(x - mn[e]) / tau);
        }
        return;
    }
    const double w = gdiag[e] * weight(x, diag[e], tau);
    gbest[s] += w;
    z->w = w;
    *rescatter = w != 0.0;
}

#define ZCAND(S, E, EXPR)                                                                \
    do {                                                                                 \
        double *gw_ = z->gw, *gv_ = z->gv;                                               \
        int again = 0;                                                                   \
        z->gw = NULL, z->gv = NULL;                                                      \
        const double c_ = (EXPR);                                                        \
        z->gw = gw_, z->gv = gv_;                                                        \
        zcandidate(z, S, E, c_, best, tau, mn, sm, diag, gdiag, gbest, &again);          \
        if (again) (void)(EXPR);                                                         \
    } while (0)

static void zedge(zsoft *z, int x, int y, int px, int py, const double *best, double tau, double *mn, double *sm,
                  const double *diag, const double *gdiag, double *gbest) {
    const ora_grid *g = z->g;
    const int L = g->L;
    const int minx = x < px ? x : px, maxx = x < px ? px : x, miny = y < py ? y : py, maxy = y < py ? py : y;
    if (x == px) {
        for (int a = 0; a < L; ++a)
            if (layer_dir(g, a)) ZCAND(a, a, zleg(z, a, 1, x, miny, maxy));
        return;
    }
    if (y == py) {
        for (int a = 0; a < L; ++a)
            if (!layer_dir(g, a)) ZCAND(a, a, zleg(z, a, 0, y, minx, maxx));
        return;
    }
    for (int a = 0; a < L; ++a)
        for (int h = 0; h < L; ++h) {
            if (layer_dir(g, a) != 1 || layer_dir(g, h) != 0) continue;
            /* vertical then horizontal through (x, py) */
            ZCAND(a, h, zleg(z, a, 1, x, y, py) + zstack(z, x, py, a, h) + zleg(z, h, 0, py, x, px));
            /* horizontal then vertical through (px, y): first layer h, last a */
            ZCAND(h, a, zleg(z, h, 0, y, x, px) + zstack(z, px, y, h, a) + zleg(z, a, 1, px, y, py));
        }
    for (int b = miny + 1; b < maxy; ++b) /* vertical, horizontal on row b, vertical */
        for (int a = 0; a < L; ++a)
            for (int h = 0; h < L; ++h)
                for (int v = 0; v < L; ++v) {
                    if (layer_dir(g, a) != 1 || layer_dir(g, h) != 0 || layer_dir(g, v) != 1) continue;
                    ZCAND(a, v, zleg(z, a, 1, x, y, b) + zstack(z, x, b, a, h) + zleg(z, h, 0, b, x, px) +
                                    zstack(z, px, b, h, v) + zleg(z, v, 1, px, b, py));
                }
    for (int b = minx + 1; b < maxx; ++b) /* horizontal, vertical on column b, horizontal */
        for (int a = 0; a < L; ++a)
            for (int v = 0; v < L; ++v)
                for (int h = 0; h < L; ++h) {
                    if (layer_dir(g, a) != 0 || layer_dir(g, v) != 1 || layer_dir(g, h) != 0) continue;
                    ZCAND(a, h, zleg(z, a, 0, y, x, b) + zstack(z, b, y, a, v) + zleg(z, v, 1, b, y, py) +
                                    zstack(z, b, py, v, h) + zleg(z, h, 0, py, b, px));
                }
}

double ora_soft_lshape_net(const ora_grid *g, const double *wcost, const double *vcost, int n, const int *nodes,
                           const int *parents, double tau, double *gw, double *gv) {
    return ora_soft_net(g, wcost, vcost, n, nodes, parents, tau, 0, gw, gv);
}

double ora_soft_net(const ora_grid *g, const double *wcost, const double *vcost, int n, const int *nodes,
                    const int *parents, double tau, int kind, double *gw, double *gv) {
    const int L = g->L, X = g->X, Y = g->Y, XY = X * Y, T = L * (L + 1) / 2;
    double *D = (double *)calloc((size_t)n * T, sizeof(double));
    double *best = (double *)calloc((size_t)n * L, sizeof(double));
    double *diag = (double *)calloc((size_t)n * L, sizeof(double));
    double *W1 = (double *)calloc((size_t)n * L, sizeof(double));
    double *W2 = (double *)calloc((size_t)n * L, sizeof(double));
    edge_geom *eg = (edge_geom *)calloc((size_t)n, sizeof(edge_geom));
    double tmp[64];
    /* via stacks */
    for (int i = 0; i < n; ++i) {
        const int v = nodes[i], fixed = v / XY, x = v / Y % X, y = v % Y;
        for (int lo = 0; lo < L; ++lo) {
            double s = 0.0;
            for (int hi = lo; hi < L; ++hi) {
                if (hi > lo) s += vcost[IDX3(g, hi - 1, x, y)];
                D[(size_t)i * T + tri(L, lo, hi)] = (fixed < L && (lo > fixed || hi < fixed)) ? ORA_INF : s;
            }
        }
    }
    /* bottom-up */
    for (int i = n - 1; i >= 1; --i) {
        const int p = parents[i];
        const int x = nodes[i] / Y % X, y = nodes[i] % Y, px = nodes[p] / Y % X, py = nodes[p] % Y;
        const int minx = x < px ? x : px, maxx = x < px ? px : x, miny = y < py ? y : py, maxy = y < py ? py : y;
        const int straight = x == px || y == py;
        edge_geom e = {x, y, px, py, straight};
        eg[i] = e;
        const double *Di = D + (size_t)i * T;
        for (int k = 0; k < L; ++k) {
            int m = 0;
            for (int lo = 0; lo <= k; ++lo)
                for (int hi = k; hi < L; ++hi) tmp[m++] = Di[tri(L, lo, hi)];
            best[(size_t)i * L + k] = smin(tmp, m, tau);
        }
        for (int c = 0; c < L; ++c) {
            const int d = layer_dir(g, c);
            double w1 = 0.0, w2 = 0.0;
            if (straight) {
                if (d ? x == px : y == py) w1 = d ? leg(g, wcost, c, 1, x, miny, maxy) : leg(g, wcost, c, 0, y, minx, maxx);
            } else {
                w1 = d ? leg(g, wcost, c, 1, x, miny, maxy) : leg(g, wcost, c, 0, y, minx, maxx);
                w2 = d ? leg(g, wcost, c, 1, px, miny, maxy) : leg(g, wcost, c, 0, py, minx, maxx);
            }
            W1[(size_t)i * L + c] = w1;
            W2[(size_t)i * L + c] = w2;
        }
        if (kind == 1) {
            double mn[32], sm[32];
            for (int e = 0; e < L; ++e) mn[e] = ORA_INF, sm[e] = 0.0;
            zsoft z = {g, wcost, vcost, NULL, NULL, 0.0};
            zedge(&z, x, y, px, py, best + (size_t)i * L, tau, mn, sm, NULL, NULL, NULL);
            for (int e = 0; e < L; ++e) diag[(size_t)i * L + e] = mn[e] >= ORA_BIG ? ORA_INF : mn[e] - tau * log(sm[e]);
        }
        for (int pl = 0; pl < L && kind == 0; ++pl) {
            const int d = layer_dir(g, pl);
            double val = ORA_INF;
            if (straight) {
                if (d ? x == px : y == py) {
                    const double b = best[(size_t)i * L + pl];
                    val = b >= ORA_BIG ? ORA_INF : b + W1[(size_t)i * L + pl];
                }
            } else {
                int m = 0;
                for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                    const double b = best[(size_t)i * L + cl];
                    const int lo = cl < pl ? cl : pl, hi = cl < pl ? pl : cl;
                    double f = 0.0;
                    for (int l = lo + 1; l < hi; ++l)
                        f += layer_dir(g, cl) ? vcost[IDX3(g, l, x, py)] : vcost[IDX3(g, l, px, y)];
                    tmp[m++] = b >= ORA_BIG ? ORA_INF : b + W1[(size_t)i * L + cl] + f + W2[(size_t)i * L + pl];
                }
                val = smin(tmp, m, tau);
            }
            diag[(size_t)i * L + pl] = val;
        }
        double *Dp = D + (size_t)p * T;
        for (int lo = 0; lo < L; ++lo)
            for (int hi = lo; hi < L; ++hi) {
                const double cval = smin(diag + (size_t)i * L + lo, hi - lo + 1, tau);
                double *t = Dp + tri(L, lo, hi);
                *t = (*t >= ORA_BIG || cval >= ORA_BIG) ? ORA_INF : *t + cval;
            }
    }
    const double F = smin(D, T, tau);
    if (gw && gv && F < ORA_BIG) {
        double *gD = (double *)calloc((size_t)n * T, sizeof(double));
        double gbest[32], gdiag[32];
        for (int e = 0; e < T; ++e) gD[e] = weight(D[e], F, tau);
        for (int c = 1; c < n; ++c) {
            const int p = parents[c];
            const edge_geom e = eg[c];
            const double *gC = gD + (size_t)p * T;
            const double *dg = diag + (size_t)c * L;
            for (int h = 0; h < L; ++h) gdiag[h] = 0.0, gbest[h] = 0.0;
            for (int lo = 0; lo < L; ++lo)
                for (int hi = lo; hi < L; ++hi) {
                    const double gc = gC[tri(L, lo, hi)];
                    if (gc == 0.0) continue;
                    const double cval = smin(dg + lo, hi - lo + 1, tau);
                    for (int h = lo; h <= hi; ++h) gdiag[h] += gc * weight(dg[h], cval, tau);
                }
            const int x = e.x, y = e.y, px = e.px, py = e.py;
            const int minx = x < px ? x : px, maxx = x < px ? px : x, miny = y < py ? y : py, maxy = y < py ? py : y;
            if (kind == 1) {
                zsoft z = {g, wcost, vcost, gw, gv, 0.0};
                zedge(&z, x, y, px, py, best + (size_t)c * L, tau, NULL, NULL, dg, gdiag, gbest);
            }
            for (int pl = 0; pl < L && kind == 0; ++pl) {
                if (gdiag[pl] == 0.0) continue;
                const int d = layer_dir(g, pl);
                if (e.straight) {
                    gbest[pl] += gdiag[pl];
                    if (d) leg_grad(g, gw, pl, 1, x, miny, maxy, gdiag[pl]);
                    else leg_grad(g, gw, pl, 0, y, minx, maxx, gdiag[pl]);
                    continue;
                }
                for (int cl = (pl & 1) ^ 1; cl < L; cl += 2) {
                    const double b = best[(size_t)c * L + cl];
                    if (b >= ORA_BIG) continue;
                    const int lo = cl < pl ? cl : pl, hi = cl < pl ? pl : cl;
                    const int dcl = layer_dir(g, cl);
                    double f = 0.0;
                    for (int l = lo + 1; l < hi; ++l) f += dcl ? vcost[IDX3(g, l, x, py)] : vcost[IDX3(g, l, px, y)];
                    const double t = b + W1[(size_t)c * L + cl] + f + W2[(size_t)c * L + pl];
                    const double gr = gdiag[pl] * weight(t, dg[pl], tau);
                    gbest[cl] += gr;
                    if (dcl) leg_grad(g, gw, cl, 1, x, miny, maxy, gr);
                    else leg_grad(g, gw, cl, 0, y, minx, maxx, gr);
                    if (d) leg_grad(g, gw, pl, 1, px, miny, maxy, gr);
                    else leg_grad(g, gw, pl, 0, py, minx, maxx, gr);
                    for (int l = lo + 1; l < hi; ++l) gv[dcl ? IDX3(g, l, x, py) : IDX3(g, l, px, y)] += gr;
                }
            }
            const double *Dc = D + (size_t)c * T;
            double *gDc = gD + (size_t)c * T;
            for (int lo = 0; lo < L; ++lo)
                for (int hi = lo; hi < L; ++hi) {
                    double s = 0.0;
                    for (int k = lo; k <= hi; ++k)
                        if (gbest[k] != 0.0) s += gbest[k] * weight(Dc[tri(L, lo, hi)], best[(size_t)c * L + k], tau);
                    gDc[tri(L, lo, hi)] = s;
                }
        }
        /* via stacks: D_i[lo][hi] holds vcost(l) for lo <= l < hi */
        for (int i = 0; i < n; ++i) {
            const int v = nodes[i], x = v / Y % X, y = v % Y;
            for (int lo = 0; lo < L; ++lo)
                for (int hi = lo + 1; hi < L; ++hi) {
                    const double gr = gD[(size_t)i * T + tri(L, lo, hi)];
                    if (gr == 0.0) continue;
                    for (int l = lo; l < hi; ++l) gv[IDX3(g, l, x, y)] += gr;
                }
        }
        free(gD);
    }
    free(D);
    free(best);
    free(diag);
    free(W1);
    free(W2);
    free(eg);
    return F;
}

double ora_soft_stage(const ora_grid *g, const float *capacity, const float *demand, double slope, double zero_slope,
                      double via_weight, double tau, int count, const int *node_off, const int *nodes,
                      const int *parents, double *gw, double *gv, double *gparams) {
    return ora_soft_stage_kind(g, capacity, demand, slope, zero_slope, via_weight, tau, 0, count, node_off, nodes,
                               parents, gw, gv, gparams);
}

double ora_soft_stage_kind(const ora_grid *g, const float *capacity, const float *demand, double slope,
                           double zero_slope, double via_weight, double tau, int kind, int count, const int *node_off,
                           const int *nodes, const int *parents, double *gw, double *gv, double *gparams) {
    const size_t cells = (size_t)g->L * g->X * g->Y;
    double *w = (double *)malloc(sizeof(double) * cells), *v = (double *)malloc(sizeof(double) * cells);
    double *ws = (double *)malloc(sizeof(double) * cells), *wz = (double *)malloc(sizeof(double) * cells);
    double *vs = (double *)malloc(sizeof(double) * cells), *vz = (double *)malloc(sizeof(double) * cells);
    ora_soft_costs(g, capacity, demand, slope, zero_slope, via_weight, w, v, ws, wz, vs, vz);
    const int grads = gw && gv;
    if (grads) {
        memset(gw, 0, sizeof(double) * cells);
        memset(gv, 0, sizeof(double) * cells);
    }
    double F = 0.0;
    for (int k = 0; k < count; ++k) {
        const int b = node_off[k];
        F += ora_soft_net(g, w, v, node_off[k + 1] - b, nodes + b, parents + b, tau, kind, grads ? gw : NULL,
                          grads ? gv : NULL);
    }
    if (grads && gparams) {
        double s = 0.0, z = 0.0, u = 0.0;
        for (size_t i = 0; i < cells; ++i) {
            s += gw[i] * ws[i] + gv[i] * vs[i];
            z += gw[i] * wz[i] + gv[i] * vz[i];
            u += gv[i] * g->unit_via_cost;
        }
        gparams[0] = s;
        gparams[1] = z;
        gparams[2] = u;
    }
    free(w), free(v), free(ws), free(wz), free(vs), free(vz);
    return F;
}
