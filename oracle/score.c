/*
 * Oracle of the evaluator score: NVR_DB::read_gr_solution and
 * update_stacked_via_counter (resources/evaluator.cpp:458-675), restated
 * over line lists instead of .out text.  TEST INFRASTRUCTURE ONLY.
 *
 * Per net the lines come in .out order: a via from metal 0 to metal 1 for
 * each pin on metal 0, the extra pairs, then each subnet's route pairs
 * shifted up one metal layer.  The net's flag grid, stacked-via counting and
 * the x-major per-layer loss sum follow the evaluator statement by
 * statement, metal 0 included: the evaluator never sets its direction, so it
 * is no routing layer (its wires are rejected, it has no overflow or
 * wirelength, its vias are costed).  Parity: pinned against the reference
 * evaluator binary built in oracle/_ref on whole .out files
 * (tests/test_score.py).
 *
 * ora_connectivity restates the evaluator's per-net completion test
 * (resources/evaluator.cpp:527-538 with check_connectivity, :677-749): the
 * same flag grid marks, then a flood fill from the first pin's access
 * points.  Parity: pinned against the same evaluator binary on solutions
 * with and without planted opens (tests/test_score.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct line_sink {
    int X, Y, M, net;
    const int *dir;
    int *flag, *wire, *stacked;
    unsigned long long *vias;
    int *via_loc, nvia, cap_via;
    int bad;
} line_sink;

static void push_via(line_sink *s, int x, int y, int z) {
    if (s->nvia == s->cap_via) {
        s->cap_via = s->cap_via ? 2 * s->cap_via : 64;
        s->via_loc = (int *)realloc(s->via_loc, sizeof(int) * 3 * (size_t)s->cap_via);
    }
    s->via_loc[3 * s->nvia] = x, s->via_loc[3 * s->nvia + 1] = y, s->via_loc[3 * s->nvia + 2] = z;
    ++s->nvia;
}

#define CELL(s, z, x, y) (((size_t)(z) * (s)->X + (x)) * (s)->Y + (y))

/* resources/evaluator.cpp:518-569: one line of a net */
static void take_line(line_sink *s, int p0, int p1) {
    if (p0 > p1) {
        const int t = p0;
        p0 = p1, p1 = t;
    }
    const int XY = s->X * s->Y;
    const int xl = p0 / s->Y % s->X, yl = p0 % s->Y, zl = p0 / XY;
    const int xh = p1 / s->Y % s->X, yh = p1 % s->Y, zh = p1 / XY;
    if (zh != zl) {
        if (xh == xl && yh == yl) {
            for (int z = zl; z < zh; z++) {
                s->vias[z]++;
                push_via(s, xl, yl, z);
            }
        } else {
            s->bad = 1;
        }
        return;
    }
    if (zl == 0) {
        s->bad = 1; /* metal 0 never gets a direction: not a routing layer (:383-388, :556-559) */
    } else if (s->dir[zl] == 0) {
        if (xh > xl && yh == yl) {
            for (int x = xl; x < xh; x++) {
                s->flag[CELL(s, zl, x, yl)] = s->net;
                s->wire[CELL(s, zl, x, yl)]++;
            }
            s->flag[CELL(s, zl, xh, yl)] = s->net;
        } else {
            s->bad = 1;
        }
    } else {
        if (yh > yl && xh == xl) {
            for (int y = yl; y < yh; y++) {
                s->flag[CELL(s, zl, xl, y)] = s->net;
                s->wire[CELL(s, zl, xl, y)]++;
            }
            s->flag[CELL(s, zl, xl, yh)] = s->net;
        } else {
            s->bad = 1;
        }
    }
}

/* resources/evaluator.cpp:635-675 */
static void stack_vias(line_sink *s) {
    for (int i = 0; i < s->nvia; ++i) {
        const int x = s->via_loc[3 * i], y = s->via_loc[3 * i + 1], z = s->via_loc[3 * i + 2];
        if (s->flag[CELL(s, z, x, y)] == s->net) continue;
        s->flag[CELL(s, z, x, y)] = s->net;
        if (s->dir[z] == 0) {
            if (x > 0 && x < s->X - 1) {
                s->stacked[CELL(s, z, x - 1, y)]++;
                s->stacked[CELL(s, z, x, y)]++;
            } else if (x > 0) {
                s->stacked[CELL(s, z, x - 1, y)] += 2;
            } else if (x < s->X - 1) {
                s->stacked[CELL(s, z, x, y)] += 2;
            }
        } else {
            if (y > 0 && y < s->Y - 1) {
                s->stacked[CELL(s, z, x, y - 1)]++;
                s->stacked[CELL(s, z, x, y)]++;
            } else if (y > 0) {
                s->stacked[CELL(s, z, x, y - 1)] += 2;
            } else if (y < s->Y - 1) {
                s->stacked[CELL(s, z, x, y)] += 2;
            }
        }
    }
    for (int i = 0; i < s->nvia; ++i) {
        const int x = s->via_loc[3 * i], y = s->via_loc[3 * i + 1], z = s->via_loc[3 * i + 2];
        s->flag[CELL(s, z, x, y)] = s->net;
        s->flag[CELL(s, z + 1, x, y)] = s->net;
    }
}

int ora_score(int X, int Y, int M, int nets, const int *pin_off, const int *pins, const int *extra_off,
              const int *extra, const int *sub_off, const int *subs, const int *routes, const int *pin_acc_num,
              int per_pin, const int *layer_dir, const double *capacity, const double *overflow_cost,
              const int *x_len, const int *y_len, double unit_wire, double unit_via, double *out,
              unsigned short *demand) {
    const size_t XY = (size_t)X * Y, cells = XY * (size_t)M;
    line_sink s;
    memset(&s, 0, sizeof(s));
    s.X = X, s.Y = Y, s.M = M, s.dir = layer_dir;
    s.flag = (int *)malloc(sizeof(int) * cells);
    s.wire = (int *)calloc(cells, sizeof(int));
    s.stacked = (int *)calloc(cells, sizeof(int));
    s.vias = (unsigned long long *)calloc((size_t)M, sizeof(unsigned long long));
    for (size_t i = 0; i < cells; ++i) s.flag[i] = -1;
    long long bad_nets = 0;
    for (int k = 0; k < nets; ++k) {
        s.net = k, s.nvia = 0, s.bad = 0;
        for (int j = pin_off[k]; j < pin_off[k + 1]; ++j)
            if ((size_t)pins[j] < XY) take_line(&s, pins[j], pins[j] + (int)XY);
        for (int j = extra_off[k]; j + 1 < extra_off[k + 1]; j += 2) take_line(&s, extra[j], extra[j + 1]);
        for (int t = sub_off[k]; t < sub_off[k + 1]; ++t) {
            const int *r = routes + (size_t)pin_acc_num[subs[t]] * per_pin;
            const int cap = (pin_acc_num[subs[t] + 1] - pin_acc_num[subs[t]]) * per_pin;
            const int words = r[0] < cap ? r[0] : cap;
            for (int j = 1; j + 1 <= words - 1; j += 2) take_line(&s, r[j] + (int)XY, r[j + 1] + (int)XY);
        }
        stack_vias(&s);
        bad_nets += s.bad;
    }
    /* resources/evaluator.cpp:575-623 */
    double wl_cost = 0, via_cost = 0, overflow_cost_total = 0;
    const double overflow_slope = 0.5;
    for (int z = 0; z < M; z++) {
        if (z == 0) { /* not a routing layer: its vias only (:577-580) */
            via_cost += (double)s.vias[z] * unit_via;
            if (demand) memset(demand, 0, sizeof(unsigned short) * XY);
            continue;
        }
        unsigned long long total_wl = 0;
        double layer_overflows = 0;
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++) {
                const size_t c = ((size_t)z * X + x) * Y + y;
                const unsigned short d = (unsigned short)(2 * s.wire[c] + s.stacked[c]);
                if (demand) demand[c] = d;
                const double cap = capacity[c];
                if (cap > 0.001) {
                    const double overflow = (double)d - 2 * cap;
                    layer_overflows += exp(overflow / 2 * overflow_slope);
                } else if (cap >= 0 && d > 0) {
                    layer_overflows += exp(1.5 * (double)d * overflow_slope);
                }
                if (layer_dir[z] == 0) {
                    if (x < X - 1) total_wl += (unsigned long long)(s.wire[c] * x_len[x]);
                } else if (y < Y - 1) {
                    total_wl += (unsigned long long)(s.wire[c] * y_len[y]);
                }
            }
        overflow_cost_total += layer_overflows * overflow_cost[z];
        via_cost += (double)s.vias[z] * unit_via;
        wl_cost += (double)total_wl * unit_wire;
    }
    out[0] = wl_cost;
    out[1] = via_cost;
    out[2] = overflow_cost_total;
    out[3] = overflow_cost_total + via_cost + wl_cost;
    out[4] = (double)bad_nets;
    free(s.flag), free(s.wire), free(s.stacked), free(s.vias), free(s.via_loc);
    return 0;
}

/* resources/evaluator.cpp:677-749: flood fill over the cells marked with
 * `mark`, from the first pin's marked access points, along each routing
 * layer's direction (metal 0 has none) and up and down; every other pin
 * needs a traced access point. */
static int connected(line_sink *s, int mark, int traced, int pin0, int npins, const int *acc_off, const int *acc,
                     int *stack) {
    const int XY = s->X * s->Y;
    int top = 0;
    for (int a = acc_off[pin0]; a < acc_off[pin0 + 1]; ++a) {
        const int c = acc[a];
        if (s->flag[c] == mark) {
            s->flag[c] = traced;
            stack[top++] = c;
        }
    }
    while (top) {
        const int c = stack[--top];
        const int z = c / XY, x = c / s->Y % s->X, y = c % s->Y;
        int nb[4], n = 0;
        if (z != 0 && s->dir[z] == 0) {
            if (x > 0) nb[n++] = c - s->Y;
            if (x < s->X - 1) nb[n++] = c + s->Y;
        } else if (z != 0) {
            if (y > 0) nb[n++] = c - 1;
            if (y < s->Y - 1) nb[n++] = c + 1;
        }
        if (z > 0) nb[n++] = c - XY;
        if (z < s->M - 1) nb[n++] = c + XY;
        for (int i = 0; i < n; ++i)
            if (s->flag[nb[i]] == mark) {
                s->flag[nb[i]] = traced;
                stack[top++] = nb[i];
            }
    }
    for (int p = 1; p < npins; ++p) {
        int ok = 0;
        for (int a = acc_off[pin0 + p]; a < acc_off[pin0 + p + 1] && !ok; ++a) ok = s->flag[acc[a]] == traced;
        if (!ok) return 0;
    }
    return 1;
}

int ora_connectivity(int X, int Y, int M, int nets, const int *pin_off, const int *pins, const int *extra_off,
                     const int *extra, const int *sub_off, const int *subs, const int *routes, const int *pin_acc_num,
                     int per_pin, const int *routes2, const int *pin_off2, int per_pin2, const int *layer_dir,
                     const int *apin_off, const int *acc_off, const int *acc, unsigned char *status) {
    const size_t XY = (size_t)X * Y, cells = XY * (size_t)M;
    line_sink s;
    memset(&s, 0, sizeof(s));
    s.X = X, s.Y = Y, s.M = M, s.dir = layer_dir;
    s.flag = (int *)malloc(sizeof(int) * cells);
    s.wire = (int *)calloc(cells, sizeof(int));
    s.stacked = (int *)calloc(cells, sizeof(int));
    s.vias = (unsigned long long *)calloc((size_t)M, sizeof(unsigned long long));
    int *stack = (int *)malloc(sizeof(int) * cells); /* every cell is pushed at most once per net */
    for (size_t i = 0; i < cells; ++i) s.flag[i] = -1;
    for (int k = 0; k < nets; ++k) {
        s.net = k, s.nvia = 0, s.bad = 0;
        for (int j = pin_off[k]; j < pin_off[k + 1]; ++j)
            if ((size_t)pins[j] < XY) take_line(&s, pins[j], pins[j] + (int)XY);
        for (int j = extra_off[k]; j + 1 < extra_off[k + 1]; j += 2) take_line(&s, extra[j], extra[j + 1]);
        for (int t = sub_off[k]; t < sub_off[k + 1]; ++t) {
            const int id = subs[t], np = pin_acc_num[id + 1] - pin_acc_num[id];
            const int two = routes2 && pin_off2 && pin_off2[id] >= 0;
            const int *r = two ? routes2 + (size_t)pin_off2[id] * per_pin2 : routes + (size_t)pin_acc_num[id] * per_pin;
            const int cap = np * (two ? per_pin2 : per_pin);
            const int words = r[0] < cap ? r[0] : cap;
            for (int j = 1; j + 1 <= words - 1; j += 2) take_line(&s, r[j] + (int)XY, r[j + 1] + (int)XY);
        }
        stack_vias(&s);
        if (s.bad) status[k] = 1; /* has_connectivity_violation: counted open (:529-530) */
        else if (apin_off[k + 1] > apin_off[k] &&
                 !connected(&s, k, k + nets, apin_off[k], apin_off[k + 1] - apin_off[k], acc_off, acc, stack))
            status[k] = 2;
        else
            status[k] = 0;
    }
    free(s.flag), free(s.wire), free(s.stacked), free(s.vias), free(s.via_loc), free(stack);
    return 0;
}
