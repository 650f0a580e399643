/*
 * CPU oracle for the GRotor routing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links or calls it.
 *
 * Every function restates the reference algorithm it names (file:line in the
 * reference tree) in plain C, single threaded, with -ffp-contract=off so that
 * every floating-point operation is one IEEE-754 rounding, as the reference's
 * HIP build compiles it (CMakeLists.txt:37).
 *
 * Grid arrays use the reference layout IDX(l, x, y) = (l * X + x) * Y + y
 * (include/router_state.hpp:267) over the L routing layers (metal layer l + 1).
 * Layer direction: dir(l) = (l & 1) ^ DIR, 0 = horizontal (edges along x),
 * 1 = vertical (edges along y) (src/preprocess.hpp:1621).
 */
#ifndef GROTOR_ORACLE_H
#define GROTOR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_grid {
    int L, X, Y, DIR;
    double unit_length_wire_cost;
    double unit_via_cost;
    const double *unit_length_short_costs; /* [L] */
    const int *x_edge_len;                 /* [X-1] */
    const int *y_edge_len;                 /* [Y-1] */
} ora_grid;

/* Project exponential: binary32 in, binary32 out.  Stands in for the CUDA
 * intrinsic __expf used by src/cost/full_update.cuh:11-20, whose numerics are
 * not reproducible here.  Restated identically in the HIP kernels. */
float ora_expf(float x);

/* src/cost/full_update.cuh:9-32 with of_cost_scale = 1.  `slope` and
 * `zero_slope` are the 0.5 and 1.5 constants of of_cost_scaled. */
double ora_of_cost(float capacity, float demand, double slope, double zero_slope);
double ora_incremental_of_cost(float capacity, float demand, double incre_demand,
                               double short_cost, double slope, double zero_slope);

/* Full cost refresh: wcost (src/cost/full_update.cuh:34-51), vcost
 * (:58-83) and the Fenwick-ordered exclusive prefix presum
 * (src/cost/fenwick.cuh:129-168).  Any of the outputs may be NULL. */
void ora_cost_refresh(const ora_grid *g, const float *capacity, const float *demand,
                      float *wcost, float *vcost, double *presum);
void ora_cost_refresh_slopes(const ora_grid *g, const float *capacity, const float *demand,
                             double slope, double zero_slope, double via_weight,
                             float *wcost, float *vcost, double *presum);

/* Fenwick prefix of one track: leaves w[0..len-1] (float), output
 * presum[0..len-1] (double), restating src/cost/fenwick.cuh:129-168. */
void ora_fenwick_track(const float *w, int len, double *presum);
/* Dense Hillis-Steele prefix of one track (src/cost/prefix_scan.cuh:74-114). */
void ora_dense_prefix_track(const float *w, int len, double *presum);

/* First-fit greedy batching (src/batching/generate_batches_cpu.hpp:7-73).
 * Nets are given in priority order k = 0..n-1.  points: 2D cells x*Y+y;
 * hsegs/vsegs: pairs (first, second) of 2D cells.  Writes the batch index of
 * every net to color[k] and returns the number of batches. */
int ora_batches(int X, int Y, int n,
                const int *pt_off, const int *pts,
                const int *hs_off, const int *hsegs,
                const int *vs_off, const int *vsegs,
                int limit, int sparsity, int *color);

/* L-shape tree DP of one net (src/routing/lshape_thread.cuh:18-171 with the
 * accumulation order of src/routing/lshape_warp.cuh:150-192).
 * nodes: 3D indices (fixed layer, or L for Steiner); parents: preorder,
 * parents[0] = -1.  Writes routes[0] = count and the segment words, returns
 * 0, or -1 if more than cap words would be written, or -2 if no finite
 * assignment exists. */
int ora_lshape_net(const ora_grid *g, const float *vcost, const double *presum,
                   int n, const int *nodes, const int *parents, int *routes, int cap);
/* As ora_lshape_net, also returning the optimal tree cost (root table minimum). */
int ora_lshape_net_cost(const ora_grid *g, const float *vcost, const double *presum,
                        int n, const int *nodes, const int *parents, int *routes, int cap, double *best);

/* Demand commit of one batch, wires then vias
 * (src/routing/demand.cuh:236-319).  routes_of[k] points at net k's words. */
void ora_commit_batch(const ora_grid *g, float *demand, int *timestamp,
                      int count, int *const *routes_of, int stamp, int K,
                      double *total_wirelength, long long *total_vias,
                      unsigned char *dirty_tracks);

/* Rip-up classification (src/routing/congestion.cuh:8-58, 114-136). */
void ora_ripup(const ora_grid *g, const float *capacity, const float *demand,
               int threshold, int net_count, int *const *routes_of,
               unsigned char *is_of_net);

/* One L-shape stage over a batch schedule (src/routing/lshape.cuh:153-179).
 * net_ids/node_off/nodes/parents are in execution order; batch_off has
 * nbatch+1 entries.  route_off[net] is the word offset of that net's route
 * area and route_cap[net] its size.  Before every batch the stage refreshes
 * the tracks whose demand changed (all tracks before the first batch).  `*stamp` is incremented once per batch.  Returns 0 or a negative
 * error from ora_lshape_net. */
int ora_lshape_stage(const ora_grid *g, const float *capacity, float *demand,
                     int *timestamp, float *wcost, float *vcost, double *presum,
                     int nbatch, const int *batch_off, const int *net_ids,
                     const int *node_off, const int *nodes, const int *parents,
                     const long long *route_off, const int *route_cap, int *routes,
                     int *stamp, double *total_wirelength, long long *total_vias,
                     int threads);

/* Wire costs rounded to integers in [0, 1e7] after the refresh, as the
 * Z-shape (GGR) build does (src/cost/full_update.cuh:47-50); off by default. */
void ora_set_integer_wire(int on);

/* ---- Z-shape DP (zshape.c), src/routing/zshape.cuh:42-404 ---- */

/* reset_nodes: planes 0-3 of every tree-node cell of the batch to infinity. */
void ora_zshape_reset(const ora_grid *g, int *map, const int *points, int count);
/* solve: points = [per-net record offsets (count), records]; map has
 * 6 * (L+1) * n * n ints and prev 16 * (L+1) * n * n (n = max(X, Y)); raw
 * receives each net's words at its record's offset; cost (may be NULL)
 * each net's DP optimum. */
void ora_zshape_solve(const ora_grid *g, const double *presum, const float *vcost, int *map, int *prev,
                      const int *points, int count, int *raw, int *cost);
/* Test probe of solve: also records where each edge's middle leg really
 * ran; ideal = 1 traces the middle leg on that layer instead of the layer
 * next to the last leg that the reference's traceback takes
 * (src/routing/zshape.cuh:188,201,311-315); quirk[k] = 1 when the two
 * differ on some edge of net k. */
void ora_zshape_probe(const ora_grid *g, const double *presum, const float *vcost, int *map, int *prev,
                      const int *points, int count, int *raw, int *cost, int ideal, int *quirk);
/* convert_routes: raw words of net k (at raw_off[k]) into the route words
 * of net_ids[k] (at route_off, capacity route_cap); returns the last error
 * code (0 ok, 1 storage, 2 no route, 3 bad coordinate). */
int ora_zshape_convert(const ora_grid *g, const int *raw, const int *raw_off, const int *net_ids, int count,
                       const long long *route_off, const int *route_cap, int *routes);

/* One Z-shape stage over a batch schedule (src/pipeline_ggr.hpp:114-127):
 * refresh, reset, solve, convert and commit per batch; stamps are
 * ++*stamp per batch.  Returns 0 or the last convert error code. */
int ora_zshape_stage(const ora_grid *g, const float *capacity, float *demand, int *timestamp, float *wcost,
                     float *vcost, double *presum, int nbatch, const int *batch_off, const int *net_ids,
                     const long long *rec_off, const int *records, const int *raw_cap,
                     const long long *route_off, const int *route_cap, int *routes, int *stamp,
                     double *total_wirelength, long long *total_vias, int threads);

/* ---- soft (log-sum-exp) DP over L-shape or Z-shape edges, FP64 (soft.c) ---- */

/* Smooth cost model: wcost, vcost per routing cell (IDX order) and their
 * partial derivatives in slope and zero_slope (any output may be NULL
 * except wcost and vcost); via_weight multiplies unit_via_cost. */
void ora_soft_costs(const ora_grid *g, const float *capacity, const float *demand, double slope, double zero_slope,
                    double via_weight, double *wcost, double *vcost, double *dw_ds, double *dw_dz, double *dv_ds,
                    double *dv_dz);

/* Soft DP value F of one net at temperature tau; with gw and gv non-NULL
 * adds dF/dwcost and dF/dvcost (IDX order) into them.  kind 0: the L-shape
 * DP's edges; kind 1: the Z-shape route family (straight, both L-shapes,
 * Z-shapes with the middle leg strictly inside the box). */
double ora_soft_net(const ora_grid *g, const double *wcost, const double *vcost, int n, const int *nodes,
                    const int *parents, double tau, int kind, double *gw, double *gv);
double ora_soft_lshape_net(const ora_grid *g, const double *wcost, const double *vcost, int n, const int *nodes,
                           const int *parents, double tau, double *gw, double *gv);

/* Sum of F over count nets (CSR node_off) at one demand snapshot; with gw
 * and gv non-NULL they receive dF/dwcost and dF/dvcost, and gparams (if
 * non-NULL) dF/d(slope, zero_slope, via_weight). */
double ora_soft_stage(const ora_grid *g, const float *capacity, const float *demand, double slope, double zero_slope,
                      double via_weight, double tau, int count, const int *node_off, const int *nodes,
                      const int *parents, double *gw, double *gv, double *gparams);
double ora_soft_stage_kind(const ora_grid *g, const float *capacity, const float *demand, double slope,
                           double zero_slope, double via_weight, double tau, int kind, int count, const int *node_off,
                           const int *nodes, const int *parents, double *gw, double *gv, double *gparams);

/* Congestion view (src/routing/congestion.cuh:60-112), X*Y cells in x*Y+y
 * order: worst demand - capacity over the layers, its exclusive prefix
 * along x (xsum) and along y (ysum) in the reference's summation order.
 * congested may be NULL. */
void ora_congestion_view(const ora_grid *g, const float *capacity, const float *demand, unsigned char *congested,
                         float *xsum, float *ysum);

/* ---- FLT DAG DP (flt.c), src/routing/flt_cost.cuh, flt_block.cuh ---- */

/* Packed FLT candidate: order-preserving float bits above a 32-bit code
 * (src/routing/flt_cost.cuh:25-40). */
unsigned long long ora_flt_pack(float cost, int code);

/* One FLT batch (src/routing/flt.cuh:142-159 without the commits): count
 * DAGs in execution order.  Net k's compact records are node_rec[4 * i] for
 * i in [node_off[k], node_off[k+1]) = {coordinate, depth, child slots,
 * parent_end} and edge_rec[2 * j] for j in [edge_off[k], edge_off[k+1]) =
 * {parent, node * 10 + slot} (net-local indices, include/flt_topology.hpp:15-16).
 * Routes of net_ids[k] go to routes + route_off[net] (capacity route_cap[net]
 * words).  layer_output (may be NULL) receives every node's chosen layer.
 * Returns 0, 1 (route storage overflow) or 2 (a node without a finite cost
 * was reached). */
int ora_flt_batch(const ora_grid *g, const double *presum, const float *vcost, int count, const int *net_ids,
                  const long long *node_off, const int *node_rec, const long long *edge_off, const int *edge_rec,
                  const long long *route_off, const int *route_cap, int *routes, int *layer_output);

/* One FLT stage (src/routing/flt.cuh:141-168): per batch, optionally remove
 * the nets' routes (K = -1 at ++*stamp; remove 1: phase-one routes, remove 2:
 * phase-two routes for nets with live2[net] set, phase-one for the others),
 * then at ++*stamp refresh, route and commit the phase-two routes.  Returns
 * the last batch error. */
int ora_flt_stage(const ora_grid *g, const float *capacity, float *demand, int *timestamp, float *wcost,
                  float *vcost, double *presum, int nbatch, const int *batch_off, const int *net_ids,
                  const long long *node_off, const int *node_rec, const long long *edge_off, const int *edge_rec,
                  const long long *route_off1, int *routes1, const long long *route_off2, const int *route_cap2,
                  int *routes2, int remove, const unsigned char *live2, int *stamp, double *total_wirelength,
                  long long *total_vias, int threads);

/* Evaluator score (resources/evaluator.cpp:458-675) of per-net line lists:
 * metal-0 pin vias, extra pairs, then subnet route pairs shifted up one metal
 * layer.  out: wirelength, via, overflow and total cost, rejected nets. */
int ora_score(int X, int Y, int M, int nets, const int *pin_off, const int *pins, const int *extra_off,
              const int *extra, const int *sub_off, const int *subs, const int *routes, const int *pin_acc_num,
              int per_pin, const int *layer_dir, const double *capacity, const double *overflow_cost,
              const int *x_len, const int *y_len, double unit_wire, double unit_via, double *out,
              unsigned short *demand);

/* The evaluator's completion test per net (resources/evaluator.cpp:527-538,
 * check_connectivity :677-749) over the same lines as ora_score; subnets with
 * pin_off2[id] >= 0 (routes2 may be NULL) take their phase-two words.  Pins
 * of net k: apin_off[k]..apin_off[k+1]; access points of pin p:
 * acc[acc_off[p]..acc_off[p+1]] as metal-layer cells.  status[k]: 0 complete,
 * 1 a line the evaluator rejects (open), 2 not connected. */
int ora_connectivity(int X, int Y, int M, int nets, const int *pin_off, const int *pins, const int *extra_off,
                     const int *extra, const int *sub_off, const int *subs, const int *routes, const int *pin_acc_num,
                     int per_pin, const int *routes2, const int *pin_off2, int per_pin2, const int *layer_dir,
                     const int *apin_off, const int *acc_off, const int *acc, unsigned char *status);

#ifdef __cplusplus
}
#endif
#endif
