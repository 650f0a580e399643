/*
 * grotor_amd: C ABI of the MI355X (gfx950) routing hot path.
 *
 * The reference routes through header-included C++/CUDA with no FFI; its
 * de-facto operator boundary is the set of host entry points its stage loops
 * call.  Each function here replaces one of them (cited file:line in the
 * reference tree) and keeps its argument meaning and error behaviour, with
 * plain pointers and sizes in place of RouterState / RoutingDeviceView.
 *
 * Conventions
 *  - Every function returns 0 (GR_OK) or a GR_E_* status; gr_last_error()
 *    returns the message of the last failure on the calling thread (the
 *    reference throws std::runtime_error("<op>: <hip error>"),
 *    include/cuda_memory.hpp:12-16).
 *  - Grid arrays exchanged through this ABI use the reference index
 *    IDX(l, x, y) = (l * X + x) * Y + y over the L routing layers
 *    (include/router_state.hpp:267).  Route words use the same encoding.
 *  - Host pointers are read or written synchronously; device state lives in
 *    an opaque context that owns all device memory
 *    (include/cuda_memory.hpp:17-53) and one HIP stream.
 *  - A context needs a HIP device; without one gr_ctx_create fails with
 *    GR_E_HIP.  There is no CPU fallback.
 */
#ifndef GROTOR_AMD_H
#define GROTOR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GR_OK = 0,
    GR_E_ARG = 1,      /* invalid argument (std::invalid_argument in the reference) */
    GR_E_HIP = 2,      /* HIP runtime failure, or no device */
    GR_E_CAPACITY = 3, /* route storage or batch limit exceeded (device assert) */
    GR_E_STATE = 4,    /* call out of order (e.g. DP before topology upload) */
    GR_E_DEVICE = 5,   /* device-side check failed (no finite DP solution, spin limit) */
    GR_E_IO = 6        /* file input/output */
};

/* Message of the last failure on this thread ("" if none). */
const char *gr_last_error(void);

/* Name and architecture of device 0, e.g. "AMD Instinct MI355X gfx950". */
int gr_device_name(char *buffer, int size);

typedef struct gr_ctx gr_ctx;

/* Routing grid of L routing layers (metal layer l + 1 of the CAP file).
 * Replaces the grid half of cudb::build_cuda_database
 * (src/preprocess.hpp:1625-1785). */
typedef struct gr_grid_desc {
    int L, X, Y;
    int DIR;                                /* direction of routing layer 0: 0 = horizontal */
    double unit_length_wire_cost;           /* db::unit_length_wire_cost */
    double unit_via_cost;                   /* db::unit_via_cost */
    const double *unit_length_short_costs;  /* [L], of metal layers 1..L */
    const int *x_edge_len;                  /* [X-1] */
    const int *y_edge_len;                  /* [Y-1] */
    const float *capacity;                  /* [L*X*Y], IDX order */
} gr_grid_desc;

/* Create / destroy a device context.  Demand and timestamps start at zero,
 * the first gr_cost_refresh computes every track
 * (graph::initialize_incremental_cost, src/cost/incremental_update.cuh:7-42). */
int gr_ctx_create(const gr_grid_desc *grid, gr_ctx **out);
int gr_ctx_destroy(gr_ctx *ctx);

/* Route storage: ROUTE_PER_PIN = 12 words per pin, net i at word
 * 12 * pin_acc_num[i] (include/router_state.hpp:31, src/preprocess.hpp:1699-1704).
 * Replaces the previous storage (cleared) and releases any phase-two storage. */
int gr_ctx_set_routes(gr_ctx *ctx, int net_count, const int *pin_acc_num);

/* Route words per pin for the next gr_ctx_set_routes: 12 by default
 * (ROUTE_PER_PIN, include/router_state.hpp:31); the Z-shape pipeline uses
 * 256 (src/pipeline_ggr.hpp:10). */
int gr_ctx_set_route_stride(gr_ctx *ctx, int words_per_pin);

/* Cost-model parameters of the refresh: overflow slopes (0.5 and 1.5,
 * src/cost/full_update.cuh:9-14), via weight (graph::via_weight = 1, :7) and
 * the Z-shape build's integer wire costs (:47-50, 0 or 1).  Marks every
 * track for refresh. */
int gr_ctx_set_cost_params(gr_ctx *ctx, double slope, double zero_slope, double via_weight,
                           int integer_wire);

/* Cost refresh: wcost, vcost and the Fenwick-ordered prefix presum of every
 * track whose demand changed since the last refresh
 * (graph::refresh_cost, src/cost/incremental_update.cuh:98-132). */
int gr_cost_refresh(gr_ctx *ctx);
/* Recompute every track regardless of dirty state (graph::update_cost plus a
 * full prefix rebuild, src/cost/full_update.cuh:93-104). */
int gr_cost_refresh_full(gr_ctx *ctx);
/* Copy the current costs to the host in IDX order; any pointer may be NULL.
 * wcost is recomputed from the current demand on request. */
int gr_cost_download(gr_ctx *ctx, float *wcost, float *vcost, double *presum);

int gr_demand_upload(gr_ctx *ctx, const float *demand);   /* IDX order; marks every track dirty */
int gr_demand_download(gr_ctx *ctx, float *demand);
int gr_timestamp_download(gr_ctx *ctx, int *timestamp);
/* Statistics accumulated by commits (RoutingStatistics, include/router_state.hpp:17-20). */
int gr_stats_download(gr_ctx *ctx, double *total_wirelength, long long *total_vias);

/* Evaluator score of the context's route words, as the reference evaluator
 * computes it from the .out text (NVR_DB::read_gr_solution and
 * update_stacked_via_counter, resources/evaluator.cpp:458-675).  Per design
 * net, in .out order: a via from metal 0 to metal 1 for each pin on metal 0
 * (pins as metal-layer cells (z * X + x) * Y + y), the extra_routes pairs,
 * then the route pairs of its subnets shifted up one metal layer.
 * layer_dir, capacity (metal-layer cells, binary64 as the evaluator reads
 * them) and overflow_cost cover the L + 1 metal layers.  out receives
 * {wirelength cost, via cost, overflow cost, total cost, nets with a line the
 * evaluator rejects, device ms}; demand, when not NULL, the evaluator's
 * 16-bit demand per metal-layer cell.  Subnets that have phase-two storage
 * (gr_ctx_set_routes_phase2, offset >= 0) are scored by their phase-two
 * words, as the reference's output prints them (src/output.hpp:98-106).
 * The connectivity flood fill is gr_check_connectivity. */
int gr_score(gr_ctx *ctx, int nets, const int *pin_off, const int *pins, const int *extra_off, const int *extra,
             const int *sub_off, const int *subs, const int *layer_dir, const double *capacity,
             const double *overflow_cost, double *out, unsigned short *demand);

/* The evaluator's completion test of every net over the same lines as
 * gr_score (NVR_DB::read_gr_solution :527-538 with check_connectivity,
 * resources/evaluator.cpp:677-749): the cells its lines mark, flood filled
 * from the first pin's marked access points along each routing layer's
 * direction and between adjacent layers, must reach an access point of
 * every other pin.  access_pin_off[nets+1]: each net's pins in NET-file
 * order; access_off[pins+1]: each pin's access points in access, as
 * metal-layer cells (z * X + x) * Y + y.  status[k] (may be NULL): 0
 * complete, 1 open (a line the evaluator rejects, counted in its "open
 * nets"), 2 not connected (the evaluator prints the net and stops).
 * counts: {open, not connected, incomplete = open + not connected}. */
int gr_check_connectivity(gr_ctx *ctx, int nets, const int *pin_off, const int *pins, const int *extra_off,
                          const int *extra, const int *sub_off, const int *subs, const int *layer_dir,
                          const int *access_pin_off, const int *access_off, const int *access, unsigned char *status,
                          long long *counts);

/* Exact first-fit greedy batching of n nets given in priority order
 * (cudb::generate_batches_rsmt, src/batching/generate_batches_cpu.hpp:7-73;
 * timed wrapper generate_batches_cpu_timed, src/batching/generate_batches.hpp:14-25).
 * points: 2D cells x*Y + y; hsegs / vsegs: (first, second) pairs of 2D cells.
 * Writes the batch index of net k to color[k] and the batch count.  Batches
 * are exactly the reference's lists: batch b holds the nets with color b in
 * priority order, no batch holds more than `limit` nets, and a net skips
 * batches that are full when it is placed.  Equals gr_batch_geometry followed
 * by gr_batch_colors. */
int gr_batch_plan(gr_ctx *ctx, int n, const int *pt_off, const int *pts,
                  const int *hs_off, const int *hsegs, const int *vs_off, const int *vsegs,
                  int limit, int sparsity, int *color, int *num_batches);

/* The same batches computed on the host, without a device or context: the
 * reference's CPU batching fallback (generate_batches_rsmt,
 * src/batching/generate_batches_cpu.hpp:7-73, timed by
 * generate_batches_cpu_timed, src/batching/generate_batches.hpp:14-25) for
 * an X x Y grid.  Same arguments, errors and result as gr_batch_plan. */
int gr_batch_plan_host(int X, int Y, int n, const int *pt_off, const int *pts,
                       const int *hs_off, const int *hsegs, const int *vs_off, const int *vsegs,
                       int limit, int sparsity, int *color, int *num_batches);

/* Validate and upload the batching geometry of n nets (priority order) into
 * the context, where it stays until the next call: the reference's GPU
 * batching uploads it before its batch interval starts
 * (src/batching/hiprt/grid_chunk.cuh:373-428). */
int gr_batch_geometry(gr_ctx *ctx, int n, const int *pt_off, const int *pts,
                      const int *hs_off, const int *hsegs, const int *vs_off, const int *vsegs);

/* Batch the uploaded geometry: the reference's batch interval
 * (generate_batch_ids, src/batching/hiprt/grid_chunk.cuh:519-524) plus the
 * download of the batch ids. */
int gr_batch_colors(gr_ctx *ctx, int limit, int sparsity, int *color, int *num_batches);

/* L-shape stage topology, uploaded once per stage (src/routing/lshape.cuh:126-138):
 * net_ids[k] is the subnet solved at stage slot k; its DFS-preorder tree is
 * nodes[node_off[k] .. node_off[k+1]) with parents (parents[first] = -1).
 * A node is a 3D index whose layer is the pin's fixed routing layer or L for
 * a Steiner node (src/preprocess.hpp:127-129). */
int gr_lshape_upload(gr_ctx *ctx, int count, const int *net_ids, const int *node_off,
                     const int *nodes, const int *parents);

/* DP of stage slots [begin, begin + count) against the current costs; writes
 * routes, does not commit (Lshape_route_warp launch, src/routing/lshape.cuh:168). */
int gr_lshape_batch(gr_ctx *ctx, int begin, int count, int stamp);

/* gr_lshape_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  Same
 * arguments and results as gr_lshape_batch (route words identical); the
 * stage topology and route storage must already be in the context.  The DP
 * runs in its persistent form with its work-queue counters, or, when a tree's
 * tables do not fit in LDS, in the size-class launches with those tables,
 * in one library-owned device array per GPU of 32 MiB (4 KiB of counters,
 * the rest tables), held until the process ends.  A batch whose global DP
 * tables exceed it fails with GR_E_CAPACITY.  Every call on a GPU shares that
 * array: calls that overlap on one GPU (from several threads or contexts)
 * race; calls on one context's stream do not. */
int gr_lshape_batch_noalloc(gr_ctx *ctx, int begin, int count, int stamp);

/* Commit the routes of stage slots [begin, begin + count): wires, then vias
 * (graph::commit_batch_demand, src/routing/demand.cuh:313-319). */
int gr_commit_batch(gr_ctx *ctx, int begin, int count, int stamp, int K);

/* The whole L-shape batch loop (src/routing/lshape.cuh:148-180): for each
 * batch b, stamp = first_stamp + b, refresh costs, solve, commit.
 * batch_off has nbatch + 1 entries over the uploaded stage slots.
 * mode: 0 = plain form, the reference's execution (src/routing/lshape.cuh:148-180):
 *           per batch one DP launch, one net per wavefront in batch order,
 *           every net's DP tables in global memory;
 *       1 = launch form: per batch the nets in two tree-size classes (the
 *           large trees on a second stream), tables in LDS, nets per
 *           wavefront from gr_lshape_tune;
 *       2 = persistent form: per batch one DP launch whose resident
 *           workgroups drain the batch's nets from a device work queue,
 *           tallest trees first, tables in LDS; the whole loop captured once
 *           into a HIP graph and replayed (the capture is reused while the
 *           stage topology, batch offsets, first stamp and nets per wavefront
 *           stay the same; capture happens before the timed interval);
 *       other values are rejected with GR_E_ARG.  Results are identical in
 *       every mode.
 * *elapsed_ms receives the synchronized interval (RoutingDagTimer,
 * include/router_state.hpp:238-265). */
int gr_lshape_run(gr_ctx *ctx, int nbatch, const int *batch_off, int first_stamp,
                  int mode, double *elapsed_ms);

/* Wait for all device work (the cudaDeviceSynchronize that opens the
 * reference's timed intervals, src/batching/generate_batches.hpp:15,
 * include/router_state.hpp:241-249). */
int gr_ctx_synchronize(gr_ctx *ctx);

/* Start and end of the last synchronized batch-loop interval (gr_lshape_run,
 * gr_flt_run or gr_zshape_run) as std::chrono::steady_clock nanoseconds since
 * its epoch, so a caller can print them against its own clock origin as the
 * reference's DAG_TIMING line does (start_us / end_us,
 * include/router_state.hpp:252-264).  Zero before the first run. */
int gr_ctx_last_interval(gr_ctx *ctx, long long *start_ns, long long *end_ns);

/* Nets per wavefront of the L-shape DP launches (1, 2, 4 or 8; each net gets
 * 64 / n lanes and its own tables) for the two tree size classes of a batch:
 * up to 16 nodes, and larger (per_wave[2]).  Results do not depend on it.
 * No reference counterpart (the reference runs one net per 32-lane warp). */
int gr_ctx_set_nets_per_wave(gr_ctx *ctx, const int *per_wave);

/* Autotuner: times every size class's DP launches over the given batches of
 * the uploaded stage at the current costs for 1, 2, 4 and 8 nets per
 * wavefront, keeps the fastest per class in the context and returns it in
 * per_wave[2]; ms (may be NULL) receives [class * 4 + candidate] times
 * (-1 where a candidate's tables do not fit in LDS).  The route words it
 * writes are rewritten by the next stage run.  No reference counterpart. */
int gr_lshape_tune(gr_ctx *ctx, int nbatch, const int *batch_off, int *per_wave, double *ms);

/* Group shapes of the persistent form (gr_lshape_run mode 2; no reference
 * counterpart): per batch the trees of at most `tiny` nodes are drained
 * from a work queue of their own with DP tables sized for them, beside the
 * rest of the small class and the large trees (those two groups take one
 * net per wavefront, whatever gr_ctx_set_nets_per_wave or gr_lshape_tune
 * chose for the launch form).  Times the persistent DP of
 * the uploaded batches for tiny = 1 (no group of its own), 2, 4 and 8 and
 * keeps the fastest; *tiny_nodes receives it, ms[4] (may be NULL) the
 * times.  Routes are the same for every choice; route words are
 * overwritten, so call it before the stage run.  gr_ctx_set_tiny_tree sets
 * the bound by hand (1..16, default 4). */
int gr_lshape_tune_groups(gr_ctx *ctx, int nbatch, const int *batch_off, int *tiny_nodes, double *ms);
int gr_ctx_set_tiny_tree(gr_ctx *ctx, int nodes);

/* Nets per wavefront of the persistent form's tiny-tree group (no reference
 * counterpart): times the persistent DP of the uploaded batches with 1, 2,
 * 4 and 8 of its trees per wavefront (each then owns 64, 32, 16 or 8 lanes)
 * at the current tiny-tree bound and keeps the fastest; *per_wave receives
 * it, ms[4] (may be NULL) the times.  Routes are the same for every choice;
 * route words are overwritten, so call it before the stage run.
 * gr_ctx_set_tiny_per_wave sets it by hand (0, the default, takes the
 * small class's value from gr_lshape_tune). */
int gr_lshape_tune_tiny(gr_ctx *ctx, int nbatch, const int *batch_off, int *per_wave, double *ms);
int gr_ctx_set_tiny_per_wave(gr_ctx *ctx, int per_wave);

/* DAG classes of the FLT DP (gr_flt_run modes 1 and 2; no reference
 * counterpart): per batch the DAGs whose LDS share exceeds `bytes` form one
 * class and the rest another, each launched with shares sized by its own
 * largest DAG (the large class on a second stream beside the other), so
 * the small DAGs are not held to the batch's largest share.  0 puts every
 * DAG in one class.  Default 20480; takes effect at the next
 * gr_flt_upload.  Routes are the same for every split. */
int gr_ctx_set_flt_split(gr_ctx *ctx, int bytes);

/* Reset the routing state to that of a fresh context: demand, timestamps,
 * statistics and route words cleared, phase-two route storage released,
 * every track marked for refresh (the DAG-throughput benchmark's reset,
 * src/experiments/lshape_benchmark.cuh:84-97). */
int gr_ctx_reset(gr_ctx *ctx);

/* Kernel classes timed by gr_profile_*. */
enum {
    GR_K_REFRESH = 0,      /* cost refresh */
    GR_K_LSHAPE = 1,       /* L-shape DP */
    GR_K_COMMIT_WIRES = 2,
    GR_K_COMMIT_VIAS = 3,
    GR_K_BATCH_MARKS = 4,  /* batching: mark expansion and per-cell lists */
    GR_K_BATCH_COLOR = 5,  /* batching: in-order colouring */
    GR_K_RIPUP = 6,
    GR_K_ZSHAPE = 7,       /* Z-shape DP (reset, solve, convert) */
    GR_K_FLT_EDGE = 8,     /* FLT layer-change search per DAG edge */
    GR_K_FLT = 9,          /* FLT DAG DP (solve and trace) */
    GR_K_COUNT = 10
};

/* Start timing every kernel launch of this context with HIP events on its
 * stream; gr_profile_end returns, per kernel class, the summed device time
 * (ms), the launch count and the algorithmic HBM bytes of those launches.
 * No reference counterpart (the reference prints stage wall times only). */
int gr_profile_begin(gr_ctx *ctx);
/* Restrict the events of gr_profile_begin to the kernel classes whose bit
 * (1 << GR_K_*) is set in kind_mask (all classes by default); the others run
 * without events and report nothing.  Fewer events, less host overhead. */
int gr_profile_select(gr_ctx *ctx, unsigned kind_mask);
int gr_profile_end(gr_ctx *ctx, double *ms, long long *launches, double *bytes, int kinds);

/* Copy `words` route words (from word 0) to the host (RouterState::download_routes,
 * include/router_state.hpp:215-222). */
int gr_routes_download(gr_ctx *ctx, int *routes, long long words);

/* Rip-up classification over all nets of the route storage
 * (graph::ripup, src/routing/congestion.cuh:114-136): is_of_net[i] = 1 if net
 * i's route crosses more than 10 edges with capacity + threshold <= demand.
 * A net an FLT stage has routed is classified by its phase-two route, as
 * mark_overflow_nets reads routes_phase2 for nets with a rip-up flag
 * (:38-45). */
int gr_ripup(gr_ctx *ctx, int threshold, unsigned char *is_of_net);

/* ---- Z-shape (GGR) stage on the device (src/pipeline_ggr.hpp:88-127) ----
 * The route storage must hold 256 words per pin (gr_ctx_set_route_stride
 * before gr_ctx_set_routes) and the cost refresh the Z-shape build's
 * integer wire costs (gr_ctx_set_cost_params). */

/* Upload a plan from gr_design_zshape_plan (validated on the host). */
int gr_zshape_upload(gr_ctx *ctx, int nbatch, const int *batch_off, const int *net_ids,
                     const long long *rec_off, const int *records, const int *raw_cap);

/* Run the uploaded Z-shape batches: per batch b (stamp first_stamp + b)
 * refresh, reset_nodes, solve, convert_routes and the wire-then-via commit;
 * *elapsed_ms = the synchronized interval (DAG_TIMING). */
int gr_zshape_run(gr_ctx *ctx, int first_stamp, double *elapsed_ms);

/* ---- FLT stage on the device (src/routing/flt.cuh:75-182) ----
 * Needs at most 9 routing layers (the reference's MAX_LAYER = 10). */

/* Phase-two route storage (src/pipeline.hpp:96-108): pin_off2[net] is the
 * net's offset in pins among the rerouted nets (pin_acc_num_phase2), -1 for
 * the others; pins2 their total; words_per_pin = ROUTE_PER_PIN_PHASE2 (100,
 * include/router_state.hpp:33).  Net i's words start at pin_off2[i] *
 * words_per_pin, capacity words_per_pin * its pin count.  Cleared to zero. */
int gr_ctx_set_routes_phase2(gr_ctx *ctx, const int *pin_off2, long long pins2, int words_per_pin);

/* Phase-two storage for a further rip-up stage (reserve_reroute_storage,
 * src/pipeline_cugr2.hpp:15-62): nets of net_ids without phase-two words get
 * them appended in the given order; nets that have them keep their offsets
 * and their words (and whether an FLT stage routed them); the stride
 * becomes max(current, words_per_pin).  Creates the storage if there is
 * none.  pin_off2 (may be NULL) receives every net's offset in pins (-1:
 * none), *pins2 (may be NULL) the pin total. */
int gr_ctx_reserve_routes_phase2(gr_ctx *ctx, int count, const int *net_ids, int words_per_pin, int *pin_off2,
                                 long long *pins2);

/* Copy `words` phase-two route words (from word 0) to the host
 * (RouterState::download_routes(true), include/router_state.hpp:215-222). */
int gr_routes2_download(gr_ctx *ctx, int *routes, long long words);

/* The congestion view the FLT DAG builder reads (extract_congestionView,
 * _xsum, _ysum, src/routing/congestion.cuh:60-112), X*Y cells in x*Y+y
 * order: congested = worst demand - capacity over the layers > 0; xsum and
 * ysum its exclusive prefix along x and along y.  Any output may be NULL. */
int gr_congestion_view(gr_ctx *ctx, unsigned char *congested, float *xsum, float *ysum);

/* Upload an FLT stage in execution order (flt::Plans, include/flt_topology.hpp:87-129,
 * and Workspace::upload, src/routing/flt_workspace.cuh:156-173): batch b
 * is nets [batch_off[b], batch_off[b+1]); net k (id net_ids[k]) has node
 * records nodes[4*i..] = {coordinate, depth, child slots, parent_end} for i
 * in [node_off[k], node_off[k+1]) and edge records edges[2*j..] = {parent,
 * node * 10 + slot} for j in [edge_off[k], edge_off[k+1]), net-local
 * indices, as flt::Topology::emit writes them (:52-70).  Every net needs
 * phase-two storage, exactly one root (no parents, depth 0), parents
 * shallower than their children, at most 6 child slots per node and
 * straight edges; violations return GR_E_ARG. */
int gr_flt_upload(gr_ctx *ctx, int nbatch, const int *batch_off, const int *net_ids,
                  const long long *node_off, const int *nodes, const long long *edge_off,
                  const int *edges);

/* Run the uploaded FLT batches (the loop of src/routing/flt.cuh:141-168):
 * per batch, the nets' routes are removed (commit K = -1 at the next stamp)
 * when `remove` is 1 (their phase-one routes: remove_phase_one) or 2 (each
 * net's current route: its phase-two words if an earlier FLT stage routed
 * it, else its phase-one words; a repeated rip-up stage); then at the next
 * stamp the cost
 * refresh, precompute_cost_edge, the DAG DP and the commit of the
 * phase-two routes.  Stamps start at first_stamp and advance by one per
 * commit.  mode: 0 = plain form, the reference's execution
 *           (route_flt_block, src/routing/flt.cuh:158): per batch one DP
 *           launch, one DAG per wavefront, the DP tables in global memory;
 *       1 = launch form: each DAG's tables in LDS (DAGs whose tables do not
 *           fit keep the global tables), DAGs per wavefront from gr_flt_tune;
 *       2 = persistent form: as 1, with per batch one DP launch whose
 *           resident workgroups drain the batch's DAGs from a device work
 *           queue, deepest first, and the whole loop captured once into a
 *           HIP graph and replayed (capture before the timed interval, reused
 *           while the uploaded stage, its storage, first stamp and DAGs per
 *           wavefront stay the same);
 *       other values are rejected with GR_E_ARG.  Results are identical in
 *       every mode.
 * *elapsed_ms = the synchronized interval (DAG_TIMING). */
int gr_flt_run(gr_ctx *ctx, int first_stamp, int remove, int mode, double *elapsed_ms);

/* DAGs per wavefront of the FLT DP (1, 2, 4 or 8; each DAG gets 64 / n
 * lanes) for batches of fewer than 512 DAGs (per_wave[0]) and of at least
 * 512 (per_wave[1]).  Results do not depend on it.  No reference
 * counterpart (the reference runs one DAG per 128-thread block). */
int gr_ctx_set_flt_per_wave(gr_ctx *ctx, const int *per_wave);

/* Autotuner of the FLT DP: per uploaded batch, the layer-change search and
 * then the DAG DP once per candidate (1, 2, 4, 8 DAGs per wavefront) at the
 * current costs; keeps the fastest per batch class (as above) and returns it
 * in per_wave[2]; ms (may be NULL) receives [class * 4 + candidate] summed
 * DP times (-1 for a class without batches).  Phase-two route storage is
 * cleared afterwards.  No reference counterpart. */
int gr_flt_tune(gr_ctx *ctx, int *per_wave, double *ms);

/* ---- soft (log-sum-exp) DP over L-shape or Z-shape edges ----
 * No counterpart in the reference: a differentiable relaxation of the
 * L-shape DP above, or of the Z-shape route family (gr_ctx_set_soft_kind),
 * in which every minimum becomes smin_tau(x) = -tau log sum exp(-x/tau),
 * over the smooth form of the cost model (src/cost/full_update.cuh:9-83
 * with exact exp) at the context's current demand, or over given costs.
 * Its gradient is the expected wire and via usage of every cell, and the
 * derivative in the cost-model parameters. */

/* Soft DP of the uploaded stage slots [begin, begin + count) with overflow
 * slopes (slope, zero_slope; 0.5 and 1.5 in the reference), via weight
 * (1.0) and temperature tau > 0, in FP32 (precision 32) or FP64 (64)
 * tables.  *total = sum of the nets' soft values; per_net (may be NULL)
 * receives each value.  The tables are checkpointed for gr_soft_backward
 * until the next forward or stage upload.  *elapsed_ms: device time. */
int gr_soft_forward(gr_ctx *ctx, int begin, int count, double slope, double zero_slope,
                    double via_weight, double tau, int precision, double *total,
                    double *per_net, double *elapsed_ms);

/* The same soft DP at caller-given costs: wcost and vcost per routing cell
 * (IDX order, host memory); the wire prefix sums are taken on the device.
 * F is then a function of these arrays, its gradient gr_soft_edge_grads. */
int gr_soft_forward_costs(gr_ctx *ctx, int begin, int count, const double *wcost, const double *vcost,
                          double tau, int precision, double *total, double *per_net, double *elapsed_ms);

/* Edges of the soft DP for the following forward passes: 0 (default) the
 * L-shape DP's (straight, or two legs with the corner vias strictly between
 * their layers); 1 the route family of the Z-shape DP
 * (src/routing/zshape.cuh:62-288): straight, both L-shapes and the
 * Z-shapes with the middle leg strictly inside the box, a full via stack
 * at each bend.  At most 10 routing layers. */
int gr_ctx_set_soft_kind(gr_ctx *ctx, int kind);

/* Adjoint of the last gr_soft_forward: grad_params[0..2] = d total /
 * d(slope, zero_slope, via_weight) (zero after gr_soft_forward_costs, whose
 * costs have no parameters); per-cell gradients are kept for
 * gr_soft_edge_grads.  *elapsed_ms: device time. */
int gr_soft_backward(gr_ctx *ctx, double *grad_params, double *elapsed_ms);

/* d total / d wcost and d total / d vcost per routing cell (IDX order):
 * the expected use of each wire edge and via landing.  Either may be NULL. */
int gr_soft_edge_grads(gr_ctx *ctx, double *grad_wcost, double *grad_vcost);

/* ---- host front end and pipeline (src/main.cu, src/pipeline.hpp) ---- */

typedef struct gr_design gr_design;

/* Read reference-format CAP/NET files (db::read, src/input.hpp:247-419). */
int gr_design_read(const char *cap_path, const char *net_path, gr_design **out);
/* Synthetic instance of workload configuration 1..6; net_scale multiplies
 * the configuration's net count, grid_side (0 = default) overrides the side. */
int gr_design_generate(int config, unsigned long long seed, double net_scale, int grid_side,
                       gr_design **out);
int gr_design_write(const gr_design *design, const char *cap_path, const char *net_path);
int gr_design_destroy(gr_design *design);

/* Access-point selection, trees, subnet split and batching geometry
 * (net::init src/input.hpp:38-243; prepare_routing_nets
 * src/preprocess.hpp:1448-1623; net::construct_rsmt src/preprocess.hpp:72-150). */
int gr_design_prepare(gr_design *design, int threads);

/* Stage topology of the prepared design in priority order (subnets sorted by
 * hpwl descending, src/routing/lshape.cuh:53-55).  Arrays are owned by the
 * design and stay valid until it is destroyed or prepared again. */
typedef struct gr_topology {
    int L, X, Y, DIR;
    int nets;               /* stage slots */
    int subnets;            /* routed subnets (route storage units) */
    const int *order;       /* [nets] subnet id per priority slot */
    const int *node_off;    /* [nets+1] */
    const int *nodes;       /* DFS preorder 3D nodes */
    const int *parents;
    const int *pt_off, *pts;    /* batching points (2D) */
    const int *hs_off, *hsegs;  /* pairs */
    const int *vs_off, *vsegs;
    const int *pin_acc_num;     /* [subnets+1] */
    const float *capacity;      /* [L*X*Y] routing layers, IDX order */
    const double *unit_length_short_costs; /* [L] */
    const int *x_edge_len, *y_edge_len;
    double unit_length_wire_cost, unit_via_cost;
} gr_topology;
int gr_design_topology(gr_design *design, gr_topology *out);

/* FLT stage input (prepare_detour_topology, src/routing/flt.cuh:48-73):
 * the nets flagged in reroute[subnet] in priority order (hpwl descending,
 * std::sort over increasing ids), one DAG per net built by
 * net::generate_detours_reconstruct (src/preprocess.hpp:769-1436) from the
 * congestion view's prefix sums (gr_congestion_view; NULL = uncongested),
 * their batching geometry (DAG points, tree segments) and the phase-two
 * route offsets of src/pipeline.hpp:96-101.  Arrays are owned by the design. */
typedef struct gr_flt_stage {
    int nets;
    const int *order;            /* [nets] subnet ids */
    const int *pt_off, *pts;     /* batching points (2D) */
    const int *hs_off, *hsegs;
    const int *vs_off, *vsegs;
    const int *pin_off2;         /* [subnets] offset in pins, -1 if not rerouted */
    long long pins2;             /* pins of the rerouted nets */
    long long nodes, edges;      /* DAG totals */
} gr_flt_stage;
int gr_design_flt_stage(gr_design *design, const unsigned char *reroute, const float *xsum,
                        const float *ysum, int threads, gr_flt_stage *out);

/* Execution plan of the FLT stage from a batch colouring of its nets
 * (color[k] for priority slot k, gr_batch_plan with limit 300000): batches
 * in reverse order, empty ones skipped (flt::Plans, include/flt_topology.hpp:92-127),
 * with the DAG records in the layout gr_flt_upload takes. */
typedef struct gr_flt_plan {
    int batches;
    const int *batch_off;        /* [batches+1] */
    const int *net_ids;          /* subnet ids, execution order */
    const long long *node_off;   /* [nets+1] */
    const int *nodes;            /* 4 ints per node */
    const long long *edge_off;   /* [nets+1] */
    const int *edges;            /* 2 ints per edge */
} gr_flt_plan;
int gr_design_flt_plan(gr_design *design, const int *color, int num_batches, gr_flt_plan *out);

/* The FLT DAG of one tree (net::generate_detours_reconstruct,
 * src/preprocess.hpp:769-1436, with `mode` of src/pipeline.hpp:88-89):
 * positions are 3D (pin layer, or L for Steiner nodes), adjacency in CSR,
 * xsum/ysum the congestion view (NULL = uncongested).  Capacities in, counts
 * out: *nnodes node records of 4 ints, *nedges edge records of 2 ints (as
 * gr_flt_upload takes them), *npoints batching cells. */
int gr_flt_dag(int X, int Y, int L, int mode, int n, const int *positions, const int *adj_off, const int *adj,
               const float *xsum, const float *ysum, int *nodes, int *nnodes, int *edges, int *nedges,
               int *points, int *npoints);

/* Metal-layer grid of a design as read from its CAP file (src/input.hpp:275-324):
 * layer count, then, where not NULL, each layer's direction, the capacity per
 * metal-layer cell (z * X + x) * Y + y and each layer's overflow unit cost. */
int gr_design_metal_grid(const gr_design *design, int *metal_layers, int *layer_dir, float *capacity,
                         double *overflow_cost);

/* Per design net of a prepared design: selected pins (3D metal indices),
 * extra_routes pairs and routed subnet ids, as CSR arrays owned by the design
 * (db::net::pins / extra_routes / subnets, src/input.hpp:15-24). */
int gr_design_nets(gr_design *design, int *count, const int **pin_off, const int **pins,
                   const int **extra_off, const int **extra, const int **subnet_off, const int **subnets);

/* Access points of every design net's pins as read from the NET file
 * (NVR_DB::read_nets, resources/evaluator.cpp:407-441; src/input.hpp:385-419):
 * per net its pins in file order (pin_off[count+1]), per pin its access
 * points (access_off[pins+1]) as metal-layer cells (z * X + x) * Y + y;
 * arrays owned by the design. */
int gr_design_access(gr_design *design, int *count, const int **pin_off, const int **access_off,
                     const int **access);

/* Batching geometry of one tree (net::construct_rsmt, src/preprocess.hpp:94-126):
 * positions are 2D cells x*Y + y, adjacency in CSR.  Capacities in, counts
 * out (*npoints points, *nhsegs / *nvsegs segment pairs). */
int gr_tree_geometry(int Y, int n, const int *positions, const int *adj_off, const int *adj,
                     int *points, int *npoints, int *hsegs, int *nhsegs, int *vsegs, int *nvsegs);

/* ---- Z-shape (GGR) stage, src/pipeline_ggr.hpp ---- */

/* Priority order of the Z-shape stage (hpwl descending, larger subnet id
 * first on ties, src/pipeline_ggr.hpp:31-33) and the batching geometry in
 * that order, for gr_batch_plan.  Arrays are owned by the design. */
typedef struct gr_zshape_stage {
    int nets;
    const int *order;
    const int *pt_off, *pts, *hs_off, *hsegs, *vs_off, *vsegs;
} gr_zshape_stage;
int gr_design_zshape_stage(gr_design *design, gr_zshape_stage *out);

/* Execution plan from the stage's batch colors: every RSMT batch refined
 * into bounding-box-disjoint batches of at most max_batch nets (250 in the
 * reference), the list reversed (src/pipeline_ggr.hpp:49-59,
 * generate_bbox_batches src/routing/zshape_topology.hpp:7-58), and each
 * net's tree encoded for the kernel (encode_batch, :59-121). */
typedef struct gr_zshape_plan {
    int batches;
    const int *batch_off;        /* [batches+1] over net_ids */
    const int *net_ids;          /* subnet per execution position */
    const long long *rec_off;    /* [batches+1] into records */
    const int *records;          /* per batch: per-net record offsets, then the records */
    const int *raw_cap;          /* raw route words per net (pins * 256) */
} gr_zshape_plan;
int gr_design_zshape_plan(gr_design *design, const int *color, int num_batches, int max_batch,
                          gr_zshape_plan *out);

/* Host post-processing of downloaded Z-shape routes (IDX words,
 * words_per_pin per pin): layer promotion of congested segments and
 * canonical merged segments (promote_congested_segments /
 * canonicalize_routes, src/routing/zshape_topology.hpp:122-232). */
int gr_design_zshape_finish(gr_design *design, int *routes, int words_per_pin, int *promoted);

/* Write the .out solution for the prepared design given its route words
 * (graph::output_nets, src/output.hpp:79-112). */
int gr_design_write_output(const gr_design *design, const int *routes, const char *out_path);
/* As gr_design_write_output after an FLT stage: subnets with pin_off2[id]
 * >= 0 print their phase-two words (routes2 at pin_off2[id] *
 * words_per_pin2), the others their phase-one words (graph::output_nets,
 * src/output.hpp:98-106). */
int gr_design_write_output2(const gr_design *design, const int *routes, const int *routes2, const int *pin_off2,
                            int words_per_pin2, const char *out_path);

#ifdef __cplusplus
}
#endif
#endif
