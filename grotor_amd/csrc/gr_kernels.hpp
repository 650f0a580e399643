// Kernel declarations and launch-argument structures of the gfx950 hot path.
#pragma once

#include "gr_device.hpp"

namespace grotor {

constexpr int kRefreshThreads = 256;
constexpr int kWave = 64;

struct CostParams {
    double slope = 0.5;       // of_cost_scaled, capacity > 0 (src/cost/full_update.cuh:11)
    double zero_slope = 1.5;  // of_cost_scaled, capacity == 0 (src/cost/full_update.cuh:12)
    double via_weight = 1.0;  // multiplies unit_via_cost (src/cost/full_update.cuh:82)
    int integer_wire = 0;     // Z-shape build: wcost rounded to an integer in [0, 1e7] (src/cost/full_update.cuh:47-50)
};

// Stage statistics, striped over kStatStripes cache lines: a wavefront adds
// into stripe (wave id mod kStatStripes), so thousands of nets per launch do
// not serialize on one address.  The host sums the stripes; every addend is
// an integer (edge lengths, counts), so the double sums are exact in any order.
struct alignas(128) Stats {
    double total_wirelength;
    unsigned long long total_vias;
    unsigned long long wire_edges;  // unit wire edges committed (for byte accounting)
};
constexpr int kStatStripes = 64;
constexpr int kCounterStride = 16;  // 64-bit counters per 128-byte stripe

// Error codes written by kernels into the context's device error word.
enum DeviceError : int {
    kDevOk = 0,
    kDevRouteOverflow = 1,
    kDevNoSolution = 2,
    kDevSpinLimit = 3,
    kDevColorOverflow = 4,
    kDevBadCoordinate = 5,
};

__global__ void refresh_tracks_kernel(DeviceGrid g, const float* capacity, const float* demand, float* vcost,
                                      double* presum, float* wcost_out, unsigned* track_dirty, int force,
                                      CostParams cp, unsigned long long* refreshed);

struct RouteStore {
    int* routes;
    const int* pin_acc_num;
    int per_pin;  // route words per pin: 12 (include/router_state.hpp:31), 256 for Z-shape stages
};

struct CommitArgs {
    DeviceGrid g;
    float* demand;
    int* timestamp;
    unsigned* track_dirty;
    Stats* stats;
    RouteStore rs;
    const int* net_ids;  // stage slots -> net id
    const int* slots;    // optional permutation of stage slots (nullptr = identity)
    int begin, count, stamp, K;
    RouteStore rs2;              // phase-two words, read for nets marked in live2
    const unsigned char* live2;  // nullptr: every net reads rs
};
// W lanes per net: 64 (a wavefront) or 16 (four nets per wavefront)
template <int W>
__global__ void commit_wires_kernel(CommitArgs a);
template <int W>
__global__ void commit_vias_kernel(CommitArgs a);

struct LshapeArgs {
    DeviceGrid g;
    const float* vcost;
    const double* presum;
    RouteStore rs;
    const int4* desc;  // per execution position: node begin, nodes, first route word, route capacity
    const int* nodes;
    const int* parents;
    const int* lvl;    // per net position q: node of the level order | its height << 16
    const int* link;   // per node: (first child + 1) | (next smaller sibling + 1) << 16
    int begin, count;
    int nmax;          // node capacity of one net's tables
    int per_wave;      // nets per wavefront: 1, 2, 4 or 8 (each gets 64 / per_wave lanes)
    double* gscratch;  // global scratch when the tree does not fit in LDS
    int* error;
    int* queue;        // persistent form: the batch's kDpQueues work-queue counters (nullptr: one workgroup per P nets)
    int cus;           // compute units, for the persistent grid
};
// Bytes of one net's DP tables for trees of up to nmax nodes (0 if L is not
// supported), and the launch of `blocks` workgroups of a.per_wave nets each
// with the tables in LDS or in a.gscratch, or, with a.queue set, of the
// persistent kernel that drains the nets with at most `blocks` resident
// workgroups (false if L is not supported).
size_t lshape_scratch_bytes(int L, int nmax);
constexpr int kDpQueues = 32;       // persistent DP: work-queue counters per batch
constexpr int kDpQueueStride = 32;  // ints between counters (one 128-byte line)
constexpr int kTreeClasses = 2;  // L-shape DP launches per batch: trees up to 16 nodes, larger
bool launch_lshape_dp(const LshapeArgs& a, int blocks, bool lds, hipStream_t stream);

struct RipupArgs {
    DeviceGrid g;
    const float* capacity;
    const float* demand;
    int* of_sum;
    int threshold;
};
__global__ void overflow_prefix_kernel(RipupArgs a);
__global__ void overflow_nets_kernel(DeviceGrid g, const int* of_sum, RouteStore rs, RouteStore rs2,
                                     const unsigned char* live2, int net_count, unsigned char* is_of_net);

// Batching
struct BatchGeom {
    int n, X, Y, sparsity;
    const int* pt_off;
    const int* pts;
    const int* hs_off;
    const int* hsegs;
    const int* vs_off;
    const int* vsegs;
};
__global__ void count_marks_kernel(BatchGeom b, int* mark_count, int* cand_count);
__global__ void emit_marks_kernel(BatchGeom b, const int* mark_off, unsigned* cell, unsigned* tag);
// 1 where sorted mark i starts a new (cell, net) run, 0 for i >= m: the
// input of the entry-index scan
struct RunHead {
    const unsigned* cell;
    const unsigned* tag;
    int m;
    __host__ __device__ int operator()(int i) const {
        if (i >= m) return 0;
        return (i == 0 || cell[i] != cell[i - 1] || (tag[i] >> 1) != (tag[i - 1] >> 1)) ? 1 : 0;
    }
};
constexpr int kEntryReader = 1;  // the cell is one of the entry's net's points
constexpr int kEntryFirst = 2;   // first entry of its cell's list
__global__ void build_entries_kernel(const unsigned* cell, const unsigned* tag, const int* eidx, int m, int* list_net,
                                     int* einfo, int* marker);
__global__ void reader_slots_kernel(const int* u, const int* list_net, const int* einfo, const int* pt_off, int* rcount,
                                    int* slot_of);
__global__ void reader_records_kernel(const int* u, const int* einfo, const int* prev_marker, const int* slot_of,
                                      int4* rec);
constexpr int kColorsPerWord = 63;  // prefix-mask word: 63 color bits + a ready bit
constexpr int kMaxMaskWords = 32;   // colors 0..2015
struct ColorArgs {
    int n;
    const int* pt_off;    // per net: first record slot
    const int* rcount;    // per net: reader records
    const int4* rec;      // per slot: (first entry read, own entry, previous reader's slot or -1, 0)
    const int* list_net;  // per entry: its net
    unsigned long long* pm;  // per slot: mask_words prefix-mask words
    int mask_words;
    int* color;           // per net, -1 until set
    int* next;            // kColorQueues dequeue counters, kColorQueueStride ints apart, zeroed
    int first;            // first net to color
    int* error;
    long long spin_limit;
    int nclosed;          // full batches: color closed_color[i] for nets >= closed_from[i]
    const int* closed_color;
    const int* closed_from;
};
// Wave w takes nets first + q + kColorQueues * c, q = w % kColorQueues, c from
// queue q's counter, one net per dequeue; the queues spread the atomics.
constexpr int kColorQueues = 16;
constexpr int kColorQueueStride = 32;  // one 128-byte line per counter
__global__ void color_kernel(ColorArgs a);
__global__ void color_hist_kernel(const int* color, int first, int n, int* counts);

// Z-shape
struct ZshapeArgs {
    DeviceGrid g;
    const double* presum;
    const float* vcost;
    int* map;    // 6 planes of layers * n * n
    int* prev;   // 16 planes
    const int* points;   // batch records: per-net record offsets, then records
    int count;
    int* raw;            // raw route words of the batch
    const int* raw_cap;  // raw words per net
    int n, layers, dirv;
    int* error;
};
__global__ void zshape_reset_kernel(ZshapeArgs a);
__global__ void zshape_solve_kernel(ZshapeArgs a);
__global__ void zshape_convert_kernel(ZshapeArgs a, const int* net_ids, RouteStore rs);

// FLT routing DAGs
constexpr int kFltStride = 10;  // layer index 1..L of the FLT tables (L <= 9), 0 unused
constexpr int kFltChangeWidth = 5;  // same-parity to-layers per (edge, from-layer)
constexpr int kFltWideBatch = 512;  // FLT batches of at least this many DAGs form the second tuning class
constexpr int kFltClasses = 2;  // FLT DAG classes by LDS share: one DP launch each, side by side
// One DAG's share of LDS in flt_route_lds_kernel, sized by the batch's largest DAG.
struct FltLdsShape {
    int slots, nodes, edges, dnodes, depths;  // child slots, nodes, edges, nodes of its widest depth, depths
};
struct FltArgs {
    DeviceGrid g;
    const double* presum;
    const float* vcost;
    // stage records (resident), net-local indices
    const int4* node_rec;   // {coordinate, first edge, end edge, first slot * 8 + slots}
    const int4* edge_rec;   // {parent, slot, parent coordinate, child coordinate}
    const int* order;       // per net: its nodes by depth, stable
    const int* depth_off;   // per net: depths + 1 offsets into its order
    // the batch
    const int4* desc;       // stage-global first node, edge, slot, depth offset
    const int4* sizes;      // nodes, edges, slots, depths
    const int* net_ids;
    int count, edges;
    int per_wave;           // DAGs per wavefront: 1, 2, 4 or 8
    int node_base, edge_base, slot_base;  // stage-global index of the batch's first node, edge, slot
    // batch tables
    double* costs;                   // [node * kFltStride + layer]
    unsigned long long* packed;      // [slot * kFltStride + layer]
    unsigned long long* best_path;   // [slot * kFltStride + layer]
    int* layer_output;
    double* cost_edges;              // [(edge * 9 + from) * 5 + to / 2]
    int* best_change;
    RouteStore routes;        // phase-two route words
    const int* pin_acc_num;   // pins per net: capacity = pins * routes.per_pin
    int* error;
    const int* eorder;      // per net: its edges by the depth of their child node (node order within a depth)
    const int* edepth_off;  // per net: depths + 1 offsets into its eorder (at the depth_off positions)
    // DP tables in LDS (flt_route_lds_kernel): bytes and shape of one DAG's
    // share, sized by the batch's largest DAG
    int lds_dag_bytes;
    FltLdsShape lds_shape;
    int* queue;  // persistent form: kFltQueues work-queue counters (nullptr: one workgroup per DAG group)
    // the batch's edges longer than 4 (stage-global indices): the only ones
    // whose layer-change tables the DP reads
    const int* far_edges;
    int far_count;
};
constexpr int kFltQueues = 32;       // persistent FLT DP: work-queue counters per batch
constexpr int kFltQueueStride = 32;  // ints between counters (one 128-byte line)
// The layer-change search of a batch's edges (flt_edge_kernel).
void launch_flt_edges(const FltArgs& a, hipStream_t stream);
// LDS bytes of one DAG's share of flt_route_lds_kernel (child slots, nodes,
// edges, nodes of its widest depth, depths; L layers)
size_t flt_lds_bytes(int L, int slots, int nodes, int edges, int dnodes, int depths);
// The FLT DP of a batch (L = 1..9): with lds, each DAG's tables in LDS
// (a.lds_dag_*, a.queue), else in the batch tables in global memory.
void launch_flt_route(const FltArgs& a, bool lds, int cus, hipStream_t stream);

// Congestion view (src/routing/congestion.cuh:60-112)
__global__ void congestion_view_kernel(DeviceGrid g, const float* capacity, const float* demand,
                                       unsigned char* congested, float* xsum, float* ysum);
__global__ void congestion_prefix_kernel(int tracks, int len, int track_stride, int pos_stride, float* sum);

// Soft DP
struct SoftCostArgs {
    DeviceGrid g;
    const float* capacity;
    const float* demand;
    double slope, zero_slope, via_weight;
    double* wcost;   // smooth wire cost per cell
    double* presum;  // exclusive prefix of wcost along each track
    double* vcost;   // smooth via cost per cell
    double *dw_ds, *dw_dz, *dv_ds, *dv_dz;  // partials in slope / zero_slope
};
struct SoftArgs {
    DeviceGrid g;
    const double* presum;
    const double* vcost;
    const int* node_off;  // stage topology (uploaded slots)
    const int* nodes;
    const int* parents;
    int begin, count;
    double tau;
    void* scratch;   // checkpointed tables of the nets [begin, begin + count)
    double* value;   // soft value per net
    double* gw;      // dF/dwcost per cell
    double* gv;      // dF/dvcost per cell
};
__global__ void soft_costs_kernel(SoftCostArgs a);
__global__ void soft_params_kernel(SoftCostArgs c, const double* gw, const double* gv, long long cells, double* out);
__global__ void soft_prefix_kernel(DeviceGrid g, const double* wcost, double* presum);
constexpr int kSoftLayers = 10;  // routing layers the soft DP supports
size_t soft_scratch_bytes(int L, long long nodes, int precision);
// kind 0: L-shape edges, 1: the Z-shape route family
void launch_soft_forward(const SoftArgs& a, int precision, int kind, int cus, hipStream_t stream);
void launch_soft_backward(const SoftArgs& a, int precision, int kind, int cus, hipStream_t stream);

// Evaluator-equivalent score (resources/evaluator.cpp:458-675)
struct ScoreArgs {
    int nets, X, Y, M;  // design nets, grid, metal layers (routing layers + 1)
    const int *pin_off, *pins;      // per net: pins as metal-layer cells; metal-0 pins print a via to metal 1
    const int *extra_off, *extra;   // per net: extra route endpoint pairs, metal-layer cells
    const int *sub_off, *subs;      // per net: its subnets, whose route words are in rs
    RouteStore rs;
    // phase-two words (after an FLT stage): subnet id with pin_off2[id] >= 0
    // has its words at routes2 + pin_off2[id] * per_pin2; nullptr: none
    const int* routes2;
    const int* pin_off2;
    int per_pin2;
    const int* dir;                 // [M] layer directions as read by the evaluator
    int* wire;                      // [M * X * Y] unit wire edges per cell
    int* stacked;                   // [M * X * Y] stacked-via demand per cell
    unsigned long long* vias;       // [kStatStripes * kScoreLayers] via layer-steps per stripe and layer
    unsigned long long* bad_nets;   // nets with a line the evaluator rejects
};
constexpr int kScoreCells = 8192;  // cells per score_cells_kernel block
constexpr int kScoreLines = 1024;  // lines of a net staged in LDS by score_nets_kernel
constexpr int kScoreLayers = 32;   // metal layers the score supports (via counters per stripe)
__global__ void score_nets_kernel(ScoreArgs a);

// The evaluator's completion test per net (resources/evaluator.cpp:527-538,
// check_connectivity :677-749) over the same lines as the score.
struct ConnArgs {
    const int* apin_off;       // [nets + 1] per net: its pins
    const int* acc_off;        // [pins + 1] per pin: its access points
    const int* acc;            // access points as metal-layer cells
    const long long* big_off;  // per net: offset of its staging in big_lines / big_par, -1: LDS
    int4* big_lines;
    int* big_par;
    unsigned char* status;        // per net: 0 complete, 1 open (a rejected line), 2 not connected
    unsigned long long* counts;   // [2] open, not connected
};
__global__ void connectivity_kernel(ScoreArgs a, ConnArgs c);
__global__ void score_cells_kernel(ScoreArgs a, const double* capacity, const int* x_len, const int* y_len,
                                   double* partial_loss, unsigned long long* partial_wl, unsigned short* demand_out);

__global__ void fill_int_kernel(int* p, int value, long long n);
__global__ void restride_routes_kernel(const int* src, int* dst, const int* ids, const int* pin_off2,
                                       const int* pin_acc_num, int old_per_pin, int per_pin);

}  // namespace grotor
