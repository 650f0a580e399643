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
};

struct Stats {
    double total_wirelength;
    unsigned long long total_vias;
    unsigned long long wire_edges;  // unit wire edges committed (for byte accounting)
};

// Error codes written by kernels into the context's device error word.
enum DeviceError : int {
    kDevOk = 0,
    kDevRouteOverflow = 1,
    kDevNoSolution = 2,
    kDevSpinLimit = 3,
    kDevColorOverflow = 4,
};

__global__ void refresh_tracks_kernel(DeviceGrid g, const float* capacity, const float* demand, float* vcost,
                                      double* presum, float* wcost_out, unsigned* track_dirty, int force,
                                      CostParams cp, unsigned long long* refreshed);

struct RouteStore {
    int* routes;
    const int* pin_acc_num;
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
};
__global__ void commit_wires_kernel(CommitArgs a);
__global__ void commit_vias_kernel(CommitArgs a);

struct LshapeArgs {
    DeviceGrid g;
    const float* vcost;
    const double* presum;
    RouteStore rs;
    const int* net_ids;
    const int* node_off;
    const int* nodes;
    const int* parents;
    const int* slots;  // permuted stage slots
    int begin, count;
    int nmax;          // node capacity of one workgroup's scratch
    double* gscratch;  // global scratch when the tree does not fit in LDS
    int* error;
};
size_t lshape_scratch_bytes(int L, int nmax);
__global__ void lshape_dp_lds_kernel(LshapeArgs a);
__global__ void lshape_dp_global_kernel(LshapeArgs a);

struct RipupArgs {
    DeviceGrid g;
    const float* capacity;
    const float* demand;
    int* of_sum;
    int threshold;
};
__global__ void overflow_prefix_kernel(RipupArgs a);
__global__ void overflow_nets_kernel(DeviceGrid g, const int* of_sum, RouteStore rs, int net_count,
                                     unsigned char* is_of_net);

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
__global__ void count_marks_kernel(BatchGeom b, int* mark_count);
__global__ void emit_keys_kernel(BatchGeom b, const int* mark_off, unsigned long long* keys);
__global__ void build_lists_kernel(const unsigned long long* ukeys, const int* unique_count, int cells, int* cell_off,
                                   int* list_net);
__global__ void mark_readers_kernel(BatchGeom b, const int* cell_off, const int* list_net, int* is_reader);
__global__ void entry_info_kernel(const unsigned long long* ukeys, const int* unique_count, const int* cell_off,
                                  const int* is_reader, const int* reader_scan, int* entry_info);
constexpr int kMaxMaskWords = 64;  // colors 0..4095
struct ColorArgs {
    BatchGeom b;
    const int* mark_off;
    const unsigned long long* keys;  // marks in emission order (cell << 32 | net)
    const int* cell_off;
    const int* list_net;
    const int* entry_info;     // 2 * reader rank + is_reader, per list entry
    int* claimed;              // per list entry: write done
    unsigned long long* mask;  // [cells * mask_words]
    int* writes;               // [cells] completed writes
    int* reads;                // [cells] completed readers
    int mask_words;
    int* color;
    int* color_count;
    int* next;
    int* max_color;
    int* error;
    long long spin_limit;
};
__global__ void color_kernel(ColorArgs a);

__global__ void fill_int_kernel(int* p, int value, long long n);

}  // namespace grotor
