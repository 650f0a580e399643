// Host side of the Z-shape (GGR) stage (src/pipeline_ggr.hpp,
// src/routing/zshape_topology.hpp).
#pragma once

#include <vector>

#include "design.hpp"

namespace grotor {

constexpr int kZshapeRoutePerPin = 256;  // ROUTE_PER_PIN of the Z-shape pipeline (src/pipeline_ggr.hpp:10)

// Stage priority order: hpwl descending, larger subnet id first on ties.
std::vector<int> zshape_order(const Design& design);

// Bounding-box refinement of one conflict-free batch (at most max_batch
// nets per refined batch, pairwise disjoint boxes).
std::vector<std::vector<int>> bbox_batches(const std::vector<int>& group, const Design& design, int max_batch);

// Encoded records of one net, written to out (2 + 6 * nodes words): raw
// offset, node count, six words per node in preorder from the first leaf.
void encode_zshape_net(const Subnet& net, int raw_offset, int X, int Y, int L, int* out);

struct ZshapePlan {
    std::vector<int> batch_off{};  // [batches + 1] over net_ids, execution order
    std::vector<int> net_ids;      // subnet id per execution position
    std::vector<long long> rec_off{};  // [batches + 1] into records
    std::vector<int> records;      // per batch: record offset of each net (relative), then records
    std::vector<int> raw_cap;      // raw words per net (pins * 256), execution order
};

// Batches in execution order from the RSMT batch colors of `order`.
ZshapePlan zshape_plan(const Design& design, const std::vector<int>& order, const int* color, int num_batches,
                       int max_batch);

// Layer promotion of congested segments (returns the number promoted) and
// canonical merged wires / via stacks, on route words of per_pin words per
// pin (IDX order).
int zshape_promote(const Design& design, int per_pin, int* routes);
void zshape_canonicalize(const Design& design, int per_pin, int* routes);

}  // namespace grotor
