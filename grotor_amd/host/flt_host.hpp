// Host side of the FLT (routing-DAG) stage: the DAG builder and the compact
// topology records the GPU stage consumes.
//
// Reference: net::generate_detours_reconstruct (src/preprocess.hpp:769-1436),
// select_root_net (:156-186), flt::Topology::emit (include/flt_topology.hpp:52-70).
#pragma once

#include <vector>

#include "design.hpp"

namespace grotor {

// One net's DAG in the reference's compact form (include/flt_topology.hpp:15-16):
// node records {coordinate, depth, child slots, parent_end} and edge records
// {parent, node * 10 + slot}, net-local indices, nodes in topological order
// (root first); points are the batching cells, sorted and unique.
struct FltDag {
    std::vector<int> nodes;  // 4 ints per node
    std::vector<int> edges;  // 2 ints per edge
    std::vector<int> points;
    int depths = 0;
};

// Builder parameters that the reference derives from the net count
// (`mode`, src/pipeline.hpp:88-89): mode 0/1 -> 6 mirror tracks, ratio 0.1,
// 7 Z-shapes with 2 selected; mode 2 -> 9 tracks, ratio 0.15, 10 Z-shapes
// with 10 selected.
struct FltParams {
    int mode = 0;
    int X = 0, Y = 0, L = 0;
    const float* xsum = nullptr;  // congestion prefix along x, [X * Y] (x * Y + y)
    const float* ysum = nullptr;  // congestion prefix along y
};

// Root of a tree: the last node popped by a breadth-first sweep that starts
// from every leaf (select_root_net, src/preprocess.hpp:156-186).
int flt_select_root(const std::vector<std::vector<int>>& rsmt);

// generate_detours_reconstruct for one subnet tree (rsmt adjacency with
// rsmt.back() holding 3D positions, Steiner nodes at layer L).
void flt_build_dag(const Subnet& net, const FltParams& p, FltDag& out);

}  // namespace grotor
