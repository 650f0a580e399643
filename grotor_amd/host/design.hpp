// Host-side design database and front end for the routing hot path.
//
// Mirrors the reference's host structures: db::net (src/input.hpp:15-24),
// cudb::net (include/router_state.hpp:156-178) and the preparation steps that
// produce the hot path's inputs: CAP/NET parsing (src/input.hpp:247-419),
// access-point selection (src/input.hpp:38-243), subnet splitting
// (src/preprocess.hpp:1448-1623), batching geometry
// (src/preprocess.hpp:72-150) and DFS-preorder trees
// (src/routing/lshape.cuh:86-102).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace grotor {

struct Layer {
    std::string name;
    int dir = 0;  // 0: horizontal (edges along x), 1: vertical
    double min_len = 0;
};

// Metal-layer grid as read from a CAP file.  capacity is indexed
// [metal layer][x][y] flattened as (l * X + x) * Y + y.
struct Grid {
    int L = 0, X = 0, Y = 0;  // metal layers and grid size
    double unit_length_wire_cost = 0, unit_via_cost = 0;
    std::vector<double> unit_length_short_costs;  // [L]
    std::vector<int> x_edge_len, y_edge_len;      // [X-1], [Y-1]
    std::vector<Layer> layers;                    // [L]
    std::vector<float> capacity;                  // [L * X * Y]

    int routing_layers() const { return L - 1; }
    int DIR() const { return layers.size() > 1 ? layers[1].dir : 0; }
};

struct DesignNet {
    std::string name;
    std::vector<std::vector<int>> access_points;  // per pin: 3D metal indices
    int minx = 0, maxx = 0, miny = 0, maxy = 0;
    std::vector<int> pins, extra_routes, subnets;
    int unfinished_subnet_count = 0;
};

// A routed (sub)net: the unit the batching and the DP operate on.
struct Subnet {
    int original = -1;
    std::vector<int> pins;  // 3D indices in routing-layer coordinates
    // rsmt[i]: adjacency of tree node i; rsmt.back(): node positions.
    std::vector<std::vector<int>> rsmt;
    std::vector<std::pair<int, int>> hsegs, vsegs;  // batching geometry, 2D
    std::vector<int> points;                        // batching points, 2D
    int minx = 0, maxx = 0, miny = 0, maxy = 0, hpwl = 0;
};

struct Design {
    Grid grid;
    std::vector<DesignNet> nets;
    std::vector<Subnet> subnets;
    std::vector<int> pin_acc_num;  // [subnets + 1]
};

// Parse reference-format CAP/NET files (src/input.hpp:247-419).  Unlike the
// reference, any metal-layer count >= 2 is accepted.
void read_design(const std::string& cap, const std::string& net, Design& design);
void write_cap(const std::string& path, const Grid& grid);
void write_net(const std::string& path, const Design& design);

// Access-point selection (src/input.hpp:38-243) for every design net.
void select_pins(Design& design, int threads);

// Build the routed subnets: tree construction, >12-pin splitting
// (src/preprocess.hpp:1448-1623) and batching geometry.
void build_subnets(Design& design, int threads);

// Geometry of src/preprocess.hpp:94-126 for a tree whose rsmt.back() holds
// 2D positions; afterwards node positions are re-encoded with their layer
// (pin layer, or L for Steiner nodes) as at src/preprocess.hpp:127-129.
void construct_geometry(Subnet& net, int X, int Y, int L);

// Batching geometry of one tree (src/preprocess.hpp:94-126): positions are
// 2D cells, adjacency lists as in rsmt; appends to the outputs.
void tree_geometry(const std::vector<std::vector<int>>& adjacency, const std::vector<int>& positions, int Y,
                   std::vector<int>& points, std::vector<std::pair<int, int>>& hsegs,
                   std::vector<std::pair<int, int>>& vsegs);

// DFS preorder of a subnet tree from node 0 (src/routing/lshape.cuh:86-102).
void preorder(const Subnet& net, std::vector<int>& nodes, std::vector<int>& parents);

// Stage order: subnet ids sorted by hpwl descending with std::sort
// (src/routing/lshape.cuh:53-55).
std::vector<int> hpwl_order(const Design& design);

// Synthetic instances for the five workload configurations; see
// generate.cpp for the distributions.
struct GenSpec {
    int config = 1;
    uint64_t seed = 0;
    double net_scale = 1.0;  // multiply the configuration's net count
    int grid_override = 0;   // square grid side, 0 = configuration default
};
void generate_design(const GenSpec& spec, Design& design);

// .out writer (src/output.hpp:79-112 with src/output_writer.hpp).
// routes_of(subnet) returns the subnet's route words ([count, a0, b0, ...]).
template <class RoutesOf>
std::string format_output(const Design& design, RoutesOf routes_of);

}  // namespace grotor

#include "output_impl.hpp"
