// C ABI of the host front end: design I/O, synthetic instances, preparation
// of the stage topology and .out output.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../csrc/gr_status.hpp"
#include "design.hpp"
#include "flt_host.hpp"
#include "zshape_host.hpp"

namespace grotor {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& message) { g_last_error = message; }
}  // namespace grotor

using namespace grotor;

struct gr_design {
    Design d;
    bool prepared = false;
    std::vector<int> order, node_off, nodes, parents, pt_off, pts, hs_off, hsegs, vs_off, vsegs;
    std::vector<float> capacity;  // routing layers, IDX order
    std::vector<double> short_costs;
    std::vector<int> net_pin_off, net_pins, net_extra_off, net_extra, net_sub_off, net_subs;
    std::vector<int> acc_pin_off, acc_off, acc;
    // Z-shape stage
    std::vector<int> zorder, zpt_off, zpts, zhs_off, zhsegs, zvs_off, zvsegs;
    ZshapePlan zplan;
    // FLT stage: DAGs of the rerouted nets (priority order), phase-two offsets, plan
    std::vector<int> forder, fpt_off, fpts, fhs_off, fhsegs, fvs_off, fvsegs, pin_off2;
    std::vector<FltDag> dags;
    std::vector<int> fp_batch_off, fp_net_ids, fp_nodes, fp_edges;
    std::vector<long long> fp_node_off, fp_edge_off;
};

extern "C" {

const char* gr_last_error(void) { return g_last_error.c_str(); }

int gr_design_read(const char* cap_path, const char* net_path, gr_design** out) {
    return guarded([&] {
        if (!cap_path || !net_path || !out) throw GrError(GR_E_ARG, "null argument");
        auto d = std::make_unique<gr_design>();
        try {
            read_design(cap_path, net_path, d->d);
        } catch (const std::invalid_argument& e) {
            throw GrError(GR_E_IO, e.what());
        }
        *out = d.release();
    });
}

int gr_design_generate(int config, unsigned long long seed, double net_scale, int grid_side, gr_design** out) {
    return guarded([&] {
        if (!out || net_scale <= 0) throw GrError(GR_E_ARG, "invalid arguments");
        auto d = std::make_unique<gr_design>();
        GenSpec spec;
        spec.config = config;
        spec.seed = seed;
        spec.net_scale = net_scale;
        spec.grid_override = grid_side;
        generate_design(spec, d->d);
        *out = d.release();
    });
}

int gr_design_write(const gr_design* design, const char* cap_path, const char* net_path) {
    return guarded([&] {
        if (!design) throw GrError(GR_E_ARG, "null design");
        if (cap_path) write_cap(cap_path, design->d.grid);
        if (net_path) write_net(net_path, design->d);
    });
}

int gr_design_destroy(gr_design* design) {
    delete design;
    return GR_OK;
}

int gr_design_prepare(gr_design* g, int threads) {
    return guarded([&] {
        if (!g) throw GrError(GR_E_ARG, "null design");
        Design& d = g->d;
        select_pins(d, threads);
        build_subnets(d, threads);
        g->order = hpwl_order(d);
        const int n = int(g->order.size());
        g->node_off.assign(1, 0);
        g->pt_off.assign(1, 0);
        g->hs_off.assign(1, 0);
        g->vs_off.assign(1, 0);
        g->nodes.clear(), g->parents.clear(), g->pts.clear(), g->hsegs.clear(), g->vsegs.clear();
        std::vector<int> nodes, parents;
        for (int k = 0; k < n; ++k) {
            const Subnet& s = d.subnets[g->order[k]];
            preorder(s, nodes, parents);
            g->nodes.insert(g->nodes.end(), nodes.begin(), nodes.end());
            g->parents.insert(g->parents.end(), parents.begin(), parents.end());
            g->node_off.push_back(int(g->nodes.size()));
            g->pts.insert(g->pts.end(), s.points.begin(), s.points.end());
            g->pt_off.push_back(int(g->pts.size()));
            for (auto& e : s.hsegs) g->hsegs.push_back(e.first), g->hsegs.push_back(e.second);
            g->hs_off.push_back(int(g->hsegs.size() / 2));
            for (auto& e : s.vsegs) g->vsegs.push_back(e.first), g->vsegs.push_back(e.second);
            g->vs_off.push_back(int(g->vsegs.size() / 2));
        }
        const Grid& gr = d.grid;
        const int L = gr.routing_layers();
        g->capacity.assign(gr.capacity.begin() + size_t(gr.X) * gr.Y, gr.capacity.end());
        g->short_costs.assign(gr.unit_length_short_costs.begin() + 1, gr.unit_length_short_costs.end());
        (void)L;
        g->prepared = true;
    });
}

int gr_design_topology(gr_design* g, gr_topology* t) {
    return guarded([&] {
        if (!g || !t) throw GrError(GR_E_ARG, "null argument");
        if (!g->prepared) throw GrError(GR_E_STATE, "design not prepared");
        const Grid& gr = g->d.grid;
        t->L = gr.routing_layers();
        t->X = gr.X;
        t->Y = gr.Y;
        t->DIR = gr.DIR();
        t->nets = int(g->order.size());
        t->subnets = int(g->d.subnets.size());
        t->order = g->order.data();
        t->node_off = g->node_off.data();
        t->nodes = g->nodes.data();
        t->parents = g->parents.data();
        t->pt_off = g->pt_off.data();
        t->pts = g->pts.data();
        t->hs_off = g->hs_off.data();
        t->hsegs = g->hsegs.data();
        t->vs_off = g->vs_off.data();
        t->vsegs = g->vsegs.data();
        t->pin_acc_num = g->d.pin_acc_num.data();
        t->capacity = g->capacity.data();
        t->unit_length_short_costs = g->short_costs.data();
        t->x_edge_len = gr.x_edge_len.data();
        t->y_edge_len = gr.y_edge_len.data();
        t->unit_length_wire_cost = gr.unit_length_wire_cost;
        t->unit_via_cost = gr.unit_via_cost;
    });
}

int gr_design_zshape_stage(gr_design* g, gr_zshape_stage* t) {
    return guarded([&] {
        if (!g || !t) throw GrError(GR_E_ARG, "null argument");
        if (!g->prepared) throw GrError(GR_E_STATE, "design not prepared");
        g->zorder = zshape_order(g->d);
        g->zpt_off.assign(1, 0), g->zhs_off.assign(1, 0), g->zvs_off.assign(1, 0);
        g->zpts.clear(), g->zhsegs.clear(), g->zvsegs.clear();
        for (int id : g->zorder) {
            const Subnet& s = g->d.subnets[id];
            g->zpts.insert(g->zpts.end(), s.points.begin(), s.points.end());
            g->zpt_off.push_back(int(g->zpts.size()));
            for (auto& e : s.hsegs) g->zhsegs.push_back(e.first), g->zhsegs.push_back(e.second);
            g->zhs_off.push_back(int(g->zhsegs.size() / 2));
            for (auto& e : s.vsegs) g->zvsegs.push_back(e.first), g->zvsegs.push_back(e.second);
            g->zvs_off.push_back(int(g->zvsegs.size() / 2));
        }
        t->nets = int(g->zorder.size());
        t->order = g->zorder.data();
        t->pt_off = g->zpt_off.data();
        t->pts = g->zpts.data();
        t->hs_off = g->zhs_off.data();
        t->hsegs = g->zhsegs.data();
        t->vs_off = g->zvs_off.data();
        t->vsegs = g->zvsegs.data();
    });
}

int gr_design_zshape_plan(gr_design* g, const int* color, int num_batches, int max_batch, gr_zshape_plan* p) {
    return guarded([&] {
        if (!g || !color || !p || num_batches < 0 || max_batch < 1) throw GrError(GR_E_ARG, "invalid arguments");
        if (g->zorder.empty() && !g->d.subnets.empty())
            throw GrError(GR_E_STATE, "no Z-shape stage order (gr_design_zshape_stage)");
        try {
            g->zplan = zshape_plan(g->d, g->zorder, color, num_batches, max_batch);
        } catch (const std::runtime_error& e) {
            throw GrError(GR_E_ARG, e.what());
        }
        const ZshapePlan& z = g->zplan;
        p->batches = int(z.batch_off.size()) - 1;
        p->batch_off = z.batch_off.data();
        p->net_ids = z.net_ids.data();
        p->rec_off = z.rec_off.data();
        p->records = z.records.data();
        p->raw_cap = z.raw_cap.data();
    });
}

int gr_design_zshape_finish(gr_design* g, int* routes, int words_per_pin, int* promoted) {
    return guarded([&] {
        if (!g || !routes || words_per_pin < 4) throw GrError(GR_E_ARG, "invalid arguments");
        const int p = zshape_promote(g->d, words_per_pin, routes);
        try {
            zshape_canonicalize(g->d, words_per_pin, routes);
        } catch (const std::runtime_error& e) {
            throw GrError(GR_E_CAPACITY, e.what());
        }
        if (promoted) *promoted = p;
    });
}

int gr_design_flt_stage(gr_design* g, const unsigned char* reroute, const float* xsum, const float* ysum, int threads,
                        gr_flt_stage* t) {
    return guarded([&] {
        if (!g || !reroute || !t) throw GrError(GR_E_ARG, "null argument");
        if (!g->prepared) throw GrError(GR_E_STATE, "design not prepared");
        const Design& d = g->d;
        const Grid& gr = d.grid;
        const int n = int(d.subnets.size());
        // prepare_detour_topology (src/routing/flt.cuh:48-73): hpwl descending
        // with std::sort over the rerouted ids in increasing order
        g->forder.clear();
        g->pin_off2.assign(n, -1);
        long long pins = 0;
        for (int i = 0; i < n; ++i)
            if (reroute[i]) {
                g->forder.push_back(i);
                g->pin_off2[i] = int(pins);  // src/pipeline.hpp:96-101
                pins += d.pin_acc_num[i + 1] - d.pin_acc_num[i];
            }
        if (pins > INT32_MAX) throw GrError(GR_E_ARG, "too many phase-two pins");
        const auto& s = d.subnets;
        std::sort(g->forder.begin(), g->forder.end(), [&](int l, int r) { return s[l].hpwl > s[r].hpwl; });
        FltParams p;
        p.mode = n > 20000000 ? 2 : n > 5000000 ? 1 : 0;  // src/pipeline.hpp:88-89
        p.X = gr.X, p.Y = gr.Y, p.L = gr.routing_layers();
        p.xsum = xsum, p.ysum = ysum;
        const int m = int(g->forder.size());
        g->dags.assign(m, FltDag{});
        std::vector<std::string> errors(std::max(threads, 1));
        std::atomic<int> next{0};
        std::vector<std::thread> pool;
        for (int w = 0; w < std::max(threads, 1); ++w)
            pool.emplace_back([&, w] {
                for (int k; (k = next++) < m;) {
                    try {
                        flt_build_dag(s[g->forder[k]], p, g->dags[k]);
                    } catch (const std::exception& e) {
                        errors[w] = e.what();
                        next = m;
                    }
                }
            });
        for (auto& th : pool) th.join();
        for (const auto& e : errors)
            if (!e.empty()) throw GrError(GR_E_ARG, "FLT DAG: " + e);
        g->fpt_off.assign(1, 0), g->fhs_off.assign(1, 0), g->fvs_off.assign(1, 0);
        g->fpts.clear(), g->fhsegs.clear(), g->fvsegs.clear();
        long long nodes = 0, edges = 0;
        for (int k = 0; k < m; ++k) {
            const Subnet& sn = s[g->forder[k]];
            const FltDag& dag = g->dags[k];
            g->fpts.insert(g->fpts.end(), dag.points.begin(), dag.points.end());
            g->fpt_off.push_back(int(g->fpts.size()));
            for (auto& e : sn.hsegs) g->fhsegs.push_back(e.first), g->fhsegs.push_back(e.second);
            g->fhs_off.push_back(int(g->fhsegs.size() / 2));
            for (auto& e : sn.vsegs) g->fvsegs.push_back(e.first), g->fvsegs.push_back(e.second);
            g->fvs_off.push_back(int(g->fvsegs.size() / 2));
            nodes += dag.nodes.size() / 4;
            edges += dag.edges.size() / 2;
        }
        t->nets = m;
        t->order = g->forder.data();
        t->pt_off = g->fpt_off.data();
        t->pts = g->fpts.data();
        t->hs_off = g->fhs_off.data();
        t->hsegs = g->fhsegs.data();
        t->vs_off = g->fvs_off.data();
        t->vsegs = g->fvsegs.data();
        t->pin_off2 = g->pin_off2.data();
        t->pins2 = pins;
        t->nodes = nodes;
        t->edges = edges;
    });
}

int gr_design_flt_plan(gr_design* g, const int* color, int num_batches, gr_flt_plan* out) {
    return guarded([&] {
        if (!g || !out || num_batches < 0 || (!color && !g->forder.empty())) throw GrError(GR_E_ARG, "invalid arguments");
        const int m = int(g->forder.size());
        std::vector<std::vector<int>> batches(num_batches);
        for (int k = 0; k < m; ++k) {
            if (color[k] < 0 || color[k] >= num_batches) throw GrError(GR_E_ARG, "batch index out of range");
            batches[color[k]].push_back(k);
        }
        // flt::Plans (include/flt_topology.hpp:92-127): batches in reverse
        // order, empty ones skipped, nets in priority order inside a batch
        g->fp_batch_off.assign(1, 0);
        g->fp_net_ids.clear(), g->fp_nodes.clear(), g->fp_edges.clear();
        g->fp_node_off.assign(1, 0), g->fp_edge_off.assign(1, 0);
        for (int b = num_batches - 1; b >= 0; --b) {
            if (batches[b].empty()) continue;
            for (int k : batches[b]) {
                const FltDag& dag = g->dags[k];
                g->fp_net_ids.push_back(g->forder[k]);
                g->fp_nodes.insert(g->fp_nodes.end(), dag.nodes.begin(), dag.nodes.end());
                g->fp_edges.insert(g->fp_edges.end(), dag.edges.begin(), dag.edges.end());
                g->fp_node_off.push_back((long long)g->fp_nodes.size() / 4);
                g->fp_edge_off.push_back((long long)g->fp_edges.size() / 2);
            }
            g->fp_batch_off.push_back(int(g->fp_net_ids.size()));
        }
        out->batches = int(g->fp_batch_off.size()) - 1;
        out->batch_off = g->fp_batch_off.data();
        out->net_ids = g->fp_net_ids.data();
        out->node_off = g->fp_node_off.data();
        out->nodes = g->fp_nodes.data();
        out->edge_off = g->fp_edge_off.data();
        out->edges = g->fp_edges.data();
    });
}

int gr_design_nets(gr_design* g, int* count, const int** pin_off, const int** pins, const int** extra_off,
                   const int** extra, const int** subnet_off, const int** subnets) {
    return guarded([&] {
        if (!g || !count) throw GrError(GR_E_ARG, "null argument");
        if (!g->prepared) throw GrError(GR_E_STATE, "design not prepared");
        g->net_pin_off.assign(1, 0), g->net_extra_off.assign(1, 0), g->net_sub_off.assign(1, 0);
        g->net_pins.clear(), g->net_extra.clear(), g->net_subs.clear();
        for (const auto& n : g->d.nets) {
            g->net_pins.insert(g->net_pins.end(), n.pins.begin(), n.pins.end());
            g->net_extra.insert(g->net_extra.end(), n.extra_routes.begin(), n.extra_routes.end());
            g->net_subs.insert(g->net_subs.end(), n.subnets.begin(), n.subnets.end());
            g->net_pin_off.push_back(int(g->net_pins.size()));
            g->net_extra_off.push_back(int(g->net_extra.size()));
            g->net_sub_off.push_back(int(g->net_subs.size()));
        }
        *count = int(g->d.nets.size());
        if (pin_off) *pin_off = g->net_pin_off.data();
        if (pins) *pins = g->net_pins.data();
        if (extra_off) *extra_off = g->net_extra_off.data();
        if (extra) *extra = g->net_extra.data();
        if (subnet_off) *subnet_off = g->net_sub_off.data();
        if (subnets) *subnets = g->net_subs.data();
    });
}

int gr_design_access(gr_design* g, int* count, const int** pin_off, const int** access_off, const int** access) {
    return guarded([&] {
        if (!g || !count) throw GrError(GR_E_ARG, "null argument");
        g->acc_pin_off.assign(1, 0), g->acc_off.assign(1, 0);
        g->acc.clear();
        for (const auto& n : g->d.nets) {
            for (const auto& pin : n.access_points) {
                g->acc.insert(g->acc.end(), pin.begin(), pin.end());
                g->acc_off.push_back(int(g->acc.size()));
            }
            g->acc_pin_off.push_back(int(g->acc_off.size()) - 1);
        }
        *count = int(g->d.nets.size());
        if (pin_off) *pin_off = g->acc_pin_off.data();
        if (access_off) *access_off = g->acc_off.data();
        if (access) *access = g->acc.data();
    });
}

int gr_tree_geometry(int Y, int n, const int* positions, const int* adj_off, const int* adj, int* points,
                     int* npoints, int* hsegs, int* nhsegs, int* vsegs, int* nvsegs) {
    return guarded([&] {
        if (n < 0 || !positions || !adj_off || !adj || !npoints || !nhsegs || !nvsegs)
            throw GrError(GR_E_ARG, "null argument");
        std::vector<std::vector<int>> adjacency(n);
        for (int i = 0; i < n; ++i) adjacency[i].assign(adj + adj_off[i], adj + adj_off[i + 1]);
        std::vector<int> pos(positions, positions + n), pts;
        std::vector<std::pair<int, int>> hs, vs;
        tree_geometry(adjacency, pos, Y, pts, hs, vs);
        if (int(pts.size()) > *npoints || int(hs.size()) > *nhsegs || int(vs.size()) > *nvsegs)
            throw GrError(GR_E_CAPACITY, "geometry output buffers too small");
        std::copy(pts.begin(), pts.end(), points);
        for (size_t i = 0; i < hs.size(); ++i) hsegs[2 * i] = hs[i].first, hsegs[2 * i + 1] = hs[i].second;
        for (size_t i = 0; i < vs.size(); ++i) vsegs[2 * i] = vs[i].first, vsegs[2 * i + 1] = vs[i].second;
        *npoints = int(pts.size()), *nhsegs = int(hs.size()), *nvsegs = int(vs.size());
    });
}

int gr_flt_dag(int X, int Y, int L, int mode, int n, const int* positions, const int* adj_off, const int* adj,
               const float* xsum, const float* ysum, int* nodes, int* nnodes, int* edges, int* nedges, int* points,
               int* npoints) {
    return guarded([&] {
        if (n < 1 || X < 1 || Y < 1 || L < 1 || !positions || !adj_off || !adj || !nodes || !nnodes || !edges ||
            !nedges || !points || !npoints)
            throw GrError(GR_E_ARG, "invalid arguments");
        Subnet net;
        net.rsmt.resize(n + 1);
        for (int i = 0; i < n; ++i) net.rsmt[i].assign(adj + adj_off[i], adj + adj_off[i + 1]);
        net.rsmt[n].assign(positions, positions + n);
        FltParams p;
        p.mode = mode, p.X = X, p.Y = Y, p.L = L, p.xsum = xsum, p.ysum = ysum;
        FltDag dag;
        try {
            flt_build_dag(net, p, dag);
        } catch (const std::runtime_error& e) {
            throw GrError(GR_E_ARG, e.what());
        }
        const int nn = int(dag.nodes.size() / 4), ne = int(dag.edges.size() / 2), np = int(dag.points.size());
        if (nn > *nnodes || ne > *nedges || np > *npoints) throw GrError(GR_E_CAPACITY, "DAG output buffers too small");
        std::copy(dag.nodes.begin(), dag.nodes.end(), nodes);
        std::copy(dag.edges.begin(), dag.edges.end(), edges);
        std::copy(dag.points.begin(), dag.points.end(), points);
        *nnodes = nn, *nedges = ne, *npoints = np;
    });
}

int gr_design_write_output2(const gr_design* g, const int* routes, const int* routes2, const int* pin_off2,
                            int words_per_pin2, const char* out_path) {
    return guarded([&] {
        if (!g || !routes || !out_path || (pin_off2 && (!routes2 || words_per_pin2 < 1)))
            throw GrError(GR_E_ARG, "invalid arguments");
        if (!g->prepared) throw GrError(GR_E_STATE, "design not prepared");
        const auto& acc = g->d.pin_acc_num;
        // src/output.hpp:98-106: rerouted subnets print their phase-two words
        const std::string text = format_output(g->d, [&](int id) {
            if (pin_off2 && pin_off2[id] >= 0) return routes2 + (size_t)pin_off2[id] * words_per_pin2;
            return routes + (size_t)acc[id] * 12;
        });
        FILE* f = fopen(out_path, "w");
        if (!f) throw GrError(GR_E_IO, std::string("cannot open output: ") + out_path);
        const size_t written = fwrite(text.data(), 1, text.size(), f);
        if (fclose(f) || written != text.size()) throw GrError(GR_E_IO, "failed to write routing output");
    });
}

int gr_design_write_output(const gr_design* g, const int* routes, const char* out_path) {
    return guarded([&] {
        if (!g || !routes || !out_path) throw GrError(GR_E_ARG, "null argument");
        if (!g->prepared) throw GrError(GR_E_STATE, "design not prepared");
        const auto& acc = g->d.pin_acc_num;
        const std::string text =
            format_output(g->d, [&](int id) { return routes + (size_t)acc[id] * 12; });
        FILE* f = fopen(out_path, "w");
        if (!f) throw GrError(GR_E_IO, std::string("cannot open output: ") + out_path);
        const size_t written = fwrite(text.data(), 1, text.size(), f);
        if (fclose(f) || written != text.size()) throw GrError(GR_E_IO, "failed to write routing output");
    });
}

int gr_design_metal_grid(const gr_design* g, int* metal_layers, int* layer_dir, float* capacity,
                         double* overflow_cost) {
    return guarded([&] {
        if (!g || !metal_layers) throw GrError(GR_E_ARG, "null argument");
        const Grid& grid = g->d.grid;
        *metal_layers = grid.L;
        const size_t cells = size_t(grid.L) * grid.X * grid.Y;
        if (layer_dir)
            for (int z = 0; z < grid.L; ++z) layer_dir[z] = grid.layers[z].dir;
        if (capacity) std::copy(grid.capacity.begin(), grid.capacity.begin() + cells, capacity);
        if (overflow_cost)
            std::copy(grid.unit_length_short_costs.begin(), grid.unit_length_short_costs.end(), overflow_cost);
    });
}

}  // extern "C"
