// grotor: the rip-up stage pipeline on one MI355X, over the C ABI only.
//
// Same command line as the reference router (src/main.cu:36-84):
//   grotor -cap FILE -net FILE -out FILE [--batching cpu|gpu] [--sparsity N]
//          [--profile-stages 0|1] [--report-score 0|1] [--stop-after lshape|routing]
//          [--lut-dir DIR] [--validate-batches 0]
// and the same stages (grotor::route, src/pipeline.hpp:14-157): L-shape
// stage over every subnet, rip-up of the nets crossing more than 10
// overflowed edges, congestion view, FLT DAGs, the FLT stage, .out output.
// Batching is the exact first-fit colouring on the GPU; `cpu` (the
// reference's default) names the same batches, which the GPU builds.
// The reference's ray-tracing backends have no counterpart on CDNA4.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grotor_amd.h"

namespace {

void ok(int rc, const char* op) {
    if (rc) throw std::runtime_error(std::string(op) + ": " + gr_last_error());
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

// Stage arrays in execution order: batch lists in priority order, the last
// batch first (src/routing/lshape.cuh:70-77).
struct Schedule {
    std::vector<int> batch_off, net_ids, node_off, nodes, parents;
};

Schedule lshape_schedule(const gr_topology& t, const std::vector<int>& color, int nb) {
    std::vector<std::vector<int>> lists(nb);
    for (int k = 0; k < t.nets; ++k) lists[color[k]].push_back(k);
    Schedule s;
    s.batch_off.push_back(0);
    s.node_off.push_back(0);
    for (int b = nb - 1; b >= 0; --b) {
        for (int k : lists[b]) {
            s.net_ids.push_back(t.order[k]);
            for (int i = t.node_off[k]; i < t.node_off[k + 1]; ++i) {
                s.nodes.push_back(t.nodes[i]);
                s.parents.push_back(t.parents[i]);
            }
            s.node_off.push_back(int(s.nodes.size()));
        }
        s.batch_off.push_back(int(s.net_ids.size()));
    }
    return s;
}

struct Options {
    std::string cap, net, out;
    int sparsity = 1;
    bool profile_stages = false, report_score = true, stop_after_lshape = false;
};

Options parse(int argc, char** argv) {
    const char* usage =
        "Usage: grotor -cap FILE -net FILE -out FILE [--batching cpu|gpu] [--validate-batches 0] [--sparsity N] "
        "[--lut-dir DIR] [--profile-stages 0|1] [--report-score 0|1] [--stop-after lshape|routing]";
    if (argc < 7 || argc % 2 != 1) throw std::invalid_argument(usage);
    Options o;
    for (int i = 1; i < argc; i += 2) {
        const std::string opt = argv[i], val = argv[i + 1];
        auto flag = [&](bool& target) {
            if (val != "0" && val != "1") throw std::invalid_argument(opt + " expects 0 or 1");
            target = val == "1";
        };
        if (opt == "-cap") o.cap = val;
        else if (opt == "-net") o.net = val;
        else if (opt == "-out") o.out = val;
        else if (opt == "--lut-dir") {
            // the tree construction needs no lookup tables
        } else if (opt == "--batching") {
            if (val == "rtcore" || val == "hiprt")
                throw std::invalid_argument("--batching " + val + " needs ray-tracing hardware; gpu builds the same batches");
            if (val != "cpu" && val != "gpu") throw std::invalid_argument("--batching expects cpu or gpu");
        } else if (opt == "--validate-batches") {
            if (val != "0" && val != "1") throw std::invalid_argument("--validate-batches expects 0 or 1");
            if (val == "1") throw std::invalid_argument("--validate-batches requires ray-tracing batching");
        } else if (opt == "--report-score") flag(o.report_score);
        else if (opt == "--profile-stages") flag(o.profile_stages);
        else if (opt == "--stop-after") {
            if (val != "lshape" && val != "routing") throw std::invalid_argument("--stop-after expects lshape or routing");
            o.stop_after_lshape = val == "lshape";
        } else if (opt == "--sparsity") {
            size_t used = 0;
            o.sparsity = std::stoi(val, &used);
            if (used != val.size() || o.sparsity < 1) throw std::invalid_argument("Sparsity must be a positive integer");
        } else {
            throw std::invalid_argument("Unknown option: " + opt);
        }
    }
    if (o.cap.empty() || o.net.empty() || o.out.empty()) throw std::invalid_argument("-cap, -net and -out are required");
    return o;
}

int run(const Options& o) {
    const auto t_start = Clock::now();
    gr_design* d = nullptr;
    ok(gr_design_read(o.cap.c_str(), o.net.c_str(), &d), "read input");
    ok(gr_design_prepare(d, 8), "prepare nets");
    gr_topology t{};
    ok(gr_design_topology(d, &t), "topology");
    char name[256];
    if (!gr_device_name(name, sizeof name)) printf("GPU_PLATFORM backend=HIP device=%s wave=64\n", name);
    gr_grid_desc g{t.L, t.X, t.Y, t.DIR, t.unit_length_wire_cost, t.unit_via_cost, t.unit_length_short_costs,
                   t.x_edge_len, t.y_edge_len, t.capacity};
    gr_ctx* ctx = nullptr;
    ok(gr_ctx_create(&g, &ctx), "create context");
    ok(gr_ctx_set_routes(ctx, t.subnets, t.pin_acc_num), "route storage");
    const double setup_ms = ms_since(t_start);

    // L-shape stage (src/routing/lshape.cuh:40-196)
    auto t0 = Clock::now();
    std::vector<int> color(std::max(t.nets, 1));
    int nb = 0;
    ok(gr_batch_plan(ctx, t.nets, t.pt_off, t.pts, t.hs_off, t.hsegs, t.vs_off, t.vsegs, 1000000, o.sparsity,
                     color.data(), &nb),
       "L-shape batching");
    const double batch_ms = ms_since(t0);
    printf("BATCH_TIMING stage=lshape backend=gpu nets=%d batches=%d elapsed_ms=%.6f\n", t.nets, nb, batch_ms);
    Schedule s = lshape_schedule(t, color, nb);
    ok(gr_lshape_upload(ctx, t.nets, s.net_ids.data(), s.node_off.data(), s.nodes.data(), s.parents.data()),
       "L-shape topology upload");
    double dag_ms = 0;
    ok(gr_lshape_run(ctx, nb, s.batch_off.data(), 1, 0, &dag_ms), "L-shape stage");
    printf("DAG_TIMING stage=lshape batch_index=-1 elapsed_ms=%.6f nets=%d batches=%d\n", dag_ms, t.nets, nb);
    const long long words = 12LL * t.pin_acc_num[t.subnets] + 1;
    std::vector<int> routes(words);
    ok(gr_routes_download(ctx, routes.data(), words), "download routes");
    double flt_ms = 0, class_ms = 0;
    int rerouted = 0;
    std::vector<int> routes2;
    const int* pin_off2 = nullptr;
    if (!o.stop_after_lshape) {
        // rip-up classification and the congestion view (src/pipeline.hpp:86-108)
        t0 = Clock::now();
        std::vector<unsigned char> flags(std::max(t.subnets, 1));
        std::vector<float> xsum(size_t(t.X) * t.Y), ysum(size_t(t.X) * t.Y);
        ok(gr_congestion_view(ctx, nullptr, xsum.data(), ysum.data()), "congestion view");
        ok(gr_ripup(ctx, 0, flags.data()), "rip-up");
        gr_flt_stage fs{};
        ok(gr_design_flt_stage(d, flags.data(), xsum.data(), ysum.data(), 8, &fs), "FLT topology");
        class_ms = ms_since(t0);
        rerouted = fs.nets;
        if (fs.nets) {
            ok(gr_ctx_set_routes_phase2(ctx, fs.pin_off2, fs.pins2, 100), "phase-two route storage");
            t0 = Clock::now();
            std::vector<int> fcolor(fs.nets);
            int fnb = 0;
            ok(gr_batch_plan(ctx, fs.nets, fs.pt_off, fs.pts, fs.hs_off, fs.hsegs, fs.vs_off, fs.vsegs, 300000,
                             o.sparsity, fcolor.data(), &fnb),
               "FLT batching");
            const double fbatch_ms = ms_since(t0);
            printf("BATCH_TIMING stage=flt backend=gpu nets=%d batches=%d elapsed_ms=%.6f\n", fs.nets, fnb, fbatch_ms);
            gr_flt_plan plan{};
            ok(gr_design_flt_plan(d, fcolor.data(), fnb, &plan), "FLT plan");
            ok(gr_flt_upload(ctx, plan.batches, plan.batch_off, plan.net_ids, plan.node_off, plan.nodes, plan.edge_off,
                             plan.edges),
               "FLT upload");
            ok(gr_flt_run(ctx, nb + 1, 1, &flt_ms), "FLT stage");
            printf("DAG_TIMING stage=detour batch_index=-1 elapsed_ms=%.6f nets=%d batches=%d\n", flt_ms, fs.nets,
                   plan.batches);
            routes2.resize(100LL * fs.pins2 + 1);
            ok(gr_routes2_download(ctx, routes2.data(), (long long)routes2.size()), "download phase-two routes");
            pin_off2 = fs.pin_off2;
        }
    }
    if (o.report_score) {
        double wl = 0;
        long long vias = 0;
        ok(gr_stats_download(ctx, &wl, &vias), "statistics");
        printf("SCORE wirelength=%.0f vias=%lld rerouted=%d\n", wl, vias, rerouted);
    }
    t0 = Clock::now();
    if (pin_off2)
        ok(gr_design_write_output2(d, routes.data(), routes2.data(), pin_off2, 100, o.out.c_str()), "write output");
    else
        ok(gr_design_write_output(d, routes.data(), o.out.c_str()), "write output");
    const double output_ms = ms_since(t0);
    if (o.profile_stages)
        printf("PIPELINE_TIMING setup_ms=%.3f lshape_ms=%.3f flt_ms=%.3f classification_ms=%.3f output_ms=%.3f\n",
               setup_ms, batch_ms + dag_ms, flt_ms, class_ms, output_ms);
    if (o.stop_after_lshape) printf("ROUTING_STOP stage=lshape flt_executed=0\n");
    ok(gr_ctx_destroy(ctx), "destroy context");
    ok(gr_design_destroy(d), "destroy design");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(parse(argc, argv));
    } catch (const std::exception& e) {
        fprintf(stderr, "GRotor error: %s\n", e.what());
        return 1;
    }
}
