// grotor: the rip-up stage pipeline on one MI355X, over the C ABI only.
//
// Same command line as the reference router (src/main.cu:36-84):
//   grotor -cap FILE -net FILE -out FILE [--batching cpu|hiprt] [--validate-batches 0|1]
//          [--sparsity N] [--lut-dir DIR] [--profile-stages 0|1] [--report-score 0|1]
//          [--stop-after lshape|routing]
// plus this build's own options:
//          [--dag-mode 0|1|2]      batch loops of both stages: plain form (0, the
//                                  reference's execution), launch form with LDS
//                                  tables (1), persistent form (2, the default)
//          [--threads N]           host threads of the front end (default 8)
//          [--flt-stages N]        rip-up -> FLT stages (default 1, as the
//                                  reference); stage k > 1 reroutes the nets
//                                  still over the threshold, removing each
//                                  net's current route, and logs its records
//                                  as flt<k> / detour<k>
// and the same stages (grotor::route, src/pipeline.hpp:14-157): L-shape stage
// over every subnet, rip-up of the nets crossing more than 10 overflowed
// edges, congestion view, FLT DAGs, the FLT stage, .out output.
//
// Batching: `cpu` (the reference's default) is the host first-fit greedy
// (gr_batch_plan_host); `hiprt` names the reference's GPU batching option,
// which here is this build's exact GPU colouring (gr_batch_geometry +
// gr_batch_colors; CDNA4 has no ray-tracing units, `gpu` is accepted as
// well).  `--validate-batches 1` recomputes the batches on the host after the
// GPU interval and requires identical lists, as the reference's HIPRT path
// does (src/batching/hiprt/grid_chunk.cuh:545-551).
//
// The run log uses the reference's record formats, so its experiment scripts
// read it: BATCH_TIMING (src/batching/generate_batches.hpp:21-23,
// src/batching/hiprt/grid_chunk.cuh:552-553), DAG_TIMING
// (include/router_state.hpp:259-264), GROTOR_CONFIG / DAG_SCHEDULING /
// SCORE_DIAGNOSTICS / GROTOR_COMPLETE (src/main.cu:126-166),
// HOST_MEMORY_POLICY and PIPELINE_TIMING (src/pipeline.hpp:44,154).
// Timestamps are microseconds of std::chrono::steady_clock since the start
// of the program (the reference's batch_clock_origin,
// include/router_state.hpp:235).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grotor_amd.h"

namespace {

using Clock = std::chrono::steady_clock;
const Clock::time_point clock_origin = Clock::now();

void ok(int rc, const char* op) {
    if (rc) throw std::runtime_error(std::string(op) + ": " + gr_last_error());
}

double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

long long us_of(Clock::time_point t) {
    return (long long)std::chrono::duration_cast<std::chrono::microseconds>(t - clock_origin).count();
}

long long us_of_ns(long long steady_ns) {
    const Clock::time_point t{std::chrono::duration_cast<Clock::duration>(std::chrono::nanoseconds(steady_ns))};
    return us_of(t);
}

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
    int sparsity = 1, dag_mode = 2, threads = 8;
    int flt_stages = 1;  // rip-up -> FLT stages (the reference runs one)
    bool gpu_batching = false, validate = false;
    bool profile_stages = false, report_score = true, stop_after_lshape = false;
};

int int_option(const std::string& opt, const std::string& val, int lo, int hi) {
    size_t used = 0;
    int v = 0;
    try {
        v = std::stoi(val, &used);
    } catch (const std::exception&) {
        used = 0;
    }
    if (used != val.size() || v < lo || v > hi)
        throw std::invalid_argument(opt + " expects an integer in [" + std::to_string(lo) + ", " + std::to_string(hi) + "]");
    return v;
}

Options parse(int argc, char** argv) {
    const char* usage =
        "Usage: grotor -cap FILE -net FILE -out FILE [--batching cpu|hiprt] [--validate-batches 0|1] [--sparsity N] "
        "[--lut-dir DIR] [--profile-stages 0|1] [--report-score 0|1] [--stop-after lshape|routing] "
        "[--dag-mode 0|1|2] [--threads N] [--flt-stages N]";
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
            // the tree construction reads no lookup tables
        } else if (opt == "--batching") {
            if (val == "rtcore") throw std::invalid_argument("HIP uses --batching hiprt");
            if (val != "cpu" && val != "hiprt" && val != "gpu")
                throw std::invalid_argument("--batching expects cpu, rtcore or hiprt");
            o.gpu_batching = val != "cpu";
        } else if (opt == "--validate-batches") flag(o.validate);
        else if (opt == "--report-score") flag(o.report_score);
        else if (opt == "--profile-stages") flag(o.profile_stages);
        else if (opt == "--stop-after") {
            if (val != "lshape" && val != "routing") throw std::invalid_argument("--stop-after expects lshape or routing");
            o.stop_after_lshape = val == "lshape";
        } else if (opt == "--sparsity") {
            size_t used = 0;
            try {
                o.sparsity = std::stoi(val, &used);
            } catch (const std::exception&) {
                used = 0;
            }
            if (used != val.size() || o.sparsity < 1) throw std::invalid_argument("Sparsity must be a positive integer");
        } else if (opt == "--dag-mode") o.dag_mode = int_option(opt, val, 0, 2);
        else if (opt == "--threads") o.threads = int_option(opt, val, 1, 256);
        else if (opt == "--flt-stages") o.flt_stages = int_option(opt, val, 1, 16);
        else throw std::invalid_argument("Unknown option: " + opt);
    }
    if (o.cap.empty() || o.net.empty() || o.out.empty()) throw std::invalid_argument("-cap, -net and -out are required");
    if (o.validate && !o.gpu_batching)
        throw std::invalid_argument("--validate-batches requires ray-tracing batching (rtcore on CUDA, hiprt on HIP)");
    if (o.gpu_batching && o.sparsity != 1) throw std::invalid_argument("Ray-tracing batching requires --sparsity 1");
    return o;
}

// Geometry of one stage's nets in priority order.
struct Geometry {
    int nets;
    const int *pt_off, *pts, *hs_off, *hsegs, *vs_off, *vsegs;
};

// One stage's batching call, logged as the reference logs it.  Returns the
// colors and the batch count.
int batch_stage(const Options& o, gr_ctx* ctx, int X, int Y, const Geometry& g, int limit, const char* stage,
                std::vector<int>& color) {
    color.assign(std::max(g.nets, 1), 0);
    int nb = 0;
    if (!o.gpu_batching) {
        // device work before the call is excluded (generate_batches.hpp:15)
        ok(gr_ctx_synchronize(ctx), "synchronize before CPU batching");
        const auto t0 = Clock::now();
        ok(gr_batch_plan_host(X, Y, g.nets, g.pt_off, g.pts, g.hs_off, g.hsegs, g.vs_off, g.vsegs, limit, o.sparsity,
                              color.data(), &nb),
           "CPU batching");
        const auto t1 = Clock::now();
        const long long a = us_of(t0), b = us_of(t1);
        printf("BATCH_BACKEND=cpu\n");
        printf("BATCH_TIMING backend=cpu start_us=%lld end_us=%lld elapsed_ms=%.6f nets=%d batches=%d used_for_routing=1\n",
               a, b, (b - a) / 1000.0, g.nets, nb);
        return nb;
    }
    // geometry upload before the interval, as the reference's scene build
    // (src/batching/hiprt/grid_chunk.cuh:373-428)
    ok(gr_batch_geometry(ctx, g.nets, g.pt_off, g.pts, g.hs_off, g.hsegs, g.vs_off, g.vsegs), "batching geometry");
    ok(gr_ctx_synchronize(ctx), "synchronize before GPU batching");
    const auto t0 = Clock::now();
    ok(gr_batch_colors(ctx, limit, o.sparsity, color.data(), &nb), "GPU batching");
    const auto t1 = Clock::now();
    if (o.validate) {
        std::vector<int> cpu(color.size());
        int cnb = 0;
        ok(gr_batch_plan_host(X, Y, g.nets, g.pt_off, g.pts, g.hs_off, g.hsegs, g.vs_off, g.vsegs, limit, 1, cpu.data(),
                              &cnb),
           "CPU batching (validation)");
        bool equal = cnb == nb;
        for (int k = 0; equal && k < g.nets; ++k) equal = cpu[k] == color[k];
        printf("BATCH_VALIDATION stage=%s nets=%d hiprt_batches=%d cpu_batches=%d exact_lists_equal=%d\n", stage, g.nets,
               nb, cnb, int(equal));
        if (!equal) throw std::runtime_error("GPU batches differ from CPU (members/order)");
    }
    const long long a = us_of(t0), b = us_of(t1);
    printf("BATCH_SCOPE backend=hiprt scope=generate_batch_ids_call_wall\n");
    printf("BATCH_TIMING backend=hiprt start_us=%lld end_us=%lld elapsed_ms=%.6f nets=%d batches=%d used_for_routing=1\n",
           a, b, (b - a) / 1000.0, g.nets, nb);
    return nb;
}

void print_dag_timing(gr_ctx* ctx, const char* stage, int nets, int batches) {
    long long s_ns = 0, e_ns = 0;
    ok(gr_ctx_last_interval(ctx, &s_ns, &e_ns), "DAG interval");
    const long long a = us_of_ns(s_ns), b = us_of_ns(e_ns);
    printf("DAG_TIMING stage=%s batch_index=-1 start_us=%lld end_us=%lld elapsed_ms=%.6f nets=%d batches=%d\n", stage, a,
           b, (b - a) / 1000.0, nets, batches);
}

int run(const Options& o) {
    const auto t_start = Clock::now();
    char name[256];
    if (!gr_device_name(name, sizeof name)) printf("GPU_PLATFORM backend=HIP device=%s wave=64\n", name);
    const char* batching = o.gpu_batching ? "hiprt" : "cpu";
    printf("GROTOR_CONFIG batching=%s cost=incremental prefix=bit memory=explicit\n", batching);
    // lshape: one wavefront per net; flt: one wavefront per DAG (the
    // reference's record names them warp and block)
    printf("DAG_SCHEDULING lshape=warp flt=block\n");
    printf("CONGESTION_REDUCTION=per-cell\n");
    printf("SCORE_DIAGNOSTICS enabled=%d\n", int(o.report_score));
    gr_design* d = nullptr;
    ok(gr_design_read(o.cap.c_str(), o.net.c_str(), &d), "read input");
    gr_grid_desc g{};
    gr_topology t{};
    gr_ctx* ctx = nullptr;
    printf("HOST_MEMORY_POLICY saving=0 bounded_preparation=0\n");
    const auto t_topo = Clock::now();
    ok(gr_design_prepare(d, o.threads), "prepare nets");
    ok(gr_design_topology(d, &t), "topology");
    const double topo_ms = ms_since(t_topo);
    g = gr_grid_desc{t.L, t.X, t.Y, t.DIR, t.unit_length_wire_cost, t.unit_via_cost, t.unit_length_short_costs,
                     t.x_edge_len, t.y_edge_len, t.capacity};
    ok(gr_ctx_create(&g, &ctx), "create context");
    ok(gr_ctx_set_routes(ctx, t.subnets, t.pin_acc_num), "route storage");
    const double setup_ms = ms_since(t_start);

    // L-shape stage (src/routing/lshape.cuh:40-196)
    printf("TOPOLOGY_TIMING stage=lshape elapsed_ms=%.6f geometry_fused=0\n", topo_ms);
    std::vector<int> color;
    const Geometry lg{t.nets, t.pt_off, t.pts, t.hs_off, t.hsegs, t.vs_off, t.vsegs};
    const int nb = batch_stage(o, ctx, t.X, t.Y, lg, 1000000, "lshape", color);
    Schedule s = lshape_schedule(t, color, nb);
    ok(gr_lshape_upload(ctx, t.nets, s.net_ids.data(), s.node_off.data(), s.nodes.data(), s.parents.data()),
       "L-shape topology upload");
    double dag_ms = 0;
    ok(gr_lshape_run(ctx, nb, s.batch_off.data(), 1, o.dag_mode, &dag_ms), "L-shape stage");
    print_dag_timing(ctx, "lshape", t.nets, nb);
    const double lshape_ms = ms_since(t_topo) - topo_ms;
    auto t0 = Clock::now();
    const long long words = 12LL * t.pin_acc_num[t.subnets] + 1;
    std::vector<int> routes(words);
    ok(gr_routes_download(ctx, routes.data(), words), "download routes");
    double download_ms = ms_since(t0);
    double flt_ms = 0, class_ms = 0;
    int rerouted = 0;
    std::vector<int> routes2, pin_off2;
    long long pins2 = 0;
    if (!o.stop_after_lshape) {
        // rip-up classification, the congestion view and the FLT stage
        // (src/pipeline.hpp:86-137); with --flt-stages N > 1 the stage
        // repeats on the nets still over the rip-up threshold, each net's
        // current route removed and phase-two storage appended for nets
        // without it (the reroute storage of src/pipeline_cugr2.hpp:15-62)
        std::vector<unsigned char> flags(std::max(t.subnets, 1));
        std::vector<float> xsum(size_t(t.X) * t.Y), ysum(size_t(t.X) * t.Y);
        int stamp = nb + 1;  // the L-shape batches took stamps 1..nb
        for (int stage = 1; stage <= o.flt_stages; ++stage) {
            t0 = Clock::now();
            ok(gr_congestion_view(ctx, nullptr, xsum.data(), ysum.data()), "congestion view");
            ok(gr_ripup(ctx, 0, flags.data()), "rip-up");
            class_ms += ms_since(t0);
            const auto t_flt = Clock::now();
            const std::string suffix = stage == 1 ? "" : std::to_string(stage);
            gr_flt_stage fs{};
            ok(gr_design_flt_stage(d, flags.data(), xsum.data(), ysum.data(), o.threads, &fs), "FLT topology");
            printf("TOPOLOGY_TIMING stage=flt%s elapsed_ms=%.6f geometry_fused=0\n", suffix.c_str(), ms_since(t_flt));
            if (stage > 1) printf("RIPUP_STAGE index=%d nets=%d\n", stage, fs.nets);
            if (!fs.nets) break;
            if (stage == 1) {
                ok(gr_ctx_set_routes_phase2(ctx, fs.pin_off2, fs.pins2, 100), "phase-two route storage");
                pin_off2.assign(fs.pin_off2, fs.pin_off2 + t.subnets);
                pins2 = fs.pins2;
            } else {
                std::vector<int> ids(fs.order, fs.order + fs.nets);
                std::sort(ids.begin(), ids.end());
                ok(gr_ctx_reserve_routes_phase2(ctx, int(ids.size()), ids.data(), 100, pin_off2.data(), &pins2),
                   "phase-two route storage");
                printf("REROUTE_STORAGE pins=%lld stride=100 bytes=%lld\n", pins2, (100 * pins2 + 1) * 4);
            }
            std::vector<int> fcolor;
            const Geometry fg{fs.nets, fs.pt_off, fs.pts, fs.hs_off, fs.hsegs, fs.vs_off, fs.vsegs};
            const int fnb = batch_stage(o, ctx, t.X, t.Y, fg, 300000, ("flt" + suffix).c_str(), fcolor);
            const auto t_prep = Clock::now();
            gr_flt_plan plan{};
            ok(gr_design_flt_plan(d, fcolor.data(), fnb, &plan), "FLT plan");
            printf("FLT_PREPARATION elapsed_ms=%.6f batches=%d\n", ms_since(t_prep), plan.batches);
            ok(gr_flt_upload(ctx, plan.batches, plan.batch_off, plan.net_ids, plan.node_off, plan.nodes, plan.edge_off,
                             plan.edges),
               "FLT upload");
            double ms = 0;
            ok(gr_flt_run(ctx, stamp, stage == 1 ? 1 : 2, o.dag_mode, &ms), "FLT stage");
            stamp += 2 * plan.batches;  // a removal and a commit per batch
            print_dag_timing(ctx, ("detour" + suffix).c_str(), fs.nets, plan.batches);
            flt_ms += ms_since(t_flt);
        }
        for (int off : pin_off2) rerouted += off >= 0;
        if (!pin_off2.empty()) {
            t0 = Clock::now();
            routes2.resize(100LL * pins2 + 1);
            ok(gr_routes2_download(ctx, routes2.data(), (long long)routes2.size()), "download phase-two routes");
            download_ms += ms_since(t0);
        }
    }
    if (o.report_score) {
        double wl = 0;
        long long vias = 0;
        ok(gr_stats_download(ctx, &wl, &vias), "statistics");
        printf("SCORE wirelength=%.0f vias=%lld rerouted=%d\n", wl, vias, rerouted);
        // the evaluator's completion test of every net (resources/evaluator.cpp:527-538, 677-749)
        int nn = 0, na = 0, metal = 0;
        const int *pin_off = nullptr, *pins = nullptr, *extra_off = nullptr, *extra = nullptr, *sub_off = nullptr,
                  *subs = nullptr, *apin = nullptr, *aoff = nullptr, *acc = nullptr;
        ok(gr_design_nets(d, &nn, &pin_off, &pins, &extra_off, &extra, &sub_off, &subs), "design nets");
        ok(gr_design_access(d, &na, &apin, &aoff, &acc), "access points");
        ok(gr_design_metal_grid(d, &metal, nullptr, nullptr, nullptr), "metal grid");
        std::vector<int> dirs(std::max(metal, 1));
        ok(gr_design_metal_grid(d, &metal, dirs.data(), nullptr, nullptr), "metal grid");
        long long counts[3] = {0, 0, 0};
        ok(gr_check_connectivity(ctx, nn, pin_off, pins, extra_off, extra, sub_off, subs, dirs.data(), apin, aoff,
                                 acc, nullptr, counts),
           "connectivity");
        printf("SCORE_CONNECTIVITY open=%lld not_connected=%lld incomplete=%lld\n", counts[0], counts[1], counts[2]);
    }
    t0 = Clock::now();
    if (!pin_off2.empty())
        ok(gr_design_write_output2(d, routes.data(), routes2.data(), pin_off2.data(), 100, o.out.c_str()),
           "write output");
    else
        ok(gr_design_write_output(d, routes.data(), o.out.c_str()), "write output");
    const double output_ms = ms_since(t0);
    if (o.stop_after_lshape) printf("ROUTING_STOP stage=lshape flt_executed=0\n");
    t0 = Clock::now();
    ok(gr_ctx_destroy(ctx), "destroy context");
    const double cleanup_ms = ms_since(t0);
    if (o.profile_stages && !o.stop_after_lshape)
        printf("PIPELINE_TIMING setup_ms=%.3f lshape_ms=%.3f flt_ms=%.3f download_ms=%.3f classification_ms=%.3f "
               "output_ms=%.3f output_wait_ms=%.3f cleanup_ms=%.3f topology_cleanup_ms=%.3f\n",
               setup_ms, lshape_ms, flt_ms, download_ms, class_ms, output_ms, output_ms, cleanup_ms, 0.0);
    ok(gr_design_destroy(d), "destroy design");
    printf("GROTOR_COMPLETE wall_seconds=%.6f\n", std::chrono::duration<double>(Clock::now() - clock_origin).count());
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
