// Golden-vector harness over the reference's own host code -- TEST
// INFRASTRUCTURE ONLY, built into oracle/_ref/ by oracle/Makefile from the
// reference sources where they lie; never shipped or linked by the product.
//
// Runs, in the reference's order: db::read (src/input.hpp:247-419, with
// net::init pin selection), flute::readLUT, prepare_routing_nets
// (src/preprocess.hpp:1448-1623), construct_rsmt for every subnet
// (src/routing/lshape.cuh:33-44), the hpwl sort (src/routing/lshape.cuh:53-55)
// and generate_batches_rsmt (src/batching/generate_batches_cpu.hpp:7-73),
// then prints one JSON object per line:
//   {"grid": ...}, {"net": ...} per design net, {"sub": ...} per subnet,
//   {"order": [...]}, {"batches": [[...], ...]}.
// With a congestion-view file (X*Y float32 xsum, then X*Y float32 ysum, in
// x*Y+y order) it then runs net::generate_detours_reconstruct
// (src/preprocess.hpp:769-1436) on every subnet tree of two or more nodes
// and prints {"dag": subnet, "nodes", "depth", "kids", "npar", "par",
// "slot", "points"} per subnet: the FLT builder's arrays.
//
// usage: ref_harness <cap> <net> <flute_lut_dir> [sparsity [view_file]]
#include <unistd.h>

#include "batching/generate_batches_cpu.hpp"
#include "preprocess.hpp"
#include "flt_topology.hpp"

static void print_json_string(const std::string& s) {
    putchar('"');
    for (char ch : s) {
        if (ch == '"' || ch == '\\') putchar('\\');
        putchar(ch);
    }
    putchar('"');
}

static void print_list(const std::vector<int>& v) {
    printf("[");
    for (size_t i = 0; i < v.size(); ++i) printf(i ? ",%d" : "%d", v[i]);
    printf("]");
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s cap net flute_lut_dir [sparsity]\n", argv[0]);
        return 2;
    }
    const int sparsity = argc > 4 ? atoi(argv[4]) : 1;
    program_start = std::chrono::high_resolution_clock::now();
    // keep the reference's progress lines off stdout
    const int saved = dup(1);
    if (!freopen("/dev/null", "w", stdout)) return 3;
    db::read(argv[1], argv[2]);
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return 3;
    if (chdir(argv[3])) return 4;
    flute::readLUT();
    if (chdir(cwd)) return 4;
    cudb::RouterState state;
    cudb::prepare_routing_nets(state);
    auto& nets = state.nets;
    for (auto& n : nets) n.construct_rsmt(state, false, false);
    std::vector<int> order(nets.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
    std::sort(order.begin(), order.end(), [&](int l, int r) { return nets[l].hpwl > nets[r].hpwl; });
    auto work = order;
    auto batches = cudb::generate_batches_rsmt(state, work, 1000000, sparsity);
    fflush(stdout);
    dup2(saved, 1);
    close(saved);
    stdout = fdopen(1, "w");
    const auto& v = state.device;
    printf("{\"grid\":{\"L\":%d,\"X\":%d,\"Y\":%d,\"DIR\":%d,\"sparsity\":%d}}\n", v.L, v.X, v.Y, v.DIR, sparsity);
    for (size_t i = 0; i < db::nets.size(); ++i) {
        const auto& n = db::nets[i];
        printf("{\"net\":%zu,\"name\":", i);
        print_json_string(n.name);
        printf(",\"pins\":");
        print_list(n.pins);
        printf(",\"extra\":");
        print_list(n.extra_routes);
        printf(",\"subnets\":");
        print_list(n.subnets);
        printf("}\n");
    }
    for (size_t i = 0; i < nets.size(); ++i) {
        const auto& n = nets[i];
        printf("{\"sub\":%zu,\"orig\":%d,\"hpwl\":%d,\"pins\":", i, n.original_net_id, n.hpwl);
        print_list(n.pins);
        printf(",\"nodes\":");
        print_list(n.rsmt.back());
        printf(",\"adj\":[");
        for (size_t j = 0; j + 1 < n.rsmt.size(); ++j) {
            if (j) printf(",");
            print_list(n.rsmt[j]);
        }
        printf("],\"points\":");
        print_list(n.points);
        std::vector<int> h, w;
        for (auto& e : n.rsmt_h_segments) h.push_back(e.first), h.push_back(e.second);
        for (auto& e : n.rsmt_v_segments) w.push_back(e.first), w.push_back(e.second);
        printf(",\"hsegs\":");
        print_list(h);
        printf(",\"vsegs\":");
        print_list(w);
        printf("}\n");
    }
    printf("{\"order\":");
    print_list(order);
    printf("}\n{\"batches\":[");
    for (size_t b = 0; b < batches.size(); ++b) {
        if (b) printf(",");
        print_list(batches[b]);
    }
    printf("]}\n");
    if (argc > 5) {
        const size_t cells = size_t(v.X) * v.Y;
        std::vector<float> xsum(cells), ysum(cells);
        std::vector<char> busy(cells, 0);
        FILE* f = fopen(argv[5], "rb");
        if (!f || fread(xsum.data(), sizeof(float), cells, f) != cells ||
            fread(ysum.data(), sizeof(float), cells, f) != cells)
            return 5;
        fclose(f);
        bool* busy_view = reinterpret_cast<bool*>(busy.data());
        float* xs = xsum.data();
        float* ys = ysum.data();
        grotor::flt::Builder b;
        for (size_t i = 0; i < nets.size(); ++i) {
            auto& n = nets[i];
            if (n.rsmt.size() < 3) continue;
            int saved_out = dup(1);
            if (!freopen("/dev/null", "w", stdout)) return 3;
            n.generate_detours_reconstruct(state, b, busy_view, xs, ys, false, false);
            fflush(stdout);
            dup2(saved_out, 1);
            close(saved_out);
            stdout = fdopen(1, "w");
            const int cnt = b.node_index_cnt;
            printf("{\"dag\":%zu,\"nodes\":", i);
            print_list(std::vector<int>(b.nodes_cpu.begin(), b.nodes_cpu.begin() + cnt));
            printf(",\"depth\":");
            print_list(std::vector<int>(b.node_depth_cpu.begin(), b.node_depth_cpu.begin() + cnt));
            printf(",\"kids\":");
            print_list(std::vector<int>(b.child_num_cpu.begin(), b.child_num_cpu.begin() + cnt));
            printf(",\"npar\":");
            print_list(std::vector<int>(b.par_num_cpu.begin(), b.par_num_cpu.begin() + cnt));
            printf(",\"par\":");
            print_list(b.par_nodes_cpu);
            printf(",\"slot\":");
            print_list(b.currentChildIDX_cpu);
            printf(",\"points\":");
            print_list(n.points);
            printf("}\n");
        }
    }
    return 0;
}
