// Synthetic CAP/NET instances for the five workload configurations.
//
// Cost header, layer names and edge lengths follow the real CAP fixture shipped
// with the reference (unit wire 0.00131579, unit via 4, overflow 5 per layer,
// edge length 3800, min_len 400).  Metal 1 is pin-only with capacity 0, the
// other layers alternate direction starting with metal 1 horizontal.  Every
// pin has one access point on metal 1 (layer index 0).  mt19937_64 drives all
// randomness.
//
//   1: 128x128x5, cap U{4..12}, 10 blockages (side U{8..24}, cap U{0..2}),
//      20k nets, pins 2 (p .6) | U{3..5} (p .3) | 6+Geom(.5) <= 50,
//      bbox side round(lognormal(median 6, sigma .75)) <= 64.
//   2: 512x512x9, cap U{6..16}, 5% of routing cells blocked, 300k nets,
//      pin counts as 1, bbox median 10 (<= 256), 8% of nets get U{20..120}
//      pins clustered with Gaussian sigma = bbox / 6.
//   3: 1024x1024x10, as 2 with 1M nets and pin counts Zipf(2.2) on [2, 200].
//   4: as 2 (the differentiable-DP timing instance).
//   5: 2048x2048x10, cap U{8..20}, 40 macros covering 3% of cells, 4M nets,
//      pins 2 (p .55) | U{3..8} (p .35) | Zipf(2.2) on [9, 300],
//      bbox median 12 (<= 1024).
//   6: 64x64x5, 2k nets, distributions of 1 (finite-difference instance).
//   7: 64x64x10, 1500 nets, distributions of 1, and 30% of pins with 2-3
//      access points on metal 1-3 around the pin (front-end fixture; the
//      reference reads 10-layer grids only).
#include <algorithm>
#include <cmath>
#include <random>
#include <set>
#include <stdexcept>

#include "design.hpp"

namespace grotor {

namespace {

struct Config {
    int side, L, nets;
    int cap_lo, cap_hi;
    int blockages;       // rectangles
    int block_side_lo, block_side_hi, block_cap_hi;
    double blocked_fraction;  // independent blocked routing cells
    double macro_fraction;    // area covered by `blockages` macros
    double bbox_median;
    int bbox_cap;
    int pin_model;  // 1: configs 1/2/4, 3: Zipf 2..200, 5: config 5
    double cluster_fraction;
    double multi_access = 0.0;  // fraction of pins with several access points
};

Config config_of(int c) {
    switch (c) {
        case 1: return {128, 5, 20000, 4, 12, 10, 8, 24, 2, 0.0, 0.0, 6, 64, 1, 0.0};
        case 2: return {512, 9, 300000, 6, 16, 0, 0, 0, 0, 0.05, 0.0, 10, 256, 1, 0.08};
        case 3: return {1024, 10, 1000000, 6, 16, 0, 0, 0, 0, 0.05, 0.0, 10, 256, 3, 0.0};
        case 4: return {512, 9, 300000, 6, 16, 0, 0, 0, 0, 0.05, 0.0, 10, 256, 1, 0.08};
        case 5: return {2048, 10, 4000000, 8, 20, 40, 0, 0, 0, 0.0, 0.03, 12, 1024, 5, 0.0};
        case 6: return {64, 5, 2000, 4, 12, 10, 8, 24, 2, 0.0, 0.0, 6, 64, 1, 0.0};
        case 7: return {64, 10, 1500, 4, 12, 6, 6, 16, 2, 0.0, 0.0, 6, 64, 1, 0.0, 0.3};
        default: throw std::invalid_argument("unknown workload configuration");
    }
}

struct Zipf {
    std::vector<double> cdf;
    int lo;
    Zipf(double s, int a, int b) : lo(a) {
        double acc = 0;
        for (int k = a; k <= b; ++k) cdf.push_back(acc += std::pow(double(k), -s));
        for (auto& v : cdf) v /= acc;
    }
    int operator()(std::mt19937_64& rng) const {
        const double u = std::uniform_real_distribution<double>(0.0, 1.0)(rng);
        return lo + int(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    }
};

}  // namespace

void generate_design(const GenSpec& spec, Design& design) {
    Config cfg = config_of(spec.config);
    if (spec.grid_override > 0) cfg.side = spec.grid_override;
    std::mt19937_64 rng(spec.seed);
    const int X = cfg.side, Y = cfg.side, L = cfg.L;
    Grid& g = design.grid;
    g = Grid{};
    g.L = L, g.X = X, g.Y = Y;
    g.unit_length_wire_cost = 0.00131579;
    g.unit_via_cost = 4;
    g.unit_length_short_costs.assign(L, 5.0);
    g.x_edge_len.assign(X - 1, 3800);
    g.y_edge_len.assign(Y - 1, 3800);
    g.layers.resize(L);
    for (int l = 0; l < L; ++l) g.layers[l] = {"metal" + std::to_string(l + 1), l & 1, 400.0};
    g.capacity.assign(size_t(L) * X * Y, 0.f);
    auto cap = [&](int l, int x, int y) -> float& { return g.capacity[(size_t(l) * X + x) * Y + y]; };
    std::uniform_int_distribution<int> ucap(cfg.cap_lo, cfg.cap_hi);
    for (int l = 1; l < L; ++l)
        for (int x = 0; x < X; ++x)
            for (int y = 0; y < Y; ++y) cap(l, x, y) = float(ucap(rng));
    if (cfg.blocked_fraction > 0) {
        std::uniform_real_distribution<double> u(0.0, 1.0);
        for (int l = 1; l < L; ++l)
            for (int x = 0; x < X; ++x)
                for (int y = 0; y < Y; ++y)
                    if (u(rng) < cfg.blocked_fraction) cap(l, x, y) = 0.f;
    }
    for (int b = 0; b < cfg.blockages; ++b) {
        int w, h, c;
        if (cfg.macro_fraction > 0) {
            const double area = cfg.macro_fraction * X * Y / cfg.blockages;
            std::uniform_real_distribution<double> aspect(0.5, 2.0);
            const double a = aspect(rng);
            w = std::max(1, std::min(X, int(std::sqrt(area * a))));
            h = std::max(1, std::min(Y, int(area / w)));
            c = 0;
        } else {
            std::uniform_int_distribution<int> side(std::min(cfg.block_side_lo, X), std::min(cfg.block_side_hi, X));
            w = side(rng), h = side(rng);
            c = std::uniform_int_distribution<int>(0, cfg.block_cap_hi)(rng);
        }
        const int x0 = std::uniform_int_distribution<int>(0, X - w)(rng);
        const int y0 = std::uniform_int_distribution<int>(0, Y - h)(rng);
        for (int l = 1; l < L; ++l)
            for (int x = x0; x < x0 + w; ++x)
                for (int y = y0; y < y0 + h; ++y) cap(l, x, y) = float(c);
    }

    const int nets = std::max(1, int(std::llround(cfg.nets * spec.net_scale)));
    std::lognormal_distribution<double> side_dist(std::log(cfg.bbox_median), 0.75);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    const Zipf zipf3(2.2, 2, 200), zipf5(2.2, 9, 300);
    design.nets.clear();
    design.nets.reserve(nets);
    for (int i = 0; i < nets; ++i) {
        int pins;
        bool clustered = false;
        if (cfg.pin_model == 3) {
            pins = zipf3(rng);
        } else if (cfg.pin_model == 5) {
            const double r = u01(rng);
            if (r < 0.55) pins = 2;
            else if (r < 0.90) pins = std::uniform_int_distribution<int>(3, 8)(rng);
            else pins = zipf5(rng);
        } else {
            const double r = u01(rng);
            if (r < 0.6) pins = 2;
            else if (r < 0.9) pins = std::uniform_int_distribution<int>(3, 5)(rng);
            else pins = std::min(50, 6 + std::geometric_distribution<int>(0.5)(rng));
            if (cfg.cluster_fraction > 0 && u01(rng) < cfg.cluster_fraction) {
                pins = std::uniform_int_distribution<int>(20, 120)(rng);
                clustered = true;
            }
        }
        auto side = [&](int limit) {
            const int s = int(std::lround(side_dist(rng)));
            return std::max(1, std::min({s, cfg.bbox_cap, limit}));
        };
        const int w = side(X), h = side(Y);
        const int x0 = std::uniform_int_distribution<int>(0, X - w)(rng);
        const int y0 = std::uniform_int_distribution<int>(0, Y - h)(rng);
        DesignNet net;
        net.name = "n" + std::to_string(i);
        std::set<int> used;
        std::normal_distribution<double> gx(x0 + (w - 1) / 2.0, std::max(0.5, w / 6.0));
        std::normal_distribution<double> gy(y0 + (h - 1) / 2.0, std::max(0.5, h / 6.0));
        int minx = X, maxx = 0, miny = Y, maxy = 0;
        for (int p = 0; p < pins; ++p) {
            int x, y;
            if (clustered) {
                x = std::min(x0 + w - 1, std::max(x0, int(std::lround(gx(rng)))));
                y = std::min(y0 + h - 1, std::max(y0, int(std::lround(gy(rng)))));
            } else {
                x = std::uniform_int_distribution<int>(x0, x0 + w - 1)(rng);
                y = std::uniform_int_distribution<int>(y0, y0 + h - 1)(rng);
            }
            if (!used.insert(x * Y + y).second) continue;  // one pin per cell
            std::vector<int> ap{1, x * Y + y};
            if (cfg.multi_access > 0 && u01(rng) < cfg.multi_access) {
                const int extra = std::uniform_int_distribution<int>(1, 2)(rng);
                for (int e = 0; e < extra; ++e) {
                    const int l = std::uniform_int_distribution<int>(0, 2)(rng);
                    const int ex = std::min(X - 1, std::max(0, x + std::uniform_int_distribution<int>(-1, 1)(rng)));
                    const int ey = std::min(Y - 1, std::max(0, y + std::uniform_int_distribution<int>(-1, 1)(rng)));
                    // access points of one pin occupy distinct 2D cells (the
                    // reference's selection asserts this, src/input.hpp:179)
                    bool seen = false;
                    for (int j = 1; j <= ap[0]; ++j) seen |= ap[j] % (X * Y) == ex * Y + ey;
                    if (seen) continue;
                    ap.push_back(l * X * Y + ex * Y + ey);
                    ap[0]++;
                    minx = std::min(minx, ex), maxx = std::max(maxx, ex), miny = std::min(miny, ey), maxy = std::max(maxy, ey);
                }
            }
            net.access_points.push_back(ap);
            minx = std::min(minx, x), maxx = std::max(maxx, x), miny = std::min(miny, y), maxy = std::max(maxy, y);
        }
        net.minx = minx, net.maxx = maxx, net.miny = miny, net.maxy = maxy;
        design.nets.push_back(std::move(net));
    }
}

}  // namespace grotor
