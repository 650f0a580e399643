// Host front end: CAP/NET I/O, access-point selection, tree construction,
// subnet split and batching geometry.  See design.hpp for the reference
// functions each step restates.
#include "design.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace grotor {

namespace {

std::vector<char> slurp(const std::string& path) {
    std::ifstream file(path, std::ios::binary | std::ios::ate);
    if (!file.good()) throw std::invalid_argument("failed to open '" + path + "'");
    const size_t size = size_t(file.tellg());
    file.seekg(0, std::ios::beg);
    std::vector<char> buffer(size + 1);
    file.read(buffer.data(), std::streamsize(size));
    buffer[size] = 0;
    return buffer;
}

struct Reader {
    const std::vector<char>& b;
    size_t p = 0;
    bool at_end() const { return p + 1 >= b.size(); }
    // src/input.hpp:251-256: skip to the next digit, read an unsigned integer
    int read_int() {
        while (!at_end() && !isdigit((unsigned char)b[p])) ++p;
        if (at_end()) throw std::runtime_error("unexpected end of input");
        int v = 0;
        for (; isdigit((unsigned char)b[p]); ++p) v = v * 10 + (b[p] - '0');
        return v;
    }
    // src/input.hpp:257-266
    double read_double() {
        while (!at_end() && !isdigit((unsigned char)b[p])) ++p;
        if (at_end()) throw std::runtime_error("unexpected end of input");
        double ans = 0, scale = 0;
        for (; isdigit((unsigned char)b[p]) || b[p] == '.'; ++p)
            if (b[p] == '.') scale = 1;
            else scale *= 10, ans = ans * 10 + (b[p] - '0');
        return scale == 0 ? ans : ans / scale;
    }
    std::string read_string() {
        while (!at_end() && (b[p] == ' ' || b[p] == '\n' || b[p] == '\r')) ++p;
        const size_t begin = p;
        while (!at_end() && b[p] != ' ' && b[p] != '\n' && b[p] != '\r') ++p;
        return std::string(b.data() + begin, p - begin);
    }
};

template <class F>
void parallel_for(int n, int threads, F&& f) {
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < n; i += threads) f(i);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

void read_design(const std::string& cap_path, const std::string& net_path, Design& design) {
    Grid& g = design.grid;
    {
        const auto buffer = slurp(cap_path);
        Reader r{buffer};
        g.L = r.read_int();
        g.X = r.read_int();
        g.Y = r.read_int();
        if (g.L < 2 || g.X < 2 || g.Y < 2) throw std::invalid_argument("CAP grid must be at least 2 x 2 x 2");
        g.unit_length_wire_cost = r.read_double();
        g.unit_via_cost = r.read_double();
        g.unit_length_short_costs.resize(g.L);
        for (int i = 0; i < g.L; ++i) g.unit_length_short_costs[i] = r.read_double();
        g.x_edge_len.resize(g.X - 1);
        g.y_edge_len.resize(g.Y - 1);
        for (auto& v : g.x_edge_len) v = r.read_int();
        for (auto& v : g.y_edge_len) v = r.read_int();
        g.layers.resize(g.L);
        g.capacity.assign(size_t(g.L) * g.X * g.Y, 0.f);
        for (int l = 0; l < g.L; ++l) {
            g.layers[l].name = r.read_string();
            g.layers[l].dir = r.read_int();
            g.layers[l].min_len = r.read_double();
            for (int y = 0; y < g.Y; ++y)
                for (int x = 0; x < g.X; ++x) g.capacity[(size_t(l) * g.X + x) * g.Y + y] = float(r.read_double());
            if (l && g.layers[l].dir == g.layers[l - 1].dir)
                throw std::invalid_argument("CAP layers must alternate direction");
        }
    }
    const auto buffer = slurp(net_path);
    const auto& b = buffer;
    size_t p = 0;
    const size_t n = b.size() - 1;
    auto skip_ws = [&] {
        while (p < n && (b[p] == ' ' || b[p] == '\n' || b[p] == '\r' || b[p] == '\t')) ++p;
    };
    const int X = g.X, Y = g.Y;
    design.nets.clear();
    skip_ws();
    while (p < n) {
        DesignNet net;
        const size_t begin = p;
        while (p < n && b[p] != ' ' && b[p] != '\n' && b[p] != '\r') ++p;
        net.name.assign(b.data() + begin, p - begin);
        int minx = X, maxx = 0, miny = Y, maxy = 0;
        // src/input.hpp:342-401: pins are "[(l, x, y), ...]" inside "( )"
        while (p < n && b[p] != '(') ++p;
        ++p;
        while (true) {
            while (p < n && b[p] != '[' && b[p] != ')') ++p;
            if (p >= n) throw std::runtime_error("unterminated net " + net.name);
            if (b[p] == ')') { ++p; break; }
            ++p;
            std::vector<int> ap(1, 0);
            while (true) {
                while (p < n && !isdigit((unsigned char)b[p]) && b[p] != ']') ++p;
                if (p >= n) throw std::runtime_error("unterminated pin in " + net.name);
                if (b[p] == ']') { ++p; break; }
                int v[3];
                for (int k = 0; k < 3; ++k) {
                    while (p < n && !isdigit((unsigned char)b[p])) ++p;
                    int a = 0;
                    for (; p < n && isdigit((unsigned char)b[p]); ++p) a = a * 10 + (b[p] - '0');
                    v[k] = a;
                }
                const int l = v[0], x = v[1], y = v[2];
                if (l >= g.L || x >= X || y >= Y) throw std::runtime_error("pin outside grid in " + net.name);
                minx = std::min(minx, x), maxx = std::max(maxx, x);
                miny = std::min(miny, y), maxy = std::max(maxy, y);
                ap.push_back(l * X * Y + x * Y + y);
                ap[0]++;
            }
            net.access_points.push_back(std::move(ap));
        }
        net.minx = minx, net.maxx = maxx, net.miny = miny, net.maxy = maxy;
        design.nets.push_back(std::move(net));
        skip_ws();
    }
}

void write_cap(const std::string& path, const Grid& g) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::string s;
    s.reserve(size_t(g.L) * g.X * g.Y * 3 + 4096);
    char buf[64];
    s += std::to_string(g.L) + " " + std::to_string(g.X) + " " + std::to_string(g.Y) + "\n";
    snprintf(buf, sizeof buf, "%.8g %.8g", g.unit_length_wire_cost, g.unit_via_cost);
    s += buf;
    for (double c : g.unit_length_short_costs) {
        snprintf(buf, sizeof buf, " %.8g", c);
        s += buf;
    }
    s += "\n";
    for (size_t i = 0; i < g.x_edge_len.size(); ++i) s += (i ? " " : "") + std::to_string(g.x_edge_len[i]);
    s += "\n";
    for (size_t i = 0; i < g.y_edge_len.size(); ++i) s += (i ? " " : "") + std::to_string(g.y_edge_len[i]);
    s += "\n";
    for (int l = 0; l < g.L; ++l) {
        snprintf(buf, sizeof buf, "%s %d %.8g\n", g.layers[l].name.c_str(), g.layers[l].dir, g.layers[l].min_len);
        s += buf;
        for (int y = 0; y < g.Y; ++y) {
            for (int x = 0; x < g.X; ++x) {
                const float c = g.capacity[(size_t(l) * g.X + x) * g.Y + y];
                if (x) s += ' ';
                if (c == float(int(c))) s += std::to_string(int(c));
                else {
                    snprintf(buf, sizeof buf, "%.6g", double(c));
                    s += buf;
                }
            }
            s += '\n';
        }
    }
    fwrite(s.data(), 1, s.size(), f);
    fclose(f);
}

void write_net(const std::string& path, const Design& design) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    const int X = design.grid.X, Y = design.grid.Y;
    std::string s;
    for (const auto& net : design.nets) {
        s += net.name;
        s += "\n(\n";
        for (const auto& ap : net.access_points) {
            s += '[';
            for (int j = 1; j <= ap[0]; ++j) {
                const int v = ap[j];
                s += (j > 1 ? ", (" : "(") + std::to_string(v / X / Y) + ", " + std::to_string(v / Y % X) + ", " +
                     std::to_string(v % Y) + ")";
            }
            s += "]\n";
        }
        s += ")\n";
        if (s.size() > (1u << 22)) {
            fwrite(s.data(), 1, s.size(), f);
            s.clear();
        }
    }
    fwrite(s.data(), 1, s.size(), f);
    fclose(f);
}

// src/input.hpp:38-243 (net::init).  The reference iterates robin_hood hash
// containers; here every container is ordered, which fixes the tie order.
static void init_net(DesignNet& net, int X, int Y) {
    const int num = int(net.access_points.size());
    const auto& ap = net.access_points;
    net.pins.clear();
    net.extra_routes.clear();
    const int center_2x = net.minx + net.maxx, center_2y = net.miny + net.maxy;
    std::set<int> p, p2D;
    if (num <= 4) {
        std::vector<int> idx(num), best(num);
        int min_diff = 1000000000, min_metric = 1000000000, min_metric2 = 1000000000, min_metric0 = 1000000000;
        std::function<void(int)> enumerate = [&](int cur) {
            if (cur == num) {
                int tmpxmin = 1000000000, tmpxmax = 0, tmpymin = 1000000000, tmpymax = 0;
                int diff = 0, metric = 0, metric2 = 0;
                for (int i = 0; i < num; i++) {
                    const int pi = ap[i][idx[i]];
                    const int xi = pi / Y % X, yi = pi % Y, li = pi / X / Y;
                    tmpxmin = std::min(tmpxmin, xi), tmpxmax = std::max(tmpxmax, xi);
                    tmpymin = std::min(tmpymin, yi), tmpymax = std::max(tmpymax, yi);
                    metric += std::abs(2 * xi - center_2x) + std::abs(2 * yi - center_2y);
                    metric2 += li;
                    for (int j = 0; j < i; j++) {
                        const int pj = ap[j][idx[j]];
                        diff += (xi != pj / Y % X) + (yi != pj % Y);
                    }
                }
                const int metric0 = tmpxmax - tmpxmin + tmpymax - tmpymin;
                if (diff < min_diff || (diff == min_diff && metric0 < min_metric0)) {
                    best = idx, min_diff = diff, min_metric0 = metric0, min_metric = metric, min_metric2 = metric2;
                } else if (diff == min_diff && metric0 == min_metric0 && metric < min_metric) {
                    best = idx, min_diff = diff, min_metric = metric, min_metric2 = metric2;
                } else if (diff == min_diff && metric0 == min_metric0 && metric == min_metric && metric2 < min_metric2) {
                    best = idx, min_metric2 = metric2;
                }
            } else {
                for (int i = 1; i <= ap[cur][0]; i++) {
                    idx[cur] = i;
                    enumerate(cur + 1);
                }
            }
        };
        enumerate(0);
        for (int i = 0; i < num; i++) {
            p.insert(ap[i][best[i]]);
            p2D.insert(ap[i][best[i]] % (X * Y));
        }
    } else {
        std::map<int, std::vector<int>> cnt_pin2d;
        std::set<int> might_overlap;
        for (int i = 0; i < num; i++)
            for (int j = 1; j <= ap[i][0]; j++) cnt_pin2d[ap[i][j] % (X * Y)].push_back(i);
        for (auto& kv : cnt_pin2d)
            for (int id : kv.second) might_overlap.insert(id);
        std::map<int, int> x_counts, y_counts;
        std::vector<int> visited(num, 0);
        for (int round = 0; round < num; round++) {
            int max_pos = -1, cnt_repeat = 0, max_coord_repeat = -1, min_dist = 1000000000;
            cnt_pin2d.clear();
            for (int i : might_overlap) {
                if (visited[i]) continue;
                for (int j = 1; j <= ap[i][0]; j++) {
                    const int pin_2d = ap[i][j] % (X * Y);
                    auto& list = cnt_pin2d[pin_2d];
                    list.push_back(i);
                    const int dist = std::abs(pin_2d / Y * 2 - center_2x) + std::abs(pin_2d % Y * 2 - center_2y);
                    const int coord_repeat = x_counts[pin_2d / Y] + y_counts[pin_2d % Y];
                    const int size = int(list.size());
                    if (size > cnt_repeat) {
                        max_pos = pin_2d, cnt_repeat = size, max_coord_repeat = coord_repeat, min_dist = dist;
                    } else if (size == cnt_repeat && coord_repeat > max_coord_repeat) {
                        max_pos = pin_2d, max_coord_repeat = coord_repeat, min_dist = dist;
                    } else if (size == cnt_repeat && coord_repeat == max_coord_repeat && dist < min_dist) {
                        max_pos = pin_2d, min_dist = dist;
                    }
                }
            }
            if (cnt_repeat <= 1) break;
            if (!p2D.count(max_pos)) {
                p2D.insert(max_pos);
                x_counts[max_pos / Y % X]++;
                y_counts[max_pos % Y]++;
            }
            for (int id : cnt_pin2d[max_pos]) {
                visited[id] = 1;
                for (int j = 1; j <= ap[id][0]; j++)
                    if (ap[id][j] % (X * Y) == max_pos) {
                        p.insert(ap[id][j]);
                        break;
                    }
            }
        }
        for (int i = 0; i < num; i++) {
            if (visited[i]) continue;
            int min_dist = 1000000000, selected = -1, max_overlap = 0;
            for (int j = 1; j <= ap[i][0]; j++) {
                const int pin = ap[i][j];
                const int dist = std::abs(pin / Y % X * 2 - center_2x) + std::abs(pin % Y * 2 - center_2y);
                const int pin_2d = pin % (X * Y);
                const int coord_repeat = x_counts[pin_2d / Y] + y_counts[pin_2d % Y];
                if (coord_repeat > max_overlap) max_overlap = coord_repeat, min_dist = dist, selected = pin;
                else if (coord_repeat == max_overlap && dist < min_dist) min_dist = dist, selected = pin;
            }
            p.insert(selected);
            if (!p2D.count(selected % (X * Y))) {
                p2D.insert(selected % (X * Y));
                x_counts[selected / Y % X]++;
                y_counts[selected % Y]++;
            }
        }
    }
    if (p2D.size() < p.size()) {
        std::map<int, int> pmax, pmin;
        for (int e : p) {
            const int pos = e % (X * Y);
            pmax[pos] = pmax.count(pos) ? std::max(pmax[pos], e) : e;
            pmin[pos] = pmin.count(pos) ? std::min(pmin[pos], e) : e;
        }
        for (int e : p2D) {
            net.pins.push_back(pmax[e]);
            if (pmin[e] != pmax[e]) {
                net.extra_routes.push_back(pmin[e]);
                net.extra_routes.push_back(pmax[e]);
            }
        }
    } else {
        net.pins.assign(p.begin(), p.end());
    }
}

void select_pins(Design& design, int threads) {
    const int X = design.grid.X, Y = design.grid.Y;
    parallel_for(int(design.nets.size()), threads, [&](int i) { init_net(design.nets[i], X, Y); });
}

// Rectilinear Steiner tree over distinct 2D positions, returned in the shape
// of my_flute (src/preprocess.hpp:21-70): nodes sorted by 2D index, adjacency
// added in (length, u, v) order.  FLUTE itself is not restated: the tree is a
// rectilinear minimum spanning tree improved by median Steiner points.
static std::vector<std::vector<int>> build_tree(const std::vector<int>& positions, int Y) {
    std::vector<int> xs, ys;
    for (int e : positions) xs.push_back(e / Y), ys.push_back(e % Y);
    int n = int(positions.size());
    std::vector<std::vector<int>> adj(n);
    if (n >= 2) {
        // Prim, ties to the lower index
        std::vector<int> best(n, INT32_MAX), link(n, -1), used(n, 0);
        best[0] = 0;
        for (int it = 0; it < n; ++it) {
            int u = -1;
            for (int v = 0; v < n; ++v)
                if (!used[v] && (u < 0 || best[v] < best[u])) u = v;
            used[u] = 1;
            if (link[u] >= 0) adj[u].push_back(link[u]), adj[link[u]].push_back(u);
            for (int v = 0; v < n; ++v) {
                const int d = std::abs(xs[u] - xs[v]) + std::abs(ys[u] - ys[v]);
                if (!used[v] && d < best[v]) best[v] = d, link[v] = u;
            }
        }
        // median Steiner points: replace (v,a),(v,b) by a star at the median
        auto dist = [&](int a, int b) { return std::abs(xs[a] - xs[b]) + std::abs(ys[a] - ys[b]); };
        for (int pass = 0; pass < 4; ++pass) {
            bool changed = false;
            for (int v = 0; v < int(adj.size()); ++v) {
                int best_gain = 0, ba = -1, bb = -1, sx = 0, sy = 0;
                for (size_t i = 0; i < adj[v].size(); ++i)
                    for (size_t j = i + 1; j < adj[v].size(); ++j) {
                        const int a = adj[v][i], b = adj[v][j];
                        const int mx = std::max(std::min(xs[v], xs[a]), std::min(std::max(xs[v], xs[a]), xs[b]));
                        const int my = std::max(std::min(ys[v], ys[a]), std::min(std::max(ys[v], ys[a]), ys[b]));
                        if ((mx == xs[v] && my == ys[v]) || (mx == xs[a] && my == ys[a]) || (mx == xs[b] && my == ys[b]))
                            continue;
                        const int before = dist(v, a) + dist(v, b);
                        const int after = std::abs(mx - xs[v]) + std::abs(my - ys[v]) + std::abs(mx - xs[a]) +
                                          std::abs(my - ys[a]) + std::abs(mx - xs[b]) + std::abs(my - ys[b]);
                        if (before - after > best_gain) best_gain = before - after, ba = a, bb = b, sx = mx, sy = my;
                    }
                if (ba < 0) continue;
                bool exists = false;
                for (int k = 0; k < int(xs.size()); ++k)
                    if (xs[k] == sx && ys[k] == sy) exists = true;
                if (exists) continue;
                const int s = int(xs.size());
                xs.push_back(sx), ys.push_back(sy);
                adj.emplace_back();
                auto erase = [&](int u, int w) { adj[u].erase(std::find(adj[u].begin(), adj[u].end(), w)); };
                erase(v, ba), erase(ba, v), erase(v, bb), erase(bb, v);
                for (int w : {v, ba, bb}) adj[s].push_back(w), adj[w].push_back(s);
                changed = true;
            }
            if (!changed) break;
        }
    }
    n = int(xs.size());
    std::vector<int> pos(n), rank(n);
    for (int i = 0; i < n; ++i) pos[i] = xs[i] * Y + ys[i];
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return pos[a] < pos[b]; });
    for (int i = 0; i < n; ++i) rank[order[i]] = i;
    std::vector<std::tuple<int, int, int>> edges;
    for (int u = 0; u < n; ++u)
        for (int w : adj[u])
            if (u < w) {
                int a = rank[u], b = rank[w];
                edges.emplace_back(std::abs(xs[u] - xs[w]) + std::abs(ys[u] - ys[w]), a, b);
            }
    std::sort(edges.begin(), edges.end());
    std::vector<std::vector<int>> graph(n);
    for (auto& e : edges) {
        graph[std::get<1>(e)].push_back(std::get<2>(e));
        graph[std::get<2>(e)].push_back(std::get<1>(e));
    }
    std::vector<int> nodes(n);
    for (int i = 0; i < n; ++i) nodes[i] = pos[order[i]];
    graph.push_back(std::move(nodes));
    return graph;
}

void tree_geometry(const std::vector<std::vector<int>>& adjacency, const std::vector<int>& pos, int Y,
                   std::vector<int>& points, std::vector<std::pair<int, int>>& hsegs,
                   std::vector<std::pair<int, int>>& vsegs) {
    // src/preprocess.hpp:100-125: every node, then per tree edge (j < i) its
    // segments (both L orientations of a bent edge) and corner / end points
    points.insert(points.end(), pos.begin(), pos.end());
    for (int i = 0; i < int(pos.size()); i++) {
        const int xi = pos[i] / Y, yi = pos[i] % Y;
        for (int j : adjacency[i])
            if (j < i) {
                const int xj = pos[j] / Y, yj = pos[j] % Y;
                const int minx = std::min(xi, xj), maxx = std::max(xi, xj), miny = std::min(yi, yj), maxy = std::max(yi, yj);
                if (xi != xj && yi != yj) {
                    hsegs.emplace_back(minx * Y + miny, maxx * Y + miny);
                    hsegs.emplace_back(minx * Y + maxy, maxx * Y + maxy);
                    vsegs.emplace_back(minx * Y + miny, minx * Y + maxy);
                    vsegs.emplace_back(maxx * Y + miny, maxx * Y + maxy);
                    points.push_back(xi * Y + yj);
                    points.push_back(xj * Y + yi);
                } else if (xi != xj) {
                    hsegs.emplace_back(minx * Y + miny, maxx * Y + miny);
                    points.push_back(minx * Y + miny);
                    points.push_back(maxx * Y + miny);
                } else if (yi != yj) {
                    vsegs.emplace_back(minx * Y + miny, minx * Y + maxy);
                    points.push_back(minx * Y + miny);
                    points.push_back(minx * Y + maxy);
                }
            }
    }
}

void construct_geometry(Subnet& net, int X, int Y, int L) {
    std::map<int, int> layer;
    std::vector<int> positions;
    for (int pin : net.pins) {
        layer[pin % (X * Y)] = pin / X / Y;
        positions.push_back(pin % (X * Y));
    }
    std::sort(positions.begin(), positions.end());
    positions.erase(std::unique(positions.begin(), positions.end()), positions.end());
    if (positions.size() != net.pins.size()) throw std::runtime_error("subnet pins share a 2D position");
    if (net.rsmt.empty()) net.rsmt = build_tree(positions, Y);
    else
        for (auto& e : net.rsmt.back()) e %= X * Y;
    net.hsegs.clear();
    net.vsegs.clear();
    net.points.clear();
    std::vector<std::vector<int>> adjacency(net.rsmt.begin(), net.rsmt.end() - 1);
    tree_geometry(adjacency, net.rsmt.back(), Y, net.points, net.hsegs, net.vsegs);
    for (auto& e : net.rsmt.back()) e += (layer.count(e) ? layer[e] : L) * X * Y;
}

// src/preprocess.hpp:1464-1574: split nets with more than 12 pins.
static void split_net(const DesignNet& db_net, int db_id, int X, int Y, int L, std::vector<Subnet>& out) {
    const int MAX_PIN_SIZE = 12;
    Subnet tmp;
    tmp.pins = db_net.pins;
    tmp.original = db_id;
    for (auto& p : tmp.pins)
        if (p >= X * Y) p -= X * Y;
    construct_geometry(tmp, X, Y, L);
    const auto& nodes = tmp.rsmt.back();
    const int n = int(nodes.size());
    std::vector<bool> not_steiner(n, false);
    for (int i = 0; i < n; i++)
        if (nodes[i] < L * X * Y) not_steiner[i] = true;
    std::vector<int> sz(n, 1), par(n);
    std::vector<std::tuple<int, int, int>> edges;
    std::function<int(int)> find = [&](int x) { return x == par[x] ? x : par[x] = find(par[x]); };
    for (int i = 0; i < n; i++) {
        par[i] = i;
        for (int j : tmp.rsmt[i])
            if (j < i) edges.emplace_back(j, i, int(not_steiner[i]) + int(not_steiner[j]));
    }
    std::sort(edges.begin(), edges.end(),
              [](std::tuple<int, int, int> l, std::tuple<int, int, int> r) { return std::get<2>(l) < std::get<2>(r); });
    bool big_steiner = false;
    for (auto e : edges) {
        int u = find(std::get<0>(e)), v = find(std::get<1>(e));
        if (u == v) continue;
        if (sz[u] + sz[v] > MAX_PIN_SIZE) {
            if (!(not_steiner[std::get<0>(e)] || not_steiner[std::get<1>(e)])) big_steiner = true;
            if (!big_steiner) continue;
        }
        if (sz[u] > sz[v]) std::swap(u, v);
        par[u] = v;
        sz[v] += sz[u];
    }
    std::vector<std::vector<int>> new_pins(n), net_edge(n);
    for (int i = 0; i < n; i++) new_pins[find(i)].push_back(nodes[i]);
    for (int i = 0; i < int(edges.size()); i++) {
        auto e = edges[i];
        int u = std::get<0>(e), v = std::get<1>(e);
        int pu = find(u), pv = find(v);
        if (pu == pv) {
            net_edge[pu].push_back(i);
            continue;
        }
        if (sz[u] > sz[v]) std::swap(u, v), std::swap(pu, pv);
        if (!not_steiner[v] && not_steiner[u]) std::swap(u, v), std::swap(pu, pv);
        sz[pu]++;
        if (sz[pv] == 1) par[pv] = pu;
        net_edge[pu].push_back(i);
        new_pins[pu].push_back(nodes[v]);
    }
    for (int i = 0; i < n; i++)
        if (find(i) == i) {
            Subnet s;
            s.rsmt.resize(new_pins[i].size() + 1);
            std::map<int, int> indx;
            for (int j = 0; j < int(new_pins[i].size()); j++) {
                if (new_pins[i][j] < L * X * Y) s.pins.push_back(new_pins[i][j]);
                indx[new_pins[i][j]] = j;
            }
            s.rsmt.back() = new_pins[i];
            for (int j : net_edge[i]) {
                auto& e = edges[j];
                int u = std::get<0>(e), v = std::get<1>(e);
                if (find(v) == i) std::swap(u, v);
                const int a = indx[nodes[u]], b = indx[nodes[v]];
                s.rsmt[a].push_back(b);
                s.rsmt[b].push_back(a);
            }
            s.original = db_id;
            out.push_back(std::move(s));
        }
}

void build_subnets(Design& design, int threads) {
    const int X = design.grid.X, Y = design.grid.Y, L = design.grid.routing_layers();
    const int MAX_PIN_SIZE = 12;
    std::vector<std::vector<Subnet>> tmp(design.nets.size());
    parallel_for(int(design.nets.size()), threads, [&](int id) {
        auto& net = design.nets[id];
        if (net.pins.size() <= 1) return;
        if (int(net.pins.size()) <= MAX_PIN_SIZE) {
            net.unfinished_subnet_count = 1;
            return;
        }
        split_net(net, id, X, Y, L, tmp[id]);
        net.unfinished_subnet_count = int(tmp[id].size());
    });
    design.subnets.clear();
    for (int id = 0; id < int(design.nets.size()); id++) {
        auto& net = design.nets[id];
        net.subnets.clear();
        const bool small = net.pins.size() > 1 && int(net.pins.size()) <= MAX_PIN_SIZE;
        if (small) {
            net.subnets.push_back(int(design.subnets.size()));
            Subnet s;
            s.pins = net.pins;
            for (auto& p : s.pins)
                if (p >= X * Y) p -= X * Y;
            s.original = id;
            design.subnets.push_back(std::move(s));
        }
        for (auto& s : tmp[id]) {
            net.subnets.push_back(int(design.subnets.size()));
            design.subnets.push_back(std::move(s));
        }
    }
    parallel_for(int(design.subnets.size()), threads, [&](int i) {
        auto& s = design.subnets[i];
        // calc_hpwl (src/preprocess.hpp:4-16) over the subnet pins
        s.minx = X, s.maxx = 0, s.miny = Y, s.maxy = 0;
        for (int p : s.pins) {
            const int x = p / Y % X, y = p % Y;
            s.minx = std::min(s.minx, x), s.maxx = std::max(s.maxx, x);
            s.miny = std::min(s.miny, y), s.maxy = std::max(s.maxy, y);
        }
        s.hpwl = s.maxx - s.minx + s.maxy - s.miny;
        construct_geometry(s, X, Y, L);
    });
    design.pin_acc_num.assign(design.subnets.size() + 1, 0);
    for (size_t i = 0; i < design.subnets.size(); ++i)
        design.pin_acc_num[i + 1] = design.pin_acc_num[i] + int(design.subnets[i].pins.size());
}

void preorder(const Subnet& net, std::vector<int>& nodes, std::vector<int>& parents) {
    const auto& graph = net.rsmt;
    const int n = int(graph.back().size());
    nodes.clear();
    parents.clear();
    nodes.reserve(n);
    parents.reserve(n);
    // iterative form of the recursive DFS: children in adjacency order
    struct Frame { int x, par, idx; size_t next; };
    std::vector<Frame> stack;
    stack.push_back({0, -1, 0, 0});
    nodes.push_back(graph.back()[0]);
    parents.push_back(-1);
    while (!stack.empty()) {
        Frame& f = stack.back();
        if (f.next >= graph[f.x].size()) {
            stack.pop_back();
            continue;
        }
        const int e = graph[f.x][f.next++];
        if (e == f.par) continue;
        const int idx = int(nodes.size());
        nodes.push_back(graph.back()[e]);
        parents.push_back(f.idx);
        stack.push_back({e, f.x, idx, 0});
    }
    if (int(nodes.size()) != n) throw std::runtime_error("subnet tree is not connected");
}

std::vector<int> hpwl_order(const Design& design) {
    std::vector<int> order(design.subnets.size());
    std::iota(order.begin(), order.end(), 0);
    const auto& s = design.subnets;
    std::sort(order.begin(), order.end(), [&](int l, int r) { return s[l].hpwl > s[r].hpwl; });
    return order;
}

}  // namespace grotor
