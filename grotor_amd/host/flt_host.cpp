// FLT DAG builder: restates net::generate_detours_reconstruct
// (src/preprocess.hpp:769-1436) for one subnet tree.
//
// The DAG offers every tree edge in several shapes: bent edges as both
// L-corners plus a few Z-shapes, long edges split into straight pieces, and,
// where a trunk of the tree crosses a congested region, "mirror" copies of
// the trunk shifted onto parallel tracks.  Nodes are renumbered in the
// reference's topological order and every node's depth is its longest path
// from the root, so that a node's parents are all shallower than the node.
#include "flt_host.hpp"

#include <algorithm>
#include <cstdlib>
#include <functional>
#include <queue>
#include <stdexcept>
#include <utility>

namespace grotor {

int flt_select_root(const std::vector<std::vector<int>>& rsmt) {
    const int n = int(rsmt.size()) - 1;
    std::vector<char> seen(std::max(n, 0), 0);
    std::queue<std::pair<int, int>> q;  // (node, node it was reached from)
    for (int i = 0; i < n; ++i)
        if (rsmt[i].size() == 1) q.push({i, -1});
    int last = -1;
    while (!q.empty()) {
        const auto [v, from] = q.front();
        q.pop();
        last = v;  // the reference records the node before the visited test
        if (seen[v]) continue;
        seen[v] = 1;
        for (int w : rsmt[v])
            if (w != from) q.push({w, v});
    }
    return last;
}

namespace {

struct DagBuilder {
    const FltParams& p;
    const std::vector<std::vector<int>>& g;  // tree adjacency, g.back() = positions
    const int X, Y, L, tracks;
    const float ratio;
    const int zshapes, zselect;
    int root = 0;

    // created nodes (creation order): 3D position, parent links
    // (parent * 10 + slot, in link order), number of incoming child links
    std::vector<int> pos;
    std::vector<std::vector<int>> ups;
    std::vector<int> incoming;
    std::vector<int> points;

    // congestion regions per direction (0: along x, 1: along y)
    std::vector<int> region[2];
    std::vector<std::pair<int, int>> range[2];
    std::vector<std::vector<int>> stems[2];

    DagBuilder(const FltParams& params, const Subnet& net)
        : p(params),
          g(net.rsmt),
          X(params.X),
          Y(params.Y),
          L(params.L),
          tracks(params.mode == 0 || params.mode == 1 ? 6 : 9),
          ratio(params.mode == 0 || params.mode == 1 ? 0.1f : 0.15f),
          zshapes(params.mode == 0 || params.mode == 1 ? 7 : 10),
          zselect(params.mode == 0 || params.mode == 1 ? 2 : 10) {}

    int px_of(int v) const { return v / Y % X; }
    int py_of(int v) const { return v % Y; }
    float xs(int x, int y) const { return p.xsum ? p.xsum[size_t(x) * Y + y] : 0.f; }
    float ys(int x, int y) const { return p.ysum ? p.ysum[size_t(x) * Y + y] : 0.f; }

    int create(int l, int x, int y) {
        if (x < 0 || y < 0) throw std::runtime_error("FLT node outside the grid");
        pos.push_back((l * X + x) * Y + y);
        ups.emplace_back();
        incoming.push_back(0);
        return int(pos.size()) - 1;
    }
    // connect_node_pre (:1002-1018)
    void link(int par, int cur, int slot) {
        const int cx = px_of(pos[cur]), cy = py_of(pos[cur]);
        const int qx = px_of(pos[par]), qy = py_of(pos[par]);
        if (cx != qx && cy != qy) throw std::runtime_error("FLT edge is not straight");
        points.push_back(cx * Y + cy);
        points.push_back(qx * Y + qy);
        ups[cur].push_back(par * 10 + slot);
        ++incoming[par];
    }
    // connect_node (:1020-1071): edges of 50 or more cells get one middle
    // node, of 150 or more two nodes at the thirds
    void connect(int par, int cur, int slot) {
        const int qx = px_of(pos[par]), qy = py_of(pos[par]);
        const int cx = px_of(pos[cur]), cy = py_of(pos[cur]);
        const int d = std::abs(qx - cx) + std::abs(qy - cy);
        if (d < 50) {
            link(par, cur, slot);
        } else if (d < 150) {
            const int m = create(L, (qx + cx) / 2, (qy + cy) / 2);
            link(par, m, slot);
            link(m, cur, 0);
        } else if (qy == cy) {
            const int m1 = create(L, qx * 2 / 3 + cx / 3, qy);
            const int m2 = create(L, qx * 1 / 3 + cx * 2 / 3, qy);
            link(par, m1, slot);
            link(m1, m2, 0);
            link(m2, cur, 0);
        } else if (qx == cx) {
            const int m1 = create(L, qx, qy * 2 / 3 + cy / 3);
            const int m2 = create(L, qx, qy * 1 / 3 + cy * 2 / 3);
            link(par, m1, slot);
            link(m1, m2, 0);
            link(m2, cur, 0);
        } else {
            throw std::runtime_error("FLT edge is not straight");
        }
    }

    int find(int v, int dir) {  // getRegionID with path compression
        if (v == -1 || region[dir][v] == -1) return -1;
        if (region[dir][v] != v) {
            const int r = find(region[dir][v], dir);
            region[dir][v] = r;
            return r;
        }
        return v;
    }

    bool congested_y(int x, int y0, int y1) const {
        return ys(x, std::max(y0, y1)) - ys(x, std::min(y0, y1)) > 0;
    }
    bool congested_x(int y, int x0, int x1) const {
        return xs(std::max(x0, x1), y) - xs(std::min(x0, x1), y) > 0;
    }

    // markCongestion (:885-986)
    void mark(int x, int par) {
        const auto& P = g.back();
        const int cx = px_of(P[x]), cy = py_of(P[x]);
        for (int e : g[x]) {
            if (e == par) continue;
            const int ex = px_of(P[e]), ey = py_of(P[e]);
            int dir = -1, kind = -1;
            if (ex == cx) {
                if (congested_y(cx, ey, cy)) dir = 1, kind = 1;
            } else if (ey == cy) {
                if (congested_x(cy, ex, cx)) dir = 0, kind = 1;
            } else {
                if (congested_y(cx, ey, cy) || congested_y(ex, ey, cy)) dir = 1, kind = 2;
                if (kind == -1 && (congested_x(cy, ex, cx) || congested_x(ey, ex, cx))) dir = 0, kind = 2;
            }
            if (kind == 1) {
                int target;
                if (x != root) {
                    if (region[dir][x] == -1) region[dir][x] = x;
                    target = find(x, dir);
                } else {
                    if (region[dir][e] == -1) region[dir][e] = e;
                    target = find(e, dir);
                }
                region[dir][e] = target;
                if (x != root)
                    for (int s : stems[dir][e]) stems[dir][target].push_back(s);
                range[dir][target].first = std::min(range[dir][target].first, range[dir][e].first);
                range[dir][target].second = std::max(range[dir][target].second, range[dir][e].second);
            } else if (kind == 2) {
                if (x != root && region[dir][x] == -1) region[dir][x] = x;
                if (region[dir][e] == -1) region[dir][e] = e;
            }
            mark(e, x);
        }
        for (int dir = 0; dir < 2; ++dir) {
            const int xr = find(x, dir);
            for (int e : g[x]) {
                if (e == par) continue;
                const int ex = px_of(P[e]), ey = py_of(P[e]);
                const int er = find(e, dir);
                if (xr == er) continue;
                if (xr >= 0) {
                    stems[dir][xr].push_back(dir ? ex : ey);
                    range[dir][xr].first = std::min(range[dir][xr].first, dir ? ey : ex);
                    range[dir][xr].second = std::max(range[dir][xr].second, dir ? ey : ex);
                }
                if (er >= 0) {
                    stems[dir][er].push_back(dir ? cx : cy);
                    range[dir][er].first = std::min(range[dir][er].first, dir ? cy : cx);
                    range[dir][er].second = std::max(range[dir][er].second, dir ? cy : cx);
                }
            }
        }
    }

    int displacement(int q, int dir, int r) const {  // calc_displace
        int s = 0;
        for (int v : stems[dir][r]) s += std::abs(v - q);
        return s;
    }

    // get_mirror_places (:1073-1117): parallel tracks whose summed stem
    // displacement stays within ratio * trunk length, thinned to about
    // `tracks` positions
    std::vector<int> mirror_places(int v, int dir) {
        const auto& P = g.back();
        const int r = region[dir][v];
        const int cx = px_of(P[v]), cy = py_of(P[v]);
        const int trunk = range[dir][r].second - range[dir][r].first;
        const int max_disp = int(ratio * float(trunk));
        const int orig = dir ? cx : cy;
        const int orig_disp = displacement(orig, dir, r);
        int lo = orig, hi = orig;
        const int bound = dir ? X : Y;
        while (lo - 1 >= 0 && displacement(lo - 1, dir, r) - orig_disp <= max_disp) --lo;
        // the reference tests hi - 1 here, not hi + 1
        while (hi + 1 < bound && displacement(hi - 1, dir, r) - orig_disp <= max_disp) ++hi;
        int step = 1;
        while ((orig - lo) / (step + 1) + (hi - orig) / (step + 1) >= tracks) ++step;
        lo = orig - (orig - lo) / step * step;
        hi = orig + (hi - orig) / step * step;
        std::vector<int> out;
        for (double q = lo; q <= hi; q += step) {
            if (int(q - orig) == 0) continue;
            out.push_back(int(q));
        }
        return out;
    }

    // Z-shape candidates of one bent edge and leg direction (:1216-1275)
    void zshapes_of(int par_node, int node, int slot, int px, int py, int cx, int cy) {
        for (int leg = 0; leg <= 1; ++leg) {
            const int len = leg ? std::max(py, cy) - std::min(py, cy) : std::max(px, cx) - std::min(px, cx);
            const int m = std::min(zshapes, len);
            std::pair<float, int> road[10] = {};
            auto mids = [&](int k, int& x1, int& y1, int& x2, int& y2) {
                x1 = leg ? px : (px * k + cx * (m - k)) / m;
                y1 = leg ? (py * k + cy * (m - k)) / m : py;
                x2 = leg ? cx : (px * k + cx * (m - k)) / m;
                y2 = leg ? (py * k + cy * (m - k)) / m : cy;
                if (x1 < 0 || y1 < 0 || x1 >= X || y1 >= Y) return false;
                if ((x1 == px && y1 == py) || (x1 == cx && y1 == cy)) return false;
                if (x2 < 0 || y2 < 0 || x2 >= X || y2 >= Y) return false;
                if ((x2 == cx && y2 == cy) || (x2 == px && y2 == py)) return false;
                return true;
            };
            for (int k = 1; k < m; ++k) {
                int x1, y1, x2, y2;
                if (!mids(k, x1, y1, x2, y2)) continue;
                // overflow of the three legs, read from the y prefix for all
                // three and accumulated into an int as the reference does
                int over = 0;
                over = int(float(over) + (ys(std::max(x1, x2), std::max(y1, y2)) - ys(std::min(x1, x2), std::min(y1, y2))));
                over = int(float(over) + (ys(std::max(x1, px), std::max(y1, py)) - ys(std::min(x1, px), std::min(y1, py))));
                over = int(float(over) + (ys(std::max(cx, x2), std::max(cy, y2)) - ys(std::min(cx, x2), std::min(cy, y2))));
                road[k - 1] = {float(over), k};
            }
            if (m - 1 > 0) std::sort(road, road + m - 1);
            for (int i = 0; i < zselect; ++i) {
                const int k = road[i].second;
                int x1, y1, x2, y2;
                if (m == 0 || !mids(k, x1, y1, x2, y2)) continue;
                const int a = create(L, x1, y1);
                connect(par_node, a, slot);
                points.push_back((x1 + x2) / 2 * Y + (y1 + y2) / 2);
                const int b = create(L, x2, y2);
                connect(a, b, 0);
                connect(b, node, 0);
            }
        }
    }

    // dfs_detours (:1119-1340)
    void dfs(int x, int par, int par_node, int slot, const std::vector<int>* mirrors_in,
             const std::vector<int>* places_in) {
        const auto& P = g.back();
        const int cl = P[x] / Y / X, cx = px_of(P[x]), cy = py_of(P[x]);
        const int node = create(cl, cx, cy);
        std::vector<int> mirrors[2], places[2];
        if (places_in) {
            for (int dir = 0; dir < 2; ++dir) {
                const int r = find(region[dir][x], dir);
                if (r < 0) continue;
                places[dir] = x == r ? mirror_places(x, dir) : places_in[dir];
                if (r != x && places[dir].size() != mirrors_in[dir].size())
                    throw std::runtime_error("FLT mirror tracks differ within a region");
                for (size_t i = 0; i < places[dir].size(); ++i) {
                    const int nx = dir ? places[dir][i] : cx, ny = dir ? cy : places[dir][i];
                    int mirror;
                    if (r != x) {
                        mirror = create(L, nx, ny);
                        connect(mirrors_in[dir][i], mirror, slot);
                    } else if (x != root) {
                        const int qx = px_of(pos[par_node]), qy = py_of(pos[par_node]);
                        if (nx != qx && ny != qy) {
                            int corner[2];
                            for (int k = 0; k <= 1; ++k) {
                                corner[k] = create(L, k ? nx : qx, k ? qy : ny);
                                connect(par_node, corner[k], slot);
                            }
                            mirror = create(L, nx, ny);
                            for (int k = 0; k <= 1; ++k) connect(corner[k], mirror, 0);
                        } else {
                            mirror = create(L, nx, ny);
                            connect(par_node, mirror, slot);
                        }
                    } else {
                        mirror = create(L, nx, ny);
                    }
                    mirrors[dir].push_back(mirror);
                }
            }
        }
        if (par_node != -1) {
            const int qx = px_of(pos[par_node]), qy = py_of(pos[par_node]);
            if (qx != cx && qy != cy) {
                for (int k = 0; k <= 1; ++k) {
                    const int c = create(L, k ? cx : qx, k ? qy : cy);
                    connect(par_node, c, slot);
                    connect(c, node, 0);
                }
                zshapes_of(par_node, node, slot, qx, qy, cx, cy);
            }
            for (int dir = 0; dir < 2; ++dir) {
                const int r = find(region[dir][x], dir);
                const int pr = find(region[dir][par], dir);
                if (pr < 0 || r == pr) continue;
                for (int m : mirrors_in[dir]) {
                    const int mx = px_of(pos[m]), my = py_of(pos[m]);
                    if (mx != cx && my != cy) {
                        for (int k = 0; k <= 1; ++k) {
                            const int c = create(L, k ? cx : mx, k ? my : cy);
                            connect(m, c, slot);
                            connect(c, node, 0);
                        }
                    } else {
                        connect(m, node, slot);
                    }
                }
            }
            if (qx == cx || qy == cy) connect(par_node, node, slot);
        }
        for (int dir = 0; dir < 2; ++dir) {
            if (find(x, dir) < 0) continue;
            if (x != root && cl < L)  // pins get a copy under every mirror
                for (int m : mirrors[dir]) {
                    const int dup = create(cl, cx, cy);
                    connect(m, dup, int(g[x].size()) - 1);
                }
        }
        int k = 0;
        for (int e : g[x])
            if (e != par) dfs(e, x, node, k++, mirrors, places);
    }
};

}  // namespace

void flt_build_dag(const Subnet& net, const FltParams& p, FltDag& out) {
    const auto& g = net.rsmt;
    const int n = int(g.size()) - 1;
    if (n < 1) throw std::runtime_error("FLT net without tree nodes");
    DagBuilder b(p, net);
    b.root = n == 1 ? 0 : flt_select_root(g);
    if (b.root < 0) throw std::runtime_error("FLT tree without leaves");
    for (int dir = 0; dir < 2; ++dir) {
        b.region[dir].assign(g.size(), -1);
        b.range[dir].resize(g.size());
        b.stems[dir].resize(g.size());
    }
    const auto& P = g.back();
    for (int x = 0; x < n; ++x) {
        const int l = P[x] / p.Y / p.X, cx = P[x] / p.Y % p.X, cy = P[x] % p.Y;
        for (int dir = 0; dir < 2; ++dir) {
            if (l < p.L) b.stems[dir][x].push_back(dir ? cx : cy);
            const int t = dir ? cy : cx;
            b.range[dir][x] = {t, t};
        }
    }
    b.mark(b.root, -1);
    b.dfs(b.root, -1, -1, -1, nullptr, nullptr);

    // determinSequence (:1374-1385): a node is pushed once all of its child
    // links have been seen, walking up from every node without children
    const int cnt = int(b.pos.size());
    std::vector<int> seen(cnt, 0), stack;
    stack.reserve(cnt);
    std::function<void(int)> push = [&](int v) {
        stack.push_back(v);
        for (int up : b.ups[v]) {
            const int q = up / 10;
            if (++seen[q] == b.incoming[q]) push(q);
        }
    };
    for (int i = 0; i < cnt; ++i)
        if (b.incoming[i] == 0) push(i);
    if (int(stack.size()) != cnt) throw std::runtime_error("FLT DAG is not acyclic and connected");
    // pop order is the node numbering (:1386-1410)
    std::vector<int> id(cnt, 0), depth(cnt, 0), kids(cnt, 0);
    out.nodes.assign(4 * size_t(cnt), 0);
    out.edges.clear();
    int edges = 0, maxdepth = 0;
    for (int k = 0; k < cnt; ++k) {
        const int v = stack[cnt - 1 - k];
        for (int up : b.ups[v]) {
            const int q = id[up / 10], s = up % 10;
            kids[q] = std::max(kids[q], s + 1);
            out.edges.push_back(q);
            out.edges.push_back(k * 10 + s);
            depth[k] = std::max(depth[k], depth[q] + 1);
        }
        edges += int(b.ups[v].size());
        id[v] = k;
        out.nodes[4 * size_t(k)] = b.pos[v];
        out.nodes[4 * size_t(k) + 3] = edges;
    }
    for (int k = 0; k < cnt; ++k) {
        if (depth[k] >= 5000) throw std::runtime_error("FLT depth exceeds supported range");
        out.nodes[4 * size_t(k) + 1] = depth[k];
        out.nodes[4 * size_t(k) + 2] = kids[k];
        maxdepth = std::max(maxdepth, depth[k] + 1);
    }
    out.depths = maxdepth;
    out.points = std::move(b.points);
    std::sort(out.points.begin(), out.points.end());
    out.points.erase(std::unique(out.points.begin(), out.points.end()), out.points.end());
}

}  // namespace grotor
