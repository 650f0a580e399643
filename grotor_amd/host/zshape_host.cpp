// Host side of the Z-shape (GGR) stage: priority order, bounding-box batch
// refinement, tree encoding for the GPU kernel and the post-processing of
// the downloaded routes.
//
// Reference: src/pipeline_ggr.hpp:8-162 and
// src/routing/zshape_topology.hpp:7-232.
#include "zshape_host.hpp"

#include <algorithm>
#include <atomic>
#include <exception>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace grotor {

std::vector<int> zshape_order(const Design& design) {
    // src/pipeline_ggr.hpp:31-33: hpwl descending, larger index first on ties
    std::vector<int> order(design.subnets.size());
    std::iota(order.begin(), order.end(), 0);
    const auto& s = design.subnets;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return s[a].hpwl == s[b].hpwl ? a > b : s[a].hpwl > s[b].hpwl; });
    return order;
}

std::vector<std::vector<int>> bbox_batches(const std::vector<int>& group, const Design& design, int max_batch) {
    // generate_bbox_batches (src/routing/zshape_topology.hpp:7-58): largest
    // boxes first (larger id first on ties), each pass a greedy set of nets
    // with pairwise disjoint boxes, at most max_batch nets per pass.  Nets
    // taken by a pass leave the pending list; a pass clears only the boxes
    // it marked.
    const auto& s = design.subnets;
    const int Y = design.grid.Y;
    auto area = [&](int id) { return (long long)(s[id].maxx - s[id].minx + 1) * (s[id].maxy - s[id].miny + 1); };
    std::vector<int> pending = group;
    std::sort(pending.begin(), pending.end(),
              [&](int a, int b) { return area(a) == area(b) ? a > b : area(a) > area(b); });
    std::vector<std::vector<int>> batches;
    // occupancy as one bit per cell, rows of x with 64 cells of y per word
    const int W = (Y + 63) / 64;
    std::vector<unsigned long long> occupied(size_t(design.grid.X) * W, 0ull);
    auto span = [](int w, int lo, int hi) {  // bits of word w covering cells lo..hi
        const int a = std::max(lo - 64 * w, 0), b = std::min(hi - 64 * w, 63);
        return (b == 63 ? ~0ull : (2ull << b) - 1) & ~((1ull << a) - 1);
    };
    auto clashes = [&](const Subnet& n) {
        for (int x = n.minx; x <= n.maxx; ++x) {
            const unsigned long long* row = &occupied[size_t(x) * W];
            for (int w = n.miny / 64; w <= n.maxy / 64; ++w)
                if (row[w] & span(w, n.miny, n.maxy)) return true;
        }
        return false;
    };
    auto mark = [&](const Subnet& n, bool set) {
        for (int x = n.minx; x <= n.maxx; ++x) {
            unsigned long long* row = &occupied[size_t(x) * W];
            for (int w = n.miny / 64; w <= n.maxy / 64; ++w) {
                if (set) row[w] |= span(w, n.miny, n.maxy);
                else row[w] &= ~span(w, n.miny, n.maxy);
            }
        }
    };
    while (!pending.empty()) {
        std::vector<int> batch;
        size_t keep = 0;
        size_t k = 0;
        for (; k < pending.size(); ++k) {
            const int id = pending[k];
            const Subnet& n = s[id];
            if (clashes(n)) {
                pending[keep++] = id;
                continue;
            }
            mark(n, true);
            batch.push_back(id);
            if (int(batch.size()) == max_batch) {
                ++k;
                break;
            }
        }
        for (; k < pending.size(); ++k) pending[keep++] = pending[k];
        pending.resize(keep);
        if (batch.empty()) break;
        for (int id : batch) mark(s[id], false);
        batches.push_back(std::move(batch));
    }
    return batches;
}

void encode_zshape_net(const Subnet& net, int raw_offset, int X, int Y, int L, int* out) {
    // encode_batch (src/routing/zshape_topology.hpp:59-121): rooted at the
    // first leaf, preorder with children in adjacency order, records of six
    // words per node, after the raw-word offset and the node count
    const auto& graph = net.rsmt;
    const int nodes = int(graph.size()) - 1;
    const int n = std::max(X, Y);
    out[0] = raw_offset;
    out[1] = nodes;
    int root = 0;
    for (int i = 0; i < nodes; ++i)
        if (graph[i].size() == 1) {
            root = i;
            break;
        }
    std::vector<int> order, parent(nodes, -1), stack{root};
    order.reserve(nodes);
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        order.push_back(v);
        for (auto it = graph[v].rbegin(); it != graph[v].rend(); ++it) {
            if (*it == parent[v]) continue;
            parent[*it] = v;
            stack.push_back(*it);
        }
    }
    if (int(order.size()) != nodes) throw std::runtime_error("Z-shape tree is not connected");
    std::vector<int> index(nodes, -1);
    for (int i = 0; i < nodes; ++i) index[order[i]] = i;
    std::fill(out + 2, out + 2 + size_t(nodes) * 6, -1);
    for (int i = 0; i < nodes; ++i) {
        const int v = order[i];
        const int pos = graph.back()[v], layer = pos / (X * Y), pos2 = pos % (X * Y);
        int* r = out + 2 + size_t(i) * 6;
        r[0] = (pos2 / Y) * n + pos2 % Y;
        if (layer < L) r[1] = r[2] = layer + 1;
        int slot = 3;
        for (int c : graph[v]) {
            if (parent[c] != v) continue;
            if (slot >= 6) throw std::runtime_error("Z-shape node with more than three children");
            r[slot++] = index[c] * 6;
        }
    }
}

ZshapePlan zshape_plan(const Design& design, const std::vector<int>& order, const int* color, int num_batches,
                       int max_batch) {
    // src/pipeline_ggr.hpp:49-90: RSMT batches in color order, each refined
    // into bounding-box batches, the whole list reversed
    std::vector<std::vector<int>> groups(num_batches);
    for (size_t k = 0; k < order.size(); ++k) {
        if (color[k] < 0 || color[k] >= num_batches) throw std::runtime_error("batch index out of range");
        groups[color[k]].push_back(order[k]);
    }
    // the RSMT batches refine independently: spread them over threads
    std::vector<std::vector<std::vector<int>>> refined(groups.size());
    {
        const int workers = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (int w = 0; w < workers; ++w)
            pool.emplace_back([&] {
                for (size_t i = next++; i < groups.size(); i = next++) refined[i] = bbox_batches(groups[i], design, max_batch);
            });
        for (auto& th : pool) th.join();
    }
    std::vector<std::vector<int>> batches;
    for (auto& r : refined)
        for (auto& b : r) batches.push_back(std::move(b));
    std::reverse(batches.begin(), batches.end());
    const Grid& g = design.grid;
    const int L = g.routing_layers();
    // layout: per batch its nets' record offsets, then their records; every
    // net's place is known up front, so the nets are encoded in parallel
    ZshapePlan plan;
    plan.batch_off.push_back(0);
    plan.rec_off.push_back(0);
    std::vector<long long> where;  // per net: its record's position in plan.records
    long long words = 0;
    for (const auto& b : batches) {
        const long long start = words;
        words += (long long)b.size();
        int raw = 0;
        for (size_t i = 0; i < b.size(); ++i) {
            const Subnet& s = design.subnets[b[i]];
            where.push_back(words);
            words += 2 + 6 * ((long long)s.rsmt.size() - 1);
            const int cap = int(s.pins.size()) * kZshapeRoutePerPin;
            plan.raw_cap.push_back(cap);
            raw += cap;
            plan.net_ids.push_back(b[i]);
        }
        (void)start;
        plan.batch_off.push_back(int(plan.net_ids.size()));
        plan.rec_off.push_back(words);
    }
    plan.records.assign(size_t(words), 0);
    for (size_t bi = 0, k = 0; bi < batches.size(); ++bi) {
        const long long start = plan.rec_off[bi];
        for (size_t i = 0; i < batches[bi].size(); ++i, ++k) plan.records[size_t(start + i)] = int(where[k] - start);
    }
    {
        const size_t nets = plan.net_ids.size();
        const int workers = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::atomic<size_t> next{0};
        std::exception_ptr failure;
        std::mutex failure_lock;
        std::vector<int> raw_off(nets);
        for (size_t bi = 0, k = 0; bi < batches.size(); ++bi) {
            int raw = 0;
            for (size_t i = 0; i < batches[bi].size(); ++i, ++k) {
                raw_off[k] = raw;
                raw += plan.raw_cap[k];
            }
        }
        std::vector<std::thread> pool;
        for (int w = 0; w < workers; ++w)
            pool.emplace_back([&] {
                try {
                    for (size_t k = next.fetch_add(256); k < nets; k = next.fetch_add(256))
                        for (size_t e = std::min(nets, k + 256); k < e; ++k)
                            encode_zshape_net(design.subnets[plan.net_ids[k]], raw_off[k], g.X, g.Y, L,
                                              plan.records.data() + where[k]);
                } catch (...) {
                    std::lock_guard<std::mutex> guard(failure_lock);
                    if (!failure) failure = std::current_exception();
                }
            });
        for (auto& th : pool) th.join();
        if (failure) std::rethrow_exception(failure);
    }
    return plan;
}

namespace {

template <class F>
void visit_edges(int p0, int p1, int X, int Y, F&& fn) {
    const int xy = X * Y, l = p0 / xy;
    if (l != p1 / xy) return;
    const int x0 = p0 / Y % X, y0 = p0 % Y, x1 = p1 / Y % X, y1 = p1 % Y;
    if (x0 == x1)
        for (int y = std::min(y0, y1); y < std::max(y0, y1); ++y) fn(l, x0, y);
    if (y0 == y1)
        for (int x = std::min(x0, x1); x < std::max(x0, x1); ++x) fn(l, x, y0);
}

}  // namespace

int zshape_promote(const Design& design, int per_pin, int* routes) {
    // promote_congested_segments (src/routing/zshape_topology.hpp:122-177)
    const Grid& g = design.grid;
    const int X = g.X, Y = g.Y, xy = X * Y, L = g.routing_layers();
    const auto& acc = design.pin_acc_num;
    const int nets = int(design.subnets.size());
    std::vector<float> usage(size_t(L) * xy, 0.0f);
    for (int net = 0; net < nets; ++net) {
        const int* r = routes + size_t(acc[net]) * per_pin;
        for (int i = 1; i + 1 < r[0]; i += 2)
            visit_edges(r[i], r[i + 1], X, Y, [&](int l, int x, int y) { usage[size_t(l) * xy + x * Y + y] += 1.0f; });
    }
    int promoted = 0;
    for (int net = 0; net < nets; ++net) {
        int* r = routes + size_t(acc[net]) * per_pin;
        const int end = r[0];
        const int limit = (acc[net + 1] - acc[net]) * per_pin;
        for (int i = 1; i + 1 < end; i += 2) {
            const int p0 = r[i], p1 = r[i + 1], old_l = p0 / xy;
            if (old_l != p1 / xy) continue;
            if (p0 % xy == p1 % xy) continue;
            auto worst = [&](int layer, float add) {
                double w = -1e30;
                visit_edges(p0 % xy + layer * xy, p1 % xy + layer * xy, X, Y, [&](int, int x, int y) {
                    const double cap = g.capacity[size_t(layer + 1) * xy + size_t(x) * Y + y];
                    w = std::max(w, double(usage[size_t(layer) * xy + x * Y + y] + add) - cap);
                });
                return w;
            };
            const double old_score = worst(old_l, 0.0f);
            if (old_score <= 0.0) continue;
            int best_l = old_l;
            double best = old_score;
            for (int l = old_l + 2; l < L; l += 2) {
                const double sc = worst(l, 1.0f);
                if (sc < best) best = sc, best_l = l;
            }
            if (best_l == old_l || best >= old_score || r[0] + 4 >= limit) continue;
            visit_edges(p0, p1, X, Y, [&](int l, int x, int y) { usage[size_t(l) * xy + x * Y + y] -= 1.0f; });
            const int q0 = p0 % xy + best_l * xy, q1 = p1 % xy + best_l * xy;
            visit_edges(q0, q1, X, Y, [&](int l, int x, int y) { usage[size_t(l) * xy + x * Y + y] += 1.0f; });
            r[i] = q0;
            r[i + 1] = q1;
            r[r[0]++] = p0;
            r[r[0]++] = q0;
            r[r[0]++] = p1;
            r[r[0]++] = q1;
            ++promoted;
        }
    }
    return promoted;
}

void zshape_canonicalize(const Design& design, int per_pin, int* routes) {
    // canonicalize_routes (src/routing/zshape_topology.hpp:179-232): merged
    // wire intervals per (layer, direction, track) in key order, then maximal
    // via stacks in (x, y, layer) order
    const Grid& g = design.grid;
    const int X = g.X, Y = g.Y, xy = X * Y;
    const auto& acc = design.pin_acc_num;
    const int nets = int(design.subnets.size());
    for (int net = 0; net < nets; ++net) {
        int* d = routes + size_t(acc[net]) * per_pin;
        if (d[0] <= 1) continue;
        std::map<std::tuple<int, int, int>, std::vector<std::pair<int, int>>> spans;
        std::set<std::tuple<int, int, int>> vias;
        for (int i = 1; i + 1 < d[0]; i += 2) {
            const int a = d[i], b = d[i + 1], la = a / xy, lb = b / xy;
            const int ax = a / Y % X, ay = a % Y, bx = b / Y % X, by = b % Y;
            if (la != lb) {
                for (int l = std::min(la, lb); l < std::max(la, lb); ++l) vias.emplace(ax, ay, l);
            } else if (ax != bx) {
                spans[{la, 0, ay}].emplace_back(std::min(ax, bx), std::max(ax, bx));
            } else if (ay != by) {
                spans[{la, 1, ax}].emplace_back(std::min(ay, by), std::max(ay, by));
            }
        }
        int out = 1;
        const int limit = (acc[net + 1] - acc[net]) * per_pin;
        auto emit = [&](int a, int b) {
            if (out + 1 >= limit) throw std::runtime_error("Z-shape route storage exceeded");
            d[out++] = a;
            d[out++] = b;
        };
        for (auto& [key, v] : spans) {
            std::sort(v.begin(), v.end());
            const auto [layer, dir, fixed] = key;
            int lo = v[0].first, hi = v[0].second;
            auto flush = [&] {
                if (dir == 0) emit(layer * xy + lo * Y + fixed, layer * xy + hi * Y + fixed);
                else emit(layer * xy + fixed * Y + lo, layer * xy + fixed * Y + hi);
            };
            for (size_t j = 1; j < v.size(); ++j) {
                if (v[j].first <= hi) hi = std::max(hi, v[j].second);
                else {
                    flush();
                    lo = v[j].first, hi = v[j].second;
                }
            }
            flush();
        }
        for (auto it = vias.begin(); it != vias.end();) {
            const auto [x, y, l0] = *it;
            int l1 = l0 + 1;
            auto next = std::next(it);
            while (next != vias.end()) {
                const auto [nx, ny, nl] = *next;
                if (nx != x || ny != y || nl != l1) break;
                ++l1;
                ++next;
            }
            emit(l0 * xy + x * Y + y, l1 * xy + x * Y + y);
            it = next;
        }
        d[0] = out;
    }
}

}  // namespace grotor
