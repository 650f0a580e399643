"""FLT oracle (oracle/flt.c) and the FLT DAG builder, on the CPU.

Pins: the builder against the reference's own generate_detours_reconstruct
(golden DAGs from tests/golden/*.flt.jsonl.gz, congestion regions and
mirrors included); the DAG DP against exhaustive enumeration of the routing
trees a DAG encodes, through the cost of the traced route words; the packed
candidate keys against their ordering contract; the congestion view against
exact prefix sums; the stage against connectivity of every rerouted net."""
import ctypes
import gzip
import itertools
import json

import numpy as np
import pytest

from tests import _golden as G
from tests._instances import topology

P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
Pf = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def build_dag(ga, X, Y, L, positions, adj, xsum=None, ysum=None, mode=0):
    """gr_flt_dag: (node records [n, 4], edge records [e, 2], points)."""
    pos = np.asarray(positions, np.int32)
    adj_off = np.concatenate([[0], np.cumsum([len(a) for a in adj])]).astype(np.int32)
    flat = np.array([v for a in adj for v in a] or [0], np.int32)
    cap = 20000
    nodes, edges, pts = np.zeros(4 * cap, np.int32), np.zeros(2 * cap, np.int32), np.zeros(cap, np.int32)
    nn, ne, npt = ctypes.c_int(cap), ctypes.c_int(cap), ctypes.c_int(cap)
    xs = None if xsum is None else np.ascontiguousarray(xsum, np.float32)
    ys = None if ysum is None else np.ascontiguousarray(ysum, np.float32)
    rc = ga.lib().gr_flt_dag(X, Y, L, mode, len(pos), P(pos), P(adj_off), P(flat),
                             None if xs is None else Pf(xs), None if ys is None else Pf(ys),
                             P(nodes), ctypes.byref(nn), P(edges), ctypes.byref(ne), P(pts), ctypes.byref(npt))
    assert rc == 0, ga.lib().gr_last_error()
    return (nodes[:4 * nn.value].reshape(-1, 4).copy(), edges[:2 * ne.value].reshape(-1, 2).copy(),
            pts[:npt.value].copy())


@pytest.mark.parametrize("name", G.NAMES)
def test_flt_builder_equals_reference(ga, name):
    """Every DAG of the reference's builder on the reference's own trees,
    with a congestion view that creates congested regions and mirrors:
    node positions, depths, child slots, parent lists, slots and batching
    points identical."""
    d = G.dump(name)
    X, Y, L = d["grid"]["X"], d["grid"]["Y"], d["grid"]["L"]
    with gzip.open(G.path(name, "flt.jsonl"), "rt") as f:
        lines = f.read().strip().split("\n")
    view = json.loads(lines[0])["view"]
    xs, ys = np.array(view["xsum"], np.float32), np.array(view["ysum"], np.float32)
    mirrored = 0
    for line in lines[1:]:
        o = json.loads(line)
        s = d["subs"][o["dag"]]
        nodes, edges, pts = build_dag(ga, X, Y, L, s["nodes"], s["adj"], xs, ys)
        assert list(nodes[:, 0]) == o["nodes"], o["dag"]
        assert list(nodes[:, 1]) == o["depth"], o["dag"]
        assert list(nodes[:, 2]) == o["kids"], o["dag"]
        assert list(np.diff(np.concatenate([[0], nodes[:, 3]]))) == o["npar"], o["dag"]
        assert list(edges[:, 0]) == o["par"], o["dag"]
        assert list(edges[:, 1] % 10) == o["slot"], o["dag"]
        owners = np.repeat(np.arange(len(nodes)), o["npar"])
        assert np.array_equal(edges[:, 1] // 10, owners)
        assert list(pts) == o["points"], o["dag"]
        mirrored += len(o["nodes"]) > 3 * len(s["nodes"])
    assert len(lines) > 1000 and mirrored > 50


def test_flt_pack_orders_costs_then_codes(ora):
    """Packed keys (src/routing/flt_cost.cuh:25-40): for non-negative costs
    the u64 order is the float order, ties broken by the smaller code; the
    empty sentinel 7208259049189261824 is float 1.01e22 with code 200000000, above
    every finite candidate cost below 1e22."""
    pack = ora.lib().ora_flt_pack
    pack.restype = ctypes.c_ulonglong
    pack.argtypes = [ctypes.c_float, ctypes.c_int]
    rng = np.random.default_rng(5)
    vals = np.concatenate([[0.0, 1e-30, 1.0, 1e21, 9.99e21], rng.uniform(0, 1e6, 200)]).astype(np.float32)
    for a, b in itertools.combinations(vals, 2):
        ka, kb = pack(float(a), 7), pack(float(b), 3)
        assert (ka < kb) == (a < b) or a == b
    assert pack(5.0, 3) < pack(5.0, 4)
    assert pack(np.float32(1.01e22), 200000000) == 7208259049189261824
    assert pack(np.float32(1e22), 2**31 - 1) < 7208259049189261824


def test_congestion_view_prefix_sums(ora):
    """Worst overflow per cell over the layers (>= 0), congested = worst > 0,
    and exclusive prefix sums along x and y; integer-valued overflow makes
    every summation order exact."""
    rng = np.random.default_rng(3)
    L, X, Y = 3, 13, 9
    grid = dict(L=L, X=X, Y=Y, DIR=0, unit_length_wire_cost=1.0, unit_via_cost=1.0,
                unit_length_short_costs=np.ones(L), x_edge_len=np.ones(X - 1, np.int32),
                y_edge_len=np.ones(Y - 1, np.int32), capacity=rng.integers(0, 6, L * X * Y).astype(np.float32))
    g = ora.Grid(grid)
    demand = rng.integers(0, 8, L * X * Y).astype(np.float32)
    cong, xs, ys = g.congestion_view(demand)
    over = (demand - grid["capacity"]).reshape(L, X, Y).max(axis=0).clip(min=0)
    assert np.array_equal(cong.reshape(X, Y), over > 0)
    xs_ref = np.concatenate([np.zeros((1, Y)), np.cumsum(over, axis=0)[:-1]], axis=0)
    ys_ref = np.concatenate([np.zeros((X, 1)), np.cumsum(over, axis=1)[:, :-1]], axis=1)
    assert np.array_equal(xs.reshape(X, Y), xs_ref.astype(np.float32))
    assert np.array_equal(ys.reshape(X, Y), ys_ref.astype(np.float32))


# ---- DAG DP against exhaustive enumeration -------------------------------

def _tiny_costs(rng, L, X, Y, DIR):
    """Integer wire costs as exclusive prefixes along each layer's tracks,
    integer via costs >= 1 (IDX order): every DP sum is exact in float."""
    w = rng.integers(0, 10, (L, X, Y)).astype(np.float64)
    presum = np.zeros((L, X, Y))
    for l in range(L):
        if (l & 1) ^ DIR:  # vertical: along y
            presum[l, :, 1:] = np.cumsum(w[l, :, :-1], axis=1)
        else:
            presum[l, 1:, :] = np.cumsum(w[l, :-1, :], axis=0)
    vcost = rng.integers(1, 6, (L, X, Y)).astype(np.float32)
    return presum.reshape(-1), vcost.reshape(-1)


def _enumerate_optimum(L, X, Y, DIR, presum, vcost, nodes, edges):
    """Minimum cost over every routing tree the DAG encodes: each node takes
    a layer; each child slot one alternative child, which reaches the node
    straight on its layer or, over more than 4 cells, changing to another
    layer of the same parity at any interior position; each node's via
    stack spans its pin layer, its own layer and its children's arrival
    layers."""
    XY = X * Y
    P3 = presum.reshape(L, X, Y)
    V3 = vcost.reshape(L, X, Y).astype(np.float64)
    xy = lambda c: (c // Y % X, c % Y)
    alts = {}
    for j, (par, code) in enumerate(edges):
        alts.setdefault((int(par), int(code % 10)), []).append((int(code // 10), j))

    def wire(l, a, b):
        (ax, ay), (bx, by) = a, b
        if ax == bx:
            return P3[l, ax, max(ay, by)] - P3[l, ax, min(ay, by)]
        return P3[l, max(ax, bx), ay] - P3[l, min(ax, bx), ay]

    def fits(l, a, b):  # layer l (0-based) runs along the edge a-b
        vertical = (l & 1) ^ DIR
        return (a[0] == b[0]) if vertical else (a[1] == b[1])

    def arrivals(c, p, lc):
        """(arrival layer, edge cost) of child c at parent p leaving on layer lc."""
        a, b = xy(int(nodes[c, 0])), xy(int(nodes[p, 0]))
        out = []
        if a == b:
            return [(lc, 0.0)]
        if not fits(lc - 1, a, b):
            return []
        out.append((lc, wire(lc - 1, a, b)))
        d = abs(a[0] - b[0]) + abs(a[1] - b[1])
        if d > 4:
            for t in range(1, L + 1):
                if t == lc or (t - lc) % 2:
                    continue
                best = None
                vertical = a[0] == b[0]
                lo, hi = (min(a[1], b[1]), max(a[1], b[1])) if vertical else (min(a[0], b[0]), max(a[0], b[0]))
                for s in range(lo + 1, hi):
                    q = (a[0], s) if vertical else (s, a[1])
                    f, tt = lc - 1, t - 1
                    cost = V3[min(f, tt):max(f, tt), q[0], q[1]].sum() + wire(f, a, q) + wire(tt, q, b)
                    best = cost if best is None else min(best, cost)
                if best is not None:
                    out.append((t, best))
        return out

    memo = {}

    def best(v, lv):
        if (v, lv) in memo:
            return memo[(v, lv)]
        c = int(nodes[v, 0])
        pin = c // XY + 1 if c // XY < L else None
        kids = int(nodes[v, 2])
        per_slot = []
        for s in range(kids):
            cand = []
            for child, _ in alts.get((v, s), []):
                for lc in range(1, L + 1):
                    sub = best(child, lc)
                    if sub is None:
                        continue
                    for t, ec in arrivals(child, v, lc):
                        cand.append((t, sub + ec))
            if not cand:
                memo[(v, lv)] = None
                return None
            per_slot.append(cand)
        res = None
        x, y = xy(c)
        for combo in itertools.product(*per_slot):
            need = [lv] + [t for t, _ in combo] + ([pin] if pin else [])
            lo, hi = min(need), max(need)
            cost = sum(sc for _, sc in combo) + V3[lo - 1:hi - 1, x, y].sum()
            res = cost if res is None else min(res, cost)
        memo[(v, lv)] = res
        return res

    vals = [best(0, l) for l in range(1, L + 1)]
    return min(v for v in vals if v is not None)


def _words_cost(L, X, Y, presum, vcost, words):
    P3 = presum.reshape(L, X, Y)
    V3 = vcost.reshape(L, X, Y).astype(np.float64)
    XY = X * Y
    total = 0.0
    for i in range(1, words[0], 2):
        a, b = int(words[i]), int(words[i + 1])
        la, lb = a // XY, b // XY
        ax, ay, bx, by = a // Y % X, a % Y, b // Y % X, b % Y
        if la != lb:
            total += V3[min(la, lb):max(la, lb), ax, ay].sum()
        else:
            total += P3[la, bx, by] - P3[la, ax, ay]
    return total


def _random_tree(rng, L, X, Y, n):
    """A random tree of n distinct cells (node 0 a pin), pins on random
    layers, degree-3 Steiner nodes kept as given, adjacency lists."""
    cells = rng.choice(X * Y, n, replace=False)
    adj = [[] for _ in range(n)]
    for i in range(1, n):
        j = int(rng.integers(0, i))
        adj[i].append(j)
        adj[j].append(i)
    pos = []
    for i, c in enumerate(cells):
        steiner = len(adj[i]) > 2 and rng.random() < 0.5
        layer = L if steiner else int(rng.integers(0, L))
        pos.append(layer * X * Y + int(c))
    return pos, adj


def _flt_plan_one(nodes, edges, net_id=0):
    return dict(batch_off=np.array([0, 1]), net_ids=np.array([net_id]),
                node_off=np.array([0, len(nodes)], np.int64), nodes=nodes.reshape(-1).astype(np.int32),
                edge_off=np.array([0, len(edges)], np.int64), edges=edges.reshape(-1).astype(np.int32))


@pytest.mark.parametrize("L,DIR", [(2, 0), (3, 1), (4, 0)])
@pytest.mark.parametrize("seed", range(4))
def test_flt_dp_equals_enumeration(ga, ora, L, DIR, seed):
    """The traced routes of the DAG DP cost exactly the optimum over all
    routing trees the DAG encodes (corner, Z-shape and mirror alternatives,
    straight and layer-changing edges, via stacks)."""
    rng = np.random.default_rng(100 * seed + L)
    X, Y = 14, 12
    presum, vcost = _tiny_costs(rng, L, X, Y, DIR)
    grid = dict(L=L, X=X, Y=Y, DIR=DIR, unit_length_wire_cost=1.0, unit_via_cost=1.0,
                unit_length_short_costs=np.ones(L), x_edge_len=np.ones(X - 1, np.int32),
                y_edge_len=np.ones(Y - 1, np.int32), capacity=np.ones(L * X * Y, np.float32))
    g = ora.Grid(grid)
    checked = 0
    for trial in range(6):
        n = int(rng.integers(2, 5))
        pos, adj = _random_tree(rng, L, X, Y, n)
        worst = (rng.random((X, Y)) < 0.15).astype(np.float32)
        xs = np.concatenate([np.zeros((1, Y)), np.cumsum(worst, 0)[:-1]]).astype(np.float32).reshape(-1)
        ys = np.concatenate([np.zeros((X, 1)), np.cumsum(worst, 1)[:, :-1]], 1).astype(np.float32).reshape(-1)
        nodes, edges, _ = build_dag(ga, X, Y, L, pos, adj, xs if trial % 2 else None, ys if trial % 2 else None)
        if len(nodes) > 40:
            continue
        pins = int(np.sum(nodes[:, 0] // (X * Y) < L))
        plan = _flt_plan_one(nodes, edges)
        acc = np.array([0, pins], np.int32)
        routes, layer, rc = g.flt_batch(presum, vcost, plan, 0, 1, acc, np.array([0], np.int32))
        assert rc == 0
        assert layer[0] > 0
        opt = _enumerate_optimum(L, X, Y, DIR, presum, vcost, nodes, edges)
        assert _words_cost(L, X, Y, presum, vcost, routes) == opt, (seed, trial)
        checked += 1
    assert checked >= 3


def _components_ok(X, Y, words, pins):
    XY = X * Y
    parent = {}

    def find(a):
        parent.setdefault(a, a)
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(1, words[0], 2):
        a, b = int(words[i]), int(words[i + 1])
        la, lb = a // XY, b // XY
        ax, ay, bx, by = a // Y % X, a % Y, b // Y % X, b % Y
        steps = []
        if la != lb:
            steps = [((l, ax, ay), (l + 1, ax, ay)) for l in range(min(la, lb), max(la, lb))]
        elif ax != bx:
            steps = [((la, x, ay), (la, x + 1, ay)) for x in range(min(ax, bx), max(ax, bx))]
        elif ay != by:
            steps = [((la, ax, y), (la, ax, y + 1)) for y in range(min(ay, by), max(ay, by))]
        for p, q in steps:
            parent[find(p)] = find(q)
    return len({find(q) for q in pins}) == 1


def flt_instance(ga, ora, config=3, seed=7, scale=0.01, side=64):
    """L-shape stage (oracle), rip-up, congestion view, FLT DAGs, batches
    (limit 300000) and the execution plan of a small congested instance."""
    d, t = topology(config, seed, scale, side)
    g = ora.Grid(t)
    color, nb = ora.batches(t["X"], t["Y"], t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    sched = ga.execution_schedule(t, color, nb)
    routes, demand, ts, wl, vias = g.lshape_stage(sched, t["pin_acc_num"], threads=4)
    flags = g.ripup(demand, routes, t["pin_acc_num"])
    cong, xs, ys = g.congestion_view(demand)
    fs = d.flt_stage(flags, xs, ys)
    geom = (fs["pt_off"], fs["pts"], fs["hs_off"], fs["hsegs"], fs["vs_off"], fs["vsegs"])
    fcol, fnb = ora.batches(t["X"], t["Y"], *geom, limit=300000)
    plan = d.flt_plan(fcol, fnb)
    return dict(d=d, t=t, g=g, sched=sched, routes=routes, demand=demand, ts=ts, wl=wl, vias=vias, flags=flags,
                view=(cong, xs, ys), fs=fs, geom=geom, color=fcol, nb=fnb, plan=plan)


def test_flt_stage_reroutes_connect_every_pin(ga, ora):
    """After the stage every rerouted net's phase-two words connect all of
    its pins; the stage removes phase one and commits phase two."""
    I = flt_instance(ga, ora)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    assert fs["nets"] > 500 and I["nb"] > 20
    first = len(I["sched"]["batch_off"])
    r2, dem2, ts2, wl2, vias2, rc, last = I["g"].flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"],
                                                            I["demand"], I["ts"], I["routes"], first,
                                                            wl=I["wl"], vias=I["vias"])
    assert rc == 0
    assert last == first + 2 * (len(plan["batch_off"]) - 1) - 1
    X, Y, XY = t["X"], t["Y"], t["X"] * t["Y"]
    d = I["d"]
    for k, sid in enumerate(plan["net_ids"]):
        words = r2[100 * fs["pin_off2"][sid]:]
        assert 1 <= words[0] <= 100 * (t["pin_acc_num"][sid + 1] - t["pin_acc_num"][sid])
        recs = plan["nodes"][4 * plan["node_off"][k]:4 * plan["node_off"][k + 1]].reshape(-1, 4)
        pins = {(int(c) // XY, int(c) // Y % X, int(c) % Y) for c in recs[:, 0] if c // XY < t["L"]}
        if len({p[1:] for p in pins}) > 1:
            assert _components_ok(X, Y, words, pins), sid
    # without rerouted nets the demand returns to phase one's less what was removed
    assert not np.array_equal(dem2, I["demand"])
    assert np.all(dem2 >= 0)
    del d
