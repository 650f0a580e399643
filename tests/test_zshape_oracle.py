"""Z-shape (GGR) stage on the CPU: the host plan (bounding-box batches,
tree encoding, post-processing) and the oracle's DP, pinned by invariants:
every traced route costs exactly the DP optimum, connects every pin, and
canonicalisation keeps the same unit edges and vias."""
import numpy as np
import pytest

from tests._instances import random_demand, topology


def _plan(ga, ora, config, seed, scale=1.0):
    d, t = topology(config, seed, net_scale=scale)
    z = d.zshape_stage()
    color, nb = ora.batches(t["X"], t["Y"], z["pt_off"], z["pts"], z["hs_off"], z["hsegs"], z["vs_off"], z["vsegs"])
    return d, t, z, color, nb, d.zshape_plan(color, nb)


def _records(plan, b):
    r0, r1 = plan["rec_off"][b], plan["rec_off"][b + 1]
    return plan["records"][r0:r1]


def _nets_of(rec, count):
    for k in range(count):
        base = rec[k]
        yield k, rec[base], rec[base + 1], rec[base + 2:base + 2 + 6 * rec[base + 1]].reshape(-1, 6)


@pytest.mark.parametrize("config,seed", [(6, 0), (1, 1)])
def test_zshape_plan_invariants(ga, ora, config, seed):
    d, t, z, color, nb, plan = _plan(ga, ora, config, seed, 1.0 if config == 6 else 0.2)
    bo, ids = plan["batch_off"], plan["net_ids"]
    assert sorted(ids.tolist()) == sorted(z["order"].tolist())
    slot = {int(v): k for k, v in enumerate(z["order"])}
    group = [color[slot[int(i)]] for i in ids]
    # execution order: RSMT groups from the last color to the first, each
    # group refined into batches of at most 250 nets with disjoint boxes
    assert all(group[i] >= group[i + 1] for i in range(len(group) - 1))
    X, Y = t["X"], t["Y"]
    n = max(X, Y)
    for b in range(len(bo) - 1):
        count = bo[b + 1] - bo[b]
        assert 1 <= count <= 250
        assert len({group[e] for e in range(bo[b], bo[b + 1])}) == 1
        boxes = []
        for k, raw, nodes, recs in _nets_of(_records(plan, b), count):
            assert nodes >= 1 and recs[0][1:3].tolist() != [0, 0]
            xs, ys = recs[:, 0] // n, recs[:, 0] % n
            boxes.append((xs.min(), xs.max(), ys.min(), ys.max()))
            for i, r in enumerate(recs):  # children after their parent, at most three
                for c in r[3:]:
                    assert c == -1 or (c % 6 == 0 and i < c // 6 < nodes)
        for i in range(len(boxes)):  # tree nodes lie in the net's box: boxes disjoint
            for j in range(i + 1, len(boxes)):
                a, c = boxes[i], boxes[j]
                assert a[1] < c[0] or c[1] < a[0] or a[3] < c[2] or c[3] < a[2]


def _raw_cost(t, w, v, p, raw):
    """Cost of raw words [start, length | -1] in the Z-shape DP's units."""
    L, X, Y, DIR = t["L"], t["X"], t["Y"], t["DIR"]
    n = max(X, Y)
    nn = n * n

    def cell(index):
        layer = index // nn - 1
        a, b = index // n % n, index % n
        vert = (layer & 1) ^ DIR
        x, y = (a, b) if vert else (b, a)
        return layer, (layer * X + x) * Y + y

    total = 0
    for i in range(1, raw[0], 2):
        start, length = int(raw[i]), int(raw[i + 1])
        layer, c0 = cell(start)
        if length == -1:
            total += int(min(max(np.float32(v[c0]), np.float32(1)), np.float32(1e6)))
        else:
            _, c1 = cell(start + length)
            total += int(min(abs(p[c1] - p[c0]), 5e8))
    return total


@pytest.mark.parametrize("config,seed", [(6, 0), (6, 3)])
def test_zshape_trace_costs_the_optimum(ga, ora, config, seed):
    """The traced route lies in the DP's route space, so it never costs less
    than the DP optimum, and it costs exactly the optimum unless the optimum
    climbs more than one layer at the bend next to the parent node: the
    traceback (src/routing/zshape.cuh:311-315) takes the middle leg's layer
    from the one via step recorded for that bend, so such routes come back
    with the middle leg on the adjacent layer."""
    d, t, z, color, nb, plan = _plan(ga, ora, config, seed)
    grid = ora.Grid(t)
    w, v, p = grid.refresh(random_demand(t, seed, high=6), integer_wire=True)
    assert np.all(w == np.floor(w))
    bo = plan["batch_off"]
    checked = exact = 0
    for b in range(len(bo) - 1):
        count = bo[b + 1] - bo[b]
        rec = _records(plan, b)
        raw, cost = grid.zshape_batch(p, v, rec, count, int(plan["raw_cap"][bo[b]:bo[b + 1]].sum()))
        for k, off, nodes, _ in _nets_of(rec, count):
            assert raw[off] > 0 and cost[k] < 10**9
            traced = _raw_cost(t, w, v, p, raw[off:])
            assert traced >= cost[k], (b, k)
            exact += traced == cost[k]
            checked += 1
    assert checked > 1000 and exact >= 0.98 * checked, (exact, checked)


def _components_ok(t, words, pins):
    X, Y = t["X"], t["Y"]
    XY = X * Y
    parent = {}

    def find(a):
        parent.setdefault(a, a)
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    def union(a, b):
        parent[find(a)] = find(b)

    for i in range(1, words[0], 2):
        a, b = int(words[i]), int(words[i + 1])
        la, lb = a // XY, b // XY
        ax, ay, bx, by = a // Y % X, a % Y, b // Y % X, b % Y
        if la != lb:
            for l in range(min(la, lb), max(la, lb)):
                union((l, ax, ay), (l + 1, ax, ay))
        elif ax != bx:
            for x in range(min(ax, bx), max(ax, bx)):
                union((la, x, ay), (la, x + 1, ay))
        elif ay != by:
            for y in range(min(ay, by), max(ay, by)):
                union((la, ax, y), (la, ax, y + 1))
    roots = {find(q) for q in pins}
    return len(roots) == 1


def test_zshape_stage_routes_connect_every_pin(ga, ora):
    d, t, z, color, nb, plan = _plan(ga, ora, 6, 2)
    routes, demand, ts, wl, vias, rc = ora.Grid(t).zshape_stage(plan, t["pin_acc_num"])
    assert rc == 0
    n = max(t["X"], t["Y"])
    acc = t["pin_acc_num"]
    bo = plan["batch_off"]
    before = routes.copy()
    for b in range(len(bo) - 1):
        rec = _records(plan, b)
        for k, off, nodes, recs in _nets_of(rec, bo[b + 1] - bo[b]):
            sid = int(plan["net_ids"][bo[b] + k])
            pins = [(int(r[1]) - 1, int(r[0]) // n, int(r[0]) % n) for r in recs if r[1] != -1]
            words = routes[256 * acc[sid]:]
            if len(pins) > 1:
                assert _components_ok(t, words, pins), sid
    # canonical form keeps the unit edges and vias of every net
    promoted = d.zshape_finish(routes)
    assert promoted >= 0
    XY = t["X"] * t["Y"]

    def units(words):
        out = set()
        for i in range(1, words[0], 2):
            a, b = int(words[i]), int(words[i + 1])
            if a // XY != b // XY:
                for l in range(min(a, b) // XY, max(a, b) // XY):
                    out.add(("v", l, a % XY))
            elif a != b:
                lo, hi = min(a, b), max(a, b)
                step = 1 if (hi - lo) < t["Y"] and lo // t["Y"] == hi // t["Y"] else t["Y"]
                for q in range(lo, hi, step):
                    out.add(("w", q))
        return out

    changed = 0
    for sid in range(len(acc) - 1):
        a, c = before[256 * acc[sid]:], routes[256 * acc[sid]:]
        if units(a) != units(c):
            changed += 1  # only nets with promoted segments move
    assert changed <= promoted
