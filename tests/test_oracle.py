"""Pins of the CPU oracle that do not depend on running the reference:
closed forms, error bounds, brute force on tiny inputs and invariants, each
chosen so that a plausible slip (a dropped term, a wrong neighbour, a wrong
sign, a transposed operand) fails it."""
import itertools

import numpy as np
import pytest

# --------------------------------------------------------------------------
# project exponential
# --------------------------------------------------------------------------


def _ulp_diff(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_expf_is_within_one_ulp_of_correctly_rounded(ora):
    """ora_expf against exp evaluated in binary64 and rounded once: at most
    1 ulp anywhere on the range the cost model uses, and exact on >99.9%."""
    xs = np.concatenate([np.linspace(-87.0, 88.5, 40001, dtype=np.float32),
                         np.random.default_rng(0).uniform(-20, 20, 20000).astype(np.float32),
                         np.arange(-40, 41, dtype=np.float32) * 0.5])
    got = np.array([ora.expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    d = _ulp_diff(got, want)
    assert d.max() <= 1
    assert (d == 0).mean() > 0.999
    assert ora.expf(0.0) == 1.0
    assert ora.expf(-100.0) == 0.0 and np.isinf(ora.expf(100.0))


# --------------------------------------------------------------------------
# cost model: src/cost/full_update.cuh:9-83
# --------------------------------------------------------------------------


def f32exp(x):
    return float(np.float32(np.exp(np.float64(np.float32(x)))))


def test_overflow_cost_closed_forms(ora):
    L = ora.lib()
    # capacity > 0.001: exp(0.5 * (d - c))
    assert L.ora_of_cost(10.0, 10.0, 0.5, 1.5) == 1.0
    assert L.ora_of_cost(10.0, 12.0, 0.5, 1.5) == f32exp(1.0)
    assert L.ora_of_cost(10.0, 4.0, 0.5, 1.5) == f32exp(-3.0)
    # capacity <= 0.001 with demand: exp(1.5 d); with no demand: 0
    assert L.ora_of_cost(0.0, 2.0, 0.5, 1.5) == f32exp(3.0)
    assert L.ora_of_cost(0.0005, 1.0, 0.5, 1.5) == f32exp(1.5)
    assert L.ora_of_cost(0.0, 0.0, 0.5, 1.5) == 0.0
    # incremental cost: (e(d + inc) - e(d)) * short cost, capped at 1e12
    v = L.ora_incremental_of_cost(10.0, 10.0, 1.0, 5.0, 0.5, 1.5)
    assert v == (f32exp(0.5) - 1.0) * 5.0
    v = L.ora_incremental_of_cost(10.0, 11.0, 0.5, 5.0, 0.5, 1.5)
    assert v == (f32exp(0.75) - f32exp(0.5)) * 5.0
    assert L.ora_incremental_of_cost(0.0, 50.0, 1.0, 5.0, 0.5, 1.5) == 1e12


def _tiny_grid(DIR=0, L=2, X=4, Y=3, cap=5.0):
    return dict(L=L, X=X, Y=Y, DIR=DIR, unit_length_wire_cost=0.5, unit_via_cost=4.0,
                unit_length_short_costs=np.full(L, 5.0), x_edge_len=np.arange(1, X, dtype=np.int32) * 10,
                y_edge_len=np.arange(1, Y, dtype=np.int32) * 100, capacity=np.full(L * X * Y, cap, np.float32))


@pytest.mark.parametrize("DIR", [0, 1])
def test_cost_refresh_neighbours_and_borders(ora, DIR):
    """A demand change at one cell moves wcost there and vcost there and at
    the next cell along the track only; border cells take the full
    increment of their single edge; wire length is added except at the last
    cell of a track; presum is the exclusive prefix of wcost."""
    g = _tiny_grid(DIR)
    grid = ora.Grid(g)
    X, Y = g["X"], g["Y"]
    idx = lambda l, x, y: (l * X + x) * Y + y
    base_w, base_v, base_p = grid.refresh(np.zeros(grid.cells, np.float32))
    demand = np.zeros(grid.cells, np.float32)
    demand[idx(0, 1, 1)] = 3.0
    w, v, p = grid.refresh(demand)
    changed_w = set(np.flatnonzero(w != base_w))
    changed_v = set(np.flatnonzero(v != base_v))
    horizontal = ((0 & 1) ^ DIR) == 0
    nxt = idx(0, 2, 1) if horizontal else idx(0, 1, 2)
    assert changed_w == {idx(0, 1, 1)}
    assert changed_v == {idx(0, 1, 1), nxt}
    inc = lambda d, a: ora.lib().ora_incremental_of_cost(5.0, d, a, 5.0, 0.5, 1.5)
    # interior cell: half + half, then the via cost
    exp_v = np.float32(np.float32(inc(3.0, 0.5) + inc(0.0, 0.5)) + 4.0)
    assert v[idx(0, 1, 1)] == exp_v
    # first cell of a track: full increment of its own edge; last: of the previous edge
    last = idx(0, X - 1, 0) if horizontal else idx(0, 0, Y - 1)
    assert v[idx(0, 0, 0)] == np.float32(np.float32(inc(0.0, 1.0)) + 4.0)
    assert v[last] == np.float32(np.float32(inc(0.0, 1.0)) + 4.0)
    # wire length on every edge but the track's last cell
    assert w[last] == np.float32(inc(0.0, 1.0))
    first_len = 10 if horizontal else 100
    assert w[idx(0, 0, 0)] == np.float32(np.float32(inc(0.0, 1.0)) + 0.5 * first_len)
    # presum: exclusive prefix along the track (integer-free here: compare
    # to the Fenwick restatement below, and check monotone + first zero)
    track = [idx(0, t, 1) for t in range(X)] if horizontal else [idx(0, 1, t) for t in range(Y)]
    assert p[track[0]] == 0.0
    assert np.all(np.diff(p[track]) > 0)


# --------------------------------------------------------------------------
# Fenwick-ordered prefix: src/cost/fenwick.cuh:129-168
# --------------------------------------------------------------------------


def _fenwick_reference(w):
    """Independent statement of the same sums: node(i) over leaves
    (i - 2^t, i] with leaf(j) = w[j-1], built as leaf(i) followed by the
    nodes i-1, i-2, i-4, ...; prefix(i) adds node(i) then the node ending
    at (i with its trailing ones cleared) - 1, and so on, starting from 0."""
    n = len(w)
    leaf = lambda j: float(np.float64(w[j - 1])) if j > 0 else 0.0
    node = {}

    def trailing_ones(i):
        t = 0
        while (i >> t) & 1:
            t += 1
        return t

    def val(j):
        return node[j] if j & 1 else leaf(j)

    for i in range(n):
        t = trailing_ones(i)
        if t == 0:
            continue
        v = leaf(i)
        for k in range(t):
            v = v + val(i - (1 << k))
        node[i] = v
    out = []
    for i in range(n):
        s, j = 0.0, i
        while j >= 0:
            s = s + val(j)
            j = (j & (j + 1)) - 1
        out.append(s)
    return np.array(out)


def test_fenwick_exact_on_integers(ora):
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 7, 8, 64, 100, 513):
        w = rng.integers(0, 1000, n).astype(np.float32)
        got = ora.fenwick_track(w)
        want = np.concatenate([[0.0], np.cumsum(w[:-1].astype(np.float64))])
        assert np.array_equal(got, want)


def test_fenwick_addition_order(ora):
    """Adversarial magnitudes make the addition order visible: the oracle
    equals the independent statement bit for bit and differs from a
    sequential sum.  The reference's dense block scan (compute_presum,
    src/cost/prefix_scan.cuh:74-114) walks the set bits of the index in the
    same ascending order, so it yields the very same values: the full-cost
    and incremental variants of the reference agree bit for bit."""
    rng = np.random.default_rng(2)
    w = (rng.choice([1e8, 1.0, 3e-3, 7.0], 257) * rng.uniform(0.5, 1.5, 257)).astype(np.float32)
    got = ora.fenwick_track(w)
    assert np.array_equal(got, _fenwick_reference(w))
    seq = np.concatenate([[0.0], np.cumsum(w[:-1].astype(np.float64))])
    assert not np.array_equal(got, seq)
    assert np.array_equal(got, ora.dense_prefix_track(w))


# --------------------------------------------------------------------------
# batching: src/batching/generate_batches_cpu.hpp:7-73
# --------------------------------------------------------------------------


def _marks(X, Y, pts, hs, vs, sparsity):
    m = set()
    for a, b in hs:
        m.update(range(a, b + 1, sparsity * Y))
    for a, b in vs:
        m.update(range(a, b + 1, sparsity))
    for p in pts:
        px, py = divmod(p, Y)
        for x, y in ((px - 1, py), (px, py - 1), (px, py), (px + 1, py)):
            if 0 <= x < X and 0 <= y < Y:
                m.add(x * Y + y)
    return m


def _random_nets(rng, X, Y, n):
    nets = []
    for _ in range(n):
        x0, y0 = rng.integers(0, X), rng.integers(0, Y)
        x1, y1 = min(X - 1, x0 + rng.integers(0, 4)), min(Y - 1, y0 + rng.integers(0, 4))
        pts = [x0 * Y + y0, x1 * Y + y1]
        hs, vs = [], []
        if x0 != x1:
            hs.append((min(x0, x1) * Y + y0, max(x0, x1) * Y + y0))
        if y0 != y1:
            vs.append((x1 * Y + min(y0, y1), x1 * Y + max(y0, y1)))
        nets.append((pts, hs, vs))
    return nets


def _pack(nets):
    pt_off, pts, hs_off, hs, vs_off, vs = [0], [], [0], [], [0], []
    for p, h, v in nets:
        pts += p
        pt_off.append(len(pts))
        for a, b in h:
            hs += [a, b]
        hs_off.append(len(hs) // 2)
        for a, b in v:
            vs += [a, b]
        vs_off.append(len(vs) // 2)
    return pt_off, pts, hs_off, hs, vs_off, vs


@pytest.mark.parametrize("seed,sparsity,limit", [(0, 1, 10**6), (1, 2, 10**6), (2, 1, 3), (3, 1, 10**6)])
def test_batching_first_fit_invariants(ora, seed, sparsity, limit):
    rng = np.random.default_rng(seed)
    X, Y = 12, 10
    nets = _random_nets(rng, X, Y, 120)
    color, nb = ora.batches(X, Y, *_pack(nets), limit=limit, sparsity=sparsity)
    marks = [_marks(X, Y, p, h, v, sparsity) for p, h, v in nets]
    members = {}
    for k, c in enumerate(color):
        # conflict-free against every earlier member of its batch
        for j in members.get(c, []):
            assert not (marks[j] & set(nets[k][0]))
        # first fit: every lower batch conflicts or was full
        for c2 in range(c):
            earlier = members.get(c2, [])
            assert len(earlier) >= limit or any(marks[j] & set(nets[k][0]) for j in earlier)
        members.setdefault(c, []).append(k)
    assert nb == len(members) and max(len(v) for v in members.values()) <= limit


# --------------------------------------------------------------------------
# L-shape DP: brute force over every layer assignment of tiny trees
# --------------------------------------------------------------------------


def _brute_force(g, vcost, presum, nodes, parents):
    """Minimum over all per-edge layer choices of the reference's cost:
    straight edge on a preferred-direction layer c: presum difference on c;
    bent edge: first leg on curl (own direction), second on parl of opposite
    parity, vias strictly between them at the corner; every node pays the
    via stack over the span of its fixed layer and incident edge layers."""
    L, X, Y, DIR = g["L"], g["X"], g["Y"], g["DIR"]
    idx = lambda l, x, y: (l * X + x) * Y + y
    d = lambda l: (l & 1) ^ DIR
    wire = lambda l, x0, x1, y0, y1: presum[idx(l, max(x0, x1), max(y0, y1))] - presum[idx(l, min(x0, x1), min(y0, y1))]
    xy = lambda v: (v // Y % X, v % Y)
    n = len(nodes)
    options = []
    for i in range(1, n):
        (x, y), (px, py) = xy(nodes[i]), xy(nodes[parents[i]])
        opts = []
        if x == px or y == py:
            for c in range(L):
                if (d(c) == 0 and y != py) or (d(c) == 1 and x != px):
                    continue
                opts.append((c, c, wire(c, x, px, y, py)))
        else:
            for cl in range(L):
                w1 = wire(cl, x, x, y, py) if d(cl) else wire(cl, x, px, y, y)
                cx, cy = (x, py) if d(cl) else (px, y)
                for pl in range((cl & 1) ^ 1, L, 2):
                    v = sum(float(vcost[idx(l, cx, cy)]) for l in range(min(cl, pl) + 1, max(cl, pl)))
                    w2 = wire(pl, px, px, y, py) if d(pl) else wire(pl, x, px, py, py)
                    opts.append((cl, pl, w1 + v + w2))
        options.append(opts)
    best = None
    for choice in itertools.product(*options):
        used = [set() for _ in range(n)]
        for i in range(n):
            if nodes[i] // (X * Y) < L:
                used[i].add(nodes[i] // (X * Y))
        total = 0.0
        for i, (cl, pl, c) in zip(range(1, n), choice):
            used[i].add(cl)
            used[parents[i]].add(pl)
            total += c
        for i in range(n):
            if used[i]:
                x, y = xy(nodes[i])
                total += sum(float(vcost[idx(l, x, y)]) for l in range(min(used[i]), max(used[i])))
        best = total if best is None else min(best, total)
    return best


def _random_tree(rng, g, n):
    L, X, Y = g["L"], g["X"], g["Y"]
    cells = rng.choice(X * Y, n, replace=False)
    parents = [-1] + [int(rng.integers(0, i)) for i in range(1, n)]
    nodes = []
    for i, c in enumerate(cells):
        layer = int(rng.integers(0, L)) if rng.random() < 0.6 else L  # pin or Steiner
        nodes.append(layer * X * Y + int(c))
    return nodes, parents


@pytest.mark.parametrize("L,DIR,seed", [(2, 0, 0), (3, 1, 1), (4, 0, 2), (4, 1, 3)])
def test_lshape_dp_equals_brute_force(ora, L, DIR, seed):
    rng = np.random.default_rng(seed)
    g = _tiny_grid(DIR, L=L, X=6, Y=5)
    grid = ora.Grid(g)
    for trial in range(25):
        demand = (rng.integers(0, 14, grid.cells) * 0.5).astype(np.float32)
        _, vcost, presum = grid.refresh(demand)
        n = int(rng.integers(2, 5))
        nodes, parents = _random_tree(rng, g, n)
        rc, routes, best = grid.lshape_net(vcost, presum, nodes, parents, with_cost=True)
        assert rc == 0
        brute = _brute_force(g, vcost, presum, nodes, parents)
        assert best == pytest.approx(brute, rel=1e-12, abs=1e-9)


def test_lshape_two_pin_straight_closed_form(ora):
    """Two pins on one horizontal track, both fixed on layer 0 of a grid whose
    layer 0 is horizontal: the optimum is the wire on layer 0 (any other
    layer adds two via stacks); routes = [3, a, b]."""
    g = _tiny_grid(0, L=3, X=6, Y=5)
    grid = ora.Grid(g)
    _, vcost, presum = grid.refresh(np.zeros(grid.cells, np.float32))
    X, Y = 6, 5
    a, b = 1 * Y + 2, 4 * Y + 2
    rc, routes, best = grid.lshape_net(vcost, presum, [a, b], [-1, 0], with_cost=True)
    assert rc == 0
    assert list(routes) == [3, a, b]
    assert best == presum[b] - presum[a]


# --------------------------------------------------------------------------
# demand commit: src/routing/demand.cuh:236-319
# --------------------------------------------------------------------------


def test_commit_wires_then_vias_timestamp_rule(ora):
    """A via only adds demand on layers whose cell carries no wire of the
    same batch; interior vias split half/half, border vias put 1 on the
    single edge; totals follow."""
    import ctypes
    g = _tiny_grid(0, L=3, X=5, Y=4)
    grid = ora.Grid(g)
    X, Y = 5, 4
    idx = lambda l, x, y: (l * X + x) * Y + y
    # wire on layer 0 (horizontal) from x=1..3 at y=2; via stack at (2,2) from layer 0 to 2;
    # via at the border cell (0,1) from layer 0 to 1 (layer 0 horizontal: x == 0 border)
    r = np.array([7, idx(0, 1, 2), idx(0, 3, 2), idx(0, 2, 2), idx(2, 2, 2), idx(0, 0, 1), idx(1, 0, 1)], np.int32)
    demand = np.zeros(grid.cells, np.float32)
    ts = np.zeros(grid.cells, np.int32)
    wl = ctypes.c_double(0)
    vias = ctypes.c_longlong(0)
    ptrs = (ctypes.POINTER(ctypes.c_int) * 1)(r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    ora.lib().ora_commit_batch(ctypes.byref(grid.c), demand.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               ts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1, ptrs, 5, 1, ctypes.byref(wl),
                               ctypes.byref(vias), None)
    want = np.zeros(grid.cells, np.float32)
    want[idx(0, 1, 2)] += 1
    want[idx(0, 2, 2)] += 1            # wire edges x=1->2, 2->3
    # via (2,2): layer 0 cell has the wire's timestamp -> skipped; layer 1 is
    # vertical (dir 1): y=2 interior -> half on (2,1) and (2,2)
    want[idx(1, 2, 1)] += 0.5
    want[idx(1, 2, 2)] += 0.5
    # via (0,1) layer 0 horizontal, x == 0 border -> full unit on (0,1)
    want[idx(0, 0, 1)] += 1
    assert np.array_equal(demand, want)
    assert wl.value == 10 * 2 + 10 * 3        # x_edge_len[1] + x_edge_len[2]
    assert vias.value == 2 + 1
    assert ts[idx(0, 1, 2)] == 5 and ts[idx(0, 3, 2)] == 5 and ts[idx(1, 2, 2)] == 0


def test_ripup_counts_overflowed_edges(ora):
    """A net is ripped up when more than 10 of its wire edges have
    capacity + threshold <= demand (src/routing/congestion.cuh:8-58)."""
    g = _tiny_grid(0, L=1, X=20, Y=2, cap=1.0)
    grid = ora.Grid(g)
    X, Y = 20, 2
    idx = lambda l, x, y: (l * X + x) * Y + y
    routes = np.zeros(12 * 2 + 1, np.int32)
    routes[0:3] = [3, idx(0, 0, 0), idx(0, 19, 0)]
    routes[12:15] = [3, idx(0, 0, 1), idx(0, 19, 1)]
    demand = np.zeros(grid.cells, np.float32)
    demand[[idx(0, x, 0) for x in range(11)]] = 1.0      # 11 overflowed edges on row 0
    demand[[idx(0, x, 1) for x in range(10)]] = 1.0      # 10 on row 1
    flags = grid.ripup(demand, routes, np.array([0, 1, 2], np.int32))
    assert list(flags) == [1, 0]
    assert list(grid.ripup(demand, routes, np.array([0, 1, 2], np.int32), threshold=1)) == [0, 0]
