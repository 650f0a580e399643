"""Evaluator score (resources/evaluator.cpp:458-675): the oracle against the
reference evaluator built in oracle/_ref on whole .out files, hand-built
cases whose demand follows from the evaluator's rules, and (GPU) gr_score
against the oracle on resident routes."""
import gzip
import os

import numpy as np
import pytest

from tests import _golden as G
from tests._instances import topology


def read_cap(path):
    """CAP text as the evaluator reads it: binary64 capacities per metal-layer
    cell (z * X + x) * Y + y, directions, unit and overflow costs, edge
    lengths (resources/evaluator.cpp:311-405)."""
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "rt") as f:
        tok = f.read().split()
    it = iter(tok)
    L, X, Y = int(next(it)), int(next(it)), int(next(it))
    unit_wire, unit_via = float(next(it)), float(next(it))
    ovf = np.array([float(next(it)) for _ in range(L)])
    xl = np.array([int(next(it)) for _ in range(X - 1)], np.int32)
    yl = np.array([int(next(it)) for _ in range(Y - 1)], np.int32)
    dirs = np.zeros(L, np.int32)
    cap = np.zeros((L, X, Y))
    for z in range(L):
        next(it)
        dirs[z] = int(next(it))
        next(it)
        rows = np.array([float(next(it)) for _ in range(X * Y)]).reshape(Y, X)
        cap[z] = rows.T
    return dict(L=L, X=X, Y=Y, unit_wire=unit_wire, unit_via=unit_via, ovf=ovf, x_len=xl, y_len=yl, dir=dirs,
                capacity=cap.reshape(-1))


def read_out_lines(path, X, Y):
    """Per net of a .out file its lines as metal-layer cell pairs, in file
    order, as extra pairs (no pins, no subnets), and the net names."""
    extra_off, extra, names, open_net = [0], [], [], False
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line == "(":
                continue
            if line == ")":
                extra_off.append(len(extra))
                open_net = False
                continue
            v = line.split()
            if len(v) == 6:
                xl, yl, zl, xh, yh, zh = map(int, v)
                extra += [(zl * X + xl) * Y + yl, (zh * X + xh) * Y + yh]
            elif not open_net:
                names.append(line)
                open_net = True
    n = len(extra_off) - 1
    z = np.zeros(n + 1, np.int32)
    return dict(pin_off=z, pins=np.zeros(0, np.int32), extra_off=np.array(extra_off, np.int32),
                extra=np.array(extra, np.int32), sub_off=z, subs=np.zeros(0, np.int32), names=names)


def read_net_access(path, X, Y, order=None):
    """Access points of every net's pins as the evaluator reads the NET file
    (resources/evaluator.cpp:407-441: "[(z, x, y), ...]" per pin), as
    metal-layer cells; nets in file order, or in the order of the names in
    `order`."""
    op = gzip.open if path.endswith(".gz") else open
    nets = {}
    names = []
    with op(path, "rt") as f:
        for line in f:
            if "(" not in line and ")" not in line and len(line) > 1:
                names.append(line.strip())
                nets[names[-1]] = []
            elif "[" in line:
                v = [int(t) for t in line.translate(str.maketrans("(),[]", "     ")).split()]
                nets[names[-1]].append([(z * X + x) * Y + y for z, x, y in zip(v[0::3], v[1::3], v[2::3])])
    pin_off, acc_off, acc = [0], [0], []
    for name in (order if order is not None else names):
        for pin in nets[name]:
            acc += pin
            acc_off.append(len(acc))
        pin_off.append(len(acc_off) - 1)
    return dict(pin_off=np.array(pin_off, np.int32), access_off=np.array(acc_off, np.int32),
                access=np.array(acc, np.int32))


def oracle_score(ora, cap, nets, routes=None, acc=None, per_pin=12):
    return ora.score(cap["X"], cap["Y"], cap["L"], nets, routes, acc, per_pin, cap["dir"], cap["capacity"], cap["ovf"],
                     cap["x_len"], cap["y_len"], cap["unit_wire"], cap["unit_via"])


@pytest.mark.parametrize("name", G.NAMES)
def test_oracle_score_equals_reference_evaluator(ga, ora, name, tmp_path):
    """The oracle pipeline's .out (L-shape, rip-up, FLT) scored by the oracle
    and by the reference evaluator: every printed cost equal."""
    if not os.path.exists(ora.evaluator_path()):
        pytest.skip("reference evaluator not built (oracle/_ref)")
    from tests.test_pipeline import oracle_pipeline
    capf, netf = G.unpacked(name, "cap"), G.unpacked(name, "net")
    out = str(tmp_path / "oracle.out")
    oracle_pipeline(ga, ora, capf, netf, out)
    cap = read_cap(capf)
    mine = oracle_score(ora, cap, read_out_lines(out, cap["X"], cap["Y"]))
    ref, res = ora.run_evaluator(capf, netf, out)
    assert res.returncode == 0, res.stdout[-2000:]
    for key in ("wirelength", "via", "overflow", "total"):
        # the evaluator prints %.4lf
        assert abs(mine[key] - ref[key]) <= 5e-5 + 1e-12 * abs(ref[key]), (key, mine[key], ref[key])
    assert mine["rejected"] == 0


def _tiny_cap(X=6, Y=5, L=3):
    cap = np.full((L, X, Y), 4.0)
    cap[0] = 0.0
    return dict(L=L, X=X, Y=Y, unit_wire=0.5, unit_via=4.0, ovf=np.array([5.0, 6.0, 7.0]),
                x_len=np.full(X - 1, 10, np.int32), y_len=np.full(Y - 1, 20, np.int32),
                dir=np.array([0, 0, 1], np.int32), capacity=cap.reshape(-1))


def _cell(cap, z, x, y):
    return (z * cap["X"] + x) * cap["Y"] + y


def _lines(cap, nets):
    off, extra = [0], []
    for lines in nets:
        for (x0, y0, z0, x1, y1, z1) in lines:
            extra += [_cell(cap, z0, x0, y0), _cell(cap, z1, x1, y1)]
        off.append(len(extra))
    z = np.zeros(len(nets) + 1, np.int32)
    return dict(pin_off=z, pins=np.zeros(0, np.int32), extra_off=np.array(off, np.int32),
                extra=np.array(extra, np.int32), sub_off=z, subs=np.zeros(0, np.int32))


def test_stacked_vias_follow_the_evaluator_rules(ora):
    """Hand-derived demand: a via under the net's own wire adds nothing, a
    repeated via location counts once, a via of another net at a covered
    cell counts, border vias add 2 to the one inner edge, and wires count
    2 per unit edge."""
    cap = _tiny_cap()
    d = lambda r, z, x, y: int(r["demand"][_cell(cap, z, x, y)])  # noqa: E731
    nets = [
        # net 0: wire on layer 1 (horizontal) from x=1 to x=3 at y=2, via 1->2 at x=2 (under the wire)
        [(1, 2, 1, 3, 2, 1), (2, 2, 1, 2, 2, 2)],
        # net 1: via 1->2 at (2, 2) twice: counted once, on layer 1 at x-1 and x
        [(2, 2, 1, 2, 2, 2), (2, 2, 1, 2, 2, 2)],
        # net 2: via 1->2 at x = 0 (border: +2 on edge x=0) and at x = X-1 (+2 on edge X-2)
        [(0, 4, 1, 0, 4, 2), (5, 4, 1, 5, 4, 2)],
        # net 3: via 0->2 at (3, 0): stacked on layer 1; metal 0 is no routing
        # layer for the evaluator (no demand there), but its via is costed
        [(3, 0, 0, 3, 0, 2)],
    ]
    r = oracle_score(ora, cap, _lines(cap, nets))
    assert r["rejected"] == 0
    # wires of net 0: edges x=1, 2 on layer 1 -> 2 each; net 1's via adds 1 at x=1 and x=2
    assert d(r, 1, 1, 2) == 2 + 1 and d(r, 1, 2, 2) == 2 + 1 and d(r, 1, 3, 2) == 0
    assert d(r, 1, 0, 4) == 2 and d(r, 1, 4, 4) == 2
    assert d(r, 0, 2, 0) == 0 and d(r, 0, 3, 0) == 0 and d(r, 1, 2, 0) == 1 and d(r, 1, 3, 0) == 1
    assert int(r["demand"].astype(np.int64).sum()) == 2 * 2 + 2 + 2 + 2 + 2
    # wirelength: 2 unit edges of length 10; vias: layer 1: 1 + 2 + 2 + 1, layer 0: 1
    assert r["wirelength"] == pytest.approx(0.5 * 20)
    assert r["via"] == pytest.approx(4.0 * (1 + 2 + 2 + 1 + 1))
    # overflow: layers 1-2 (capacity 4) only
    dem = r["demand"].reshape(3, cap["X"], cap["Y"]).astype(np.float64)
    want = 0.0
    for z in (1, 2):
        want += cap["ovf"][z] * np.exp((dem[z] - 8.0) / 2 * 0.5).sum()
    assert r["overflow"] == pytest.approx(want, rel=1e-12)


def test_rejected_lines_and_demand_wrap(ora):
    """A wire against its layer's direction, a wire on metal 0 and a via that
    moves are rejected (nothing counted for them); demand is kept in 16 bits."""
    cap = _tiny_cap()
    nets = [[(1, 1, 2, 3, 1, 2)], [(1, 1, 1, 2, 2, 2)], [(0, 0, 1, 1, 0, 1)] * 40000, [(0, 3, 0, 2, 3, 0)]]
    r = oracle_score(ora, cap, _lines(cap, nets))
    assert r["rejected"] == 3
    assert r["via"] == 0.0
    # 40000 wires over one edge: 80000 mod 65536
    assert int(r["demand"][_cell(cap, 1, 0, 0)]) == 80000 % 65536


@pytest.mark.gpu
@pytest.mark.parametrize("config,seed", [(1, 0), (6, 4)])
def test_gpu_score_matches_oracle(ga, ora, config, seed):
    """gr_score on the resident routes of a GPU L-shape stage against the
    oracle on the same routes: demand and rejected nets identical, wirelength
    and via cost equal, overflow cost within 1e-10 relative: the device exp
    and the order of each layer's sum differ from the evaluator's sequential
    sum of up to millions of terms, whose own rounding error is of that
    order."""
    d, t = topology(config, seed)
    print(f"score config={config} seed={seed}", flush=True)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    color, nb = ctx.batch_plan(*geom)
    sched = ga.execution_schedule(t, color, nb)
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
    nets = d.nets()
    mg = d.metal_grid()
    cap = np.asarray(mg["capacity"], np.float64)
    got = ctx.score(nets, mg["dir"], cap, mg["overflow_cost"], demand=True)
    M, X, Y = mg["layers"], int(t["X"]), int(t["Y"])
    want = ora.score(X, Y, M, nets, ctx.routes(), t["pin_acc_num"], 12, mg["dir"], cap, mg["overflow_cost"],
                     t["x_edge_len"], t["y_edge_len"], t["unit_length_wire_cost"], t["unit_via_cost"])
    assert np.array_equal(got["demand"], want["demand"])
    assert got["rejected"] == want["rejected"] == 0
    assert got["wirelength"] == want["wirelength"] and got["via"] == want["via"]
    assert got["overflow"] == pytest.approx(want["overflow"], rel=1e-10)
    assert got["ms"] > 0
    ctx.close()


@pytest.mark.gpu
def test_gpu_score_hand_cases_match_oracle(ga, ora):
    """gr_score on extra-route lines only (no subnets routed): stacking under
    the net's own wire, repeated and border vias, metal-0 wires, moving vias
    and wires against their layer's direction (rejected), 16-bit demand wrap,
    and a net of 40,000 lines that exceeds the LDS line table and takes the
    global-memory path; demand, costs and rejected nets against the oracle."""
    d, t = topology(6, 0)
    print("score hand cases", flush=True)
    mg = d.metal_grid()
    M, X, Y = mg["layers"], int(t["X"]), int(t["Y"])
    dirs = [int(v) for v in mg["dir"]]
    hz = next(z for z in range(1, M) if dirs[z] == 0)  # a horizontal routing layer
    vz = next(z for z in range(1, M) if dirs[z] == 1)  # a vertical one
    lo, hi = min(hz, vz), max(hz, vz)
    cell = lambda z, x, y: (z * X + x) * Y + y  # noqa: E731
    nets = [
        [(1, 2, hz, 5, 2, hz), (3, 2, lo, 3, 2, hi)],      # via under the net's own wire on hz (when lo == hz)
        [(3, 2, lo, 3, 2, hi), (3, 2, lo, 3, 2, hi)],      # the same via twice
        [(0, 4, lo, 0, 4, hi), (X - 1, 4, lo, X - 1, 4, hi), (6, 0, lo, 6, 0, hi), (6, Y - 1, lo, 6, Y - 1, hi)],
        [(2, 2, 0, 2, 2, hi)],                              # from metal 0
        [(1, 1, 0, 4, 1, 0)],                               # wire on metal 0: rejected
        [(1, 1, lo, 2, 2, hi)],                             # moving via: rejected
        [(1, 1, vz, 4, 1, vz)],                             # along x on a vertical layer: rejected
        [(7, 7, hz, 8, 7, hz)] * 40000,                     # 80,000 demand on one edge: wraps
    ]
    off, extra = [0], []
    for lines in nets:
        for (x0, y0, z0, x1, y1, z1) in lines:
            extra += [cell(z0, x0, y0), cell(z1, x1, y1)]
        off.append(len(extra))
    z = np.zeros(len(nets) + 1, np.int32)
    lines = dict(pin_off=z, pins=np.zeros(0, np.int32), extra_off=np.array(off, np.int32),
                 extra=np.array(extra, np.int32), sub_off=z, subs=np.zeros(0, np.int32))
    cap = np.asarray(mg["capacity"], np.float64)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    got = ctx.score(lines, mg["dir"], cap, mg["overflow_cost"], demand=True)
    want = ora.score(X, Y, M, lines, ctx.routes(), t["pin_acc_num"], 12, mg["dir"], cap, mg["overflow_cost"],
                     t["x_edge_len"], t["y_edge_len"], t["unit_length_wire_cost"], t["unit_via_cost"])
    assert want["rejected"] == 3
    assert got["rejected"] == want["rejected"]
    assert np.array_equal(got["demand"], want["demand"])
    assert int(got["demand"][cell(hz, 7, 7)]) == 80000 % 65536
    assert got["wirelength"] == want["wirelength"] and got["via"] == want["via"]
    assert got["overflow"] == pytest.approx(want["overflow"], rel=1e-10)
    ctx.close()


def oracle_connectivity(ora, cap, nets, access):
    return ora.connectivity(cap["X"], cap["Y"], cap["L"], nets, None, None, 12, cap["dir"], access)


def _out_blocks(text):
    """(name line index, [wire line indices]) per net of .out text lines."""
    blocks, cur = [], None
    for i, line in enumerate(text):
        s = line.strip()
        if not s or s == "(":
            continue
        if s == ")":
            blocks.append(cur)
            cur = None
        elif cur is None:
            cur = (i, [])
        else:
            cur[1].append(i)
    return blocks


@pytest.mark.parametrize("name", G.NAMES)
def test_oracle_connectivity_equals_reference_evaluator(ga, ora, name, tmp_path):
    """The completion test against the reference evaluator: the oracle
    pipeline's .out is complete for both (no open or incomplete net); with
    one line dropped from a multi-pin net, the evaluator stops at the first
    net, in .out order, that the oracle finds not connected (printing its
    name, no costs), or completes when the oracle finds every net connected;
    a via that moves is an open line for the oracle and fails the
    evaluator's assertion."""
    if not os.path.exists(ora.evaluator_path()):
        pytest.skip("reference evaluator not built (oracle/_ref)")
    from tests.test_pipeline import oracle_pipeline
    capf, netf = G.unpacked(name, "cap"), G.unpacked(name, "net")
    out = str(tmp_path / "oracle.out")
    oracle_pipeline(ga, ora, capf, netf, out)
    cap = read_cap(capf)
    X, Y = cap["X"], cap["Y"]
    lines = read_out_lines(out, X, Y)
    access = read_net_access(netf, X, Y, order=lines["names"])
    assert not oracle_connectivity(ora, cap, lines, access).any()
    ref, res = ora.run_evaluator(capf, netf, out)
    assert res.returncode == 0 and ref["open"] == 0 and ref["incomplete"] == 0
    text = open(out).read().split("\n")
    blocks = _out_blocks(text)
    pins = np.diff(access["pin_off"])
    cand = [k for k, b in enumerate(blocks) if pins[k] >= 2 and len(b[1]) >= 2]
    rng = np.random.default_rng(11)
    opened = 0
    for trial in range(8):
        k = int(rng.choice(cand))
        drop = int(rng.choice(blocks[k][1]))
        planted = str(tmp_path / f"planted{trial}.out")
        with open(planted, "w") as f:
            f.write("\n".join(t for i, t in enumerate(text) if i != drop))
        st = oracle_connectivity(ora, cap, read_out_lines(planted, X, Y), access)
        assert not (st == 1).any()
        ref, res = ora.run_evaluator(capf, netf, planted)
        assert res.returncode == 0, res.stderr[-500:]
        bad = np.flatnonzero(st == 2)
        if len(bad):
            opened += 1
            assert "total cost" not in res.stdout
            assert res.stderr.splitlines()[0] == lines["names"][bad[0]], (res.stderr[:200], bad[:3])
        else:
            assert ref["open"] == 0 and ref["incomplete"] == 0
    assert opened >= 3
    # a via that moves in x: rejected (open) for the oracle, the evaluator's assertion fails
    i = blocks[cand[0]][1][0]
    v = text[i].split()
    bad_line = f"{v[0]} {v[1]} 1 {int(v[0]) + 1} {v[1]} 2"
    planted = str(tmp_path / "moving_via.out")
    with open(planted, "w") as f:
        f.write("\n".join(bad_line if j == i else t for j, t in enumerate(text)))
    st = oracle_connectivity(ora, cap, read_out_lines(planted, X, Y), access)
    assert st[cand[0]] == 1 and (st == 1).sum() == 1
    ref, res = ora.run_evaluator(capf, netf, planted)
    assert res.returncode != 0 and "Assertion" in res.stderr


def _conn_cases(X, Y):
    """Hand-built nets (lines, pins as lists of access points, expected
    status) on layers 0 (metal 0: no direction), 1 (horizontal), 2
    (vertical), each expectation following from the flood fill's rules."""
    c = lambda x, y, z: (z * X + x) * Y + y  # noqa: E731
    return [
        # wires on one track with adjacent ends (x = 2 and 3): neighbours on a horizontal layer
        ([(1, 2, 1, 2, 2, 1), (3, 2, 1, 4, 2, 1)], [[c(1, 2, 1)], [c(4, 2, 1)]], 0),
        # a horizontal and a vertical wire crossing at (1, 1) on adjacent layers, no via
        ([(0, 1, 1, 2, 1, 1), (1, 0, 2, 1, 3, 2)], [[c(0, 1, 1)], [c(1, 3, 2)]], 0),
        # a via column next to a wire's end along layer 1
        ([(3, 3, 1, 3, 3, 2), (4, 3, 1, 5, 3, 1)], [[c(3, 3, 2)], [c(5, 3, 1)]], 0),
        # one unmarked cell between two wires of a track
        ([(0, 4, 1, 1, 4, 1), (3, 4, 1, 4, 4, 1)], [[c(0, 4, 1)], [c(4, 4, 1)]], 2),
        # neighbours on metal 0 (no direction) and across layer 1's direction
        ([(0, 0, 0, 0, 0, 1), (0, 1, 0, 0, 1, 1)], [[c(0, 0, 0)], [c(0, 1, 0)]], 2),
        # a pin whose access point no line marks
        ([(1, 1, 2, 1, 3, 2)], [[c(1, 1, 2)], [c(2, 2, 2)]], 2),
        # a pin with two access points, the second on the route
        ([(2, 0, 2, 2, 2, 2)], [[c(2, 0, 2)], [c(4, 4, 1), c(2, 2, 2)]], 0),
        # one pin, no lines
        ([], [[c(3, 1, 1)]], 0),
        # a wire across its layer's direction: rejected, never flood filled
        ([(0, 0, 2, 2, 0, 2), (0, 0, 1, 2, 0, 1)], [[c(0, 0, 1)], [c(2, 0, 1)]], 1),
        # the first pin on no marked cell: nothing is traced
        ([(1, 1, 1, 3, 1, 1)], [[c(5, 0, 1)], [c(2, 1, 1)]], 2),
        # a via column through three layers joins its ends
        ([(4, 1, 0, 4, 1, 2)], [[c(4, 1, 0)], [c(4, 1, 2)]], 0),
        # two wires sharing a cell on one layer, pins at the far ends
        ([(0, 3, 1, 3, 3, 1), (2, 3, 1, 5, 3, 1)], [[c(0, 3, 1)], [c(5, 3, 1)], [c(2, 3, 1)]], 0),
    ]


def _conn_inputs(cases, X, Y):
    off, extra, pin_off, acc_off, acc = [0], [], [0], [0], []
    for lines, pins, _ in cases:
        for (x0, y0, z0, x1, y1, z1) in lines:
            extra += [(z0 * X + x0) * Y + y0, (z1 * X + x1) * Y + y1]
        off.append(len(extra))
        for p in pins:
            acc += p
            acc_off.append(len(acc))
        pin_off.append(len(acc_off) - 1)
    z = np.zeros(len(cases) + 1, np.int32)
    nets = dict(pin_off=z, pins=np.zeros(0, np.int32), extra_off=np.array(off, np.int32),
                extra=np.array(extra, np.int32), sub_off=z, subs=np.zeros(0, np.int32))
    access = dict(pin_off=np.array(pin_off, np.int32), access_off=np.array(acc_off, np.int32),
                  access=np.array(acc, np.int32))
    return nets, access


def test_connectivity_hand_cases(ora):
    """The flood fill's rules on hand-built nets: neighbours along a layer's
    direction and between layers count, metal 0 has no direction, any access
    point of a pin suffices, the first pin starts the fill."""
    cap = _tiny_cap()
    cases = _conn_cases(cap["X"], cap["Y"])
    nets, access = _conn_inputs(cases, cap["X"], cap["Y"])
    st = oracle_connectivity(ora, cap, nets, access)
    assert list(st) == [e for _, _, e in cases]


def _serpentine(X, Y, rows, gap=None):
    """A net of more lines than the LDS staging holds: unit wires along
    layer 1 (horizontal) row after row, rows joined by vertical wires on
    layer 2 at alternating ends; `gap` drops two consecutive unit wires, so
    that one cell of the row is unmarked (one dropped wire would leave its
    end cells marked, and neighbours along the layer are connected)."""
    lines = []
    for r in range(rows):  # row r on y = 2r: connectors at one end stay one cell apart
        for x in range(X - 1):
            lines.append((x, 2 * r, 1, x + 1, 2 * r, 1))
        if r + 1 < rows:
            xe = X - 1 if r % 2 == 0 else 0
            lines.append((xe, 2 * r, 2, xe, 2 * r + 2, 2))
    if gap is not None:
        del lines[gap:gap + 2]
    return (lines, [[(1 * X + 0) * Y + 0], [(1 * X + 5) * Y + 2 * rows - 2]], 0 if gap is None else 2)


@pytest.mark.gpu
def test_gpu_connectivity_matches_oracle(ga, ora, tmp_path):
    """gr_check_connectivity against the oracle, status per net: the hand
    cases, serpentine nets of more lines than the LDS staging holds (global
    staging), the golden pipeline's .out lines with single lines dropped,
    and the resident routes of an L-shape stage and of an FLT stage (its
    nets scored by their phase-two words)."""
    from tests.test_flt_oracle import flt_instance
    from tests.test_pipeline import oracle_pipeline
    d, t = topology(6, 0)
    X, Y, M = int(t["X"]), int(t["Y"]), int(t["L"]) + 1
    print(f"connectivity hand cases on {X}x{Y}x{M}", flush=True)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    dirs = np.array([0] + [(z + 1) % 2 for z in range(1, M)], np.int32)  # 0, 0, 1, 0, ...
    cases = _conn_cases(X, Y) + [_serpentine(X, Y, 20), _serpentine(X, Y, 20, gap=700)]
    nets, access = _conn_inputs(cases, X, Y)
    st, counts = ctx.connectivity(nets, dirs, access)
    want = ora.connectivity(X, Y, M, nets, None, None, 12, dirs, access)
    assert list(want) == [e for _, _, e in cases]
    assert np.array_equal(st, want)
    assert counts == dict(open=1, disconnected=int((want == 2).sum()), incomplete=int((want > 0).sum()))
    ctx.close()
    # the golden pipeline's lines, whole and with lines dropped
    for name in G.NAMES:
        capf, netf = G.unpacked(name, "cap"), G.unpacked(name, "net")
        print(f"connectivity golden {name}", flush=True)
        out = str(tmp_path / f"{name}.out")
        oracle_pipeline(ga, ora, capf, netf, out)
        cap = read_cap(capf)
        gd = ga.Design.read(capf, netf)
        gd.prepare(4)
        gt = gd.topology()
        gctx = ga.Context(gt)
        gctx.set_routes(gt["pin_acc_num"])
        text = open(out).read().split("\n")
        blocks = _out_blocks(text)
        rng = np.random.default_rng(3)
        drops = [None] + [int(rng.choice(blocks[k][1])) for k in rng.choice(len(blocks), 6) if blocks[k][1]]
        lines = read_out_lines(out, cap["X"], cap["Y"])
        access = read_net_access(netf, cap["X"], cap["Y"], order=lines["names"])
        for drop in drops:
            planted = str(tmp_path / "planted.out")
            with open(planted, "w") as f:
                f.write("\n".join(t_ for i, t_ in enumerate(text) if i != drop))
            pl = read_out_lines(planted, cap["X"], cap["Y"])
            st, _ = gctx.connectivity(pl, cap["dir"], access)
            assert np.array_equal(st, oracle_connectivity(ora, cap, pl, access)), (name, drop)
        gctx.close()
    # resident routes: L-shape stage, then the FLT stage's phase-two words
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    t, fs, plan, d = I["t"], I["fs"], I["plan"], I["d"]
    print("connectivity of resident routes", flush=True)
    X, Y, M = int(t["X"]), int(t["Y"]), int(t["L"]) + 1
    nets, access, mg = d.nets(), d.access(), d.metal_grid()
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    s = I["sched"]
    ctx.lshape_upload(s["net_ids"], s["node_off"], s["nodes"], s["parents"])
    ctx.lshape_run(s["batch_off"], first_stamp=1, mode=2)
    st, counts = ctx.connectivity(nets, mg["dir"], access)
    want = ora.connectivity(X, Y, M, nets, ctx.routes(), t["pin_acc_num"], 12, mg["dir"], access)
    assert np.array_equal(st, want) and counts["incomplete"] == 0
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.flt_upload(plan)
    ctx.flt_run(first_stamp=len(s["batch_off"]), remove_phase_one=True, mode=2)
    st, counts = ctx.connectivity(nets, mg["dir"], access)
    want = ora.connectivity(X, Y, M, nets, ctx.routes(), t["pin_acc_num"], 12, mg["dir"], access,
                            routes2=ctx.routes2(), pin_off2=fs["pin_off2"])
    assert np.array_equal(st, want) and counts["incomplete"] == 0
    ctx.close()
