"""FLT stage on the GPU against the CPU oracle: the L-shape stage, rip-up
flags, congestion view, FLT batches and the FLT stage itself (phase-one
removal, DAG DP, phase-two commit).  Every result is compared exactly:
route words (appended in the same deterministic order), demand bits,
timestamps and statistics."""
import numpy as np
import pytest

from tests.test_flt_oracle import flt_instance

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _gpu_phase_one(ga, I):
    t, s = I["t"], I["sched"]
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(s["net_ids"], s["node_off"], s["nodes"], s["parents"])
    ctx.lshape_run(s["batch_off"], first_stamp=1)
    return ctx


@pytest.mark.parametrize("config,seed,scale,side,mode", [(3, 7, 0.01, 64, 0), (3, 7, 0.01, 64, 1), (3, 7, 0.01, 64, 2),
                                                         (1, 0, 0.5, 0, 0), (1, 0, 0.5, 0, 2)])
def test_flt_stage_matches_oracle(ga, ora, config, seed, scale, side, mode):
    I = flt_instance(ga, ora, config, seed, scale, side)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    print(f"flt stage config={config} seed={seed} scale={scale} side={side} mode={mode} nets={fs['nets']} "
          f"nodes={fs['nodes']} batches={len(plan['batch_off']) - 1}", flush=True)
    ctx = _gpu_phase_one(ga, I)
    assert np.array_equal(ctx.routes(), I["routes"])
    flags = ctx.ripup(len(t["pin_acc_num"]) - 1)
    assert np.array_equal(flags, I["flags"])
    cong, xs, ys = ctx.congestion_view()
    assert np.array_equal(cong, I["view"][0])
    assert np.array_equal(_bits(xs), _bits(I["view"][1])) and np.array_equal(_bits(ys), _bits(I["view"][2]))
    color, nb = ctx.batch_plan(*I["geom"], limit=300000)
    assert nb == I["nb"] and np.array_equal(color, I["color"])
    first = len(I["sched"]["batch_off"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.flt_upload(plan)
    if mode == 2:
        # new phase-two storage after the upload: the run binds the current buffers
        ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ms = ctx.flt_run(first_stamp=first, remove_phase_one=True, mode=mode)
    r2, dem, ts, wl, vias, rc, _ = I["g"].flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"], I["demand"],
                                                     I["ts"], I["routes"], first, wl=I["wl"], vias=I["vias"])
    assert rc == 0
    assert np.array_equal(ctx.routes2(), r2)
    assert np.array_equal(_bits(ctx.demand()), _bits(dem))
    assert np.array_equal(ctx.timestamps(), ts)
    gwl, gvias = ctx.stats()
    assert gwl == wl and gvias == vias
    assert ms > 0
    ctx.close()


@pytest.mark.parametrize("per_wave", [(8, 8), (2, 4), None])
def test_flt_dags_per_wavefront_do_not_change_results(ga, ora, per_wave):
    """Several DAGs per wavefront (or the autotuner's choice) give the
    oracle's phase-two routes, demand and statistics."""
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    print(f"flt per wave {per_wave}", flush=True)
    ctx = _gpu_phase_one(ga, I)
    first = len(I["sched"]["batch_off"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.flt_upload(plan)
    if per_wave is None:
        chosen, ms = ctx.flt_tune()
        assert all(p in (1, 2, 4, 8) for p in chosen) and max(max(r) for r in ms) > 0
    else:
        ctx.set_flt_per_wave(per_wave)
    ctx.flt_run(first_stamp=first, remove_phase_one=True, mode=1)
    r2, dem, ts, wl, vias, rc, _ = I["g"].flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"], I["demand"],
                                                     I["ts"], I["routes"], first, wl=I["wl"], vias=I["vias"])
    assert rc == 0
    assert np.array_equal(ctx.routes2(), r2)
    assert np.array_equal(_bits(ctx.demand()), _bits(dem))
    assert np.array_equal(ctx.timestamps(), ts)
    assert ctx.stats() == (wl, vias)
    ctx.close()


def test_flt_stage_without_phase_one_removal(ga, ora):
    """remove_phase_one = 0 commits phase two on top of phase one, one stamp per batch."""
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    print(f"flt stage keep phase one nets={fs['nets']}", flush=True)
    ctx = _gpu_phase_one(ga, I)
    first = len(I["sched"]["batch_off"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.flt_upload(plan)
    ctx.flt_run(first_stamp=first, remove_phase_one=False)
    r2, dem, ts, wl, vias, rc, last = I["g"].flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"],
                                                        I["demand"], I["ts"], I["routes"], first,
                                                        remove_phase_one=False, wl=I["wl"], vias=I["vias"])
    assert rc == 0 and last == first + len(plan["batch_off"]) - 2
    assert np.array_equal(ctx.routes2(), r2)
    assert np.array_equal(_bits(ctx.demand()), _bits(dem))
    assert np.array_equal(ctx.timestamps(), ts)
    ctx.close()


def test_flt_small_route_storage_is_reported(ga, ora):
    """Phase-two storage of 8 words per pin overflows: the run returns
    GR_E_CAPACITY (the reference asserts, src/routing/flt_block.cuh:222)."""
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    fs, plan = I["fs"], I["plan"]
    print("flt small route storage", flush=True)
    ctx = _gpu_phase_one(ga, I)
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"], words_per_pin=8)
    ctx.flt_upload(plan)
    with pytest.raises(ga.GrError) as e:
        ctx.flt_run(first_stamp=len(I["sched"]["batch_off"]))
    assert e.value.code == ga.GR_E_CAPACITY
    ctx.close()


def test_flt_upload_rejects_malformed_dags(ga, ora):
    """Host validation of the compact records before any launch."""
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    fs, plan = I["fs"], I["plan"]
    print("flt malformed uploads", flush=True)
    ctx = ga.Context(I["t"])
    ctx.set_routes(I["t"]["pin_acc_num"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    k = 0
    n0, n1 = int(plan["node_off"][k]), int(plan["node_off"][k + 1])
    e0, e1 = int(plan["edge_off"][k]), int(plan["edge_off"][k + 1])
    one = dict(batch_off=np.array([0, 1]), net_ids=plan["net_ids"][:1],
               node_off=np.array([0, n1 - n0], np.int64), nodes=plan["nodes"][4 * n0:4 * n1].copy(),
               edge_off=np.array([0, e1 - e0], np.int64), edges=plan["edges"][2 * e0:2 * e1].copy())
    ctx.flt_upload(one)  # well formed
    bad = {k: v.copy() for k, v in one.items()}
    bad["nodes"][4 * 1 + 1] = 0  # a child at the root's depth
    with pytest.raises(ga.GrError):
        ctx.flt_upload(bad)
    bad = {k: v.copy() for k, v in one.items()}
    bad["edges"][1] += 10 * 3  # an edge record owned by another node
    with pytest.raises(ga.GrError):
        ctx.flt_upload(bad)
    bad = {k: v.copy() for k, v in one.items()}
    bad["nodes"][4 * 1] = I["t"]["X"] * I["t"]["Y"] * (I["t"]["L"] + 1)  # off the grid
    with pytest.raises(ga.GrError):
        ctx.flt_upload(bad)
    bad = {k: v.copy() for k, v in one.items()}
    bad["net_ids"] = np.array([int(np.flatnonzero(fs["pin_off2"] < 0)[0])])  # no phase-two storage
    with pytest.raises(ga.GrError):
        ctx.flt_upload(bad)
    ctx.close()


def test_flt_graph_mode_replays(ga, ora):
    """Mode 2 replays the captured FLT loop: a second stage on the same
    context (phase one rebuilt, storage replaced) and a replay without
    changes equal the oracle."""
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    s = I["sched"]
    first = len(s["batch_off"])
    r2, dem, ts, wl, vias, rc, _ = I["g"].flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"], I["demand"],
                                                     I["ts"], I["routes"], first, wl=I["wl"], vias=I["vias"])
    assert rc == 0
    print("flt graph replays", flush=True)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(s["net_ids"], s["node_off"], s["nodes"], s["parents"])
    for rep in range(3):
        if rep < 2:
            ctx.reset()
            ctx.lshape_run(s["batch_off"], first_stamp=1, mode=2)
            ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
            ctx.flt_upload(plan)
        else:
            # the same phase one and stage again: the captured loop is replayed
            ctx.reset()
            ctx.lshape_run(s["batch_off"], first_stamp=1, mode=2)
            ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
        if rep == 2:
            ctx.flt_upload(plan)
        ctx.flt_run(first_stamp=first, remove_phase_one=True, mode=2)
        assert np.array_equal(ctx.routes2(), r2), rep
        assert np.array_equal(_bits(ctx.demand()), _bits(dem)), rep
        assert np.array_equal(ctx.timestamps(), ts), rep
        assert ctx.stats() == (wl, vias), rep
    ctx.close()


def _crafted_flt_plan(X, Y, L):
    """Three hand-made DAGs: a star whose root has six child slots, a
    two-level star whose nodes have five (the general solve), and a chain
    of 2,600 nodes whose tables exceed a workgroup's LDS (the global-table
    kernel).  Node records {coordinate, depth, child slots, parent_end},
    edge records {parent, node * 10 + slot}."""
    XY = X * Y

    def c(layer, x, y):
        return layer * XY + x * Y + y

    dags = []
    # star: root pin with six leaves on both axes
    nodes = [[c(0, 20, 20), 0, 6, 0]]
    edges = []
    for i, (x, y) in enumerate([(25, 20), (14, 20), (20, 27), (20, 12), (31, 20), (20, 33)]):
        nodes.append([c(i % L, x, y), 1, 0, i + 1])
        edges.append([0, (i + 1) * 10 + i])
    dags.append((nodes, edges))
    # two levels of five child slots, Steiner nodes at the centres
    nodes = [[c(L, 40, 40), 0, 5, 0]]
    edges = []
    level1 = [(46, 40, L, 5), (34, 40, 0, 0), (40, 47, 1 % L, 0), (40, 33, 0, 0), (52, 40, 2 % L, 0)]
    for i, (x, y, layer, slots) in enumerate(level1):
        nodes.append([c(layer, x, y), 1, slots, i + 1])
        edges.append([0, (i + 1) * 10 + i])
    for i, (x, y) in enumerate([(46, 46), (46, 34), (50, 40), (46, 52), (46, 28)]):
        nodes.append([c(i % L, x, y), 2, 0, len(edges) + 1])
        edges.append([1, (len(nodes) - 1) * 10 + i])
    dags.append((nodes, edges))
    # a snake through the grid, one cell per node
    path = []
    for y in range(2, Y - 2):
        xs = range(2, X - 2) if (y % 2 == 0) else range(X - 3, 1, -1)
        path.extend((x, y) for x in xs)
    n = 2600
    assert len(path) >= n
    nodes, edges = [], []
    for i in range(n):
        x, y = path[i]
        pin = i == 0 or i == n - 1 or i % 13 == 0
        nodes.append([c((i // 13) % L if pin else L, x, y), i, 1 if i + 1 < n else 0, i])
        if i:
            edges.append([i - 1, i * 10])
    dags.append((nodes, edges))
    node_off, edge_off = [0], [0]
    for nd, ed in dags:
        node_off.append(node_off[-1] + len(nd))
        edge_off.append(edge_off[-1] + len(ed))
    return dict(batch_off=np.array([0, 3]), net_ids=np.array([0, 1, 2]),
                node_off=np.array(node_off, np.int64),
                nodes=np.array([r for nd, _ in dags for r in nd], np.int32).reshape(-1),
                edge_off=np.array(edge_off, np.int64),
                edges=np.array([r for _, ed in dags for r in ed], np.int32).reshape(-1))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_flt_wide_nodes_and_long_chains_match_oracle(ga, ora, mode):
    """Nodes with five and six child slots and a DAG too large for LDS,
    against the oracle, in both execution forms (each DAG's tables in LDS,
    the large one in global memory)."""
    from tests._instances import random_demand, topology
    _, t = topology(6, 0)
    L = t["L"]
    plan = _crafted_flt_plan(t["X"], t["Y"], L)
    pin_acc_num = np.array([0, 40, 52, 352], np.int32)  # route capacity: 100 words per pin
    pin_off2 = np.array([0, 40, 52], np.int32)
    demand = random_demand(t, 41)
    print(f"flt crafted DAGs mode={mode}", flush=True)
    ctx = ga.Context(t)
    ctx.set_routes(pin_acc_num)
    ctx.upload_demand(demand)
    ctx.set_routes_phase2(pin_off2, 352)
    ctx.flt_upload(plan)
    ctx.flt_run(first_stamp=1, remove_phase_one=False, mode=mode)
    g = ora.Grid(t)
    r2, dem, ts, wl, vias, rc, _ = g.flt_stage(plan, pin_acc_num, pin_off2, 352, demand, np.zeros(g.cells, np.int32),
                                               np.zeros(12 * 352 + 1, np.int32), 1, remove_phase_one=False)
    assert rc == 0
    assert np.array_equal(ctx.routes2(), r2)
    assert np.array_equal(_bits(ctx.demand()), _bits(dem))
    assert np.array_equal(ctx.timestamps(), ts)
    assert ctx.stats() == (wl, vias)
    ctx.close()


def test_reserve_phase_two_keeps_routed_words_at_a_wider_stride(ga, ora):
    """gr_ctx_reserve_routes_phase2 after an FLT stage: nets routed by the
    stage keep their offsets and words (copied to the wider stride), new
    nets are appended in the order given, with empty words; reserving nets
    that already have words changes nothing; the tuner leaves earlier
    stages' words in place."""
    I = flt_instance(ga, ora, 3, 7, 0.01, 64)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    print("reserve phase-two storage after an FLT stage", flush=True)
    pin_acc = np.asarray(t["pin_acc_num"], np.int64)
    ctx = _gpu_phase_one(ga, I)
    first = len(I["sched"]["batch_off"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.flt_upload(plan)
    ctx.flt_run(first_stamp=first, remove_phase_one=True, mode=1)
    before = ctx.routes2()
    ctx.flt_tune()
    assert np.array_equal(ctx.routes2(), before)
    off0 = np.asarray(fs["pin_off2"], np.int64)
    kept = np.flatnonzero(off0 >= 0)
    fresh = np.flatnonzero(off0 < 0)[::-3][:40]  # appended in this (decreasing) order
    ids = np.concatenate([fresh, kept[:5]])
    off, pins2 = ctx.reserve_routes_phase2(ids, words_per_pin=128)
    assert np.array_equal(off[kept], off0[kept])
    expect, p = off0.copy(), int(fs["pins2"])
    for i in fresh:
        expect[i] = p
        p += int(pin_acc[i + 1] - pin_acc[i])
    assert np.array_equal(off, expect) and pins2 == p
    after = ctx.routes2()
    assert len(after) == 128 * p + 1
    for i in kept:
        n = int(pin_acc[i + 1] - pin_acc[i]) * 100
        assert np.array_equal(after[128 * off[i]:128 * off[i] + n], before[100 * off0[i]:100 * off0[i] + n]), i
    for i in fresh:
        assert after[128 * off[i]] == 0
    off2, pins3 = ctx.reserve_routes_phase2(kept[:3], words_per_pin=100)
    assert np.array_equal(off2, off) and pins3 == p
    assert np.array_equal(ctx.routes2(), after)
    ctx.close()
