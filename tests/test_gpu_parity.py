"""GPU path (libgrotor_amd.so on a HIP device) against the CPU oracle.

Integer and index results must be identical; floating-point results are
bit-identical by construction (same operations in the same order, no
contraction), so every comparison here is exact."""
import numpy as np
import pytest

from tests._instances import random_demand, topology

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("config,seed", [(1, 0), (6, 3)])
def test_refresh_matches_oracle(ga, ora, config, seed):
    _, topo = topology(config, seed)
    print(f"refresh config={config} seed={seed}", flush=True)
    ctx = ga.Context(topo)
    grid = ora.Grid(topo)
    for s in range(3):
        demand = random_demand(topo, seed * 10 + s)
        ctx.upload_demand(demand)
        ctx.refresh()
        w, v, p = ctx.costs()
        wo, vo, po = grid.refresh(demand)
        assert np.array_equal(_bits(w), _bits(wo))
        assert np.array_equal(_bits(v), _bits(vo))
        assert np.array_equal(_bits(p), _bits(po))
    ctx.close()


@pytest.mark.parametrize("config,seed", [(1, 0), (6, 1)])
def test_batch_plan_matches_oracle(ga, ora, config, seed):
    _, t = topology(config, seed)
    print(f"batch_plan config={config} seed={seed} nets={t['nets']}", flush=True)
    ctx = ga.Context(t)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ctx.batch_plan(*geom)
    ocolor, onb = ora.batches(t["X"], t["Y"], *geom)
    assert nb == onb
    assert np.array_equal(color, ocolor)
    for sparsity in (2, 5):
        color, nb = ctx.batch_plan(*geom, sparsity=sparsity)
        ocolor, onb = ora.batches(t["X"], t["Y"], *geom, sparsity=sparsity)
        assert nb == onb and np.array_equal(color, ocolor)
    ctx.close()


@pytest.mark.parametrize("limit", [40, 150, 600])
def test_batch_limit_matches_oracle(ga, ora, limit):
    """Batches that fill up close for later nets (generate_batches_cpu.hpp:50)."""
    _, t = topology(1, 0)
    print(f"batch limit={limit} nets={t['nets']}", flush=True)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    ocolor, onb = ora.batches(t["X"], t["Y"], *geom, limit=limit)
    ctx = ga.Context(t)
    ctx.batch_geometry(*geom)
    color, nb = ctx.batch_colors(limit=limit)
    assert nb == onb and np.array_equal(color, ocolor)
    assert np.bincount(color).max() <= limit
    color2, nb2 = ctx.batch_colors()  # same resident geometry, no limit
    ucolor, unb = ora.batches(t["X"], t["Y"], *geom)
    assert nb2 == unb and np.array_equal(color2, ucolor)
    ctx.close()


def test_batch_edge_cases(ga, ora):
    """Nets without points or segments, one net, and wide color masks."""
    _, t = topology(6, 0)
    ctx = ga.Context(t)
    n = 5
    empty = np.zeros(n + 1, np.int32)
    color, nb = ctx.batch_plan(empty, np.zeros(0, np.int32), empty, np.zeros(0, np.int32), empty,
                               np.zeros(0, np.int32))
    assert nb == 1 and np.array_equal(color, np.zeros(n, np.int32))
    # every net on the same cell: one batch per net, beyond one mask word
    k = 150
    off = np.arange(k + 1, dtype=np.int32)
    pts = np.full(k, t["Y"] * 3 + 4, np.int32)
    color, nb = ctx.batch_plan(off, pts, np.zeros(k + 1, np.int32), np.zeros(0, np.int32),
                               np.zeros(k + 1, np.int32), np.zeros(0, np.int32))
    assert nb == k and np.array_equal(color, np.arange(k, dtype=np.int32))
    with pytest.raises(ga.GrError):
        ctx.batch_plan(off, np.full(k, t["X"] * t["Y"], np.int32), np.zeros(k + 1, np.int32),
                       np.zeros(0, np.int32), np.zeros(k + 1, np.int32), np.zeros(0, np.int32))
    ctx.close()


@pytest.mark.parametrize("config,seed", [(1, 0), (6, 2)])
def test_lshape_stage_matches_oracle(ga, ora, config, seed):
    _, t = topology(config, seed)
    print(f"lshape stage config={config} seed={seed}", flush=True)
    color, nb = ora.batches(t["X"], t["Y"], t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    sched = ga.execution_schedule(t, color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, t["pin_acc_num"], threads=8)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ms = ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=0)
    assert ms > 0
    routes = ctx.routes()
    assert np.array_equal(routes, routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert np.array_equal(ctx.timestamps(), ts_o)
    wl, vias = ctx.stats()
    assert wl == wl_o and vias == vias_o
    flags = ctx.ripup(t["subnets"], 0)
    flags_o = ora.Grid(t).ripup(demand_o, routes_o, t["pin_acc_num"], 0)
    assert np.array_equal(flags, flags_o)
    ctx.close()


@pytest.mark.parametrize("per_wave", [(1, 1), (8, 8), (4, 2)])
def test_nets_per_wavefront_do_not_change_results(ga, ora, per_wave):
    """Any nets-per-wavefront setting of the two tree size classes gives the
    oracle's routes, demand, timestamps and statistics; the autotuner picks
    a valid setting and the stage still matches after it ran."""
    _, t = topology(6, 2)
    print(f"nets per wavefront {per_wave}", flush=True)
    color, nb = ora.batches(t["X"], t["Y"], t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    sched = ga.execution_schedule(t, color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, t["pin_acc_num"], threads=8)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.set_nets_per_wave(per_wave)
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=1)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert np.array_equal(ctx.timestamps(), ts_o)
    assert ctx.stats() == (wl_o, vias_o)
    if per_wave == (4, 2):
        chosen, ms = ctx.lshape_tune(sched["batch_off"])
        assert all(p in (1, 2, 4, 8) for p in chosen)
        assert any(v > 0 for row in ms for v in row)
        ctx.reset()
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
        assert np.array_equal(ctx.routes(), routes_o)
        assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    with pytest.raises(ga.GrError):
        ctx.set_nets_per_wave((3, 1))
    ctx.close()


def test_graph_mode_replays_match_oracle(ga, ora):
    """Mode 2 (captured HIP graph): every replay, with resets and eager work
    between replays, equals the oracle; a changed schedule recaptures."""
    _, t = topology(6, 5)
    print("graph mode config=6 seed=5", flush=True)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    sched = ga.execution_schedule(t, color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, t["pin_acc_num"], threads=8)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    for rep in range(3):
        ctx.reset()
        ctx.batch_plan(*geom)  # eager device work between replays
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
        assert np.array_equal(ctx.routes(), routes_o), rep
        assert np.array_equal(_bits(ctx.demand()), _bits(demand_o)), rep
        assert np.array_equal(ctx.timestamps(), ts_o), rep
        assert ctx.stats() == (wl_o, vias_o), rep
    # the same stage uploaded again (fresh device buffers) with the same
    # schedule: the slot permutation must be uploaded again, not assumed
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.reset()
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    # a different schedule (first half of the batches only)
    half = dict(sched)
    half["batch_off"] = sched["batch_off"][:len(sched["batch_off"]) // 2 + 1]
    routes_h, demand_h, _, _, _ = ora.Grid(t).lshape_stage(half, t["pin_acc_num"], threads=8)
    ctx.reset()
    ctx.lshape_run(half["batch_off"], first_stamp=1, mode=2)
    assert np.array_equal(ctx.routes(), routes_h)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_h))
    with pytest.raises(ga.GrError):
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=3)
    ctx.close()


@pytest.mark.parametrize("per_wave,tiny", [((1, 1), (4, 0)), ((4, 1), (4, 0)), ((1, 1), (8, 4)), ((2, 1), (2, 8)),
                                           ((2, 1), (8, 2))])
def test_persistent_mode_matches_oracle(ga, ora, per_wave, tiny):
    """Mode 2 (one persistent DP launch per batch draining a work queue,
    the loop captured in a HIP graph): replays and a switch back and forth
    with mode 1 equal the oracle, for several tiny-tree groups (bound, nets
    per wavefront)."""
    _, t = topology(6, 5)
    print(f"persistent mode config=6 seed=5 per_wave={per_wave} tiny={tiny}", flush=True)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    sched = ga.execution_schedule(t, color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, t["pin_acc_num"], threads=8)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.set_nets_per_wave(per_wave)
    ctx.set_tiny_tree(tiny[0])
    ctx.set_tiny_per_wave(tiny[1])
    for rep, mode in enumerate((2, 2, 1, 0, 2)):
        ctx.reset()
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=mode)
        assert np.array_equal(ctx.routes(), routes_o), (rep, mode)
        assert np.array_equal(_bits(ctx.demand()), _bits(demand_o)), (rep, mode)
        assert np.array_equal(ctx.timestamps(), ts_o), (rep, mode)
        assert ctx.stats() == (wl_o, vias_o), (rep, mode)
    ctx.close()


def _large_tree_stage(t, sizes, seed):
    """A stage of random trees (pins as leaves on layer 0, Steiner points
    inside), some too large for one workgroup's LDS tables."""
    rng = np.random.default_rng(seed)
    X, Y, L = t["X"], t["Y"], t["L"]
    nodes, parents, node_off = [], [], [0]
    for n in sizes:
        par = np.array([-1] + [int(rng.integers(0, i)) for i in range(1, n)], np.int32)
        leaf = np.ones(n, bool)
        leaf[par[1:]] = False
        leaf[0] = True  # the root is a pin
        x = rng.integers(0, X, n)
        y = rng.integers(0, Y, n)
        lay = np.where(leaf, 0, L)
        nodes.append(((lay * X + x) * Y + y).astype(np.int32))
        parents.append(par)
        node_off.append(node_off[-1] + n)
    return dict(net_ids=np.arange(len(sizes), dtype=np.int32), node_off=np.array(node_off, np.int32),
                nodes=np.concatenate(nodes), parents=np.concatenate(parents))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_trees_beyond_lds_match_oracle(ga, ora, mode):
    """Trees whose DP tables do not fit in LDS (300 and 240 nodes on four
    layers) take the global-scratch kernel, in the launch and persistent
    forms alike, and equal the oracle."""
    _, t = topology(6, 2)
    print(f"large trees mode={mode}", flush=True)
    sched = _large_tree_stage(t, [300, 12, 240, 60, 3, 97], 11)
    sched["batch_off"] = np.array([0, 3, 6], np.int32)
    pin_acc_num = np.arange(0, 250 * 7, 250, dtype=np.int32)  # 3,000 route words per net
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, pin_acc_num, threads=8)
    ctx = ga.Context(t)
    ctx.set_routes(pin_acc_num)
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=mode)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert np.array_equal(ctx.timestamps(), ts_o)
    assert ctx.stats() == (wl_o, vias_o)
    ctx.close()


def test_single_batch_entry_points(ga, ora):
    """gr_lshape_batch + gr_commit_batch over one batch equal the oracle."""
    _, t = topology(6, 4)
    color, nb = ora.batches(t["X"], t["Y"], t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    sched = ga.execution_schedule(t, color, nb)
    bo = sched["batch_off"]
    first = {k: sched[k] for k in sched}
    first["batch_off"] = bo[:2]
    routes_o, demand_o, _, _, _ = ora.Grid(t).lshape_stage(first, t["pin_acc_num"])
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.refresh()
    ctx.lshape_batch(0, int(bo[1]), 1)
    ctx.commit_batch(0, int(bo[1]), 1)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    ctx.close()


@pytest.mark.parametrize("name", ["syn10", "mempool"])
def test_golden_batches_and_stage_on_gpu(ga, ora, name):
    """On the reference's own trees and geometry: GPU batching equals the
    reference's CPU batch lists, and the GPU L-shape stage equals the oracle."""
    from tests import _golden as G
    t = G.topology(name)
    print(f"golden {name}: {t['nets']} nets", flush=True)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    color, nb = ctx.batch_plan(t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    ref_color, ref_nb = G.reference_colors(name)
    assert nb == ref_nb and np.array_equal(color, ref_color)
    sched = ga.execution_schedule(t, color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, t["pin_acc_num"], threads=4)
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.lshape_run(sched["batch_off"], first_stamp=1)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert ctx.stats() == (wl_o, vias_o)
    ctx.close()


def test_repeat_runs_are_deterministic(ga):
    """Two full stages from a reset context give identical routes and demand
    (atomics only ever add exact binary32 multiples of 0.5)."""
    _, t = topology(1, 5)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    outs = []
    for _ in range(2):
        ctx.reset()
        color, nb = ctx.batch_plan(*geom)
        sched = ga.execution_schedule(t, color, nb)
        ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
        ctx.lshape_run(sched["batch_off"], first_stamp=1)
        outs.append((color.copy(), ctx.routes(), ctx.demand()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(_bits(outs[0][2]), _bits(outs[1][2]))
    ctx.close()


def test_gpu_expf_matches_oracle(ga, ora):
    """The refresh kernel's exponential equals the oracle's on every demand
    value the cost model sees: demand sweeps over 0..200 in steps of 0.5
    against several capacities, compared through the resulting vcost."""
    _, t = topology(6, 0)
    ctx = ga.Context(t)
    grid = ora.Grid(t)
    rng = np.random.default_rng(11)
    for s in range(4):
        demand = (rng.integers(0, 400, ctx.cells) * 0.5).astype(np.float32)
        ctx.upload_demand(demand)
        ctx.refresh()
        w, v, p = ctx.costs()
        wo, vo, po = grid.refresh(demand)
        assert np.array_equal(_bits(w), _bits(wo)) and np.array_equal(_bits(v), _bits(vo))
    ctx.close()


@pytest.mark.parametrize("slope,zero_slope", [(0.5, 1.5), (0.37, 2.2)])
def test_gpu_refresh_wide_arguments_match_oracle(ga, ora, slope, zero_slope):
    """The refresh's table and step exponentials over arguments across the
    whole [-88, 89] range and beyond: capacities zero, fractional and up to
    400, demands arbitrary binary32 values and exact halves; the default
    slopes take the step path, the others (steps not 1/4, 1/2, 3/4, 3/2)
    the table path for all three demands."""
    _, t = topology(6, 0)
    t = dict(t)
    rng = np.random.default_rng(23)
    cells = int(t["L"]) * int(t["X"]) * int(t["Y"])
    cap = rng.uniform(0.0, 400.0, cells).astype(np.float32)
    cap[rng.random(cells) < 0.25] = 0.0
    ints = rng.random(cells) < 0.25
    cap[ints] = np.round(cap[ints])
    t["capacity"] = cap
    print(f"refresh wide slope={slope} zero_slope={zero_slope}", flush=True)
    ctx = ga.Context(t)
    ctx.set_cost_params(slope=slope, zero_slope=zero_slope)
    grid = ora.Grid(t)
    for s in range(3):
        demand = rng.uniform(0.0, 420.0, cells).astype(np.float32)
        halves = rng.random(cells) < 0.5
        demand[halves] = np.round(demand[halves] * 2.0) * 0.5
        ctx.upload_demand(demand)
        ctx.refresh(full=True)
        w, v, p = ctx.costs()
        wo, vo, po = grid.refresh(demand, slope=slope, zero_slope=zero_slope)
        assert np.array_equal(_bits(w), _bits(wo)), s
        assert np.array_equal(_bits(v), _bits(vo)), s
        assert np.array_equal(_bits(p), _bits(po)), s
    ctx.close()


def test_graph_replay_after_storage_changes(ga, ora):
    """A captured batch loop is dropped when the buffers it holds are
    reallocated: replays after gr_ctx_set_routes, after a single-batch DP
    over a wider range (global scratch growth) and after a stride change
    (which waits for the next gr_ctx_set_routes) equal the oracle."""
    _, t = topology(6, 5)
    print("graph replay after storage changes config=6 seed=5", flush=True)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    sched = ga.execution_schedule(t, color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o = ora.Grid(t).lshape_stage(sched, t["pin_acc_num"], threads=8)

    def check(ctx, tag):
        assert np.array_equal(ctx.routes(), routes_o), tag
        assert np.array_equal(_bits(ctx.demand()), _bits(demand_o)), tag
        assert np.array_equal(ctx.timestamps(), ts_o), tag
        assert ctx.stats() == (wl_o, vias_o), tag

    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    for mode in (2, 1):
        ctx.reset()
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=mode)
        check(ctx, f"first mode {mode}")
        ctx.set_routes(t["pin_acc_num"])  # new route storage
        ctx.reset()
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=mode)
        check(ctx, f"after set_routes mode {mode}")
        ctx.reset()
        ctx.refresh()
        ctx.lshape_batch(0, int(sched["batch_off"][-1]), 1)  # every net at once: scratch may grow
        ctx.reset()
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=mode)
        check(ctx, f"after wide lshape_batch mode {mode}")
    # a stride change applies to the next route storage only
    ga._call("gr_ctx_set_route_stride", ctx._h, ga.ctypes.c_int(256))
    ctx.reset()
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
    check(ctx, "stride pending")
    ctx.close()
