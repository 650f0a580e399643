"""The HIP path at the full sizes of the benchmark configurations, against
the oracle on the same seeded instances.

Configuration 2 (512x512x9, 300k nets), 3 (1024x1024x10, 1M nets) and 5
(2048x2048x10, 4M nets) have routing tracks of 512, 1024 and 2048 cells, so
the refresh kernel's workgroup covers several cells per thread (the
cross-thread via-cost hand-off), and the DP, FLT and Z-shape kernels run
their 8- and 9-layer instantiations.  Every comparison is exact: colors,
route words, demand and cost bits, timestamps and statistics."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 8)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _geom(t):
    return (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])


@pytest.fixture(scope="module", params=[2, 3, 5])
def full(request, ga, ora):
    """One full-size instance per configuration (seed 0), its oracle
    batches and the oracle's L-shape stage on them, built once per module."""
    cfg = request.param
    print(f"\n[full] generating configuration {cfg}", flush=True)
    d = ga.Design.generate(cfg, seed=0)
    d.prepare(THREADS)
    t = d.topology()
    g = ora.Grid(t)
    color, nb = ora.batches(t["X"], t["Y"], *_geom(t))
    sched = ga.execution_schedule(t, color, nb)
    print(f"[full] configuration {cfg}: {t['nets']} subnets, {nb} batches, grid {t['X']}x{t['Y']}x{t['L']}",
          flush=True)
    state = dict(cfg=cfg, d=d, t=t, g=g, color=color, nb=nb, sched=sched, stage=None)
    yield state
    state.clear()


def _oracle_stage(full):
    if full["stage"] is None:
        full["stage"] = full["g"].lshape_stage(full["sched"], full["t"]["pin_acc_num"], threads=THREADS)
    return full["stage"]


@pytest.mark.timeout(300)
def test_full_size_refresh_matches_oracle(ga, ora, full):
    """Incremental and full refresh after uploads of random demand: wire
    cost, via cost and Fenwick prefix bits equal the oracle's on every
    cell of tracks 512-2048 cells long."""
    t, g = full["t"], full["g"]
    print(f"full refresh configuration {full['cfg']}", flush=True)
    rng = np.random.default_rng(100 + full["cfg"])
    cells = g.cells
    ctx = ga.Context(t)
    for s, full_refresh in enumerate((False, True)):
        if s == 0:
            demand = (rng.integers(0, 48, cells) * 0.5).astype(np.float32)
        else:
            demand = rng.uniform(0.0, 40.0, cells).astype(np.float32)
        ctx.upload_demand(demand)
        ctx.refresh(full=full_refresh)
        w, v, p = ctx.costs()
        wo, vo, po = g.refresh(demand)
        assert np.array_equal(_bits(w), _bits(wo)), s
        assert np.array_equal(_bits(v), _bits(vo)), s
        assert np.array_equal(_bits(p), _bits(po)), s
    ctx.close()


@pytest.mark.timeout(300)
def test_full_size_batching_matches_oracle(ga, ora, full):
    """GPU first-fit colours of every subnet equal the oracle's greedy."""
    t = full["t"]
    print(f"full batching configuration {full['cfg']}", flush=True)
    ctx = ga.Context(t)
    color, nb = ctx.batch_plan(*_geom(t))
    assert nb == full["nb"]
    assert np.array_equal(color, full["color"])
    ctx.close()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("mode", [2, 1, 0])
def test_full_size_lshape_stage_matches_oracle(ga, ora, full, mode):
    """The whole L-shape stage (refresh, DP, commit per batch) in the
    persistent captured form and in plain launches: route words, demand
    bits, timestamps, statistics and rip-up flags equal the oracle's."""
    t = full["t"]
    print(f"full lshape stage configuration {full['cfg']} mode {mode}", flush=True)
    routes_o, demand_o, ts_o, wl_o, vias_o = _oracle_stage(full)
    s = full["sched"]
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(s["net_ids"], s["node_off"], s["nodes"], s["parents"])
    ctx.lshape_run(s["batch_off"], first_stamp=1, mode=mode)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert np.array_equal(ctx.timestamps(), ts_o)
    assert ctx.stats() == (wl_o, vias_o)
    if mode == 2:
        flags = ctx.ripup(len(t["pin_acc_num"]) - 1, 0)
        flags_o = full["g"].ripup(demand_o, routes_o, t["pin_acc_num"], 0)
        assert np.array_equal(flags, flags_o)
    ctx.close()


@pytest.mark.timeout(300)
def test_full_size_flt_stage_matches_oracle(ga, ora, full):
    """Configuration 3's FLT rip-up stage at full size: rip-up flags,
    congestion view, the FLT batches of the rerouted nets and the FLT batch
    loop (phase-one removal, layer-change search, DAG DP, phase-two commit)
    equal the oracle's."""
    if full["cfg"] != 3:
        pytest.skip("the FLT stage is configuration 3's workload")
    t, g, d = full["t"], full["g"], full["d"]
    print("full flt stage configuration 3", flush=True)
    routes_o, demand_o, ts_o, wl_o, vias_o = _oracle_stage(full)
    flags_o = g.ripup(demand_o, routes_o, t["pin_acc_num"])
    cong_o, xs_o, ys_o = g.congestion_view(demand_o)
    fs = d.flt_stage(flags_o, xs_o, ys_o, threads=THREADS)
    fgeom = _geom(fs)
    fcol, fnb = ora.batches(t["X"], t["Y"], *fgeom, limit=300000)
    plan = d.flt_plan(fcol, fnb)
    first = len(full["sched"]["batch_off"])
    s = full["sched"]
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(s["net_ids"], s["node_off"], s["nodes"], s["parents"])
    ctx.lshape_run(s["batch_off"], first_stamp=1, mode=2)
    flags = ctx.ripup(len(t["pin_acc_num"]) - 1)
    assert np.array_equal(flags, flags_o)
    cong, xs, ys = ctx.congestion_view()
    assert np.array_equal(cong, cong_o)
    assert np.array_equal(_bits(xs), _bits(xs_o)) and np.array_equal(_bits(ys), _bits(ys_o))
    color, nb = ctx.batch_plan(*fgeom, limit=300000)
    assert nb == fnb and np.array_equal(color, fcol)
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.flt_upload(plan)
    ctx.flt_run(first_stamp=first, remove_phase_one=True, mode=2)
    r2, dem, ts, wl, vias, rc, _ = g.flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"], demand_o, ts_o,
                                               routes_o, first, wl=wl_o, vias=vias_o, threads=THREADS)
    assert rc == 0
    assert np.array_equal(ctx.routes2(), r2)
    assert np.array_equal(_bits(ctx.demand()), _bits(dem))
    assert np.array_equal(ctx.timestamps(), ts)
    assert ctx.stats() == (wl, vias)
    ctx.close()


@pytest.mark.timeout(300)
def test_full_grid_zshape_stage_matches_oracle(ga, ora):
    """The Z-shape stage on configuration 2's full 512x512x9 grid (a quarter
    of its nets, so the oracle's sequential loop stays short): the 8-layer
    solve, route words, demand, timestamps and statistics equal the
    oracle's."""
    d = ga.Design.generate(2, seed=0, net_scale=0.25)
    d.prepare(THREADS)
    t = d.topology()
    z = d.zshape_stage()
    geom = _geom(z)
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    plan = d.zshape_plan(color, nb)
    print(f"full-grid zshape stage: {z['nets']} nets, {len(plan['batch_off']) - 1} batches", flush=True)
    routes_o, demand_o, ts_o, wl_o, vias_o, rc = ora.Grid(t).zshape_stage(plan, t["pin_acc_num"], threads=THREADS)
    assert rc == 0
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"], 256)
    ctx.set_cost_params(integer_wire=True)
    gcolor, gnb = ctx.batch_plan(*geom)
    assert gnb == nb and np.array_equal(gcolor, color)
    ctx.zshape_upload(plan)
    ctx.zshape_run(first_stamp=1)
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert np.array_equal(ctx.timestamps(), ts_o)
    assert ctx.stats() == (wl_o, vias_o)
    ctx.close()


@pytest.mark.timeout(600)
def test_full_size_repeated_ripup_stages_match_oracle(ga, ora, full):
    """Three rip-up -> FLT stages (configuration 5's pipeline), at
    configurations 3 and 5: per stage the rip-up flags (phase-two routes of
    nets an earlier stage rerouted), congestion view, FLT batches, appended
    phase-two storage and the stage itself (each net's current route
    removed) equal the oracle's; the stages run in the persistent, launch
    and plain forms in turn."""
    from tests import _ripup as R
    if full["cfg"] == 2:
        pytest.skip("configuration 2 is the Z-shape workload")
    t, g, d, s = full["t"], full["g"], full["d"], full["sched"]
    print(f"full repeated rip-up stages configuration {full['cfg']}", flush=True)
    routes_o, demand_o, ts_o, wl_o, vias_o = _oracle_stage(full)
    nets = len(t["pin_acc_num"]) - 1
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(s["net_ids"], s["node_off"], s["nodes"], s["parents"])
    ctx.lshape_run(s["batch_off"], first_stamp=1, mode=2)
    done = []

    def on_stage(o):
        k, fs = o["stage"], o["fs"]
        print(f"full repeated rip-up configuration {full['cfg']} stage {k}: {fs['nets']} nets", flush=True)
        flags = ctx.ripup(nets)
        assert np.array_equal(flags, o["flags"]), k
        cong, xs, ys = ctx.congestion_view()
        assert np.array_equal(cong, o["cong"])
        assert np.array_equal(_bits(xs), _bits(o["xs"])) and np.array_equal(_bits(ys), _bits(o["ys"]))
        color, nb = ctx.batch_plan(*_geom(fs), limit=300000)
        assert nb == o["nb"] and np.array_equal(color, o["color"])
        if k == 1:
            ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
        else:
            off, pins2 = ctx.reserve_routes_phase2(np.flatnonzero(flags))
            assert np.array_equal(off, o["off"]) and pins2 == o["pins2"]
        ctx.flt_upload(o["plan"])
        ctx.flt_run(first_stamp=o["first"], remove=1 if k == 1 else 2, mode={1: 2, 2: 1, 3: 0}[k])
        assert np.array_equal(ctx.routes2(), o["r2"]), k
        assert np.array_equal(_bits(ctx.demand()), _bits(o["demand"])), k
        assert np.array_equal(ctx.timestamps(), o["ts"]), k
        assert ctx.stats() == (o["wl"], o["vias"]), k
        done.append(k)

    st = R.oracle_stages(ora, d, t, g, full["nb"], routes_o, demand_o, ts_o, wl_o, vias_o, 3, threads=THREADS,
                         on_stage=on_stage)
    assert done == [1, 2, 3] and min(st["nets"]) > 0
    ctx.close()
