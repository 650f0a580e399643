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


@pytest.mark.parametrize("config,seed,scale,side", [(3, 7, 0.01, 64), (1, 0, 0.5, 0)])
def test_flt_stage_matches_oracle(ga, ora, config, seed, scale, side):
    I = flt_instance(ga, ora, config, seed, scale, side)
    t, fs, plan = I["t"], I["fs"], I["plan"]
    print(f"flt stage config={config} seed={seed} scale={scale} side={side} nets={fs['nets']} "
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
    ms = ctx.flt_run(first_stamp=first, remove_phase_one=True)
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
    ctx.flt_run(first_stamp=first, remove_phase_one=True)
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
