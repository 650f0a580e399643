"""GPU Z-shape stage (reset, solve, convert, commit per batch) against the
oracle's restatement on the same plan: route words, demand, timestamps and
statistics identical."""
import numpy as np
import pytest

from tests._instances import topology

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("config,seed,scale", [(6, 0, 1.0), (6, 5, 1.0), (1, 2, 0.25)])
def test_zshape_stage_matches_oracle(ga, ora, config, seed, scale):
    d, t = topology(config, seed, net_scale=scale)
    z = d.zshape_stage()
    print(f"zshape stage config={config} seed={seed} nets={z['nets']}", flush=True)
    geom = (z["pt_off"], z["pts"], z["hs_off"], z["hsegs"], z["vs_off"], z["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    plan = d.zshape_plan(color, nb)
    routes_o, demand_o, ts_o, wl_o, vias_o, rc = ora.Grid(t).zshape_stage(plan, t["pin_acc_num"], threads=8)
    assert rc == 0
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"], 256)
    ctx.set_cost_params(integer_wire=True)
    gcolor, gnb = ctx.batch_plan(*geom)
    assert gnb == nb and np.array_equal(gcolor, color)
    ctx.zshape_upload(plan)
    ms = ctx.zshape_run(first_stamp=1)
    assert ms > 0
    assert np.array_equal(ctx.routes(), routes_o)
    assert np.array_equal(_bits(ctx.demand()), _bits(demand_o))
    assert np.array_equal(ctx.timestamps(), ts_o)
    assert ctx.stats() == (wl_o, vias_o)
    # host post-processing on the GPU's routes gives the oracle's result
    r_gpu, r_o = ctx.routes(), routes_o.copy()
    assert d.zshape_finish(r_gpu) == d.zshape_finish(r_o)
    assert np.array_equal(r_gpu, r_o)
    ctx.close()
