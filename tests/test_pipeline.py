"""The rip-up stage pipeline command line (grotor_amd/grotor): flags and
errors on the CPU; on the GPU, the .out it writes for the golden instances
equals, byte for byte, the one the CPU oracle's pipeline writes, and the
reference evaluator scores it with no open or incomplete net."""
import os
import subprocess

import numpy as np
import pytest

from tests import _golden as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "grotor_amd", "grotor")


def _run(args, timeout=300):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("args,message", [
    (["-cap", "a"], "Usage"),
    (["-cap", "a", "-net", "b", "-out", "c", "--batching", "hiprt"], "ray-tracing"),
    (["-cap", "a", "-net", "b", "-out", "c", "--sparsity", "0"], "positive integer"),
    (["-cap", "a", "-net", "b", "-out", "c", "--stop-after", "maze"], "lshape or routing"),
    (["-cap", "a", "-net", "b", "-out", "c", "--report-score", "2"], "0 or 1"),
    (["-cap", "a", "-net", "b", "-out", "c", "--colour", "1"], "Unknown option"),
    (["-cap", "/nonexistent.cap", "-net", "/nonexistent.net", "-out", "/tmp/x.out"], "read input"),
])
def test_cli_rejects_bad_arguments(args, message):
    """Argument errors exit 1 with 'GRotor error: ...' as src/main.cu:168-178."""
    if not os.path.exists(CLI):
        pytest.skip("grotor not built")
    res = _run(args)
    assert res.returncode == 1
    assert res.stderr.startswith("GRotor error:") and message in res.stderr, res.stderr


def oracle_pipeline(ga, ora, cap, net, out):
    """The pipeline of src/pipeline.hpp on the CPU oracle: L-shape stage,
    congestion view, rip-up, FLT DAGs and the FLT stage, then .out."""
    d = ga.Design.read(cap, net)
    d.prepare(8)
    t = d.topology()
    g = ora.Grid(t)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    sched = ga.execution_schedule(t, color, nb)
    routes, demand, ts, wl, vias = g.lshape_stage(sched, t["pin_acc_num"], threads=4)
    cong, xs, ys = g.congestion_view(demand)
    flags = g.ripup(demand, routes, t["pin_acc_num"])
    fs = d.flt_stage(flags, xs, ys)
    if fs["nets"] == 0:
        d.write_output(routes, out)
        return wl, vias, 0
    fgeom = (fs["pt_off"], fs["pts"], fs["hs_off"], fs["hsegs"], fs["vs_off"], fs["vsegs"])
    fcol, fnb = ora.batches(t["X"], t["Y"], *fgeom, limit=300000)
    plan = d.flt_plan(fcol, fnb)
    r2, _, _, wl2, vias2, rc, _ = g.flt_stage(plan, t["pin_acc_num"], fs["pin_off2"], fs["pins2"], demand, ts, routes,
                                              nb + 1, wl=wl, vias=vias)
    assert rc == 0
    d.write_output(routes, out, r2, fs["pin_off2"])
    return wl2, vias2, fs["nets"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", G.NAMES)
def test_cli_pipeline_equals_oracle_pipeline(ga, ora, name, tmp_path):
    cap, net = G.unpacked(name, "cap"), G.unpacked(name, "net")
    mine, ref = str(tmp_path / "gpu.out"), str(tmp_path / "oracle.out")
    print(f"pipeline {name}", flush=True)
    res = _run(["-cap", cap, "-net", net, "-out", mine, "--profile-stages", "1"])
    assert res.returncode == 0, res.stderr[-2000:]
    wl, vias, rerouted = oracle_pipeline(ga, ora, cap, net, ref)
    assert rerouted > 0 or name == "mempool"  # the mempool window has no net over the rip-up threshold
    score = [line for line in res.stdout.splitlines() if line.startswith("SCORE")][0]
    assert f"wirelength={wl:.0f} vias={vias} rerouted={rerouted}" in score, score
    assert open(mine).read() == open(ref).read()
    if not os.path.exists(ora.evaluator_path()):
        pytest.skip("reference evaluator not built (oracle/_ref)")
    scores, ev = ora.run_evaluator(cap, net, mine)
    assert ev.returncode == 0, ev.stdout[-2000:]
    assert scores["open"] == 0 and scores["incomplete"] == 0


