"""The rip-up stage pipeline command line (grotor_amd/grotor): flags and
errors on the CPU; on the GPU, the .out it writes for the golden instances
equals, byte for byte, the one the CPU oracle's pipeline writes, with host
(`--batching cpu`) and GPU (`--batching hiprt --validate-batches 1`)
batching, and its run log carries the reference's records in the formats
the reference prints them (src/batching/generate_batches.hpp:21-23,
src/batching/hiprt/grid_chunk.cuh:549-553, include/router_state.hpp:259-264,
src/main.cu:126-166, src/pipeline.hpp:44,154), checked the way the
reference's experiment script reads them (script/experiments/run_amd.py:80-113)."""
import os
import re
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
    (["-cap", "a", "-net", "b", "-out", "c", "--batching", "rtcore"], "HIP uses --batching hiprt"),
    (["-cap", "a", "-net", "b", "-out", "c", "--batching", "optix"], "cpu, rtcore or hiprt"),
    (["-cap", "a", "-net", "b", "-out", "c", "--validate-batches", "1"], "requires ray-tracing batching"),
    (["-cap", "a", "-net", "b", "-out", "c", "--batching", "hiprt", "--sparsity", "2"], "--sparsity 1"),
    (["-cap", "a", "-net", "b", "-out", "c", "--dag-mode", "3"], "--dag-mode"),
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


BATCH_RE = re.compile(r"^BATCH_TIMING backend=(\w+) start_us=(\d+) end_us=(\d+) elapsed_ms=([\d.]+) nets=(\d+) "
                      r"batches=(\d+) used_for_routing=1$", re.M)
DAG_RE = re.compile(r"^DAG_TIMING stage=(\w+) batch_index=-1 start_us=(\d+) end_us=(\d+) elapsed_ms=([\d.]+) "
                    r"nets=(\d+) batches=(\d+)$", re.M)
COMPLETE_RE = re.compile(r"^GROTOR_COMPLETE wall_seconds=([\d.]+)$", re.M)
POLICY_RE = re.compile(r"^HOST_MEMORY_POLICY saving=([01]) bounded_preparation=([01])$", re.M)
SCOPE_RE = re.compile(r"^BATCH_SCOPE backend=hiprt scope=generate_batch_ids_call_wall$", re.M)
VALID_RE = re.compile(r"^BATCH_VALIDATION stage=(\w+) nets=(\d+) hiprt_batches=(\d+) cpu_batches=(\d+) "
                      r"exact_lists_equal=([01])$", re.M)
PIPE_RE = re.compile(r"^PIPELINE_TIMING setup_ms=[\d.]+ lshape_ms=[\d.]+ flt_ms=[\d.]+ download_ms=[\d.]+ "
                     r"classification_ms=[\d.]+ output_ms=[\d.]+ output_wait_ms=[\d.]+ cleanup_ms=[\d.]+ "
                     r"topology_cleanup_ms=[\d.]+$", re.M)


def check_run_log(text, backend, validated):
    """The two-stage timing records of a run, read as the reference's
    experiment script reads them: one BATCH_TIMING per stage with the
    expected backend, DAG_TIMING for lshape then detour with the same nets
    and batches, consistent microsecond stamps, each DAG interval after its
    batching interval, one completion and one memory-policy record, the
    declared configuration and scheduling.  Returns per-stage rows."""
    batches, dags = BATCH_RE.findall(text), DAG_RE.findall(text)
    assert "SCORE_DIAGNOSTICS enabled=0" in text
    assert len(batches) == 2 and [b[0] for b in batches] == [backend] * 2, batches
    assert [d[0] for d in dags] == ["lshape", "detour"], dags
    assert len(COMPLETE_RE.findall(text)) == 1 and len(POLICY_RE.findall(text)) == 1
    assert "cost=incremental prefix=bit" in text and "lshape=warp flt=block" in text
    assert f"GROTOR_CONFIG batching={backend} " in text
    assert len(PIPE_RE.findall(text)) == 1
    if backend == "hiprt":
        assert len(SCOPE_RE.findall(text)) == 2
    valid = VALID_RE.findall(text)
    if validated:
        assert [v[0] for v in valid] == ["lshape", "flt"] and all(v[4] == "1" for v in valid), valid
        for v, b in zip(valid, batches):
            assert v[1] == b[4] and v[2] == v[3] == b[5]
    else:
        assert not valid
    rows = []
    for b, d in zip(batches, dags):
        for item in (b, d):
            start, end, ms = int(item[1]), int(item[2]), float(item[3])
            assert end >= start and abs((end - start) / 1000 - ms) <= 1e-5, item
        assert b[4:] == d[4:], (b, d)
        assert int(d[1]) >= int(b[2]), (b, d)
        rows.append({"nets": int(b[4]), "batches": int(b[5]), "batching_ms": float(b[3]), "dag_ms": float(d[3])})
    return rows


@pytest.mark.gpu
@pytest.mark.parametrize("name", G.NAMES)
def test_cli_pipeline_equals_oracle_pipeline(ga, ora, name, tmp_path):
    cap, net = G.unpacked(name, "cap"), G.unpacked(name, "net")
    ref = str(tmp_path / "oracle.out")
    print(f"pipeline {name}", flush=True)
    wl, vias, rerouted = oracle_pipeline(ga, ora, cap, net, ref)
    assert rerouted > 0 or name == "mempool"  # the mempool window has no net over the rip-up threshold
    for backend, extra in (("cpu", []), ("hiprt", ["--validate-batches", "1"]), ("hiprt", ["--dag-mode", "0"])):
        mine = str(tmp_path / f"gpu_{backend}_{len(extra)}.out")
        print(f"pipeline {name} --batching {backend} {' '.join(extra)}", flush=True)
        res = _run(["-cap", cap, "-net", net, "-out", mine, "--batching", backend, "--profile-stages", "1"] + extra)
        assert res.returncode == 0, res.stderr[-2000:]
        assert open(mine).read() == open(ref).read()
        res_s = _run(["-cap", cap, "-net", net, "-out", mine, "--batching", backend, "--report-score", "1"] + extra)
        assert res_s.returncode == 0, res_s.stderr[-2000:]
        score = [line for line in res_s.stdout.splitlines() if line.startswith("SCORE wirelength=")][0]
        assert f"wirelength={wl:.0f} vias={vias} rerouted={rerouted}" in score, score
        if rerouted == 0:
            continue
        res = _run(["-cap", cap, "-net", net, "-out", mine, "--batching", backend, "--profile-stages", "1",
                    "--report-score", "0"] + extra)
        assert res.returncode == 0, res.stderr[-2000:]
        rows = check_run_log(res.stdout, backend, "--validate-batches" in extra)
        assert rows[1]["nets"] == rerouted
