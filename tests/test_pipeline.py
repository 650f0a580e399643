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
from tests import _ripup as R

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


def oracle_pipeline(ga, ora, cap, net, out, stages=1, counts=None):
    """The pipeline of src/pipeline.hpp on the CPU oracle: L-shape stage,
    congestion view, rip-up, FLT DAGs and the FLT stage, then .out; with
    stages > 1 the rip-up -> FLT stage repeats (tests/_ripup.py).  Returns
    (wirelength, vias, nets holding phase-two routes); counts, when a list,
    receives each stage's nets."""
    d = ga.Design.read(cap, net)
    d.prepare(8)
    t = d.topology()
    g = ora.Grid(t)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    sched = ga.execution_schedule(t, color, nb)
    routes, demand, ts, wl, vias = g.lshape_stage(sched, t["pin_acc_num"], threads=4)
    st = R.oracle_stages(ora, d, t, g, nb, routes, demand, ts, wl, vias, stages)
    if counts is not None:
        counts.extend(st["nets"])
    if st["off"] is None:
        d.write_output(routes, out)
        return st["wl"], st["vias"], 0
    d.write_output(routes, out, st["r2"], st["off"])
    return st["wl"], st["vias"], int(np.count_nonzero(st["off"] >= 0))


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
        assert "SCORE_CONNECTIVITY open=0 not_connected=0 incomplete=0" in res_s.stdout
        if rerouted == 0:
            continue
        res = _run(["-cap", cap, "-net", net, "-out", mine, "--batching", backend, "--profile-stages", "1",
                    "--report-score", "0"] + extra)
        assert res.returncode == 0, res.stderr[-2000:]
        rows = check_run_log(res.stdout, backend, "--validate-batches" in extra)
        assert rows[1]["nets"] == rerouted


@pytest.mark.gpu
@pytest.mark.parametrize("name", G.NAMES)
def test_cli_repeated_ripup_stages_equal_oracle(ga, ora, name, tmp_path):
    """--flt-stages 3: the later stages reroute the nets whose current
    (phase-two) routes still cross more than 10 overflowed edges; the .out,
    statistics and per-stage records equal the oracle's three-stage
    pipeline."""
    cap, net = G.unpacked(name, "cap"), G.unpacked(name, "net")
    ref = str(tmp_path / "oracle3.out")
    print(f"pipeline {name} --flt-stages 3", flush=True)
    counts = []
    wl, vias, rerouted = oracle_pipeline(ga, ora, cap, net, ref, stages=3, counts=counts)
    if name == "syn10":
        assert counts[1] > 0  # the instance reroutes nets again in stage 2
    mine = str(tmp_path / "gpu3.out")
    res = _run(["-cap", cap, "-net", net, "-out", mine, "--batching", "hiprt", "--flt-stages", "3"])
    assert res.returncode == 0, res.stderr[-2000:]
    assert open(mine).read() == open(ref).read()
    score = [line for line in res.stdout.splitlines() if line.startswith("SCORE wirelength=")][0]
    assert f"wirelength={wl:.0f} vias={vias} rerouted={rerouted}" in score, score
    assert "SCORE_CONNECTIVITY open=0 not_connected=0 incomplete=0" in res.stdout
    dags = DAG_RE.findall(res.stdout)
    expect = ["lshape"] + ["detour" + ("" if k == 0 else str(k + 1)) for k, n in enumerate(counts) if n]
    assert [x[0] for x in dags] == expect, dags
    assert [int(x[4]) for x in dags[1:]] == [n for n in counts if n]


def test_oracle_repeated_stages_keep_statistics_of_current_routes(ga, ora):
    """After every rip-up stage the oracle's wirelength and via statistics
    equal those of the routes the nets hold now (phase-two words for nets an
    FLT stage routed): a removal that read a net's stale phase-one route
    after its phase-two route replaced it would leave both off."""
    cap, net = G.unpacked("syn10", "cap"), G.unpacked("syn10", "net")
    d = ga.Design.read(cap, net)
    d.prepare(4)
    t = d.topology()
    g = ora.Grid(t)
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ora.batches(t["X"], t["Y"], *geom)
    sched = ga.execution_schedule(t, color, nb)
    routes, demand, ts, wl, vias = g.lshape_stage(sched, t["pin_acc_num"], threads=4)
    assert R.current_route_totals(t, routes, dict(live=None)) == (wl, vias)
    seen = []

    def check(s):
        live = np.zeros(len(s["off"]), np.uint8)
        live[s["off"] >= 0] = 1  # every net with phase-two words has been routed there by now
        totals = R.current_route_totals(t, routes, dict(live=live, off=s["off"], r2=s["r2"]))
        assert totals == (s["wl"], s["vias"]), (s["stage"], totals, s["wl"], s["vias"])
        seen.append(s["stage"])

    st = R.oracle_stages(ora, d, t, g, nb, routes, demand, ts, wl, vias, 3, on_stage=check)
    assert seen == [1, 2, 3] and st["nets"][1] > 0
    # the rip-up of stage 2 read phase-two routes: the nets it flags are
    # classified by them, not by their phase-one routes
    flags_now = g.ripup(st["demand"], routes, t["pin_acc_num"], routes2=st["r2"], pin_off2=st["off"], live2=st["live"])
    flags_stale = g.ripup(st["demand"], routes, t["pin_acc_num"])
    assert not np.array_equal(flags_now, flags_stale)

    # the same run with the stale removal planted in the oracle's stage
    # (phase-one routes removed again in stage 2) breaks the invariant
    class StaleRemoval:
        def __getattr__(self, name):
            return getattr(g, name)

        def flt_stage(self, *args, **kw):
            kw["live2"] = None
            return g.flt_stage(*args, **kw)

    seen.clear()
    with pytest.raises(AssertionError, match=r"^\(2, "):
        R.oracle_stages(ora, d, t, StaleRemoval(), nb, routes, demand, ts, wl, vias, 2, on_stage=check)
    assert seen == [1]
