#!/usr/bin/env python3
"""Routing hot-path benchmark: routed nets/s (batching + DAG) on one MI355X.

Workload (default): configuration 3 -- 1024x1024 grid, 10 metal layers (9
routing layers), 1M synthetic nets with heavy-tailed pin counts (Zipf,
s = 2.2, up to 200 pins), seed 0.  The front end (pin selection, trees,
>12-pin split) runs once outside the timed region.  A step is the
reference's rip-up pipeline (src/pipeline.hpp:14-157) over the instance:
the L-shape stage over every routed subnet, then rip-up of the nets that
cross more than 10 overflowed edges, and the FLT stage over them (their
Steiner-topology routing DAGs).  As in the reference's metric
(docs/amd.md:64-65, script/experiments/run_amd.py:98-108) a stage's time is
its batching call's wall time plus its synchronized DAG interval; the FLT
interval includes the upload of the stage's DAG records, as the reference's
does (src/routing/flt.cuh:131-137).  Work between the intervals (host
regrouping, rip-up classification, DAG construction, plan building) is
outside both, as it is in the reference.  `value` is the routed nets of both
stages per second of their summed intervals, in the persistent form.

The same steps run in three forms on the same context, with identical
routes and evaluator cost checked: plain (the reference's execution: one
launch per kernel per batch, one net or DAG per wavefront, DP tables in
global memory), launch (DP tables in LDS, autotuned nets per wavefront) and
persistent (work-queue DP launches draining each batch's nets tallest
first, tables in LDS, each stage's batch loop captured in a HIP graph).

Prints one JSON line.  `roofline` reports the kernel class with the largest
device time in the persistent form, timed with HIP events on the context's
stream; `cpu_baseline` times the CPU oracle (the restated reference path) on
the same instance, all host threads and one thread, and checks that its
routes equal the GPU's; `config.soft_dp` carries the soft-DP forward + adjoint
times on configuration 4 (512x512x9, 300k nets) for tau in {0.1, 1, 10}.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KIND_NAMES = ["refresh_tracks_kernel", "lshape_dp_kernel", "commit_wires_kernel", "commit_vias_kernel",
              "batch_marks_kernels", "color_kernel", "ripup_kernels", "zshape_kernels", "flt_edge_kernel",
              "flt_route_kernel"]
CONFIG_TEXT = {
    1: "128x128x5, 20k nets",
    2: "512x512x9, 300k nets",
    3: "1024x1024x10, 1M nets",
    4: "512x512x9, 300k nets (soft DP)",
    5: "2048x2048x10, 4M nets",
    6: "64x64x5, 2k nets",
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=None,
                    help="workload configuration (default 3 for the pipeline, 2 for single-stage runs)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--net-scale", type=float, default=1.0)
    ap.add_argument("--mode", type=int, default=2, choices=(0, 1, 2),
                    help="single-stage runs: 0 plain form (the reference's execution: one launch per batch, "
                         "DP tables in global memory), 1 launch form (DP tables in LDS), 2 persistent form "
                         "(work-queue DP launches, tables in LDS, the batch loop captured in a HIP graph)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-kernel HIP events in the timed steps (roofline from one extra profiled step)")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 8))
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the nets-per-wavefront autotuner (one net / DAG per wavefront), e.g. for counter runs")
    ap.add_argument("--stage", choices=("pipeline", "lshape", "zshape", "soft", "flt", "score"), default="pipeline",
                    help="pipeline: L-shape stage + FLT rip-up stage, plain and persistent forms (the headline); "
                         "lshape: the L-shape stage alone; zshape: the Z-shape stage; soft: soft-DP fwd+bwd; "
                         "flt: the FLT rip-up stage after the L-shape stage; score: the evaluator score of the "
                         "L-shape stage's routes")
    ap.add_argument("--no-soft", action="store_true", help="pipeline: skip the configuration-4 soft-DP timings")
    ap.add_argument("--cpu-sample", type=float, default=0.2,
                    help="pipeline: fraction of the L-shape stage's nets (leading batches) in the one-thread CPU run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reference for HBM counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy of known size: rocprofv3 FETCH_SIZE / WRITE_SIZE of this
    kernel calibrate the counters against exact bytes read and written."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                                           "launches": 5}}), flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` (configN, or
    configN-launches for the L-shape stage's mode 0) from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE table
    (profiles/hbm_traffic.txt), scaled by its plain-copy calibration; None
    when not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 6 and not line.startswith("#") and f[0] == kernel and f[1] == workload:
            fetch, write, fs, ws = map(float, f[2:])
            return round((fetch * fs + write * ws) * 1024.0)
    return None


def main():
    args = parse()
    if args.config is None:
        args.config = 3 if args.stage == "pipeline" else 2
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU path is out of scope).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import grotor_amd as ga

    t0 = time.perf_counter()
    design = ga.Design.generate(args.config, seed=args.seed, net_scale=args.net_scale)
    design.prepare(args.threads)
    topo = design.topology()
    prep_s = time.perf_counter() - t0
    print(f"[bench] instance config={args.config} subnets={topo['nets']} nodes={int(topo['node_off'][-1])} "
          f"prepare={prep_s:.2f}s device={ga.device_name()}", file=sys.stderr, flush=True)

    if args.stage == "pipeline":
        return bench_pipeline(args, ga, design, topo)
    if args.stage == "zshape":
        return bench_zshape(args, ga, design, topo)
    if args.stage == "soft":
        return bench_soft(args, ga, design, topo)
    if args.stage == "flt":
        return bench_flt(args, ga, design, topo)
    if args.stage == "score":
        return bench_score(args, ga, design, topo)
    ctx = ga.Context(topo)
    ctx.set_routes(topo["pin_acc_num"])
    geom = (topo["pt_off"], topo["pts"], topo["hs_off"], topo["hsegs"], topo["vs_off"], topo["vsegs"])

    # autotuner: nets per wavefront of the DP launches, once, before any step
    ctx.batch_geometry(*geom)
    tcolor, tnb = ctx.batch_colors()
    tsched = ga.execution_schedule(topo, tcolor, tnb)
    ctx.lshape_upload(tsched["net_ids"], tsched["node_off"], tsched["nodes"], tsched["parents"])
    t0 = time.perf_counter()
    per_wave, tune_ms = ([1, 1], [[-1.0] * 4] * 2) if args.no_tune else ctx.lshape_tune(tsched["batch_off"])
    tune_s = time.perf_counter() - t0
    print(f"[bench] autotuner: nets per wavefront {per_wave} (classes <=16 nodes, larger); "
          f"DP ms per candidate 1/2/4/8: {[[round(v, 2) for v in row] for row in tune_ms]} ({tune_s:.2f} s)",
          file=sys.stderr, flush=True)

    def step(mode=args.mode):
        ctx.reset()
        u0 = time.perf_counter()
        ctx.batch_geometry(*geom)
        upload_ms = (time.perf_counter() - u0) * 1e3
        b0 = time.perf_counter()
        color, nb = ctx.batch_colors()
        batch_ms = (time.perf_counter() - b0) * 1e3
        sched = ga.execution_schedule(topo, color, nb)
        ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
        dag_ms = ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=mode)
        return batch_ms, dag_ms, nb, color, sched, upload_ms

    # The last warmup step runs with HIP events on every kernel: it gives the
    # per-class breakdown and names the dominant class.  The timed steps then
    # carry events on that class only (mode 0), so the events add little host
    # work to the timed region; a captured graph (modes 1 and 2) cannot carry
    # them, so its roofline comes from one extra step after the timed ones,
    # in which the library issues the same launches directly.
    breakdown = None
    for i in range(args.warmup):
        last = i == args.warmup - 1
        if last:
            ctx.profile_select(None)
            ctx.profile_begin()
        bm, dm, nb, _, _, um = step()
        if last:
            breakdown = ctx.profile_end()
        print(f"[bench] warmup {i}: batching {bm:.2f} ms (+ geometry upload {um:.2f} ms), DAG {dm:.2f} ms, "
              f"{nb} batches", file=sys.stderr, flush=True)
    dom_kind = None
    if breakdown is not None:
        dom_kind = max(range(len(breakdown["ms"])), key=lambda k: breakdown["ms"][k])
    inline_profile = args.mode == 0 and not args.no_profile
    if inline_profile:
        ctx.profile_select(None if dom_kind is None else [dom_kind])
        ctx.profile_begin()
    total_ms = 0.0
    batch_total = dag_total = upload_total = 0.0
    for i in range(args.steps):
        bm, dm, nb, color, sched, um = step()
        batch_total += bm
        dag_total += dm
        upload_total += um
        total_ms += bm + dm
        print(f"[bench] step {i}: batching {bm:.2f} ms (+ geometry upload {um:.2f} ms), DAG {dm:.2f} ms",
              file=sys.stderr, flush=True)
    routes = ctx.routes()
    wl, vias = ctx.stats()
    if inline_profile:
        prof = ctx.profile_end()
    else:
        ctx.profile_select(None if dom_kind is None else [dom_kind])
        ctx.profile_begin()
        step()
        prof = ctx.profile_end()
    if breakdown is None:
        breakdown = prof

    nets = topo["nets"]
    value = nets * args.steps / (total_ms / 1e3)
    ms_per_step = total_ms / args.steps
    dom = max(range(len(prof["ms"])), key=lambda k: prof["ms"][k])
    dom_ms = prof["ms"][dom]
    dom_launches = prof["launches"][dom]
    achieved = prof["bytes"][dom] / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6),
                "traffic": measured_traffic(KIND_NAMES[dom], {2: f"config{args.config}",
                                                              0: f"config{args.config}-launches"}.get(args.mode, "")),
                "events": "timed steps (this class only)" if inline_profile
                          else "one extra step after the timed steps, its launches issued directly",
                "kernel": KIND_NAMES[dom], "avg_launch_us": round(dom_ms * 1e3 / max(dom_launches, 1), 3),
                "launches": int(dom_launches),
                "alg_bytes_per_launch": round(prof["bytes"][dom] / max(dom_launches, 1), 1)}
    kernels = {KIND_NAMES[k]: {"ms": round(breakdown["ms"][k], 4), "launches": int(breakdown["launches"][k]),
                               "GBs": round(breakdown["bytes"][k] / (breakdown["ms"][k] / 1e3) / 1e9, 2)
                               if breakdown["ms"][k] else 0}
               for k in range(len(breakdown["ms"])) if breakdown["launches"][k]}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as ora
        c0 = time.perf_counter()
        ocolor, onb = ora.batches(topo["X"], topo["Y"], *geom)
        cpu_batch_s = time.perf_counter() - c0
        osched = ga.execution_schedule(topo, ocolor, onb)
        c1 = time.perf_counter()
        oroutes, _, _, owl, ovias = ora.Grid(topo).lshape_stage(osched, topo["pin_acc_num"], threads=args.threads)
        cpu_dag_s = time.perf_counter() - c1
        same = bool(onb == nb and np.array_equal(ocolor, color) and np.array_equal(oroutes, routes)
                    and owl == wl and ovias == vias)
        cpu = {"value": round(nets / (cpu_batch_s + cpu_dag_s), 1), "unit": "nets/s", "cores": args.threads,
               "kind": "port",
               "sample": f"whole workload: {nets} subnets; CPU oracle batching on 1 thread "
                         f"({cpu_batch_s * 1e3:.1f} ms) + DAG loop on {args.threads} OpenMP threads "
                         f"({cpu_dag_s * 1e3:.1f} ms)",
               "batching_ms": round(cpu_batch_s * 1e3, 3), "dag_ms": round(cpu_dag_s * 1e3, 3),
               "routes_identical_to_gpu": same}
        print(f"[bench] cpu: batching {cpu_batch_s * 1e3:.1f} ms, DAG {cpu_dag_s * 1e3:.1f} ms, identical={same}",
              file=sys.stderr, flush=True)

    line = {
        "metric": "routed nets/s (batching+DAG) at equal evaluator cost",
        "value": round(value, 1),
        "unit": "nets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"configuration {args.config} L-shape stage ({CONFIG_TEXT.get(args.config, '')}, "
                               f"{nets} routed subnets): GPU batching + L-shape batch loop",
                   "routed_subnets": nets, "batches": nb, "mode": args.mode,
                   "nets_per_wavefront": {"classes": ["<=16 nodes", "larger"],
                                          "chosen": per_wave,
                                          "dp_ms_1_2_4_8": [[round(v, 3) for v in row] for row in tune_ms]},
                   "batching_ms_per_step": round(batch_total / args.steps, 4),
                   "geometry_upload_ms_per_step": round(upload_total / args.steps, 4),
                   "dag_ms_per_step": round(dag_total / args.steps, 4),
                   "total_wirelength": wl, "total_vias": vias, "parallelism": "single GPU",
                   "kernel_breakdown": {"steps": 1 if args.warmup else args.steps, "classes": kernels}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def host_info(threads):
    """Host CPU model, sockets and thread settings for the CPU baseline."""
    import subprocess
    info = {"nproc": os.cpu_count(), "threads_used": threads, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def bench_pipeline(args, ga, design, topo):
    """The rip-up pipeline (L-shape stage, rip-up, FLT stage) in the plain
    and persistent forms; see the module docstring."""
    import numpy as np
    nsub = len(topo["pin_acc_num"]) - 1
    geom = (topo["pt_off"], topo["pts"], topo["hs_off"], topo["hsegs"], topo["vs_off"], topo["vsegs"])
    ctx = ga.Context(topo)
    ctx.set_routes(topo["pin_acc_num"])
    forms = {"plain": {"lshape_mode": 0, "flt_mode": 0}, "launch": {"lshape_mode": 1, "flt_mode": 1},
             "persistent": {"lshape_mode": 2, "flt_mode": 2}}

    # autotuners, once, before any step: nets per wavefront of the L-shape DP
    # and DAGs per wavefront of the FLT DP
    ctx.batch_geometry(*geom)
    tcolor, tnb = ctx.batch_colors()
    tsched = ga.execution_schedule(topo, tcolor, tnb)
    ctx.lshape_upload(tsched["net_ids"], tsched["node_off"], tsched["nodes"], tsched["parents"])
    l_tune = ([1, 1], [[-1.0] * 4] * 2) if args.no_tune else ctx.lshape_tune(tsched["batch_off"])
    g_tune = (4, [-1.0] * 4) if args.no_tune else ctx.lshape_tune_groups(tsched["batch_off"])
    t_tune = (0, [-1.0] * 4) if args.no_tune else ctx.lshape_tune_tiny(tsched["batch_off"])
    # the bound again at the group's chosen nets per wavefront
    g_tune = g_tune if args.no_tune else ctx.lshape_tune_groups(tsched["batch_off"])
    ctx.reset()
    ctx.lshape_run(tsched["batch_off"], first_stamp=1, mode=0)
    flags = ctx.ripup(nsub)
    _, xs, ys = ctx.congestion_view()
    fs = design.flt_stage(flags, xs, ys, threads=args.threads)
    fgeom = (fs["pt_off"], fs["pts"], fs["hs_off"], fs["hsegs"], fs["vs_off"], fs["vsegs"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    ctx.batch_geometry(*fgeom)
    fcolor, fnb = ctx.batch_colors(limit=300000)
    ctx.flt_upload(design.flt_plan(fcolor, fnb))
    f_tune = ([1, 1], [[-1.0] * 4] * 2) if args.no_tune else ctx.flt_tune()
    print(f"[bench] autotuners: L-shape nets per wavefront {l_tune[0]}, persistent tiny-tree group <= {g_tune[0]} "
          f"nodes {[round(v, 3) for v in g_tune[1]]} at {t_tune[0]} per wavefront {[round(v, 3) for v in t_tune[1]]}, "
          f"FLT DAGs per wavefront {f_tune[0]}; "
          f"{fs['nets']} rerouted nets, {fs['nodes']} DAG nodes, {fs['edges']} DAG edges",
          file=sys.stderr, flush=True)

    def step(form, profile=False):
        """One pipeline pass; returns its intervals (ms) and results."""
        f = forms[form]
        ctx.reset()
        if profile:
            ctx.profile_select(None)
            ctx.profile_begin()
        ctx.batch_geometry(*geom)
        t0 = time.perf_counter()
        color, nb = ctx.batch_colors()
        lb = (time.perf_counter() - t0) * 1e3
        sched = ga.execution_schedule(topo, color, nb)
        ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
        ld = ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=f["lshape_mode"])
        flg = ctx.ripup(nsub)
        _, sx, sy = ctx.congestion_view()
        st = design.flt_stage(flg, sx, sy, threads=args.threads)
        assert st["nets"] == fs["nets"], "rip-up set changed between steps"
        ctx.set_routes_phase2(st["pin_off2"], st["pins2"])
        ctx.batch_geometry(*fgeom)
        t0 = time.perf_counter()
        fc, fb_n = ctx.batch_colors(limit=300000)
        fb = (time.perf_counter() - t0) * 1e3
        plan = design.flt_plan(fc, fb_n)
        t0 = time.perf_counter()
        ctx.flt_upload(plan)
        fu = (time.perf_counter() - t0) * 1e3
        fd = ctx.flt_run(first_stamp=nb + 1, remove_phase_one=True, mode=f["flt_mode"])
        prof = ctx.profile_end() if profile else None
        return dict(lb=lb, ld=ld, fb=fb, fu=fu, fd=fd, nb=nb, fnb=len(plan["batch_off"]) - 1, color=color, fc=fc,
                    prof=prof)

    results = {}
    for form in ("plain", "launch", "persistent"):
        for i in range(args.warmup):
            r = step(form)
            print(f"[bench] {form} warmup {i}: L-shape {r['lb']:.2f} + {r['ld']:.2f} ms, FLT {r['fb']:.2f} + "
                  f"{r['fu']:.2f} + {r['fd']:.2f} ms", file=sys.stderr, flush=True)
        acc = dict(lb=0.0, ld=0.0, fb=0.0, fu=0.0, fd=0.0)
        for i in range(args.steps):
            r = step(form)
            for k in acc:
                acc[k] += r[k]
            print(f"[bench] {form} step {i}: L-shape {r['lb']:.2f} + {r['ld']:.2f} ms, FLT {r['fb']:.2f} + "
                  f"{r['fu']:.2f} + {r['fd']:.2f} ms", file=sys.stderr, flush=True)
        per = {k: v / args.steps for k, v in acc.items()}
        wl, vias = ctx.stats()
        nets = design.nets()
        mg = design.metal_grid()
        sc = ctx.score(nets, mg["dir"], np.asarray(mg["capacity"], np.float64), mg["overflow_cost"])
        results[form] = dict(per=per, nb=r["nb"], fnb=r["fnb"], routes=ctx.routes(), routes2=ctx.routes2(),
                             color=r["color"], fc=r["fc"], wl=wl, vias=vias,
                             score={k: sc[k] for k in ("wirelength", "via", "overflow", "total")},
                             rejected=sc["rejected"])

    # per-kernel device time of the persistent form: one more step whose
    # launches are issued directly (a replayed graph carries no events)
    prof = step("persistent", profile=True)["prof"]
    ln, fn = topo["nets"], fs["nets"]

    def rate(per):
        lms, fms = per["lb"] + per["ld"], per["fb"] + per["fu"] + per["fd"]
        return dict(ms=lms + fms, lshape_ms=lms, flt_ms=fms, value=(ln + fn) / ((lms + fms) / 1e3),
                    lshape_nets_per_s=ln / (lms / 1e3), flt_nets_per_s=fn / (fms / 1e3) if fn else None,
                    batching_ms={"lshape": round(per["lb"], 4), "flt": round(per["fb"], 4)},
                    dag_ms={"lshape": round(per["ld"], 4), "flt_upload": round(per["fu"], 4),
                            "flt": round(per["fd"], 4)})

    pl, la, pe = (rate(results[f]["per"]) for f in ("plain", "launch", "persistent"))
    b = results["persistent"]
    identical = all(np.array_equal(results[f]["routes"], b["routes"]) and np.array_equal(results[f]["routes2"], b["routes2"])
                    and results[f]["score"] == b["score"] and results[f]["wl"] == b["wl"] and results[f]["vias"] == b["vias"]
                    for f in ("plain", "launch"))
    dom = max(range(len(prof["ms"])), key=lambda k: prof["ms"][k])
    dom_ms, dom_launches = prof["ms"][dom], prof["launches"][dom]
    achieved = prof["bytes"][dom] / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6),
                "traffic": measured_traffic(KIND_NAMES[dom], f"config{args.config}-pipeline"),
                "events": "one extra persistent-form step after the timed steps, its launches issued directly",
                "kernel": KIND_NAMES[dom], "avg_launch_us": round(dom_ms * 1e3 / max(dom_launches, 1), 3),
                "launches": int(dom_launches), "alg_bytes_per_launch": round(prof["bytes"][dom] / max(dom_launches, 1), 1)}
    kernels = {KIND_NAMES[k]: {"ms": round(prof["ms"][k], 4), "launches": int(prof["launches"][k]),
                               "GBs": round(prof["bytes"][k] / (prof["ms"][k] / 1e3) / 1e9, 2) if prof["ms"][k] else 0}
               for k in range(len(prof["ms"])) if prof["launches"][k]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_pipeline_baseline(args, ga, design, topo, fs, results["persistent"])
    soft = None if args.no_soft else soft_timings(args, ga)
    line = {
        "metric": "routed nets/s (batching+DAG) at equal evaluator cost",
        "value": round(pe["value"], 1),
        "unit": "nets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(pe["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"configuration {args.config} rip-up pipeline ({CONFIG_TEXT.get(args.config, '')}): "
                               f"L-shape stage over {ln} routed subnets, then the FLT stage over the {fn} rerouted "
                               "nets; GPU batching + DAG loop per stage, persistent form",
                   "routed_nets": {"lshape": ln, "flt": fn}, "batches": {"lshape": b["nb"], "flt": b["fnb"]},
                   "dag_nodes_flt": fs["nodes"], "dag_edges_flt": fs["edges"],
                   "persistent": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in pe.items()},
                   "plain": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in pl.items()},
                   "launch": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in la.items()},
                   "persistent_over_plain": {"pipeline": round(pl["ms"] / pe["ms"], 4),
                                             "lshape": round(pl["lshape_ms"] / pe["lshape_ms"], 4),
                                             "flt": round(pl["flt_ms"] / pe["flt_ms"], 4)},
                   "persistent_over_launch": {"pipeline": round(la["ms"] / pe["ms"], 4),
                                              "lshape": round(la["lshape_ms"] / pe["lshape_ms"], 4),
                                              "flt": round(la["flt_ms"] / pe["flt_ms"], 4)},
                   "forms": {"plain": "the reference's execution: per batch one launch per kernel, one net or DAG "
                                      "per wavefront, DP tables in global memory",
                             "launch": "per batch one launch per kernel (L-shape DP: one per tree-size class, "
                                       "on two streams), DP tables in LDS, autotuned nets (DAGs) per wavefront",
                             "persistent": "per batch (L-shape: per tree-size class, on two streams) one "
                                           "work-queue DP launch draining the nets (DAGs) tallest first, DP tables "
                                           "in LDS, autotuned nets (DAGs) per wavefront, each stage's batch loop "
                                           "captured in a HIP graph"},
                   "forms_identical": identical,
                   "evaluator_cost": b["score"], "evaluator_rejected_nets": b["rejected"],
                   "total_wirelength": b["wl"], "total_vias": b["vias"],
                   "nets_per_wavefront": {"lshape": {"classes": ["<=16 nodes", "larger"], "chosen": l_tune[0],
                                                     "dp_ms_1_2_4_8": [[round(v, 3) for v in row] for row in l_tune[1]]},
                                          "flt": {"classes": ["batches < 512 DAGs", ">= 512 DAGs"], "chosen": f_tune[0],
                                                  "dp_ms_1_2_4_8": [[round(v, 3) for v in row] for row in f_tune[1]]},
                                          "lshape_persistent_groups": {
                                              "tiny_tree_bound": g_tune[0],
                                              "dp_ms_bound_1_2_4_8": [round(v, 3) for v in g_tune[1]],
                                              "tiny_nets_per_wavefront": t_tune[0],
                                              "dp_ms_tiny_per_wave_1_2_4_8": [round(v, 3) for v in t_tune[1]]}},
                   "soft_dp": soft, "parallelism": "single GPU",
                   "kernel_breakdown": {"steps": 1, "form": "persistent", "classes": kernels}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def cpu_pipeline_baseline(args, ga, design, topo, fs, gpu):
    """The oracle on the same instance: both stages with the DAG loop on all
    threads (batching is sequential), routes checked against the GPU's, and
    the L-shape DAG loop on one thread over the leading batches."""
    import numpy as np
    from oracle import oracle as ora
    geom = (topo["pt_off"], topo["pts"], topo["hs_off"], topo["hsegs"], topo["vs_off"], topo["vsegs"])
    fgeom = (fs["pt_off"], fs["pts"], fs["hs_off"], fs["hsegs"], fs["vs_off"], fs["vsegs"])
    g = ora.Grid(topo)
    c0 = time.perf_counter()
    color, nb = ora.batches(topo["X"], topo["Y"], *geom)
    lb = time.perf_counter() - c0
    sched = ga.execution_schedule(topo, color, nb)
    c0 = time.perf_counter()
    routes, demand, ts, wl, vias = g.lshape_stage(sched, topo["pin_acc_num"], threads=args.threads)
    ld = time.perf_counter() - c0
    c0 = time.perf_counter()
    fcol, fnb = ora.batches(topo["X"], topo["Y"], *fgeom, limit=300000)
    fb = time.perf_counter() - c0
    plan = design.flt_plan(fcol, fnb)
    c0 = time.perf_counter()
    r2, _, _, wl2, vias2, rc, _ = g.flt_stage(plan, topo["pin_acc_num"], fs["pin_off2"], fs["pins2"], demand, ts,
                                              routes, nb + 1, wl=wl, vias=vias, threads=args.threads)
    fd = time.perf_counter() - c0
    same = bool(rc == 0 and np.array_equal(color, gpu["color"]) and np.array_equal(fcol, gpu["fc"])
                and np.array_equal(routes, gpu["routes"]) and np.array_equal(r2, gpu["routes2"])
                and wl2 == gpu["wl"] and vias2 == gpu["vias"])
    # one thread: the L-shape DAG loop over the leading batches holding about
    # args.cpu_sample of the stage's nets
    bo = sched["batch_off"]
    k = int(np.searchsorted(bo, args.cpu_sample * bo[-1])) if args.cpu_sample > 0 else 1
    k = max(1, min(len(bo) - 1, k))
    sample = dict(sched)
    sample["batch_off"] = bo[:k + 1]
    c0 = time.perf_counter()
    g.lshape_stage(sample, topo["pin_acc_num"], threads=1)
    one = time.perf_counter() - c0
    ln, fn = topo["nets"], fs["nets"]
    total = lb + ld + fb + fd
    print(f"[bench] cpu: L-shape batching {lb * 1e3:.1f} ms + DAG {ld * 1e3:.1f} ms, FLT batching {fb * 1e3:.1f} ms "
          f"+ DAG {fd * 1e3:.1f} ms; one thread {int(bo[k])} nets in {one:.2f} s; identical={same}",
          file=sys.stderr, flush=True)
    return {"value": round((ln + fn) / total, 1), "unit": "nets/s", "cores": args.threads, "kind": "port",
            "sample": f"whole workload: both stages ({ln} + {fn} nets); CPU oracle batching on 1 thread, DAG loops "
                      f"with the nets of a batch over {args.threads} OpenMP threads",
            "lshape": {"batching_ms": round(lb * 1e3, 3), "dag_ms": round(ld * 1e3, 3),
                       "nets_per_s": round(ln / (lb + ld), 1)},
            "flt": {"batching_ms": round(fb * 1e3, 3), "dag_ms": round(fd * 1e3, 3),
                    "nets_per_s": round(fn / (fb + fd), 1) if fn else None},
            "one_thread": {"sample": f"L-shape DAG loop over the first {k} of {len(bo) - 1} batches "
                                     f"({int(bo[k])} nets)", "dag_ms": round(one * 1e3, 3),
                           "nets_per_s": round(int(bo[k]) / one, 1)},
            "host": host_info(args.threads), "routes_identical_to_gpu": same}


def soft_timings(args, ga):
    """Soft-DP forward + adjoint over configuration 4 (the 512x512x9,
    300k-net instance) at the demand its L-shape stage leaves, FP32 tables,
    tau in {0.1, 1, 10}; ms per pass over every routed subnet."""
    d = ga.Design.generate(4, seed=args.seed)
    d.prepare(args.threads)
    t = d.topology()
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    geom = (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    color, nb = ctx.batch_plan(*geom)
    sched = ga.execution_schedule(t, color, nb)
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
    ctx.lshape_upload(t["order"], t["node_off"], t["nodes"], t["parents"])
    nets = t["nets"]
    out = {"workload": f"configuration 4 ({nets} routed subnets, 512x512x9)", "precision": "f32 tables",
           "edges": "L-shape DP edges (the tau_* entries) and the Z-shape route family (zshape)"}
    for kind in (0, 1):
        ctx.set_soft_kind(kind)
        dest = out if kind == 0 else out.setdefault("zshape", {})
        for tau in (0.1, 1.0, 10.0):
            ctx.soft_forward(0, nets, tau=tau, precision=32)
            ctx.soft_backward()
            f = b = 0.0
            reps = 3
            for _ in range(reps):
                _, _, fm = ctx.soft_forward(0, nets, tau=tau, precision=32)
                _, bm = ctx.soft_backward()
                f += fm
                b += bm
            dest[f"tau_{tau:g}"] = {"fwd_ms": round(f / reps, 3), "bwd_ms": round(b / reps, 3),
                                    "fwd_bwd_ms": round((f + b) / reps, 3)}
            print(f"[bench] soft {'zshape' if kind else 'lshape'} tau={tau}: fwd {f / reps:.2f} ms, "
                  f"bwd {b / reps:.2f} ms", file=sys.stderr, flush=True)
    ctx.close()
    return out


def bench_flt(args, ga, design, topo):
    """FLT rip-up stage (src/routing/flt.cuh:75-182): after the L-shape
    stage, rip-up and the congestion view, GPU batching of the rerouted
    nets (limit 300000) and the FLT batch loop (phase-one removal, cost
    refresh, layer-change search, DAG DP, phase-two commit).  The L-shape
    stage that rebuilds the phase-one state before every step is untimed."""
    import numpy as np
    ctx = ga.Context(topo)
    ctx.set_routes(topo["pin_acc_num"])
    geom = (topo["pt_off"], topo["pts"], topo["hs_off"], topo["hsegs"], topo["vs_off"], topo["vsegs"])
    nsub = len(topo["pin_acc_num"]) - 1

    def phase_one():
        ctx.reset()
        ctx.batch_geometry(*geom)
        color, nb = ctx.batch_colors()
        sched = ga.execution_schedule(topo, color, nb)
        ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
        ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
        return nb

    nb1 = phase_one()
    flags = ctx.ripup(nsub)
    cong, xs, ys = ctx.congestion_view()
    t0 = time.perf_counter()
    fs = design.flt_stage(flags, xs, ys, threads=args.threads)
    topo_ms = (time.perf_counter() - t0) * 1e3
    fgeom = (fs["pt_off"], fs["pts"], fs["hs_off"], fs["hsegs"], fs["vs_off"], fs["vsegs"])
    ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])
    nets = fs["nets"]
    # autotuner: DAGs per wavefront of the FLT DP, once, on the first plan
    ctx.batch_geometry(*fgeom)
    tcolor, tnb = ctx.batch_colors(limit=300000)
    ctx.flt_upload(design.flt_plan(tcolor, tnb))
    per_wave, tune_ms = ([1, 1], [[-1.0] * 4] * 2) if args.no_tune else ctx.flt_tune()
    print(f"[bench] flt autotuner: DAGs per wavefront {per_wave} (batches <512, >=512 DAGs); "
          f"DP ms per candidate 1/2/4/8: {[[round(v, 2) for v in row] for row in tune_ms]}",
          file=sys.stderr, flush=True)
    print(f"[bench] flt: {nets} rerouted nets, {fs['nodes']} DAG nodes, {fs['edges']} DAG edges, "
          f"DAG build {topo_ms:.1f} ms", file=sys.stderr, flush=True)

    def step(profile_kinds=False):
        phase_one()
        ctx.set_routes_phase2(fs["pin_off2"], fs["pins2"])  # a reset releases it
        ctx.batch_geometry(*fgeom)
        if profile_kinds is not False:
            ctx.profile_select(profile_kinds)
            ctx.profile_begin()
        b0 = time.perf_counter()
        color, nb = ctx.batch_colors(limit=300000)
        batch_ms = (time.perf_counter() - b0) * 1e3
        u0 = time.perf_counter()
        plan = design.flt_plan(color, nb)  # flt::Plans, outside the reference's DAG interval too
        ctx.flt_upload(plan)
        upload_ms = (time.perf_counter() - u0) * 1e3
        dag_ms = ctx.flt_run(first_stamp=nb1 + 1, remove_phase_one=True, mode=args.mode)
        prof = ctx.profile_end() if profile_kinds is not False else None
        return batch_ms, dag_ms, upload_ms, plan, prof

    breakdown = None
    for i in range(args.warmup):
        last = i == args.warmup - 1
        bm, dm, um, plan, prof = step(None if last else False)
        if last:
            breakdown = prof
        print(f"[bench] flt warmup {i}: batching {bm:.2f} ms, DAG {dm:.2f} ms (+ plan and upload {um:.2f} ms), "
              f"{len(plan['batch_off']) - 1} batches", file=sys.stderr, flush=True)
    dom_kind = None
    if breakdown is not None:
        dom_kind = max(range(len(breakdown["ms"])), key=lambda k: breakdown["ms"][k])
    bt = dt = ut = 0.0
    prof_sum = None
    for i in range(args.steps):
        bm, dm, um, plan, prof = step(False if args.no_profile or dom_kind is None else [dom_kind])
        bt += bm
        dt += dm
        ut += um
        if prof is not None:
            prof_sum = prof if prof_sum is None else {k: [a + b for a, b in zip(prof_sum[k], prof[k])]
                                                      for k in prof}
        print(f"[bench] flt step {i}: batching {bm:.2f} ms, DAG {dm:.2f} ms (+ plan and upload {um:.2f} ms)",
              file=sys.stderr, flush=True)
    routes2 = ctx.routes2()
    wl, vias = ctx.stats()
    if prof_sum is None:
        _, _, _, _, prof_sum = step(None if dom_kind is None else [dom_kind])
    if breakdown is None:
        breakdown = prof_sum
    dom = max(range(len(prof_sum["ms"])), key=lambda k: prof_sum["ms"][k])
    dom_ms, dom_launches = prof_sum["ms"][dom], prof_sum["launches"][dom]
    achieved = prof_sum["bytes"][dom] / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(KIND_NAMES[dom], f"config{args.config}"),
                "kernel": KIND_NAMES[dom], "avg_launch_us": round(dom_ms * 1e3 / max(dom_launches, 1), 3),
                "launches": int(dom_launches), "alg_bytes_per_launch": round(prof_sum["bytes"][dom] / max(dom_launches, 1), 1)}
    kernels = {KIND_NAMES[k]: {"ms": round(breakdown["ms"][k], 4), "launches": int(breakdown["launches"][k]),
                               "GBs": round(breakdown["bytes"][k] / (breakdown["ms"][k] / 1e3) / 1e9, 2)
                               if breakdown["ms"][k] else 0}
               for k in range(len(breakdown["ms"])) if breakdown["launches"][k]}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as ora
        # the oracle continues from the GPU's phase-one state (bit-identical to its own)
        phase_one()
        routes1, demand1, ts1 = ctx.routes(), ctx.demand(), ctx.timestamps()
        wl1, vias1 = ctx.stats()
        c0 = time.perf_counter()
        ocolor, onb = ora.batches(topo["X"], topo["Y"], *fgeom, limit=300000)
        oplan = design.flt_plan(ocolor, onb)
        cb = time.perf_counter() - c0
        c1 = time.perf_counter()
        o2, _, _, owl, ovias, rc, _ = ora.Grid(topo).flt_stage(oplan, topo["pin_acc_num"], fs["pin_off2"], fs["pins2"],
                                                              demand1, ts1, routes1, nb1 + 1, wl=wl1, vias=vias1,
                                                              threads=args.threads)
        cd = time.perf_counter() - c1
        same = bool(rc == 0 and np.array_equal(o2, routes2) and owl == wl and ovias == vias)
        cpu = {"value": round(nets / (cb + cd), 1), "unit": "nets/s", "cores": 1, "kind": "port",
               "sample": f"whole FLT stage: {nets} rerouted nets; CPU oracle batching ({cb * 1e3:.1f} ms) + "
                         f"sequential FLT loop ({cd * 1e3:.1f} ms, cost refresh on {args.threads} threads)",
               "routes_identical_to_gpu": same}
        print(f"[bench] flt cpu: batching {cb * 1e3:.1f} ms, DAG {cd * 1e3:.1f} ms, identical={same}",
              file=sys.stderr, flush=True)
    ms = (bt + dt) / args.steps
    print(json.dumps({
        "metric": "routed nets/s (batching+DAG) at equal evaluator cost",
        "value": round(nets / (ms / 1e3), 1), "unit": "nets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"configuration {args.config} FLT rip-up stage ({CONFIG_TEXT.get(args.config, '')}): "
                               f"{nets} rerouted nets after the L-shape stage; GPU batching + FLT batch loop",
                   "rerouted_nets": nets, "dag_nodes": fs["nodes"], "dag_edges": fs["edges"],
                   "batches": len(plan["batch_off"]) - 1, "batching_ms_per_step": round(bt / args.steps, 4),
                   "dag_ms_per_step": round(dt / args.steps, 4), "plan_upload_ms_per_step": round(ut / args.steps, 4),
                   "dag_build_ms": round(topo_ms, 2), "total_wirelength": wl, "total_vias": vias,
                   "dags_per_wavefront": {"classes": ["batches < 512 DAGs", ">= 512 DAGs"], "chosen": per_wave,
                                          "dp_ms_1_2_4_8": [[round(v, 3) for v in row] for row in tune_ms]},
                   "parallelism": "single GPU", "kernel_breakdown": {"steps": 1, "classes": kernels}},
        "roofline": roofline, "cpu_baseline": cpu}), flush=True)
    ctx.close()


def bench_score(args, ga, design, topo):
    """Evaluator score (resources/evaluator.cpp:458-675) of the routes of one
    GPU L-shape stage (untimed), resident on the device: a step is one
    gr_score call over every design net, device time from its HIP events."""
    import numpy as np
    ctx = ga.Context(topo)
    ctx.set_routes(topo["pin_acc_num"])
    geom = (topo["pt_off"], topo["pts"], topo["hs_off"], topo["hsegs"], topo["vs_off"], topo["vsegs"])
    color, nb = ctx.batch_plan(*geom)
    sched = ga.execution_schedule(topo, color, nb)
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.lshape_run(sched["batch_off"], first_stamp=1, mode=2)
    nets = design.nets()
    mg = design.metal_grid()
    cap = np.asarray(mg["capacity"], np.float64)
    n = len(nets["pin_off"]) - 1
    for i in range(args.warmup):
        r = ctx.score(nets, mg["dir"], cap, mg["overflow_cost"])
        print(f"[bench] score warmup {i}: {r['ms']:.3f} ms", file=sys.stderr, flush=True)
    tot = 0.0
    for i in range(args.steps):
        r = ctx.score(nets, mg["dir"], cap, mg["overflow_cost"])
        tot += r["ms"]
        print(f"[bench] score step {i}: {r['ms']:.3f} ms", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as ora
        c0 = time.perf_counter()
        o = ora.score(int(topo["X"]), int(topo["Y"]), mg["layers"], nets, ctx.routes(), topo["pin_acc_num"], 12,
                      mg["dir"], cap, mg["overflow_cost"], topo["x_edge_len"], topo["y_edge_len"],
                      topo["unit_length_wire_cost"], topo["unit_via_cost"])
        cs = time.perf_counter() - c0
        same = bool(o["wirelength"] == r["wirelength"] and o["via"] == r["via"] and o["rejected"] == r["rejected"]
                    and abs(o["overflow"] - r["overflow"]) <= 1e-10 * abs(o["overflow"]))
        cpu = {"value": round(n / cs, 1), "unit": "nets/s", "cores": 1, "kind": "port",
               "sample": f"whole workload: {n} design nets scored by the CPU oracle ({cs * 1e3:.1f} ms, "
                         "route words already in host memory)", "costs_equal_to_gpu": same}
    ms = tot / args.steps
    print(json.dumps({
        "metric": "scored nets/s (evaluator score of the L-shape stage routes)", "value": round(n / (ms / 1e3), 1),
        "unit": "nets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64", "data": "synthetic",
        "config": {"workload": f"configuration {args.config} ({CONFIG_TEXT.get(args.config, '')}): evaluator score "
                               f"of the L-shape stage routes, {n} design nets",
                   "costs": {k: r[k] for k in ("wirelength", "via", "overflow", "total")}, "rejected": r["rejected"],
                   "parallelism": "single GPU"},
        "cpu_baseline": cpu}))


def bench_zshape(args, ga, design, topo):
    """Z-shape stage (src/pipeline_ggr.hpp): GPU RSMT batching plus the
    host bounding-box refinement and tree encoding as the batching call,
    then the synchronized Z-shape batch loop."""
    import numpy as np
    z = design.zshape_stage()
    geom = (z["pt_off"], z["pts"], z["hs_off"], z["hsegs"], z["vs_off"], z["vsegs"])
    ctx = ga.Context(topo)
    ctx.set_routes(topo["pin_acc_num"], 256)
    ctx.set_cost_params(integer_wire=True)
    nets = z["nets"]

    def step():
        ctx.reset()
        ctx.batch_geometry(*geom)
        b0 = time.perf_counter()
        color, nb = ctx.batch_colors()
        plan = design.zshape_plan(color, nb)
        batch_ms = (time.perf_counter() - b0) * 1e3
        ctx.zshape_upload(plan)
        dag_ms = ctx.zshape_run(first_stamp=1)
        return batch_ms, dag_ms, plan, color, nb

    for i in range(args.warmup):
        bm, dm, plan, _, _ = step()
        print(f"[bench] zshape warmup {i}: batching {bm:.2f} ms, DAG {dm:.2f} ms, "
              f"{len(plan['batch_off']) - 1} batches", file=sys.stderr, flush=True)
    bt = dt = 0.0
    for i in range(args.steps):
        bm, dm, plan, color, nb = step()
        bt += bm
        dt += dm
        print(f"[bench] zshape step {i}: batching {bm:.2f} ms, DAG {dm:.2f} ms", file=sys.stderr, flush=True)
    routes = ctx.routes()
    wl, vias = ctx.stats()
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as ora
        c0 = time.perf_counter()
        ocolor, onb = ora.batches(topo["X"], topo["Y"], *geom)
        oplan = design.zshape_plan(ocolor, onb)
        cb = time.perf_counter() - c0
        c1 = time.perf_counter()
        oroutes, _, _, owl, ovias, _ = ora.Grid(topo).zshape_stage(oplan, topo["pin_acc_num"], threads=args.threads)
        cd = time.perf_counter() - c1
        same = bool(np.array_equal(ocolor, color) and np.array_equal(oroutes, routes) and owl == wl and ovias == vias)
        cpu = {"value": round(nets / (cb + cd), 1), "unit": "nets/s", "cores": 1, "kind": "port",
               "sample": f"whole workload: {nets} subnets; CPU oracle batching + bbox refinement "
                         f"({cb * 1e3:.1f} ms) + sequential Z-shape loop ({cd * 1e3:.1f} ms)",
               "routes_identical_to_gpu": same}
    ms = (bt + dt) / args.steps
    print(json.dumps({
        "metric": "routed nets/s (batching+DAG), Z-shape stage", "value": round(nets / (ms / 1e3), 1),
        "unit": "nets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"configuration {args.config} Z-shape stage ({CONFIG_TEXT.get(args.config, '')}, "
                               f"{nets} routed subnets)", "batches": len(plan["batch_off"]) - 1,
                   "batching_ms_per_step": round(bt / args.steps, 4), "dag_ms_per_step": round(dt / args.steps, 4),
                   "total_wirelength": wl, "total_vias": vias},
        "cpu_baseline": cpu}), flush=True)
    ctx.close()


def bench_soft(args, ga, design, topo):
    """Soft-DP forward + adjoint over every routed net at the demand the
    L-shape stage leaves, for tau in {0.1, 1, 10}, FP32 tables."""
    from oracle import oracle as ora  # noqa: F401  (the CPU leg below)
    ctx = ga.Context(topo)
    ctx.set_routes(topo["pin_acc_num"])
    geom = (topo["pt_off"], topo["pts"], topo["hs_off"], topo["hsegs"], topo["vs_off"], topo["vsegs"])
    color, nb = ctx.batch_plan(*geom)
    sched = ga.execution_schedule(topo, color, nb)
    ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
    ctx.lshape_run(sched["batch_off"], first_stamp=1)
    ctx.lshape_upload(topo["order"], topo["node_off"], topo["nodes"], topo["parents"])
    nets = topo["nets"]
    per_tau, zshape = {}, {}
    for kind, dest in ((1, zshape), (0, per_tau)):
        ctx.set_soft_kind(kind)
        for tau in (0.1, 1.0, 10.0):
            for _ in range(args.warmup):
                ctx.soft_forward(0, nets, tau=tau, precision=32)
                ctx.soft_backward()
            f = b = 0.0
            for _ in range(args.steps):
                total, _, fm = ctx.soft_forward(0, nets, tau=tau, precision=32)
                gp, bm = ctx.soft_backward()
                f += fm
                b += bm
            dest[str(tau)] = {"fwd_ms": round(f / args.steps, 3), "bwd_ms": round(b / args.steps, 3),
                              "total": total, "grad_params": [float(x) for x in gp]}
            print(f"[bench] soft {'zshape' if kind else 'lshape'} tau={tau}: fwd {f / args.steps:.2f} ms, "
                  f"bwd {b / args.steps:.2f} ms", file=sys.stderr, flush=True)
    ms = sum(v["fwd_ms"] + v["bwd_ms"] for v in per_tau.values()) / len(per_tau)
    print(json.dumps({
        "metric": "soft-DP fwd+bwd ms per stage", "value": round(ms, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"configuration {args.config} soft L-shape DP over {nets} routed subnets at the "
                               f"L-shape stage's final demand", "per_tau": per_tau,
                   "zshape_per_tau": zshape}}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
