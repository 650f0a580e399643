"""Repeated rip-up -> FLT stages on the oracle (shared by the pipeline and
full-size tests).

Stage 1 is the reference's FLT stage (src/pipeline.hpp:86-137): rip-up of
the nets whose phase-one routes cross more than 10 overflowed edges, their
FLT DAGs on the congestion view, their phase-one routes removed batch by
batch.  A later stage classifies every net by its current route (phase-two
words for nets an FLT stage routed, as mark_overflow_nets reads
routes_phase2 for flagged nets, src/routing/congestion.cuh:38-45), appends
phase-two words for flagged nets without them in increasing net order (the
reroute storage of src/pipeline_cugr2.hpp:15-62) and removes each net's
current route."""
import numpy as np


def _geom(s):
    return (s["pt_off"], s["pts"], s["hs_off"], s["hsegs"], s["vs_off"], s["vsegs"])


def oracle_stages(ora, d, t, g, nb, routes, demand, ts, wl, vias, stages, threads=4, on_stage=None):
    """Runs up to `stages` rip-up -> FLT stages after an L-shape stage of nb
    batches (its routes, demand, timestamps and statistics).  on_stage, when
    given, is called after each stage with a dict of that stage's inputs and
    results.  Returns the final state dict: r2, off, live, pins2, demand, ts,
    wl, vias, stamp, nets (per stage)."""
    pin_acc = np.asarray(t["pin_acc_num"], np.int64)
    st = dict(r2=None, off=None, live=None, pins2=0, demand=demand, ts=ts, wl=wl, vias=vias, stamp=nb + 1, nets=[])
    for stage in range(1, stages + 1):
        cong, xs, ys = g.congestion_view(st["demand"])
        flags = g.ripup(st["demand"], routes, t["pin_acc_num"], routes2=st["r2"], pin_off2=st["off"],
                        live2=st["live"])
        fs = d.flt_stage(flags, xs, ys, threads=threads)
        st["nets"].append(fs["nets"])
        if fs["nets"] == 0:
            break
        if stage == 1:
            st["off"], st["pins2"] = np.array(fs["pin_off2"], np.int32), int(fs["pins2"])
            st["live"] = np.zeros(len(st["off"]), np.uint8)
            st["r2"] = np.zeros(100 * st["pins2"] + 1, np.int32)
        else:
            off = st["off"]
            for i in np.flatnonzero(flags):
                if off[i] < 0:
                    off[i] = st["pins2"]
                    st["pins2"] += int(pin_acc[i + 1] - pin_acc[i])
            grown = np.zeros(100 * st["pins2"] + 1, np.int32)
            grown[:len(st["r2"])] = st["r2"]
            st["r2"] = grown
        fcol, fnb = ora.batches(t["X"], t["Y"], *_geom(fs), limit=300000)
        plan = d.flt_plan(fcol, fnb)
        first = st["stamp"]
        r2, dem, ts2, wl2, vias2, rc, last = g.flt_stage(
            plan, t["pin_acc_num"], st["off"], st["pins2"], st["demand"], st["ts"], routes, first, wl=st["wl"],
            vias=st["vias"], threads=threads, routes2=st["r2"], live2=st["live"] if stage > 1 else None)
        assert rc == 0
        assert last == first + 2 * (len(plan["batch_off"]) - 1) - 1
        st.update(r2=r2, demand=dem, ts=ts2, wl=wl2, vias=vias2, stamp=last + 1)
        if on_stage is not None:
            on_stage(dict(stage=stage, flags=flags, cong=cong, xs=xs, ys=ys, fs=fs, color=fcol, nb=fnb, plan=plan,
                          first=first, off=st["off"].copy(), pins2=st["pins2"], r2=r2, demand=dem, ts=ts2, wl=wl2,
                          vias=vias2))
        st["live"][np.asarray(plan["net_ids"], np.int64)] = 1
    return st


def current_route_totals(t, routes, st, words_per_pin=12):
    """Wirelength and via count of every net's current route (phase-two
    words for live nets), counted as the commit kernels count them
    (src/routing/demand.cuh:236-319): wire edges by their edge lengths,
    |layer difference| per via pair."""
    X, Y = t["X"], t["Y"]
    XY = X * Y
    xl = np.asarray(t["x_edge_len"], np.float64)
    yl = np.asarray(t["y_edge_len"], np.float64)
    xc = np.concatenate([[0.0], np.cumsum(xl)])
    yc = np.concatenate([[0.0], np.cumsum(yl)])
    pin_acc = np.asarray(t["pin_acc_num"], np.int64)
    wl, vias = 0.0, 0
    for i in range(len(pin_acc) - 1):
        if st["live"] is not None and st["live"][i]:
            base, words = 100 * int(st["off"][i]), st["r2"]
        else:
            base, words = words_per_pin * int(pin_acc[i]), routes
        n = int(words[base])
        w = np.asarray(words[base + 1:base + n], np.int64)
        a, b = w[0::2], w[1::2]
        la, lb = a // XY, b // XY
        x0, y0, x1, y1 = a // Y % X, a % Y, b // Y % X, b % Y
        same = la == lb
        wl += float(np.sum(np.abs(xc[x1[same]] - xc[x0[same]])) + np.sum(np.abs(yc[y1[same]] - yc[y0[same]])))
        vias += int(np.sum(np.abs(la - lb)[~same]))
    return wl, vias
