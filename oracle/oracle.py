"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker; the product never imports it.  Grid arrays use the
reference IDX(l, x, y) order.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
ROUTE_PER_PIN = 12

_lib = None
_i32p = ctypes.POINTER(ctypes.c_int)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


class OraGrid(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("X", ctypes.c_int), ("Y", ctypes.c_int), ("DIR", ctypes.c_int),
                ("unit_length_wire_cost", ctypes.c_double), ("unit_via_cost", ctypes.c_double),
                ("unit_length_short_costs", _f64p), ("x_edge_len", _i32p), ("y_edge_len", _i32p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.ora_expf.restype = ctypes.c_float
        _lib.ora_expf.argtypes = [ctypes.c_float]
        _lib.ora_of_cost.restype = ctypes.c_double
        _lib.ora_of_cost.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_double, ctypes.c_double]
        _lib.ora_incremental_of_cost.restype = ctypes.c_double
        _lib.ora_incremental_of_cost.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_double]
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Grid:
    """Grid parameters shared by every oracle call."""

    def __init__(self, grid):
        self.L, self.X, self.Y, self.DIR = int(grid["L"]), int(grid["X"]), int(grid["Y"]), int(grid["DIR"])
        self.short = np.ascontiguousarray(grid["unit_length_short_costs"], np.float64)
        self.xlen = _i32(grid["x_edge_len"])
        self.ylen = _i32(grid["y_edge_len"])
        self.capacity = np.ascontiguousarray(grid["capacity"], np.float32)
        self.wire = float(grid["unit_length_wire_cost"])
        self.via = float(grid["unit_via_cost"])
        self.c = OraGrid(self.L, self.X, self.Y, self.DIR, self.wire, self.via, _p(self.short, ctypes.c_double),
                         _p(self.xlen, ctypes.c_int), _p(self.ylen, ctypes.c_int))
        self.cells = self.L * self.X * self.Y

    def refresh(self, demand, slope=0.5, zero_slope=1.5, via_weight=1.0, integer_wire=False):
        demand = np.ascontiguousarray(demand, np.float32)
        w = np.empty(self.cells, np.float32)
        v = np.empty(self.cells, np.float32)
        p = np.empty(self.cells, np.float64)
        lib().ora_set_integer_wire(1 if integer_wire else 0)
        try:
            lib().ora_cost_refresh_slopes(ctypes.byref(self.c), _p(self.capacity, ctypes.c_float),
                                          _p(demand, ctypes.c_float), ctypes.c_double(slope),
                                          ctypes.c_double(zero_slope), ctypes.c_double(via_weight),
                                          _p(w, ctypes.c_float), _p(v, ctypes.c_float), _p(p, ctypes.c_double))
        finally:
            lib().ora_set_integer_wire(0)
        return w, v, p

    def lshape_net(self, vcost, presum, nodes, parents, cap=10**6, with_cost=False):
        nodes, parents = _i32(nodes), _i32(parents)
        vcost = np.ascontiguousarray(vcost, np.float32)
        presum = np.ascontiguousarray(presum, np.float64)
        routes = np.zeros(cap, np.int32)
        best = ctypes.c_double()
        rc = lib().ora_lshape_net_cost(ctypes.byref(self.c), _p(vcost, ctypes.c_float), _p(presum, ctypes.c_double),
                                       ctypes.c_int(len(nodes)), _p(nodes, ctypes.c_int), _p(parents, ctypes.c_int),
                                       _p(routes, ctypes.c_int), ctypes.c_int(cap), ctypes.byref(best))
        r = routes[:routes[0]] if rc == 0 else routes
        return (rc, r, best.value) if with_cost else (rc, r)

    def lshape_stage(self, sched, pin_acc_num, threads=1, first_stamp=0):
        """Run the L-shape batch loop; returns (routes, demand, timestamps, wl, vias)."""
        pin_acc_num = _i32(pin_acc_num)
        nsub = len(pin_acc_num) - 1
        route_off = (ROUTE_PER_PIN * pin_acc_num[:-1].astype(np.int64)).astype(np.int64)
        route_cap = (ROUTE_PER_PIN * np.diff(pin_acc_num)).astype(np.int32)
        routes = np.zeros(ROUTE_PER_PIN * int(pin_acc_num[-1]) + 1, np.int32)
        demand = np.zeros(self.cells, np.float32)
        ts = np.zeros(self.cells, np.int32)
        w = np.empty(self.cells, np.float32)
        v = np.empty(self.cells, np.float32)
        p = np.empty(self.cells, np.float64)
        stamp = ctypes.c_int(first_stamp)
        wl = ctypes.c_double(0)
        vias = ctypes.c_longlong(0)
        bo = _i32(sched["batch_off"])
        arrs = [_i32(sched[k]) for k in ("net_ids", "node_off", "nodes", "parents")]
        rc = lib().ora_lshape_stage(ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float),
                                    _p(ts, ctypes.c_int), _p(w, ctypes.c_float), _p(v, ctypes.c_float),
                                    _p(p, ctypes.c_double), ctypes.c_int(len(bo) - 1), _p(bo, ctypes.c_int),
                                    *[_p(a, ctypes.c_int) for a in arrs], _p(route_off, ctypes.c_longlong),
                                    _p(route_cap, ctypes.c_int), _p(routes, ctypes.c_int), ctypes.byref(stamp),
                                    ctypes.byref(wl), ctypes.byref(vias), ctypes.c_int(threads))
        if rc:
            raise RuntimeError(f"oracle L-shape stage failed ({rc})")
        del nsub
        return routes, demand, ts, wl.value, vias.value

    def ripup(self, demand, routes, pin_acc_num, threshold=0, routes2=None, pin_off2=None, live2=None,
              words_per_pin=100):
        """Rip-up flags (src/routing/congestion.cuh:114-136).  A net marked in
        live2 is classified by its phase-two route (routes2 at
        words_per_pin * pin_off2[net]), as mark_overflow_nets reads
        routes_phase2 for nets with a rip-up flag (:38-45)."""
        pin_acc_num = _i32(pin_acc_num)
        n = len(pin_acc_num) - 1
        routes = _i32(routes)
        base = routes.ctypes.data
        if live2 is not None:
            routes2 = _i32(routes2)
            base2 = routes2.ctypes.data
        ptrs = (ctypes.POINTER(ctypes.c_int) * max(n, 1))()
        for i in range(n):
            if live2 is not None and live2[i]:
                ptrs[i] = ctypes.cast(base2 + 4 * words_per_pin * int(pin_off2[i]), _i32p)
            else:
                ptrs[i] = ctypes.cast(base + 4 * ROUTE_PER_PIN * int(pin_acc_num[i]), _i32p)
        flags = np.zeros(max(n, 1), np.uint8)
        demand = np.ascontiguousarray(demand, np.float32)
        lib().ora_ripup(ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float),
                        ctypes.c_int(threshold), ctypes.c_int(n), ptrs, _p(flags, ctypes.c_ubyte))
        return flags[:n]


    def congestion_view(self, demand):
        """(congested, xsum, ysum) over X*Y cells in x*Y+y order."""
        demand = np.ascontiguousarray(demand, np.float32)
        n = self.X * self.Y
        cong, xs, ys = np.zeros(n, np.uint8), np.zeros(n, np.float32), np.zeros(n, np.float32)
        lib().ora_congestion_view(ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float),
                                  _p(cong, ctypes.c_ubyte), _p(xs, ctypes.c_float), _p(ys, ctypes.c_float))
        return cong.astype(bool), xs, ys

    def flt_batch(self, presum, vcost, plan, begin, count, pin_acc_num, pin_off2, words_per_pin=100, routes=None):
        """One FLT batch of plan nets [begin, begin + count) at the given costs;
        returns (phase-two routes, layer_output of the batch nodes, status)."""
        pin_acc_num, pin_off2 = _i32(pin_acc_num), _i32(pin_off2)
        route_off = (words_per_pin * pin_off2.astype(np.int64)).astype(np.int64)
        route_cap = (words_per_pin * np.diff(pin_acc_num)).astype(np.int32)
        if routes is None:
            routes = np.zeros(words_per_pin * (int(np.max(pin_off2 + np.diff(pin_acc_num), initial=0))) + 1, np.int32)
        node_off = np.ascontiguousarray(plan["node_off"], np.int64)
        edge_off = np.ascontiguousarray(plan["edge_off"], np.int64)
        nodes_b = int(node_off[begin + count] - node_off[begin])
        layer = np.zeros(max(nodes_b, 1), np.int32)
        presum = np.ascontiguousarray(presum, np.float64)
        vcost = np.ascontiguousarray(vcost, np.float32)
        rc = lib().ora_flt_batch(ctypes.byref(self.c), _p(presum, ctypes.c_double), _p(vcost, ctypes.c_float),
                                 ctypes.c_int(count), _p(_i32(plan["net_ids"][begin:begin + count]), ctypes.c_int),
                                 _p(node_off[begin:], ctypes.c_longlong), _p(_i32(plan["nodes"]), ctypes.c_int),
                                 _p(edge_off[begin:], ctypes.c_longlong), _p(_i32(plan["edges"]), ctypes.c_int),
                                 _p(route_off, ctypes.c_longlong), _p(route_cap, ctypes.c_int),
                                 _p(routes, ctypes.c_int), _p(layer, ctypes.c_int))
        return routes, layer[:nodes_b], rc

    def flt_stage(self, plan, pin_acc_num, pin_off2, pins2, demand, ts, routes1, first_stamp, remove_phase_one=True,
                  words_per_pin=100, wl=0.0, vias=0, threads=1, routes2=None, live2=None):
        """FLT stage continuing from a phase-one state (demand, timestamps,
        phase-one routes at 12 words per pin, updated in place copies);
        returns (routes2, demand, ts, wl, vias, status, last stamp).  With
        routes2 (phase-two words of earlier stages, laid out by pin_off2) and
        live2 (nets whose current route is there), the stage continues a
        multi-stage rip-up: each net's current route is removed."""
        pin_acc_num, pin_off2 = _i32(pin_acc_num), _i32(pin_off2)
        route_off1 = (ROUTE_PER_PIN * pin_acc_num[:-1].astype(np.int64)).astype(np.int64)
        route_off2 = (words_per_pin * pin_off2.astype(np.int64)).astype(np.int64)
        route_cap2 = (words_per_pin * np.diff(pin_acc_num)).astype(np.int32)
        routes1 = _i32(routes1).copy()
        if routes2 is None:
            routes2 = np.zeros(words_per_pin * int(pins2) + 1, np.int32)
        else:
            routes2 = _i32(routes2).copy()
            assert len(routes2) == words_per_pin * int(pins2) + 1
        remove = (2 if live2 is not None else 1) if remove_phase_one else 0
        live = np.ascontiguousarray(live2 if live2 is not None else np.zeros(1), np.uint8)
        demand = np.ascontiguousarray(demand, np.float32).copy()
        ts = _i32(ts).copy()
        w = np.empty(self.cells, np.float32)
        v = np.empty(self.cells, np.float32)
        p = np.empty(self.cells, np.float64)
        stamp = ctypes.c_int(first_stamp - 1)
        wlc, vc = ctypes.c_double(wl), ctypes.c_longlong(vias)
        bo = _i32(plan["batch_off"])
        rc = lib().ora_flt_stage(
            ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float), _p(ts, ctypes.c_int),
            _p(w, ctypes.c_float), _p(v, ctypes.c_float), _p(p, ctypes.c_double), ctypes.c_int(len(bo) - 1),
            _p(bo, ctypes.c_int), _p(_i32(plan["net_ids"]), ctypes.c_int),
            _p(np.ascontiguousarray(plan["node_off"], np.int64), ctypes.c_longlong), _p(_i32(plan["nodes"]), ctypes.c_int),
            _p(np.ascontiguousarray(plan["edge_off"], np.int64), ctypes.c_longlong), _p(_i32(plan["edges"]), ctypes.c_int),
            _p(route_off1, ctypes.c_longlong), _p(routes1, ctypes.c_int), _p(route_off2, ctypes.c_longlong),
            _p(route_cap2, ctypes.c_int), _p(routes2, ctypes.c_int), ctypes.c_int(remove), _p(live, ctypes.c_ubyte),
            ctypes.byref(stamp), ctypes.byref(wlc), ctypes.byref(vc), ctypes.c_int(threads))
        return routes2, demand, ts, wlc.value, vc.value, rc, stamp.value

    def zshape_stage(self, plan, pin_acc_num, words_per_pin=256, threads=1, first_stamp=0):
        """Z-shape stage over a plan; returns (routes, demand, timestamps, wl, vias, status)."""
        pin_acc_num = _i32(pin_acc_num)
        route_off = (words_per_pin * pin_acc_num[:-1].astype(np.int64)).astype(np.int64)
        route_cap = (words_per_pin * np.diff(pin_acc_num)).astype(np.int32)
        routes = np.zeros(words_per_pin * int(pin_acc_num[-1]) + 1, np.int32)
        demand = np.zeros(self.cells, np.float32)
        ts = np.zeros(self.cells, np.int32)
        w = np.empty(self.cells, np.float32)
        v = np.empty(self.cells, np.float32)
        p = np.empty(self.cells, np.float64)
        stamp = ctypes.c_int(first_stamp)
        wl = ctypes.c_double(0)
        vias = ctypes.c_longlong(0)
        bo = _i32(plan["batch_off"])
        ro = np.ascontiguousarray(plan["rec_off"], np.int64)
        lib().ora_set_integer_wire(1)
        try:
            rc = lib().ora_zshape_stage(
                ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float),
                _p(ts, ctypes.c_int), _p(w, ctypes.c_float), _p(v, ctypes.c_float), _p(p, ctypes.c_double),
                ctypes.c_int(len(bo) - 1), _p(bo, ctypes.c_int), _p(_i32(plan["net_ids"]), ctypes.c_int),
                _p(ro, ctypes.c_longlong), _p(_i32(plan["records"]), ctypes.c_int), _p(_i32(plan["raw_cap"]), ctypes.c_int),
                _p(route_off, ctypes.c_longlong), _p(route_cap, ctypes.c_int), _p(routes, ctypes.c_int),
                ctypes.byref(stamp), ctypes.byref(wl), ctypes.byref(vias), ctypes.c_int(threads))
        finally:
            lib().ora_set_integer_wire(0)
        return routes, demand, ts, wl.value, vias.value, rc

    def zshape_batch(self, presum, vcost, records, count, raw_words):
        """One Z-shape batch on fresh maps: (raw words, DP optimum per net)."""
        n = max(self.X, self.Y)
        plane = (self.L + 1) * n * n
        zmap = np.zeros(6 * plane, np.int32)
        zprev = np.zeros(16 * plane, np.int32)
        raw = np.zeros(max(raw_words, 1), np.int32)
        cost = np.zeros(max(count, 1), np.int32)
        records = _i32(records)
        presum = np.ascontiguousarray(presum, np.float64)
        vcost = np.ascontiguousarray(vcost, np.float32)
        lib().ora_zshape_reset(ctypes.byref(self.c), _p(zmap, ctypes.c_int), _p(records, ctypes.c_int), ctypes.c_int(count))
        lib().ora_zshape_solve(ctypes.byref(self.c), _p(presum, ctypes.c_double), _p(vcost, ctypes.c_float),
                               _p(zmap, ctypes.c_int), _p(zprev, ctypes.c_int), _p(records, ctypes.c_int),
                               ctypes.c_int(count), _p(raw, ctypes.c_int), _p(cost, ctypes.c_int))
        return raw, cost[:count]

    def zshape_probe(self, presum, vcost, records, count, raw_words, ideal):
        """zshape_batch through the probe: (raw words, DP optimum, quirk flag
        per net); ideal traces each middle leg on the layer it really ran on."""
        n = max(self.X, self.Y)
        plane = (self.L + 1) * n * n
        zmap = np.zeros(6 * plane, np.int32)
        zprev = np.zeros(16 * plane, np.int32)
        raw = np.zeros(max(raw_words, 1), np.int32)
        cost = np.zeros(max(count, 1), np.int32)
        quirk = np.zeros(max(count, 1), np.int32)
        records = _i32(records)
        presum = np.ascontiguousarray(presum, np.float64)
        vcost = np.ascontiguousarray(vcost, np.float32)
        lib().ora_zshape_reset(ctypes.byref(self.c), _p(zmap, ctypes.c_int), _p(records, ctypes.c_int), ctypes.c_int(count))
        lib().ora_zshape_probe(ctypes.byref(self.c), _p(presum, ctypes.c_double), _p(vcost, ctypes.c_float),
                               _p(zmap, ctypes.c_int), _p(zprev, ctypes.c_int), _p(records, ctypes.c_int),
                               ctypes.c_int(count), _p(raw, ctypes.c_int), _p(cost, ctypes.c_int),
                               ctypes.c_int(1 if ideal else 0), _p(quirk, ctypes.c_int))
        return raw, cost[:count], quirk[:count]

    def soft_costs(self, demand, slope=0.5, zero_slope=1.5, via_weight=1.0):
        """Smooth cost model: (wcost, vcost, dw_ds, dw_dz, dv_ds, dv_dz), FP64, IDX order."""
        demand = np.ascontiguousarray(demand, np.float32)
        out = [np.empty(self.cells, np.float64) for _ in range(6)]
        lib().ora_soft_costs(ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float),
                             ctypes.c_double(slope), ctypes.c_double(zero_slope), ctypes.c_double(via_weight),
                             *[_p(a, ctypes.c_double) for a in out])
        return out

    def soft_net(self, wcost, vcost, nodes, parents, tau, grads=False, kind=0):
        """Soft DP value of one net (kind 0: L-shape edges, 1: Z-shape route
        family); with grads also (dF/dwcost, dF/dvcost)."""
        wcost = np.ascontiguousarray(wcost, np.float64)
        vcost = np.ascontiguousarray(vcost, np.float64)
        nodes, parents = _i32(nodes), _i32(parents)
        gw = np.zeros(self.cells, np.float64) if grads else None
        gv = np.zeros(self.cells, np.float64) if grads else None
        lib().ora_soft_net.restype = ctypes.c_double
        F = lib().ora_soft_net(ctypes.byref(self.c), _p(wcost, ctypes.c_double), _p(vcost, ctypes.c_double),
                               ctypes.c_int(len(nodes)), _p(nodes, ctypes.c_int), _p(parents, ctypes.c_int),
                               ctypes.c_double(tau), ctypes.c_int(kind), _p(gw, ctypes.c_double),
                               _p(gv, ctypes.c_double))
        return (F, gw, gv) if grads else F

    def soft_stage(self, demand, node_off, nodes, parents, tau, slope=0.5, zero_slope=1.5, via_weight=1.0,
                   grads=False, kind=0):
        """Sum of soft DP values over the nets; with grads also (gw, gv, [dslope, dzero_slope, dvia_weight])."""
        demand = np.ascontiguousarray(demand, np.float32)
        node_off, nodes, parents = _i32(node_off), _i32(nodes), _i32(parents)
        gw = np.zeros(self.cells, np.float64) if grads else None
        gv = np.zeros(self.cells, np.float64) if grads else None
        gp = np.zeros(3, np.float64) if grads else None
        lib().ora_soft_stage_kind.restype = ctypes.c_double
        F = lib().ora_soft_stage_kind(ctypes.byref(self.c), _p(self.capacity, ctypes.c_float), _p(demand, ctypes.c_float),
                                 ctypes.c_double(slope), ctypes.c_double(zero_slope), ctypes.c_double(via_weight),
                                 ctypes.c_double(tau), ctypes.c_int(kind), ctypes.c_int(len(node_off) - 1),
                                 _p(node_off, ctypes.c_int),
                                 _p(nodes, ctypes.c_int), _p(parents, ctypes.c_int), _p(gw, ctypes.c_double),
                                 _p(gv, ctypes.c_double), _p(gp, ctypes.c_double))
        return (F, gw, gv, gp) if grads else F


def batches(X, Y, pt_off, pts, hs_off, hsegs, vs_off, vsegs, limit=1000000, sparsity=1):
    arrs = [_i32(a) for a in (pt_off, pts, hs_off, hsegs, vs_off, vsegs)]
    n = len(arrs[0]) - 1
    color = np.zeros(max(n, 1), np.int32)
    nb = lib().ora_batches(ctypes.c_int(X), ctypes.c_int(Y), ctypes.c_int(n), *[_p(a, ctypes.c_int) for a in arrs],
                           ctypes.c_int(limit), ctypes.c_int(sparsity), _p(color, ctypes.c_int))
    return color[:n], nb


def fenwick_track(w):
    w = np.ascontiguousarray(w, np.float32)
    out = np.empty(len(w), np.float64)
    lib().ora_fenwick_track(_p(w, ctypes.c_float), ctypes.c_int(len(w)), _p(out, ctypes.c_double))
    return out


def dense_prefix_track(w):
    w = np.ascontiguousarray(w, np.float32)
    out = np.empty(len(w), np.float64)
    lib().ora_dense_prefix_track(_p(w, ctypes.c_float), ctypes.c_int(len(w)), _p(out, ctypes.c_double))
    return out


def expf(x):
    return lib().ora_expf(float(x))


def score(X, Y, M, nets, routes, pin_acc_num, per_pin, layer_dir, capacity, overflow_cost, x_len, y_len,
          unit_wire, unit_via):
    """Evaluator score of per-net lines (ora_score): dict of costs, rejected
    nets and the 16-bit demand per metal-layer cell."""
    out = np.zeros(5, np.float64)
    dem = np.zeros(M * X * Y, np.uint16)
    arrs = [_i32(nets[k]) for k in ("pin_off", "pins", "extra_off", "extra", "sub_off", "subs")]
    routes = _i32(routes if routes is not None else np.zeros(1, np.int32))
    acc = _i32(pin_acc_num if pin_acc_num is not None else np.zeros(1, np.int32))
    dirs, xl, yl = _i32(layer_dir), _i32(x_len), _i32(y_len)
    cap = np.ascontiguousarray(capacity, np.float64)
    ovf = np.ascontiguousarray(overflow_cost, np.float64)
    lib().ora_score(ctypes.c_int(X), ctypes.c_int(Y), ctypes.c_int(M), ctypes.c_int(len(arrs[0]) - 1),
                    *[_p(a, ctypes.c_int) for a in arrs], _p(routes, ctypes.c_int), _p(acc, ctypes.c_int),
                    ctypes.c_int(per_pin), _p(dirs, ctypes.c_int), _p(cap, ctypes.c_double),
                    _p(ovf, ctypes.c_double), _p(xl, ctypes.c_int), _p(yl, ctypes.c_int), ctypes.c_double(unit_wire),
                    ctypes.c_double(unit_via), _p(out, ctypes.c_double), _p(dem, ctypes.c_ushort))
    return dict(wirelength=out[0], via=out[1], overflow=out[2], total=out[3], rejected=int(out[4]), demand=dem)


def connectivity(X, Y, M, nets, routes, pin_acc_num, per_pin, layer_dir, access, routes2=None, pin_off2=None,
                 per_pin2=100):
    """The evaluator's completion test per net (ora_connectivity): status per
    net, 0 complete, 1 open (a rejected line), 2 not connected."""
    arrs = [_i32(nets[k]) for k in ("pin_off", "pins", "extra_off", "extra", "sub_off", "subs")]
    n = len(arrs[0]) - 1
    routes = _i32(routes if routes is not None else np.zeros(1, np.int32))
    acc = _i32(pin_acc_num if pin_acc_num is not None else np.zeros(1, np.int32))
    r2 = _i32(routes2) if routes2 is not None else None
    o2 = _i32(pin_off2) if pin_off2 is not None else None
    ap = [_i32(access[k]) for k in ("pin_off", "access_off", "access")]
    status = np.zeros(max(n, 1), np.uint8)
    lib().ora_connectivity(ctypes.c_int(X), ctypes.c_int(Y), ctypes.c_int(M), ctypes.c_int(n),
                           *[_p(a, ctypes.c_int) for a in arrs], _p(routes, ctypes.c_int), _p(acc, ctypes.c_int),
                           ctypes.c_int(per_pin), _p(r2, ctypes.c_int) if r2 is not None else None,
                           _p(o2, ctypes.c_int) if o2 is not None else None, ctypes.c_int(per_pin2),
                           _p(_i32(layer_dir), ctypes.c_int), *[_p(a, ctypes.c_int) for a in ap],
                           _p(status, ctypes.c_ubyte))
    return status[:n]


def evaluator_path():
    return os.path.join(REF_DIR, "evaluator")


def run_evaluator(cap, net, out, timeout=600):
    """Score a solution with the reference evaluator built into oracle/_ref/."""
    res = subprocess.run([evaluator_path(), cap, net, out], capture_output=True, text=True, timeout=timeout)
    scores = {}
    for line in res.stdout.splitlines():
        for key in ("wirelength cost", "via cost", "overflow cost", "total cost"):
            if line.startswith(key):
                scores[key.split()[0]] = float(line.split()[-1])
        if line.startswith("Number of open nets"):
            scores["open"] = int(line.split()[-1])
        if line.startswith("Number of incompleted nets"):
            scores["incomplete"] = int(line.split()[-1])
    return scores, res
