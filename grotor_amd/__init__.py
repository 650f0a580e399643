"""grotor_amd: the MI355X (gfx950) routing hot path of GRotor.

Python mirror of the C ABI in include/grotor_amd.h, used by the tests and the
benchmark.  All compute runs in libgrotor_amd.so (hand-written gfx950
kernels); there is no CPU fallback: without the built library every entry
point raises, and without a HIP device Context() raises.

Reference entry points mirrored here (reference tree file:line):
  Context.refresh          graph::refresh_cost         src/cost/incremental_update.cuh:98
  Context.batch_plan       generate_batches_cpu_timed  src/batching/generate_batches.hpp:14
  Context.lshape_batch     Lshape_route_warp launch    src/routing/lshape.cuh:168
  Context.commit_batch     graph::commit_batch_demand  src/routing/demand.cuh:313
  Context.lshape_run       L-shape batch loop          src/routing/lshape.cuh:148-180
  Context.ripup            graph::ripup                src/routing/congestion.cuh:114
  Context.congestion_view  extract_congestionView*     src/routing/congestion.cuh:60-112
  Context.flt_run          FLT batch loop              src/routing/flt.cuh:141-168
  Design.flt_stage         prepare_detour_topology     src/routing/flt.cuh:48-73
  Design.prepare           prepare_routing_nets etc.   src/preprocess.hpp:1448
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GROTOR_AMD_LIB") or os.path.join(_HERE, "libgrotor_amd.so")
ROUTE_PER_PIN = 12

GR_OK, GR_E_ARG, GR_E_HIP, GR_E_CAPACITY, GR_E_STATE, GR_E_DEVICE, GR_E_IO = range(7)


class GrError(RuntimeError):
    """A C ABI call returned a non-zero status."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


class GridDesc(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("X", ctypes.c_int), ("Y", ctypes.c_int), ("DIR", ctypes.c_int),
                ("unit_length_wire_cost", ctypes.c_double), ("unit_via_cost", ctypes.c_double),
                ("unit_length_short_costs", _f64p), ("x_edge_len", _i32p), ("y_edge_len", _i32p),
                ("capacity", _f32p)]


class TopologyDesc(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("X", ctypes.c_int), ("Y", ctypes.c_int), ("DIR", ctypes.c_int),
                ("nets", ctypes.c_int), ("subnets", ctypes.c_int),
                ("order", _i32p), ("node_off", _i32p), ("nodes", _i32p), ("parents", _i32p),
                ("pt_off", _i32p), ("pts", _i32p), ("hs_off", _i32p), ("hsegs", _i32p),
                ("vs_off", _i32p), ("vsegs", _i32p), ("pin_acc_num", _i32p), ("capacity", _f32p),
                ("unit_length_short_costs", _f64p), ("x_edge_len", _i32p), ("y_edge_len", _i32p),
                ("unit_length_wire_cost", ctypes.c_double), ("unit_via_cost", ctypes.c_double)]


class ZshapeStageDesc(ctypes.Structure):
    _fields_ = [("nets", ctypes.c_int), ("order", _i32p), ("pt_off", _i32p), ("pts", _i32p), ("hs_off", _i32p),
                ("hsegs", _i32p), ("vs_off", _i32p), ("vsegs", _i32p)]


class FltStageDesc(ctypes.Structure):
    _fields_ = [("nets", ctypes.c_int), ("order", _i32p), ("pt_off", _i32p), ("pts", _i32p), ("hs_off", _i32p),
                ("hsegs", _i32p), ("vs_off", _i32p), ("vsegs", _i32p), ("pin_off2", _i32p),
                ("pins2", ctypes.c_longlong), ("nodes", ctypes.c_longlong), ("edges", ctypes.c_longlong)]


class FltPlanDesc(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int), ("batch_off", _i32p), ("net_ids", _i32p),
                ("node_off", ctypes.POINTER(ctypes.c_longlong)), ("nodes", _i32p),
                ("edge_off", ctypes.POINTER(ctypes.c_longlong)), ("edges", _i32p)]


class ZshapePlanDesc(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int), ("batch_off", _i32p), ("net_ids", _i32p),
                ("rec_off", ctypes.POINTER(ctypes.c_longlong)), ("records", _i32p), ("raw_cap", _i32p)]


def lib():
    """Load libgrotor_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GrError(GR_E_STATE, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.gr_last_error.restype = ctypes.c_char_p
    return _lib


def _call(name, *args):
    status = getattr(lib(), name)(*args)
    if status != GR_OK:
        raise GrError(status, f"{name}: {lib().gr_last_error().decode()}")


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def device_name():
    buf = ctypes.create_string_buffer(256)
    _call("gr_device_name", buf, 256)
    return buf.value.decode()


class Design:
    """Host design database: CAP/NET input, synthetic instances, front end."""

    def __init__(self, handle):
        self._h = handle
        self._topo = None

    @classmethod
    def read(cls, cap, net):
        h = ctypes.c_void_p()
        _call("gr_design_read", cap.encode(), net.encode(), ctypes.byref(h))
        return cls(h)

    @classmethod
    def generate(cls, config, seed=0, net_scale=1.0, grid_side=0):
        h = ctypes.c_void_p()
        _call("gr_design_generate", ctypes.c_int(config), ctypes.c_ulonglong(seed), ctypes.c_double(net_scale),
              ctypes.c_int(grid_side), ctypes.byref(h))
        return cls(h)

    def write(self, cap=None, net=None):
        _call("gr_design_write", self._h, cap.encode() if cap else None, net.encode() if net else None)

    def prepare(self, threads=8):
        _call("gr_design_prepare", self._h, ctypes.c_int(threads))
        self._topo = None
        return self

    def topology(self):
        """Stage topology as numpy arrays (copies)."""
        if self._topo is None:
            t = TopologyDesc()
            _call("gr_design_topology", self._h, ctypes.byref(t))
            n = t.nets

            def arr(p, count, dtype=np.int32):
                if count == 0:
                    return np.zeros(0, dtype)
                return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype, copy=True)

            node_off = arr(t.node_off, n + 1)
            pt_off = arr(t.pt_off, n + 1)
            hs_off = arr(t.hs_off, n + 1)
            vs_off = arr(t.vs_off, n + 1)
            cells = t.L * t.X * t.Y
            self._topo = dict(
                L=t.L, X=t.X, Y=t.Y, DIR=t.DIR, nets=n, subnets=t.subnets,
                order=arr(t.order, n), node_off=node_off, nodes=arr(t.nodes, int(node_off[-1])),
                parents=arr(t.parents, int(node_off[-1])), pt_off=pt_off, pts=arr(t.pts, int(pt_off[-1])),
                hs_off=hs_off, hsegs=arr(t.hsegs, 2 * int(hs_off[-1])), vs_off=vs_off,
                vsegs=arr(t.vsegs, 2 * int(vs_off[-1])), pin_acc_num=arr(t.pin_acc_num, t.subnets + 1),
                capacity=arr(t.capacity, cells, np.float32),
                unit_length_short_costs=arr(t.unit_length_short_costs, t.L, np.float64),
                x_edge_len=arr(t.x_edge_len, t.X - 1), y_edge_len=arr(t.y_edge_len, t.Y - 1),
                unit_length_wire_cost=t.unit_length_wire_cost, unit_via_cost=t.unit_via_cost)
        return self._topo

    def zshape_stage(self):
        """Z-shape stage priority order and its batching geometry (numpy copies)."""
        t = ZshapeStageDesc()
        _call("gr_design_zshape_stage", self._h, ctypes.byref(t))
        n = t.nets

        def arr(p, count):
            return np.zeros(0, np.int32) if count == 0 else np.ctypeslib.as_array(p, shape=(count,)).copy()

        pt_off, hs_off, vs_off = arr(t.pt_off, n + 1), arr(t.hs_off, n + 1), arr(t.vs_off, n + 1)
        return dict(nets=n, order=arr(t.order, n), pt_off=pt_off, pts=arr(t.pts, int(pt_off[-1])), hs_off=hs_off,
                    hsegs=arr(t.hsegs, 2 * int(hs_off[-1])), vs_off=vs_off, vsegs=arr(t.vsegs, 2 * int(vs_off[-1])))

    def zshape_plan(self, color, nb, max_batch=250):
        """Bounding-box batches in execution order with encoded trees."""
        color = _i32(color)
        p = ZshapePlanDesc()
        _call("gr_design_zshape_plan", self._h, _ptr(color, ctypes.c_int), ctypes.c_int(nb), ctypes.c_int(max_batch),
              ctypes.byref(p))
        nb2 = p.batches
        batch_off = np.ctypeslib.as_array(p.batch_off, shape=(nb2 + 1,)).copy()
        rec_off = np.ctypeslib.as_array(p.rec_off, shape=(nb2 + 1,)).copy()
        nets, words = int(batch_off[-1]), int(rec_off[-1])
        take = lambda q, c: np.zeros(0, np.int32) if c == 0 else np.ctypeslib.as_array(q, shape=(c,)).copy()
        return dict(batch_off=batch_off, net_ids=take(p.net_ids, nets), rec_off=rec_off,
                    records=take(p.records, words), raw_cap=take(p.raw_cap, nets))

    def zshape_finish(self, routes, words_per_pin=256):
        """Promote congested segments and canonicalise, in place; returns the promotions."""
        assert routes.dtype == np.int32 and routes.flags.c_contiguous
        promoted = ctypes.c_int()
        _call("gr_design_zshape_finish", self._h, _ptr(routes, ctypes.c_int), ctypes.c_int(words_per_pin),
              ctypes.byref(promoted))
        return promoted.value

    def flt_stage(self, reroute, xsum=None, ysum=None, threads=8):
        """FLT stage input for the subnets flagged in reroute: priority order,
        batching geometry, phase-two pin offsets (numpy copies)."""
        reroute = np.ascontiguousarray(reroute, np.uint8)
        xs = None if xsum is None else np.ascontiguousarray(xsum, np.float32)
        ys = None if ysum is None else np.ascontiguousarray(ysum, np.float32)
        t = FltStageDesc()
        _call("gr_design_flt_stage", self._h, _ptr(reroute, ctypes.c_ubyte), _ptr(xs, ctypes.c_float),
              _ptr(ys, ctypes.c_float), ctypes.c_int(threads), ctypes.byref(t))
        n = t.nets

        def arr(p, count):
            return np.zeros(0, np.int32) if count == 0 else np.ctypeslib.as_array(p, shape=(count,)).copy()

        pt_off, hs_off, vs_off = arr(t.pt_off, n + 1), arr(t.hs_off, n + 1), arr(t.vs_off, n + 1)
        return dict(nets=n, order=arr(t.order, n), pt_off=pt_off, pts=arr(t.pts, int(pt_off[-1])), hs_off=hs_off,
                    hsegs=arr(t.hsegs, 2 * int(hs_off[-1])), vs_off=vs_off, vsegs=arr(t.vsegs, 2 * int(vs_off[-1])),
                    pin_off2=arr(t.pin_off2, len(reroute)), pins2=int(t.pins2), nodes=int(t.nodes),
                    edges=int(t.edges))

    def flt_plan(self, color, nb):
        """FLT execution plan (batches reversed) with the DAG records of every net."""
        color = _i32(color)
        p = FltPlanDesc()
        _call("gr_design_flt_plan", self._h, _ptr(color, ctypes.c_int), ctypes.c_int(nb), ctypes.byref(p))
        nb2 = p.batches
        batch_off = np.ctypeslib.as_array(p.batch_off, shape=(nb2 + 1,)).copy()
        nets = int(batch_off[-1])
        node_off = np.ctypeslib.as_array(p.node_off, shape=(nets + 1,)).copy()
        edge_off = np.ctypeslib.as_array(p.edge_off, shape=(nets + 1,)).copy()
        take = lambda q, c: np.zeros(0, np.int32) if c == 0 else np.ctypeslib.as_array(q, shape=(c,)).copy()
        return dict(batch_off=batch_off, net_ids=take(p.net_ids, nets), node_off=node_off,
                    nodes=take(p.nodes, 4 * int(node_off[-1])), edge_off=edge_off,
                    edges=take(p.edges, 2 * int(edge_off[-1])))

    def nets(self):
        """Per design net CSR arrays: pins (metal-layer cells), extra route
        pairs and subnet ids (db::net, src/input.hpp:15-24)."""
        count = ctypes.c_int()
        ptrs = [ctypes.POINTER(ctypes.c_int)() for _ in range(6)]
        _call("gr_design_nets", self._h, ctypes.byref(count), *[ctypes.byref(p) for p in ptrs])
        n = count.value

        def arr(p, k):
            return np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, np.int32)

        pin_off, extra_off, sub_off = arr(ptrs[0], n + 1), arr(ptrs[2], n + 1), arr(ptrs[4], n + 1)
        return dict(pin_off=pin_off, pins=arr(ptrs[1], int(pin_off[-1])), extra_off=extra_off,
                    extra=arr(ptrs[3], int(extra_off[-1])), sub_off=sub_off, subs=arr(ptrs[5], int(sub_off[-1])))

    def access(self):
        """Per design net its pins' access points as read from the NET file:
        dict(pin_off, access_off, access) (metal-layer cells)."""
        count = ctypes.c_int()
        ptrs = [ctypes.POINTER(ctypes.c_int)() for _ in range(3)]
        _call("gr_design_access", self._h, ctypes.byref(count), *[ctypes.byref(p) for p in ptrs])
        n = count.value
        pin_off = np.ctypeslib.as_array(ptrs[0], shape=(n + 1,)).copy()
        acc_off = np.ctypeslib.as_array(ptrs[1], shape=(int(pin_off[-1]) + 1,)).copy()
        acc = (np.ctypeslib.as_array(ptrs[2], shape=(int(acc_off[-1]),)).copy() if acc_off[-1]
               else np.zeros(0, np.int32))
        return dict(pin_off=pin_off, access_off=acc_off, access=acc)

    def metal_grid(self):
        """Metal-layer directions, capacities (float, as parsed) and overflow unit costs."""
        m = ctypes.c_int()
        _call("gr_design_metal_grid", self._h, ctypes.byref(m), None, None, None)
        t = self.topology()
        M = m.value
        dirs = np.zeros(M, np.int32)
        cap = np.zeros(M * int(t["X"]) * int(t["Y"]), np.float32)
        ovf = np.zeros(M, np.float64)
        _call("gr_design_metal_grid", self._h, ctypes.byref(m), _ptr(dirs, ctypes.c_int), _ptr(cap, ctypes.c_float),
              _ptr(ovf, ctypes.c_double))
        return dict(layers=M, dir=dirs, capacity=cap, overflow_cost=ovf)

    def write_output(self, routes, path, routes2=None, pin_off2=None, words_per_pin2=100):
        """.out solution; with routes2/pin_off2 the rerouted subnets print
        their phase-two words."""
        routes = _i32(routes)
        if routes2 is None:
            _call("gr_design_write_output", self._h, _ptr(routes, ctypes.c_int), path.encode())
            return
        routes2, pin_off2 = _i32(routes2), _i32(pin_off2)
        _call("gr_design_write_output2", self._h, _ptr(routes, ctypes.c_int), _ptr(routes2, ctypes.c_int),
              _ptr(pin_off2, ctypes.c_int), ctypes.c_int(words_per_pin2), path.encode())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gr_design_destroy(self._h)
            self._h = None


class Context:
    """Device state of one routing run (the device half of RouterState)."""

    def __init__(self, grid):
        g = grid
        self._geom_n = 0
        self.L, self.X, self.Y, self.DIR = int(g["L"]), int(g["X"]), int(g["Y"]), int(g["DIR"])
        self._keep = [np.ascontiguousarray(g["unit_length_short_costs"], np.float64), _i32(g["x_edge_len"]),
                      _i32(g["y_edge_len"]), np.ascontiguousarray(g["capacity"], np.float32)]
        desc = GridDesc(self.L, self.X, self.Y, self.DIR, float(g["unit_length_wire_cost"]),
                        float(g["unit_via_cost"]), _ptr(self._keep[0], ctypes.c_double),
                        _ptr(self._keep[1], ctypes.c_int), _ptr(self._keep[2], ctypes.c_int),
                        _ptr(self._keep[3], ctypes.c_float))
        h = ctypes.c_void_p()
        _call("gr_ctx_create", ctypes.byref(desc), ctypes.byref(h))
        self._h = h
        self.cells = self.L * self.X * self.Y
        self.route_words = 0
        self.soft_generation = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().gr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def set_routes(self, pin_acc_num, words_per_pin=ROUTE_PER_PIN):
        """Route storage of words_per_pin words per pin (12, or 256 for Z-shape stages)."""
        pin_acc_num = _i32(pin_acc_num)
        _call("gr_ctx_set_route_stride", self._h, ctypes.c_int(words_per_pin))
        _call("gr_ctx_set_routes", self._h, ctypes.c_int(len(pin_acc_num) - 1), _ptr(pin_acc_num, ctypes.c_int))
        self.route_words = words_per_pin * int(pin_acc_num[-1]) + 1
        self.words_per_pin = words_per_pin
        self.nets = len(pin_acc_num) - 1
        self.words_per_pin2 = 0

    def set_cost_params(self, slope=0.5, zero_slope=1.5, via_weight=1.0, integer_wire=False):
        _call("gr_ctx_set_cost_params", self._h, ctypes.c_double(slope), ctypes.c_double(zero_slope),
              ctypes.c_double(via_weight), ctypes.c_int(1 if integer_wire else 0))

    def set_routes_phase2(self, pin_off2, pins2, words_per_pin=100):
        """Phase-two route storage (ROUTE_PER_PIN_PHASE2 = 100 words per pin)."""
        pin_off2 = _i32(pin_off2)
        _call("gr_ctx_set_routes_phase2", self._h, _ptr(pin_off2, ctypes.c_int), ctypes.c_longlong(pins2),
              ctypes.c_int(words_per_pin))
        self.route2_words = words_per_pin * int(pins2) + 1
        self.words_per_pin2 = words_per_pin

    def reserve_routes_phase2(self, net_ids, words_per_pin=100):
        """Phase-two storage for a further rip-up stage: appends words for
        the nets in net_ids that have none, keeps the others' words; returns
        (pin_off2 per net, pins2)."""
        ids = _i32(net_ids)
        off = np.empty(max(self.nets, 1), np.int32)
        pins = ctypes.c_longlong()
        _call("gr_ctx_reserve_routes_phase2", self._h, ctypes.c_int(len(ids)), _ptr(ids, ctypes.c_int),
              ctypes.c_int(words_per_pin), _ptr(off, ctypes.c_int), ctypes.byref(pins))
        self.words_per_pin2 = max(self.words_per_pin2, words_per_pin)
        self.route2_words = self.words_per_pin2 * pins.value + 1
        return off[:self.nets], pins.value

    def routes2(self):
        r = np.empty(self.route2_words, np.int32)
        _call("gr_routes2_download", self._h, _ptr(r, ctypes.c_int), ctypes.c_longlong(self.route2_words))
        return r

    def congestion_view(self):
        """(congested bool[X*Y], xsum float[X*Y], ysum float[X*Y]) in x*Y+y order."""
        n = self.X * self.Y
        cong, xs, ys = np.empty(n, np.uint8), np.empty(n, np.float32), np.empty(n, np.float32)
        _call("gr_congestion_view", self._h, _ptr(cong, ctypes.c_ubyte), _ptr(xs, ctypes.c_float),
              _ptr(ys, ctypes.c_float))
        return cong.astype(bool), xs, ys

    def flt_upload(self, plan):
        """Upload a Design.flt_plan() plan."""
        self._fplan = plan
        _call("gr_flt_upload", self._h, ctypes.c_int(len(plan["batch_off"]) - 1),
              _ptr(_i32(plan["batch_off"]), ctypes.c_int), _ptr(_i32(plan["net_ids"]), ctypes.c_int),
              _ptr(np.ascontiguousarray(plan["node_off"], np.int64), ctypes.c_longlong),
              _ptr(_i32(plan["nodes"]), ctypes.c_int),
              _ptr(np.ascontiguousarray(plan["edge_off"], np.int64), ctypes.c_longlong),
              _ptr(_i32(plan["edges"]), ctypes.c_int))

    def flt_run(self, first_stamp=1, remove_phase_one=True, mode=0, remove=None):
        """Run the uploaded FLT batches (mode 0 plain, 1 launch, 2 persistent
        form); remove (0 none, 1 phase-one routes, 2 current routes) defaults
        to remove_phase_one; returns the DAG interval in ms."""
        if remove is None:
            remove = 1 if remove_phase_one else 0
        ms = ctypes.c_double()
        _call("gr_flt_run", self._h, ctypes.c_int(first_stamp), ctypes.c_int(remove), ctypes.c_int(mode),
              ctypes.byref(ms))
        return ms.value

    def set_flt_per_wave(self, per_wave):
        """FLT DP DAGs per wavefront for batches of <512 and >=512 DAGs."""
        pw = _i32(per_wave)
        _call("gr_ctx_set_flt_per_wave", self._h, _ptr(pw, ctypes.c_int))

    def flt_tune(self):
        """Autotune FLT DAGs per wavefront: (chosen per class, ms[class][candidate 1, 2, 4, 8])."""
        pw = np.zeros(2, np.int32)
        ms = np.zeros(8, np.float64)
        _call("gr_flt_tune", self._h, _ptr(pw, ctypes.c_int), _ptr(ms, ctypes.c_double))
        return pw.tolist(), ms.reshape(2, 4).tolist()

    def zshape_upload(self, plan):
        """Upload a Design.zshape_plan() plan."""
        self._zplan = plan
        _call("gr_zshape_upload", self._h, ctypes.c_int(len(plan["batch_off"]) - 1),
              _ptr(_i32(plan["batch_off"]), ctypes.c_int), _ptr(_i32(plan["net_ids"]), ctypes.c_int),
              _ptr(np.ascontiguousarray(plan["rec_off"], np.int64), ctypes.c_longlong),
              _ptr(_i32(plan["records"]), ctypes.c_int), _ptr(_i32(plan["raw_cap"]), ctypes.c_int))

    def zshape_run(self, first_stamp=1):
        """Run the uploaded Z-shape batches; returns the DAG interval in ms."""
        ms = ctypes.c_double()
        _call("gr_zshape_run", self._h, ctypes.c_int(first_stamp), ctypes.byref(ms))
        return ms.value

    def refresh(self, full=False):
        _call("gr_cost_refresh_full" if full else "gr_cost_refresh", self._h)

    def costs(self, wcost=True, vcost=True, presum=True):
        w = np.empty(self.cells, np.float32) if wcost else None
        v = np.empty(self.cells, np.float32) if vcost else None
        p = np.empty(self.cells, np.float64) if presum else None
        _call("gr_cost_download", self._h, _ptr(w, ctypes.c_float), _ptr(v, ctypes.c_float), _ptr(p, ctypes.c_double))
        return w, v, p

    def upload_demand(self, demand):
        d = np.ascontiguousarray(demand, np.float32)
        _call("gr_demand_upload", self._h, _ptr(d, ctypes.c_float))

    def demand(self):
        d = np.empty(self.cells, np.float32)
        _call("gr_demand_download", self._h, _ptr(d, ctypes.c_float))
        return d

    def timestamps(self):
        t = np.empty(self.cells, np.int32)
        _call("gr_timestamp_download", self._h, _ptr(t, ctypes.c_int))
        return t

    def stats(self):
        wl = ctypes.c_double()
        vias = ctypes.c_longlong()
        _call("gr_stats_download", self._h, ctypes.byref(wl), ctypes.byref(vias))
        return wl.value, vias.value

    def batch_plan(self, pt_off, pts, hs_off, hsegs, vs_off, vsegs, limit=1000000, sparsity=1):
        """Batch index of every net (priority order) and the batch count."""
        arrs = [_i32(a) for a in (pt_off, pts, hs_off, hsegs, vs_off, vsegs)]
        n = len(arrs[0]) - 1
        self._geom_n = n
        color = np.empty(max(n, 1), np.int32)
        nb = ctypes.c_int()
        _call("gr_batch_plan", self._h, ctypes.c_int(n), *[_ptr(a, ctypes.c_int) for a in arrs],
              ctypes.c_int(limit), ctypes.c_int(sparsity), _ptr(color, ctypes.c_int), ctypes.byref(nb))
        return color[:n], nb.value

    def batch_geometry(self, pt_off, pts, hs_off, hsegs, vs_off, vsegs):
        """Upload the batching geometry (kept until the next call)."""
        arrs = [_i32(a) for a in (pt_off, pts, hs_off, hsegs, vs_off, vsegs)]
        self._geom_n = len(arrs[0]) - 1
        _call("gr_batch_geometry", self._h, ctypes.c_int(self._geom_n), *[_ptr(a, ctypes.c_int) for a in arrs])

    def batch_colors(self, limit=1000000, sparsity=1):
        """Batch the uploaded geometry: (color per net, batch count)."""
        n = self._geom_n
        color = np.empty(max(n, 1), np.int32)
        nb = ctypes.c_int()
        _call("gr_batch_colors", self._h, ctypes.c_int(limit), ctypes.c_int(sparsity), _ptr(color, ctypes.c_int),
              ctypes.byref(nb))
        return color[:n], nb.value

    def lshape_upload(self, net_ids, node_off, nodes, parents):
        arrs = [_i32(a) for a in (net_ids, node_off, nodes, parents)]
        _call("gr_lshape_upload", self._h, ctypes.c_int(len(arrs[0])), *[_ptr(a, ctypes.c_int) for a in arrs])

    def lshape_batch(self, begin, count, stamp):
        _call("gr_lshape_batch", self._h, ctypes.c_int(begin), ctypes.c_int(count), ctypes.c_int(stamp))

    def lshape_batch_noalloc(self, begin, count, stamp):
        """gr_lshape_batch without device allocation (library-owned scratch)."""
        _call("gr_lshape_batch_noalloc", self._h, ctypes.c_int(begin), ctypes.c_int(count), ctypes.c_int(stamp))

    def commit_batch(self, begin, count, stamp, K=1):
        _call("gr_commit_batch", self._h, ctypes.c_int(begin), ctypes.c_int(count), ctypes.c_int(stamp),
              ctypes.c_int(K))

    def lshape_run(self, batch_off, first_stamp=1, mode=0):
        """Run the batch loop; returns the synchronized interval in ms."""
        batch_off = _i32(batch_off)
        ms = ctypes.c_double()
        _call("gr_lshape_run", self._h, ctypes.c_int(len(batch_off) - 1), _ptr(batch_off, ctypes.c_int),
              ctypes.c_int(first_stamp), ctypes.c_int(mode), ctypes.byref(ms))
        return ms.value

    def set_nets_per_wave(self, per_wave):
        """L-shape DP nets per wavefront for the tree size classes (<=16 nodes, larger)."""
        pw = _i32(per_wave)
        _call("gr_ctx_set_nets_per_wave", self._h, _ptr(pw, ctypes.c_int))

    def lshape_tune(self, batch_off):
        """Autotune nets per wavefront over the uploaded stage's batches:
        (chosen per class, ms[class][candidate 1, 2, 4, 8] with -1 where not tried)."""
        batch_off = _i32(batch_off)
        pw = np.zeros(2, np.int32)
        ms = np.zeros(8, np.float64)
        _call("gr_lshape_tune", self._h, ctypes.c_int(len(batch_off) - 1), _ptr(batch_off, ctypes.c_int),
              _ptr(pw, ctypes.c_int), _ptr(ms, ctypes.c_double))
        return pw.tolist(), ms.reshape(2, 4).tolist()

    def lshape_tune_groups(self, batch_off):
        """Autotune the persistent form's group shape (the tiny-tree bound):
        (chosen bound, ms for bounds 1, 2, 4, 8)."""
        batch_off = _i32(batch_off)
        chosen = ctypes.c_int()
        ms = np.zeros(4, np.float64)
        _call("gr_lshape_tune_groups", self._h, ctypes.c_int(len(batch_off) - 1), _ptr(batch_off, ctypes.c_int),
              ctypes.byref(chosen), _ptr(ms, ctypes.c_double))
        return chosen.value, ms.tolist()

    def set_tiny_tree(self, nodes):
        _call("gr_ctx_set_tiny_tree", self._h, ctypes.c_int(nodes))

    def lshape_tune_tiny(self, batch_off):
        """Autotune the persistent form's tiny-tree group's nets per
        wavefront: (chosen, ms for 1, 2, 4, 8)."""
        batch_off = _i32(batch_off)
        chosen = ctypes.c_int()
        ms = np.zeros(4, np.float64)
        _call("gr_lshape_tune_tiny", self._h, ctypes.c_int(len(batch_off) - 1), _ptr(batch_off, ctypes.c_int),
              ctypes.byref(chosen), _ptr(ms, ctypes.c_double))
        return chosen.value, ms.tolist()

    def set_tiny_per_wave(self, per_wave):
        _call("gr_ctx_set_tiny_per_wave", self._h, ctypes.c_int(per_wave))

    def set_flt_split(self, nbytes):
        """LDS-share bound (bytes) between the FLT DP's two DAG classes, for
        the next flt_upload (0: one class)."""
        _call("gr_ctx_set_flt_split", self._h, ctypes.c_int(nbytes))

    def score(self, nets, layer_dir, capacity, overflow_cost, demand=False):
        """Evaluator score of the resident routes (gr_score): dict of
        wirelength, via, overflow and total cost, rejected nets, device ms,
        and the 16-bit demand per metal-layer cell when asked."""
        out = np.zeros(6, np.float64)
        n = len(nets["pin_off"]) - 1
        arrs = [_i32(nets[k]) for k in ("pin_off", "pins", "extra_off", "extra", "sub_off", "subs")]
        dirs = _i32(layer_dir)
        cap = np.ascontiguousarray(capacity, np.float64)
        ovf = np.ascontiguousarray(overflow_cost, np.float64)
        dem = np.zeros(len(cap), np.uint16) if demand else None
        _call("gr_score", self._h, ctypes.c_int(n), *[_ptr(a, ctypes.c_int) for a in arrs], _ptr(dirs, ctypes.c_int),
              _ptr(cap, ctypes.c_double), _ptr(ovf, ctypes.c_double), _ptr(out, ctypes.c_double),
              _ptr(dem, ctypes.c_ushort))
        r = dict(wirelength=out[0], via=out[1], overflow=out[2], total=out[3], rejected=int(out[4]), ms=out[5])
        if demand:
            r["demand"] = dem
        return r

    def connectivity(self, nets, layer_dir, access):
        """The evaluator's completion test of every net (gr_check_connectivity):
        (status per net: 0 complete, 1 open, 2 not connected; counts of open,
        not connected and incomplete nets)."""
        n = len(nets["pin_off"]) - 1
        arrs = [_i32(nets[k]) for k in ("pin_off", "pins", "extra_off", "extra", "sub_off", "subs")]
        acc = [_i32(access[k]) for k in ("pin_off", "access_off", "access")]
        status = np.zeros(max(n, 1), np.uint8)
        counts = np.zeros(3, np.int64)
        _call("gr_check_connectivity", self._h, ctypes.c_int(n), *[_ptr(a, ctypes.c_int) for a in arrs],
              _ptr(_i32(layer_dir), ctypes.c_int), *[_ptr(a, ctypes.c_int) for a in acc],
              _ptr(status, ctypes.c_ubyte), _ptr(counts, ctypes.c_longlong))
        return status[:n], dict(open=int(counts[0]), disconnected=int(counts[1]), incomplete=int(counts[2]))

    def routes(self):
        r = np.empty(self.route_words, np.int32)
        _call("gr_routes_download", self._h, _ptr(r, ctypes.c_int), ctypes.c_longlong(self.route_words))
        return r

    def reset(self):
        _call("gr_ctx_reset", self._h)

    KINDS = 10

    def profile_begin(self):
        _call("gr_profile_begin", self._h)

    def profile_select(self, kinds=None):
        """Give events only to the kernel classes listed (None: all)."""
        mask = 0xFFFFFFFF if kinds is None else sum(1 << int(k) for k in kinds)
        _call("gr_profile_select", self._h, ctypes.c_uint(mask))

    def profile_end(self):
        """Per kernel class: summed device ms, launches and algorithmic bytes."""
        ms = np.zeros(self.KINDS, np.float64)
        n = np.zeros(self.KINDS, np.int64)
        by = np.zeros(self.KINDS, np.float64)
        _call("gr_profile_end", self._h, _ptr(ms, ctypes.c_double), _ptr(n, ctypes.c_longlong),
              _ptr(by, ctypes.c_double), ctypes.c_int(self.KINDS))
        return {"ms": ms.tolist(), "launches": n.tolist(), "bytes": by.tolist()}

    def soft_forward(self, begin, count, slope=0.5, zero_slope=1.5, via_weight=1.0, tau=1.0, precision=32,
                     per_net=False):
        """Soft L-shape DP over stage slots [begin, begin + count): (total, per-net values or None, ms)."""
        total, ms = ctypes.c_double(), ctypes.c_double()
        vals = np.empty(max(count, 1), np.float64) if per_net else None
        self.soft_generation += 1
        _call("gr_soft_forward", self._h, ctypes.c_int(begin), ctypes.c_int(count), ctypes.c_double(slope),
              ctypes.c_double(zero_slope), ctypes.c_double(via_weight), ctypes.c_double(tau), ctypes.c_int(precision),
              ctypes.byref(total), _ptr(vals, ctypes.c_double), ctypes.byref(ms))
        return total.value, (vals[:count] if per_net else None), ms.value

    def soft_forward_costs(self, begin, count, wcost, vcost, tau=1.0, precision=64, per_net=False):
        """Soft DP over stage slots [begin, begin + count) at given wire and via
        costs (IDX order): (total, per-net values or None, ms)."""
        total, ms = ctypes.c_double(), ctypes.c_double()
        vals = np.empty(max(count, 1), np.float64) if per_net else None
        w = np.ascontiguousarray(wcost, np.float64)
        v = np.ascontiguousarray(vcost, np.float64)
        if w.size != self.cells or v.size != self.cells:
            raise ValueError("cost arrays must hold one value per routing cell")
        self.soft_generation += 1
        _call("gr_soft_forward_costs", self._h, ctypes.c_int(begin), ctypes.c_int(count), _ptr(w, ctypes.c_double),
              _ptr(v, ctypes.c_double), ctypes.c_double(tau), ctypes.c_int(precision), ctypes.byref(total),
              _ptr(vals, ctypes.c_double), ctypes.byref(ms))
        return total.value, (vals[:count] if per_net else None), ms.value

    def set_soft_kind(self, kind):
        """Edges of the soft DP: 0 the L-shape DP's, 1 the Z-shape route family."""
        _call("gr_ctx_set_soft_kind", self._h, ctypes.c_int(kind))
        self.soft_generation += 1

    def soft_backward(self):
        """Adjoint of the last soft_forward: (d/d(slope, zero_slope, via_weight), ms)."""
        gp = np.zeros(3, np.float64)
        ms = ctypes.c_double()
        _call("gr_soft_backward", self._h, _ptr(gp, ctypes.c_double), ctypes.byref(ms))
        return gp, ms.value

    def soft_edge_grads(self):
        """Per-cell d total / d wcost and d total / d vcost (IDX order)."""
        gw = np.empty(self.cells, np.float64)
        gv = np.empty(self.cells, np.float64)
        _call("gr_soft_edge_grads", self._h, _ptr(gw, ctypes.c_double), _ptr(gv, ctypes.c_double))
        return gw, gv

    def ripup(self, net_count, threshold=0):
        flags = np.empty(max(net_count, 1), np.uint8)
        _call("gr_ripup", self._h, ctypes.c_int(threshold), _ptr(flags, ctypes.c_ubyte))
        return flags[:net_count]


def batch_plan_host(X, Y, pt_off, pts, hs_off, hsegs, vs_off, vsegs, limit=1000000, sparsity=1):
    """The same first-fit batches computed on the host (the reference's CPU
    batching fallback): (color per net, batch count).  Needs no device."""
    arrs = [_i32(a) for a in (pt_off, pts, hs_off, hsegs, vs_off, vsegs)]
    n = len(arrs[0]) - 1
    color = np.empty(max(n, 1), np.int32)
    nb = ctypes.c_int()
    _call("gr_batch_plan_host", ctypes.c_int(int(X)), ctypes.c_int(int(Y)), ctypes.c_int(n),
          *[_ptr(a, ctypes.c_int) for a in arrs], ctypes.c_int(limit), ctypes.c_int(sparsity),
          _ptr(color, ctypes.c_int), ctypes.byref(nb))
    return color[:n], nb.value


def batches_from_colors(color, nb):
    """Reference batch lists (vector<vector<int>>) from per-slot batch indices:
    batch b holds the priority slots with color b in priority order."""
    color = np.asarray(color)
    order = np.argsort(color, kind="stable")
    counts = np.bincount(color, minlength=nb)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return [order[off[b]:off[b + 1]] for b in range(nb)]


def execution_schedule(topo, color, nb):
    """Stage arrays in execution order: batches reversed (last color first,
    src/routing/lshape.cuh:70) and nets grouped by batch (:72-77)."""
    lists = batches_from_colors(color, nb)[::-1]
    slots = np.concatenate(lists) if lists else np.zeros(0, np.int64)
    batch_off = np.concatenate([[0], np.cumsum([len(b) for b in lists])]).astype(np.int32)
    node_off = topo["node_off"]
    sizes = node_off[slots + 1] - node_off[slots]
    new_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = np.concatenate([np.arange(node_off[s], node_off[s + 1]) for s in slots]) if len(slots) else np.zeros(0, np.int64)
    return dict(net_ids=topo["order"][slots].astype(np.int32), node_off=new_off,
                nodes=topo["nodes"][idx].astype(np.int32), parents=topo["parents"][idx].astype(np.int32),
                batch_off=batch_off, slots=slots)
