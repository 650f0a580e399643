"""C ABI and host front end, on the CPU: the library loads, exports every
function include/grotor_amd.h declares, reports errors through status codes,
and the front end (generator, CAP/NET I/O, preparation) is deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "grotor_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gr_\w+)\s*\(", text)))


def test_library_exports_every_declared_function(ga):
    lib = ga.lib()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_oracle_library_is_separate(ga, ora):
    """The product library neither links nor exports the oracle."""
    lib = ga.lib()
    assert not hasattr(lib, "ora_lshape_net")
    import subprocess
    out = subprocess.run(["ldd", ga.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out


def test_status_codes_and_messages(ga):
    lib = ga.lib()
    h = ctypes.c_void_p()
    assert lib.gr_design_read(b"/nonexistent.cap", b"/nonexistent.net", ctypes.byref(h)) == ga.GR_E_IO
    assert b"failed to open" in lib.gr_last_error()
    assert lib.gr_design_generate(99, 0, ctypes.c_double(1.0), 0, ctypes.byref(h)) != ga.GR_OK
    assert lib.gr_ctx_create(None, ctypes.byref(h)) == ga.GR_E_ARG
    with pytest.raises(ga.GrError):
        ga.Design.generate(1, net_scale=-1.0)


def test_context_needs_a_device(ga):
    """Without a HIP device the context refuses to start: no CPU fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    d = ga.Design.generate(6, seed=0, net_scale=0.05).prepare(2)
    with pytest.raises(ga.GrError) as e:
        ga.Context(d.topology())
    assert e.value.code == ga.GR_E_HIP


def test_generator_is_deterministic(ga, tmp_path):
    paths = []
    for k in range(2):
        d = ga.Design.generate(1, seed=7, net_scale=0.1)
        cap, net = str(tmp_path / f"a{k}.cap"), str(tmp_path / f"a{k}.net")
        d.write(cap, net)
        paths.append((open(cap).read(), open(net).read()))
    assert paths[0] == paths[1]
    other = ga.Design.generate(1, seed=8, net_scale=0.1)
    other.write(str(tmp_path / "b.cap"), str(tmp_path / "b.net"))
    assert open(tmp_path / "b.net").read() != paths[0][1]


def test_cap_net_round_trip(ga, tmp_path):
    """Writing and re-reading an instance gives the same prepared topology."""
    a = ga.Design.generate(6, seed=3)
    cap, net = str(tmp_path / "x.cap"), str(tmp_path / "x.net")
    a.write(cap, net)
    b = ga.Design.read(cap, net)
    ta, tb = a.prepare(4).topology(), b.prepare(4).topology()
    for k in ("order", "node_off", "nodes", "parents", "pts", "hsegs", "vsegs", "pin_acc_num", "capacity"):
        assert np.array_equal(ta[k], tb[k]), k
    assert ta["DIR"] == tb["DIR"] == 1


def test_prepared_topology_shapes(ga):
    d = ga.Design.generate(1, seed=0, net_scale=0.2).prepare(4)
    t = d.topology()
    n = t["nets"]
    assert len(t["order"]) == n and sorted(t["order"]) == list(range(t["subnets"]))
    no = t["node_off"]
    for k in range(0, n, max(1, n // 50)):
        par = t["parents"][no[k]:no[k + 1]]
        assert par[0] == -1 and all(0 <= par[i] < i for i in range(1, len(par)))
    # hpwl order: descending
    X, Y = t["X"], t["Y"]
    assert t["capacity"].shape == (t["L"] * X * Y,)
