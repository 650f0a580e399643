"""gr_lshape_batch_noalloc (the DP of one batch without device allocation,
its scratch in a library-owned device array) against gr_lshape_batch on the
same inputs: route words identical, called through the C ABI on two
contexts, for batches of small trees (the persistent form, whose queue
counters the previous call left non-zero) and of trees too large for LDS
(global tables in the array), each followed by a smaller call."""
import ctypes

import numpy as np
import pytest

from tests._instances import random_demand, topology
from tests.test_gpu_parity import _large_tree_stage

pytestmark = pytest.mark.gpu


def _pair(ga, t, pin_acc_num, sched, seed):
    """Two contexts with the same demand, costs, routes storage and stage."""
    demand = random_demand(t, seed)
    out = []
    for _ in range(2):
        ctx = ga.Context(t)
        ctx.set_routes(pin_acc_num)
        ctx.upload_demand(demand)
        ctx.refresh()
        ctx.lshape_upload(sched["net_ids"], sched["node_off"], sched["nodes"], sched["parents"])
        out.append(ctx)
    return out


def _both(ga, a, b, pin_acc_num, begin, count, tag):
    print(f"no-alloc {tag}: slots [{begin}, {begin + count})", flush=True)
    for ctx in (a, b):
        ctx.set_routes(pin_acc_num)  # cleared route words: each call must write its own
    before = b.routes()
    assert ga.lib().gr_lshape_batch(a._h, ctypes.c_int(begin), ctypes.c_int(count), ctypes.c_int(1)) == 0, \
        ga.lib().gr_last_error()
    assert ga.lib().gr_lshape_batch_noalloc(b._h, ctypes.c_int(begin), ctypes.c_int(count), ctypes.c_int(1)) == 0, \
        ga.lib().gr_last_error()
    ra, rb = a.routes(), b.routes()
    assert not np.array_equal(rb, before), tag  # the call wrote route words
    assert np.array_equal(ra, rb), tag


def test_noalloc_batch_equals_lshape_batch(ga, ora):
    _, t = topology(6, 2)
    color, nb = ora.batches(t["X"], t["Y"], t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])
    sched = ga.execution_schedule(t, color, nb)
    bo = sched["batch_off"]
    sizes = np.diff(bo)
    big, small = int(np.argmax(sizes)), int(np.argmin(sizes))
    assert sizes[big] > sizes[small] > 0
    a, b = _pair(ga, t, t["pin_acc_num"], sched, 31)
    _both(ga, a, b, t["pin_acc_num"], int(bo[big]), int(sizes[big]), "small trees, largest batch")
    _both(ga, a, b, t["pin_acc_num"], int(bo[small]), int(sizes[small]), "small trees, smallest batch")
    a.close()
    b.close()

    # trees beyond LDS (300 and 240 nodes on four layers): tables in the array
    lt = _large_tree_stage(t, [300, 12, 240, 60, 3, 97], 11)
    pin_acc_num = np.arange(0, 250 * 7, 250, dtype=np.int32)
    a, b = _pair(ga, t, pin_acc_num, lt, 32)
    _both(ga, a, b, pin_acc_num, 0, 6, "large trees")
    _both(ga, a, b, pin_acc_num, 0, 2, "large trees, fewer")
    _both(ga, a, b, pin_acc_num, 3, 2, "medium trees after large ones")
    a.close()
    b.close()


def test_noalloc_batch_rejects_bad_ranges(ga):
    _, t = topology(6, 2)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    with pytest.raises(ga.GrError) as e:
        ctx.lshape_batch_noalloc(0, 1, 1)
    assert e.value.code == ga.GR_E_STATE  # no stage topology
    ctx.lshape_upload(t["order"][:4], t["node_off"][:5], t["nodes"][:t["node_off"][4]],
                      t["parents"][:t["node_off"][4]])
    with pytest.raises(ga.GrError) as e:
        ctx.lshape_batch_noalloc(2, 5, 1)
    assert e.value.code == ga.GR_E_ARG
    ctx.lshape_batch_noalloc(0, 0, 1)  # an empty batch is a no-op
    ctx.close()
