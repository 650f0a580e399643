"""Host first-fit batching (gr_batch_plan_host, the product's `--batching
cpu` path) against the reference's own batch lists on the golden instances
and against the oracle's greedy on synthetic instances, with batch limits
and sparsities; argument errors come back as status codes."""
import numpy as np
import pytest

from tests._instances import topology


def _geom(t):
    return (t["pt_off"], t["pts"], t["hs_off"], t["hsegs"], t["vs_off"], t["vsegs"])


@pytest.mark.parametrize("name", ["syn10", "mempool"])
def test_host_batches_equal_reference_lists(ga, name):
    from tests import _golden as G
    t = G.topology(name)
    color, nb = ga.batch_plan_host(t["X"], t["Y"], *_geom(t))
    ref_color, ref_nb = G.reference_colors(name)
    assert nb == ref_nb and np.array_equal(color, ref_color)


@pytest.mark.parametrize("config,seed,limit,sparsity", [(1, 0, 1000000, 1), (1, 3, 40, 1), (1, 1, 150, 2),
                                                         (6, 2, 1000000, 5), (6, 4, 7, 1)])
def test_host_batches_equal_oracle(ga, ora, config, seed, limit, sparsity):
    _, t = topology(config, seed)
    color, nb = ga.batch_plan_host(t["X"], t["Y"], *_geom(t), limit=limit, sparsity=sparsity)
    ocolor, onb = ora.batches(t["X"], t["Y"], *_geom(t), limit=limit, sparsity=sparsity)
    assert nb == onb and np.array_equal(color, ocolor)
    assert np.bincount(color).max() <= limit


def test_host_batches_beyond_one_mask_word(ga):
    """150 nets on one cell: one batch each, across three 64-batch planes;
    a limit of 1 gives the same; a point off the grid is an argument error."""
    X, Y, k = 8, 8, 150
    off = np.arange(k + 1, dtype=np.int32)
    pts = np.full(k, 3 * Y + 4, np.int32)
    z = np.zeros(k + 1, np.int32)
    e = np.zeros(0, np.int32)
    color, nb = ga.batch_plan_host(X, Y, off, pts, z, e, z, e)
    assert nb == k and np.array_equal(color, np.arange(k))
    # disjoint nets, limit 1: still one batch per net
    pts2 = np.arange(k, dtype=np.int32) % (X * Y)
    color, nb = ga.batch_plan_host(X, Y, off, pts2, z, e, z, e, limit=1)
    assert nb == k and np.array_equal(color, np.arange(k))
    empty = np.zeros(4, np.int32)
    color, nb = ga.batch_plan_host(X, Y, empty, e, empty, e, empty, e)
    assert nb == 1 and np.array_equal(color, np.zeros(3, np.int32))
    with pytest.raises(ga.GrError) as err:
        ga.batch_plan_host(X, Y, off, np.full(k, X * Y, np.int32), z, e, z, e)
    assert err.value.code == ga.GR_E_ARG
    with pytest.raises(ga.GrError):
        ga.batch_plan_host(X, Y, off, pts, z, e, z, e, sparsity=0)
