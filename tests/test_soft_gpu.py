"""GPU soft (log-sum-exp) DP, over L-shape edges and over the Z-shape route
family, against the FP64 oracle (oracle/soft.c) on the 64x64x5, 2k-net
instance: values, per-cell gradients, parameter gradients against FP64
central finite differences (h = 1e-4), and the torch.autograd.Functions
(over the cost-model parameters, and over the per-cell cost tensors, whose
gradients are checked against FP64 central differences of the GPU forward,
h = 1e-4, relative error at most 1e-4)."""
import numpy as np
import pytest

from tests._instances import random_demand, topology

pytestmark = pytest.mark.gpu

TAUS = (0.1, 1.0, 10.0)
THETA = (0.5, 1.5, 1.0)


def _setup(ga, seed=0):
    _, t = topology(6, seed)
    ctx = ga.Context(t)
    ctx.set_routes(t["pin_acc_num"])
    ctx.lshape_upload(t["order"], t["node_off"], t["nodes"], t["parents"])
    demand = random_demand(t, seed + 11, high=4)
    ctx.upload_demand(demand)
    return t, ctx, demand


def _fd(grid, t, demand, tau, j, h=1e-4, kind=0):
    tp, tm = list(THETA), list(THETA)
    tp[j] += h
    tm[j] -= h
    args = (demand, t["node_off"], t["nodes"], t["parents"], tau)
    return (grid.soft_stage(*args, *tp, kind=kind) - grid.soft_stage(*args, *tm, kind=kind)) / (2 * h)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("tau", TAUS)
def test_soft_fp64_matches_oracle(ga, ora, tau, kind):
    t, ctx, demand = _setup(ga)
    n = t["nets"]
    print(f"soft fp64 kind={kind} tau={tau} nets={n}", flush=True)
    grid = ora.Grid(t)
    F_o, gw_o, gv_o, gp_o = grid.soft_stage(demand, t["node_off"], t["nodes"], t["parents"], tau, *THETA, grads=True,
                                            kind=kind)
    ctx.set_soft_kind(kind)
    total, per_net, _ = ctx.soft_forward(0, n, *THETA, tau=tau, precision=64, per_net=True)
    assert abs(total - F_o) <= 1e-10 * abs(F_o)
    gp, _ = ctx.soft_backward()
    gw, gv = ctx.soft_edge_grads()
    assert np.allclose(gw, gw_o, rtol=1e-8, atol=1e-10)
    assert np.allclose(gv, gv_o, rtol=1e-8, atol=1e-10)
    assert np.allclose(gp, gp_o, rtol=1e-8, atol=1e-10)
    for j in range(3):
        fd = _fd(grid, t, demand, tau, j, kind=kind)
        assert abs(gp[j] - fd) <= 1e-4 * abs(fd), (j, gp[j], fd)
    ctx.close()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("tau", TAUS)
def test_soft_fp32_backward_matches_finite_differences(ga, ora, tau, kind):
    """FP32 tables: relative error of d total / d params against FP64 central
    differences of the oracle, bound 1e-4."""
    t, ctx, demand = _setup(ga, 1)
    n = t["nets"]
    print(f"soft fp32 kind={kind} tau={tau} nets={n}", flush=True)
    grid = ora.Grid(t)
    F_o = grid.soft_stage(demand, t["node_off"], t["nodes"], t["parents"], tau, *THETA, kind=kind)
    ctx.set_soft_kind(kind)
    total, _, _ = ctx.soft_forward(0, n, *THETA, tau=tau, precision=32)
    assert abs(total - F_o) <= 1e-4 * abs(F_o)
    gp, _ = ctx.soft_backward()
    errs = []
    for j in range(3):
        fd = _fd(grid, t, demand, tau, j, kind=kind)
        errs.append(abs(gp[j] - fd) / abs(fd))
    print("fp32 relative errors", errs, flush=True)
    assert max(errs) <= 1e-4, errs
    ctx.close()


def test_soft_autograd_function(ga, ora):
    torch = pytest.importorskip("torch")
    from grotor_amd.soft import soft_lshape_cost
    t, ctx, demand = _setup(ga, 2)
    n = t["nets"]
    params = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    total = soft_lshape_cost(ctx, params, 0, n, tau=1.0, precision=64)
    total.backward()
    grid = ora.Grid(t)
    _, _, _, gp_o = grid.soft_stage(demand, t["node_off"], t["nodes"], t["parents"], 1.0, *THETA, grads=True)
    assert np.allclose(params.grad.numpy(), gp_o, rtol=1e-8)
    # a later forward invalidates an earlier graph's tables
    p2 = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    first = soft_lshape_cost(ctx, p2, 0, n, tau=1.0, precision=32)
    soft_lshape_cost(ctx, p2.detach(), 0, n, tau=1.0, precision=32)
    with pytest.raises(RuntimeError):
        first.backward()
    ctx.close()


def test_soft_batches_sum_to_stage(ga, ora):
    """Forward over two halves equals the whole (per-net values are independent)."""
    t, ctx, demand = _setup(ga, 3)
    n = t["nets"]
    whole, per, _ = ctx.soft_forward(0, n, *THETA, tau=1.0, precision=64, per_net=True)
    a, pa, _ = ctx.soft_forward(0, n // 2, *THETA, tau=1.0, precision=64, per_net=True)
    b, pb, _ = ctx.soft_forward(n // 2, n - n // 2, *THETA, tau=1.0, precision=64, per_net=True)
    assert np.array_equal(np.concatenate([pa, pb]), per)
    assert abs(a + b - whole) <= 1e-12 * abs(whole)
    ctx.close()


@pytest.mark.parametrize("kind", [0, 1])
def test_soft_dag_cost_gradients_match_finite_differences(ga, ora, kind):
    """SoftDagCost over the per-cell cost tensors, FP64: the autograd
    gradients equal the oracle's adjoint, and on cells the nets use they
    agree with central differences of the GPU forward (h = 1e-4) within 1e-4
    relative; a cell no net touches has gradient 0."""
    torch = pytest.importorskip("torch")
    from grotor_amd.soft import soft_dag_cost
    t, ctx, demand = _setup(ga, 4)
    # a slice of the nets keeps F small enough that the central difference's
    # rounding error (eps * |F| / h) stays far below the 1e-4 bound
    n = 64
    print(f"soft dag cost kind={kind} nets={n}", flush=True)
    grid = ora.Grid(t)
    w0, v0, *_ = grid.soft_costs(demand, *THETA)
    tau = 1.0
    wt = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    vt = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    total = soft_dag_cost(ctx, wt, vt, 0, n, tau=tau, precision=64, kind=kind)
    total.backward()
    gw, gv = wt.grad.numpy(), vt.grad.numpy()
    off = t["node_off"][:n + 1]
    F_o, gw_o, gv_o, _ = grid.soft_stage(demand, off, t["nodes"][:off[-1]], t["parents"][:off[-1]], tau, *THETA,
                                         grads=True, kind=kind)
    assert abs(total.item() - F_o) <= 1e-10 * abs(F_o)
    assert np.allclose(gw, gw_o, rtol=1e-8, atol=1e-10) and np.allclose(gv, gv_o, rtol=1e-8, atol=1e-10)
    rng = np.random.default_rng(9)
    h = 1e-4
    for name, base, grad in (("w", w0, gw), ("v", v0, gv)):
        used = np.flatnonzero(np.abs(grad) > 1e-2)
        for cell in rng.choice(used, size=min(8, len(used)), replace=False):
            a = base.copy()
            a[cell] += h
            fp = ctx.soft_forward_costs(0, n, a if name == "w" else w0, a if name == "v" else v0, tau, 64)[0]
            a[cell] -= 2 * h
            fm = ctx.soft_forward_costs(0, n, a if name == "w" else w0, a if name == "v" else v0, tau, 64)[0]
            fd = (fp - fm) / (2 * h)
            rounding = 8 * np.finfo(np.float64).eps * abs(F_o) / h
            assert abs(fd - grad[cell]) <= 1e-4 * abs(fd) + rounding, (name, cell, fd, grad[cell], rounding)
        unused = np.flatnonzero(grad == 0)
        assert len(unused) > 0
    ctx.close()
