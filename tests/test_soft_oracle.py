"""The FP64 soft-DP oracle (oracle/soft.c), on the CPU: its adjoint against
central finite differences of its own forward, against torch autograd of an
independent dense torch restatement, and its tau -> 0 limit against the
hard DP.  These pin the oracle; the GPU soft DP is compared with it in
test_soft_gpu.py."""
import numpy as np
import pytest

from tests._instances import random_demand, topology


def _nets(t, count, min_nodes=2, max_nodes=12):
    off = t["node_off"]
    out = []
    for k in range(t["nets"]):
        n = off[k + 1] - off[k]
        if min_nodes <= n <= max_nodes:
            out.append((t["nodes"][off[k]:off[k + 1]], t["parents"][off[k]:off[k + 1]]))
        if len(out) == count:
            break
    return out


def _costs(ora, t, seed):
    # moderate demand keeps F small enough that FP64 differences of F
    # resolve the gradient
    g = ora.Grid(t)
    w, v, *_ = g.soft_costs(random_demand(t, seed, high=4))
    return g, w, v


def _fd_tol(F, h):
    """Rounding error of a central difference of F (FP64) with step h."""
    return 8 * np.finfo(np.float64).eps * max(1.0, abs(F)) / h


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("tau", [0.1, 1.0, 10.0])
def test_soft_adjoint_matches_finite_differences(ora, tau, kind):
    _, t = topology(6, 0)
    g, w, v = _costs(ora, t, 1)
    rng = np.random.default_rng(7)
    checked = {"w": 0, "v": 0}
    for nodes, parents in _nets(t, 8, min_nodes=3):
        F, gw, gv = g.soft_net(w, v, nodes, parents, tau, grads=True, kind=kind)
        for name, arr, grad in (("w", w, gw), ("v", v, gv)):
            used = np.flatnonzero(np.abs(grad) > 1e-3)
            checked[name] += min(6, len(used))
            for cell in rng.choice(used, size=min(6, len(used)), replace=False):
                h = 1e-5 * max(1.0, abs(arr[cell]))
                a = arr.copy()
                a[cell] += h
                fp = g.soft_net(a if arr is w else w, a if arr is v else v, nodes, parents, tau, kind=kind)
                a[cell] -= 2 * h
                fm = g.soft_net(a if arr is w else w, a if arr is v else v, nodes, parents, tau, kind=kind)
                fd = (fp - fm) / (2 * h)
                assert abs(fd - grad[cell]) <= 1e-6 * max(1.0, abs(fd)) + _fd_tol(F, h), (cell, fd, grad[cell])
    assert checked["w"] >= 20 and checked["v"] >= 10, checked


@pytest.mark.parametrize("kind", [0, 1])
def test_soft_gradients_are_expected_usage(ora, kind):
    """dF/dcost is a probability-weighted count: nonnegative, and on a
    two-pin net the wire gradients sum to the expected number of unit edges,
    the Manhattan length (every route of both families is monotone)."""
    _, t = topology(6, 1)
    g, w, v = _costs(ora, t, 2)
    for nodes, parents in _nets(t, 20, min_nodes=2, max_nodes=2):
        F, gw, gv = g.soft_net(w, v, nodes, parents, 1.0, grads=True, kind=kind)
        assert gw.min() >= -1e-12 and gv.min() >= -1e-12
        a, b = int(nodes[0]), int(nodes[1])
        Y, X = t["Y"], t["X"]
        manhattan = abs(a // Y % X - b // Y % X) + abs(a % Y - b % Y)
        assert abs(gw.sum() - manhattan) < 1e-9


def test_soft_tends_to_hard_dp(ora):
    """As tau -> 0 the soft value tends to the hard DP's optimum from below,
    within tau * log(number of alternatives)."""
    _, t = topology(6, 2)
    grid = ora.Grid(t)
    demand = random_demand(t, 3, high=12)
    w32, v32, p = grid.refresh(demand)
    w = w32.astype(np.float64)
    v = v32.astype(np.float64)
    for nodes, parents in _nets(t, 15, min_nodes=2, max_nodes=8):
        rc, _, hard = grid.lshape_net(v32, p, nodes, parents, with_cost=True)
        assert rc == 0
        for tau in (1e-2, 1e-3):
            soft = grid.soft_net(w, v, nodes, parents, tau)
            bound = tau * np.log(64.0) * len(nodes) * 3
            assert soft <= hard + 1e-6 * abs(hard)
            assert hard - soft <= bound + 1e-6 * abs(hard), (hard, soft, tau)


def test_soft_parameter_gradients_match_finite_differences(ora):
    """dF/d(slope, zero_slope, via_weight) over a batch of nets against FP64
    central differences with h = 1e-4."""
    _, t = topology(6, 3)
    grid = ora.Grid(t)
    demand = random_demand(t, 4, high=4)
    k = 200
    off = t["node_off"][:k + 1]
    nodes, parents = t["nodes"][:off[-1]], t["parents"][:off[-1]]
    theta = np.array([0.5, 1.5, 1.0])
    for tau in (0.1, 1.0, 10.0):
        F, gw, gv, gp = grid.soft_stage(demand, off, nodes, parents, tau, *theta, grads=True)
        for j in range(3):
            h = 1e-4
            tp, tm = theta.copy(), theta.copy()
            tp[j] += h
            tm[j] -= h
            fd = (grid.soft_stage(demand, off, nodes, parents, tau, *tp) -
                  grid.soft_stage(demand, off, nodes, parents, tau, *tm)) / (2 * h)
            assert abs(fd - gp[j]) <= 1e-4 * max(1e-3, abs(fd)) + _fd_tol(F, h), (tau, j, fd, gp[j])


def _torch_soft_net(torch, t, w, v, nodes, parents, tau):
    """Dense torch restatement of the soft DP (independent of soft.c):
    interval tables as L x L tensors with +inf outside lo <= hi."""
    L, X, Y = t["L"], t["X"], t["Y"]
    XY = X * Y
    DIR = t["DIR"]
    inf = torch.tensor(float("inf"), dtype=torch.float64)

    def idx(l, x, y):
        return (l * X + x) * Y + y

    def smin(x):
        x = x.reshape(-1)
        x = x[torch.isfinite(x)]
        if x.numel() == 0:
            return inf
        return -tau * torch.logsumexp(-x / tau, dim=0)

    def leg(l, along_y, fixed, lo, hi):
        if hi <= lo:
            return torch.zeros((), dtype=torch.float64)
        cells = [idx(l, fixed, p) if along_y else idx(l, p, fixed) for p in range(lo, hi)]
        return w[cells].sum()

    def ldir(l):
        return (l & 1) ^ DIR

    n = len(nodes)
    D = []
    for i in range(n):
        fixed, x, y = int(nodes[i]) // XY, int(nodes[i]) // Y % X, int(nodes[i]) % Y
        rows = []
        for lo in range(L):
            row = []
            for hi in range(L):
                if hi < lo or (fixed < L and (lo > fixed or hi < fixed)):
                    row.append(inf)
                else:
                    row.append(sum((v[idx(l, x, y)] for l in range(lo, hi)), torch.zeros((), dtype=torch.float64)))
            rows.append(torch.stack(row))
        D.append(torch.stack(rows))
    for i in range(n - 1, 0, -1):
        p = int(parents[i])
        x, y = int(nodes[i]) // Y % X, int(nodes[i]) % Y
        px, py = int(nodes[p]) // Y % X, int(nodes[p]) % Y
        minx, maxx, miny, maxy = min(x, px), max(x, px), min(y, py), max(y, py)
        straight = x == px or y == py
        best = [smin(D[i][:k + 1, k:]) for k in range(L)]
        diag = []
        for pl in range(L):
            d = ldir(pl)
            if straight:
                ok = (x == px) if d else (y == py)
                w1 = leg(pl, 1, x, miny, maxy) if d else leg(pl, 0, y, minx, maxx)
                diag.append(best[pl] + w1 if ok else inf)
            else:
                cand = []
                for cl in range(L):
                    if not (cl ^ pl) & 1:
                        continue
                    dcl = ldir(cl)
                    w1 = leg(cl, 1, x, miny, maxy) if dcl else leg(cl, 0, y, minx, maxx)
                    w2 = leg(pl, 1, px, miny, maxy) if d else leg(pl, 0, py, minx, maxx)
                    lo, hi = min(cl, pl), max(cl, pl)
                    f = sum((v[idx(l, x, py)] if dcl else v[idx(l, px, y)] for l in range(lo + 1, hi)),
                            torch.zeros((), dtype=torch.float64))
                    cand.append(best[cl] + w1 + f + w2)
                diag.append(smin(torch.stack(cand)))
        diag = torch.stack(diag)
        C = torch.stack([torch.stack([smin(diag[lo:hi + 1]) if hi >= lo else inf for hi in range(L)])
                         for lo in range(L)])
        D[p] = D[p] + C
    return smin(D[0])


def test_soft_adjoint_matches_torch_autograd(ora):
    torch = pytest.importorskip("torch")
    _, t = topology(6, 4)
    g, w, v = _costs(ora, t, 5)
    for nodes, parents in _nets(t, 4, min_nodes=2, max_nodes=5):
        F, gw, gv = g.soft_net(w, v, nodes, parents, 1.0, grads=True)
        wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
        vt = torch.tensor(v, dtype=torch.float64, requires_grad=True)
        Ft = _torch_soft_net(torch, t, wt, vt, nodes, parents, 1.0)
        Ft.backward()
        assert abs(Ft.item() - F) <= 1e-9 * max(1.0, abs(F))
        assert np.allclose(wt.grad.numpy(), gw, rtol=1e-9, atol=1e-12)
        assert np.allclose(vt.grad.numpy(), gv, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("L", [2, 3, 4])
def test_soft_zshape_tends_to_enumerated_optimum(ora, L):
    """The Z-shape family (kind 1) at tau -> 0 against the exhaustive
    minimum over the Z-shape route space on tiny boxes (the enumeration of
    test_zshape_oracle.py: every node interval, every edge shape, bend and
    leg layer): with integer wire steps and via costs of at least 1 its
    integer costs are the real ones, so the soft value lies within
    tau * log(number of alternatives) below the optimum."""
    from tests.test_zshape_oracle import _brute_force, _random_tree, _tiny_instance
    rng = np.random.default_rng(40 + L)
    done = 0
    for trial in range(30):
        X, Y = int(rng.integers(3, 6)), int(rng.integers(3, 6))
        DIR = int(rng.integers(0, 2))
        grid, presum, vcost = _tiny_instance(rng, X, Y, L, DIR)
        vcost = np.maximum(vcost, 1).astype(np.float32)
        tree = _random_tree(rng, X, Y, L, int(rng.integers(2, 5)))
        want = _brute_force(grid, presum, vcost, tree)
        # wcost[p] is the step from track position p to p + 1
        ps = presum.reshape(L, X, Y)
        w = np.zeros((L, X, Y))
        for l in range(L):
            if (l & 1) ^ DIR:  # along y
                w[l, :, :-1] = ps[l, :, 1:] - ps[l, :, :-1]
            else:
                w[l, :-1, :] = ps[l, 1:, :] - ps[l, :-1, :]
        t = dict(grid, nets=0)
        g = ora.Grid(t)
        nodes = [(L if pin is None else pin - 1) * X * Y + x * Y + y for x, y, pin, _, _ in tree]
        parents = [max(p, 0) for _, _, _, p, _ in tree]
        for tau in (1e-2, 1e-3):
            soft = g.soft_net(w.reshape(-1), vcost.astype(np.float64), nodes, parents, tau, kind=1)
            alternatives = len(tree) * np.log(L * L) + (len(tree) - 1) * np.log(L ** 3 * (X + Y + 2))
            assert soft <= want + 1e-9 and want - soft <= tau * alternatives + 1e-9, (trial, tau, soft, want)
        done += 1
    assert done == 30
