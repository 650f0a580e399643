"""Shared synthetic instances for the tests (seeded, small)."""
import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def topology(config=1, seed=0, net_scale=1.0, grid_side=0):
    import grotor_amd as ga
    d = ga.Design.generate(config, seed=seed, net_scale=net_scale, grid_side=grid_side)
    d.prepare(8)
    return d, d.topology()


def random_demand(topo, seed, high=24):
    """Demand values as commits produce them: multiples of 0.5."""
    rng = np.random.default_rng(seed)
    cells = topo["L"] * topo["X"] * topo["Y"]
    return (rng.integers(0, 2 * high, cells) * 0.5).astype(np.float32)
