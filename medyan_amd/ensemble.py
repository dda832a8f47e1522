"""Batched ensembles of independent networks (the batched minimization form).

An ensemble of networks whose cylinder counts are log-uniform in
[500, 20000] (seeded), 10 cylinders per filament, cubic boxes at a fixed
density of 60 cylinders per um^3, 0.5 linkers and 0.05 motors per cylinder,
with the potentials of the single-network workloads; minimized together by
mdy_minimize_batch, one cluster of resident blocks per network.
"""
import math

import numpy as np

from . import _abi, mechanics, network

SEED = 0x4D454459 + 2


def ensemble_sizes(n=256, lo=500, hi=20000, seed=SEED):
    """Cylinder counts of the ensemble (multiples of 10)."""
    rng = np.random.default_rng(seed)
    nc = np.exp(rng.uniform(math.log(lo), math.log(hi), n))
    return (np.maximum(1, np.round(nc / 10)) * 10).astype(int)


def ensemble_network(lib, n_cyl, index, seed=SEED, density=60.0):
    """Network `index` of the ensemble: n_cyl cylinders in a box holding
    `density` cylinders per um^3."""
    box = 1000.0 * (n_cyl / density) ** (1.0 / 3.0)
    return network.generate(lib, None, seed=seed + 7919 * (index + 1), n_filaments=int(n_cyl // 10),
                            cyl_per_filament=10, box=box, n_linkers=int(0.5 * n_cyl),
                            n_motors=int(0.05 * n_cyl), n_branches=0)


def make_ensemble(lib, n=256, seed=SEED, sizes=None):
    sizes = ensemble_sizes(n, seed=seed) if sizes is None else sizes
    return [ensemble_network(lib, int(c), i, seed) for i, c in enumerate(sizes)]


def autotune(ctxs, d_coords, coords0, params, candidates=(64, 128, 256), copy=None):
    """Threads per block of the batched launch: each candidate minimizes the
    ensemble once from the same coordinates; the fastest is returned with
    all timings. copy(i, src) restores network i's coordinates."""
    times = {}
    for tpb in candidates:
        for i in range(len(ctxs)):
            copy(i, coords0[i])
        _, _, tot = mechanics.minimize_batch(ctxs, params, d_coords, threads_per_block=tpb)
        times[tpb] = tot
    best = min(times, key=times.get)
    return best, times
