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


BONDED_TABLES = ("fs_k", "fb_k", "ls_k", "ms_k", "bs_k", "bb_k", "bdh_k", "bp_k")


def work(ctx):
    """A network's per-evaluation work as the batched launch sizes its block
    cluster by default: its interactions (bonded, excluded volume, boundary)
    plus 2 per bead."""
    n = ctx.net
    bonded = sum(len(getattr(n, f)) for f in BONDED_TABLES)
    return bonded + len(ctx.exvol_table()[1]) + len(ctx.boundary_table()[1]) + 2 * n.n_beads


def pass_rule(first_pass_calls):
    """The pass rule a first_pass_calls value of mdy_minimize_batch names."""
    f = int(first_pass_calls)
    if f > 0:
        return f"passes capped at {f} x 4^k force calls"
    if f < 0:
        return f"passes end when 1/{-f} of their networks have finished" if f < -1 else "one pass"
    return "one pass"


def autotune(ctxs, params, d_coords, restore, threads=(64, 128, 256), first_pass=(-8, -4, -2, 250, 500, 1000, 2000),
             exponents=(1.0, 1.25), log=None):
    """Block and tile shapes of the batched launch (mdy_minimize_batch) for
    one ensemble, by coordinate search over measured runs:
      1. threads per block, the unit every network's cluster is built from
         (64: more, smaller blocks per network; 256: fewer, wider ones), with
         the first pass rule;
      2. the pass rule: a force-call cap for the first pass (how long each
         network keeps its first share of blocks before the blocks of
         finished networks are moved to the ones still running), or, for a
         negative value -k, passes that end once 1/k of their networks have
         finished;
      3. the exponent a of the cluster shares w_i = work_i ** a (a > 1 gives
         the large networks more than their share of the blocks).
    Each stage keeps the best value of the previous ones. Each candidate
    minimizes the whole ensemble once from the same coordinates (restore()
    puts them back); every candidate gives the same results bit for bit,
    only the time differs. Returns (settings for minimize_batch: dict with
    threads_per_block, first_pass_calls, weights; {(threads, cap, a): ms})."""
    wk = np.array([work(c) for c in ctxs], dtype=np.float64)
    times = {}

    def run(tpb, first, a):
        key = (tpb, first, a)
        if key not in times:
            restore()
            _, _, tot = mechanics.minimize_batch(ctxs, params, d_coords, threads_per_block=tpb,
                                                 weights=wk ** a, first_pass_calls=first)
            times[key] = tot
            if log:
                log(f"autotune: {tpb} threads, {pass_rule(first)}, shares work^{a}: {tot:.1f} ms")
        return times[key]

    first0 = first_pass[0]
    tpb = min(threads, key=lambda t: run(t, first0, 1.0))
    first = min(first_pass, key=lambda f: run(tpb, f, 1.0))
    a = min(exponents, key=lambda e: run(tpb, first, e))
    return dict(threads_per_block=tpb, first_pass_calls=first, weights=wk ** a), times
