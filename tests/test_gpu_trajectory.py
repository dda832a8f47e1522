"""The whole net40k minimization (20,875 CG iterations to 1 pN) on the
device against the exact-mode oracle's, bit for bit: SHA-256 of every
per-iteration trace array (accepted lambda, energy, max force, beta,
safe-mode flag), of the final coordinates and forces, and the result's
counts, from tests/golden/net40k_trace.json (written on the CPU by
tests/golden/make_net40k_trace.py; the oracle needs about 50 minutes for it).
On a mismatch the fixture's every-500th entries locate the first difference."""
import json
import os

import numpy as np
import pytest

from medyan_amd import _abi, mechanics, network

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def digest(a, dtype):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def test_net40k_full_trajectory_bitwise(lib):
    with open(os.path.join(HERE, "golden", "net40k_trace.json")) as fh:
        ref = json.load(fh)
    net = network.generate(lib, "net40k")
    print(f"[gpu] net40k full trajectory: {net.n_cyl} cylinders", flush=True)
    ctx = mechanics.Context(0)
    ctx.set_network(net)
    ctx.build_neighbors()
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    x, f, res = ctx.minimize(p, trace_cap=200000)
    tr = ctx.trace()
    keys = ("lambda_", "energy", "maxf", "beta")
    for n in sorted(int(k) for k in ref["samples"]):
        got = [float(tr[k][n]).hex() for k in keys] if n < len(tr["lambda_"]) else None
        assert got == ref["samples"][str(n)], f"first sampled difference at iteration {n}"
    assert (res.num_iter, res.num_energy_call, res.num_force_call) == (
        ref["num_iter"], ref["num_energy_call"], ref["num_force_call"])
    assert len(tr["lambda_"]) == ref["trace_len"]
    for k in keys:
        assert digest(tr[k], np.float64) == ref["trace_sha256"][k], k
    assert digest(tr["safe"], np.int32) == ref["trace_sha256"]["safe"]
    assert digest(x, np.float64) == ref["coord_sha256"]
    assert digest(f, np.float64) == ref["force_sha256"]
    assert float(res.energies_after[_abi.MDY_NFF]).hex() == ref["final_energy_hex"]
    assert float(res.grad_inf_norm).hex() == ref["grad_inf_norm_hex"]
    print(f"[gpu] {res.num_iter} iterations bit for bit")
    ctx.close()
