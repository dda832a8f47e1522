"""Writes net40k_trace.json: the exact-mode oracle's complete minimization of
the net40k network (4,000 filaments x 10 cylinders, seed 0x4D454459 + 1) to
1 pN with the benchmark's CG parameters (PR + backtracking, maxDist 1,
lambdaMax 0.01, energy criterion off), as checksums: SHA-256 of the
per-iteration trace (lambda, energy, max force, beta as float64 bytes; the
safe-mode flags as int32 bytes) and of the final coordinates and forces,
the result record's counts, and every 500th trace entry in hex for
locating a first difference. The oracle is the CPU restatement in oracle/
(parity build, exact mode: the shared atan/exp and exact sums); a run takes
10-15 minutes on one core.

    python tests/golden/make_net40k_trace.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SAMPLE_EVERY = 500
OUT = os.path.join(HERE, "net40k_trace.json")


def digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def trace_record(tr, x, f, res):
    n = len(tr["lambda_"])
    keys = ("lambda_", "energy", "maxf", "beta")
    return {
        "num_iter": int(res.num_iter),
        "num_energy_call": int(res.num_energy_call),
        "num_force_call": int(res.num_force_call),
        "trace_len": n,
        "trace_sha256": {k: digest(tr[k], np.float64) for k in keys} | {"safe": digest(tr["safe"], np.int32)},
        "coord_sha256": digest(x, np.float64),
        "force_sha256": digest(f, np.float64),
        "final_energy_hex": float(res.energies_after[-1]).hex(),
        "grad_inf_norm_hex": float(res.grad_inf_norm).hex(),
        "samples": {str(i): [float(tr[k][i]).hex() for k in keys] for i in range(0, n, SAMPLE_EVERY)},
    }


def main():
    import pyoracle
    from medyan_amd import _abi, mechanics, network

    lib = mechanics.load_library()
    net = network.generate(lib, "net40k")
    o = pyoracle.Oracle()
    o.set_exact(1)
    s = o.system(net)
    p = _abi.default_cg_params(net.n_beads, grad_tol=1.0, energy_rel_tol=0.0)
    t0 = time.perf_counter()
    x, f, res, tr = s.minimize(p, trace_cap=200000)
    dt = time.perf_counter() - t0
    rec = trace_record(tr, x, f, res)
    rec["workload"] = "net40k"
    rec["oracle_seconds"] = round(dt, 1)
    with open(OUT, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(f"{res.num_iter} iterations in {dt:.0f} s -> {OUT}")


if __name__ == "__main__":
    main()
