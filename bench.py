"""Benchmark: CG minimizations/s to 1 pN max force at 40k cylinders on one GPU.

One step = one full mechanical equilibration of the 40k-cylinder network
(4,000 filaments x 10 cylinders, 4 um box, 20k linkers, 2k motors, 2k branch
points; seeded synthetic geometry): neighbor-list rebuild from the initial
coordinates, then Polak-Ribiere CG with backtracking line search until the
largest force component is below 1 pN (energy-change criterion off, so every
run ends on the force criterion). Inputs are resident in device memory before
the timed region; the step copies the initial coordinates device-to-device,
rebuilds the lists and minimizes.

Prints one JSON line. The roofline block is for the persistent CG kernel
(`k_cg`, one launch per minimization): algorithmic bytes per launch from the
counts the minimization reports, over the launch's duration measured with HIP
events on the stream it runs on. The CPU baseline is the oracle (a C
restatement of the reference minimizer, built with the reference's release
flags) timed single-threaded on a bounded number of CG iterations and scaled
to minimizations/s with the GPU run's iteration count per minimization.
"""
import argparse
import glob
import json
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "net40k"
# per-interaction algorithmic bytes: int32 bead offsets + FP64 parameters
B_FS, B_FB, B_LM, B_BS, B_BB, B_EV, B_BD = 24, 28, 48, 36, 32, 40, 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget (0: skip)")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false", help="time the FP64 form only")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def algorithmic_bytes(net, n_ev, n_bd, res):
    """Minimal unique bytes the minimization touches, from its own counts:
    every evaluation reads coordinates and every interaction table; every
    force evaluation writes one force per bead; each iteration makes 7
    vector passes and each line-search trial 3 (x, s, F, Fprev, coordinates)."""
    nb = net.n_beads
    ndof = 3 * nb
    tables = (B_FS * len(net.fs_k) + B_FB * len(net.fb_k) + B_LM * (len(net.ls_k) + len(net.ms_k))
              + B_BS * len(net.bs_k) + B_BB * len(net.bb_k) + B_EV * n_ev + B_BD * n_bd)
    b_energy = 24 * nb + tables + 8
    n_eval = res.num_energy_eval
    trials = max(n_eval - 2, 0)
    return (n_eval * b_energy + res.num_force_call * 24 * nb
            + 8 * ndof * (7 * res.num_iter + 3 * trials))


def traffic_from_profiles(workload):
    """HBM bytes per k_cg launch from the committed counter summary, if any."""
    for path in sorted(glob.glob(os.path.join(HERE, "profiles", f"k_cg_{workload}_traffic*.md"))):
        with open(path) as fh:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", fh.read())
        if m:
            return float(m.group(1))
    return None


def cpu_baseline(net, gpu_iters, budget_s):
    """The oracle minimizer (reference release flags), one thread, on the same
    network for a bounded number of iterations."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    from medyan_amd import _abi

    o = pyoracle.Oracle(fast=True)
    o.set_exact(0)
    s = o.system(net)
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    t0 = time.perf_counter()
    _, _, r, _ = s.minimize(p, iter_cap=10)
    dt = time.perf_counter() - t0
    per_iter = dt / max(r.num_iter, 1)
    n = max(10, int(budget_s / per_iter))
    t0 = time.perf_counter()
    _, _, r, _ = s.minimize(p, iter_cap=n)
    dt = time.perf_counter() - t0
    per_iter = dt / max(r.num_iter, 1)
    return dict(value=1.0 / (per_iter * gpu_iters), unit="minimizations/s", cores=1, kind="port",
                sample=(f"{WORKLOAD}: first {r.num_iter} CG iterations of the oracle minimizer "
                        f"(-O2 -ffast-math, 1 thread) in {dt:.1f} s = {per_iter * 1e3:.2f} ms/iteration, "
                        f"scaled by the GPU's {gpu_iters} iterations per minimization; "
                        f"host {_cpu_model()}, 1 of {os.cpu_count()} cores"))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}): the single-network CG path "
              "runs on one device; multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from medyan_amd import _abi, mechanics, network

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    if a.copy_probe_mb:
        n = a.copy_probe_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}))
        return

    lib = mechanics.load_library()
    net = network.generate(lib, a.workload)
    x0 = torch.from_numpy(np.ascontiguousarray(net.coord)).to(dev)
    params = _abi.default_cg_params(net.n_beads, grad_tol=1.0, energy_rel_tol=0.0)

    def run_form(precision):
        """Times a.steps minimizations (after a.warmup) in one precision form."""
        ctx = mechanics.Context(precision=precision)
        ctx.set_network(net)
        x = torch.empty_like(x0)

        def step():
            x.copy_(x0)
            torch.cuda.synchronize()
            ctx.build_neighbors_dev(x.data_ptr())
            return ctx.minimize_dev(params, x.data_ptr())

        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        kernel_ms, results = [], []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            results.append(step())
            kernel_ms.append(ctx.last_ms())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        r = results[-1]
        if r.grad_inf_norm >= 1.0 or r.status != 0:
            print(f"minimization did not reach 1 pN: maxF={r.grad_inf_norm} status={r.status}", file=sys.stderr)
            sys.exit(1)
        return ctx, wall, float(np.mean(kernel_ms)), r

    ctx, wall, k_ms, r = run_form(_abi.MDY_PREC_FP64)
    n_ev = len(ctx.exvol_table()[1])
    n_bd = len(ctx.boundary_table()[1])
    abytes = algorithmic_bytes(net, n_ev, n_bd, r)
    achieved = abytes / (k_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(a.workload)
    mixed = None
    if a.mixed:
        mctx, mwall, mk_ms, mr = run_form(_abi.MDY_PREC_MIXED)
        e64, emx = r.energies_after[_abi.MDY_NFF], mr.energies_after[_abi.MDY_NFF]
        mixed = {
            "value": a.steps / mwall,
            "unit": "minimizations/s",
            "ms_per_step": 1e3 * mwall / a.steps,
            "dtype": "f32 pair forces (well-conditioned pairs) + f64 energies, reductions, directions",
            "cg_iterations": int(mr.num_iter),
            "us_per_cg_iteration": 1e3 * mk_ms / max(int(mr.num_iter), 1),
            "final_max_force_pN": float(mr.grad_inf_norm),
            "final_energy": float(emx),
            "final_energy_rel_diff_vs_fp64": abs(emx - e64) / abs(e64),
            "note": ("both forms stop below 1 pN; they follow different trajectories and settle in "
                     "different local minima, so final energies differ by percent, not by rounding"),
        }
        del mctx
    out = {
        "metric": "CG minimizations/s to 1 pN max-force at 40k cylinders; mixed vs FP64 on GPU",
        "value": a.steps / wall,
        "unit": "minimizations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded network generator)",
        "config": {
            "workload": (f"{a.workload}: {net.n_cyl} cylinders, {net.n_beads} beads, "
                         f"{n_ev} excluded-volume pairs, {n_bd} boundary terms; PR-CG + backtracking, "
                         "gradTol 1 pN, maxDist 1, lambdaMax 0.01, energy criterion off; "
                         "neighbor-list rebuild + minimization per step"),
            "precision": "fp64",
            "cg_iterations": int(r.num_iter),
            "us_per_cg_iteration": 1e3 * k_ms / max(int(r.num_iter), 1),
            "energy_evaluations": int(r.num_energy_eval),
            "force_calls": int(r.num_force_call),
            "final_max_force_pN": float(r.grad_inf_norm),
            "final_energy": float(r.energies_after[_abi.MDY_NFF]),
        },
        "mixed": mixed,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_cg (persistent CG, one launch per minimization)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": abytes,
            "launch_ms": k_ms,
        },
        "cpu_baseline": None,
    }
    if a.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(net, int(r.num_iter), a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
