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
flags), single-threaded: its seconds per energy evaluation, per force
evaluation and per CG iteration for the rest are measured (the last on a
bounded sample of the same minimization) and applied to the energy calls,
force calls and iterations of the GPU's minimization. `--cpu-validate` checks
that model against one full CPU minimization (no GPU needed).

Other workloads (--workload): `ensemble256`, 256 independent networks
minimized by one batched launch (networks/s, per-network p50/p99); `net200k`,
successive minimizations of a 200k-cylinder network with the neighbor list
rebuilt before each (minimizations/s mixed vs FP64, GB/s per kernel).
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
B_BDH, B_BP = 32, 28


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget (0: skip)")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false", help="time the FP64 form only")
    ap.add_argument("--cpu-validate", action="store_true",
                    help="CPU only: one full oracle minimization of --workload against the CPU cost model")
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
              + B_BS * len(net.bs_k) + B_BB * len(net.bb_k) + B_BDH * len(net.bdh_k) + B_BP * len(net.bp_k)
              + B_EV * n_ev + B_BD * n_bd)
    b_energy = 24 * nb + tables + 8
    n_eval = res.num_energy_eval
    trials = max(n_eval - 2, 0)
    return (n_eval * b_energy + res.num_force_call * 24 * nb
            + 8 * ndof * (7 * res.num_iter + 3 * trials))


def generator_rule(lib):
    """The seeded generator's placement rule, stated with the workload."""
    from medyan_amd import network
    g = network.gen_params(lib)
    return (f"generator: filament start uniform in the box, direction uniform on the sphere, a placement "
            f"rejected if a segment comes within {g.min_sep:g} nm of a placed one (avoids singular "
            f"excluded-volume overlaps), then N(0, {g.jitter:g} nm) bead jitter")


def traffic_from_profiles(workload):
    """HBM bytes per k_cg launch from the committed counter summary, if any."""
    for path in sorted(glob.glob(os.path.join(HERE, "profiles", f"k_cg_{workload}_traffic*.md"))):
        with open(path) as fh:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", fh.read())
        if m:
            return float(m.group(1))
    return None


def _timed(fn, budget_s):
    """Seconds per call of fn, repeated for about budget_s."""
    fn()
    n, t0 = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            return dt / n


def cpu_cost_model(net, budget_s):
    """The oracle minimizer (reference release flags, reference-order
    arithmetic), one thread, on a bounded sample of the minimization: its
    seconds per evaluation unit, where an energy evaluation is one unit and a
    force evaluation c units (c = the ratio of the two, timed separately at
    the start coordinates); the line-search vector passes and CG updates of
    the sample are spread over its evaluations. Returns (model, sample
    result, sample seconds)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    from medyan_amd import _abi

    o = pyoracle.Oracle(fast=True)
    o.set_exact(0)
    s = o.system(net)
    t_e = _timed(lambda: s.energy(), min(1.0, 0.1 * budget_s))
    t_f = _timed(lambda: s.forces(), min(1.0, 0.1 * budget_s))
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    t0 = time.perf_counter()
    _, _, r, _ = s.minimize(p, iter_cap=5)
    per_iter = (time.perf_counter() - t0) / max(r.num_iter, 1)
    n = max(10, int(0.8 * budget_s / per_iter))
    t0 = time.perf_counter()
    _, _, r, _ = s.minimize(p, iter_cap=n)
    dt = time.perf_counter() - t0
    c = t_f / t_e
    unit = dt / (r.num_energy_call + c * r.num_force_call)
    return dict(force_over_energy=c, seconds_per_unit=unit), r, dt


def cpu_model_seconds(m, n_energy_call, n_force_call, n_iter=None):
    return m["seconds_per_unit"] * (n_energy_call + m["force_over_energy"] * n_force_call)


def cpu_baseline(net, gpu_res, budget_s, name=WORKLOAD):
    """CPU minimizations/s for the GPU's minimization: the cost model of
    cpu_cost_model applied to the energy calls, force calls and iterations the
    reference's algorithm makes on that path (the GPU's counts in the
    reference's convention: every line search recomputes E(x))."""
    m, r, dt = cpu_cost_model(net, budget_s)
    secs = cpu_model_seconds(m, gpu_res.num_energy_call, gpu_res.num_force_call, gpu_res.num_iter)
    return dict(value=1.0 / secs, unit="minimizations/s", cores=1, kind="port",
                sample=(f"{name}: oracle minimizer (-O2 -ffast-math, 1 thread): "
                        f"{m['seconds_per_unit'] * 1e3:.2f} ms per energy evaluation and "
                        f"{m['force_over_energy'] * m['seconds_per_unit'] * 1e3:.2f} ms per force evaluation, "
                        f"other work included (first {r.num_iter} iterations, {r.num_energy_call} energy and "
                        f"{r.num_force_call} force calls in {dt:.1f} s); applied to the GPU run's "
                        f"{gpu_res.num_energy_call} energy calls, {gpu_res.num_force_call} force calls, "
                        f"{gpu_res.num_iter} iterations = {secs:.0f} s per minimization; "
                        f"host {_cpu_model()}, 1 of {os.cpu_count()} cores"))


def _mean_counts(results):
    """Mean energy calls, force calls and iterations of several minimizations."""
    from types import SimpleNamespace
    import numpy as np
    return SimpleNamespace(**{k: int(round(np.mean([getattr(r, k) for r in results])))
                              for k in ("num_energy_call", "num_force_call", "num_iter")})


def cpu_validate(workload, budget_s):
    """A full CPU minimization (oracle, reference release flags, one thread)
    against the cost model's estimate from a bounded sample of the same
    minimization, with the CPU run's own counts. No GPU needed."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    from medyan_amd import _abi, mechanics, network

    lib = mechanics.load_library()
    net = network.generate(lib, workload)
    m, r_s, dt_s = cpu_cost_model(net, budget_s)
    o = pyoracle.Oracle(fast=True)
    o.set_exact(0)
    s = o.system(net)
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    t0 = time.perf_counter()
    _, _, r, _ = s.minimize(p)
    full = time.perf_counter() - t0
    est = cpu_model_seconds(m, r.num_energy_call, r.num_force_call, r.num_iter)
    naive = dt_s / max(r_s.num_iter, 1) * r.num_iter
    print(json.dumps({
        "workload": workload, "host": _cpu_model(), "threads": 1,
        "full_seconds": full, "model_seconds": est, "model_over_full": est / full,
        "per_iteration_scaling_seconds": naive, "per_iteration_over_full": naive / full,
        "num_iter": int(r.num_iter), "num_energy_call": int(r.num_energy_call),
        "num_force_call": int(r.num_force_call), "grad_inf_norm": float(r.grad_inf_norm),
        "status": int(r.status), "model": m, "sample_iterations": int(r_s.num_iter), "sample_seconds": dt_s,
        "sample_energy_calls": int(r_s.num_energy_call), "sample_force_calls": int(r_s.num_force_call),
    }))


def cpu_baseline_ensemble(nets, iters, budget_s):
    """The oracle minimizer, one thread, on three networks of the ensemble
    (smallest, median, largest) for a bounded number of iterations each; the
    per-iteration cost per interaction is fitted and scaled to the whole
    ensemble with the GPU's iteration count of every network."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import numpy as np
    import pyoracle
    from medyan_amd import _abi

    o = pyoracle.Oracle(fast=True)
    o.set_exact(0)
    order = np.argsort([n.n_cyl for n in nets])
    pick = [order[0], order[len(order) // 2], order[-1]]
    rates, done = [], []
    for i in pick:
        s = o.system(nets[i])
        p = _abi.default_cg_params(nets[i].n_beads, energy_rel_tol=0.0)
        t0 = time.perf_counter()
        _, _, r, _ = s.minimize(p, iter_cap=5)
        per = (time.perf_counter() - t0) / max(r.num_iter, 1)
        n = max(5, int(budget_s / 3 / per))
        t0 = time.perf_counter()
        _, _, r, _ = s.minimize(p, iter_cap=n)
        dt = time.perf_counter() - t0
        rates.append(dt / max(r.num_iter, 1) / nets[i].n_cyl)
        done.append((nets[i].n_cyl, int(r.num_iter), dt))
    per_cyl_iter = float(np.mean(rates))
    total = sum(per_cyl_iter * n.n_cyl * it for n, it in zip(nets, iters))
    return dict(value=len(nets) / total, unit="networks/s", cores=1, kind="port",
                sample=("oracle minimizer (-O2 -ffast-math, 1 thread) on the smallest, median and largest "
                        "networks: " + "; ".join(f"{c} cylinders, {k} iterations in {t:.1f} s" for c, k, t in done)
                        + f"; {per_cyl_iter * 1e9:.0f} ns per cylinder-iteration, scaled by every network's GPU "
                        f"iteration count; host {_cpu_model()}, 1 of {os.cpu_count()} cores"))


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
    if a.cpu_validate:
        return cpu_validate(a.workload, a.cpu_seconds)
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
    if a.workload == "ensemble256":
        return bench_ensemble(a, lib)
    if a.workload == "net200k":
        return bench_successive(a, lib)
    if a.workload == "mixed_tight":
        return bench_mixed_tight(a, lib)
    if a.workload == "net50k_grad":
        return bench_gradient(a, lib)
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
                     "different local minima, so final energies differ by percent, not by rounding, and the "
                     "minimizations/s of the two forms are not a precision speed-up: compare "
                     "us_per_cg_iteration (the iteration is latency-bound in both). At tighter tolerances the "
                     "final-energy difference stays at the size of FP64's own sensitivity to a 1e-9 nm "
                     "change of the start (--workload mixed_tight)"),
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
                         "neighbor-list rebuild + minimization per step; " + generator_rule(lib)),
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
        out["cpu_baseline"] = cpu_baseline(net, r, a.cpu_seconds)
    print(json.dumps(out))


def bench_ensemble(a, lib):
    """256 independent networks (cylinder counts log-uniform in [500, 20000],
    60 cylinders per um^3, 0.5 linkers and 0.05 motors per cylinder) minimized
    to 1 pN together by one batched launch in the mixed form. One step = one
    batched minimization of the whole ensemble from its initial coordinates
    (tables and neighbor lists resident), in passes that move the blocks of
    finished networks to the slow ones (passes ended by force-call caps or
    by completions). Threads per block, the pass rule and the cluster shares
    come from the library's autotuner (medyan_amd.ensemble.autotune) before
    the warmup."""
    import numpy as np
    import torch
    from medyan_amd import _abi, ensemble, mechanics

    nets = ensemble.make_ensemble(lib, 256)
    ctxs = []
    t0 = time.perf_counter()
    for net in nets:
        c = mechanics.Context(precision=_abi.MDY_PREC_MIXED)
        c.set_network(net)
        c.build_neighbors()
        ctxs.append(c)
    setup_s = time.perf_counter() - t0
    x0 = [torch.from_numpy(np.ascontiguousarray(n.coord)).cuda() for n in nets]
    xs = [x.clone() for x in x0]
    ptrs = [x.data_ptr() for x in xs]
    params = [_abi.default_cg_params(n.n_beads, grad_tol=1.0, energy_rel_tol=0.0) for n in nets]
    ncyl = np.array([n.n_cyl for n in nets], dtype=np.float64)
    print(f"[ensemble256] {len(nets)} networks ready", file=sys.stderr, flush=True)

    def restore():
        for x, x_0 in zip(xs, x0):
            x.copy_(x_0)
        torch.cuda.synchronize()

    def run(prm=params):
        restore()
        return mechanics.minimize_batch(ctxs, prm, ptrs, **best)

    best, tune = ensemble.autotune(ctxs, params, ptrs, restore,
                                   log=lambda m: print(f"[ensemble256] {m}", file=sys.stderr, flush=True))
    tpb, first = best["threads_per_block"], best["first_pass_calls"]
    share_exp = min((k for k in tune if k[0] == tpb and k[1] == first), key=tune.get)[2]
    for _ in range(a.warmup):
        run()
    ms_steps, per_net = [], []
    for _ in range(a.steps):
        res, ms, tot = run()
        ms_steps.append(tot)
        per_net.append(ms)
        print(f"[ensemble256] step: {tot:.1f} ms", file=sys.stderr, flush=True)
    bad = [i for i, r in enumerate(res) if r.status != 0 or r.grad_inf_norm >= 1.0]
    if bad:
        print(f"networks {bad[:8]} did not reach 1 pN", file=sys.stderr)
        sys.exit(1)
    per_net = np.concatenate(per_net)
    batch_ms = float(np.mean(ms_steps))
    abytes = 0
    for net, c, r in zip(nets, ctxs, res):
        abytes += algorithmic_bytes(net, len(c.exvol_table()[1]), len(c.boundary_table()[1]), r)
    achieved = abytes / (batch_ms * 1e-3) / 1e9
    iters = [int(r.num_iter) for r in res]
    traffic = traffic_from_profiles("ensemble256")
    # the reference's default stopping rule (force OR relative energy change < 1e-9)
    dparams = [_abi.default_cg_params(n.n_beads, grad_tol=1.0) for n in nets]
    _, _, default_ms = run(dparams)
    # the same networks one after another, each on the whole GPU (single-network kernel)
    seq_ms = 0.0
    for c, x, x_0, p in zip(ctxs, xs, x0, params):
        x.copy_(x_0)
        torch.cuda.synchronize()
        c.minimize_dev(p, x.data_ptr())
        seq_ms += c.last_ms()
    out = {
        "metric": "networks/s to 1 pN, 256-network batched ensemble (mixed form), one launch",
        "value": 256.0 / (batch_ms * 1e-3),
        "unit": "networks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": batch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces (well-conditioned pairs) + f64 energies, reductions, directions",
        "data": "synthetic (seeded network generator)",
        "config": {
            "workload": (f"ensemble256: 256 networks, {int(ncyl.sum())} cylinders in total "
                         f"({int(ncyl.min())}..{int(ncyl.max())} per network, log-uniform), 60 cylinders per "
                         "um^3, 0.5 linkers and 0.05 motors per cylinder; PR-CG + backtracking to 1 pN, "
                         "energy criterion off; tables and neighbor lists resident; " + generator_rule(lib)),
            "threads_per_block": tpb,
            "first_pass_calls": first,
            "pass_rule": ensemble.pass_rule(first),
            "cluster_share_exponent": share_exp,
            "autotune_ms": {f"{k[0]} threads, {ensemble.pass_rule(k[1])}, shares work^{k[2]}": round(v, 1)
                            for k, v in tune.items()},
            "per_network_ms_p50": float(np.percentile(per_net, 50)),
            "per_network_ms_p99": float(np.percentile(per_net, 99)),
            "per_network_ms_max": float(per_net.max()),
            "cg_iterations_median": float(np.median(iters)),
            "cg_iterations_max": int(max(iters)),
            "sequential_single_network_ms": seq_ms,
            "sequential_networks_per_s": 256.0 / (seq_ms * 1e-3),
            "default_stopping_rule_batch_ms": default_ms,
            "default_stopping_rule_networks_per_s": 256.0 / (default_ms * 1e-3),
            "setup_s_contexts_and_neighbor_lists": setup_s,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_cg_batch (one cooperative launch, one block cluster per network)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": abytes,
            "launch_ms": batch_ms,
        },
        "cpu_baseline": cpu_baseline_ensemble(nets, iters, a.cpu_seconds) if a.cpu_seconds > 0 else None,
    }
    print(json.dumps(out))


def bench_successive(a, lib):
    """200k cylinders: successive minimizations, each from the previous
    minimum displaced by seeded N(0, 5 nm) bead noise, with the neighbor list
    rebuilt before each. One step = rebuild + minimization. Mixed form is the
    headline; the FP64 form runs the same chain from the same start."""
    import numpy as np
    import torch
    from medyan_amd import _abi, mechanics, network

    net = network.generate(lib, "net200k")
    params = _abi.default_cg_params(net.n_beads, grad_tol=1.0, energy_rel_tol=0.0)
    x_init = torch.from_numpy(np.ascontiguousarray(net.coord)).cuda()

    def chain(precision):
        ctx = mechanics.Context(precision=precision)
        ctx.set_network(net)
        x = x_init.clone()
        gen = torch.Generator(device="cuda")
        gen.manual_seed(0x4D454459 + 4)
        nl_ms, cg_ms, res, starts = [], [], [], []
        for m in range(a.warmup + a.steps):
            x.add_(torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float64), alpha=5.0)
            if m >= a.warmup:
                starts.append(x.clone())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.build_neighbors_dev(x.data_ptr())
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r = ctx.minimize_dev(params, x.data_ptr())
            if r.status != 0 or r.grad_inf_norm >= 1.0:
                print(f"minimization {m} did not reach 1 pN: {r.grad_inf_norm} status {r.status}", file=sys.stderr)
                sys.exit(1)
            print(f"[net200k] precision {precision} minimization {m}: {r.num_iter} iterations, "
                  f"{ctx.last_ms():.0f} ms", file=sys.stderr, flush=True)
            if m >= a.warmup:
                nl_ms.append(1e3 * (t1 - t0))
                cg_ms.append(ctx.last_ms())
                res.append(r)
        return ctx, nl_ms, cg_ms, res, starts

    mctx, m_nl, m_cg, m_res, _ = chain(_abi.MDY_PREC_MIXED)
    fctx, f_nl, f_cg, f_res, f_starts = chain(_abi.MDY_PREC_FP64)
    # paired: the mixed form from each FP64 step's own start, final energies
    # compared; control: the FP64 form from the same start displaced by
    # N(0, 1e-9 nm), the resolution of the 1 pN stopping rule itself
    nff = _abi.MDY_NFF
    paired, control = [], []
    cgen = torch.Generator(device="cuda")
    cgen.manual_seed(7)
    for xs, rf in zip(f_starts, f_res):
        ef = rf.energies_after[nff]
        x = xs.clone()
        mctx.build_neighbors_dev(x.data_ptr())
        r = mctx.minimize_dev(params, x.data_ptr())
        paired.append(abs(r.energies_after[nff] - ef) / abs(ef))
        x = xs.clone()
        x.add_(torch.randn(x.shape, generator=cgen, device="cuda", dtype=torch.float64), alpha=1e-9)
        fctx.build_neighbors_dev(x.data_ptr())
        r = fctx.minimize_dev(params, x.data_ptr())
        control.append(abs(r.energies_after[nff] - ef) / abs(ef))
    n_ev = len(mctx.exvol_table()[1])
    n_bd = len(mctx.boundary_table()[1])
    m_bytes = [algorithmic_bytes(net, n_ev, n_bd, r) for r in m_res]
    f_bytes = [algorithmic_bytes(net, n_ev, n_bd, r) for r in f_res]
    m_gbs = [b / (t * 1e-3) / 1e9 for b, t in zip(m_bytes, m_cg)]
    f_gbs = [b / (t * 1e-3) / 1e9 for b, t in zip(f_bytes, f_cg)]
    # neighbor-list rebuild: coordinates and cylinder tables read, pairs written
    nl_bytes = 24 * net.n_beads + 4 * 8 * net.n_cyl + 8 * n_ev
    m_step = np.array(m_nl) + np.array(m_cg)
    f_step = np.array(f_nl) + np.array(f_cg)
    achieved = float(np.mean(m_gbs))
    out = {
        "metric": "minimizations/s to 1 pN at 200k cylinders (successive, NL rebuilt each), mixed vs FP64",
        "value": 1e3 / float(np.mean(m_step)),
        "unit": "minimizations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": float(np.mean(m_step)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces (well-conditioned pairs) + f64 energies, reductions, directions",
        "data": "synthetic (seeded network generator)",
        "config": {
            "workload": (f"net200k: {net.n_cyl} cylinders, {net.n_beads} beads, {len(net.ls_k)} linkers, "
                         f"{len(net.ms_k)} motors, {len(net.bs_k)} branch points, {n_ev} excluded-volume pairs "
                         "at the last rebuild; each step: +N(0, 5 nm) bead displacement of the previous minimum, "
                         "neighbor-list rebuild, PR-CG to 1 pN; " + generator_rule(lib)),
            "mixed_cg_iterations": [int(r.num_iter) for r in m_res],
            "fp64_cg_iterations": [int(r.num_iter) for r in f_res],
            "mixed_ms_per_step": float(np.mean(m_step)),
            "fp64_ms_per_step": float(np.mean(f_step)),
            "fp64_minimizations_per_s": 1e3 / float(np.mean(f_step)),
            "mixed_vs_fp64_speedup": float(np.mean(f_step) / np.mean(m_step)),
            "paired_final_energy_rel_diff_mixed_vs_fp64": paired,
            "control_final_energy_rel_diff_fp64_start_moved_1e-9nm": control,
            "kernel_GBps": {
                "k_cg mixed": achieved,
                "k_cg fp64": float(np.mean(f_gbs)),
                "neighbor-list rebuild (all its kernels, wall)": nl_bytes / (float(np.mean(m_nl)) * 1e-3) / 1e9,
            },
            "neighbor_list_ms": float(np.mean(m_nl)),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_cg (mixed form)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("net200k"),
            "algorithmic_bytes_per_launch": float(np.mean(m_bytes)),
            "launch_ms": float(np.mean(m_cg)),
        },
        "cpu_baseline": (cpu_baseline(net, _mean_counts(f_res), a.cpu_seconds, "net200k")
                         if a.cpu_seconds > 0 else None),
    }
    print(json.dumps(out))


def bench_mixed_tight(a, lib):
    """Mixed vs FP64 at equal final energy: from one FP64 minimum at 1 pN
    (identical start for both forms), each form minimizes to tighter
    gradient tolerances; the final energies, their relative difference and
    the iterations are reported per tolerance, with an FP64 control started
    from the same coordinates moved by N(0, 1e-9 nm). net500 and net40k."""
    import numpy as np
    import torch
    from medyan_amd import _abi, mechanics, network

    nff = _abi.MDY_NFF
    rows = []
    for wl in ("net500", "net40k"):
        net = network.generate(lib, wl)
        ctxs = {}
        for name, prec in (("fp64", _abi.MDY_PREC_FP64), ("mixed", _abi.MDY_PREC_MIXED)):
            c = mechanics.Context(precision=prec)
            c.set_network(net)
            ctxs[name] = c
        x0 = torch.from_numpy(np.ascontiguousarray(net.coord)).cuda()
        x1 = x0.clone()
        ctxs["fp64"].build_neighbors_dev(x1.data_ptr())
        ctxs["mixed"].build_neighbors_dev(x1.data_ptr())
        p1 = _abi.default_cg_params(net.n_beads, grad_tol=1.0, energy_rel_tol=0.0)
        r1 = ctxs["fp64"].minimize_dev(p1, x1.data_ptr())
        gen = torch.Generator(device="cuda")
        gen.manual_seed(5)
        for g in (1.0, 1e-1, 1e-2):
            p = _abi.default_cg_params(net.n_beads, grad_tol=g, energy_rel_tol=0.0, force_call_limit=400000)
            out = {"workload": wl, "grad_tol_pN": g}
            for name in ("fp64", "mixed", "fp64_control"):
                c = ctxs["mixed" if name == "mixed" else "fp64"]
                x = x1.clone()
                if name == "fp64_control":
                    x.add_(torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float64), alpha=1e-9)
                t0 = time.perf_counter()
                r = c.minimize_dev(p, x.data_ptr())
                torch.cuda.synchronize()
                out[name] = {"energy": float(r.energies_after[nff]), "iterations": int(r.num_iter),
                             "max_force_pN": float(r.grad_inf_norm), "reached": bool(r.grad_inf_norm < g),
                             "status": int(r.status), "seconds": time.perf_counter() - t0,
                             "us_per_iteration": 1e3 * c.last_ms() / max(int(r.num_iter), 1)}
            ef = out["fp64"]["energy"]
            out["mixed_vs_fp64_rel"] = abs(out["mixed"]["energy"] - ef) / abs(ef)
            out["control_vs_fp64_rel"] = abs(out["fp64_control"]["energy"] - ef) / abs(ef)
            print(f"[mixed_tight] {wl} gradTol {g}: mixed {out['mixed_vs_fp64_rel']:.2e} "
                  f"control {out['control_vs_fp64_rel']:.2e}", file=sys.stderr, flush=True)
            rows.append(out)
        rows.append({"workload": wl, "start": "FP64 minimum at 1 pN from the generated coordinates",
                     "start_iterations": int(r1.num_iter), "start_energy": float(r1.energies_after[nff])})
        for c in ctxs.values():
            c.close()
    print(json.dumps({"metric": "final-energy difference, mixed vs FP64, from identical starts by gradient tolerance",
                      "rows": rows}))


def bench_gradient(a, lib):
    """Differentiable equilibrium at 50k cylinders (5,000 filaments x 10
    cylinders, 5 um box, 20k linkers, 2k motors), FP64, in the quadrature
    mode (ill-conditioned excluded-volume pairs by quadrature, so second
    derivatives are not rounding noise). Minimization to 1 pN; the timed step
    is one Hessian-vector product enqueued back to back (mdy_hvp_dev_async;
    the exact mode's time and both modes waiting for each product beside it); the
    curvature that conjugate gradients on H meet at the 1 pN structure, with
    the energy's second difference along it (negative: not a minimum); and the
    implicit derivative of the extent (sum of squared filament end-to-end
    distances) with respect to the stretching, bending and linker stiffness at
    the tethered minimum (every bead held to the 1 pN structure by a spring),
    against central differences (step 1e-4 relative) of tethered
    re-minimizations."""
    import numpy as np
    import torch
    from medyan_amd import _abi, autodiff, mechanics, network

    def log(msg):
        print(f"[net50k_grad] {msg}", file=sys.stderr, flush=True)

    net = network.generate(lib, "net50k")
    ctx = mechanics.Context(coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    ctx.set_network(net)
    ctx.build_neighbors()
    params = _abi.default_cg_params(net.n_beads, grad_tol=1.0, energy_rel_tol=0.0)
    x = torch.from_numpy(np.ascontiguousarray(net.coord)).cuda()
    r = ctx.minimize_dev(params, x.data_ptr())
    min_ms = ctx.last_ms()
    log(f"minimized: {r.num_iter} iterations, {min_ms:.0f} ms, max|F| {r.grad_inf_norm:.3f}")
    x1 = x.clone()
    x1h = x1.cpu().numpy()
    n_quad = ctx.quadrature_pair_count(x1h)   # 0 in this mode; counted again below
    # H v throughput on a side stream: mdy_hvp_dev_async enqueues without a host
    # synchronization per call (failures checked once after the timed calls);
    # mdy_hvp_dev, which waits for each product, is timed beside it
    v = torch.randn(x1.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    hv = torch.empty_like(v)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())

    def time_hvp(call):
        with torch.cuda.stream(side):
            for _ in range(a.warmup):
                call(x1.data_ptr(), v.data_ptr(), hv.data_ptr(), side.cuda_stream)
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record(side)
            for _ in range(a.steps):
                call(x1.data_ptr(), v.data_ptr(), hv.data_ptr(), side.cuda_stream)
            t1.record(side)
        ctx.sync_errors(side.cuda_stream)
        side.synchronize()
        return t0.elapsed_time(t1) / a.steps

    exact_ms = time_hvp(ctx.hvp_dev_async)
    exact_sync_ms = time_hvp(ctx.hvp_dev)
    ctx.set_opts(coplanar_mode=_abi.MDY_COPLANAR_QUADRATURE)
    n_quad = ctx.quadrature_pair_count(x1h)
    hvp_ms = time_hvp(ctx.hvp_dev_async)
    hvp_sync_ms = time_hvp(ctx.hvp_dev)
    n_ev = len(ctx.exvol_table()[1])
    n_bd = len(ctx.boundary_table()[1])
    tables = (B_FS * len(net.fs_k) + B_FB * len(net.fb_k) + B_LM * (len(net.ls_k) + len(net.ms_k))
              + B_BS * len(net.bs_k) + B_BB * len(net.bb_k) + B_EV * n_ev + B_BD * n_bd)
    hvp_bytes = tables + 48 * net.n_beads + 24 * net.n_beads
    log(f"H v: {hvp_ms:.3f} ms in the quadrature mode ({n_quad} of {n_ev} pairs), {exact_ms:.3f} ms in the exact mode; "
        f"waiting for each product: {hvp_sync_ms:.3f} / {exact_sync_ms:.3f} ms")

    fil = net.cyl_filament
    first = np.searchsorted(fil, np.arange(fil.max() + 1))
    last = np.r_[first[1:], len(fil)] - 1
    ends = np.stack([net.cyl_bead1[first] // 3, net.cyl_bead2[last] // 3], 1)

    def extent(z):
        Z = z.reshape(-1, 3)
        d = Z[ends[:, 1]] - Z[ends[:, 0]]
        return (d * d).sum()

    def second_difference(z, p, h=0.3):
        e0, _ = ctx.energy(z)
        ep, _ = ctx.energy(z + h * p)
        em, _ = ctx.energy(z - h * p)
        return (ep + em - 2 * e0) / h ** 2

    # not a minimum: conjugate gradients on H (rigid motions of free clusters projected out) meet negative curvature
    xg = x1.clone().requires_grad_(True)
    (gl,) = torch.autograd.grad(extent(xg), xg)
    curvature = None
    try:
        autodiff.stiffness_sensitivity(ctx, x1, gl, rtol=1e-10, maxiter=20000)
    except autodiff.NotAMinimum as e:
        curvature = dict(hessian=e.curvature, energy_second_difference=second_difference(x1h, e.direction.cpu().numpy()))
    log(f"curvature at the 1 pN structure: {curvature}")
    # tethered minimum, implicit derivative, central differences of re-minimizations
    tethered = {}
    for tether in (1.0, 0.1):
        t0 = time.perf_counter()
        try:
            xs, fm, info = autodiff.tethered_minimum(ctx, x1, tether, ftol=1e-9)
            t_min = time.perf_counter() - t0
            xg = xs.clone().requires_grad_(True)
            (gl,) = torch.autograd.grad(extent(xg), xg)
            t0 = time.perf_counter()
            d, inf = autodiff.stiffness_sensitivity(ctx, xs, gl, rtol=1e-12, maxiter=200000, tether=tether)
            t_imp = time.perf_counter() - t0
            d = d.cpu().numpy()
            rec = dict(newton_steps=info["steps"], newton_cg_iterations=info["cg_iterations"], max_force=fm,
                       minimization_s=t_min, implicit_s=t_imp, implicit_cg_iterations=inf["cg_iterations"],
                       relative_residual=inf["relative_residual"])
            for k, name in ((0, "stretching"), (1, "bending"), (2, "linker")):
                vals = []
                for sg in (1, -1):
                    sc = np.ones(_abi.MDY_NFF)
                    sc[k] += sg * 1e-4
                    c2 = mechanics.Context(coplanar_mode=_abi.MDY_COPLANAR_QUADRATURE)
                    c2.set_network(autodiff.scale_stiffness(net, sc))
                    c2.build_neighbors(net.coord)
                    x2, f2, _ = autodiff.tethered_minimum(c2, x1, tether, x0=xs, ftol=1e-9)
                    vals.append(float(extent(x2)))
                    c2.close()
                fd = (vals[0] - vals[1]) / 2e-4
                rec[name] = dict(implicit=float(d[k]), central_difference=fd,
                                 relative_difference=abs(fd - d[k]) / np.abs(d[:3]).max())
            tethered[str(tether)] = rec
        except autodiff.NotAMinimum as e:
            tethered[str(tether)] = dict(error=str(e))
        log(f"tether {tether} pN/nm: {tethered[str(tether)]}")
    ctx.set_opts(coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    out = {
        "metric": "Hessian-vector products/s at 50k cylinders (FP64, quadrature mode)",
        "value": 1e3 / hvp_ms,
        "unit": "HVP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": hvp_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded network generator)",
        "config": {
            "workload": (f"net50k: {net.n_cyl} cylinders, {net.n_beads} beads, {len(net.ls_k)} linkers, "
                         f"{len(net.ms_k)} motors, {n_ev} excluded-volume pairs ({n_quad} by quadrature at the "
                         f"1 pN structure)"),
            "minimization_ms": min_ms,
            "minimization_iterations": int(r.num_iter),
            "hvp_ms_exact_mode": exact_ms,
            "hvp_ms_waiting_each_call": hvp_sync_ms,
            "hvp_ms_exact_mode_waiting_each_call": exact_sync_ms,
            "extent_nm2": float(extent(x1)),
            "curvature_at_1pN": curvature,
            "tethered_implicit_derivative": tethered,
            "note": ("the 1 pN structure is not a minimum (negative curvature of H, confirmed by the energy), so "
                     "dx*/ds exists only for a regularized problem: the tethered minimum (every bead held to the "
                     "1 pN structure by a spring of the stated stiffness), whose implicit derivative is checked "
                     "against central differences of tethered re-minimizations"),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_hvp + k_hvp_gather (+ k_ev_classify, k_quad)",
            "achieved": hvp_bytes / (hvp_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": hvp_bytes / (hvp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": hvp_bytes,
            "launch_ms": hvp_ms,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    try:
        main()
    finally:
        # device memory released while the HIP runtime is still up
        if "medyan_amd.mechanics" in sys.modules:
            sys.modules["medyan_amd.mechanics"].close_all()
