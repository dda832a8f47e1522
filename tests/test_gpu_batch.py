"""Batched minimization (mdy_minimize_batch): independent networks in one
cooperative launch, one cluster of blocks per network.

Every decision of the minimizer comes from exact sums and a fixed-order force
gather, so a network's trajectory does not depend on how many blocks run it:
each batched result must equal the single-network minimization bit for bit
(coordinates, forces, total energy, call counts), in FP64 and in the mixed
form. The per-force-field energy reports are plain double sums whose
association follows the block count; they agree to 1e-12.
The single-network path is itself pinned to the oracle bit for bit.
"""
import numpy as np
import pytest
import torch

from medyan_amd import _abi, ensemble, mechanics, network

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


def nets_small(lib):
    return [network.generate(lib, "net500"),
            network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=2),
            network.generate(lib, "net500", n_filaments=60, cyl_per_filament=3, box=500.0, n_linkers=40,
                             n_motors=4, seed=3),
            ensemble.ensemble_network(lib, 1200, 0),
            ensemble.ensemble_network(lib, 3000, 1),
            network.generate(lib, None, seed=11, n_filaments=3, cyl_per_filament=4, box=800.0, n_linkers=2,
                             n_motors=0, n_branches=0)]


def ctx_for(net, precision):
    c = mechanics.Context(precision=precision)
    c.set_network(net)
    c.build_neighbors()
    return c


def same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("precision,tpb,passes,ls", [(_abi.MDY_PREC_FP64, 64, 0, 0), (_abi.MDY_PREC_FP64, 256, 0, 0),
                                                     (_abi.MDY_PREC_MIXED, 128, 0, 0), (_abi.MDY_PREC_FP64, 64, 40, 0),
                                                     (_abi.MDY_PREC_MIXED, 64, 40, 0),
                                                     (_abi.MDY_PREC_FP64, 64, -2, 0),
                                                     (_abi.MDY_PREC_MIXED, 128, -4, 0),
                                                     (_abi.MDY_PREC_FP64, 128, 0, _abi.MDY_LS_QUADRATIC)])
def test_batch_equals_single_minimizations(lib, precision, tpb, passes, ls):
    """passes > 0: a first pass capped at 40 force calls (six networks get
    64 blocks or more each, so the pass runs uncapped). passes = -k: each
    pass ends once max(1, m / k) of its m networks have finished (-2: after
    3, then 1, ...; -4: after every completion, six passes). ls: line search."""
    nets = nets_small(lib)
    print(f"[gpu] batch of {len(nets)}, precision {precision}, {tpb} threads per block, first pass {passes}, "
          f"line search {ls}", flush=True)
    singles = []
    for net in nets:
        c = ctx_for(net, precision)
        x, f, r = c.minimize(_abi.default_cg_params(net.n_beads, line_search=ls))
        singles.append((x, f, r))
    ctxs = [ctx_for(net, precision) for net in nets]
    xs = [torch.tensor(net.coord, device="cuda") for net in nets]
    fs = [torch.zeros_like(x) for x in xs]
    params = [_abi.default_cg_params(net.n_beads, line_search=ls) for net in nets]
    res, ms, tot = mechanics.minimize_batch(ctxs, params, [x.data_ptr() for x in xs], [f.data_ptr() for f in fs],
                                            threads_per_block=tpb, first_pass_calls=passes)
    torch.cuda.synchronize()
    for i, (x1, f1, r1) in enumerate(singles):
        r = res[i]
        assert r.status == 0 and r.error_flags == r1.error_flags
        assert (r.num_iter, r.num_energy_call, r.num_force_call, r.force_call_limit) == (
            r1.num_iter, r1.num_energy_call, r1.num_force_call, r1.force_call_limit)
        # the total is an exact sum; the per-force-field reports are plain
        # double sums whose order follows the cluster's block count
        assert r.energies_after[NFF] == r1.energies_after[NFF]
        np.testing.assert_allclose(np.array(r.energies_after[:NFF]), np.array(r1.energies_after[:NFF]),
                                   rtol=1e-12, atol=1e-12 * abs(r1.energies_after[NFF]))
        assert r.grad_inf_norm == r1.grad_inf_norm
        assert same(xs[i].cpu().numpy(), x1), f"network {i}: coordinates differ"
        assert same(fs[i].cpu().numpy(), f1), f"network {i}: forces differ"
        assert 0 < ms[i]
    print(f"[gpu] batch {tot:.1f} ms; per network {np.round(ms, 2)}")


@pytest.mark.parametrize("passes", [0, 3, -2])
def test_batch_larger_than_resident_blocks(lib, passes):
    """More networks than resident blocks: the call runs several launches;
    with passes, capped networks are suspended and resumed (3: caps 3, 12,
    48, ... force calls; -2: each launch's pass ends when half of its
    networks have finished)."""
    base = [network.generate(lib, None, seed=100 + i, n_filaments=3, cyl_per_filament=3, box=700.0, n_linkers=2,
                             n_motors=0, n_branches=0) for i in range(4)]
    n = 256 + 8   # above one 256-thread block per CU on 256 CUs
    nets = [base[i % 4] for i in range(n)]
    ctxs = [ctx_for(net, _abi.MDY_PREC_FP64) for net in nets]
    xs = [torch.tensor(net.coord, device="cuda") for net in nets]
    params = [_abi.default_cg_params(net.n_beads) for net in nets]
    res, ms, tot = mechanics.minimize_batch(ctxs, params, [x.data_ptr() for x in xs], threads_per_block=256,
                                            first_pass_calls=passes)
    torch.cuda.synchronize()
    ref = [ctx_for(b, _abi.MDY_PREC_FP64).minimize(_abi.default_cg_params(b.n_beads)) for b in base]
    for i in range(n):
        r, r1 = res[i], ref[i % 4][2]
        assert r.status == 0
        assert (r.num_iter, r.num_force_call, r.force_call_limit) == (r1.num_iter, r1.num_force_call,
                                                                      r1.force_call_limit)
        assert same(xs[i].cpu().numpy(), ref[i % 4][0])
    print(f"[gpu] {n} networks in {tot:.1f} ms")


def test_batch_argument_errors(lib):
    net = network.generate(lib, None, seed=5, n_filaments=3, cyl_per_filament=3, box=700.0, n_linkers=2,
                           n_motors=0, n_branches=0)
    c = ctx_for(net, _abi.MDY_PREC_FP64)
    x = torch.tensor(net.coord, device="cuda")
    p = _abi.default_cg_params(net.n_beads)
    with pytest.raises(mechanics.MdyError) as e:
        mechanics.minimize_batch([c, c], [p, p], [x.data_ptr(), x.data_ptr()])
    assert e.value.code == _abi.MDY_ERR_ARG
    with pytest.raises(mechanics.MdyError) as e:
        mechanics.minimize_batch([c], [p], [x.data_ptr()], threads_per_block=96)
    assert e.value.code == _abi.MDY_ERR_ARG
    m = mechanics.Context(precision=_abi.MDY_PREC_MIXED)
    m.set_network(net)
    m.build_neighbors()
    with pytest.raises(mechanics.MdyError) as e:
        mechanics.minimize_batch([c, m], [p, p], [x.data_ptr(), x.data_ptr()])
    assert e.value.code == _abi.MDY_ERR_ARG


def test_autotuner_choice_leaves_results_unchanged(lib):
    """medyan_amd.ensemble.autotune times each candidate (threads per block,
    pass rule, cluster-share exponent) on the ensemble and returns the
    fastest; whatever it picks, the minimizations equal the default launch's
    bit for bit."""
    nets = nets_small(lib)
    ctxs = [ctx_for(net, _abi.MDY_PREC_MIXED) for net in nets]
    x0 = [torch.tensor(net.coord, device="cuda") for net in nets]
    xs = [x.clone() for x in x0]
    ptrs = [x.data_ptr() for x in xs]
    params = [_abi.default_cg_params(net.n_beads, energy_rel_tol=0.0) for net in nets]

    def restore():
        for x, x_0 in zip(xs, x0):
            x.copy_(x_0)
        torch.cuda.synchronize()

    best, times = ensemble.autotune(ctxs, params, ptrs, restore, threads=(64, 128), first_pass=(-2, 40),
                                    exponents=(1.0, 1.5), log=lambda m: print(f"[gpu] {m}", flush=True))
    assert len(times) == 2 + 1 + 1 and best["threads_per_block"] in (64, 128)
    assert best["first_pass_calls"] in (-2, 40) and len(best["weights"]) == len(nets)
    restore()
    res_b, _, _ = mechanics.minimize_batch(ctxs, params, ptrs, **best)
    got = [x.cpu().numpy() for x in xs]
    restore()
    res_d, _, _ = mechanics.minimize_batch(ctxs, params, ptrs)
    for i, x in enumerate(xs):
        assert same(got[i], x.cpu().numpy()), i
        assert res_b[i].num_iter == res_d[i].num_iter and res_b[i].grad_inf_norm == res_d[i].grad_inf_norm
