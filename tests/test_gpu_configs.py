"""The large workloads on the device against the oracle: the 200k-cylinder
network of the successive-minimization benchmark, the 256-network batched
ensemble and the 50k-cylinder network of the differentiable form.

* net200k: neighbor list and excluded-volume / boundary tables bit-exact;
  energies and forces bit for bit against the exact-mode oracle and within
  1e-10 of the reference-order oracle (outside ill-conditioned pairs); the
  first 10 CG iterations bit for bit; the minimized state's forces equal the
  oracle's at the returned coordinates; then one successive step (5 nm bead
  displacement, neighbor-list rebuild, minimization) checked the same way.
* ensemble256: one batched launch in the mixed form; every network's
  returned forces against the oracle's at its returned coordinates (the
  mixed form's stated error, 1e-5 of the largest pair force scale), and the
  1 pN criterion by the oracle's own FP64 forces; then the FP64 batched form,
  whose returned forces equal the exact-mode oracle's bit for bit.
* net50k: H v at the minimum against central differences of the oracle's
  exact-gradient forces along three random directions.
"""
import numpy as np
import pytest
import torch

from medyan_amd import _abi, autodiff, ensemble, mechanics, network
from test_gpu_parity import bits_equal, ill_conditioned_beads

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


def log(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def net200k(lib):
    log("generating net200k")
    return network.generate(lib, "net200k")


def ctx_for(net, coord=None, precision=_abi.MDY_PREC_FP64, coplanar=_abi.MDY_COPLANAR_REFERENCE):
    ctx = mechanics.Context(0, precision=precision, coplanar_mode=coplanar)
    ctx.set_network(net)
    ctx.build_neighbors(coord)
    return ctx


def check_tables(ctx, oracle, net, coord):
    np.testing.assert_array_equal(ctx.pairs(), oracle.nl_pairs(net, coord))
    s = oracle.system(net, coord)
    bs, kr = ctx.exvol_table()
    obs, okr, oeq = s.ev_table
    np.testing.assert_array_equal(bs, obs)
    assert bits_equal(kr, okr * oeq[:, 0] * oeq[:, 1])
    nn, bbs = ctx.boundary_table()
    onn, obbs = s.bd_table
    np.testing.assert_array_equal(nn, onn)
    np.testing.assert_array_equal(bbs, obbs)
    return s


def test_net200k_tables_energies_forces(oracle, net200k):
    net = net200k
    log(f"net200k: {net.n_cyl} cylinders, tables")
    ctx = ctx_for(net)
    s = check_tables(ctx, oracle, net, net.coord)
    e, per = ctx.energy(net.coord)
    f = ctx.forces(net.coord)
    eo, pero = s.energy()
    fo = s.forces()
    assert abs(e - eo) <= 1e-10 * abs(eo)
    np.testing.assert_allclose(per, pero, rtol=1e-10, atol=1e-10 * abs(eo))
    ok = np.ones(net.n_beads, bool)
    ill = ill_conditioned_beads(net, s)
    ok[ill] = False
    diff = np.abs(f - fo).reshape(-1, 3).max(1)
    assert diff[ok].max() <= 1e-10 * np.abs(fo).max()
    oracle.set_exact(1)
    try:
        eo, _ = s.energy()
        assert bits_equal([e], [eo])
        assert bits_equal(f, s.forces())
    finally:
        oracle.set_exact(0)
    log(f"net200k: {len(ctx.exvol_table()[1])} pairs, {len(ill)} beads in ill-conditioned pairs")
    ctx.close()


def test_net200k_minimization_and_successive_step(oracle, net200k):
    net = net200k
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    ctx = ctx_for(net)
    log("net200k: minimizing")
    x, f, res = ctx.minimize(p, trace_cap=200000)
    tr = ctx.trace()
    assert res.status == 0 and res.error_flags == 0 and res.grad_inf_norm < 1.0
    oracle.set_exact(1)
    try:
        s = oracle.system(net)
        log(f"net200k: {res.num_iter} iterations; oracle's first 10")
        _, _, _, tro = s.minimize(p, iter_cap=10, trace_cap=20)
        for key in ("lambda_", "energy", "maxf", "beta"):
            assert bits_equal(tr[key][:10], tro[key]), key
        assert bits_equal(f, s.forces(x))
        eo, _ = s.energy(x, individual=True)
        assert bits_equal([res.energies_after[NFF]], [eo])
        # one successive step: seeded 5 nm displacement of the minimum, lists rebuilt there
        x1 = x + np.random.default_rng(11).normal(0.0, 5.0, x.shape)
        ctx.build_neighbors(x1)
        s1 = check_tables(ctx, oracle, net, x1)
        log("net200k: successive step")
        x2, f2, res2 = ctx.minimize(p, x1)
        assert res2.status == 0 and res2.error_flags == 0 and res2.grad_inf_norm < 1.0
        assert bits_equal(f2, s1.forces(x2))
        assert np.abs(f2).max() == res2.grad_inf_norm
    finally:
        oracle.set_exact(0)
    log(f"net200k successive: {res2.num_iter} iterations, max|F| {res2.grad_inf_norm:.3f}")
    ctx.close()


@pytest.fixture(scope="module")
def ens(lib):
    log("generating ensemble256")
    return ensemble.make_ensemble(lib, 256)


def batch_minimize(nets, precision):
    ctxs = [ctx_for(n, precision=precision) for n in nets]
    xs = [torch.tensor(np.ascontiguousarray(n.coord), device="cuda") for n in nets]
    fs = [torch.zeros_like(x) for x in xs]
    params = [_abi.default_cg_params(n.n_beads, energy_rel_tol=0.0) for n in nets]
    res, ms, tot = mechanics.minimize_batch(ctxs, params, [x.data_ptr() for x in xs], [f.data_ptr() for f in fs],
                                            threads_per_block=64, first_pass_calls=500)
    out = [(x.cpu().numpy(), f.cpu().numpy()) for x, f in zip(xs, fs)]
    for c in ctxs:
        c.close()
    return res, out, tot


def test_ensemble256_mixed_batch_against_oracle(oracle, ens):
    log(f"ensemble256 mixed: {sum(n.n_cyl for n in ens)} cylinders")
    res, out, tot = batch_minimize(ens, _abi.MDY_PREC_MIXED)
    log(f"ensemble256 mixed: one batched minimization in {tot:.0f} ms")
    worst_err, worst_f = 0.0, 0.0
    for i, (net, r, (x, f)) in enumerate(zip(ens, res, out)):
        assert r.status == 0 and r.error_flags == 0 and r.grad_inf_norm < 1.0, i
        s = oracle.system(net)
        fo = s.forces(x)
        # FP32 pair forces: error relative to the pair-force scale at the start
        scale = max(np.abs(s.forces(net.coord)).max(), 1.0)
        err = np.abs(f - fo).max() / scale
        worst_err = max(worst_err, err)
        worst_f = max(worst_f, np.abs(fo).max())
        assert err <= 1e-5, (i, err)
        assert np.abs(fo).max() < 1.0 + 1e-5 * scale, i
    log(f"ensemble256 mixed: worst force error {worst_err:.2e} of the pair-force scale, "
        f"worst FP64 max|F| at the returned coordinates {worst_f:.4f} pN")


def test_ensemble256_fp64_batch_bitwise(oracle, ens):
    res, out, tot = batch_minimize(ens, _abi.MDY_PREC_FP64)
    log(f"ensemble256 FP64: one batched minimization in {tot:.0f} ms")
    oracle.set_exact(1)
    try:
        for i, (net, r, (x, f)) in enumerate(zip(ens, res, out)):
            assert r.status == 0 and r.error_flags == 0 and r.grad_inf_norm < 1.0, i
            s = oracle.system(net)
            assert bits_equal(f, s.forces(x)), i
            assert np.abs(f).max() == r.grad_inf_norm, i
    finally:
        oracle.set_exact(0)


def test_net50k_hvp_against_force_differences(oracle, lib):
    net = network.generate(lib, "net50k")
    log(f"net50k: {net.n_cyl} cylinders, minimizing")
    ctx = ctx_for(net, coplanar=_abi.MDY_COPLANAR_EXACT)
    x = torch.tensor(np.ascontiguousarray(net.coord), device="cuda")
    r = ctx.minimize_dev(_abi.default_cg_params(net.n_beads, energy_rel_tol=0.0), x.data_ptr())
    assert r.status == 0 and r.grad_inf_norm < 1.0
    xm = x.cpu().numpy()
    s = oracle.system(net, net.coord, coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    ok = np.ones(net.n_beads, bool)
    ok[ill_conditioned_beads(_at(net, xm), s)] = False
    rng = np.random.default_rng(17)
    h = 1e-3
    for k in range(3):
        v = rng.normal(size=xm.size)
        hv = autodiff.hvp(ctx, x, torch.tensor(v, device="cuda")).cpu().numpy()
        fd = -(s.forces(xm + h * v) - s.forces(xm - h * v)) / (2 * h)
        err = np.abs(hv - fd).reshape(-1, 3)[ok].max() / np.abs(fd).max()
        log(f"net50k direction {k}: max |Hv - CD| / max|CD| = {err:.2e} on {ok.sum()} of {net.n_beads} beads")
        assert np.all(np.isfinite(hv))
        assert err < 2e-6
    ctx.close()


def _at(net, x):
    n = net.copy()
    n.coord = np.ascontiguousarray(x)
    return n
