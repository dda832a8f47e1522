"""The quadrature mode (MDY_COPLANAR_QUADRATURE) on the device: excluded-volume
pairs whose closed form is ill-conditioned (squared normalized triple
product q < 1e-3) are evaluated by composite Gauss-Legendre quadrature of the
defining integral. No reference counterpart: checked against mathematics.

* energies and forces: every force field but excluded volume equals the
  exact mode bit for bit; the excluded-volume energy and the forces equal the
  oracle's closed form on the unflagged pairs plus an independent numpy
  quadrature (a different rule, tests/ev_quadrature.py) on the flagged ones;
* H v equals (extrapolated) central differences of the mode's own forces on
  every bead (the exact mode's must leave out the beads of ill-conditioned
  pairs), is
  symmetric, and grad E_k . w equals (extrapolated) central differences of
  the energies;
* at the 50k-cylinder network stopped at 1 pN: curvatures of H (a Lanczos
  eigenvector; the direction where conjugate gradients meet negative
  curvature) agree with second differences of the energy, so the structure
  is not a minimum; the implicit derivative of the extent with respect to
  the stretching, bending and linker stiffness at the tethered minimum
  equals central differences of tethered re-minimizations;
* minimizations and the allocation-free energy refuse the mode.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import ev_quadrature as evq
from coplanar_fixture import as_network
from medyan_amd import _abi, autodiff, mechanics, network

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF
EXACT, QUAD = _abi.MDY_COPLANAR_EXACT, _abi.MDY_COPLANAR_QUADRATURE


def log(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def nets(lib):
    cop = as_network()
    cop.exvol_cutoff = 150.0
    return {"net500": network.generate(lib, "net500"), "coplanar": cop}


def ctx_for(net, mode, coord=None):
    ctx = mechanics.Context(0, coplanar_mode=mode)
    ctx.set_network(net)
    ctx.build_neighbors(coord)
    return ctx


def dev(a):
    return torch.tensor(np.ascontiguousarray(a, np.float64), device="cuda")


@pytest.mark.parametrize("case", ["net500", "coplanar"])
def test_energy_forces_against_closed_form_plus_quadrature(oracle, nets, case):
    net = nets[case]
    x = net.coord
    cq, ce = ctx_for(net, QUAD), ctx_for(net, EXACT)
    e_q, per_q = cq.energy(x)
    e_e, per_e = ce.energy(x)
    f_q, f_e = cq.forces(x), ce.forces(x)
    s = oracle.system(net, coplanar_mode=EXACT)
    bs, kr, eq = s.ev_table
    fl = evq.flagged(bs, x)
    log(f"{case}: {fl.sum()} of {len(fl)} pairs by quadrature")
    assert fl.sum() > 0 and cq.quadrature_pair_count(x) == fl.sum() and ce.quadrature_pair_count(x) == 0
    for k in range(NFF):
        if k != _abi.MDY_FF_EV:
            assert per_q[k] == per_e[k], k
    X = x.reshape(-1, 3)
    e_ev = 0.0
    f_quad = np.zeros_like(x).reshape(-1, 3)
    oracle.set_exact(1)
    try:
        for j in np.nonzero(~fl)[0]:
            c = X[bs[j] // 3]
            e_ev += oracle.ev_pair_energy(c[0], c[1], c[2], c[3], kr[j] * eq[j, 0] * eq[j, 1])
        for j in np.nonzero(fl)[0]:
            c = X[bs[j] // 3]
            E, F = evq.pair(c[0], c[1], c[2], c[3], kr[j] * eq[j, 0] * eq[j, 1])
            e_ev += E
            f_quad[bs[j] // 3] += F
        # the exact mode's forces without any excluded volume, then the unflagged pairs' closed form
        ev_all = oracle.ev_forces(x, bs.ravel(), kr, eq.ravel(), coplanar_mode=EXACT)
        ev_ok = oracle.ev_forces(x, bs[~fl].ravel(), kr[~fl], eq[~fl].ravel(), coplanar_mode=EXACT)
    finally:
        oracle.set_exact(0)
    assert per_q[_abi.MDY_FF_EV] == pytest.approx(e_ev, rel=1e-12)
    assert e_q == pytest.approx(sum(per_q), rel=1e-14)
    expect = f_e - ev_all + ev_ok + f_quad.ravel()
    err = np.abs(f_q - expect).max() / np.abs(expect).max()
    log(f"{case}: forces within {err:.1e} of max|F|; exact mode differs by "
        f"{np.abs(f_e - expect).max() / np.abs(expect).max():.1e}")
    assert err <= 1e-11
    cq.close()
    ce.close()


def _hvp_check(ctx, x, rng, ndir=3, h=1e-3, tol=1e-6):
    """H v against Richardson-extrapolated central differences of the forces
    (steps h and h/2: truncation O(h^4))."""
    xd = dev(x)

    def cd(v, step):
        return -(ctx.forces(x + step * v) - ctx.forces(x - step * v)) / (2 * step)

    for k in range(ndir):
        v = rng.normal(size=x.size)
        hv = autodiff.hvp(ctx, xd, dev(v)).cpu().numpy()
        fd = (4.0 * cd(v, h / 2) - cd(v, h)) / 3.0
        err = np.abs(hv - fd).max() / np.abs(fd).max()
        log(f"direction {k}: max |Hv - CD| / max|CD| = {err:.2e} over all {x.size // 3} beads")
        assert np.all(np.isfinite(hv)) and err < tol
    u, w = rng.normal(size=x.size), rng.normal(size=x.size)
    hu = autodiff.hvp(ctx, xd, dev(u)).cpu().numpy()
    hw = autodiff.hvp(ctx, xd, dev(w)).cpu().numpy()
    assert abs(u @ hw - w @ hu) <= 1e-10 * np.abs(u).sum() * np.abs(hw).max()


@pytest.mark.parametrize("case", ["net500", "coplanar"])
def test_hvp_and_jvp_are_derivatives(nets, case):
    net = nets[case]
    x = net.coord
    ctx = ctx_for(net, QUAD)
    rng = np.random.default_rng(23)
    _hvp_check(ctx, x, rng)
    w = rng.normal(size=x.size)
    g = autodiff.energy_jvp(ctx, dev(x), dev(w)).cpu().numpy()

    def cd(step):
        _, ep = ctx.energy(x + step * w)
        _, em = ctx.energy(x - step * w)
        return (np.asarray(ep) - np.asarray(em)) / (2 * step)

    fd = (4.0 * cd(5e-4) - cd(1e-3)) / 3.0
    scale = np.abs(fd).max()
    assert np.abs(g[:NFF] - fd).max() <= 1e-6 * scale
    ctx.close()


@pytest.fixture(scope="module")
def net50k_1pn(lib):
    net = network.generate(lib, "net50k")
    ce = ctx_for(net, EXACT)
    x = dev(net.coord)
    r = ce.minimize_dev(_abi.default_cg_params(net.n_beads, energy_rel_tol=0.0), x.data_ptr())
    assert r.status == 0 and r.grad_inf_norm < 1.0
    ce.close()
    return net, x.cpu().numpy()


def test_net50k_hvp_all_beads(net50k_1pn):
    net, x = net50k_1pn
    ctx = ctx_for(net, QUAD, net.coord)
    log(f"net50k: {ctx.quadrature_pair_count(x)} of {len(ctx.exvol_table()[1])} pairs by quadrature")
    _hvp_check(ctx, x, np.random.default_rng(29))
    ctx.close()


def _second_difference(ctx, x, v, h):
    e0, _ = ctx.energy(x)
    ep, _ = ctx.energy(x + h * v)
    em, _ = ctx.energy(x - h * v)
    return (ep + em - 2 * e0) / h ** 2


def test_net50k_curvatures_match_energy(net50k_1pn):
    """The Hessian's curvatures along the Lanczos eigenvector and along the
    direction where conjugate gradients on H meet negative curvature agree
    with second differences of the energy: the structure stopped at 1 pN is
    not a minimum (H is indefinite), and that is a property of the energy,
    not of rounding."""
    net, x = net50k_1pn
    ctx = ctx_for(net, QUAD, net.coord)
    xd = dev(x)
    lam, vec = autodiff.lowest_curvature(ctx, xd, steps=300)
    sd = _second_difference(ctx, x, vec.cpu().numpy(), 0.3)
    log(f"net50k at 1 pN: Lanczos eigenvalue {lam:.6e}, energy second difference {sd:.6e}")
    assert abs(sd - lam) <= 1e-2 * abs(lam) + 1e-6
    ends = _filament_ends(net)
    xg = xd.clone().requires_grad_(True)
    (gl,) = torch.autograd.grad(_extent(xg, ends), xg)
    with pytest.raises(autodiff.NotAMinimum) as e:
        autodiff.stiffness_sensitivity(ctx, xd, gl, rtol=1e-10, maxiter=20000)
    p = e.value.direction.cpu().numpy()
    sd = _second_difference(ctx, x, p, 0.3)
    log(f"net50k at 1 pN: conjugate-gradient curvature {e.value.curvature:.6e}, energy second difference {sd:.6e}")
    assert e.value.curvature < 0 and sd < 0
    assert abs(sd - e.value.curvature) <= 2e-2 * abs(e.value.curvature)
    ctx.close()


def _filament_ends(net):
    fil = net.cyl_filament
    first = np.searchsorted(fil, np.arange(fil.max() + 1))
    last = np.r_[first[1:], len(fil)] - 1
    return np.stack([net.cyl_bead1[first] // 3, net.cyl_bead2[last] // 3], 1)


def _extent(z, ends):
    """Sum of squared filament end-to-end distances."""
    Z = z.reshape(-1, 3)
    d = Z[ends[:, 1]] - Z[ends[:, 0]]
    return (d * d).sum()


def test_net50k_tethered_implicit_derivative_matches_reminimization(net50k_1pn):
    """d(extent)/d(k_stretch, k_bend, k_linker) at 5,000 x 10 cylinders, by
    implicit differentiation of the tethered minimum (every bead held to the
    1 pN structure by 1 pN/nm), against central differences (step 1e-4
    relative) of tethered re-minimizations converged to 1e-9 pN."""
    net, x1 = net50k_1pn
    tether = 1.0
    ctx = ctx_for(net, QUAD, net.coord)
    xr = dev(x1)
    xs, fm, info = autodiff.tethered_minimum(ctx, xr, tether, ftol=1e-9)
    assert fm < 1e-9
    ends = _filament_ends(net)
    xg = xs.clone().requires_grad_(True)
    (gl,) = torch.autograd.grad(_extent(xg, ends), xg)
    d, inf = autodiff.stiffness_sensitivity(ctx, xs, gl, rtol=1e-12, maxiter=50000, tether=tether)
    d = d.cpu().numpy()
    log(f"net50k tethered: {info['steps']} Newton steps; implicit dL/ds {d[:3]} ({inf['cg_iterations']} CG iterations)")
    for k in (0, 1, 2):
        vals = []
        for sg in (1, -1):
            sc = np.ones(NFF)
            sc[k] += sg * 1e-4
            c2 = ctx_for(autodiff.scale_stiffness(net, sc), QUAD, net.coord)
            x2, f2, _ = autodiff.tethered_minimum(c2, xr, tether, x0=xs, ftol=1e-9)
            assert f2 < 1e-9
            vals.append(float(_extent(x2, ends)))
            c2.close()
        fd = (vals[0] - vals[1]) / 2e-4
        log(f"force field {k}: implicit {d[k]:.10g}, central difference {fd:.10g}")
        assert abs(d[k] - fd) <= 1e-6 * np.abs(d[:3]).max()
    ctx.close()


def test_mode_refused_where_unsupported(nets):
    net = nets["net500"]
    ctx = ctx_for(net, QUAD)
    with pytest.raises(mechanics.MdyError):
        ctx.minimize(_abi.default_cg_params(net.n_beads))
    x = dev(net.coord)
    out = torch.empty(NFF + 1, dtype=torch.float64, device="cuda")
    rc = ctx.lib.mdy_energy_dev_noalloc(ctx.h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), None)
    assert rc == _abi.MDY_ERR_ARG
    with pytest.raises(mechanics.MdyError):
        ctx.set_opts(precision=_abi.MDY_PREC_MIXED)
    ctx.close()
    with pytest.raises(mechanics.MdyError):
        mechanics.Context(0, precision=_abi.MDY_PREC_MIXED, coplanar_mode=QUAD)
