"""Differentiable mechanics: torch.autograd functions over the device energy
and forces, the Hessian-vector-product kernel, per-force-field energy
derivatives and implicit derivatives of minimized structures.

Second derivatives have no counterpart in the reference, so they are pinned
by mathematics, against the oracle's forces and energies:
* H v equals central differences of the oracle's exact-gradient forces
  (outside ill-conditioned excluded-volume pairs, where FP64 force
  differences lose their digits); H is symmetric, linear, and maps rigid
  translations to zero;
* grad E_k . w equals central differences of the oracle's individual
  energies (w zero on the beads of ill-conditioned pairs), and its sum
  equals -F . w;
* autograd's first derivative is the force kernel's output, negated, bit
  for bit; its second derivative is the HVP kernel's output bit for bit;
* the implicit derivative of a minimized structure's end-to-end extent with
  respect to the stretching, bending, linker and motor stiffness equals
  central differences (step 1e-4 relative) of re-minimized structures,
  converged to max|F| < 1e-10 pN by Newton steps whose fixed point is set by
  the force kernel alone.
"""
import numpy as np
import pytest
import torch

from medyan_amd import _abi, autodiff, mechanics, network
from test_gpu_parity import ill_conditioned_beads

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


def ladder(seed=0, nb=6):
    """Two filaments (5 cylinders each) joined by 8 linkers pulled to 80% of
    their initial length and one motor: rigid up to global motions, away from
    the boundary and out of excluded-volume range."""
    rng = np.random.default_rng(seed)
    t = np.array([1230 + 108.0 * j for j in range(nb)])
    a = np.stack([t, np.full(nb, 1500.0), np.full(nb, 1500.0)], 1) + rng.normal(0, 4.0, (nb, 3))
    b = np.stack([t + 20, 1650.0 + 0.05 * (t - 1500), np.full(nb, 1520.0)], 1) + rng.normal(0, 4.0, (nb, 3))
    links = []
    for (ca, pa, cb, pb) in [(0, 0.2, 0, 0.7), (1, 0.3, 0, 0.9), (1, 0.8, 2, 0.1), (2, 0.5, 2, 0.5),
                             (3, 0.6, 3, 0.2), (3, 0.1, 2, 0.9), (4, 0.4, 4, 0.6), (4, 0.9, 3, 0.5)]:
        ma = a[ca] + pa * (a[ca + 1] - a[ca])
        mb = b[cb] + pb * (b[cb + 1] - b[cb])
        links.append((0, ca, pa, 1, cb, pb, 8.0, 0.8 * float(np.linalg.norm(ma - mb))))
    motors = [(0, 2, 0.3, 1, 1, 0.4, 2.5, 120.0)]
    return network.from_filaments([a, b], 3000.0, linkers=links, motors=motors)


def extent(x, ends):
    """Sum of squared end-to-end distances of the filaments."""
    X = x.reshape(-1, 3)
    d = X[ends[:, 1]] - X[ends[:, 0]]
    return (d * d).sum()


def filament_ends(net):
    fil = net.cyl_filament
    ends = []
    for f in np.unique(fil):
        cs = np.nonzero(fil == f)[0]
        ends.append((net.cyl_bead1[cs[0]] // 3, net.cyl_bead2[cs[-1]] // 3))
    return np.array(ends)


def exact_ctx(net, coord=None):
    ctx = mechanics.Context(coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    ctx.set_network(net)
    ctx.build_neighbors(coord)
    return ctx


def dev(a):
    return torch.tensor(np.ascontiguousarray(a, np.float64), device="cuda")


@pytest.mark.parametrize("case", ["ladder", "net500", "branched", "branched_full"])
def test_hvp_matches_force_differences(oracle, lib, case):
    if case == "ladder":
        net = ladder()
    elif case == "net500":
        net = network.generate(lib, "net500")
    elif case == "branched":
        net = network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=2)
    else:   # with the branching dihedral and position terms
        net = network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=4,
                               k_br_dihedral=10.0, k_br_position=20.0)
    print(f"[gpu] hvp vs differences, {case}", flush=True)
    ctx = exact_ctx(net)
    s = oracle.system(net, coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    rng = np.random.default_rng(7)
    v = rng.normal(size=3 * net.n_beads)
    hv = autodiff.hvp(ctx, dev(net.coord), dev(v)).cpu().numpy()
    h = 1e-3
    fd = -(s.forces(net.coord + h * v) - s.forces(net.coord - h * v)) / (2 * h)
    ok = np.ones(net.n_beads, bool)
    ok[ill_conditioned_beads(net, s)] = False
    err = np.abs(hv - fd).reshape(-1, 3)[ok].max() / np.abs(fd).max()
    print(f"[gpu] {case}: max |Hv - CD| / max|CD| = {err:.2e} on {ok.sum()} of {net.n_beads} beads")
    assert np.all(np.isfinite(hv))
    assert err < 2e-6


def test_hvp_symmetric_linear_translation_free():
    net = ladder()
    ctx = exact_ctx(net)
    x = dev(net.coord)
    rng = np.random.default_rng(3)
    u, v = dev(rng.normal(size=x.numel())), dev(rng.normal(size=x.numel()))
    hu, hv = autodiff.hvp(ctx, x, u), autodiff.hvp(ctx, x, v)
    a, b = float(torch.dot(u, hv)), float(torch.dot(v, hu))
    assert abs(a - b) <= 1e-11 * (float(hu.norm() * v.norm()))
    comb = autodiff.hvp(ctx, x, 2.5 * u - 0.75 * v)
    assert float((comb - (2.5 * hu - 0.75 * hv)).abs().max()) <= 1e-12 * float(hu.abs().max() + hv.abs().max())
    t = dev(np.tile([0.3, -1.2, 0.7], net.n_beads))   # no boundary terms: translations are free
    ht = autodiff.hvp(ctx, x, t)
    assert float(ht.abs().max()) <= 1e-12 * float(hu.abs().max()) * float(t.norm()) / float(u.norm())


@pytest.mark.parametrize("case", ["ladder", "net500"])
def test_energy_jvp_matches_energy_differences(oracle, lib, case):
    net = ladder() if case == "ladder" else network.generate(lib, "net500")
    ctx = exact_ctx(net)
    s = oracle.system(net, coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    w = np.random.default_rng(5).normal(size=3 * net.n_beads)
    w.reshape(-1, 3)[ill_conditioned_beads(net, s)] = 0.0   # well-conditioned pairs only
    jv = autodiff.energy_jvp(ctx, dev(net.coord), dev(w)).cpu().numpy()
    h = 1e-4
    _, ep = s.energy(net.coord + h * w, individual=True)
    _, em = s.energy(net.coord - h * w, individual=True)
    cd = (ep - em) / (2 * h)
    scale = np.abs(cd).max()
    print(f"[gpu] {case}: jvp {jv[:NFF]} cd {cd}")
    np.testing.assert_allclose(jv[:NFF], cd, rtol=0, atol=1e-6 * scale)
    fw = -float(np.dot(s.forces(), w))
    assert abs(jv[NFF] - fw) <= 1e-11 * np.abs(s.forces() * w).sum()


def test_autograd_first_and_second_derivatives(exact_oracle):
    net = ladder()
    ctx = exact_ctx(net)
    x = dev(net.coord).requires_grad_(True)
    e = autodiff.energy(ctx, x)
    s = exact_oracle.system(net, coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    assert e.detach().item() == s.energy()[0]
    (g,) = torch.autograd.grad(e, x, create_graph=True)
    fo = s.forces()
    assert np.array_equal(g.detach().cpu().numpy(), -fo)       # -forces, bit for bit
    v = dev(np.random.default_rng(11).normal(size=x.numel()))
    (hv,) = torch.autograd.grad(g, x, grad_outputs=v)
    assert torch.equal(hv, autodiff.hvp(ctx, x.detach(), v))
    with pytest.raises(ValueError):
        autodiff.energy(mechanics.Context(), x)                 # coplanar perturbation: not a gradient


@pytest.fixture
def exact_oracle(oracle):
    oracle.set_exact(1)
    yield oracle
    oracle.set_exact(0)


def minimized(net, x0):
    ctx = exact_ctx(net, x0)
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0, grad_tol=1e-5)
    xs, _, res = ctx.minimize(p, x0)
    assert res.status == 0 and res.error_flags == 0
    x, fmax, steps = autodiff.newton_polish(ctx, dev(xs), ftol=1e-10)
    assert fmax < 1e-10
    return ctx, x


def test_implicit_stiffness_derivative_matches_reminimization():
    net = ladder()
    ends = filament_ends(net)
    ctx, xs = minimized(net, net.coord)
    x = xs.clone().requires_grad_(True)
    L = extent(x, ends)
    (gl,) = torch.autograd.grad(L, x)
    d, info = autodiff.stiffness_sensitivity(ctx, xs, gl, rtol=1e-12)
    d = d.cpu().numpy()
    print(f"[gpu] implicit dL/ds {d} {info}")
    assert info["relative_residual"] < 1e-12 and info["symmetry_modes"] == 6
    x0 = xs.cpu().numpy()
    for k in (0, 1, 2, 3):
        vals = []
        for sg in (1, -1):
            sc = np.ones(NFF)
            sc[k] += sg * 1e-4
            n2 = autodiff.scale_stiffness(net, sc)
            _, x2 = minimized(n2, x0)
            vals.append(float(extent(x2, ends)))
        fd = (vals[0] - vals[1]) / 2e-4
        print(f"[gpu] force field {k}: implicit {d[k]:.8g}  differences {fd:.8g}")
        assert abs(d[k] - fd) <= 1e-5 * np.abs(d).max()
    assert all(d[k] == 0 for k in range(4, NFF))   # no branch points, pairs or boundary terms


@pytest.mark.parametrize("side_stream", [False, True])
def test_device_calls_follow_torch_stream(lib, side_stream):
    """The device-pointer calls read coordinates that torch has just written:
    on torch's null stream (which the C ABI cannot name, so the calls run on
    the context stream between synchronizations) and on a side stream (passed
    through), each result equals the host-path evaluation at the same
    coordinates bit for bit."""
    net = network.generate(lib, "net500")
    ctx = exact_ctx(net)
    x0 = dev(net.coord)
    rng = np.random.default_rng(5)
    s = torch.cuda.Stream() if side_stream else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        for k in range(6):
            step = dev(rng.normal(0.0, 0.5, x0.numel()))
            xn = x0 + step * (k + 1)       # written by torch immediately before each call
            e = autodiff.energy(ctx, xn).item()
            f = (-autodiff.gradient(ctx, xn)).cpu().numpy()
            xh = xn.cpu().numpy()
            eh, _ = ctx.energy(xh)
            assert e == eh, (k, e, eh)
            assert np.array_equal(f, ctx.forces(xh)), k
    torch.cuda.synchronize()
    ctx.close()


def test_async_hvp_equals_waiting_hvp(lib):
    """mdy_hvp_dev_async enqueues the same kernels as mdy_hvp_dev without
    waiting: products bit for bit equal, no failure left to report."""
    net = network.generate(lib, "net500")
    ctx = mechanics.Context(0, coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    ctx.set_network(net)
    ctx.build_neighbors()
    x = torch.tensor(net.coord, device="cuda")
    g = torch.Generator("cuda").manual_seed(5)
    vs = [torch.randn(x.shape, dtype=torch.float64, device="cuda", generator=g) for _ in range(4)]
    want = []
    for v in vs:
        hv = torch.empty_like(v)
        ctx.hvp_dev(x.data_ptr(), v.data_ptr(), hv.data_ptr())
        want.append(hv.cpu().numpy())
    outs = [torch.empty_like(v) for v in vs]
    for v, hv in zip(vs, outs):
        ctx.hvp_dev_async(x.data_ptr(), v.data_ptr(), hv.data_ptr())
    ctx.sync_errors()
    for w, hv in zip(want, outs):
        assert np.array_equal(hv.cpu().numpy().view(np.int64), w.view(np.int64))
    ctx.close()


def test_async_hvp_failure_reported_by_sync_errors(lib):
    """Exactly parallel cylinders (the reference's 2filaments example) have
    no analytic second derivative: the waiting call raises MDY_ERR_FORCE at
    once, the asynchronous one leaves it for sync_errors, which clears it."""
    import os
    from medyan_amd import io
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    net = io.network_from_filament_file(os.path.join(golden, "filament_input_2filaments.txt"), 2000.0)
    ctx = mechanics.Context(0, coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    ctx.set_network(net)
    ctx.build_neighbors()
    x = torch.tensor(net.coord, device="cuda")
    v = torch.ones_like(x)
    hv = torch.empty_like(x)
    with pytest.raises(mechanics.MdyError) as e:
        ctx.hvp_dev(x.data_ptr(), v.data_ptr(), hv.data_ptr())
    assert e.value.code == _abi.MDY_ERR_FORCE
    ctx.hvp_dev_async(x.data_ptr(), v.data_ptr(), hv.data_ptr())
    with pytest.raises(mechanics.MdyError) as e:
        ctx.sync_errors()
    assert e.value.code == _abi.MDY_ERR_FORCE
    ctx.sync_errors()   # cleared
    ctx.close()
