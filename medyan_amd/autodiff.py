"""Differentiable mechanics on the device.

torch.autograd functions over the energy, force, Hessian-vector-product and
energy-directional-derivative kernels of libmdy_hip.so, and the implicit
derivative of a minimized structure with respect to the stiffness constants
of each force field.

The reference has no differentiable interface; these functions expose its
force field (ForceFieldManager::computeEnergy / computeForces,
src/Mechanics/ForceField/ForceFieldManager.cpp:96-290) to torch. Derivatives
are those of the exact-gradient force field, so the context must use
MDY_COPLANAR_EXACT: the reference's coplanar excluded-volume force
(CylinderExclVolRepulsion.cpp:610-626) perturbs the geometry and is not the
gradient of the energy there. MDY_COPLANAR_QUADRATURE additionally
evaluates the excluded-volume pairs whose closed form is ill-conditioned
(nearly coplanar or parallel cylinders) by quadrature of the defining
integral, so their second derivatives are not rounding noise; that is the
mode for Hessians of large random networks. Coordinates are float64 CUDA tensors of
3 * n_beads entries, on the context's device; work is ordered on torch's
current stream.

Implicit differentiation: at a minimum x*(s) of E(x; s) = sum_k s_k E_k(x)
(every force field's energy is linear in its stiffness constant, so s_k
scales force field k), grad_x E(x*, s) = 0 gives
    H dx*/ds_k = -grad E_k(x*),
and for a quantity L(x*)
    dL/ds_k = -w . grad E_k(x*),   H w = grad L(x*),
with w from conjugate gradients on Hessian-vector products. H is positive
semi-definite at a minimum; rigid motions of free clusters are its null
space, so grad L must not move along them (L invariant under those motions,
e.g. distances within a filament).

Tethered minima: a structure stopped at a force tolerance need not be a
minimum (its Hessian can have negative curvature), and then dx*/ds does not
exist. Tethering every bead to reference coordinates x_ref by a spring of
stiffness t, E_t(x; s) = E(x; s) + t/2 |x - x_ref|^2, gives a minimum with
(H + t I) positive definite once t exceeds the most negative curvature, and
    dL/ds_k = -w . grad E_k(x_t*),   (H + t I) w = grad L(x_t*),
on the full space (the tether removes the rigid-motion symmetries).
"""
import math

import numpy as np
import torch

from . import _abi
from .mechanics import Context
from .network import Network

NFF = _abi.MDY_NFF

# stiffness arrays of each force field, in force-field order (FS, FB, LS, MS,
# BS, BB, BDH, BP, EV, BD); every energy is linear in these
STIFFNESS_FIELDS = ("fs_k", "fb_k", "ls_k", "ms_k", "bs_k", "bb_k", "bdh_k", "bp_k", "cyl_exvol_k", "plane_k")


def _ordered(call):
    """Runs call(stream) ordered with torch's current stream. A stream with a
    handle is passed through; torch's null stream cannot be named through the
    C ABI (NULL there means the context's own non-blocking stream), so the
    call then runs on the context stream between two device synchronizations:
    its inputs are complete before it starts and its outputs before torch
    reads them."""
    s = torch.cuda.current_stream().cuda_stream
    if s:
        call(s)
    else:
        torch.cuda.synchronize()
        call(0)
        torch.cuda.synchronize()


def _check(ctx: Context, x: torch.Tensor, name="x"):
    if ctx.opts.coplanar_mode not in (_abi.MDY_COPLANAR_EXACT, _abi.MDY_COPLANAR_QUADRATURE):
        raise ValueError("derivatives need a context with coplanar_mode MDY_COPLANAR_EXACT or "
                         "MDY_COPLANAR_QUADRATURE")
    n = 3 * ctx.net.n_beads
    if x.dtype != torch.float64 or not x.is_cuda or x.numel() != n or not x.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous float64 CUDA tensor of {n} entries")


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(fc, x, ctx):
        out = torch.empty(NFF + 2, dtype=torch.float64, device=x.device)
        _ordered(lambda s: ctx.energy_dev(x.data_ptr(), out.data_ptr(), s))
        fc.save_for_backward(x)
        fc.mech = ctx
        return out[NFF].clone()

    @staticmethod
    def backward(fc, g):
        (x,) = fc.saved_tensors
        return g * _Gradient.apply(x, fc.mech), None


class _Gradient(torch.autograd.Function):
    @staticmethod
    def forward(fc, x, ctx):
        f = torch.empty_like(x)
        _ordered(lambda s: ctx.forces_dev(x.data_ptr(), f.data_ptr(), accumulate=False, stream=s))
        fc.save_for_backward(x)
        fc.mech = ctx
        return f.neg_()

    @staticmethod
    def backward(fc, v):
        (x,) = fc.saved_tensors
        return hvp(fc.mech, x, v), None


def energy(ctx: Context, x: torch.Tensor) -> torch.Tensor:
    """computeEnergy total at x, differentiable twice (gradient = -forces,
    second derivative through Hessian-vector products)."""
    _check(ctx, x)
    return _Energy.apply(x, ctx)


def gradient(ctx: Context, x: torch.Tensor) -> torch.Tensor:
    """grad E(x) = -computeForces(x), differentiable (Hessian-vector products)."""
    _check(ctx, x)
    return _Gradient.apply(x, ctx)


def hvp(ctx: Context, x: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """H(x) v."""
    _check(ctx, x)
    v = v.detach().to(torch.float64).contiguous()
    _check(ctx, v, "v")
    hv = torch.empty_like(v)
    _ordered(lambda s: ctx.hvp_dev(x.data_ptr(), v.data_ptr(), hv.data_ptr(), s))
    return hv


def energy_jvp(ctx: Context, x: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """[grad E_k(x) . w for each force field k] + [their sum]."""
    _check(ctx, x)
    w = w.detach().to(torch.float64).contiguous()
    _check(ctx, w, "w")
    out = torch.empty(NFF + 1, dtype=torch.float64, device=x.device)
    _ordered(lambda s: ctx.energy_jvp_dev(x.data_ptr(), w.data_ptr(), out.data_ptr(), s))
    return out


class SymmetryModes:
    """Orthonormal basis of the exact symmetries of the energy with the
    context's current tables: the rigid motions of every cluster of beads
    joined by interactions (bonded terms and excluded-volume pairs), limited
    by the boundary planes the cluster touches (a translation must be
    parallel to every touched plane, a rotation axis normal to all of them).
    Every force field's energy is unchanged to first order along these
    directions, so its gradient is orthogonal to them; at a minimum H maps
    them to zero (translations at any coordinates). They are projected out
    of the Hessian solves."""

    def __init__(self, ctx: Context, x: torch.Tensor):
        import scipy.sparse as sp
        from scipy.sparse.csgraph import connected_components

        net = ctx.net
        nb = net.n_beads
        X = x.detach().cpu().numpy().reshape(nb, 3)
        groups = [(net.fs_beadset, 2), (net.fb_beadset, 3), (net.ls_beadset, 4), (net.ms_beadset, 4),
                  (net.bs_beadset, 3), (net.bb_beadset, 4), (ctx.exvol_table()[0].ravel(), 4)]
        src, dst = [], []
        for bs, w in groups:
            b = np.asarray(bs, np.int64).reshape(-1, w) // 3
            for k in range(1, w):
                src.append(b[:, 0]); dst.append(b[:, k])
        src = np.concatenate(src) if src else np.zeros(0, np.int64)
        dst = np.concatenate(dst) if dst else np.zeros(0, np.int64)
        g = sp.coo_matrix((np.ones(len(src)), (src, dst)), shape=(nb, nb))
        ncomp, comp = connected_components(g, directed=False)
        # boundary planes touched by each cluster
        nn, bbs = ctx.boundary_table()
        npl = len(nn)
        touched = np.zeros((ncomp, max(npl, 1)), bool)
        off = 0
        for p in range(npl):
            beads = bbs[off:off + nn[p]] // 3
            touched[comp[beads], p] = True
            off += nn[p]
        normals = np.asarray(net.plane, np.float64).reshape(-1, 4)[:, :3] if npl else np.zeros((0, 3))
        cnt = np.bincount(comp, minlength=ncomp).astype(np.float64)
        cen = np.stack([np.bincount(comp, X[:, d], ncomp) for d in range(3)], 1) / cnt[:, None]
        M = np.zeros((nb, 3, 6))
        keys = {}
        for c in range(ncomp):
            keys.setdefault(touched[c, :npl].tobytes(), []).append(c)
        for key, cs in keys.items():
            t = np.frombuffer(key, bool) if npl else np.zeros(0, bool)
            N = normals[t] if npl else np.zeros((0, 3))
            if len(N):
                _, sv, vt = np.linalg.svd(N)
                rank = int(np.sum(sv > 1e-12 * sv[0]))
                trans = vt[rank:]
                axes = vt[:1] if rank == 1 else np.zeros((0, 3))
            else:
                trans, axes = np.eye(3), np.eye(3)
            sel = np.isin(comp, cs)
            R = X[sel] - cen[comp[sel]]
            for j, tv in enumerate(trans):
                M[sel, :, j] = tv
            for j, a in enumerate(axes):
                M[sel, :, 3 + j] = np.cross(a, R)
        # per-cluster Gram-Schmidt, twice
        def segsum(v):
            return np.bincount(comp, v, ncomp)

        Q = np.zeros_like(M)
        for k in range(6):
            m = M[:, :, k].copy()
            ref = np.sqrt(segsum((m * m).sum(1)))
            for _ in range(2):
                for j in range(k):
                    d = segsum((Q[:, :, j] * m).sum(1))
                    m -= Q[:, :, j] * d[comp][:, None]
            nrm = np.sqrt(segsum((m * m).sum(1)))
            keep = (ref > 0) & (nrm > 1e-8 * np.where(ref > 0, ref, 1.0))
            scale = np.where(keep, 1.0 / np.where(nrm > 0, nrm, 1.0), 0.0)
            Q[:, :, k] = m * scale[comp][:, None]
        self.n_clusters = int(ncomp)
        self.n_modes = int(sum(((np.bincount(comp, (Q[:, :, k] ** 2).sum(1), ncomp) > 0.5).sum()) for k in range(6)))
        self._comp = torch.as_tensor(comp, device=x.device)
        self._q = torch.as_tensor(Q, device=x.device)
        self._ncomp = ncomp

    def project(self, v: torch.Tensor) -> torch.Tensor:
        """v minus its components along the symmetry directions."""
        V = v.reshape(-1, 3)
        per = (V[:, :, None] * self._q).sum(1)
        c = torch.zeros(self._ncomp, 6, dtype=v.dtype, device=v.device).index_add_(0, self._comp, per)
        return (V - (self._q * c[self._comp][:, None, :]).sum(2)).reshape(v.shape)


class NotAMinimum(RuntimeError):
    """Conjugate gradients on H met a direction of non-positive curvature:
    `direction` (a device tensor, unit length) with `curvature` = p.Hp / p.p."""

    def __init__(self, msg, direction, curvature):
        super().__init__(msg)
        self.direction = direction
        self.curvature = curvature


def solve_hessian(ctx: Context, x: torch.Tensor, b: torch.Tensor, rtol=1e-10, maxiter=20000,
                  modes: "SymmetryModes | None" = None, shift=0.0):
    """w with (H(x) + shift I) w = b by conjugate gradients on the
    complement of the symmetry directions `modes` (H is positive definite
    there at a well-determined minimum; b is projected first). shift > 0
    regularizes near-null mechanisms (Tikhonov). Returns (w, iterations,
    relative residual); raises NotAMinimum where H + shift I is not positive
    definite along a conjugate direction."""
    P = modes.project if modes is not None else (lambda v: v)
    b = P(b.detach().to(torch.float64).contiguous())
    w = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = float(torch.dot(r, r))
    bb = rr
    if bb == 0.0:
        return w, 0, 0.0
    it = 0
    while it < maxiter and rr > rtol * rtol * bb:
        hp = hvp(ctx, x, p)
        if shift:
            hp.add_(p, alpha=shift)
        hp = P(hp)
        php = float(torch.dot(p, hp))
        if not php > 0.0:
            pp = float(torch.dot(p, p))
            raise NotAMinimum(f"Hessian not positive along a conjugate direction (p.Hp = {php:g}, "
                              f"|p|^2 = {pp:g}, iteration {it}): not a minimum",
                              p / math.sqrt(pp), php / pp - shift)
        a = rr / php
        w.add_(p, alpha=a)
        r.add_(hp, alpha=-a)
        rr_new = float(torch.dot(r, r))
        p.mul_(rr_new / rr).add_(r)
        rr = rr_new
        it += 1
    return w, it, math.sqrt(rr / bb)


def stiffness_sensitivity(ctx: Context, x_star: torch.Tensor, grad_l: torch.Tensor, rtol=1e-10, maxiter=20000,
                          shift=0.0, tether=0.0):
    """dL/ds_k (k over the force fields) of a quantity L of the minimized
    coordinates x_star, given grad L(x_star); s_k scales force field k's
    stiffness constants (STIFFNESS_FIELDS). L must be invariant under the
    symmetry directions (rigid motions of free clusters). shift > 0 solves
    with H + shift I (a regularized derivative, for networks with floppy
    mechanisms). tether > 0: x_star is the minimum of the tethered energy
    (tethered_minimum with that stiffness) and the solve is (H + tether I)
    w = grad L on the full space. Returns (dL/ds [NFF], info)."""
    if tether > 0.0:
        w, it, res = solve_hessian(ctx, x_star, grad_l, rtol, maxiter, None, tether)
        d = -energy_jvp(ctx, x_star, w)[:NFF]
        return d, dict(cg_iterations=it, relative_residual=res, tether=tether)
    modes = SymmetryModes(ctx, x_star)
    w, it, res = solve_hessian(ctx, x_star, grad_l, rtol, maxiter, modes, shift)
    d = -energy_jvp(ctx, x_star, w)[:NFF]
    return d, dict(cg_iterations=it, relative_residual=res, clusters=modes.n_clusters, symmetry_modes=modes.n_modes)


def tethered_minimum(ctx: Context, x_ref: torch.Tensor, tether: float, x0: "torch.Tensor | None" = None,
                     ftol=1e-9, maxsteps=50, rtol=1e-12, maxiter=50000, maxstep=1.0):
    """Minimum of E(x) + tether/2 |x - x_ref|^2 by Newton steps
    (H + tether I) dx = F - tether (x - x_ref), F from the force kernel and
    the solves by conjugate gradients on Hessian-vector products. A step
    longer than `maxstep` nm on some bead is scaled down to it, and a step
    that raises the tethered energy is halved until it does not (at most 30
    times). A step is kept when the tethered energy does not rise beyond
    rounding or, where energy differences are below the energy's own rounding
    (close to the minimum), when the 2-norm of the tethered force falls.
    Starts from x0 (default x_ref); stops when the largest tethered force is
    below ftol. Returns (x, max|F_t|, info). Raises NotAMinimum if the tether
    does not make H + tether I positive definite."""
    _check(ctx, x_ref, "x_ref")
    if not tether > 0.0:
        raise ValueError("tether stiffness must be positive")
    x_ref = x_ref.detach()
    x = (x0 if x0 is not None else x_ref).detach().clone()
    _check(ctx, x)

    def tethered(z):
        d = z - x_ref
        return float(energy(ctx, z)) + 0.5 * tether * float(torch.dot(d, d))

    def force(z):
        return -gradient(ctx, z) - tether * (z - x_ref)

    e = tethered(x)
    f = force(x)
    fmax = float(f.abs().max())
    fnorm = float(torch.linalg.vector_norm(f))
    steps = cg_total = halvings = 0
    while fmax >= ftol and steps < maxsteps:
        dx, it, _ = solve_hessian(ctx, x, f, rtol=rtol, maxiter=maxiter, modes=None, shift=tether)
        cg_total += it
        big = float(dx.abs().max())
        if big > maxstep:
            dx.mul_(maxstep / big)
        for _ in range(30):
            xn = x + dx
            en = tethered(xn)
            fn = force(xn)
            fnn = float(torch.linalg.vector_norm(fn))
            if math.isfinite(en) and math.isfinite(fnn) and (en <= e + 1e-13 * abs(e) or fnn < fnorm):
                break
            dx.mul_(0.5)
            halvings += 1
        else:
            break
        x, e, f, fnorm = xn, en, fn, fnn
        fmax = float(f.abs().max())
        steps += 1
    return x, fmax, dict(steps=steps, cg_iterations=cg_total, halvings=halvings, tethered_energy=e)


def newton_polish(ctx: Context, x: torch.Tensor, ftol=1e-10, maxiter=30, modes: "SymmetryModes | None" = None):
    """Refines a minimum by Newton steps H dx = F (forces from the force
    kernel, solves on Hessian-vector products) until max|F| < ftol. The
    fixed point is F = 0 whatever the accuracy of H, which only sets the
    convergence rate. Returns (x, max|F|, steps)."""
    _check(ctx, x)
    x = x.detach().clone()
    modes = modes if modes is not None else SymmetryModes(ctx, x)
    f = -gradient(ctx, x)
    fmax = float(f.abs().max())
    steps = 0
    while fmax >= ftol and steps < maxiter:
        dx, _, _ = solve_hessian(ctx, x, f, rtol=1e-13, maxiter=20000, modes=modes)
        x.add_(dx)
        f = -gradient(ctx, x)
        fmax = float(f.abs().max())
        steps += 1
    return x, fmax, steps


def lm_polish(ctx: Context, x: torch.Tensor, ftol=1e-9, mu=0.0, maxsteps=60, rtol=1e-10, maxiter=20000,
              modes: "SymmetryModes | None" = None, maxstep=1.0, log=None):
    """Converges a near-minimum to a true local minimum by damped Newton
    (Levenberg-Marquardt) steps (H + mu I) dx = F, with F from the force
    kernel and the solves by conjugate gradients on Hessian-vector products
    outside the symmetry directions. A step is kept when no bead moves more
    than `maxstep` nm, the forces stay finite, and it lowers the energy or,
    at equal energy to rounding, the largest force; then mu shrinks (x 1/4),
    otherwise it grows (x 8). mu starts at least at max|F| / maxstep. Where
    the Hessian has negative curvature (a saddle of a stopped minimization)
    mu stays above it, so every kept step descends until the Hessian is
    positive semi-definite. Stops when max|F| < ftol. Returns (x, max|F|,
    info)."""
    _check(ctx, x)
    x = x.detach().clone()
    modes = modes if modes is not None else SymmetryModes(ctx, x)
    e = float(energy(ctx, x))
    f = modes.project(-gradient(ctx, x))
    fraw = float((-gradient(ctx, x)).abs().max())
    mu = max(mu, fraw / maxstep)
    steps = solves = cg_total = 0
    while fraw >= ftol and steps < maxsteps:
        steps += 1
        try:
            dx, it, res = solve_hessian(ctx, x, f, rtol=rtol, maxiter=maxiter, modes=modes, shift=mu)
            cg_total += it
            solves += 1
        except RuntimeError:
            mu *= 8.0
            continue
        if float(dx.abs().max()) > maxstep:
            mu *= 8.0
            continue
        xn = x + dx
        en = float(energy(ctx, xn))
        fn_raw = float((-gradient(ctx, xn)).abs().max())
        if math.isfinite(en) and math.isfinite(fn_raw) and (en < e or (en <= e + 1e-13 * abs(e) and fn_raw < fraw)):
            x, e, fraw = xn, en, fn_raw
            f = modes.project(-gradient(ctx, x))
            mu = max(mu * 0.25, 1e-12)
        else:
            mu *= 8.0
        if log:
            log(f"lm step {steps}: mu {mu:.2e}, E {e:.15e}, max|F| {fraw:.3e}, cg {it}")
    return x, fraw, dict(steps=steps, solves=solves, cg_iterations=cg_total, final_mu=mu)


def lowest_curvature(ctx: Context, x: torch.Tensor, steps=200, seed=0):
    """Estimate of the smallest eigenvalue of H(x) and its eigenvector, by
    `steps` Lanczos steps on Hessian-vector products (full
    re-orthogonalization, start vector seeded). A negative value shows that
    x is not a minimum; the returned unit vector is a direction of negative
    curvature that the energy alone can confirm (second differences).
    Returns (eigenvalue, vector)."""
    from scipy.linalg import eigh_tridiagonal

    _check(ctx, x)
    n = x.numel()
    g = torch.Generator(device=x.device).manual_seed(seed)
    q = torch.randn(n, dtype=torch.float64, device=x.device, generator=g)
    q /= torch.linalg.vector_norm(q)
    V = torch.empty(steps, n, dtype=torch.float64, device=x.device)
    alpha, beta = [], []
    m = steps
    for j in range(steps):
        V[j] = q
        w = hvp(ctx, x, q)
        a = float(torch.dot(w, q))
        alpha.append(a)
        for _ in range(2):
            w -= V[:j + 1].T @ (V[:j + 1] @ w)
        b = float(torch.linalg.vector_norm(w))
        if j == steps - 1 or b < 1e-12 * max(abs(a), 1.0):
            m = j + 1
            break
        beta.append(b)
        q = w / b
    lam, Y = eigh_tridiagonal(np.array(alpha[:m]), np.array(beta[:m - 1]), select="i", select_range=(0, 0))
    vec = V[:m].T @ torch.as_tensor(Y[:, 0], device=x.device)
    return float(lam[0]), vec / torch.linalg.vector_norm(vec)


def scale_stiffness(net: Network, s) -> Network:
    """A copy of net with force field k's stiffness constants times s[k]."""
    s = np.asarray(s, dtype=np.float64)
    if s.shape != (NFF,):
        raise ValueError(f"expected {NFF} scales")
    out = net.copy()
    for k, name in enumerate(STIFFNESS_FIELDS):
        setattr(out, name, getattr(out, name) * s[k])
    return out
