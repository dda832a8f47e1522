"""Independent numerical evaluation of the cylinder-cylinder excluded-volume
integral, for checking the library's quadrature mode (MDY_COPLANAR_QUADRATURE):

    E = kr * Int_0^1 Int_0^1 ds dt / |P(s) - Q(t)|^4,
    P(s) = c1 + s (c2 - c1),  Q(t) = c3 + t (c4 - c3),

the integral that CylinderExclVolRepulsion::energy evaluates in closed form
(kr = k * l1 * l2 with the cylinders' equilibrium lengths). A composite
Gauss-Legendre rule of a different order (32 points per panel) and with
twice the library's panel count per cylinder; forces and H v are the exact
derivatives of that rule. numpy only."""
from functools import lru_cache

import numpy as np

Q_MAX = 1e-3   # pairs with squared normalized triple product below this are integrated


@lru_cache(maxsize=None)
def _gl01(n):
    t, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (t + 1.0), 0.5 * w


def composite_rule(m, n=32):
    t, w = _gl01(n)
    return ((np.arange(m)[:, None] + t[None, :]) / m).ravel(), np.tile(w, m) / m


def triple_q(c1, c2, c3, c4):
    """(v1 x v2 . v3)^2 / (|v1|^2 |v2|^2 |v3|^2) per pair (arrays [..., 3])."""
    v1, v2, v3 = c2 - c1, c4 - c3, c3 - c1
    n = np.cross(v1, v2)
    tp = (n * v3).sum(-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        return tp * tp / ((v1 * v1).sum(-1) * (v2 * v2).sum(-1) * (v3 * v3).sum(-1))


def flagged(bs, x):
    """Pairs of an excluded-volume table (rows of 4 bead offsets x3) that the
    quadrature mode integrates numerically at coordinates x."""
    X = np.asarray(x).reshape(-1, 3)
    c = X[np.asarray(bs).reshape(-1, 4) // 3]
    q = triple_q(c[:, 0], c[:, 1], c[:, 2], c[:, 3])
    return ~(q >= Q_MAX)


def segment_distance(c1, c2, c3, c4, samples=257):
    """Distance between two segments by dense sampling plus local refinement
    (a lower bound within the sampling step is all the panel count needs)."""
    t = np.linspace(0.0, 1.0, samples)
    P = c1[None, :] + t[:, None] * (c2 - c1)[None, :]
    Q = c3[None, :] + t[:, None] * (c4 - c3)[None, :]
    d = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1))
    step = max(np.linalg.norm(c2 - c1), np.linalg.norm(c4 - c3)) / (samples - 1)
    return max(d.min() - step, 1e-3 * step)


def pair(c1, c2, c3, c4, kr, v=None, n=32):
    """(E, F[4, 3]) of one pair, or with v ([4, 3], the direction at the four
    beads) also H v [4, 3]."""
    c1, c2, c3, c4 = (np.asarray(c, np.float64) for c in (c1, c2, c3, c4))
    L = max(np.linalg.norm(c2 - c1), np.linalg.norm(c4 - c3))
    m = int(min(128, max(1, 2 * np.ceil(L / segment_distance(c1, c2, c3, c4)))))
    s, w = composite_rule(m, n)
    v1, v2, c13 = c2 - c1, c4 - c3, c1 - c3
    r = c13[None, None, :] + s[:, None, None] * v1[None, None, :] - s[None, :, None] * v2[None, None, :]
    d2 = (r * r).sum(-1)
    W = w[:, None] * w[None, :]
    E = kr * float((W / d2 ** 2).sum())
    g = (4.0 * kr * W / d2 ** 3)[..., None] * r            # -grad_r of kr W |r|^-4
    S, T = s[:, None, None], s[None, :, None]
    F = np.stack([(g * (1 - S)).sum((0, 1)), (g * S).sum((0, 1)),
                  -(g * (1 - T)).sum((0, 1)), -(g * T).sum((0, 1))])
    if v is None:
        return E, F
    v = np.asarray(v, np.float64)
    dr = (v[0] - v[2])[None, None, :] + S * (v[1] - v[0])[None, None, :] - T * (v[3] - v[2])[None, None, :]
    rd = (r * dr).sum(-1)[..., None]
    h = kr * W[..., None] * (-4.0 * dr / d2[..., None] ** 3 + 24.0 * rd * r / d2[..., None] ** 4)
    Hv = np.stack([(h * (1 - S)).sum((0, 1)), (h * S).sum((0, 1)),
                   -(h * (1 - T)).sum((0, 1)), -(h * T).sum((0, 1))])
    return E, F, Hv
