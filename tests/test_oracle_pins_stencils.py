"""Pins of the oracle's level-set stencils against what the mathematics fixes (CPU only).

Each check is chosen so that a plausible slip in the oracle (a swapped coefficient,
a wrong branch, a wrong index or ghost layer) fails it:

* tube_construction (levelset.f90:87-197): its saturated-sum test is equivalent to
  "some point of the 13-point cross stencil has |phi| < NarrowBand_2", because every
  clamped term |phi| - NB2 is <= 0 and a sum of non-positive doubles is zero only if
  every term is; checked with stale z ghost layers, which the wall rows read.
* HOUC5_nb (hyperPDE.f90:397-495): the 6-point upwind stencil is the unique one exact
  on quintics, the 4-point wall stencil (HOUC3) the unique one exact on cubics.
* WENO5_nb (hyperPDE.f90:503-800): rebuilt from its published definition (Jiang and
  Shu 1996; Jiang and Peng 2000) -- cell-average quadratic reconstructions of the
  one-sided differences, smoothness indicators as integrals of the squared first and
  second derivatives, linear weights from the 5-cell reconstruction -- with numpy
  linear algebra, not from the reference's coefficients.
* WENO5_reinit (hyperPDE.f90:2646-2746): on piecewise-linear profiles with unequal
  slopes the upwind selection has closed forms, one per branch; the smoothed sign
  phi / sqrt(phi^2 + dz^2) inside |phi| <= dz (levelset.f90:660-669).
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import fortran_coords, small_drop
from iclsamd import params as P


# ---------------------------------------------------------------- tube_construction
def _periodic_xy(a):
    """Fill the three-wide x/y halos of an LS-shaped array with periodic images."""
    n0, n1 = a.shape[0] - 6, a.shape[1] - 6
    ix = (np.arange(a.shape[0]) - 3) % n0 + 3
    iy = (np.arange(a.shape[1]) - 3) % n1 + 3
    return a[ix][:, iy]


def test_tube_mask_is_the_cross_stencil_rule_with_stale_ghosts():
    p = small_drop(16, nz=12)
    o = O.Oracle(p)
    o.init_program()
    nb2 = 6.0 * p.lz / p.ktot
    rng = np.random.default_rng(11)
    dims = P.field_dims(p, P.F_LVSET)
    sgn = np.where(rng.random(dims) < 0.5, -1.0, 1.0)
    phi = sgn * nb2 * (1.0 + rng.random(dims))                 # saturated everywhere ...
    near = rng.random(dims) < 0.012
    phi[near] = rng.uniform(-nb2, nb2, size=near.sum())       # ... except scattered band points
    phi = _periodic_xy(phi)
    # stale z ghost layers (k = -2..0 and k1..k1+2), unrelated to the interior
    for kk in (0, 1, 2, dims[2] - 3, dims[2] - 2, dims[2] - 1):
        g = nb2 * (1.0 + rng.random(dims[:2])) * np.where(rng.random(dims[:2]) < 0.5, -1.0, 1.0)
        hit = rng.random(dims[:2]) < 0.05
        g[hit] = rng.uniform(-nb2, nb2, size=hit.sum())
        phi[:, :, kk] = g
    phi = _periodic_xy(phi)
    # one column whose only band points are a ghost below k = 1 and one above k = kmax
    for d in (-2, -1, 1, 2):
        phi[3 + 5 + d, 3 + 7, :] = nb2 * 1.5
        phi[3 + 5, 3 + 7 + d, :] = nb2 * 1.5
    phi[3 + 5, 3 + 7, :] = nb2 * 1.5
    phi[3 + 5, 3 + 7, 1] = 0.25 * nb2            # k = -1: seen by k = 1 only
    phi[3 + 5, 3 + 7, -2] = -0.5 * nb2           # k = k1 + 1: seen by k = kmax only
    phi = _periodic_xy(phi)
    o.set(P.F_LVSET, phi)
    o.tube_construction(0)
    mask = o.band_mask()

    inb = np.abs(phi) < nb2
    n, nz = p.itot, p.ktot
    want = np.zeros((n, n, nz), dtype=bool)
    I, J, K = np.meshgrid(np.arange(n) + 3, np.arange(n) + 3, np.arange(nz) + 3, indexing="ij")
    for d in (-2, -1, 0, 1, 2):
        want |= inb[I + d, J, K] | inb[I, J + d, K] | inb[I, J, K + d]
    assert np.array_equal(mask.astype(bool), want)
    assert mask[5, 7, 0] == 1 and mask[5, 7, 1] == 0 and mask[5, 7, -1] == 1 and mask[5, 7, -2] == 0
    # the rows next to the walls really depend on the ghosts
    want_no_ghost = want.copy()
    inb2 = inb.copy()
    inb2[:, :, :3] = False
    inb2[:, :, -3:] = False
    want_no_ghost[:] = False
    for d in (-2, -1, 0, 1, 2):
        want_no_ghost |= inb2[I + d, J, K] | inb2[I, J + d, K] | inb2[I, J, K + d]
    assert (want != want_no_ghost).sum() > 10
    # phi leaves clamped to +-NB2, with the line-extrapolated wall ghost (boundsls)
    out = o.get(P.F_LVSET)
    clamped = np.clip(phi, -nb2, nb2)
    assert np.array_equal(out[3:-3, 3:-3, 3:-3], clamped[3:-3, 3:-3, 3:-3])
    assert np.allclose(out[3:-3, 3:-3, 2], 2 * clamped[3:-3, 3:-3, 3] - clamped[3:-3, 3:-3, 4], rtol=0, atol=1e-15)
    assert np.allclose(out[3:-3, 3:-3, -3], 2 * clamped[3:-3, 3:-3, -4] - clamped[3:-3, 3:-3, -5], rtol=0, atol=1e-15)


# ---------------------------------------------------------------- HOUC5 / HOUC3 / first order
def _adv_setup(n=16, nz=16, lz=1.0, scheme=P.ADV_HOUC5, seed=5):
    p = small_drop(n, nz=nz)
    p.lz = lz
    p.ls_adv_space = scheme
    o = O.Oracle(p)
    o.init_program()
    rng = np.random.default_rng(seed)
    u, v, w = (rng.standard_normal((n, n, nz)) for _ in range(3))
    o.set(P.F_UCC, u)
    o.set(P.F_VCC, v)
    o.set(P.F_WCC, w)
    return p, o, u, v, w


def _eval_on(o, phi, scheme):
    """Advection RHS of phi over the whole interior: the tube is built on phi (all
    cells lie within the band) and phi is then restored with its non-periodic halos."""
    o.set(P.F_LVSET, phi)
    o.tube_construction(0)
    assert o.band_count() == o.p.itot * o.p.jtot * o.p.ktot
    o.set(P.F_LVSET, phi)
    return o.eval_adv(0, scheme)


def _poly(rng, deg, t):
    c = rng.standard_normal(deg + 1)
    t0 = t - 0.5
    val = sum(c[d] * t0 ** d for d in range(deg + 1))
    der = sum(d * c[d] * t0 ** (d - 1) for d in range(1, deg + 1))
    return val, der


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_houc5_exact_on_quintics(axis):
    p, o, u, v, w = _adv_setup()
    X = fortran_coords(p, P.F_LVSET)
    rng = np.random.default_rng(40 + axis)
    amp = 0.01
    val, der = _poly(rng, 5, X[axis])
    dfdt = _eval_on(o, amp * val, P.ADV_HOUC5)
    vel = (u, v, w)[axis]
    ex = -vel * amp * der[3:-3, 3:-3, 3:-3]
    ks = slice(None) if axis < 2 else slice(2, -2)   # z: HOUC5 on k = 3 .. kmax-2
    err = np.max(np.abs(dfdt[:, :, ks] - ex[:, :, ks]))
    assert err <= 1e-10 * np.max(np.abs(ex))
    if axis == 2:
        # the wall rows are not quintic-exact: they are the lower-order stencils
        assert np.max(np.abs(dfdt[:, :, 1] - ex[:, :, 1])) > 1e-6 * np.max(np.abs(ex))


@pytest.mark.parametrize("scheme", [P.ADV_HOUC5, P.ADV_WENO5])
def test_wall_rows_houc3_cubic_and_first_order_linear(scheme):
    """k = 2 and kmax-1 use HOUC3 (exact on cubics, and only on them); k = 1 and kmax use
    first-order upwinding (exact on linear data), in both schemes (hyperPDE.f90:459-480,
    763-784)."""
    p, o, u, v, w = _adv_setup(scheme=scheme, lz=2.0, nz=16)
    X = fortran_coords(p, P.F_LVSET)
    rng = np.random.default_rng(77)
    amp = 0.01
    zs = X[2] / p.lz                      # keep phi inside the band on the 1 x 1 x 2 box
    val, der = _poly(rng, 3, zs)
    dfdt = _eval_on(o, amp * val, scheme)
    ex = -w * amp * der[3:-3, 3:-3, 3:-3] / p.lz
    for k in (1, p.ktot - 2):
        assert np.max(np.abs(dfdt[:, :, k] - ex[:, :, k])) <= 1e-10 * np.max(np.abs(ex))
    val, der = _poly(rng, 4, zs)
    dfdt = _eval_on(o, amp * val, scheme)
    ex = -w * amp * der[3:-3, 3:-3, 3:-3] / p.lz
    assert np.max(np.abs(dfdt[:, :, 1] - ex[:, :, 1])) > 1e-6 * np.max(np.abs(ex))
    val, der = _poly(rng, 1, zs)
    dfdt = _eval_on(o, amp * val, scheme)
    ex = -w * amp * der[3:-3, 3:-3, 3:-3] / p.lz
    for k in (0, p.ktot - 1):
        assert np.max(np.abs(dfdt[:, :, k] - ex[:, :, k])) <= 1e-10 * np.max(np.abs(ex))


# ---------------------------------------------------------------- WENO5 from its definition
def _cell_average_quadratic(cells, avgs):
    """Coefficients (a, b, c) of a + b x + c x^2 whose averages over the unit cells
    centred at ``cells`` are ``avgs``."""
    M = np.array([[1.0, m, m * m + 1.0 / 12.0] for m in cells])
    return np.linalg.solve(M, avgs)


def _cell_average_quartic(cells, avgs):
    M = np.array([[((m + 0.5) ** (d + 1) - (m - 0.5) ** (d + 1)) / (d + 1) for d in range(5)] for m in cells])
    return np.linalg.solve(M, avgs)


def _weno_definition():
    """Per sub-stencil the value at the right cell edge x = 1/2 as a linear functional
    of the five averages, the smoothness indicator as a quadratic form, and the linear
    weights, all derived numerically from the definitions."""
    subs = [(-2, -1, 0), (-1, 0, 1), (0, 1, 2)]
    E = np.eye(5)
    val = np.zeros((3, 5))
    beta = np.zeros((3, 5, 5))
    xs, wq = np.polynomial.legendre.leggauss(6)
    xs = 0.5 * xs
    wq = 0.5 * wq
    for s, cells in enumerate(subs):
        for q in range(5):
            a, b, c = _cell_average_quadratic(cells, E[q, [m + 2 for m in cells]])
            val[s, q] = a + b * 0.5 + c * 0.25
        # beta = int_{-1/2}^{1/2} p'(x)^2 + p''(x)^2 dx, a quadratic form in the averages
        G = np.zeros((5, len(xs) + 1))
        for q in range(5):
            a, b, c = _cell_average_quadratic(cells, E[q, [m + 2 for m in cells]])
            G[q, :-1] = (b + 2 * c * xs) * np.sqrt(wq)
            G[q, -1] = 2 * c
        beta[s] = G @ G.T
    full = np.zeros(5)
    for q in range(5):
        co = _cell_average_quartic(range(-2, 3), E[q])
        full[q] = sum(co[d] * 0.5 ** d for d in range(5))
    d, *_ = np.linalg.lstsq(val.T, full, rcond=None)
    assert np.allclose(val.T @ d, full, atol=1e-13)
    return val, beta, d


def _weno_derivative(vs, eps=1e-6):
    """WENO derivative from five one-sided differences ``vs`` (shape (..., 5))."""
    val, beta, d = _weno_definition()
    q = np.einsum("sq,...q->...s", val, vs)
    b = np.einsum("...q,sqr,...r->...s", vs, beta, vs)
    alpha = d / (eps + b) ** 2
    w = alpha / alpha.sum(axis=-1, keepdims=True)
    return (w * q).sum(axis=-1)


def test_weno_definition_gives_the_published_constants():
    """Sanity of the derivation itself: linear weights 1/10, 6/10, 3/10 and the
    reconstruction of sub-stencil 1 = (2 v1 - 7 v2 + 11 v3) / 6."""
    val, beta, d = _weno_definition()
    assert np.allclose(d, [0.1, 0.6, 0.3], atol=1e-14)
    assert np.allclose(val[0], [1 / 3, -7 / 6, 11 / 6, 0, 0], atol=1e-14)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_weno5_matches_its_definition(axis):
    n, nz = 16, 16
    p, o, u, v, w = _adv_setup(n=n, nz=nz, scheme=P.ADV_WENO5, seed=9 + axis)
    dims = P.field_dims(p, P.F_LVSET)
    h = (p.lx / n, p.ly / n, p.lz / nz)[axis]
    rng = np.random.default_rng(100 + axis)
    # a profile along one axis mixing smooth stretches, kinks and noise, inside the band
    m = dims[axis]
    t = np.arange(m) * h
    prof = 0.05 * np.sin(7 * t) + 0.03 * np.abs(t - 0.41) + 0.004 * rng.standard_normal(m)
    shape = [1, 1, 1]
    shape[axis] = m
    phi = np.broadcast_to(prof.reshape(shape), dims).copy()
    dfdt = _eval_on(o, phi, P.ADV_WENO5)
    vel = (u, v, w)[axis]
    got = dfdt
    idx = np.arange(m - 6) + 3                       # interior LS indices
    dm = (prof[1:] - prof[:-1]) / h                  # dm[c] = (prof[c+1] - prof[c]) / h
    vm = np.stack([dm[idx + s - 1] for s in (-2, -1, 0, 1, 2)], axis=-1)      # v1..v5, upwind from the left
    vp = np.stack([dm[idx - s] for s in (-2, -1, 0, 1, 2)], axis=-1)          # mirrored, from the right
    dminus = _weno_derivative(vm)
    dplus = _weno_derivative(vp)
    sel = (slice(None),) * 3
    if axis == 2:
        sel = (slice(None), slice(None), slice(2, -2))   # WENO5 in z on k = 3 .. kmax-2
    der = np.where(vel >= 0, np.broadcast_to(dminus.reshape([-1 if a == axis else 1 for a in range(3)]), vel.shape),
                   np.broadcast_to(dplus.reshape([-1 if a == axis else 1 for a in range(3)]), vel.shape))
    ex = -vel * der
    assert np.max(np.abs(got[sel] - ex[sel])) <= 1e-11 * np.max(np.abs(ex[sel]))
    # the weights really are nonlinear here: the linear-weight stencil differs
    lin = np.einsum("sq,...q->...s", _weno_definition()[0], vm) @ np.array([0.1, 0.6, 0.3])
    assert np.max(np.abs(lin - dminus)) > 1e-3 * np.max(np.abs(dminus))


def test_weno5_picks_the_smooth_side_of_a_kink():
    """ENO property: at a kink the stencils that cross it get negligible weight, so the
    one-sided derivative equals the slope of its own side."""
    p, o, u, v, w = _adv_setup(scheme=P.ADV_WENO5)
    X = fortran_coords(p, P.F_LVSET)
    x0 = (8 - 0.5) / 16.0                              # the centre of cell i = 8
    phi = np.where(X[0] < x0, -0.3 * (X[0] - x0), 0.45 * (X[0] - x0)) - 0.02
    dfdt = _eval_on(o, phi, P.ADV_WENO5)
    # cell 9 (index 8) upwinded from the left: its left stencil (cells 8..11) is on the right branch
    i = 8
    ex = -u[i] * 0.45
    sel = u[i] >= 0
    assert np.max(np.abs(dfdt[i][sel] - ex[sel])) <= 1e-9
    # cell 7 (index 6) upwinded from the right uses cells 4..7: the left branch
    i = 6
    sel = u[i] < 0
    assert np.max(np.abs(dfdt[i][sel] - (-u[i][sel] * -0.3))) <= 1e-9


# ---------------------------------------------------------------- WENO5 re-initialisation RHS
def _reinit_case(kind, a, b, c, axis=0):
    """phi along one axis: a V (slopes -b left, +a right, minimum -c) or a Lambda
    (slopes +b left, -a right, maximum c), kink at the centre of cell 8."""
    n = 16
    p = small_drop(n)
    if axis == 2:
        p.lz = 2.0
    o = O.Oracle(p)
    o.init_program()
    X = fortran_coords(p, P.F_LVSET)
    h = (p.lx / n, p.ly / n, p.lz / n)[axis]
    x0 = 7.5 * h
    t = X[axis] - x0
    if kind == "V":
        phi = np.where(t < 0, -b * t, a * t) - c
    else:
        phi = np.where(t < 0, b * t, -a * t) + c
    o.set(P.F_LVSET, phi)
    rhs = o.eval_reinit_rhs(0)[3:-3, 3:-3, 3:-3]
    rhs = np.moveaxis(rhs, axis, 0)[:, 5, 5]
    phic = np.moveaxis(phi[3:-3, 3:-3, 3:-3], axis, 0)[:, 5, 5]
    return p, rhs, phic, p.lz / p.ktot


def _sign(phi, dz):
    return np.where(phi < -dz, -1.0, np.where(phi > dz, 1.0, phi / np.sqrt(phi * phi + dz * dz)))


@pytest.mark.parametrize("axis", [0, 2])
@pytest.mark.parametrize("kind,a,b,c,branch", [
    ("V", 1.2, 0.8, 0.1, "plus"),        # s < 0 at the kink, right slope steeper: B
    ("V", 0.8, 1.2, 0.1, "minus"),       # s < 0, left slope steeper: A
    ("V", 1.2, 0.8, -0.1, "mean"),       # s > 0 at a minimum: the mean of A and B
    ("L", 1.2, 0.8, 0.1, "plus"),        # s > 0 at a maximum, right slope steeper: B
    ("L", 0.8, 1.2, 0.1, "minus"),
    ("L", 1.2, 0.8, -0.1, "mean"),       # s < 0 at a maximum
])
def test_reinit_rhs_upwind_branches_closed_form(axis, kind, a, b, c, branch):
    p, rhs, phi, dz = _reinit_case(kind, a, b, c, axis)
    nb2 = 6 * dz
    s = _sign(phi, dz)
    m = 7                                  # the kink cell (cell 8)
    if axis == 2:
        assert 3 <= m + 1 <= p.ktot - 2
    left, right = (-b, a) if kind == "V" else (b, -a)
    want = s * (1.0 - np.where(np.arange(p.itot) < m, abs(left), abs(right)))
    k_sel = {"plus": right, "minus": left, "mean": 0.5 * (left + right)}[branch]
    want[m] = s[m] * (1.0 - abs(k_sel))
    want = np.where(np.abs(phi) < nb2 + dz, want, 0.0)
    assert np.max(np.abs(rhs - want)) <= 1e-11


def test_reinit_smoothed_sign_inside_one_cell():
    """Where |phi0| <= dz the sign is phi0 / sqrt(phi0^2 + dz^2); a slope-2 profile has
    RHS = S(phi0) (1 - 2) there."""
    n = 16
    p = small_drop(n)
    o = O.Oracle(p)
    o.init_program()
    X = fortran_coords(p, P.F_LVSET)
    dz = p.lz / p.ktot
    phi = 2.0 * (X[2] - 0.5 - 0.3 * dz)
    o.set(P.F_LVSET, phi)
    rhs = o.eval_reinit_rhs(0)[3:-3, 3:-3, 3:-3][4, 4, :]
    ph = phi[3:-3, 3:-3, 3:-3][4, 4, :]
    inside = (np.abs(ph) <= dz)
    inside[:2] = inside[-2:] = False
    assert inside.sum() >= 1
    s = ph / np.sqrt(ph * ph + dz * dz)
    assert np.max(np.abs(rhs[inside] - s[inside] * (1 - 2.0))) <= 1e-12
    assert np.all(np.abs(s[inside]) < 1 / np.sqrt(2) + 1e-15)
