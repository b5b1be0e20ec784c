"""The mixed-precision mode (ICLS_MODE_MIXED) against the FP64 GPU path and the oracle.

Mixed mode stores the field arrays (level set, velocities, momentum history,
materials, curvature, GFM jumps) as float and computes on them in float; the
pressures, the Thomas elimination, the Poisson residual with one refinement solve and
the volume sums stay double.  There is no bit-exact target: these tests bound the
error of every compared field against FP64 after N steps and print it.

Tolerances (relative L-infinity, max|a - b| / max|b|):
  solver2d on a float-representable right-hand side, against the oracle: 1e-9 (the
    refinement step takes the float-transform error, ~1e-6, down to ~1e-12);
  10 steps of the 32^3 drop and 5 steps of the 32x32x64 bubble against the FP64 GPU
    path: level set 1e-5, velocities and pressure 1e-3 (the velocity after a few
    steps is the small difference of large pressure and surface-tension terms);
  droplet volume after 100 steps of the 64^3 drop: |V_mixed - V_fp64| / V <= 1e-6.
"""
import numpy as np
import pytest

from helpers import rel_linf, small_drop
from iclsamd import params as P

pytestmark = pytest.mark.gpu


def _solver(p):
    from iclsamd import _lib
    return _lib.Solver(p)


def _mixed(p):
    q = P.IclsParams.from_buffer_copy(p)
    q.mode = P.MODE_MIXED
    return q


@pytest.mark.parametrize("n,nz", [(512, 8), (256, 16), (64, 32), (16, 8)])
def test_mixed_solver2d_refined_matches_oracle(n, nz):
    import pyoracle
    p = _mixed(small_drop(n, nz=nz))
    o = pyoracle.Oracle(p)
    g = _solver(p)
    try:
        o.init_program()
        g.init_program()
        rng = np.random.default_rng(7 * n + nz)
        f = np.zeros(P.field_dims(p, P.F_P), order="F")
        r = rng.standard_normal((n, n, nz))
        r -= r.mean()
        f[1:-1, 1:-1, 1:-1] = r.astype(np.float32).astype(np.float64)
        o.set(P.F_P, f)
        g.upload(P.F_P, f)
        o.solver2d()
        g.solver2d()
        a = g.download(P.F_P)[1:-1, 1:-1, 1:-1]
        b = o.get(P.F_P)[1:-1, 1:-1, 1:-1]
        e = rel_linf(a, b)
        print("mixed solver2d %dx%dx%d: rel Linf vs oracle %.3e" % (n, n, nz, e))
        assert e <= 1e-9
    finally:
        g.close()


@pytest.mark.parametrize("nx,ny,nz", [(512, 64, 8), (64, 512, 8), (256, 32, 8), (32, 256, 8)])
def test_mixed_solver2d_rectangular_matches_oracle(nx, ny, nz):
    """Lines of 256 or 512 along one axis and of another length along the other: the
    register transforms on one axis, the LDS transform on the other (and the y register
    kernels' fallback where nx is not a multiple of 64)."""
    import pyoracle
    p = _mixed(small_drop(max(nx, ny), nz=nz))
    p.itot, p.jtot = nx, ny
    o = pyoracle.Oracle(p)
    g = _solver(p)
    try:
        o.init_program()
        g.init_program()
        rng = np.random.default_rng(nx + 3 * ny + nz)
        f = np.zeros(P.field_dims(p, P.F_P), order="F")
        r = rng.standard_normal((nx, ny, nz))
        r -= r.mean()
        f[1:-1, 1:-1, 1:-1] = r.astype(np.float32).astype(np.float64)
        o.set(P.F_P, f)
        g.upload(P.F_P, f)
        o.solver2d()
        g.solver2d()
        a = g.download(P.F_P)[1:-1, 1:-1, 1:-1]
        b = o.get(P.F_P)[1:-1, 1:-1, 1:-1]
        e = rel_linf(a, b)
        print("mixed solver2d %dx%dx%d: rel Linf vs oracle %.3e" % (nx, ny, nz, e))
        assert e <= 1e-9
    finally:
        g.close()


FIELDS = [P.F_LVSET, P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW]
TOLS = {P.F_LVSET: 1e-5, P.F_UNEW: 1e-3, P.F_VNEW: 1e-3, P.F_WNEW: 1e-3, P.F_PNEW: 1e-3}


def _run(p, nsteps):
    """Fields and volume after nsteps.  The library holds one context at a time, so
    the two modes run one after the other."""
    g = _solver(p)
    try:
        g.init_program()
        for s in range(1, nsteps + 1):
            g.step(s, 0)
        g.check_nan()
        return {f: g.download(f) for f in FIELDS}, g.volume(0)[0]
    finally:
        g.close()


def _run_pair(p, nsteps):
    f64, v64 = _run(p, nsteps)
    f32, v32 = _run(_mixed(p), nsteps)
    errs = {P.FIELD_NAMES[f]: rel_linf(f32[f], f64[f]) for f in FIELDS}
    return errs, abs(v32 - v64) / abs(v64)


@pytest.mark.parametrize("scheme", [P.ADV_HOUC5, P.ADV_WENO5])
def test_mixed_steps_drop32(scheme):
    p = small_drop(32, ls_adv_space=scheme, mass_correction=0)
    errs, dv = _run_pair(p, 10)
    print("mixed vs fp64, drop32 x10 (scheme %d):" % scheme, {k: "%.2e" % v for k, v in errs.items()}, "dV %.2e" % dv)
    for f in FIELDS:
        assert errs[P.FIELD_NAMES[f]] <= TOLS[f], (P.FIELD_NAMES[f], errs)
    assert dv <= 1e-6


def test_mixed_steps_bubble():
    p = P.bubble256(32)
    p.ktot = 64
    p.timestep = 2e-4
    p.cfl = 0.0
    p.mass_correction = 0
    errs, dv = _run_pair(p, 5)
    print("mixed vs fp64, bubble 32x32x64 x5:", {k: "%.2e" % v for k, v in errs.items()}, "dV %.2e" % dv)
    for f in FIELDS:
        assert errs[P.FIELD_NAMES[f]] <= TOLS[f], (P.FIELD_NAMES[f], errs)
    assert dv <= 1e-6


@pytest.mark.slow
def test_mixed_volume_drop64_100_steps():
    """Droplet volume of the 64^3 drop after 100 steps, mixed against FP64."""
    errs, dv = _run_pair(P.drop64(), 100)
    print("mixed vs fp64, drop64 x100:", {k: "%.2e" % v for k, v in errs.items()}, "dV %.2e" % dv)
    assert dv <= 1e-6
