"""The workload configurations on the GPU: the branches and configurations the other
parity tests do not reach, against the CPU oracle where it finishes in seconds, and
through properties the physics fixes at the full sizes.

* Couette walls (bound.f90:29-35, ns_np = 0) and the 'cou' start (init.f90:191-210)
  with two drops, FP64, every field including the z ghost layers at 1e-10.
* chkdiv (chkdiv.f90:10-90): divtot, divmax and the MAXLOC cell on the same velocity
  fields in both precisions.
* A reduced many-drop case (128 x 128 x 64, 12 seeded drops of radius 3 .. 20 cells,
  per-drop windows) against the oracle's whole-grid level sets, FP64 at 1e-10.
* The mixed-precision step against the oracle's FP64 step on the drop geometry (64^3)
  and the rising bubble (32 x 32 x 64): the L-infinity and L2 errors of phi, u, v, w
  and p are printed, and bounded by the stated tolerances.
* Full-size property checks: the 512^3 mixed drop, the 256^3 64-drop learned case and
  the 512 x 512 x 256 Couette emulsion with 200 drops, a few steps each.
"""
import numpy as np
import pytest

from helpers import rel_linf, small_drop
from iclsamd import params as P

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS = [P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW, P.F_POLD, P.F_MIU, P.F_RHO, P.F_DUOLD, P.F_DVOLD,
          P.F_DWOLD, P.F_PX, P.F_PY, P.F_PZ, P.F_UCC, P.F_VCC, P.F_WCC]


def _pair(p):
    import pyoracle
    from iclsamd import _lib
    return pyoracle.Oracle(p), _lib.Solver(p)


def _compare(o, g, fields, tol, tag):
    bad = {}
    for f in fields:
        a, b = g.download(f), o.get(f)
        e = rel_linf(a, b) if np.max(np.abs(b)) > 0 else float(np.max(np.abs(a)))
        if not e <= tol:
            bad[P.FIELD_NAMES[f]] = e
    assert not bad, "%s: fields beyond %.1e: %s" % (tag, tol, bad)


def _clip_ls(a, p):
    nb2 = 6.0 * p.lz / p.ktot
    return np.clip(a, -nb2, nb2)


# ---------------------------------------------------------------- Couette walls
def test_couette_two_drops_parity():
    p = P.base_params(32, nz=16)
    p.lz = 0.5
    p.ns_np = 0
    p.iniu = P.INIU_COU
    p.wall_speed = 1.0
    p.mass_correction = 0
    p.set_drops([(0.3, 0.4, 0.25, 0.1), (0.7, 0.6, 0.22, 0.08)])
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        _compare(o, g, [P.F_VNEW], 1e-15, "Couette start")
        for s in range(1, 6):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, FIELDS, TOL, "Couette, 5 steps")
        for l in range(2):
            assert rel_linf(g.download(P.F_LVSET, l), o.get(P.F_LVSET, l)) <= TOL, l
        v = g.download(P.F_VNEW)
        # the wall values (mean of the two layers about each wall) stay at -1 and +1
        assert np.max(np.abs(0.5 * (v[:, :, 0] + v[:, :, 1]) + 1.0)) <= 1e-12
        assert np.max(np.abs(0.5 * (v[:, :, -1] + v[:, :, -2]) - 1.0)) <= 1e-12
    finally:
        g.close()


# ---------------------------------------------------------------- chkdiv
@pytest.mark.parametrize("mode", [P.MODE_FP64, P.MODE_MIXED])
def test_chkdiv_matches_oracle(mode):
    """The same (float-representable in the mixed mode) velocity fields on both sides:
    the divergence of every cell is formed in the same order in double, so divmax and
    its cell agree exactly and divtot to the rounding of its sum."""
    p = small_drop(32, nz=24)
    p.mode = mode
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        rng = np.random.default_rng(17 + mode)
        shape = P.field_dims(p, P.F_UNEW)
        for f in (P.F_UNEW, P.F_VNEW, P.F_WNEW):
            a = rng.standard_normal(shape).astype(np.float32).astype(np.float64)
            a[0], a[-1] = a[-2], a[1]                    # periodic x/y images
            a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
            o.set(f, a)
            g.upload(f, a)
        rc_o, tot_o, max_o, ijk_o = o.chkdiv(5)
        rc_g, tot_g, max_g, ijk_g = g.chkdiv(5)
        assert rc_o == 1 and rc_g == 4              # both report the divergence
        assert max_g == max_o and ijk_g == ijk_o, (max_g, max_o, ijk_g, ijk_o)
        n = p.itot * p.jtot * p.ktot
        assert abs(tot_g - tot_o) <= 4 * np.sqrt(n) * np.finfo(float).eps * n * max_o
        # and after a real step the divergence-free velocity passes in both
        g.init_program()
        g.step(1, 0)
        rc, tot, dmax, ijk = g.chkdiv(1)
        assert rc == 0 and dmax < (1e-12 if mode == P.MODE_FP64 else 1e-3)
    finally:
        g.close()


# ---------------------------------------------------------------- many drops, reduced
def test_many_drops_reduced_parity():
    """A reduced version of the 64-drop configuration: 128 x 128 x 64 over 1 x 1 x 0.5,
    12 drops of radius 3 .. 20 cells from the seeded sampler, ratios 1:1, sigma = 1,
    each level set stepped in its own window (refreshed every 4 steps); 5 steps."""
    p = P.base_params(128, nz=64)
    p.lz = 0.5
    p.rho1 = p.rho2 = 1.0
    p.miu1 = p.miu2 = 1.0
    p.timestep = 1e-5
    p.mass_correction = 0
    p.set_drops(P.random_drops(12, 128, 3.0, 20.0, seed=20261202, lz=0.5))
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 6):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, FIELDS, TOL, "12 drops, 5 steps")
        for l in range(p.lmax):
            e = rel_linf(_clip_ls(g.download(P.F_LVSET, l), p), _clip_ls(o.get(P.F_LVSET, l), p))
            assert e <= TOL, (l, e)
            # V = 1 - sum(H) / N, summed in another order: about sqrt(N) eps apart
            assert abs(g.volume(l)[0] - o.volume(l)[0]) <= 4 * np.sqrt(p.itot * p.jtot * p.ktot) * 2.2e-16, l
    finally:
        g.close()


# ---------------------------------------------------------------- mixed precision vs the FP64 oracle
def _errors(o, g, p):
    out = {}
    for name, f, sl in (("phi", P.F_LVSET, np.s_[3:-3, 3:-3, 3:-3]), ("u", P.F_UNEW, np.s_[1:-1, 1:-1, 1:-1]),
                        ("v", P.F_VNEW, np.s_[1:-1, 1:-1, 1:-1]), ("w", P.F_WNEW, np.s_[1:-1, 1:-1, 1:-1]),
                        ("p", P.F_PNEW, np.s_[1:-1, 1:-1, 1:-1])):
        a, b = g.download(f)[sl], o.get(f)[sl]
        if f == P.F_LVSET:
            a, b = _clip_ls(a, p), _clip_ls(b, p)
        scale = max(float(np.max(np.abs(b))), 1e-300)
        out[name] = (float(np.max(np.abs(a - b))) / scale, float(np.sqrt(np.mean((a - b) ** 2))) / scale)
    return out


# Relative L-infinity bounds of the mixed step against the FP64 oracle after 20 steps
# (3x what the MI355X gives: phi 6.5e-7 and 7.7e-7, p 1.7e-6 and 5.9e-7; the static
# drop's velocity is its spurious current, ~1e-5 of the bubble's, so its relative error
# is the float rounding of terms that nearly cancel: 8.6e-4 against 7.6e-6).
MIXED_BOUNDS = {"drop64": {"phi": 2e-6, "u": 3e-3, "v": 3e-3, "w": 3e-3, "p": 5e-6},
                "bubble32x64": {"phi": 2.5e-6, "u": 2.5e-5, "v": 2.5e-5, "w": 1e-5, "p": 2e-6}}


@pytest.mark.parametrize("case", ["drop64", "bubble32x64"])
def test_mixed_step_against_fp64_oracle(case):
    """The precision form (FP32 level set, momentum and transforms; FP64 Thomas, Poisson
    residual with one refinement, volumes) against the oracle's FP64 step, 20 steps."""
    if case == "drop64":
        p = small_drop(64, mass_correction=0)
    else:
        p = P.bubble256(32)
        p.ktot = 64
        p.cfl = 0.0
        p.timestep = 2e-4
        p.mass_correction = 0
    p.mode = P.MODE_MIXED
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 21):
            o.step(s, 0)
            g.step(s, 0)
        err = _errors(o, g, p)
        v, v0 = g.volume(0)
        vo, _ = o.volume(0)
        print("%s mixed vs FP64 oracle after 20 steps: %s; volume |V - V_oracle| / V = %.2e" %
              (case, {k: "Linf %.2e L2 %.2e" % e for k, e in err.items()}, abs(v - vo) / vo))
        for k, bound in MIXED_BOUNDS[case].items():
            assert err[k][0] <= bound, (k, err[k])
        assert abs(v - vo) <= 1e-6 * vo
    finally:
        g.close()


# ---------------------------------------------------------------- full-size properties
def _run_props(p, steps, div_tol, dv_tol):
    from iclsamd import _lib
    g = _lib.Solver(p)
    try:
        g.init_program()
        for s in range(1, steps + 1):
            g.step(s, 0)
        g.check_nan()
        rc, tot, dmax, ijk = g.chkdiv(steps)
        assert rc == 0 and dmax < div_tol, dmax
        dv = [abs(v / v0 - 1.0) for v, v0 in (g.volume(l) for l in range(p.lmax))]
        assert max(dv) <= dv_tol, max(dv)
        listed = g.mlp_listed() if p.mode == P.MODE_MIXED_LEARNED else 0
        vel = g.max_vel_at_center()
        return g, max(dv), listed, vel
    except Exception:
        g.close()
        raise


def test_config_drop512_mixed_properties():
    p = P.drop64(512)
    p.mode = P.MODE_MIXED
    p.mass_correction = 0
    g, dv, _, vel = _run_props(p, 3, 1e-3, 1e-6)
    g.close()
    print("512^3 mixed drop, 3 steps: max |V/V0 - 1| = %.2e, max |u_c| = %.2e" % (dv, vel))


def test_config_drops256_learned_properties():
    p = P.drops256(256, 64, mode=P.MODE_MIXED_LEARNED)
    p.mass_correction = 2
    p.cfl = 0.3
    g, dv, listed, vel = _run_props(p, 3, 1e-3, 5e-5)
    g.close()
    print("256^3, 64 drops, learned, 3 steps: max |V/V0 - 1| = %.2e, closure cells %d, max |u_c| = %.2e" %
          (dv, listed, vel))
    assert listed > 0


def test_config_emulsion512_learned_properties():
    p = P.emulsion512(512, 256, 200, mode=P.MODE_MIXED_LEARNED)
    g, dv, listed, vel = _run_props(p, 3, 1e-3, 5e-5)
    try:
        v = g.download(P.F_VNEW)
        wall_lo = 0.5 * (v[:, :, 0] + v[:, :, 1])
        wall_hi = 0.5 * (v[:, :, -1] + v[:, :, -2])
        assert np.max(np.abs(wall_lo + 1.0)) <= 1e-5 and np.max(np.abs(wall_hi - 1.0)) <= 1e-5
    finally:
        g.close()
    print("512x512x256 emulsion, 200 drops, learned, 3 steps: max |V/V0 - 1| = %.2e, closure cells %d" %
          (dv, listed))
