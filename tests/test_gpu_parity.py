"""GPU (libicls.so through the C ABI) against the CPU oracle on the same inputs.

The tolerance of the FP64 path is the one fixed for the plain FP64 GPU run: 1e-10
relative L-infinity (max|gpu - oracle| / max|oracle|), here on every field compared.

The mass-correction trigger compares |V - V0|/V0 with 1e-5 and scales the correction
velocity with V - V0 (levelset.f90:386-395, 509-514), a difference of two nearly equal
volume sums: near the trigger its relative error is the reduction-order error of V
(~1e-16 |V|) divided by |V - V0|.  The step tests therefore run with the trigger off,
and the inflation path has its own test with |V - V0|/V0 = 1e-3, where the same
1e-10 bound applies.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, rel_linf, small_drop
from iclsamd import params as P

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(p):
    import pyoracle
    from iclsamd import _lib
    o = pyoracle.Oracle(p)
    g = _lib.Solver(p)
    return o, g


FIELDS = [P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW, P.F_POLD, P.F_LVSET, P.F_MIU, P.F_RHO, P.F_DUOLD, P.F_DVOLD,
          P.F_DWOLD, P.F_PX, P.F_PY, P.F_PZ, P.F_UCC, P.F_VCC, P.F_WCC]


def _compare(o, g, fields, tol, tag):
    worst = {}
    for f in fields:
        a = g.download(f)
        b = o.get(f)
        e = rel_linf(a, b) if np.max(np.abs(b)) > 0 else float(np.max(np.abs(a)))
        worst[P.FIELD_NAMES[f]] = e
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, "%s: fields beyond %.1e: %s" % (tag, tol, bad)
    return worst


def test_init_program_parity():
    p = small_drop(32)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        _compare(o, g, FIELDS, 1e-14, "init")
        v, v0 = g.volume(0)
        vo, vo0 = o.volume(0)
        # volume = 1 - sum(H)/N: reduction order differs, and the 1 - S/N cancellation
        # amplifies its ulp-level difference by about 1/volume
        assert abs(v - vo) <= 1e-12 * abs(vo)
    finally:
        g.close()


@pytest.mark.parametrize("scheme", [P.ADV_HOUC5, P.ADV_WENO5])
def test_steps_parity_drop32(scheme):
    p = small_drop(32, ls_adv_space=scheme, mass_correction=0)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 11):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, FIELDS, TOL, "10 steps")
        assert rel_linf([g.scalar(P.S_REINIT_RES)], [o.scalar(P.S_REINIT_RES)]) <= TOL
    finally:
        g.close()


def test_steps_parity_anisotropic_bubble():
    """Rising-bubble physics (density 100:1, gravity, WENO5) on a 32x32x64 grid over
    1 x 1 x 2 with dz = 2 dx, so that every dx / dy / dz is exercised distinctly."""
    p = P.bubble256(32)
    p.ktot = 64
    p.timestep = 2e-4
    p.cfl = 0.0
    p.mass_correction = 0
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 6):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, FIELDS, TOL, "bubble 5 steps")
    finally:
        g.close()


def test_steps_parity_adaptive_dt_bubble():
    """cfl = 0.3 with dt set from chkdt at init and after every second step, so that
    steps 3 and 5 run the variable-step AB2 and pressure extrapolation (r != 1).
    dt follows rho_avr, a device reduction, so it agrees to rounding, not bitwise."""
    p = P.bubble256(32)
    p.ktot = 64
    p.mass_correction = 0
    p.cfl = 0.3
    p.ioutchk = 2
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        dts = []
        for s in range(0, 7):
            if s:
                o.step(s, 0)
                g.step(s, 0)
            dg, do = g.scalar(P.S_DT), o.scalar(P.S_DT)
            assert abs(dg - do) <= 1e-13 * do, (s, dg, do)
            dts.append(do)
        assert dts[0] != p.timestep and len(set(dts)) >= 2
        _compare(o, g, FIELDS, TOL, "adaptive dt, 6 steps")
    finally:
        g.close()


def test_spurious_currents_matches_oracle():
    """spurious_currents (misc.f90:802-843) after 5 steps of the static drop: the
    reference's Ca = 0.1 max|u_c| / 1, with max|u_c| against the oracle's."""
    p = small_drop(32, mass_correction=0)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        o.run(5)
        for s in range(1, 6):
            g.step(s, 0)
        ca, vm = g.spurious_currents(), g.max_vel_at_center()
        assert ca == 0.1 * vm
        want = o.max_vel_at_center()
        assert want > 0 and abs(vm - want) <= TOL * want, (vm, want)
    finally:
        g.close()


def test_steps_parity_three_drops():
    """lmax = 3 ('mdp', init.f90:901-986): one level set per drop, materials from the
    minimum Heaviside (misc.f90:102-120), GFM jumps accumulated over the drops
    (gfm.f90:219-233) with fillps and p_jump added separately (main.f90:139-142), and
    ipot's interface distances (interface.f90:225-277)."""
    p = P.base_params(32)
    p.set_drops([(0.3, 0.3, 0.5, 0.12), (0.62, 0.45, 0.5, 0.10), (0.45, 0.7, 0.35, 0.09)])
    p.mass_correction = 0
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 6):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, [f for f in FIELDS if f != P.F_LVSET], TOL, "three drops, 5 steps")
        for l in range(3):
            e = rel_linf(g.download(P.F_LVSET, l), o.get(P.F_LVSET, l))
            assert e <= TOL, (l, e)
        dmin = g.ipot()
        for m in range(3):
            for n in range(3):
                want = o.ipot_dmin(m, n)
                assert abs(dmin[m, n] - want) <= TOL * max(abs(want), 1e-300), (m, n, dmin[m, n], want)
        assert dmin[0, 1] != 0.0 and dmin[0, 1] < 1.0      # drops 1 and 2 lie within the band
    finally:
        g.close()


def _window_drops(mode=P.MODE_FP64, mc=0):
    """Three drops of radius 4 cells on 128 x 128 x 32: the first crosses the periodic
    x boundary and keeps the whole x axis, the other two get 64-cell windows in x and y."""
    p = P.base_params(128, nz=32)
    p.lz = 0.25
    p.rho1 = p.rho2 = 1.0
    p.miu1 = p.miu2 = 1.0
    p.timestep = 1e-5
    p.mass_correction = mc
    p.mode = mode
    r = 4.0 / 128
    p.set_drops([(0.01, 0.45, 0.13, r), (0.25, 0.26, 0.125, r), (0.70, 0.66, 0.12, r)])
    return p


def _clip_ls(a, p):
    nb2 = 6.0 * p.lz / p.ktot
    return np.clip(a, -nb2, nb2)


def test_windows_match_oracle():
    """Level sets stepped inside per-drop windows (two window refreshes in 9 steps)
    against the oracle's whole-grid level sets: every field at 1e-10, the level sets
    after clipping to the narrow band (cells beyond it keep their old values outside
    a window)."""
    p = _window_drops()
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 10):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, [f for f in FIELDS if f != P.F_LVSET], TOL, "windows, 9 steps")
        for l in range(3):
            e = rel_linf(_clip_ls(g.download(P.F_LVSET, l), p), _clip_ls(o.get(P.F_LVSET, l), p))
            assert e <= TOL, (l, e)
    finally:
        g.close()


def test_windowed_inflation_matches_oracle():
    """The normalised mass correction firing for the last drop inside its window (maxchk
    is the last level set's, levelset.f90:380): the volume counts the cells outside the
    window as Heaviside 1.  The drop covers ~270 of 524288 cells (V ~ 5e-4), so the
    rounding of V = 1 - sum(H)/N (~1e-16, summed over the window here and over the grid
    in the oracle) is ~2e-11 of V - V0 at a deficit of 1e-2."""
    p = _window_drops(mc=2)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        v, _ = o.volume(2)
        o.set_volume(2, v, v * (1 + 1e-2))
        g.set_volume(2, v, v * (1 + 1e-2))
        o.step(1, 0)
        g.step(1, 0)
        mo, mg = o.scalar(P.S_MAXCHK), g.scalar(P.S_MAXCHK)
        assert mo != 0.0 and abs(mg - mo) <= 1e-10 * abs(mo), (mg, mo)
        _compare(o, g, [f for f in FIELDS if f != P.F_LVSET], TOL, "windowed inflation")
        e = rel_linf(_clip_ls(g.download(P.F_LVSET, 2), p), _clip_ls(o.get(P.F_LVSET, 2), p))
        assert e <= TOL, e
    finally:
        g.close()


def test_one_colour_batch_matches_oracle():
    """Four drops whose 64 x 64 windows tile the 128 x 128 plane: one colour, so tube,
    advection and re-initialisation run as single launches over four windows (a window
    table in the kernel arguments), and two of the drops are 1% short of volume so that
    their corrections fire inside the colour.  Against the oracle after 3 steps: every
    field and the four level sets at 1e-10."""
    p = P.base_params(128, nz=32)
    p.lz = 0.25
    p.rho1 = p.rho2 = 1.0
    p.miu1 = p.miu2 = 1.0
    p.timestep = 1e-5
    p.mass_correction = 2
    r = 4.0 / 128
    p.set_drops([(0.25, 0.25, 0.125, r), (0.75, 0.25, 0.125, r), (0.25, 0.75, 0.125, r), (0.75, 0.75, 0.125, r)])
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for l in (1, 2):
            v, _ = o.volume(l)
            o.set_volume(l, v, v * (1 + 1e-2))
            g.set_volume(l, v, v * (1 + 1e-2))
        for s in range(1, 4):
            o.step(s, 0)
            g.step(s, 0)
        _compare(o, g, [f for f in FIELDS if f != P.F_LVSET], TOL, "one colour, four windows")
        for l in range(4):
            e = rel_linf(_clip_ls(g.download(P.F_LVSET, l), p), _clip_ls(o.get(P.F_LVSET, l), p))
            assert e <= TOL, (l, e)
    finally:
        g.close()


def test_overlapping_windows_correct_in_own_band():
    """Two drops whose windows overlap (different colours), both short of volume by 1%
    so that both corrections fire: each correction advects inside its own tube's band
    mask (levelset.f90:47-68), although the other drop's tube writes the shared mask
    array where the windows overlap.  Against the oracle: every field and both level
    sets after step 2 at 1e-10, and maxchk of the last level set after step 1, where
    the 1% deficit sets it.  maxchk is proportional to V - V0 with V = 1 - sum(H)/N
    (levelset.f90:505): the sum of N ~ 5e5 terms near 1 is taken in another order here
    than in the oracle, which moves V by about sqrt(N) eps, so maxchk agrees to about
    sqrt(N) eps / |V - V0| (4x that is the bound; ~6e-8 here)."""
    p = P.base_params(128, nz=32)
    p.lz = 0.25
    p.rho1 = p.rho2 = 1.0
    p.miu1 = p.miu2 = 1.0
    p.timestep = 1e-5
    p.mass_correction = 2
    r = 5.0 / 128
    p.set_drops([(0.40, 0.50, 0.125, r), (0.525, 0.52, 0.125, r)])
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for l in range(2):
            v, _ = o.volume(l)
            o.set_volume(l, v, v * (1 + 1e-2))
            g.set_volume(l, v, v * (1 + 1e-2))
        o.step(1, 0)
        g.step(1, 0)
        mo, mg = o.scalar(P.S_MAXCHK), g.scalar(P.S_MAXCHK)
        n = p.itot * p.jtot * p.ktot
        bound = 4 * np.sqrt(n) * np.finfo(float).eps / (1e-2 * v)
        assert mo != 0.0 and abs(mg - mo) <= bound * abs(mo), (mg, mo, bound)
        o.step(2, 0)
        g.step(2, 0)
        _compare(o, g, [f for f in FIELDS if f != P.F_LVSET], TOL, "overlapping windows, corrections")
        for l in range(2):
            e = rel_linf(_clip_ls(g.download(P.F_LVSET, l), p), _clip_ls(o.get(P.F_LVSET, l), p))
            assert e <= TOL, (l, e)
    finally:
        g.close()


def test_windows_equal_whole_grid_learned():
    """MIXED_LEARNED (FP32 fields, learned closure): windowed and whole-grid level
    sets give the same fields after 9 steps, and the same volumes bit for bit (the
    windowed materials kernel sums each drop's Heaviside over the blocks its window
    meets and counts the cells of the others)."""
    from iclsamd import _lib
    out = []
    for full in (0, 1):
        p = _window_drops(mode=P.MODE_MIXED_LEARNED)
        p.ls_full = full
        g = _lib.Solver(p)
        try:
            g.init_program()
            for s in range(1, 10):
                g.step(s, 0)
            out.append({f: g.download(f) for f in (P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW)})
            out[-1]["ls"] = [_clip_ls(g.download(P.F_LVSET, l), p) for l in range(3)]
            out[-1]["vol"] = [g.volume(l) for l in range(3)]
        finally:
            g.close()
    for f in (P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW):
        assert rel_linf(out[0][f], out[1][f]) <= 1e-12, P.FIELD_NAMES[f]
    for l in range(3):
        assert rel_linf(out[0]["ls"][l], out[1]["ls"][l]) <= 1e-12, l
    assert out[0]["vol"] == out[1]["vol"], (out[0]["vol"], out[1]["vol"])


def test_restart_from_fld_matches_oracle(tmp_path):
    """Checkpoint in the reference's fld format (loadd.f90:16-108) and restart
    (init.f90:44-51): the file holds u, v, w, p, each level set and jiemian (interiors,
    FP64, x fastest), then time, nr, dt; the restarted GPU run and the oracle loaded
    from the same file and re-derived (init.f90:63-110) agree step by step."""
    import pyoracle
    from iclsamd import _lib
    p = small_drop(32, mass_correction=0)
    n = p.itot * p.jtot * p.ktot
    path = str(tmp_path / "fld0000004")
    g = _lib.Solver(p)
    try:
        g.init_program()
        for s in range(1, 5):
            g.step(s, 0)
        u_before = g.download(P.F_UNEW)
        g.write_fld(path, 4)
    finally:
        g.close()
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == (4 + p.lmax + 1) * n + 3
    time_, nr, dt = raw[-3:]
    assert nr == 4.0 and dt == p.timestep and abs(time_ - 4 * p.timestep) <= 1e-15
    blocks = [raw[b * n:(b + 1) * n].reshape((p.itot, p.jtot, p.ktot), order="F") for b in range(4 + p.lmax + 1)]
    assert np.array_equal(blocks[0], u_before[1:-1, 1:-1, 1:-1])

    o = pyoracle.Oracle(p)
    o.init_program()
    for b, f in enumerate([P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW]):
        a = np.zeros(P.field_dims(p, f), order="F")
        a[1:-1, 1:-1, 1:-1] = blocks[b]
        o.set(f, a)
    a = np.zeros(P.field_dims(p, P.F_LVSET), order="F")
    a[3:-3, 3:-3, 3:-3] = blocks[4]
    o.set(P.F_LVSET, a)
    o.set_scalar(P.S_TIME, time_)
    o.init_derived()
    g = _lib.Solver(p)
    try:
        begin = g.init_program_restart(path)
        assert begin == 4
        _compare(o, g, FIELDS, 1e-13, "restart state")
        for s in range(begin + 1, begin + 5):
            o.step(s, begin)
            g.step(s, begin)
        _compare(o, g, FIELDS, TOL, "4 steps after restart")
    finally:
        g.close()


@pytest.mark.parametrize("mc", [1, 2])
def test_inflation_parity(mc):
    """One step with the mass correction firing (|V - V0|/V0 = 1e-3): the reference's
    normalisation (1) and the normalised variant (2)."""
    p = small_drop(32, mass_correction=mc)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        v, v0 = o.volume(0)
        o.set_volume(0, v, v * (1 + 1e-3))
        g.set_volume(0, v, v * (1 + 1e-3))
        o.step(1, 0)
        g.step(1, 0)
        mo, mg = o.scalar(P.S_MAXCHK), g.scalar(P.S_MAXCHK)
        assert mo != 0.0 and abs(mg - mo) <= 1e-10 * abs(mo), (mg, mo)
        _compare(o, g, FIELDS, TOL, "inflation step")
    finally:
        g.close()


def test_lowered_mass_trigger_parity():
    """mass_trigger_exp = 7: a deficit of 3e-7, below the reference's 1e-5 trigger, fires
    the normalised correction on both sides and not with the default trigger.  maxchk is
    proportional to V - V0, so it agrees to about sqrt(N) eps / |V - V0| (4x that is the
    bound, as for the windowed corrections)."""
    for exp, fires in ((0, False), (7, True)):
        p = small_drop(32, mass_correction=2)
        p.mass_trigger_exp = exp
        o, g = _pair(p)
        try:
            o.init_program()
            g.init_program()
            v, v0 = o.volume(0)
            o.set_volume(0, v, v * (1 + 3e-7))
            g.set_volume(0, v, v * (1 + 3e-7))
            o.step(1, 0)
            g.step(1, 0)
            mo, mg = o.scalar(P.S_MAXCHK), g.scalar(P.S_MAXCHK)
            assert (mo != 0.0) == fires and (mg != 0.0) == fires, (exp, mo, mg)
            bound = 4 * np.sqrt(32 ** 3) * np.finfo(float).eps / (3e-7 * v)
            assert abs(mg - mo) <= bound * abs(mo), (mg, mo, bound)
            _compare(o, g, FIELDS, 1e-9 if fires else TOL, "trigger 10^-%d" % exp)
        finally:
            g.close()


@pytest.mark.slow
def test_drop64_100_steps_parity():
    """The CPU-runnable workload: 64^3 static drop, 100 AB2/FFT-projection steps
    (mass correction on; it does not fire within 100 steps here)."""
    p = P.drop64()
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        for s in range(1, 101):
            o.step(s, 0)
            g.step(s, 0)
        assert o.scalar(P.S_MAXCHK) == 0.0
        worst = _compare(o, g, [P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW, P.F_LVSET], TOL, "100 steps")
        print("drop64 x100 worst relative Linf:", worst)
    finally:
        g.close()


def test_fortran_driver_matches_oracle(tmp_path):
    """The Fortran main (main.f90 call order through iso_c_binding) on the GPU."""
    exe = os.path.join(ROOT, "icls-release_amd", "icls_fortran")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    n, steps = 32, 10
    out = tmp_path / "fields.bin"
    nml = tmp_path / "run.nml"
    nml.write_text("&icls\n itot=%d, jtot=%d, ktot=%d, nstep=%d, ioutchk=5, drop_r(1)=0.2,\n"
                   " mass_correction=0, out_file='%s'\n/\n" % (n, n, n, steps, out))
    r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation completed" in r.stdout
    p = small_drop(n, mass_correction=0)
    import pyoracle
    o = pyoracle.Oracle(p)
    o.init_program()
    for s in range(1, steps + 1):
        o.step(s, 0)
    raw = np.fromfile(str(out), dtype=np.float64)
    nc = (n + 2) ** 3
    got = {P.F_UNEW: raw[0:nc], P.F_VNEW: raw[nc:2 * nc], P.F_WNEW: raw[2 * nc:3 * nc],
           P.F_PNEW: raw[3 * nc:4 * nc], P.F_LVSET: raw[4 * nc:]}
    for f, a in got.items():
        b = o.get(f).ravel(order="F")
        assert rel_linf(a, b) <= TOL, P.FIELD_NAMES[f]


@pytest.mark.parametrize("n,nz", [(512, 8), (256, 16), (64, 32), (16, 8)])
def test_solver2d_parity(n, nz):
    """solver2d alone on a random zero-mean right-hand side: FFT x, FFT y, Thomas z,
    inverse FFTs (solver.f90:68-196), at the line lengths of the workloads."""
    p = small_drop(n, nz=nz)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        rng = np.random.default_rng(n + nz)
        f = np.zeros(P.field_dims(p, P.F_P), order="F")
        f[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, nz))
        f[1:-1, 1:-1, 1:-1] -= f[1:-1, 1:-1, 1:-1].mean()
        o.set(P.F_P, f)
        g.upload(P.F_P, f)
        o.solver2d()
        g.solver2d()
        a = g.download(P.F_P)[1:-1, 1:-1, 1:-1]
        b = o.get(P.F_P)[1:-1, 1:-1, 1:-1]
        assert rel_linf(a, b) <= TOL
    finally:
        g.close()


@pytest.mark.parametrize("bcy", [P.BCY_WALLS, P.BCY_INOUT, P.BCY_OUTOUT])
@pytest.mark.parametrize("n,nz,mode", [(16, 8, P.MODE_FP64), (64, 24, P.MODE_FP64), (32, 80, P.MODE_FP64),
                                       (256, 16, P.MODE_FP64), (32, 40, P.MODE_MIXED), (16, 128, P.MODE_FP64),
                                       (16, 160, P.MODE_FP64)])
def test_solver2d_cos_parity(bcy, n, nz, mode):
    """solver2d_cos (solver.f90:202-317) with each BC_in_y rule (initsolver.f90:84-124):
    FFT x, VCOSQB z, Gauss elimination y, VCOSQF z, inverse FFT x.  nz = 24, 40 and 80
    leave partial 16-layer passes and 64-output tiles in the dense z transform; nz = 128
    and 160 take the halved (even/odd) products, 160 with partial tiles; the mixed mode
    runs the same FP64 path on its double pressure."""
    p = small_drop(n, nz=nz, mode=mode)
    o, g = _pair(p)
    try:
        o.init_program()
        g.init_program()
        rng = np.random.default_rng(100 * bcy + n + nz)
        f = np.zeros(P.field_dims(p, P.F_P), order="F")
        f[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, nz))
        if bcy == P.BCY_WALLS:
            f[1:-1, 1:-1, 1:-1] -= f[1:-1, 1:-1, 1:-1].mean()
        o.set(P.F_P, f)
        g.upload(P.F_P, f)
        assert o.solver2d_cos(bcy) == 0
        g.solver2d_cos(bcy)
        a = g.download(P.F_P)[1:-1, 1:-1, 1:-1]
        b = o.get(P.F_P)[1:-1, 1:-1, 1:-1]
        e = rel_linf(a, b)
        print("solver2d_cos bc_y=%d %dx%dx%d mode %d: rel Linf vs oracle %.3e" % (bcy, n, n, nz, mode, e))
        assert e <= TOL
        # twice on the same context: the tables built on the first call are reused
        g.upload(P.F_P, f)
        g.solver2d_cos(bcy)
        assert np.array_equal(g.download(P.F_P)[1:-1, 1:-1, 1:-1], a)
    finally:
        g.close()


def test_solver2d_cos_rejects_bad_rule():
    from iclsamd import _lib
    g = _lib.Solver(small_drop(16, nz=8))
    try:
        g.init_program()
        with pytest.raises(_lib.IclsError):
            g.solver2d_cos(0)
        with pytest.raises(_lib.IclsError):
            g.solver2d_cos(4)
    finally:
        g.close()
