"""Runtime parameter block of the ICLS step and the workload presets.

``IclsParams`` mirrors ``struct icls_params`` of include/icls.h field for field; it
replaces the compile-time PARAMETERs of the reference's src/param.f90:12-172 and
mod_param_ls (src/param.f90:199-275).

Presets (phase 1 = phi > 0 = outer fluid, phase 2 = drop or bubble, as in
src/param.f90:58-60 and src/misc.f90:124,146):

* ``drop64``   -- 64^3 static drop, R = 0.2, density and viscosity of the drop 10x the
                  outer fluid, sigma = 1 (We = 1), Re = 10, HOUC5, re-initialisation and
                  mass correction every step, dt = 1e-4.
* ``bubble256`` -- 256^3 rising bubble in a 1 x 1 x 2 box (dz = 2 dx), R = 0.125 at
                  z = 0.5, density ratio 100:1, viscosity ratio 50:1, gravity -z,
                  Eotvos ~ 10, WENO5 advection, dt = 0.3 / (chkdt's combined speed),
                  updated every ioutchk = 20 steps (``cfl``; the reference keeps dt
                  fixed, so this is an extension).
* ``drops256`` -- 256^3 box with seeded random non-overlapping static drops, ratios 1:1.
* ``emulsion512`` -- 512 x 512 x 256 plane Couette flow (walls at -1 and +1) with 200
                  seeded drops of radius 4 .. 16 cells, ratios 1:1, Ca = 0.2.
"""
import ctypes
import math

MAX_DROPS = 1024

MODE_FP64, MODE_MIXED, MODE_MIXED_LEARNED = 0, 1, 2
ADV_HOUC5, ADV_WENO5 = 0, 1
BCY_WALLS, BCY_INOUT, BCY_OUTOUT = 1, 2, 3   # BC_in_y of solver2d_cos (initsolver.f90:84-124)
INIU_ZER, INIU_COU = 0, 1

# field ids (enum icls_field)
(F_UNEW, F_VNEW, F_WNEW, F_DUDT, F_DVDT, F_DWDT, F_PNEW, F_POLD, F_P, F_PJUMP,
 F_DUOLD, F_DVOLD, F_DWOLD, F_PX, F_PY, F_PZ, F_PXOLD, F_PYOLD, F_PZOLD,
 F_MIU, F_RHO, F_JIEMIAN, F_CURV, F_UCC, F_VCC, F_WCC, F_RHO_U, F_RHO_V, F_RHO_W,
 F_LVSET, F_NORMAL_M, F_COUNT) = range(32)

SHAPE_C, SHAPE_LS, SHAPE_I = 0, 1, 2

FIELD_NAMES = ["unew", "vnew", "wnew", "dudt", "dvdt", "dwdt", "pnew", "pold", "p", "p_jump",
               "duold", "dvold", "dwold", "p_x", "p_y", "p_z", "p_xold", "p_yold", "p_zold",
               "miu", "rho", "jiemian", "curv", "ucc", "vcc", "wcc", "rho_u", "rho_v", "rho_w",
               "lvset", "normal_m"]

# scalar ids (enum icls_scalar)
S_TIME, S_DT, S_RHO_AVR, S_V_BULK, S_MAXCHK, S_REINIT_RES, S_NORM_OLD = range(7)


def field_shape_kind(field):
    if field in (F_DUOLD, F_DVOLD, F_DWOLD, F_UCC, F_VCC, F_WCC, F_RHO_U, F_RHO_V, F_RHO_W, F_NORMAL_M):
        return SHAPE_I
    if field == F_LVSET:
        return SHAPE_LS
    return SHAPE_C


def field_dims(p, field):
    """Fortran extents (nx, ny, nz) of a field's host array, i fastest."""
    kind = field_shape_kind(field)
    if kind == SHAPE_C:
        return (p.itot + 2, p.jtot + 2, p.ktot + 2)
    if kind == SHAPE_LS:
        return (p.itot + 6, p.jtot + 6, p.ktot + 6)
    return (p.itot, p.jtot, p.ktot)


class IclsParams(ctypes.Structure):
    _fields_ = [
        ("itot", ctypes.c_int32), ("jtot", ctypes.c_int32), ("ktot", ctypes.c_int32),
        ("lmax", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("timestep", ctypes.c_double),
        ("rho1", ctypes.c_double), ("rho2", ctypes.c_double),
        ("miu1", ctypes.c_double), ("miu2", ctypes.c_double),
        ("Re", ctypes.c_double), ("We", ctypes.c_double), ("iFr", ctypes.c_double),
        ("g_unit", ctypes.c_double * 3),
        ("ls_adv_space", ctypes.c_int32),
        ("re_initialization", ctypes.c_int32),
        ("reinit_step", ctypes.c_int32),
        ("reinit_iter_max", ctypes.c_int32),
        ("mass_correction", ctypes.c_int32),
        ("inflate_step", ctypes.c_int32),
        ("ns_np", ctypes.c_int32),
        ("wall_speed", ctypes.c_double),
        ("constPoi", ctypes.c_int32),
        ("iniu", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("ioutchk", ctypes.c_int32),
        ("cfl", ctypes.c_double),
        ("ls_full", ctypes.c_int32),
        ("mass_trigger_exp", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
        ("drop_xc", ctypes.c_double * MAX_DROPS),
        ("drop_yc", ctypes.c_double * MAX_DROPS),
        ("drop_zc", ctypes.c_double * MAX_DROPS),
        ("drop_r", ctypes.c_double * MAX_DROPS),
    ]

    def set_drops(self, drops):
        if len(drops) > MAX_DROPS:
            raise ValueError("at most %d drops" % MAX_DROPS)
        self.lmax = len(drops)
        for l, (x, y, z, r) in enumerate(drops):
            self.drop_xc[l], self.drop_yc[l], self.drop_zc[l], self.drop_r[l] = x, y, z, r

    def copy(self):
        q = IclsParams()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(IclsParams))
        return q


def base_params(n=64, lx=1.0, ly=1.0, lz=1.0, nz=None):
    p = IclsParams()
    p.itot = n
    p.jtot = n
    p.ktot = n if nz is None else nz
    p.lx, p.ly, p.lz = lx, ly, lz
    p.timestep = 1e-4
    p.rho1, p.rho2 = 1.0, 10.0
    p.miu1, p.miu2 = 1.0, 10.0
    p.Re, p.We, p.iFr = 10.0, 1.0, 0.0
    p.g_unit[0], p.g_unit[1], p.g_unit[2] = 0.0, 0.0, 0.0
    p.ls_adv_space = ADV_HOUC5
    p.re_initialization = 1
    p.reinit_step = 1
    p.reinit_iter_max = 1
    p.mass_correction = 1
    p.inflate_step = 1
    p.ns_np = 1
    p.wall_speed = 1.5
    p.constPoi = 0
    p.iniu = INIU_ZER
    p.mode = MODE_FP64
    p.ioutchk = 20
    p.set_drops([(0.5 * lx, 0.5 * ly, 0.5 * lz, 0.2)])
    return p


def drop64(n=64, steps_note=None):
    """Static drop: 64^3, R = 0.2 at the centre, 100 steps at dt = 1e-4."""
    return base_params(n)


def bubble256(n=256, mode=MODE_FP64):
    """Rising bubble: n x n x n cells over 1 x 1 x 2, R = 0.125 at z = 0.5."""
    p = base_params(n, 1.0, 1.0, 2.0)
    p.rho1, p.rho2 = 1.0, 0.01
    p.miu1, p.miu2 = 1.0, 0.02
    p.Re = 100.0
    p.iFr = 1.0
    p.g_unit[2] = -1.0
    # Eotvos = (rho1 - rho2) g (2R)^2 / sigma = 10, with sigma = 1/We
    sigma = (p.rho1 - p.rho2) * 1.0 * (2 * 0.125) ** 2 / 10.0
    p.We = 1.0 / sigma
    p.ls_adv_space = ADV_WENO5
    p.timestep = 2.0e-5        # unused while cfl > 0
    p.cfl = 0.3
    p.mode = mode
    p.set_drops([(0.5, 0.5, 0.5, 0.125)])
    return p


def splitmix64(seed):
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield (z ^ (z >> 31)) / float(1 << 64)


def random_drops(ndrops, n, rmin_cells, rmax_cells, seed=20261202, gap_cells=6, lx=1.0, ly=1.0, lz=1.0,
                 max_tries=200000):
    """Seeded non-overlapping drops (centres uniform in the box, radii uniform in
    [rmin, rmax] cells), rejection with a gap of ``gap_cells`` cells between
    interfaces and from the z walls.

    The level sets are initialised with the +lx and +ly periodic images only
    (init.f90:945-980), so a drop whose narrow band (radius + 7 cells) crosses the
    x = lx (y = ly) seam is given the centre x - lx (y - ly), whose +lx image then
    covers the far side; a drop crossing an x and a y seam at once would need the
    diagonal image as well and is drawn again."""
    dx = lx / n
    rng = splitmix64(seed)
    drops = []
    tries = 0
    while len(drops) < ndrops:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d drops" % ndrops)
        r = (rmin_cells + (rmax_cells - rmin_cells) * next(rng)) * dx
        x = lx * next(rng)
        y = ly * next(rng)
        z = r + gap_cells * dx + (lz - 2 * (r + gap_cells * dx)) * next(rng)
        ok = True
        for (x2, y2, z2, r2) in drops:
            ddx = min(abs(x - x2), lx - abs(x - x2))
            ddy = min(abs(y - y2), ly - abs(y - y2))
            if math.sqrt(ddx * ddx + ddy * ddy + (z - z2) ** 2) < r + r2 + gap_cells * dx:
                ok = False
                break
        if ok:
            m = r + 7 * dx
            cross_x = x + m > lx or x - m < 0
            cross_y = y + m > ly or y - m < 0
            if cross_x and cross_y:
                continue
            drops.append((x - lx if x + m > lx else x, y - ly if y + m > ly else y, z, r))
    return drops


def drops256(n=256, ndrops=64, mode=MODE_FP64, seed=20261202):
    """Static drops with density and viscosity ratio 1:1 (spurious-current test)."""
    p = base_params(n)
    p.rho1 = p.rho2 = 1.0
    p.miu1 = p.miu2 = 1.0
    p.mode = mode
    p.set_drops(random_drops(ndrops, n, 3.0, 20.0, seed=seed))
    return p


def emulsion512(n=512, nz=256, ndrops=200, mode=MODE_MIXED_LEARNED, seed=20261202, ca=0.2):
    """Drops in plane Couette flow: n x n x nz cells over 1 x 1 x nz/n, z walls moving at
    -1 and +1 along y (ns_np = 0, wall_speed = 1, Couette start), seeded drops of radius
    4 .. 16 cells, density and viscosity ratio 1:1, Re = 10, capillary number
    Ca = nu gamma R_ref / sigma = ca with gamma = 2 / lz and R_ref = 10 cells, dt from chkdt
    at CFL 0.3, the normalised mass correction."""
    p = base_params(n, 1.0, 1.0, nz / n, nz=nz)
    p.rho1 = p.rho2 = 1.0
    p.miu1 = p.miu2 = 1.0
    p.Re = 10.0
    p.ns_np = 0
    p.wall_speed = 1.0
    p.iniu = INIU_COU
    gamma = 2.0 / p.lz
    sigma = (1.0 / p.Re) * gamma * (10.0 / n) / ca
    p.We = 1.0 / sigma
    p.cfl = 0.3
    p.ioutchk = 20
    p.mass_correction = 2
    p.mode = mode
    p.set_drops(random_drops(ndrops, n, 4.0, 16.0, seed=seed, lx=1.0, ly=1.0, lz=p.lz))
    return p


PRESETS = {"drop64": drop64, "bubble256": bubble256, "drops256": drops256, "emulsion512": emulsion512}
