"""ctypes binding of libicls.so, the C ABI declared in include/icls.h.

The library is built in place by icls-release_amd/Makefile.  There is no CPU
fallback: if the shared object is missing or no GPU is visible every entry point
raises.
"""
import ctypes
import os

import numpy as np

from . import params as P

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libicls.so")

ERRORS = {0: "ICLS_OK", 1: "ICLS_ERR_PARAM", 2: "ICLS_ERR_HIP", 3: "ICLS_ERR_NAN", 4: "ICLS_ERR_DIVERGENCE",
          5: "ICLS_ERR_STATE", 6: "ICLS_ERR_NODEVICE", 7: "ICLS_ERR_UNSUPPORTED"}

# every symbol include/icls.h declares
EXPORTS = [
    "icls_last_error", "icls_version", "icls_init", "icls_finalize", "icls_field_shape", "icls_init_program",
    "icls_init_derived", "icls_upload", "icls_download", "icls_get_scalar", "icls_set_scalar", "icls_get_volume",
    "icls_set_volume",
    "icls_sync", "icls_ls_main", "icls_average_stream_vel", "icls_ab2", "icls_fillps", "icls_gfm4p",
    "icls_add_pjump", "icls_solver2d", "icls_solver2d_cos", "icls_boundp", "icls_correc", "icls_bounduvw",
    "icls_update_pressure",
    "icls_average_vel_at_center", "icls_check_nan", "icls_chkdiv", "icls_chkdiv_noalloc", "icls_chkdt",
    "icls_max_vel_at_center",
    "icls_spurious_currents",
    "icls_ipot", "icls_write_fld", "icls_init_program_restart", "icls_write_vtk", "icls_post2d",
    "icls_step", "icls_profile_enable", "icls_profile_only", "icls_profile_reset", "icls_kernel_times",
    "icls_mlp_defaults", "icls_mlp_train", "icls_mlp_set_weights", "icls_mlp_get_weights", "icls_mlp_set_gate",
    "icls_mlp_forward", "icls_mlp_gradient", "icls_mlp_samples", "icls_mlp_listed",
]

MLP_NPARAM = 6337
MLP_NPAD = 32


class MlpTrainOpts(ctypes.Structure):
    """icls_mlp_train_opts (include/icls.h)."""
    _fields_ = [
        ("n_samples", ctypes.c_int64), ("epochs", ctypes.c_int32), ("batch", ctypes.c_int32),
        ("lr", ctypes.c_float), ("r_min", ctypes.c_float), ("r_max", ctypes.c_float),
        ("seed", ctypes.c_uint64), ("loss_first", ctypes.c_float), ("loss_last", ctypes.c_float),
    ]


class IclsError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__("%s returned %s: %s" % (where, ERRORS.get(code, code), msg))
        self.code = code


_LIB = None


def load():
    """Load libicls.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libicls.so not built: run `make -C icls-release_amd` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    i32, dbl, dp = ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_double)
    L.icls_last_error.restype = ctypes.c_char_p
    L.icls_version.restype = i32
    L.icls_init.argtypes = [ctypes.POINTER(P.IclsParams)]
    L.icls_field_shape.argtypes = [i32, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_int64)]
    L.icls_upload.argtypes = [i32, i32, dp]
    L.icls_download.argtypes = [i32, i32, dp]
    L.icls_get_scalar.argtypes = [i32, dp]
    L.icls_set_scalar.argtypes = [i32, dbl]
    L.icls_get_volume.argtypes = [i32, dp, dp]
    L.icls_set_volume.argtypes = [i32, dbl, dbl]
    for name in ["icls_ls_main", "icls_solver2d_cos", "icls_boundp", "icls_bounduvw", "icls_update_pressure",
                 "icls_profile_enable"]:
        getattr(L, name).argtypes = [i32]
    L.icls_average_stream_vel.argtypes = [dp]
    L.icls_chkdiv.argtypes = [i32, dp, dp, ctypes.POINTER(i32)]
    vp = ctypes.c_void_p
    L.icls_chkdiv_noalloc.argtypes = [vp, vp, vp, i32, i32, i32, i32, dbl, dbl, dbl, vp, vp]
    L.icls_chkdt.argtypes = [dp]
    L.icls_max_vel_at_center.argtypes = [dp]
    L.icls_spurious_currents.argtypes = [dp]
    L.icls_ipot.argtypes = [dp]
    L.icls_write_fld.argtypes = [ctypes.c_char_p, i32]
    L.icls_init_program_restart.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]
    L.icls_write_vtk.argtypes = [ctypes.c_char_p]
    L.icls_post2d.argtypes = [ctypes.c_char_p]
    L.icls_step.argtypes = [i32, i32]
    L.icls_profile_only.argtypes = [ctypes.c_char_p]
    L.icls_kernel_times.argtypes = [i32, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(i32)]
    fp, u16p = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint16)
    L.icls_mlp_defaults.argtypes = [ctypes.POINTER(MlpTrainOpts)]
    L.icls_mlp_train.argtypes = [ctypes.POINTER(MlpTrainOpts)]
    L.icls_mlp_set_weights.argtypes = [fp, i32]
    L.icls_mlp_get_weights.argtypes = [fp, i32]
    L.icls_mlp_set_gate.argtypes = [ctypes.c_float, ctypes.c_float]
    L.icls_mlp_forward.argtypes = [u16p, i32, fp]
    L.icls_mlp_gradient.argtypes = [u16p, fp, i32, fp]
    L.icls_mlp_samples.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_float, u16p, fp, fp]
    L.icls_mlp_listed.argtypes = [ctypes.POINTER(i32)]
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is ctypes.c_int:
            fn.restype = i32
    _LIB = L
    return L


def check(code, where):
    if code != 0:
        msg = load().icls_last_error()
        raise IclsError(code, where, msg.decode() if msg else "")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    """The device-resident ICLS state for one parameter block (one per process)."""

    def __init__(self, p):
        self.L = load()
        self.p = p.copy()
        check(self.L.icls_init(ctypes.byref(self.p)), "icls_init")

    def close(self):
        if self.L is not None:
            self.L.icls_finalize()
            self.L = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, name, *args):
        check(getattr(self.L, name)(*args), name)

    # --- lifecycle
    def init_program(self):
        self._call("icls_init_program")

    def init_derived(self):
        self._call("icls_init_derived")

    # --- fields
    def download(self, fid, l=0):
        dims = P.field_dims(self.p, fid)
        out = np.zeros(dims, dtype=np.float64, order="F")
        self._call("icls_download", fid, l, _dp(out))
        return out

    def upload(self, fid, arr, l=0):
        dims = P.field_dims(self.p, fid)
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64))
        if a.shape != dims:
            raise ValueError("field %d expects shape %s, got %s" % (fid, dims, a.shape))
        self._call("icls_upload", fid, l, _dp(a))

    def scalar(self, which):
        v = ctypes.c_double()
        self._call("icls_get_scalar", which, ctypes.byref(v))
        return v.value

    def set_scalar(self, which, value):
        self._call("icls_set_scalar", which, float(value))

    def volume(self, l=0):
        v, v0 = ctypes.c_double(), ctypes.c_double()
        self._call("icls_get_volume", l, ctypes.byref(v), ctypes.byref(v0))
        return v.value, v0.value

    def set_volume(self, l, volume, init_volume):
        self._call("icls_set_volume", l, float(volume), float(init_volume))

    def sync(self):
        self._call("icls_sync")

    # --- operators (main.f90 order)
    def step(self, istep, begin=0):
        self._call("icls_step", istep, begin)

    def run(self, nsteps, begin=0):
        for istep in range(begin + 1, begin + nsteps + 1):
            self.step(istep, begin)

    def ls_main(self, istep):
        self._call("icls_ls_main", istep)

    def average_stream_vel(self):
        v = ctypes.c_double()
        self._call("icls_average_stream_vel", ctypes.byref(v))
        return v.value

    def ab2(self):
        self._call("icls_ab2")

    def fillps(self):
        self._call("icls_fillps")

    def gfm4p(self):
        self._call("icls_gfm4p")

    def add_pjump(self):
        self._call("icls_add_pjump")

    def solver2d(self):
        self._call("icls_solver2d")

    def solver2d_cos(self, bc_y):
        """solver2d_cos (solver.f90:202-317) on p with BC_in_y = bc_y (P.BCY_*)."""
        self._call("icls_solver2d_cos", bc_y)

    def boundp(self, fid):
        self._call("icls_boundp", fid)

    def correc(self):
        self._call("icls_correc")

    def bounduvw(self, which):
        self._call("icls_bounduvw", which)

    def update_pressure(self, first_step):
        self._call("icls_update_pressure", 1 if first_step else 0)

    def average_vel_at_center(self):
        self._call("icls_average_vel_at_center")

    def check_nan(self):
        self._call("icls_check_nan")

    def chkdiv(self, istep=1):
        dt, dm = ctypes.c_double(), ctypes.c_double()
        ijk = (ctypes.c_int32 * 3)()
        code = self.L.icls_chkdiv(istep, ctypes.byref(dt), ctypes.byref(dm), ijk)
        if code not in (0, 4):
            check(code, "icls_chkdiv")
        return code, dt.value, dm.value, tuple(ijk)

    def chkdt(self):
        v = ctypes.c_double()
        self._call("icls_chkdt", ctypes.byref(v))
        return v.value

    def max_vel_at_center(self):
        v = ctypes.c_double()
        self._call("icls_max_vel_at_center", ctypes.byref(v))
        return v.value

    def spurious_currents(self):
        """spurious_currents (misc.f90:802-843): Ca = 0.1 max|u_c| / 1."""
        v = ctypes.c_double()
        self._call("icls_spurious_currents", ctypes.byref(v))
        return v.value

    def ipot(self):
        """dmin (lmax x lmax) of interface.f90:225-277."""
        L = self.p.lmax
        d = np.zeros(L * L)
        self._call("icls_ipot", _dp(d))
        return d.reshape((L, L), order="F")

    # --- checkpoint / restart (src/loadd.f90)
    def write_fld(self, path, nr):
        self._call("icls_write_fld", os.fsencode(path), nr)

    def init_program_restart(self, path):
        """Restart from an fld file; returns the step number stored in it."""
        b = ctypes.c_int32()
        self._call("icls_init_program_restart", os.fsencode(path), ctypes.byref(b))
        return b.value

    # --- output files (src/write_vtk.f90, src/output.f90)
    def write_vtk(self, path):
        self._call("icls_write_vtk", os.fsencode(path))

    def post2d(self, path):
        self._call("icls_post2d", os.fsencode(path))

    # --- learned curvature closure
    def mlp_train(self, **kw):
        """Train on the device; keyword overrides of icls_mlp_train_opts.  Returns the
        options with loss_first / loss_last filled in, as a dict."""
        o = MlpTrainOpts()
        self.L.icls_mlp_defaults(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        self._call("icls_mlp_train", ctypes.byref(o))
        return {k: getattr(o, k) for k, _ in MlpTrainOpts._fields_}

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._call("icls_mlp_set_weights", w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.size)

    def mlp_get_weights(self):
        w = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._call("icls_mlp_get_weights", w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.size)
        return w

    def mlp_set_gate(self, gate, tol):
        self._call("icls_mlp_set_gate", gate, tol)

    def mlp_forward(self, x_bf16):
        """x_bf16: (n, 32) uint16 bfloat16 bit patterns."""
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        y = np.zeros(x.shape[0], dtype=np.float32)
        self._call("icls_mlp_forward", x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), x.shape[0],
                   y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return y

    def mlp_gradient(self, x_bf16, t):
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_NPARAM + 1, dtype=np.float32)
        self._call("icls_mlp_gradient", x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                   t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), x.shape[0],
                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return g[:-1], float(g[-1])

    def mlp_samples(self, n, seed=20261202, r_min=1.5, r_max=8.0):
        """(features (n, 32) bf16 bits, targets = correction * dz, coarse kappa * dz)"""
        x = np.zeros((n, MLP_NPAD), dtype=np.uint16)
        t = np.zeros(n, dtype=np.float32)
        fd = np.zeros(n, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._call("icls_mlp_samples", n, seed, r_min, r_max, x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                   t.ctypes.data_as(fp), fd.ctypes.data_as(fp))
        return x, t, fd

    def mlp_listed(self):
        v = ctypes.c_int32()
        self._call("icls_mlp_listed", ctypes.byref(v))
        return v.value

    # --- profiling
    def profile(self, on=True):
        self._call("icls_profile_enable", 1 if on else 0)

    def profile_only(self, name=None):
        self._call("icls_profile_only", name.encode() if name else None)

    def profile_reset(self):
        self._call("icls_profile_reset")

    def kernel_times(self, n=256):
        names = ctypes.create_string_buffer(64 * n)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        nout = ctypes.c_int32()
        self._call("icls_kernel_times", n, names, ms, cnt, ctypes.byref(nout))
        res = {}
        for q in range(nout.value):
            nm = names.raw[64 * q:64 * q + 64].split(b"\0", 1)[0].decode()
            res[nm] = (ms[q], cnt[q])
        return res
