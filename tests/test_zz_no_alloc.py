"""icls_chkdiv_noalloc, the allocation-free divergence check, against icls_chkdiv.

The entry takes caller-owned device arrays in the library's layout and keeps its
block partials in a fixed device array, so it must give exactly what icls_chkdiv
gives on the same velocity fields: the same kernels run over the same reduction
blocks.  Two calls in a row, the second on a smaller grid and in float, then a grid
beyond the fixed array's capacity, which is refused before anything is launched.
The device buffers here are allocated by the test through the HIP runtime the
library itself links.
"""
import ctypes

import numpy as np
import pytest

from helpers import small_drop
from iclsamd import params as P

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from iclsamd import _lib
    L = _lib.load()
    H = ctypes.CDLL("libamdhip64.so.7")     # the runtime libicls.so is linked against
    vp = ctypes.c_void_p
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    H.hipFree.argtypes = [vp]
    H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    return L, H


def _dev(H, a):
    ptr = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(ptr), a.nbytes) == 0
    assert H.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0
    return ptr


def _fields(p, seed):
    """Random velocity fields (float-representable) with periodic x/y halo images."""
    rng = np.random.default_rng(seed)
    shape = P.field_dims(p, P.F_UNEW)
    out = []
    for _ in range(3):
        a = rng.standard_normal(shape).astype(np.float32).astype(np.float64)
        a[0], a[-1] = a[-2], a[1]
        a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
        out.append(a)
    return out


def _device_layout(a, p, dtype):
    """C-halo host array -> the library's device layout (x, y without halo; z from
    k = -2 to ktot + 3, of which k = 0 .. ktot + 1 are filled)."""
    nx, ny, nz = p.itot, p.jtot, p.ktot
    d = np.zeros((nz + 6, ny, nx), dtype=dtype)
    d[2:nz + 4] = np.transpose(a[1:-1, 1:-1, :], (2, 1, 0))
    return np.ascontiguousarray(d)


def _noalloc(L, H, p, fields, esz):
    dtype = np.float64 if esz == 8 else np.float32
    bufs = [_dev(H, _device_layout(a, p, dtype)) for a in fields]
    out = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(out), 24) == 0
    try:
        rc = L.icls_chkdiv_noalloc(bufs[0], bufs[1], bufs[2], esz, p.itot, p.jtot, p.ktot, p.lx, p.ly, p.lz, out,
                                   None)
        assert rc == 0, L.icls_last_error()
        res = np.zeros(3)
        assert H.hipMemcpy(res.ctypes.data, out, 24, D2H) == 0      # synchronous on the null stream
        # beyond the fixed partials array (1024 x 1024 x 512 needs 262144 blocks): refused on the host
        assert L.icls_chkdiv_noalloc(bufs[0], bufs[1], bufs[2], esz, 1024, 1024, 512, 1.0, 1.0, 0.5, out, None) == 1
        assert b"reduction blocks" in L.icls_last_error()
    finally:
        for b in bufs + [out]:
            H.hipFree(b)
    return res


def _with_context(p, fields):
    from iclsamd import _lib
    g = _lib.Solver(p)
    try:
        g.init_program()
        for f, a in zip((P.F_UNEW, P.F_VNEW, P.F_WNEW), fields):
            g.upload(f, a)
        return g.chkdiv(1)
    finally:
        g.close()


def test_chkdiv_noalloc_matches_chkdiv_twice_second_smaller():
    L, H = _hip()
    for n, nz, mode, esz, seed in ((32, 24, P.MODE_FP64, 8, 5), (16, 8, P.MODE_MIXED, 4, 6)):
        p = small_drop(n, nz=nz)
        p.mode = mode
        fields = _fields(p, seed)
        res = _noalloc(L, H, p, fields, esz)
        rc, tot, dmax, ijk = _with_context(p, fields)
        loc = int(res[2])
        got_ijk = (loc % n + 1, (loc // n) % n + 1, loc // (n * n) + 1)
        assert rc == 4 and dmax > 0
        assert res[1] == dmax and got_ijk == ijk, (res, dmax, ijk)
        assert res[0] == tot, (res[0], tot)
