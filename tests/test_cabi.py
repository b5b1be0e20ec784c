"""C ABI of libicls.so (CPU: no compute calls).

Checks that the library loads, exports every entry point include/icls.h declares,
that the Python and Fortran mirrors of struct icls_params have the C layout, and
that entry points fail with status codes (never a crash) without a GPU or context.
"""
import ctypes
import os
import re
import subprocess
import sys
import tempfile

import pytest

from helpers import ROOT
from iclsamd import _lib
from iclsamd import params as P

HEADER = os.path.join(ROOT, "include", "icls.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(icls_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = _declared()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared


def _c_layout():
    code = r"""
#include <stdio.h>
#include <stddef.h>
#include "icls.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(icls_params), offsetof(icls_params, g_unit),
         offsetof(icls_params, wall_speed), offsetof(icls_params, cfl), offsetof(icls_params, mass_trigger_exp),
         offsetof(icls_params, reserved), offsetof(icls_params, drop_xc), offsetof(icls_params, drop_r));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(code)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c])
        return [int(x) for x in subprocess.check_output([exe]).split()]


def test_params_struct_layout_matches_c():
    S = P.IclsParams
    ours = [ctypes.sizeof(S), S.g_unit.offset, S.wall_speed.offset, S.cfl.offset, S.mass_trigger_exp.offset,
            S.reserved.offset, S.drop_xc.offset, S.drop_r.offset]
    assert ours == _c_layout()


def test_calls_without_context_fail_cleanly():
    L = _lib.load()
    L.icls_finalize()
    assert L.icls_ab2() == 5          # ICLS_ERR_STATE
    assert L.icls_step(1, 0) == 5
    assert b"icls_init" in L.icls_last_error()


def test_bad_parameters_rejected_before_device_use():
    L = _lib.load()
    p = P.drop64()
    p.itot = 48                        # not a power of two
    assert L.icls_init(ctypes.byref(p)) == 1
    p = P.drop64()
    p.reinit_iter_max = 3
    assert L.icls_init(ctypes.byref(p)) == 7


def test_init_without_gpu_reports_nodevice():
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    L = _lib.load()
    p = P.drop64()
    assert L.icls_init(ctypes.byref(p)) == 6
    with pytest.raises(_lib.IclsError):
        _lib.Solver(p)


def test_fortran_driver_links_the_library():
    exe = os.path.join(ROOT, "icls-release_amd", "icls_fortran")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (make -C icls-release_amd fortran)")
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libicls.so" in out


def test_bench_traffic_table_names_profiled_kernels():
    """Every kernel the bench reports traffic for is a key of the committed
    profiles/traffic.json for both bench workloads (a renamed template would otherwise
    leave roofline.traffic empty)."""
    import json
    from iclsamd import roofline as R
    table = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    for workload, fp32 in (("bubble256_learned", True), ("bubble256", False)):
        for scope in R._TRAFFIC_KERNEL:
            key = R.traffic_key(scope, fp32)
            assert key in table[workload], (workload, scope, key)
            assert table[workload][key]["bytes_per_launch"] > 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with a message and a nonzero status (the
    solver is single-GPU), before touching the library or the device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
