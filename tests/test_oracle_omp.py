"""The OpenMP build of the oracle (liboracle_omp.so, the CPU baseline on all host
cores) against the serial build: every field and scalar bit for bit after a few
steps, on cases that reach the parallel loops of every stage (band advection with
HOUC5 and WENO5, re-initialisation, curvature, inflation, several level sets with
ipot, momentum, GFM, the solve, correc and the pressure update)."""
import os

import numpy as np
import pytest

import pyoracle
from helpers import small_drop
from iclsamd import params as P

FIELDS = [f for f in range(len(P.FIELD_NAMES)) if f != P.F_LVSET]


def _case(name):
    if name == "drop_houc5":
        return small_drop(32)
    p = small_drop(32, nz=16)
    p.lz = 0.5
    p.ls_adv_space = P.ADV_WENO5
    p.mass_correction = 1
    p.inflate_step = 1
    p.mass_trigger_exp = 12   # inflation every step
    p.set_drops([(0.3, 0.4, 0.25, 0.1), (0.7, 0.6, 0.22, 0.08), (0.5, 0.2, 0.25, 0.09)])
    return p


@pytest.mark.parametrize("name", ["drop_houc5", "three_drops_weno5_inflation"])
def test_omp_build_matches_serial(name):
    if (os.cpu_count() or 1) < 2:
        pytest.skip("one host core: the OpenMP build runs serially")
    p = _case(name)
    a, b = pyoracle.Oracle(p), pyoracle.Oracle(p, omp=True)
    try:
        a.init_program()
        b.init_program()
        for s in range(1, 5):
            a.step(s, 0)
            b.step(s, 0)
        for f in FIELDS:
            assert np.array_equal(a.get(f), b.get(f)), P.FIELD_NAMES[f]
        for l in range(p.lmax):
            assert np.array_equal(a.get(P.F_LVSET, l), b.get(P.F_LVSET, l)), l
            assert a.volume(l) == b.volume(l), l
        for s in range(7):
            assert a.scalar(s) == b.scalar(s) or (np.isnan(a.scalar(s)) and np.isnan(b.scalar(s))), s
    finally:
        a.close()
        b.close()
