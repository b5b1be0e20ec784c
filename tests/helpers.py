"""Shared helpers of the test suite: oracle/GPU pairing and error norms."""
import os

import numpy as np

from iclsamd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel_linf(a, b):
    """max|a-b| / max(max|b|, 1e-300)"""
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den


def interior(a, shape_kind):
    """Interior view (1:imax,1:jmax,1:kmax) of a host array in Fortran layout."""
    if shape_kind == P.SHAPE_C:
        return a[1:-1, 1:-1, 1:-1]
    if shape_kind == P.SHAPE_LS:
        return a[3:-3, 3:-3, 3:-3]
    return a


def small_drop(n=32, nz=None, r=0.2, **kw):
    p = P.base_params(n, nz=nz)
    p.set_drops([(0.5, 0.5, 0.5 * p.lz, r)])
    for k, v in kw.items():
        setattr(p, k, v)
    return p
