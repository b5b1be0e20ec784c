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


def fortran_coords(p, field):
    """Coordinates of the cell centres of a host array of ``field`` (Fortran bounds
    0, -2 or 1; cell (i,j,k) is centred at ((i-1/2) dx, (j-1/2) dy, (k-1/2) dz))."""
    dims = P.field_dims(p, field)
    lb = {P.SHAPE_C: 0, P.SHAPE_LS: -2, P.SHAPE_I: 1}[P.field_shape_kind(field)]
    dx, dy, dz = p.lx / p.itot, p.ly / p.jtot, p.lz / p.ktot
    i = (np.arange(dims[0]) + lb - 0.5) * dx
    j = (np.arange(dims[1]) + lb - 0.5) * dy
    k = (np.arange(dims[2]) + lb - 0.5) * dz
    return np.meshgrid(i, j, k, indexing="ij")
