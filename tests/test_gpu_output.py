"""Output files of the GPU library against the CPU oracle: the legacy VTK file of
write_vtk (src/write_vtk.f90:214-432) and the Tecplot mid-yz plane of post2d
(src/output.f90:179-427), called through the C ABI and through the Fortran driver.

The VTK reader below follows the legacy VTK file-format specification (ASCII header
lines, big-endian binary payload per SCALARS block), not the writer's code.  The
numeric header fields are checked against strings derived by hand from the Fortran
edit descriptors the reference uses: Iw right-justifies an integer in w columns and
Ew.dE2 writes 0.d1d2..dd E+xx (d significant digits, mantissa in [0.1, 1)).
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, rel_linf, small_drop
from iclsamd import params as P

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _read_vtk(path):
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode("ascii")
        pos = e + 1
        return s

    head = [line() for _ in range(10)]
    npts = int(head[9].split()[1])
    blocks = {}
    while pos < len(data) - 1:
        sc = line()
        assert sc.startswith("SCALARS ") and sc.split()[2] == "double" and sc.split()[3] == "1", sc
        assert line() == "LOOKUP_TABLE default"
        blocks[sc.split()[1]] = np.frombuffer(data, dtype=">f8", count=npts, offset=pos).astype(np.float64)
        pos += 8 * npts
    assert data[pos:] == b"\n"
    return head, blocks


def _interior(a, lb):
    """(1:n,1:n,1:n) of a Fortran-order array whose first index is lb, x fastest."""
    s = tuple(slice(1 - lb, a.shape[d] - (1 - lb)) for d in range(3))
    return a[s].ravel(order="F")


def _state(n=32, nz=None, steps=4):
    import pyoracle
    from iclsamd import _lib
    p = small_drop(n, nz=nz, mass_correction=0)
    if nz is not None:
        p.lz = nz / n
    o = pyoracle.Oracle(p)
    g = _lib.Solver(p)
    o.init_program()
    g.init_program()
    for s in range(1, steps + 1):
        o.step(s, 0)
        g.step(s, 0)
    return p, o, g


def test_write_vtk_header_and_blocks(tmp_path):
    p, o, g = _state(32, nz=64)
    try:
        path = str(tmp_path / "para0000004.vtk")
        g.write_vtk(path)
        head, b = _read_vtk(path)
        # dx = dy = dz = 1/32: dx/2 = 0.015625, dx = 0.03125
        assert head == ["# vtk DataFile Version 3.0", "test file", "BINARY", "", "DATASET STRUCTURED_POINTS",
                        "DIMENSIONS          32          32          64",
                        "ORIGIN  0.156250E-01  0.156250E-01  0.156250E-01",
                        "SPACING  0.312500E-01  0.312500E-01  0.312500E-01", "",
                        "POINT_DATA       65536"]
        assert list(b) == ["U", "V", "W", "abs(Vel)", "V_r", "Pressure", "Interface"]
        u = g.download(P.F_UCC).ravel(order="F")
        v = g.download(P.F_VCC).ravel(order="F")
        w = g.download(P.F_WCC).ravel(order="F")
        assert np.array_equal(b["U"], u) and np.array_equal(b["V"], v) and np.array_equal(b["W"], w)
        assert np.array_equal(b["abs(Vel)"], np.sqrt(u * u + v * v + w * w))
        z = ((np.arange(1, 65) - 0.5) * (1.0 / 32) / p.lz)
        vu = np.repeat(6.0 * z * (1.0 - z), 32 * 32)
        assert np.allclose(b["V_r"], v - vu, rtol=0, atol=1e-15)
        assert np.array_equal(b["Pressure"], _interior(g.download(P.F_PNEW), 0))
        assert np.array_equal(b["Interface"], _interior(g.download(P.F_JIEMIAN), 0))
        # and the written values are the oracle's to the FP64 parity bound
        assert rel_linf(b["U"], o.get(P.F_UCC).ravel(order="F")) <= TOL
        assert rel_linf(b["W"], o.get(P.F_WCC).ravel(order="F")) <= TOL
        assert rel_linf(b["Pressure"], _interior(o.get(P.F_PNEW), 0)) <= TOL
        assert rel_linf(b["Interface"], _interior(o.get(P.F_JIEMIAN), 0)) <= TOL
    finally:
        g.close()


def test_post2d_columns(tmp_path):
    p, o, g = _state(32)
    try:
        path = str(tmp_path / "yz_mid0000004.dat")
        g.post2d(path)
        lines = open(path).read().split("\n")
        assert lines[0] == ' VARIABLES = "y","z","u","v","w","p","s","indie","fixhv"'
        assert lines[1] == ' ZONE T="Zone1" I= 32  J= 32 , F=POINT'
        rows = lines[2:-1]
        assert lines[-1] == "" and len(rows) == 32 * 32
        for r in rows[:64]:
            assert len(r) == 9 * 15
            for q in range(9):
                f = r[15 * q:15 * (q + 1)]
                m, e = f.strip().split("E")
                assert m.lstrip("-").startswith("0.") and len(m.lstrip("-")) == 9 and len(e) == 3, f
        got = np.array([[float(r[15 * q:15 * (q + 1)]) for q in range(9)] for r in rows]).reshape(32, 32, 9)
        # expected from the oracle's fields (arrays (0:i1,0:j1,0:k1) and (-2:..) for lvset)
        U, V, W, Pn = (o.get(f) for f in (P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW))
        L, J = o.get(P.F_LVSET), o.get(P.F_JIEMIAN)
        i, n = 16, 32
        want = np.zeros((n, n, 9))
        for k in range(1, n + 1):
            for j in range(1, n + 1):
                jm = j - 1 if j > 1 else n
                wkm = W[i, j, k - 1] if k > 1 else 0.0
                want[k - 1, j - 1] = [(j - 0.5) / n, (k - 0.5) / n, 0.5 * (U[i, j, k] + U[i - 1, j, k]),
                                      0.5 * (V[i, j, k] + V[i, jm, k]), 0.5 * (W[i, j, k] + wkm), Pn[i, j, k],
                                      L[i + 2, j + 2, k + 2], J[i, j, k], 0.0]
        # E15.7 keeps 7 significant digits with a mantissa in [0.1, 1): relative rounding
        # <= 5e-7 per value, on top of the GPU/oracle parity bound relative to each source
        # field's maximum (the u and v columns lie next to the drop's symmetry plane)
        fmax = np.array([0, 0, np.abs(U).max(), np.abs(V).max(), np.abs(W).max(), np.abs(Pn).max(),
                         np.abs(L).max(), np.abs(J).max(), 0])
        bound = 5.0001e-7 * np.abs(want) + TOL * fmax
        assert (np.abs(got - want) <= bound).all(), (np.abs(got - want) / np.maximum(bound, 1e-300)).max(axis=(0, 1))
    finally:
        g.close()


def test_fortran_driver_writes_output_files(tmp_path):
    """iout2d / iout3d / ioutfld of the Fortran main program (main.f90:262-277) write yz_mid, para
    .vtk and fld files into data/ under the reference's names."""
    exe = os.path.join(ROOT, "icls-release_amd", "icls_fortran")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    n = 32
    (tmp_path / "data").mkdir()
    nml = tmp_path / "run.nml"
    nml.write_text("&icls\n itot=%d, jtot=%d, ktot=%d, nstep=6, ioutchk=3, drop_r(1)=0.2,\n"
                   " mass_correction=0, iout2d=3, iout3d=6, ioutfld=6\n/\n" % (n, n, n))
    r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    names = sorted(os.listdir(tmp_path / "data"))
    assert names == ["fld0000006", "para0000006.vtk", "yz_mid0000003.dat", "yz_mid0000006.dat"], names
    import pyoracle
    o = pyoracle.Oracle(small_drop(n, mass_correction=0))
    o.init_program()
    o.run(6)
    head, b = _read_vtk(str(tmp_path / "data" / "para0000006.vtk"))
    assert rel_linf(b["Pressure"], _interior(o.get(P.F_PNEW), 0)) <= TOL
    assert rel_linf(b["V"], o.get(P.F_VCC).ravel(order="F")) <= TOL
    raw = np.fromfile(str(tmp_path / "data" / "fld0000006"), dtype=np.float64)
    assert raw[-3] == pytest.approx(6e-4, rel=1e-12) and raw[-2] == 6.0


def test_fortran_driver_restarts_from_fld(tmp_path):
    """begin > 0 in the Fortran main program: init_program reads data/fld<begin>
    (init.f90:44-51, loadflds(0, begin)) and re-derives the state; the restarted run's
    fld after 4 more steps against the oracle loaded from the same file."""
    exe = os.path.join(ROOT, "icls-release_amd", "icls_fortran")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    import shutil
    import pyoracle
    n = 32
    base = " itot=%d, jtot=%d, ktot=%d, ioutchk=2, drop_r(1)=0.2, mass_correction=0,\n" % (n, n, n)
    a, b = tmp_path / "a", tmp_path / "b"
    for d in (a, b):
        (d / "data").mkdir(parents=True)
    (a / "run.nml").write_text("&icls\n" + base + " nstep=4, ioutfld=4\n/\n")
    r = subprocess.run([exe, str(a / "run.nml")], capture_output=True, text=True, timeout=120, cwd=str(a))
    assert r.returncode == 0, r.stdout + r.stderr
    shutil.copy(a / "data" / "fld0000004", b / "data" / "fld0000004")
    (b / "run.nml").write_text("&icls\n" + base + " begin=4, nstep=8, ioutfld=8\n/\n")
    r = subprocess.run([exe, str(b / "run.nml")], capture_output=True, text=True, timeout=120, cwd=str(b))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Restarting from step" in r.stdout
    p = small_drop(n, mass_correction=0)
    cells = n * n * n
    raw = np.fromfile(str(a / "data" / "fld0000004"), dtype=np.float64)
    blocks = [raw[q * cells:(q + 1) * cells].reshape((n, n, n), order="F") for q in range(6)]
    o = pyoracle.Oracle(p)
    o.init_program()
    for q, f in enumerate([P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW]):
        x = np.zeros(P.field_dims(p, f), order="F")
        x[1:-1, 1:-1, 1:-1] = blocks[q]
        o.set(f, x)
    x = np.zeros(P.field_dims(p, P.F_LVSET), order="F")
    x[3:-3, 3:-3, 3:-3] = blocks[4]
    o.set(P.F_LVSET, x)
    o.set_scalar(P.S_TIME, raw[-3])
    o.init_derived()
    for s in range(5, 9):
        o.step(s, 4)
    out = np.fromfile(str(b / "data" / "fld0000008"), dtype=np.float64)
    assert out[-2] == 8.0
    # the restarted run keeps the time it read (loadd.f90:61) and adds one dt per step
    assert abs(out[-3] - (raw[-3] + 4 * p.timestep)) <= 1e-15 and raw[-3] > 0
    got = [out[q * cells:(q + 1) * cells].reshape((n, n, n), order="F") for q in range(5)]
    want = [o.get(f)[1:-1, 1:-1, 1:-1] for f in (P.F_UNEW, P.F_VNEW, P.F_WNEW, P.F_PNEW)]
    want.append(o.get(P.F_LVSET)[3:-3, 3:-3, 3:-3])
    for q in range(5):
        assert rel_linf(got[q], want[q]) <= TOL, q
