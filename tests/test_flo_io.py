"""The reference's file formats on the host: decks, geometry.inp, FLO.RES,
the FLO.DAT restart, XYCpCf.dat and tecplt.dat, and the cell vertex walks
(CPU only: nothing here touches the device)."""
import os

import numpy as np
import pytest


def test_fortran_edit_descriptors():
    """Ew.d and Fw.d as a Fortran compiler writes them (values checked by
    hand against the standard's rules: 0.d...dE+xx mantissa, the E dropped
    for three exponent digits, asterisks when the field is too narrow)."""
    from unsflow2d_amd.flo_io import fortran_e, fortran_f
    assert fortran_e(1.0, 16, 6) == "    0.100000E+01"
    assert fortran_e(-1.2345678e-5, 16, 6) == "   -0.123457E-04"
    assert fortran_e(0.0, 16, 6) == "    0.000000E+00"
    assert fortran_e(9.9999996, 16, 6) == "    0.100000E+02"
    assert fortran_e(1e-120, 16, 6) == "    0.100000-119"
    assert fortran_e(3.14159265358979, 20, 10) == "    0.3141592654E+01"
    assert fortran_f(-3.0, 15, 6) == "      -3.000000"
    assert fortran_f(1e12, 15, 6) == "*" * 15


def test_flo_res_records():
    from unsflow2d_amd.flo_io import flo_res_lines
    s = flo_res_lines([1, 12], [0.0, -2.5], [0.8, 0.8])
    assert s.splitlines() == ["     1    0.000000E+00    0.800000E+00",
                              "    12   -0.250000E+01    0.800000E+00"]
    s = flo_res_lines([7], [-1.0], [50.0], log_fresi=[0.5])
    assert s == "     7   -0.100000E+01    0.500000E+00    0.500000E+02\n"


def test_restart_file_round_trip(tmp_path):
    """FLO.DAT keeps ten significant digits per value (4e20.10): a mantissa
    0.d1...d10 rounds to within 5e-10 relative."""
    from unsflow2d_amd.flo_io import read_flow, save_flow
    rng = np.random.default_rng(1)
    q = rng.uniform(-3, 3, (57, 4)) * 10.0 ** rng.integers(-5, 5, (57, 4))
    p = str(tmp_path / "FLO.DAT")
    save_flow(p, q)
    lines = open(p).read().splitlines()
    assert len(lines) == 57 and all(len(line) == 80 for line in lines)
    r = read_flow(p, 57)
    assert np.allclose(r, q, rtol=5.01e-10, atol=0)
    assert not np.array_equal(r, q)
    with pytest.raises(ValueError):
        read_flow(p, 58)


def test_cc_deck_round_trip(tmp_path):
    from unsflow2d_amd import flo_io, meshgen
    d = meshgen.Deck(flow_type="laminar", iwall=0, t_wall=288.5, mach=0.63, aoa_deg=2.5,
                     reynolds=5000.0, cfl_type="tanh", cfl_min=0.1, cfl_max=3.5,
                     cfl_ramp_steps=400, timemode="lusgs", grad_type="ggfc", maxiter=77,
                     minres=1e-9, flux_type="ausm", ilimit=0, irs=1, xref=0.25, yref=-0.5)
    p = str(tmp_path / "flo.inp")
    open(p, "w").write(d.text())
    r, extra = flo_io.read_cc_deck(p)
    assert r == d
    assert extra["gridfile"] == "geometry.inp" and extra["istart"] == 1
    bad = d.text().replace("ausm", "kfvs")
    open(p, "w").write(bad)
    with pytest.raises(ValueError, match="flux"):
        flo_io.read_cc_deck(p)


def test_nc_deck_round_trip(tmp_path):
    from unsflow2d_amd import flo_io, nc_mesh
    d = nc_mesh.NCDeck(mach=0.45, aoa_deg=1.25, cfl_max=25.0, timemode="lusgs", maxiter=321,
                       minres=1e-7, ilimit=1)
    p = str(tmp_path / "flo.inp")
    open(p, "w").write(d.text())
    r, extra = flo_io.read_nc_deck(p)
    assert r == d
    assert extra["gmres"] == (20, 1, 0.1)
    d2 = nc_mesh.NCDeck(mach=0.45, cfl_max=25.0, timemode="rk3")
    open(p, "w").write(d2.text())
    with pytest.raises(ValueError, match="less than 1"):
        flo_io.read_nc_deck(p)


def test_nc_geometry_round_trip(tmp_path):
    from unsflow2d_amd import nc_mesh
    m = nc_mesh.o_grid(24, 9, 2, stretch=1.2)
    p = str(tmp_path / "geometry.inp")
    nc_mesh.write_geometry_inp(m, p)
    r = nc_mesh.read_geometry_inp(p)
    for k in ("x", "y", "pbc", "cen", "area", "edge", "ebc"):
        assert np.array_equal(getattr(r, k), getattr(m, k)), k


def test_nc_cell_walks_are_the_triangles_counter_clockwise():
    """Every walk visits the three vertices of its triangle, starts on the
    cell's lowest-numbered edge, and goes round counter-clockwise (the
    centroid on the left of each step)."""
    from unsflow2d_amd import flo_io, nc, nc_mesh
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(30, 12, 4, stretch=1.2))
    a = m.arrays()
    t = flo_io.nc_cell_walks(a)
    x, y = a["pt_x"], a["pt_y"]
    i, j, k = t[:, 0] - 1, t[:, 1] - 1, t[:, 2] - 1
    area2 = (x[j] - x[i]) * (y[k] - y[i]) - (x[k] - x[i]) * (y[j] - y[i])
    assert np.all(area2 > 0)
    assert np.allclose(0.5 * area2, a["cell_cv"], rtol=1e-9)
    assert np.allclose((x[i] + x[j] + x[k]) / 3, a["cell_cenx"], rtol=1e-9, atol=1e-12)
    # the first step is the cell's lowest-numbered edge
    e = {}
    for n, (p1, p2) in enumerate(zip(a["fc_p1"], a["fc_p2"])):
        e.setdefault((min(p1, p2), max(p1, p2)), n)
    for c in range(0, len(t), 37):
        es = [e[(min(u, v), max(u, v))] for u, v in ((t[c, 0], t[c, 1]), (t[c, 1], t[c, 2]),
                                                     (t[c, 2], t[c, 0]))]
        assert es[0] == min(es)


def test_tecplot_writers(tmp_path):
    from unsflow2d_amd import flo_io
    x, y = np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])
    prim = np.array([[1.0, 0.5, 0.0, 2.0, 1.0]] * 3)
    p = str(tmp_path / "t.dat")
    flo_io.write_tecplot_cc(p, x, y, prim, np.zeros(3), [4], [1, 2, 3, 1], 2.0)
    lines = open(p).read().splitlines()
    assert lines[3].split() == ["N=", "3", ",E=", "1"]
    # 11 values per vertex in (1x,7(f15.6,2x)): seven, then four on the next line
    assert len(lines) == 4 + 2 * 3 + 1
    assert lines[4].startswith(" ") and len(lines[4]) == 1 + 7 * 17
    assert len(lines[5]) == 4 * 17
    assert lines[-1].split() == ["1", "2", "3", "3"]
    qp = np.array([[1.0, 0.5, 0.0, 2.0]] * 3)
    flo_io.write_tecplot_nc(p, x, y, qp, [[1, 2, 3]])
    lines = open(p).read().splitlines()
    assert "ET=TRIANGLE" in lines[2] and len(lines) == 4 + 3 + 1
    vals = [float(v) for v in lines[4].split()]
    assert vals[-1] == pytest.approx(np.sqrt(0.25 * 1.0 / (1.4 * 2.0)), abs=1e-6)
