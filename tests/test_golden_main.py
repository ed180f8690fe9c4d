"""The oracle against runs of the reference's own main program.

The fixtures tests/golden/ccmain_*.npz were written by
tests/golden/make_golden.py from CCFVM/main.F90 compiled unmodified (no
floating-point contraction): FLO.RES, the XYCpCf.dat that find_cl_cd writes
(CCFVM/main.F90:263-355, list-directed, so every double is kept to its
shortest round-trip digits), the printed force coefficients, and what the
reference prints when it STOPs -- in roe_flux on a non-positive Roe-averaged
sound speed (CCFVM/roe_flux.F90:73-87, the two states it received) and in
check_positivity (CCFVM/main.F90:384-411, the first cell).  The oracle must
reproduce them: the wall rows bit for bit, the printed coefficients to their
printed digits, the STOP at the same iteration with the same values.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT
from test_golden_cc import cfl_seq, oracle_for, raw_of

GOLD = os.path.join(ROOT, "tests", "golden")
FORCES = ["ccmain_forces_euler", "ccmain_forces_laminar", "ccmain_forces_isothermal"]


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    meta["deck"].setdefault("timemode", "rk3")
    return z, meta


def run_oracle(meta, niter, check_positivity=False):
    """The oracle through main's loop for `niter` iterations; stops early at
    an error.  Returns (oracle, arrays, iterations done, positivity cell)."""
    from unsflow2d_amd import cc
    mesh = cc.Mesh.from_raw(raw_of(meta))
    arrays = mesh.arrays()
    orc = oracle_for(meta, arrays)
    tm = meta["deck"]["timemode"]
    cfls = cfl_seq(meta, niter)
    bad = 0
    for it in range(1, niter + 1):
        try:
            orc.iteration(it, cfls[it - 1], tm)
        except RuntimeError:
            return orc, arrays, it, 0
        if check_positivity:
            bad = orc.check_positivity()
            if bad:
                return orc, arrays, it, bad
    return orc, arrays, niter, bad


@pytest.mark.parametrize("case", FORCES)
def test_find_cl_cd_matches_reference(case):
    """XYCpCf.dat bit for bit; Cl, Cd, Cl1, Cd1, Cm to the printed digits;
    the FLO.RES history to its printed digits and the first-iteration
    residue exactly."""
    z, meta = load(case)
    d = meta["deck"]
    orc, arrays, done, _ = run_oracle(meta, d["maxiter"])
    assert done == d["maxiter"]
    f = orc.find_cl_cd(d.get("xref", 0.0), d.get("yref", 0.0))
    ref = z["xycpcf"]
    assert ref.shape == (len(f["x"]), 4)
    for k, col in (("x", 0), ("y", 1), ("mcp", 2), ("cf", 3)):
        np.testing.assert_array_equal(f[k], ref[:, col], err_msg=k)
    if d.get("flow_type", "inviscid") != "inviscid":
        assert np.any(ref[:, 3] != 0.0)
    # f12.6 prints: the printed value is the oracle's rounded to 6 decimals
    half = 0.5e-6 + 1e-12
    cl, cd = z["printed_cl_cd"]
    assert abs(f["cl"] - cl) <= half and abs(f["cd"] - cd) <= half
    cl1, cd1 = z["printed_cl1_cd1"]
    assert abs(f["cl1"] - cl1) <= half and abs(f["cd1"] - cd1) <= half
    # the screen block prints the flux-based pair again, with Cm
    np.testing.assert_array_equal(z["printed_cl_cd_cm"][:2], z["printed_cl1_cd1"])
    assert abs(f["cm"] - z["printed_cl_cd_cm"][2]) <= half
    assert orc.s.fres1 == z["fres1"][0]


def test_force_sums_follow_from_the_reference_rows():
    """Cl, Cd and Cm recomputed with numpy from the reference's own XYCpCf
    rows (wall pressure p = p_inf + Cp q_inf, moment arm from the written
    midpoints) and the mesh's wall normals agree with the oracle's sums to
    rounding: a dropped term, a swapped component or a wrong sign in the
    angle of attack or the moment arm would not."""
    z, meta = load("ccmain_forces_euler")
    d = meta["deck"]
    orc, arrays, _, _ = run_oracle(meta, d["maxiter"])
    f = orc.find_cl_cd(d["xref"], d["yref"])
    nfi = orc.nfi
    wall = np.flatnonzero(arrays["fc_bc"][nfi:] == 1001) + nfi
    nx, ny, area = (arrays[k][wall] for k in ("fc_nx", "fc_ny", "fc_area"))
    m = d["mach"]
    p_inf = 1.0 / (1.4 * m * m)
    q_inf = 0.5 * 1.4 * p_inf * m * m
    cp = -z["xycpcf"][:, 2]
    p = p_inf + cp * q_inf
    fx, fy = np.sum(p * nx * area), np.sum(p * ny * area)
    a = math.radians(d["aoa_deg"])
    cl = (fy * math.cos(a) - fx * math.sin(a)) / q_inf
    cd = (fy * math.sin(a) + fx * math.cos(a)) / q_inf
    x, y = z["xycpcf"][:, 0], z["xycpcf"][:, 1]
    cm = np.sum(cp * nx * area * (y - d["yref"]) - cp * ny * area * (x - d["xref"]))
    assert abs(cl - f["cl"]) <= 1e-12 * abs(cl)
    assert abs(cd - f["cd"]) <= 1e-12 * abs(fx / q_inf) + 1e-14
    assert abs(cm - f["cm"]) <= 1e-12 * np.sum(np.abs(cp * area))
    # and the flux-based pair from the harness's Fx1, Fy1 the same way
    cl1 = (orc.s.fy1 * math.cos(a) - orc.s.fx1 * math.sin(a)) / q_inf
    assert abs(cl1 - f["cl1"]) <= 1e-13 * abs(cl1)


def test_flo_res_history_matches_reference():
    """log10 of the normalised residual per iteration, as FLO.RES prints it
    (e16.6), and the CFL."""
    z, meta = load("ccmain_forces_euler")
    d = meta["deck"]
    from unsflow2d_amd import cc
    orc = oracle_for(meta, cc.Mesh.from_raw(raw_of(meta)).arrays())
    cfls = cfl_seq(meta, d["maxiter"])
    hist = [math.log10(orc.iteration(it, cfls[it - 1])) for it in range(1, d["maxiter"] + 1)]
    fr = z["flo_res"]
    np.testing.assert_array_equal(fr[:, 0], np.arange(1, d["maxiter"] + 1))
    # e16.6 keeps 6 mantissa digits 0.dddddd: half a unit in the last one
    v = fr[:, 1]
    half = np.where(v == 0.0, 0.0, 5.0001e-6 * 10.0 ** np.floor(np.log10(np.abs(v) + 1e-300)))
    assert np.all(np.abs(np.array(hist) - v) <= half)


def test_roe_stop_matches_reference():
    """The reference STOPs in roe_flux at the first face (boundary faces
    first, then interior faces) whose Roe-averaged sound speed squared is not
    positive; the oracle fails in the same iteration on a face whose two
    states equal the printed ones bit for bit."""
    z, meta = load("ccmain_stop_roe")
    stop = z["stop_roe"]
    it_ref = int(stop[0])
    orc, arrays, it, _ = run_oracle(meta, it_ref + 5)
    assert it == it_ref
    e = orc.error_info
    assert e["kind"] == 1
    np.testing.assert_array_equal(e["ql"], stop[1:5])
    np.testing.assert_array_equal(e["qr"], stop[5:9])
    f = e["face"] - 1
    assert arrays["fc_in"][f] == e["cell"]
    assert len(z["flo_res"]) == it_ref - 1


def test_positivity_stop_matches_reference():
    """check_positivity runs after the update of every iteration
    (CCFVM/main.F90:113) on the primitives of the iteration's start; the
    reference STOPs at iteration 2 on the first non-positive cell."""
    z, meta = load("ccmain_stop_positivity")
    it_ref, cell_ref = (int(v) for v in z["stop_positivity"])
    orc, _, it, bad = run_oracle(meta, it_ref + 5, check_positivity=True)
    assert (it, bad) == (it_ref, cell_ref)
    # the first failing cell is the lowest of the non-positive ones
    qp = orc.qp[:orc.noc]
    first = np.flatnonzero((qp[:, 0] <= 0) | (qp[:, 3] <= 0))
    assert first[0] + 1 == cell_ref


def test_check_positivity_known_answers():
    """check_positivity on planted states: 1-based first interior cell whose
    density or pressure is <= 0 (zero counts, NaN does not, ghosts are not
    scanned), 0 when all pass."""
    from unsflow2d_amd import cc, meshgen
    from oracle.oracle import CCOracle
    mesh = cc.Mesh.from_raw(meshgen.bump_channel(8, 4, 3, h=0.05))
    orc = CCOracle(mesh.arrays())
    orc.prim()
    noc = orc.noc
    assert orc.check_positivity() == 0
    qp0 = orc.qp.copy()
    cases = [({17: (0, -1.0), 5: (3, -2.0)}, 6),     # density at 18, pressure at 6 -> 6
             ({9: (3, 0.0)}, 10),                     # zero pressure fails
             ({9: (0, 0.0)}, 10),                     # zero density fails
             ({2: (0, float("nan")), 40: (3, -1e-300)}, 41),  # NaN passes
             ({noc + 1: (0, -1.0)}, 0),               # a ghost is not scanned
             ({noc - 1: (3, -1.0)}, noc)]             # the last interior cell
    for plant, want in cases:
        orc.qp[:] = qp0
        for c, (k, v) in plant.items():
            orc.qp[c, k] = v
        assert orc.check_positivity() == want, plant
