"""The device path against the reference's own outputs (golden fixtures).

Same fixtures as test_golden_cc.py, where the oracle matches them bit for
bit; the device path (FMA contraction, per-cell gathers, tree reductions) is
held to 1e-12 normwise per routine and to the stated tolerances over whole
runs.  The Fortran host program is run on one case as well.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, relerr
from test_golden_cc import CASES, TINY, cfl_seq, load, raw_of

pytestmark = pytest.mark.gpu


def solver_for(meta, precision="fp64"):
    from unsflow2d_amd import cc
    d = meta["deck"]
    cfg = cc.Config(mach=d.get("mach", 0.5), aoa_deg=d.get("aoa_deg", 0.0),
                    reynolds=d.get("reynolds", 1000.0), t_wall=d.get("t_wall", 300.0),
                    flow_type=d.get("flow_type", "inviscid"), iwall=d.get("iwall", 1),
                    flux=d.get("flux_type", "roe"), grad=d.get("grad_type", "lsqr"),
                    ilimit=d.get("ilimit", 1), irs=d.get("irs", 0),
                    timemode=d.get("timemode", "rk3"), precision=precision,
                    cfl_type=d.get("cfl_type", "ramp"), cfl_min=d.get("cfl_min", 0.8),
                    cfl_max=d.get("cfl_max", 0.8), cfl_ramp_steps=d.get("cfl_ramp_steps", 1))
    mesh = cc.Mesh.from_raw(raw_of(meta))
    s = cc.Solver(mesh, cfg)
    if meta["phi0"] >= 0:
        s.set("phi", np.full((s.counts["noc"], 4), meta["phi0"]))
    return s


def _check(name, got, want, tol):
    e = relerr(got, want)
    print(f"  {name:14s} rel {e:.2e}", flush=True)
    assert e <= tol, f"{name}: {e:.3e} > {tol}"


@pytest.mark.parametrize("case", TINY)
def test_device_matches_reference_fixture(case):
    print(f"case: {case}", flush=True)
    z, meta = load(case)
    s = solver_for(meta)
    noc = s.counts["noc"]
    lusgs = meta["deck"].get("timemode") == "lusgs"
    cfls = cfl_seq(meta, meta["niter"])
    for it in range(1, meta["ndump"] + 1):
        t = lambda n, k=0: z[f"{n}_it{it:04d}_{k}"].reshape(-1)
        s.prim()
        _check("qp", s.get("qp")[:noc], t("qp").reshape(-1, 5)[:noc], 1e-12)
        s.avg_c2v()
        _check("pt_prim", s.get("pt_prim"), t("ptprim"), 1e-12)
        s.time_step(cfls[it - 1])
        _check("dt", s.get("dt"), t("dt"), 1e-12)
        s.save_old()
        s.gradient()
        _check("grad", s.get("grad"), t("grad"), 1e-12)
        _check("phi", s.get("phi"), t("phi"), 1e-12)
        for k in ((1,) if lusgs else (1, 2, 3)):
            if lusgs:
                s.lusgs()
            else:
                s.fvresidual()
            _check("res", s.get("res"), t("res", k), 1e-12)
            if not lusgs:
                s.rk_update(k)
            _check("qc", s.get("qc")[:noc], t("qc", k).reshape(-1, 4), 1e-12)
        s.residue(it)
    hist = z["hist"].reshape(-1, 8)
    rest = meta["niter"] - meta["ndump"]
    if rest > 0:
        s.iter, s.cfl = meta["ndump"], cfls[meta["ndump"] - 1]
        h = s.iterate(rest)
        _check("history", h, hist[meta["ndump"]:, 1], 1e-10)
    if "final_qc" in z.files:
        _check("final qc", s.get("qc"), z["final_qc"], 1e-10)


@pytest.mark.parametrize("case,tol", [("cc_c1_first_order", 1e-9), ("cc_c1_second_order", 1e-7)])
def test_device_config1_run(case, tol):
    """The configuration-1 workload (20 000 triangles), 1000 RK3 iterations:
    residual history and final state against the reference."""
    print(f"case: {case}", flush=True)
    z, meta = load(case)
    s = solver_for(meta)
    h = s.iterate(meta["niter"])
    ref = z["hist"].reshape(-1, 8)
    _check("history", h, ref[:, 1], tol)
    _check("final qc", s.get("qc"), z["final_qc"], tol)


def test_fortran_host_program(tmp_path):
    """flo2d_gpu (Fortran, over the ISO_C_BINDING interface) reproduces the
    reference's residual history on the fixture case."""
    exe = os.path.join(ROOT, "unsflow2d_amd", "fortran", "flo2d_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (amdflang absent at build time)")
    from unsflow2d_amd import meshgen
    z, meta = load("cc_tiny_roe_lsqr")
    raw = raw_of(meta)
    meshgen.write_geometry_inp(raw, str(tmp_path / "geometry.inp"))
    deck = meshgen.Deck(**dict(meta["deck"], maxiter=meta["niter"]))
    (tmp_path / "flo.inp").write_text(deck.text())
    print("case: fortran host, cc_tiny_roe_lsqr", flush=True)
    r = subprocess.run([exe, "flo.inp"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.loadtxt(tmp_path / "FLO.RES")
    ref = z["hist"].reshape(-1, 8)
    assert res.shape[0] == meta["niter"]
    _check("log10 fres", 10.0 ** res[:, 1], ref[:, 1], 1e-10)
