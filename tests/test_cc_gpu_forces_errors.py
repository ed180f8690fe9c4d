"""Force integration and the device's STOP reports against the oracle and the
reference's own main program (fixtures tests/golden/ccmain_*, where the oracle
is pinned bit for bit, tests/test_golden_main.py).

* ufx_cc_forces (find_cl_cd, CCFVM/main.F90:263-355): the oracle's state of
  the start of the last iteration is injected, the device runs that
  iteration, and Cl, Cd, Cl1, Cd1, Cm, -Cp and Cf must agree with the
  oracle's to 1e-12; a whole run from the free stream is held to the
  reference's XYCpCf.dat rows.
* The error word: a negative Roe-averaged sound speed (CCFVM/roe_flux.F90:
  73-87) and a non-positive density or pressure (check_positivity,
  CCFVM/main.F90:358-414) end a device run at the reference's iteration,
  and ufx_cc_error names the face or cell the reference's sequential loops
  stop at.
"""
import numpy as np
import pytest

from conftest import relerr
from test_golden_cc import cfl_seq, oracle_for, raw_of
from test_golden_main import FORCES, load, run_oracle

pytestmark = pytest.mark.gpu


def device(meta, precision="fp64"):
    from unsflow2d_amd import cc
    d = meta["deck"]
    cfg = cc.Config(mach=d.get("mach", 0.5), aoa_deg=d.get("aoa_deg", 0.0),
                    reynolds=d.get("reynolds", 1000.0), t_wall=d.get("t_wall", 300.0),
                    flow_type=d.get("flow_type", "inviscid"), iwall=d.get("iwall", 1),
                    flux=d.get("flux_type", "roe"), grad=d.get("grad_type", "lsqr"),
                    ilimit=d.get("ilimit", 1), irs=d.get("irs", 0),
                    timemode=d.get("timemode", "rk3"), precision=precision,
                    cfl_type=d.get("cfl_type", "ramp"), cfl_min=d.get("cfl_min", 0.8),
                    cfl_max=d.get("cfl_max", 0.8), cfl_ramp_steps=d.get("cfl_ramp_steps", 1),
                    xref=d.get("xref", 0.0), yref=d.get("yref", 0.0))
    return cc.Solver(cc.Mesh.from_raw(raw_of(meta)), cfg)


def inject(s, orc, it_done, cfl):
    """Hand the oracle's state after `it_done` iterations to the device: the
    conserved state (ghosts too), the primitives (ghost ones feed the vertex
    averages), the limiter values (ratchet) and the cell and face viscosities
    of the last viscous pass (the time step reads the face ones,
    CCFVM/timestep.F90:147-171)."""
    s.set("qc", orc.qc)
    s.set("qp", orc.qp)
    s.set("phi", orc.phi)
    s.set("cell_mul", orc.cmul)
    s.set("fc_mu", orc.fmu)
    s.iter, s.cfl = it_done, cfl


def _rel(name, got, want, tol):
    e = relerr(got, want)
    print(f"  {name:6s} rel {e:.2e}", flush=True)
    assert e <= tol, f"{name}: {e:.3e} > {tol}"


@pytest.mark.parametrize("case", FORCES)
def test_forces_match_oracle(case):
    print(f"case: {case} (forces, injected last iteration)", flush=True)
    z, meta = load(case)
    d = meta["deck"]
    n = d["maxiter"]
    orc, _, done, _ = run_oracle(meta, n - 1)
    assert done == n - 1
    cfls = cfl_seq(meta, n)
    s = device(meta)
    inject(s, orc, n - 1, cfls[n - 2])
    s.iterate(1)
    orc.iteration(n, cfls[n - 1], d.get("timemode", "rk3"))
    want = orc.find_cl_cd(d.get("xref", 0.0), d.get("yref", 0.0))
    got = s.forces(wall=True)
    for k in ("cl", "cd", "cl1", "cd1", "cm"):
        _rel(k, got[k], want[k], 1e-12)
    _rel("-Cp", got["mcp"], want["mcp"], 1e-12)
    if d.get("flow_type", "inviscid") != "inviscid":
        _rel("Cf", got["cf"], want["cf"], 1e-12)
    else:
        assert np.all(got["cf"] == 0.0)


@pytest.mark.parametrize("case", FORCES)
def test_forces_whole_run_match_reference(case):
    """From the free stream through the reference's iterations, the wall rows
    of the reference's XYCpCf.dat and its printed coefficients."""
    print(f"case: {case} (whole run)", flush=True)
    z, meta = load(case)
    s = device(meta)
    s.iterate(meta["deck"]["maxiter"])
    got = s.forces(wall=True)
    ref = z["xycpcf"]
    _rel("-Cp", got["mcp"], ref[:, 2], 1e-10)
    if meta["deck"].get("flow_type", "inviscid") != "inviscid":
        _rel("Cf", got["cf"], ref[:, 3], 1e-9)
    half = 0.5e-6 + 1e-9
    assert abs(got["cl"] - z["printed_cl_cd"][0]) <= half
    assert abs(got["cd"] - z["printed_cl_cd"][1]) <= half
    assert abs(got["cl1"] - z["printed_cl1_cd1"][0]) <= half
    assert abs(got["cd1"] - z["printed_cl1_cd1"][1]) <= half
    assert abs(got["cm"] - z["printed_cl_cd_cm"][2]) <= half


def test_roe_stop_reported_at_reference_face():
    """M = 3 over the steep bump: the reference STOPs in roe_flux in
    iteration 7.  The device runs iterations 1-6, then reports the error in
    iteration 7 at the face whose states the reference printed (the oracle,
    bit-exact with those printed states, names it)."""
    from unsflow2d_amd.cc import UfxError
    z, meta = load("ccmain_stop_roe")
    it_ref = int(z["stop_roe"][0])
    orc, _, it, _ = run_oracle(meta, it_ref)
    assert it == it_ref
    want = orc.error_info
    print(f"case: ccmain_stop_roe, reference face {want['face']}, cell {want['cell']}",
          flush=True)
    s = device(meta)
    s.iterate(it_ref - 1)
    assert s.error()["kind"] is None
    with pytest.raises(UfxError) as ei:
        s.iterate(1)
    assert ei.value.code == 4
    assert "Sonic speed is negative" in str(ei.value)
    e = s.error()
    print(f"  device: {e}", flush=True)
    assert e == dict(kind="roe", face=want["face"], cell=want["cell"])
    # later calls keep reporting it and do no work, as after the reference's STOP
    with pytest.raises(UfxError):
        s.iterate(1)
    assert s.error() == e


def test_roe_stop_injected_state():
    """The same through the routine-by-routine entry points from the oracle's
    state of the start of iteration 7: the residual call that fails reports
    the reference's face."""
    from unsflow2d_amd.cc import UfxError
    z, meta = load("ccmain_stop_roe")
    it_ref = int(z["stop_roe"][0])
    orc, _, _, _ = run_oracle(meta, it_ref - 1)
    cfls = cfl_seq(meta, it_ref)
    s = device(meta)
    inject(s, orc, it_ref - 1, cfls[it_ref - 2])
    try:
        orc.iteration(it_ref, cfls[it_ref - 1])
    except RuntimeError:
        pass
    want = orc.error_info
    print(f"case: ccmain_stop_roe injected, face {want['face']}", flush=True)
    s.prim(); s.avg_c2v(); s.time_step(cfls[it_ref - 1]); s.save_old(); s.gradient()
    failed = False
    for k in (1, 2, 3):
        s.fvresidual()
        try:
            s.sync()
        except UfxError:
            failed = True
            break
        s.rk_update(k)
    assert failed
    assert s.error() == dict(kind="roe", face=want["face"], cell=want["cell"])


def test_positivity_stop_reported_at_reference_cell():
    """AUSM+-up under LU-SGS at CFL 50: after iteration 1 the state is
    negative in places; the reference's check_positivity STOPs in iteration
    2 at cell 70, after the iteration's update.  The device reports the same
    cell at the end of iteration 2."""
    from unsflow2d_amd.cc import UfxError
    z, meta = load("ccmain_stop_positivity")
    it_ref, cell_ref = (int(v) for v in z["stop_positivity"])
    print(f"case: ccmain_stop_positivity, reference cell {cell_ref}", flush=True)
    s = device(meta)
    s.iterate(it_ref - 1)
    with pytest.raises(UfxError) as ei:
        s.iterate(1)
    assert "negative at point %d" % cell_ref in str(ei.value)
    assert s.error() == dict(kind="positivity", face=0, cell=cell_ref)


def test_check_positivity_entry_point():
    """ufx_cc_check_positivity on the oracle's state after iteration 1 names
    the reference's cell, and on planted primitives the first interior cell
    with density or pressure <= 0 (NaN passes, ghosts are not scanned).  It
    only reports: the context's error word stays clear."""
    z, meta = load("ccmain_stop_positivity")
    it_ref, cell_ref = (int(v) for v in z["stop_positivity"])
    orc, _, _, _ = run_oracle(meta, it_ref - 1)
    s = device(meta)
    s.set("qc", orc.qc)
    s.prim()
    assert s.check_positivity() == cell_ref
    assert s.error()["kind"] is None
    qp = s.get("qp")
    noc = s.counts["noc"]
    qp[:noc, 0] = 1.0
    qp[:noc, 3] = 1.0
    for plant, want in [({17: (0, -1.0), 5: (3, -2.0)}, 6), ({9: (3, 0.0)}, 10),
                        ({2: (0, float("nan")), 40: (3, -1e-300)}, 41),
                        ({noc + 1: (0, -1.0)}, 0), ({}, 0)]:
        q = qp.copy()
        for c, (k, v) in plant.items():
            q[c, k] = v
        s.set("qp", q)
        assert s.check_positivity() == want, plant
    assert s.error()["kind"] is None


def test_qold_refused_after_fused_iterations():
    """Whole inviscid RK3 iterations keep their start state in qc and do not
    write qold; reading qold or a routine that uses it is refused until
    save_old writes it again."""
    from unsflow2d_amd.cc import UfxError
    z, meta = load("ccmain_forces_euler")
    s = device(meta)
    s.iterate(2)
    for call in (lambda: s.get("qold"), lambda: s.rk_update(2), lambda: s.rk_stage(1),
                 lambda: s.lusgs()):
        with pytest.raises(UfxError) as ei:
            call()
        assert ei.value.code == 4 and "save_old" in str(ei.value)
    s.save_old()
    np.testing.assert_array_equal(s.get("qold"), s.get("qc")[:s.counts["noc"]])
