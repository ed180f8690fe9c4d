"""The device path at the workloads' real sizes.

* The 2 M-triangle bump channel (1000 x 1000 quads, h = 0.08, M = 0.75,
  least-squares gradient, Venkatakrishnan limiter, Roe, RK3 at CFL 0.8): one
  iteration routine by routine against the oracle (the C restatement,
  bit-exact with the reference on every golden case), 1e-12 normwise per
  routine, and whole iterations against the routine-by-routine ones.
* The batched workload (256 bump channels of 5 000 - 200 000 cells from
  unsflow2d_amd.batch.cases): the four smallest cases batched in FP64
  against the oracle for 20 iterations; all 256 batched in mixed precision
  for one 50-iteration chunk, checked by properties (finite states, sampled
  cases equal to their own single-case graph replay bit for bit, batched
  norms equal to each case's own).
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _rel(name, got, want, tol):
    e = relerr(got, want)
    print(f"  {name:10s} rel {e:.2e}", flush=True)
    assert e <= tol, f"{name}: {e:.3e} > {tol}"


def test_bump2m_one_iteration_routines_match_oracle():
    from unsflow2d_amd import cc, meshgen
    from oracle.oracle import CCOracle
    print("case: bump channel 1000x1000 h 0.08 M 0.75 seed 2, one iteration", flush=True)
    raw = meshgen.bump_channel(1000, 1000, 2, h=0.08)
    mesh = cc.Mesh.from_raw(raw)
    cfg = cc.Config(mach=0.75)
    s = cc.Solver(mesh, cfg)
    orc = CCOracle(mesh.arrays(), mach=0.75)
    noc = s.counts["noc"]
    s.prim(); orc.prim()
    _rel("qp", s.get("qp")[:noc], orc.qp[:noc], 1e-12)
    s.avg_c2v(); orc.avg_c2v()
    _rel("pt_prim", s.get("pt_prim"), orc.pprim, 1e-12)
    s.time_step(0.8); orc.time_step(0.8)
    _rel("dt", s.get("dt"), orc.dt, 1e-12)
    _rel("fc_la", s.get("fc_la"), orc.fla, 1e-12)
    s.save_old(); orc.save_old()
    s.gradient(); orc.gradient()
    _rel("grad", s.get("grad"), orc.grad, 1e-12)
    _rel("phi", s.get("phi"), orc.phi, 1e-12)
    for k in (1, 2, 3):
        s.fvresidual(); orc.fvresidual()
        _rel(f"res {k}", s.get("res"), orc.res, 1e-12)
        _rel(f"ghost qc {k}", s.get("qc")[noc:], orc.qc[noc:], 1e-12)
        s.rk_update(k); orc.rk_update(k)
        _rel(f"qc {k}", s.get("qc")[:noc], orc.qc[:noc], 1e-12)
    f, _ = s.residue(1)
    orc.residue(1)
    assert f == 1.0 and orc.s.fres == 1.0
    # whole iterations (fused stages) from the same start equal the routines
    w = cc.Solver(mesh, cfg)
    w.iterate(1)
    _rel("whole qc", w.get("qc"), s.get("qc"), 1e-13)
    s.close(); w.close()


def _case_solver(c, precision):
    from unsflow2d_amd import cc, meshgen
    m = cc.Mesh.from_raw(meshgen.bump_channel(c.nx, c.ny, c.seed, h=c.h))
    cfg = cc.Config(mach=c.mach, aoa_deg=c.aoa_deg, flux="roe", grad="lsqr", ilimit=1,
                    precision=precision, cfl_type="ramp", cfl_min=0.8, cfl_max=0.8,
                    cfl_ramp_steps=1)
    s = cc.Solver(m, cfg)
    s.keep_residual(True)
    return m, s


def test_batched_smallest_cases_match_oracle():
    """The four smallest of the 256 batched cases (FP64), iteration 1 each
    alone and 19 more in one batched chunk: residual history and state against
    the oracle."""
    from unsflow2d_amd import batch, cc
    from oracle.oracle import CCOracle
    cs = sorted(batch.cases(256), key=lambda c: (c.cells, c.k))[:4]
    print("case: batched cases " + ", ".join(f"k={c.k} cells={c.cells} M={c.mach:.3f}"
                                             for c in cs), flush=True)
    runs = [_case_solver(c, "fp64") for c in cs]
    hist = []
    for m, s in runs:
        hist.append([s.iterate(1)[0]])
    fb = cc.Batch([s for _, s in runs])
    # the history of batched iterations lands in each case's own buffer
    fb.iterate(19)
    for (m, s), c, h1 in zip(runs, cs, hist):
        orc = CCOracle(m.arrays(), mach=c.mach, aoa_deg=c.aoa_deg)
        ref = np.array([orc.iteration(it, 0.8) for it in range(1, 21)])
        h = np.concatenate([h1, s.history(19)])
        _rel(f"hist k={c.k}", h, ref, 1e-11)
        _rel(f"qc k={c.k}", s.get("qc"), orc.qc, 1e-11)
    fb.close()
    for m, s in runs:
        s.close(); m.close()


def test_batched_256_cases_one_chunk():
    """All 256 cases (13.4 M cells, mixed precision): one 50-iteration chunk
    after iteration 1.  Every state is finite, eight sampled cases equal
    their own single-case graph replay bit for bit, and the batched density
    norms equal each case's own."""
    from unsflow2d_amd import batch, cc
    cs = batch.cases(256)
    print(f"case: 256 batched cases, {sum(c.cells for c in cs)} cells, mixed", flush=True)
    runs = [_case_solver(c, "mixed") for c in cs]
    for _, s in runs:
        s.iterate(1, hist=False)
    fb = cc.Batch([s for _, s in runs])
    fb.iterate(50)
    norms = fb.residual_norms(1)
    print("  batched chunk done", flush=True)
    for i, (_, s) in enumerate(runs):
        q = s.get("qc")
        assert np.all(np.isfinite(q)), f"case {i}"
        assert norms[i] == s.residual_norm(1), f"case {i}: batched norm"
    rng = np.random.default_rng(0)
    for i in sorted(rng.choice(256, 8, replace=False)):
        m, s = _case_solver(cs[i], "mixed")
        s.iterate(1, hist=False)
        s.iterate_graph(50, hist=False)
        for f in ("qc", "phi", "res"):
            assert np.array_equal(s.get(f), runs[i][1].get(f)), f"case {i}: {f}"
        s.close(); m.close()
    fb.close()
    for m, s in runs:
        s.close(); m.close()
