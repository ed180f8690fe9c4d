"""Whole RK3 iterations whose last stage refreshes the next iteration's
primitives.

The reference refreshes the primitives at the top of every iteration
(con2prim, CCFVM/main.F90:62-67) from the state the previous iteration's
last stage wrote.  The device's last stage of a whole inviscid RK3 iteration
does that refresh itself from the state it has just computed (the compact
primitives, the temperature, the limiter's range partials and the positivity
note), so the next iteration starts without a pass of its own for it;
UFX_ITER_FUSION=0 selects the separate pass.  Checked here:
* both schedules give the same state, limiter values, gradients,
  primitives, residual, time steps, history and forces bit for bit, eagerly and from
  replayed graphs, over the fluxes, both precisions, both stage forms and
  the gradient methods that use it;
* a routine call between whole iterations sees the reference's current
  primitives (those the last iteration started from);
* a state that turns negative is reported at the reference's iteration and
  cell (check_positivity, CCFVM/main.F90:358-414), eagerly and from a graph.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _solver(mesh, monkeypatch, epi, form=None, **cfg):
    from unsflow2d_amd import cc
    monkeypatch.setenv("UFX_ITER_FUSION", "1" if epi else "0")
    s = cc.Solver(mesh, cc.Config(**cfg))
    monkeypatch.delenv("UFX_ITER_FUSION")
    if form:
        s.set_stage_form(form)
    s.keep_residual(True)
    return s


def _state(s):
    f = s.forces()
    return {"qc": s.get("qc"), "phi": s.get("phi"), "grad": s.get("grad"), "qp": s.get("qp"),
            "res": s.get("res"), "dt": s.get("dt"), "la": s.get("la"), "forces": np.array([f[k] for k in ("cl", "cd", "cl1", "cd1")])}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("flux,grad,form", [("roe", "lsqr", "gather_lds"), ("roe", "lsqr", "shared"),
                                            ("vanleer", "gg", "gather"), ("vanleer", "lsqr", "shared_lds"),
                                            ("roe", "gg", "gather_lds")])
def test_epilogue_equals_separate_refresh(flux, grad, form, precision, monkeypatch):
    from unsflow2d_amd import cc, meshgen
    print(f"case: prim epilogue vs pass, {flux} {grad} {form} {precision}, bump 40x16 seed 9",
          flush=True)
    mesh = cc.Mesh.from_raw(meshgen.bump_channel(40, 16, 9, h=0.08))
    cfg = dict(mach=0.65, flux=flux, grad=grad, precision=precision)
    a = _solver(mesh, monkeypatch, True, form, **cfg)
    b = _solver(mesh, monkeypatch, False, form, **cfg)
    ha = np.concatenate([a.iterate(6), a.iterate(5)])
    hb = np.concatenate([b.iterate(6), b.iterate(5)])
    assert np.all(np.isfinite(ha))
    assert np.array_equal(ha, hb)
    sa, sb = _state(a), _state(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    # graph replays (a new capture, replays, a shorter eager remainder)
    for n in (7, 7, 4):
        assert np.array_equal(a.iterate_graph(n), b.iterate_graph(n))
    sa, sb = _state(a), _state(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), f"after graphs: {k}"


def test_routine_after_iterations_sees_current_primitives(monkeypatch):
    from unsflow2d_amd import cc, meshgen
    from oracle.oracle import CCOracle
    print("case: routine after whole iterations, bump 30x12 seed 11", flush=True)
    mesh = cc.Mesh.from_raw(meshgen.bump_channel(30, 12, 11, h=0.08))
    a = _solver(mesh, monkeypatch, True, mach=0.55)
    b = _solver(mesh, monkeypatch, False, mach=0.55)
    orc = CCOracle(mesh.arrays(), mach=0.55)
    for s in (a, b):
        s.iterate(5)
    for it in range(1, 6):
        orc.iteration(it, 0.8)
    # the reference's qp after its loop: the last iteration's start primitives
    noc = a.counts["noc"]
    e = relerr(a.get("qp")[:noc], orc.qp[:noc])
    print(f"  qp vs oracle rel {e:.2e}", flush=True)
    assert e <= 1e-11
    # a gradient call in between: on the current primitives it recomputes
    # the gradients and limiter values the last iteration used
    a.gradient()
    b.gradient()
    assert np.array_equal(a.get("grad"), b.get("grad"))
    assert np.array_equal(a.get("phi"), b.get("phi"))
    ha, hb = a.iterate(4), b.iterate(4)
    assert np.array_equal(ha, hb)
    assert np.array_equal(a.get("qc"), b.get("qc"))


@pytest.mark.parametrize("graph", [False, True])
def test_positivity_stop_through_the_last_stage(graph, monkeypatch):
    """van Leer, RK3 at CFL 3 on the 30x12 channel: the primitives the
    oracle's iteration 11 starts from are negative at reference cell 412, so
    the reference stops at the end of iteration 11; the device too, with the
    note taken by iteration 10's last stage."""
    from unsflow2d_amd import cc, meshgen
    from unsflow2d_amd.cc import UfxError
    from oracle.oracle import CCOracle
    mesh = cc.Mesh.from_raw(meshgen.bump_channel(30, 12, 11, h=0.08))
    orc = CCOracle(mesh.arrays(), mach=0.55, flux="vanleer", cfl_max=3.0)
    stop = None
    for it in range(1, 30):
        orc.iteration(it, 3.0)
        bad = orc.check_positivity()
        if bad:
            stop = (it, bad)
            break
    assert stop is not None
    print(f"case: positivity stop via epilogue, oracle iteration {stop[0]} cell {stop[1]}, "
          f"graph={graph}", flush=True)
    s = _solver(mesh, monkeypatch, True, mach=0.55, flux="vanleer", cfl_min=3.0, cfl_max=3.0)
    if graph:
        s.iterate(1)
        s.iterate_graph(stop[0] - 2)
        with pytest.raises(UfxError) as ei:
            s.iterate_graph(1)
    else:
        s.iterate(stop[0] - 1)
        with pytest.raises(UfxError) as ei:
            s.iterate(1)
    assert "negative at point %d" % stop[1] in str(ei.value)
    assert s.error() == dict(kind="positivity", face=0, cell=stop[1])
