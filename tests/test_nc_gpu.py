"""Device parity of the vertex-centred (median-dual) hot path.

The device kernels gather per vertex over its edges in ascending edge order,
which is the order in which the reference's edge loops scatter into that
vertex, and are compiled without a*b+c contraction, so every per-vertex
quantity (primitives, gradients, limiter values, residuals, spectral radii,
time steps, states) must equal the C oracle's -- and through it the
reference's (test_golden_nc.py) -- bit for bit.  Only the residual norm is a
parallel reduction in another order: 1e-12.  One operation is not exactly
reproducible: pow in farfield_flux1, where the device's and the host C
library's pow (neither correctly rounded in every case) can differ in the
last bit; on the 20 000-vertex case the first such difference comes at
iteration 221, after which the states agree to about 1e-13.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, relerr

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden")


def _check(name, got, want, tol):
    """tol 0: bit for bit."""
    e = relerr(got, want)
    print(f"  {name:14s} rel {e:.2e}", flush=True)
    if tol == 0:
        g, w = np.asarray(got, np.float64).ravel(), np.asarray(want, np.float64).ravel()
        nd = int(np.sum(g.view(np.int64) != w.view(np.int64)))
        assert nd == 0, f"{name}: {nd} of {g.size} values differ (max rel {e:.3e})"
    else:
        assert e <= tol, f"{name}: {e:.3e} > {tol}"


def _setup(ni, nj, seed, stretch, grad, timemode, cfl_max, mach=0.4, aoa=0.0, flux="vanleer",
           limiter="vkn"):
    from unsflow2d_amd import nc, nc_mesh
    from oracle.oracle import NCOracle
    raw = nc_mesh.o_grid(ni, nj, seed, stretch=stretch)
    m = nc.NCMesh.from_raw(raw)
    dev = nc.NCSolver(m, mach=mach, aoa_deg=aoa, cfl_max=cfl_max, grad=grad, timemode=timemode,
                      flux=flux, limiter=limiter)
    orc = NCOracle(m.arrays(), mach=mach, aoa_deg=aoa, cfl_max=cfl_max, grad=grad, flux=flux,
                   limiter=limiter)
    return dev, orc


@pytest.mark.parametrize("grad,timemode,flux,limiter", [
    ("gg", "lusgs", "vanleer", "vkn"), ("lsqr", "lusgs", "vanleer", "vkn"),
    ("gg", "rk3", "vanleer", "vkn"), ("lsqr", "rk3", "vanleer", "vkn"),
    ("gg", "rk3", "vanleer", "bj"), ("lsqr", "lusgs", "ausm", "vkn"),
    ("gg", "rk3", "ausm", "bj")])
def test_nc_routines_match_oracle(grad, timemode, flux, limiter):
    """Every routine from the same developed state: the oracle runs 20
    iterations, its state (qc and the limiter ratchet) is injected into the
    device, and both run two more iterations routine by routine.  (From a
    free-stream start the gradients are differences of nearly equal values,
    whose relative rounding is amplified by the conditioning.)"""
    print(f"case: nc routines {grad} {timemode} {flux} {limiter} o-grid 30x12 seed 4", flush=True)
    cfl_max = 50.0 if timemode == "lusgs" else 0.8
    dev, orc = _setup(30, 12, 4, 1.2, grad, timemode, cfl_max, flux=flux, limiter=limiter)
    for it in range(1, 21):
        orc.iteration(it, timemode)
    fres1 = orc.s.fres1
    dev.set("qc", orc.qc)
    dev.set("phi", orc.phi)
    for it in (21, 22):
        cfl = orc.cfl(it)
        dev.save_old(); orc.save_old()
        _check("qp", dev.get("qp"), orc.qp, 0)
        for k in ((1, 2, 3) if timemode == "rk3" else (1,)):
            if timemode == "rk3":
                dev.rk_stage(k, cfl)
                orc.fvresidual(cfl)
            else:
                dev.lusgs(cfl)
                orc.lusgs(cfl)
            _check("grad", dev.get("grad"), orc.grad, 0)
            _check("phi", dev.get("phi"), orc.phi, 0)
            _check("res", dev.get("res"), orc.res, 0)
            _check("la", dev.get("la"), orc.la, 0)
            _check("dt", dev.get("dt"), orc.dt, 0)
            _check("fc_la", dev.get("fc_la"), orc.fla, 0)
            if timemode == "rk3":
                orc.rk_update(k)
            _check("qc", dev.get("qc"), orc.qc, 0)
        h = dev.residue(it)
        orc.residue(it)
        _check("fres", h[[1, 2]], [orc.s.fres * fres1, orc.s.fresi], 1e-12)
        assert int(h[3]) == orc.s.iresi


def test_nc_first_iteration_from_free_stream():
    """Iteration 1 from the free stream, residual and update."""
    print("case: nc first iteration o-grid 30x12 seed 4", flush=True)
    dev, orc = _setup(30, 12, 4, 1.2, "gg", "lusgs", 50.0)
    cfl = orc.cfl(1)
    dev.save_old(); orc.save_old()
    dev.lusgs(cfl); orc.lusgs(cfl)
    _check("res", dev.get("res"), orc.res, 0)
    _check("qc", dev.get("qc"), orc.qc, 0)
    h = dev.residue(1)
    orc.residue(1)
    _check("fres", h[[0, 1, 2]], [orc.s.fres, orc.s.fres1, orc.s.fresi], 1e-12)


@pytest.mark.parametrize("grad,timemode,flux", [("gg", "lusgs", "vanleer"), ("lsqr", "rk3", "vanleer"),
                                                ("gg", "lusgs", "ausm")])
def test_nc_iterations_match_oracle(grad, timemode, flux):
    n = 80
    print(f"case: nc iterate {grad} {timemode} {flux} o-grid 40x16 seed 6 n={n}", flush=True)
    cfl_max = 50.0 if timemode == "lusgs" else 0.8
    dev, orc = _setup(40, 16, 6, 1.15, grad, timemode, cfl_max, mach=0.5, aoa=1.5, flux=flux)
    h = dev.iterate(n)
    ref = np.array([orc.iteration(it, timemode) for it in range(1, n + 1)])
    _check("history", h[:, 0], ref, 1e-12)
    _check("final qc", dev.get("qc"), orc.qc, 0)


def _load(case):
    z = np.load(os.path.join(GOLD, case + ".npz"))
    return z, json.loads(bytes(z["meta"]).decode())


@pytest.mark.parametrize("case,tol", [("nc_tiny_lusgs_gg", 0), ("nc_tiny_lusgs_lsqr", 0),
                                      ("nc_tiny_rk3_gg", 0), ("nc_tiny_rk3_lsqr", 0),
                                      ("nc_tiny_lusgs_gg_bj", 0), ("nc_tiny_rk3_lsqr_bj", 0),
                                      ("nc_c3_small", 1e-12)])
def test_nc_device_matches_reference_fixture(case, tol):
    """Per-routine outputs of the first iterations bit for bit against the
    reference's dumps, and the final state (bit for bit on the short cases;
    on the 500-iteration case the farfield pow differs in the last bit from
    iteration 221 on, so 1e-12); the residual history (a norm) to 1e-12."""
    from unsflow2d_amd import nc, nc_mesh
    print(f"case: {case}", flush=True)
    z, meta = _load(case)
    mkw = dict(meta["mesh"])
    raw = nc_mesh.o_grid(mkw.pop("ni"), mkw.pop("nj"), mkw.pop("seed"), **mkw)
    d = meta["deck"]
    grad, _, lim = meta["grad"].partition(":")
    dev = nc.NCSolver(nc.NCMesh.from_raw(raw), mach=d["mach"], aoa_deg=d.get("aoa_deg", 0.0),
                      cfl_max=d["cfl_max"], grad=grad, timemode=d["timemode"],
                      limiter=lim or "vkn")
    tm = d["timemode"]
    hist = z["hist"].reshape(-1, 8)
    nd = meta["ndump"]
    for it in range(1, nd + 1):
        t = lambda n, k: z[f"{n}_it{it:04d}_{k}"]
        cfl = dev.cfl(it)
        dev.save_old()
        _check("qp", dev.get("qp").ravel(), t("qp", 0), 0)
        for k in ((1, 2, 3) if tm == "rk3" else (1,)):
            if tm == "rk3":
                dev.rk_stage(k, cfl)
            else:
                dev.lusgs(cfl)
            _check("grad", dev.get("grad").ravel(), t("grad", k), 0)
            _check("res", dev.get("res").ravel(), t("res", k), 0)
            _check("phi", dev.get("phi"), t("phi", k).reshape(-1, 12)[:, :4], 0)
            _check("la", dev.get("la"), t("la", k), 0)
            _check("dt", dev.get("dt"), t("dt", k), 0)
            _check("fc_la", dev.get("fc_la"), t("fcla", k), 0)
            _check("qc", dev.get("qc").ravel(), t("qc", k), 0)
        h = dev.residue(it)
        _check("fres", h[0], hist[it - 1, 1], 1e-12)
    dev.iter = nd
    h = dev.iterate(meta["niter"] - nd)
    _check("history", h[:, 0], hist[nd:, 1], 1e-12)
    _check("final qc", dev.get("qc").ravel(), z["final_qc"], tol)


def test_nc_sweep_forms_agree(monkeypatch):
    """The four forms of the LU-SGS sweeps -- one launch per level, the same
    launches captured once in a graph and replayed every iteration (with the
    residual and norm kernels launched eagerly between replays), one
    workgroup walking all levels, and one cooperative launch of several
    workgroups with a grid barrier between levels, and one workgroup over
    level-ordered data -- give the same iterates, bit for bit, every
    iteration."""
    from unsflow2d_amd import nc, nc_mesh
    print("case: nc sweep forms o-grid 40x16 seed 6", flush=True)
    raw = nc_mesh.o_grid(40, 16, 6, stretch=1.15)
    m = nc.NCMesh.from_raw(raw)
    runs = {}
    for mode in ("levels", "graph", "single", "grid", "stream"):
        monkeypatch.setenv("UFX_NC_SWEEP", mode)
        s = nc.NCSolver(m, mach=0.5, aoa_deg=1.5, cfl_max=50.0, grad="gg", timemode="lusgs")
        hs, qs = [], []
        for _ in range(6):
            hs.append(s.iterate(1)[0, 0])
            qs.append(s.get("qc"))
        runs[mode] = (np.array(hs), qs)
    monkeypatch.delenv("UFX_NC_SWEEP")
    h0, q0 = runs["levels"]
    for mode in ("graph", "single", "grid", "stream"):
        h, q = runs[mode]
        for k in range(6):
            _check(f"{mode} qc it{k + 1}", q[k], q0[k], 0)
        _check(f"{mode} history", h, h0, 0)


@pytest.mark.parametrize("timemode,order,precision,grad", [
    ("lusgs", "multicolour", "fp64", "gg"), ("lusgs", "multicolour", "mixed", "gg"),
    ("lusgs", "reference", "fp64", "lsqr"), ("rk3", "reference", "fp64", "gg")])
def test_nc_storage_orders_and_residual_forms_agree(monkeypatch, timemode, order, precision, grad):
    """Vertices stored in the reference's breadth-first numbering
    (UFX_NC_ORDER=reference) or along a Hilbert curve (the default), and the
    residual as a per-vertex gather (UFX_NC_RESIDUAL=gather) or with each
    edge's flux evaluated once per block (the default): every per-vertex sum
    runs in the reference's order in all four, so states, limiter values,
    residuals and spectral radii agree bit for bit every iteration.  The
    residual norm is a parallel reduction over blocks of stored vertices,
    whose order follows the storage: 1e-13."""
    from unsflow2d_amd import nc, nc_mesh
    print(f"case: nc storage orders / residual forms, {timemode} {order} {precision} {grad}, "
          f"o-grid 40x16 seed 6", flush=True)
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(40, 16, 6, stretch=1.15))
    cfl = 50.0 if timemode == "lusgs" else 0.8
    runs = {}
    for store in ("reference", "hilbert"):
        for form in ("gather", "shared"):
            monkeypatch.setenv("UFX_NC_ORDER", store)
            monkeypatch.setenv("UFX_NC_RESIDUAL", form)
            s = nc.NCSolver(m, mach=0.5, aoa_deg=1.5, cfl_max=cfl, grad=grad, timemode=timemode,
                            lusgs_order=order, precision=precision)
            hs, out = [], []
            for _ in range(8):
                hs.append(s.iterate(1)[0, 0])
                out.append({k: s.get(k) for k in ("qc", "phi", "res", "dt", "fc_la")})
            runs[(store, form)] = (np.array(hs), out)
            s.close()
    monkeypatch.delenv("UFX_NC_ORDER")
    monkeypatch.delenv("UFX_NC_RESIDUAL")
    h0, o0 = runs[("reference", "gather")]
    for key, (h, o) in runs.items():
        for k in range(8):
            for name in o0[k]:
                _check(f"{key} {name} it{k + 1}", o[k][name], o0[k][name], 0)
        _check(f"{key} history", h, h0, 1e-13)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nc_serpentine_traversal_equals_forward(monkeypatch, precision):
    """The gradient pass and every other colour launch of the multicolour
    sweeps walk their blocks from the far end of each XCD's slice (so each
    starts on the data the pass before touched last); every per-vertex
    result is independent of the order in which blocks run, and the norm's
    partials are combined in vertex-block order, so everything equals the
    all-forward passes (UFX_SERPENTINE=0) bit for bit."""
    from unsflow2d_amd import nc, nc_mesh
    print(f"case: nc serpentine vs forward, multicolour {precision}, o-grid 40x16 seed 6",
          flush=True)
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(40, 16, 6, stretch=1.15))
    runs = {}
    for serp in ("1", "0"):
        monkeypatch.setenv("UFX_SERPENTINE", serp)
        s = nc.NCSolver(m, mach=0.5, aoa_deg=1.5, cfl_max=50.0, grad="gg", timemode="lusgs",
                        lusgs_order="multicolour", precision=precision)
        h = s.iterate(10)
        runs[serp] = (h, {k: s.get(k) for k in ("qc", "phi", "res", "dt", "fc_la")})
        s.close()
    monkeypatch.delenv("UFX_SERPENTINE")
    (h1, a1), (h0, a0) = runs["1"], runs["0"]
    for k in a0:
        _check(k, a1[k], a0[k], 0)
    _check("history", h1, h0, 0)


@pytest.mark.parametrize("flux", ["vanleer", "ausm"])
def test_nc_wall_forces_match_oracle(flux):
    """The wall force (the wall closure of solid_flux summed in edge order)
    and -Cp per wall edge of a developed state, bit for bit with the
    oracle's sequential sum."""
    print(f"case: nc forces {flux} o-grid 40x16 seed 6", flush=True)
    dev, orc = _setup(40, 16, 6, 1.15, "gg", "lusgs", 50.0, mach=0.5, aoa=1.5, flux=flux)
    f0, _ = dev.forces()
    g0, _ = orc.forces()
    _check("free-stream forces", f0, g0, 0)
    dev.iterate(60)
    for it in range(1, 61):
        orc.iteration(it, "lusgs")
    f, cp = dev.forces()
    g, cq = orc.forces()
    print(f"  cl {f[0]:.6f} cd {f[1]:.6f}", flush=True)
    _check("cl cd fx fy", f, g, 0)
    _check("-cp", cp, cq, 0)


def test_stream_sweep_beyond_the_lds_ring(monkeypatch):
    """On a 180 000-vertex O-grid some swept neighbours lie more than the
    LDS ring's 5 120 positions back (0.21 % of the forward and 0.26 % of the
    backward reads; test_nc_sweep_schedule.py counts them on the host), so
    the streamed sweep reads those from HBM and the rest from the ring; its
    iterates equal the level-launch form's (per-vertex records, no ring) bit
    for bit."""
    from unsflow2d_amd import nc, nc_mesh
    print("case: nc stream sweep vs levels, o-grid 600x300 seed 7", flush=True)
    raw = nc_mesh.o_grid(600, 300, 7, stretch=1.005 ** (999 / 299))
    m = nc.NCMesh.from_raw(raw)
    qs = {}
    for mode in ("levels", "stream"):
        monkeypatch.setenv("UFX_NC_SWEEP", mode)
        s = nc.NCSolver(m, mach=0.4, cfl_max=50.0, grad="gg", timemode="lusgs")
        s.iterate(3, hist=False)
        qs[mode] = s.get("qc")
        s.close()
    monkeypatch.delenv("UFX_NC_SWEEP")
    _check("stream vs levels qc", qs["stream"], qs["levels"], 0)
