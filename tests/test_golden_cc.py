"""The oracle and the host preprocessing against the reference's own outputs.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py,
which runs the reference solver compiled from its own sources (no
floating-point contraction) on seeded meshes.  The C oracle (same expression
order, also without contraction) must reproduce every dumped array bit for
bit; so must the preprocessing of libufx (host code, no GPU needed).
"""
import glob
import json
import hashlib
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
CASES = sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLD, "cc_*.npz")))
TINY = [c for c in CASES if not c.startswith("cc_c1")]


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def raw_of(meta):
    from unsflow2d_amd import meshgen
    mk = dict(meta["mesh"])
    raw = meshgen.bump_channel(mk.pop("nx"), mk.pop("ny"), mk.pop("seed"), **mk)
    h = hashlib.sha256()
    for a in (raw.x, raw.y, raw.cen, raw.area, raw.fint, raw.fbnd, raw.fbn, raw.fbc):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == meta["raw_sha"], "mesh generator no longer reproduces the fixture mesh"
    return raw


def oracle_for(meta, arrays):
    from oracle.oracle import CCOracle
    d = meta["deck"]
    return CCOracle(arrays, mach=d.get("mach", 0.5), aoa_deg=d.get("aoa_deg", 0.0),
                    reynolds=d.get("reynolds", 1000.0), t_wall=d.get("t_wall", 300.0),
                    flow_type=d.get("flow_type", "inviscid"), iwall=d.get("iwall", 1),
                    flux=d.get("flux_type", "roe"), grad=d.get("grad_type", "lsqr"),
                    ilimit=d.get("ilimit", 1), irs=d.get("irs", 0), cfl_max=d.get("cfl_max", 0.8))


def cfl_seq(meta, n):
    from unsflow2d_amd import cc
    d = meta["deck"]
    c = cc.Config(cfl_type=d.get("cfl_type", "ramp"), cfl_min=d.get("cfl_min", 0.8),
                  cfl_max=d.get("cfl_max", 0.8), cfl_ramp_steps=d.get("cfl_ramp_steps", 1))
    out, prev = [], float("nan")
    for it in range(1, n + 1):
        prev = c.cfl(it, prev)
        out.append(prev)
    return out


def same(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("case", CASES)
def test_geometric_matches_reference(case):
    """ufx_cc_geometric reproduces the reference's preprocessed mesh."""
    from unsflow2d_amd import cc
    z, meta = load(case)
    mesh = cc.Mesh.from_raw(raw_of(meta))
    a = mesh.arrays()
    checked = 0
    for n in cc.MESH_I32 + cc.MESH_F64:
        if n in ("oldnum",):
            continue
        if n == "counts":
            continue
        if n in z.files:
            assert same(a[n], z[n]), n
            checked += 1
        elif "sha_" + n in z.files:
            dt = np.int32 if n in cc.MESH_I32 else np.float64
            dig = hashlib.sha256(np.ascontiguousarray(a[n], dtype=dt).tobytes()).digest()
            assert bytes(z["sha_" + n]) == dig, n
            checked += 1
    assert checked >= 25


@pytest.mark.parametrize("case", TINY)
def test_oracle_routines_match_reference(case):
    """Every dumped routine output of the first iterations, bit for bit."""
    from unsflow2d_amd import cc
    z, meta = load(case)
    mesh = cc.Mesh.from_raw(raw_of(meta))
    orc = oracle_for(meta, mesh.arrays())
    if meta["phi0"] >= 0:
        orc.phi[:] = meta["phi0"]
    lusgs = meta["deck"].get("timemode") == "lusgs"
    cfls = cfl_seq(meta, max(1, meta["ndump"]))
    noc = orc.noc
    for it in range(1, meta["ndump"] + 1):
        t = lambda n, k=0: z[f"{n}_it{it:04d}_{k}"]
        orc.prim()
        assert same(orc.qp, t("qp")), f"qp it{it}"
        orc.avg_c2v()
        assert same(orc.pprim, t("ptprim")), f"pt prim it{it}"
        assert same(orc.fqp, t("fcqp")), f"fc qp it{it}"
        orc.time_step(cfls[it - 1])
        assert same(orc.la, t("la")) and same(orc.dt, t("dt")) and same(orc.fla, t("fcla"))
        orc.save_old()
        orc.gradient()
        assert same(orc.grad, t("grad")), f"grad it{it}"
        assert same(orc.phi, t("phi")), f"phi it{it}"
        stages = (1,) if lusgs else (1, 2, 3)
        for k in stages:
            if lusgs:
                orc.lusgs()
            else:
                orc.fvresidual()
            assert same(orc.res, t("res", k)), f"res it{it} stage {k}"
            assert same(orc.qc[noc:], t("gqc", k)) and same(orc.qp[noc:], t("gqp", k))
            assert same(orc.forces, t("force", k)), f"forces it{it} stage {k}"
            if f"sgrad_it{it:04d}_{k}" in z.files:
                assert same(orc.grad, t("sgrad", k)) and same(orc.phi, t("sphi", k))
            if not lusgs:
                orc.rk_update(k)
            assert same(orc.qc[:noc], t("qc", k)), f"qc it{it} stage {k}"
        orc.residue(it)


@pytest.mark.parametrize("case", CASES)
def test_oracle_history_matches_reference(case):
    """Residual history (log10 of the normalised L2 norm each iteration) and
    the final state over the fixture's full run, bit for bit."""
    from unsflow2d_amd import cc
    z, meta = load(case)
    n = meta["niter"]
    if case.startswith("cc_c1") and os.environ.get("UFX_FULL_GOLDEN") != "1":
        n = 150   # the full 1000 iterations at 20 000 cells take ~20 s each
    mesh = cc.Mesh.from_raw(raw_of(meta))
    orc = oracle_for(meta, mesh.arrays())
    if meta["phi0"] >= 0:
        orc.phi[:] = meta["phi0"]
    hist = z["hist"].reshape(-1, 8)
    mode = "lusgs" if meta["deck"].get("timemode") == "lusgs" else "rk3"
    for it, cfl in enumerate(cfl_seq(meta, n), start=1):
        f = orc.iteration(it, cfl, mode)
        assert same(f, hist[it - 1, 1]), f"fres at iteration {it}"
        assert same(orc.forces, hist[it - 1, 4:6]), f"forces at iteration {it}"
    if n == meta["niter"] and "final_qc" in z.files:
        assert same(orc.qc, z["final_qc"])


def test_fixture_headline_numbers():
    """The configuration-1 fixtures carry the reference's residual drops."""
    z, meta = load("cc_c1_first_order")
    h = z["hist"].reshape(-1, 8)
    assert h.shape[0] == 1000 and np.all(np.isfinite(h[:, 1]))
    z2, _ = load("cc_c1_second_order")
    h2 = z2["hist"].reshape(-1, 8)
    # both converge (the normalised residual falls by more than an order)
    assert np.log10(h[-1, 1]) < -1.0 and np.log10(h2[-1, 1]) < -1.0
