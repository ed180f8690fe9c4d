#!/usr/bin/env python3
"""Write the golden fixtures of the hot paths from the reference.

Runs the reference solvers built from their own sources (oracle/_ref/cc and
oracle/_ref/nc, no floating-point contraction; `make -C oracle ref`) through
the harness drivers oracle/ref_cc_driver.F90 and oracle/ref_nc_driver.F90 on
seeded meshes from unsflow2d_amd.meshgen / unsflow2d_amd.nc_mesh, and stores
what they dump -- the preprocessed mesh, every routine's output for the first
iterations, the residual history and the final state -- in one .npz per case
under tests/golden/.  Needs /root/reference (this container only); the
fixtures themselves are data and travel with the repository.

The vertex-centred reference reads freshly allocated memory before writing
it (NCFVM/geometric.f90:101-104, :906-951); its driver runs with
MALLOC_MMAP_THRESHOLD_=0 so that memory is zero, the meaning the code
intends.  Its preprocessing is checked against the library's restatement
(bit for bit) in tests/test_golden_nc.py.

    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from unsflow2d_amd import meshgen, nc_mesh  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "cc", "ref_cc_driver")
NC_DRIVER = os.path.join(ROOT, "oracle", "_ref", "nc", "ref_nc_driver")

# name -> (mesh kwargs, deck kwargs, niter, ndump, phi0, keep_mesh, keep_final)
TINY = dict(nx=12, ny=6, seed=7, h=0.08)
VISC = dict(nx=16, ny=10, seed=4, h=0.05, first=2e-3, growth=1.3, jitter_y=False)
CASES = {
    "cc_tiny_roe_lsqr": (TINY, dict(mach=0.5), 30, 3, -1.0, True, True),
    "cc_tiny_first_order": (TINY, dict(mach=0.5), 30, 2, 0.0, False, True),
    "cc_tiny_vanleer": (TINY, dict(mach=0.5, flux_type="vanleer"), 10, 2, -1.0, False, True),
    "cc_tiny_rusanov": (TINY, dict(mach=0.5, flux_type="rusanov"), 2, 2, -1.0, False, True),
    "cc_tiny_ausm": (TINY, dict(mach=0.5, flux_type="ausm", cfl_min=0.05, cfl_max=0.05),
                     10, 2, -1.0, False, True),
    "cc_tiny_gg": (TINY, dict(mach=0.5, grad_type="gg"), 10, 2, -1.0, False, True),
    # quadrilateral cells, and quadrilaterals mixed with triangles
    "cc_tiny_quad": (dict(TINY, cells="quad"), dict(mach=0.5), 10, 2, -1.0, True, True),
    "cc_tiny_mixed_cells": (dict(TINY, cells="mixed"), dict(mach=0.5, grad_type="gg"), 10, 2, -1.0,
                            True, True),
    "cc_tiny_ggfc": (TINY, dict(mach=0.5, grad_type="ggfc"), 10, 2, -1.0, False, True),
    "cc_visc_ggfc_irs": (VISC, dict(mach=0.3, flow_type="laminar", reynolds=1000.0, iwall=1,
                                    grad_type="ggfc", irs=1, cfl_min=0.5, cfl_max=0.8,
                                    cfl_ramp_steps=5), 10, 2, -1.0, True, True),
    "cc_visc_lsqr_iso": (VISC, dict(mach=0.3, flow_type="laminar", reynolds=1000.0, iwall=0,
                                    grad_type="lsqr", irs=0, cfl_min=0.5, cfl_max=0.8,
                                    cfl_ramp_steps=5), 10, 2, -1.0, False, True),
    "cc_lusgs_gg": (dict(nx=20, ny=10, seed=8, h=0.08),
                    dict(mach=0.5, grad_type="gg", timemode="lusgs", cfl_min=1.0, cfl_max=5.0,
                         cfl_ramp_steps=10), 15, 2, -1.0, True, True),
    # the configuration-1 workload: 100 x 100 quads, first order (limiter values
    # start at 0), and the same with the second-order scheme
    "cc_c1_first_order": (dict(nx=100, ny=100, seed=1, h=0.05), dict(mach=0.5), 1000, 0, 0.0,
                          False, ("final_qc",)),
    "cc_c1_second_order": (dict(nx=100, ny=100, seed=1, h=0.05), dict(mach=0.5), 1000, 0, -1.0,
                           False, ("final_qc",)),
}

# runs of the reference's own main program (CCFVM/main.F90, unmodified):
# name -> (mesh kwargs, deck kwargs).  With ScrInt = MaxIter the last
# iteration calls screen -> find_cl_cd, which writes XYCpCf.dat (wall-face
# midpoint, -Cp, Cf; list-directed, i.e. every double to its shortest
# round-trip digits) and prints the force coefficients; the "stop" cases end
# at one of the reference's STOPs (Roe's non-positive sound speed,
# check_positivity), whose printed values are kept.
STEEP = dict(nx=12, ny=6, seed=7, h=0.2)
MAIN_CASES = {
    "ccmain_forces_euler": (TINY, dict(mach=0.5, aoa_deg=1.5, xref=0.5, yref=0.1, maxiter=30,
                                       scrint=30)),
    "ccmain_forces_laminar": (VISC, dict(mach=0.3, aoa_deg=1.0, flow_type="laminar",
                                         reynolds=1000.0, iwall=1, grad_type="ggfc", irs=1,
                                         cfl_min=0.5, cfl_max=0.8, cfl_ramp_steps=5, xref=0.3,
                                         yref=-0.05, maxiter=10, scrint=10)),
    "ccmain_forces_isothermal": (VISC, dict(mach=0.3, flow_type="laminar", reynolds=1000.0,
                                            iwall=0, grad_type="lsqr", cfl_min=0.5, cfl_max=0.8,
                                            cfl_ramp_steps=5, maxiter=10, scrint=10)),
    # M = 3 over a steep bump: a reconstructed state with negative pressure
    # makes Roe's averaged sound speed imaginary (CCFVM/roe_flux.F90:73-87)
    "ccmain_stop_roe": (STEEP, dict(mach=3.0, cfl_min=0.99, cfl_max=0.99, maxiter=300)),
    # AUSM+-up under LU-SGS at CFL 50 leaves negative states after iteration 1
    # (CCFVM/main.F90:384-411)
    "ccmain_stop_positivity": (STEEP, dict(mach=0.8, flux_type="ausm", timemode="lusgs",
                                           cfl_min=50.0, cfl_max=50.0, maxiter=300)),
}
MAIN = os.path.join(ROOT, "oracle", "_ref", "cc", "ref_main")

# vertex-centred cases: name -> (o_grid kwargs, deck kwargs, gradient, niter,
# ndump, keep_mesh)
NC_TINY = dict(ni=24, nj=10, seed=3, stretch=1.2)
NC_CASES = {
    "nc_tiny_lusgs_gg": (NC_TINY, dict(mach=0.4, cfl_max=50.0, timemode="lusgs"), "gg", 100, 2,
                         True),
    "nc_tiny_lusgs_lsqr": (NC_TINY, dict(mach=0.4, cfl_max=50.0, timemode="lusgs"), "lsqr", 60,
                           2, False),
    "nc_tiny_rk3_gg": (NC_TINY, dict(mach=0.4, cfl_max=0.8, timemode="rk3"), "gg", 60, 2, False),
    "nc_tiny_rk3_lsqr": (dict(ni=20, nj=8, seed=5, stretch=1.25),
                         dict(mach=0.5, aoa_deg=2.0, cfl_max=0.8, timemode="rk3"), "lsqr", 40, 2,
                         True),
    # limiter_min_max (Barth-Jespersen) in place of limiter_VKN; under LU-SGS
    # at CFL 50 the reference's own solution turns NaN at iteration 9 (its
    # phi can go negative: dumax starts at -eps), so the case stops at 8
    "nc_tiny_lusgs_gg_bj": (NC_TINY, dict(mach=0.4, cfl_max=50.0, timemode="lusgs"), "gg:bj",
                            8, 2, False),
    "nc_tiny_rk3_lsqr_bj": (dict(ni=20, nj=8, seed=5, stretch=1.25),
                            dict(mach=0.5, aoa_deg=2.0, cfl_max=0.8, timemode="rk3"), "lsqr:bj",
                            30, 2, False),
    # the vertex-centred workload's generator at 200 x 100 points (20 000
    # vertices), LU-SGS to CFL 50, 500 iterations
    "nc_c3_small": (dict(ni=200, nj=100, seed=3, stretch=1.005 ** (999 / 99)),
                    dict(mach=0.4, cfl_max=50.0, timemode="lusgs"), "gg", 500, 0, False),
}

NC_MESH_I = ["fc_p1", "fc_p2", "fc_in", "fc_out", "fc_bc", "pt_nv2v", "pt_v2v", "pt_nv2f",
             "pt_v2f", "counts"]
NC_MESH_F = ["fc_nx", "fc_ny", "fc_area", "pt_x", "pt_y", "pt_cv", "pt_wxy"]

MESH_I = ["fc_p1", "fc_p2", "fc_in", "fc_out", "fc_bc", "cell_nc2c", "cell_nc2f", "cell_c2f",
          "pt_nv2c", "pt_v2c", "counts"]
MESH_F = ["fc_nx", "fc_ny", "fc_area", "fc_cenx", "fc_ceny", "fc_cov", "cell_cenx", "cell_ceny",
          "cell_cv", "cell_cov", "cell_r11", "cell_r12", "cell_r22", "pt_x", "pt_y", "pt_wt"]


def raw_sha(raw: meshgen.RawMesh) -> str:
    h = hashlib.sha256()
    for a in (raw.x, raw.y, raw.cen, raw.area, raw.fint, raw.fbnd, raw.fbn, raw.fbc):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def make(name: str) -> str:
    mkw, dkw, niter, ndump, phi0, keep_mesh, keep_final = CASES[name]
    mkw = dict(mkw)
    raw = meshgen.bump_channel(mkw.pop("nx"), mkw.pop("ny"), mkw.pop("seed"), **mkw)
    deck = meshgen.Deck(**dkw)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        meshgen.write_geometry_inp(raw, os.path.join(tmp, "geometry.inp"))
        with open(os.path.join(tmp, "flo.inp"), "w") as f:
            f.write(deck.text())
        os.makedirs(os.path.join(tmp, "d"))
        r = subprocess.run([DRIVER, "flo.inp", "d", str(niter), str(ndump), repr(phi0)], cwd=tmp,
                           capture_output=True, text=True, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        ddir = os.path.join(tmp, "d")
        for fn in sorted(os.listdir(ddir)):
            base, ext = os.path.splitext(fn)
            dt = np.int32 if ext == ".i32" else np.float64
            a = np.fromfile(os.path.join(ddir, fn), dtype=dt)
            is_mesh = base in MESH_I or base in MESH_F
            if is_mesh and not keep_mesh:
                out["sha_" + base] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), np.uint8)
                continue
            if base.startswith("final_") and not (keep_final is True or base in keep_final):
                continue
            out[base] = a
    meta = dict(mesh=CASES[name][0], deck=dkw, niter=niter, ndump=ndump, phi0=phi0,
                raw_sha=raw_sha(raw), reference="CCFVM, amdflang -O2 -ffp-contract=off",
                driver="oracle/ref_cc_driver.F90")
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def _floats(text: str) -> list:
    return [float(t) for t in text.split()]


def make_main(name: str) -> str:
    """Run the reference's unmodified main program and keep what it writes:
    FLO.RES (iteration, log10 of the normalised residual, CFL), XYCpCf.dat,
    and from the screen output the first-iteration residue, the printed force
    coefficients and, for a STOP, the iteration it happened in with the
    values printed there."""
    mkw, dkw = MAIN_CASES[name]
    mkw = dict(mkw)
    raw = meshgen.bump_channel(mkw.pop("nx"), mkw.pop("ny"), mkw.pop("seed"), **mkw)
    deck = meshgen.Deck(**dkw)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        meshgen.write_geometry_inp(raw, os.path.join(tmp, "geometry.inp"))
        with open(os.path.join(tmp, "flo.inp"), "w") as f:
            f.write(deck.text())
        r = subprocess.run([MAIN, "flo.inp"], cwd=tmp, capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        lines = r.stdout.splitlines()
        res = open(os.path.join(tmp, "FLO.RES")).read().split()
        out["flo_res"] = np.array([float(t) for t in res]).reshape(-1, 3)
        for ln in lines:
            if "Residue in first iteration" in ln:
                out["fres1"] = np.array([float(ln.split("=")[1])])
            elif ln.strip().startswith("Cl, Cd, Cm"):
                out["printed_cl_cd_cm"] = np.array(_floats(ln.split("=")[1]))
            elif ln.strip().startswith("Cl1, Cd1"):
                out["printed_cl1_cd1"] = np.array(_floats(ln.split("=")[1]))
            elif ln.strip().startswith("Cl, Cd"):
                out["printed_cl_cd"] = np.array(_floats(ln.split("=")[1]))
        xy = os.path.join(tmp, "XYCpCf.dat")
        if os.path.exists(xy):
            out["xycpcf"] = np.array(_floats(open(xy).read())).reshape(-1, 4)
        text = r.stdout
        stop_it = len(out["flo_res"]) + 1
        if "Sonic speed is negative" in text:
            seg = text.split("Left/right cell values")[1].split("Left/right vertex values")[0]
            v = _floats(seg)
            assert len(v) == 8
            out["stop_roe"] = np.array([stop_it] + v)
        if "Density/pressure is negative at point" in text:
            cell = int(text.split("Density/pressure is negative at point")[1].split()[0])
            out["stop_positivity"] = np.array([stop_it, cell], dtype=np.int64)
    meta = dict(mesh=MAIN_CASES[name][0], deck=dkw, raw_sha=raw_sha(raw),
                reference="CCFVM main program unmodified, amdflang -O2 -ffp-contract=off")
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def nc_raw_sha(m: nc_mesh.NCRawMesh) -> str:
    h = hashlib.sha256()
    for a in (m.x, m.y, m.pbc, m.cen, m.area, m.edge, m.ebc):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def make_nc(name: str) -> str:
    mkw, dkw, grad, niter, ndump, keep_mesh = NC_CASES[name]
    mkw = dict(mkw)
    m = nc_mesh.o_grid(mkw.pop("ni"), mkw.pop("nj"), mkw.pop("seed"), **mkw)
    deck = nc_mesh.NCDeck(**dkw)
    out = {}
    env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="0")
    with tempfile.TemporaryDirectory() as tmp:
        nc_mesh.write_geometry_inp(m, os.path.join(tmp, "geometry.inp"))
        with open(os.path.join(tmp, "flo.inp"), "w") as f:
            f.write(deck.text())
        os.makedirs(os.path.join(tmp, "d"))
        r = subprocess.run([NC_DRIVER, "flo.inp", "d", str(niter), str(ndump), grad], cwd=tmp,
                           capture_output=True, text=True, timeout=3600, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        ddir = os.path.join(tmp, "d")
        for fn in sorted(os.listdir(ddir)):
            base, ext = os.path.splitext(fn)
            dt = np.int32 if ext == ".i32" else np.float64
            a = np.fromfile(os.path.join(ddir, fn), dtype=dt)
            if (base in NC_MESH_I or base in NC_MESH_F) and not keep_mesh:
                out["sha_" + base] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), np.uint8)
                continue
            out[base] = a
    meta = dict(mesh=NC_CASES[name][0], deck=dkw, grad=grad, niter=niter, ndump=ndump,
                raw_sha=nc_raw_sha(m), reference="NCFVM, amdflang -O2 -ffp-contract=off",
                driver="oracle/ref_nc_driver.F90", env="MALLOC_MMAP_THRESHOLD_=0")
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def main(argv):
    if not all(os.path.exists(p) for p in (DRIVER, NC_DRIVER, MAIN)):
        raise SystemExit("build the reference first: make -C oracle ref")
    for name in (argv or list(CASES) + list(NC_CASES) + list(MAIN_CASES)):
        p = (make_nc(name) if name in NC_CASES else make_main(name) if name in MAIN_CASES
             else make(name))
        print(f"{name}: {os.path.getsize(p) / 1024:.0f} KiB", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
