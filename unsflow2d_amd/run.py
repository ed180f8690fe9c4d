"""Whole runs with the behaviour of the reference's main programs.

``python -m unsflow2d_amd.run cc DECK [MESH] [--out DIR] [--precision mixed]``
``python -m unsflow2d_amd.run nc DECK [MESH] [--out DIR] [--grad lsqr] [--limiter bj]``

Each reads the reference's deck and ``geometry.inp``, preprocesses the mesh,
iterates on the device until ``MaxIter`` or the normalised residual reaches
``MinRes`` (CCFVM/main.F90:57-128, NCFVM/main.f90:51-88), writing ``FLO.RES``
as it goes, and at the end writes ``tecplt.dat``, ``XYCpCf.dat`` and the
restart file ``FLO.DAT``.  A deck whose start option is 2 (restart) begins
from ``FLO.DAT`` in the output directory; the reference declares that option
and the save_flow/read_flow routines but never calls them, so the restart
carries the conserved state only (limiter values and the residual
normalisation start afresh).
"""
from __future__ import annotations

import argparse
import math
import os
import sys

import numpy as np

from . import cc, flo_io, meshgen, nc, nc_mesh


def _mesh_path(deck: str, mesh: str | None, gridfile: str) -> str:
    if mesh:
        return mesh
    return os.path.join(os.path.dirname(os.path.abspath(deck)), gridfile)


def run_cc(deck: str, mesh: str | None = None, out: str = ".", precision: str = "fp64",
           log=None) -> dict:
    """The cell-centred main program on the device."""
    d, extra = flo_io.read_cc_deck(deck)
    raw = meshgen.read_geometry_inp(_mesh_path(deck, mesh, extra["gridfile"]))
    m = cc.Mesh.from_raw(raw)
    cfg = cc.Config(mach=d.mach, aoa_deg=d.aoa_deg, reynolds=d.reynolds, t_wall=d.t_wall,
                    flow_type=d.flow_type, iwall=d.iwall, flux=d.flux_type, grad=d.grad_type,
                    ilimit=d.ilimit, irs=d.irs, timemode=d.timemode, precision=precision,
                    cfl_type=d.cfl_type, cfl_min=d.cfl_min, cfl_max=d.cfl_max,
                    cfl_ramp_steps=d.cfl_ramp_steps, xref=d.xref, yref=d.yref)
    s = cc.Solver(m, cfg)
    counts = m.counts
    os.makedirs(out, exist_ok=True)
    if extra["istart"] == 2:
        q = s.get("qc")
        q[:counts["noc"]] = flo_io.read_flow(os.path.join(out, "FLO.DAT"), counts["noc"])
        s.set("qc", q)
    fres = 1.0
    with open(os.path.join(out, "FLO.RES"), "w") as f:
        while s.iter < d.maxiter and fres > d.minres:
            fres = float(s.iterate(1)[0])
            f.write(flo_io.flo_res_lines([s.iter], [math.log10(fres)], [s.cfl]))
            if log and s.iter % 100 == 0:
                log(f"iter {s.iter} fres {fres:.6e}")
    fo = s.forces(wall=True)
    a = {k: m.get(k) for k in ("fc_p1", "fc_p2", "fc_bc", "pt_x", "pt_y")}
    w = np.nonzero(a["fc_bc"] == 1001)[0]
    xm = 0.5 * (a["pt_x"][a["fc_p1"][w] - 1] + a["pt_x"][a["fc_p2"][w] - 1])
    ym = 0.5 * (a["pt_y"][a["fc_p1"][w] - 1] + a["pt_y"][a["fc_p2"][w] - 1])
    flo_io.write_xycpcf(os.path.join(out, "XYCpCf.dat"), xm, ym, fo["mcp"], fo["cf"])
    s.avg_c2v()
    # ent_inf = p_inf / r_inf**gamma with r_inf = 1 (CCFVM/solution.F90:24-32)
    p_inf = 1.0 / (flo_io.GAMMA * d.mach * d.mach)
    flo_io.write_tecplot_cc(os.path.join(out, "tecplt.dat"), a["pt_x"], a["pt_y"],
                            s.get("pt_prim"), s.get("pt_mu").ravel(), m.get("cell_nc2v"),
                            m.get("cell_c2v"), p_inf)
    flo_io.save_flow(os.path.join(out, "FLO.DAT"), s.get("qc")[:counts["noc"]])
    return dict(iter=s.iter, fres=fres, cl=fo["cl"], cd=fo["cd"], cl1=fo["cl1"], cd1=fo["cd1"],
                cm=fo["cm"])


def run_nc(deck: str, mesh: str | None = None, out: str = ".", grad: str = "gg",
           limiter: str = "vkn", log=None) -> dict:
    """The vertex-centred main program on the device.  The deck's flux field
    is not read by the reference (van Leer is hard-wired); "ausm" selects the
    build-defined AUSM+-up option here."""
    d, extra = flo_io.read_nc_deck(deck)
    raw = nc_mesh.read_geometry_inp(_mesh_path(deck, mesh, extra["gridfile"]))
    m = nc.NCMesh.from_raw(raw)
    flux = "ausm" if extra["flux_type"] == "ausm" else "vanleer"
    s = nc.NCSolver(m, mach=d.mach, aoa_deg=d.aoa_deg, cfl_max=d.cfl_max, grad=grad,
                    timemode=d.timemode, flux=flux, limiter=limiter)
    nop = m.counts["nop"]
    os.makedirs(out, exist_ok=True)
    if extra["istart"] == 2:
        s.set("qc", flo_io.read_flow(os.path.join(out, "FLO.DAT"), nop))
    fres = 1.0
    with open(os.path.join(out, "FLO.RES"), "w") as f:
        while s.iter < d.maxiter and fres > d.minres:
            h = s.iterate(1)[0]
            fres = float(h[0])
            f.write(flo_io.flo_res_lines([s.iter], [math.log10(fres)], [s.cfl(s.iter)],
                                         [math.log10(h[2])]))
            if log and s.iter % 100 == 0:
                log(f"iter {s.iter} fres {fres:.6e}")
    fo, mcp = s.forces()
    a = m.arrays()
    w = np.nonzero(a["fc_bc"] == 1001)[0]
    p1 = a["fc_p1"][w] - 1
    flo_io.write_xycpcf(os.path.join(out, "XYCpCf.dat"), a["pt_x"][p1], a["pt_y"][p1], mcp,
                        np.zeros_like(mcp))
    flo_io.write_tecplot_nc(os.path.join(out, "tecplt.dat"), a["pt_x"], a["pt_y"], s.get("qp"),
                            flo_io.nc_cell_walks(a))
    flo_io.save_flow(os.path.join(out, "FLO.DAT"), s.get("qc"))
    return dict(iter=s.iter, fres=fres, cl=float(fo[0]), cd=float(fo[1]))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m unsflow2d_amd.run")
    ap.add_argument("solver", choices=["cc", "nc"])
    ap.add_argument("deck")
    ap.add_argument("mesh", nargs="?")
    ap.add_argument("--out", default=".")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--grad", default="gg", choices=["gg", "lsqr"])
    ap.add_argument("--limiter", default="vkn", choices=["vkn", "bj"])
    a = ap.parse_args(argv)
    log = lambda msg: print(msg, flush=True)
    if a.solver == "cc":
        r = run_cc(a.deck, a.mesh, a.out, a.precision, log=log)
    else:
        r = run_nc(a.deck, a.mesh, a.out, a.grad, a.limiter, log=log)
    print(" ".join(f"{k}={v:.10g}" if isinstance(v, float) else f"{k}={v}" for k, v in r.items()),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
