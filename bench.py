#!/usr/bin/env python3
"""Benchmark of the cell-centred residual hot path on one MI355X.

Workload: the second-order Euler case on the seeded bump channel at
1000 x 1000 quads (2 000 000 triangles), bump height 0.08, Mach 0.75,
least-squares gradients with the Venkatakrishnan limiter, Roe flux, three
Runge-Kutta stages per iteration, CFL 0.8 with local time stepping.  A step
is one iteration of the reference's main loop (CCFVM/main.F90:57-128) on the
device: primitive refresh, vertex averages, time step, gradient + limiter and
three residual stages with their updates, then the residual norm.  Inputs are
resident in HBM before the timed region.

Metric: million cell-RK3-stages per second = 3 * cells * iterations / s / 1e6.

Prints one JSON line with, besides the contract keys, the roofline of the
dominant kernel (the residual stage, timed with HIP events on the library's
stream) and a CPU baseline (the reference solver compiled from its own
sources, timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_bytes(noc: int, nof: int, w: int, d: int = 8, i: int = 4) -> float:
    """Algorithmic bytes of one residual-stage launch, averaged over the three
    stages: face connectivity and geometry once per face, cell state once per
    cell (primitives, gradients, limiter values, centroid), the old and the
    current conserved state, dt/area, and the written state; the last stage
    also writes the next iteration's primitives and temperature (5w per
    cell, the primitive refresh it does for the next iteration)."""
    faces = nof * (2 * i + 5 * w)
    cells = noc * (2 * w + 4 * w + 8 * w + 4 * w) + noc * 4 * d + noc * w + noc * 4 * d
    prev_q = noc * 4 * d * 2.0 / 3.0
    epilogue = noc * 5 * w / 3.0
    return faces + cells + prev_q + epilogue


def pass_bytes(noc: int, nof: int, nop: int, nv2c: int, w: int, d: int = 8, i: int = 4) -> dict:
    """Algorithmic bytes per launch of the passes of a whole RK3 iteration, by
    profiling category: the vertex averages (P2), the face spectral radii
    (P3), the gradient pass with the limiter and the time step (P4) and a
    residual stage (P5, averaged over the three)."""
    p2 = nop * i + nv2c * (i + w) + noc * 4 * w + nop * 4 * w
    p3 = nof * (4 * i + 3 * w) + nop * 4 * w + noc * 2 * w
    p4 = nof * 2 * i + noc * (2 * w + 3 * w + 4 * w + 4 * w) + nof * 2 * w + noc * (8 * w + 4 * w)
    return {"avg_c2v": p2, "time_step": p3, "gradient": p4, "stage": stage_bytes(noc, nof, w, d, i)}


def iteration_bytes(noc: int, nof: int, nop: int, nv2c: int, w: int, d: int = 8, i: int = 4) -> float:
    """Algorithmic bytes of one RK3 iteration (all passes of the reference's
    call structure: primitive refresh P1, vertex averages P2, spectral radii
    P3, gradient and limiter P4, three residual stages P5)."""
    p1 = noc * (4 * d + 4 * w)
    p2 = nop * i + nv2c * (i + w) + noc * 4 * w + nop * 4 * w
    p3 = nof * (4 * i + 3 * w) + nop * 4 * w + noc * 2 * w
    p4 = nof * 2 * i + noc * (2 * w + 3 * w + 4 * w + 4 * w) + nof * 2 * w + noc * (8 * w + 4 * w)
    faces = nof * (2 * i + 5 * w)
    cells = noc * (2 * w + 4 * w + 8 * w + 4 * w) + noc * 4 * d + noc * w + noc * 4 * d
    return p1 + p2 + p3 + p4 + 3 * (faces + cells) + 2 * noc * 4 * d


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_stage_bump2M_hbm_traffic.md")


STAGE_KERNELS = {"gather": ("k_stage", "false"), "gather_lds": ("k_stage", "true"),
                 "shared": ("k_stage_blk", "false"), "shared_lds": ("k_stage_blk", "true")}


def kernel_name(category: str, precision: str, form: str) -> str:
    """The kernel that runs a profiling category of the Roe/LSQ/limited
    whole iteration, as rocprofv3 names it."""
    r = "double" if precision == "fp64" else "float"
    if category == "stage":
        k, lds = STAGE_KERNELS[form]
        return f"ufx::{k}<{r},1,{lds}>"
    if category == "gradient":
        return f"ufx::k_cell<{r},1,false,true,true>"
    if category == "time_step":
        return f"ufx::k_face_lambda<{r}>"
    return f"ufx::k_avg_c2v_range<{r},false>"


def measured_traffic(precision: str, kernel: str) -> int | None:
    """HBM bytes per launch of `kernel`, from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the same
    counter reads for a plain copy of known size): lines
    ``traffic <precision> <kernel> <bytes>``.  None when not measured."""
    import re
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                m = re.match(r"^traffic (\S+) (\S+) (\d+)\s*$", line.strip())
                if m and m.group(1) == precision and m.group(2) == kernel:
                    return int(m.group(3))
    except OSError:
        pass
    return None


def cpu_baseline(budget_s: float = 20.0) -> dict | None:
    """Time the reference solver (oracle/_ref, built from the reference's own
    sources) on a bounded sample of the workload: the same generator, flow
    and scheme on a 200 x 200-quad channel (80 000 triangles), warm-up then a
    timed window, single thread."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cc_fast", "ref_cc_driver")
    from unsflow2d_amd import meshgen
    nq = 200
    raw = meshgen.bump_channel(nq, nq, 2, h=0.08)
    cells = raw.noc
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as tmp:
            meshgen.write_geometry_inp(raw, os.path.join(tmp, "geometry.inp"))
            deck = meshgen.Deck(mach=0.75, maxiter=10)
            with open(os.path.join(tmp, "flo.inp"), "w") as f:
                f.write(deck.text())
            # a short probe sizes the window to the budget
            warm = 3
            out = subprocess.run([exe, "flo.inp", ".", str(warm + 3), "0", "-1", str(warm)],
                                 cwd=tmp, capture_output=True, text=True, timeout=600)
            ms = float(out.stdout.split("ms_per_iteration")[-1].split()[0])
            nit = max(5, int(budget_s * 1000.0 / ms))
            out = subprocess.run([exe, "flo.inp", ".", str(warm + nit), "0", "-1", str(warm)],
                                 cwd=tmp, capture_output=True, text=True, timeout=1200)
            ms = float(out.stdout.split("ms_per_iteration")[-1].split()[0])
        kind, what = "reference", "reference CCFVM (amdflang -O3 -march=x86-64-v3)"
    else:
        from oracle.oracle import CCOracle
        from unsflow2d_amd import cc
        mesh = cc.Mesh.from_raw(raw)
        orc = CCOracle(mesh.arrays(), mach=0.75)
        for it in range(1, 4):
            orc.iteration(it, 0.8)
        t0 = time.perf_counter()
        nit = 0
        while time.perf_counter() - t0 < budget_s:
            nit += 1
            orc.iteration(3 + nit, 0.8)
        ms = (time.perf_counter() - t0) * 1000.0 / nit
        kind, what = "port", "C restatement of the reference (gcc -O2)"
    value = 3.0 * cells / (ms / 1000.0) / 1e6
    return {"value": round(value, 3), "unit": "M cell-RK3-stages/s", "cores": 1, "kind": kind,
            "sample": f"{what}, bump channel {nq}x{nq} quads ({cells} triangles), M=0.75, LSQ + "
                      f"limiter, Roe, RK3, {nit} timed iterations after 3 warm-up, "
                      f"{ms:.2f} ms/iteration, 1 thread"}


def batch_cpu_baseline(cs, maxiter: int, budget_s: float = 20.0) -> dict | None:
    """The reference (oracle/_ref/cc_fast) on a bounded sample of the batched
    workload: P evenly spaced cases (P = min(16, host cores)), each in its own
    single-threaded process, all P at once, each timed over a window sized to
    about `budget_s`.  The sample's cost per cell-iteration, times the whole
    batch's cells x MaxIter spread over P cores, gives the host's solves/s."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cc_fast", "ref_cc_driver")
    if not os.path.exists(exe):
        return None
    from unsflow2d_amd import meshgen
    ncore = max(1, min(16, os.cpu_count() or 1))
    sample = [cs[i * len(cs) // ncore] for i in range(min(ncore, len(cs)))]
    warm = 2
    with tempfile.TemporaryDirectory() as tmp:
        procs = []
        for c in sample:
            d = os.path.join(tmp, f"case{c.k}")
            os.makedirs(d)
            meshgen.write_geometry_inp(meshgen.bump_channel(c.nx, c.ny, c.seed, h=c.h),
                                       os.path.join(d, "geometry.inp"))
            with open(os.path.join(d, "flo.inp"), "w") as f:
                f.write(meshgen.Deck(mach=c.mach, aoa_deg=c.aoa_deg, maxiter=10).text())
            # window sized from a nominal 3 M cell-stages/s per core
            nit = int(min(20000, max(5, budget_s * 1000.0 / (c.cells * 3.0 / 3.0e6 * 1000.0))))
            procs.append((c, nit, subprocess.Popen(
                [exe, "flo.inp", ".", str(warm + nit), "0", "-1", str(warm)], cwd=d,
                stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        ms_sum = cells_sum = 0.0
        for c, nit, pr in procs:
            out, _ = pr.communicate(timeout=1200)
            ms_sum += float(out.split("ms_per_iteration")[-1].split()[0])
            cells_sum += c.cells
    ms_per_cell_it = ms_sum / cells_sum
    total_s = sum(c.cells for c in cs) * maxiter * ms_per_cell_it / 1000.0 / ncore
    return {"value": round(len(cs) / total_s, 4), "unit": "solves/s", "cores": ncore,
            "kind": "reference",
            "sample": f"reference CCFVM (amdflang -O3 -march=x86-64-v3, FP64) on {len(sample)} of "
                      f"the {len(cs)} cases (k = 0, {sample[1].k if len(sample) > 1 else 0}, ...; "
                      f"{int(cells_sum)} cells), one single-threaded process per case, "
                      f"{len(sample)} at once, each timed over a window of about {budget_s:.0f} s "
                      f"after {warm} warm-up iterations; {ms_per_cell_it * 1e6:.1f} ns per "
                      f"cell-iteration scaled to all {len(cs)} cases x {maxiter} iterations "
                      f"over {ncore} cores"}


def copy_probe(mib: int) -> int:
    """A plain device-to-device copy of known size, five times: the reference
    point that scales the profiler's HBM read/write counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"# copy probe: {mib} MiB read + {mib} MiB written per launch, 5 launches", flush=True)
    return 0


def nc_cpu_baseline(budget_s: float = 20.0) -> dict | None:
    """The reference vertex-centred solver (oracle/_ref/nc_fast) on a bounded
    sample of the vertex-centred workload: the same generator at 500 x 250
    points (125 000 vertices, a working set well beyond one core's 32 MiB of
    L3), LU-SGS, GG, single thread."""
    exe = os.path.join(ROOT, "oracle", "_ref", "nc_fast", "ref_nc_driver")
    if not os.path.exists(exe):
        return None
    from unsflow2d_amd import nc_mesh
    m = nc_mesh.o_grid(500, 250, 3, stretch=1.005 ** (999 / 249))
    env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="0")
    with tempfile.TemporaryDirectory() as tmp:
        nc_mesh.write_geometry_inp(m, os.path.join(tmp, "geometry.inp"))
        with open(os.path.join(tmp, "flo.inp"), "w") as f:
            f.write(nc_mesh.NCDeck(mach=0.4, cfl_max=50.0, timemode="lusgs").text())
        warm = 3
        out = subprocess.run([exe, "flo.inp", ".", str(warm + 10), "0", "gg", str(warm)], cwd=tmp,
                             capture_output=True, text=True, timeout=600, env=env)
        ms = float(out.stdout.split("ms_per_iteration")[-1].split()[0])
        nit = max(10, int(budget_s * 1000.0 / ms))
        out = subprocess.run([exe, "flo.inp", ".", str(warm + nit), "0", "gg", str(warm)], cwd=tmp,
                             capture_output=True, text=True, timeout=1200, env=env)
        ms = float(out.stdout.split("ms_per_iteration")[-1].split()[0])
    value = m.nop / (ms / 1000.0) / 1e6
    return {"value": round(value, 4), "unit": "M vertex-iterations/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference NCFVM (amdflang -O3 -march=x86-64-v3), O-grid 500x250 points "
                      f"({m.nop} vertices), M=0.4, GG + limiter, van Leer, LU-SGS, {nit} timed "
                      f"iterations after {warm} warm-up, {ms:.2f} ms/iteration, 1 thread"}


def nc_bytes(nop: int, nof: int, w: int = 8, d: int = 8, i: int = 4) -> float:
    """Algorithmic bytes of one residual pass of the vertex-centred solver:
    edge connectivity and geometry once per edge, each vertex's record
    (primitives, gradient, limiter, coordinates), dual volume, and the
    written residual, spectral radius and time step."""
    return nof * (2 * i + 3 * w + w) + nop * (4 * w + 8 * w + 4 * w + 2 * w + w + 4 * d + 2 * d)


def nc_pass_bytes(nop: int, nof: int, w: int = 8, d: int = 8, i: int = 4) -> dict:
    """Algorithmic bytes per iteration of each pass of the vertex-centred
    LU-SGS iteration (the per-pass model: every array a pass needs read once,
    every output written once): save_old, gradient + limiter, residual, and
    the two sweeps (per swept edge its ends, geometry and spectral radius; per
    vertex the residual, the three states it reads, its time-step terms and
    the written dq and state)."""
    return {
        "save_old": nop * (4 * d + 4 * w + 4 * d),
        "gradient": nof * (2 * i + 3 * w) + nop * (4 * w + w + 8 * w)
                    + nof * 2 * i + nop * (2 * w + 4 * w + 8 * w + w + 4 * w + 4 * w),
        "residual": nc_bytes(nop, nof, w, d, i),
        "lusgs": 2 * (nof * (2 * i + 3 * w + w) + nop * (4 * d * 3 + 4 * d + 3 * w + 4 * d * 2)),
    }


def main_nc(args) -> int:
    from unsflow2d_amd import nc, nc_mesh
    t0 = time.time()
    raw = nc_mesh.config3_mesh()
    mesh = nc.NCMesh.from_raw(raw)
    cnt = mesh.counts
    solver = nc.NCSolver(mesh, mach=0.4, cfl_max=50.0, grad="gg", timemode="lusgs",
                         lusgs_order=args.lusgs_order, precision=args.precision,
                         flux=args.nc_flux, limiter=args.nc_limiter)
    ncol = int(mesh.get("pt_colour").max()) + 1
    print(f"# setup {time.time() - t0:.1f}s: {cnt['nop']} vertices, {cnt['nof']} edges",
          file=sys.stderr, flush=True)
    solver.iterate(args.warmup, hist=False)
    solver.sync()
    t1 = time.perf_counter()
    solver.iterate(args.steps, hist=False)
    solver.sync()
    wall = time.perf_counter() - t1
    # per-kernel times from more iterations with HIP events around every
    # launch (kept out of the timed region; in the reference order the
    # LU-SGS sweeps are thousands of small launches)
    nprof = 1 if args.lusgs_order == "reference" else min(args.steps, 10)
    solver.profile(True)
    h = solver.iterate(nprof)
    prof = solver.profile_read()
    solver.profile(False)
    prof = {k: (v[0] * args.steps / nprof, v[1] * args.steps / nprof) for k, v in prof.items()}
    if not np.all(np.isfinite(h)):
        print("# residual is not finite", file=sys.stderr)
        return 1
    nop, nof = cnt["nop"], cnt["nof"]
    ms_step = wall * 1000.0 / args.steps
    kern = {k: round(v[0] / max(1, v[1]), 4) for k, v in prof.items() if v[1]}
    print(f"# kernel avg ms: {kern}; LU-SGS total per iteration "
          f"{prof['lusgs'][0] / args.steps:.3f} ms; iteration {ms_step:.3f} ms",
          file=sys.stderr, flush=True)
    # the roofline of every pass (algorithmic bytes per iteration / its time
    # per iteration); the line's roofline is the pass that takes longest
    pb = nc_pass_bytes(nop, nof, w=8 if args.precision == "fp64" else 4)
    passes = {}
    for k, b in pb.items():
        ms_it = prof[k][0] / args.steps
        if ms_it > 0:
            passes[k] = {"ms_per_iteration": round(ms_it, 4), "bytes_per_iteration": int(b),
                         "achieved_gbs": round(b / (ms_it / 1000.0) / 1e9, 1),
                         "frac": round(b / (ms_it / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                         "launches_per_iteration": round(prof[k][1] / args.steps, 1)}
    dom = max(passes, key=lambda k: passes[k]["ms_per_iteration"])
    kname = {"lusgs": "k_nc_lusgs_mc" if args.lusgs_order == "multicolour" else "k_nc_lusgs_stream",
             "residual": "k_nc_residual_blk", "gradient": "k_nc_grad_lim",
             "save_old": "k_nc_save_old"}[dom]
    d_ms = passes[dom]["ms_per_iteration"]
    d_n = passes[dom]["launches_per_iteration"]
    achieved = passes[dom]["achieved_gbs"]
    cpu = None if args.no_cpu else nc_cpu_baseline(args.cpu_budget)
    flux_name = {"vanleer": "van Leer", "ausm": "AUSM+-up"}[args.nc_flux]
    lim_name = {"vkn": "Venkatakrishnan", "bj": "Barth-Jespersen"}[args.nc_limiter]
    sweep = ("multicolour: one launch per colour per sweep, %d colours; the first colour's "
             "forward update in the residual pass, the last colour's reverse update in its "
             "forward launch" % ncol
             if args.lusgs_order == "multicolour" else
             "level-scheduled, reference order, one workgroup per sweep")
    line = {
        "metric": "M vertex-iterations/s @2M vertices, NCFVM LU-SGS",
        "value": round(nop * args.steps / wall / 1e6, 3),
        "unit": "M vertex-iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"NCFVM median dual, seeded O-grid 2000x1000 points around the "
                               f"unit circle ({nop} vertices, {nof} edges), M=0.4, GG + "
                               f"{lim_name}, {flux_name}, LU-SGS ({sweep}), tanh CFL ramp to 50",
                   "vertices": nop, "edges": nof, "lusgs_order": args.lusgs_order,
                   "precision": args.precision, "flux": args.nc_flux,
                   "limiter": args.nc_limiter, "colours": ncol},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("nc-" + args.precision, "ufx::" + kname),
                     "kernel": kname, "bytes_per_launch": int(pb[dom] / max(1.0, d_n)),
                     "avg_launch_ms": round(d_ms / max(1.0, d_n), 5),
                     "basis": f"{dom} pass: algorithmic bytes per iteration over its "
                              f"{d_n:g} launches per iteration"},
        "passes": passes,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


# algorithmic HBM bytes per cell-stage of the second-order inviscid RK3
# iteration (LSQ + limiter + Roe), the per-pass model of the cell-centred
# path: 468.8 B in FP64, 296.7 B in the mixed form
MODEL_BYTES_PER_CELL_STAGE = {"fp64": 468.8, "mixed": 296.7}


def sequential_baseline(cs, precision: str, maxiter: int, window: int) -> dict:
    """The batched workload solved the plain way: one case after another,
    each in its own context with its kernels launched eagerly
    (ufx_cc_iterate), as a user of the single-mesh interface would.  Every
    case runs iteration 1, then `window` timed iterations; its time per
    iteration times (MaxIter - 1) plus iteration 1 gives the case's solve
    time, summed over the cases (setup excluded)."""
    from unsflow2d_amd import cc, meshgen
    total = 0.0
    for c in cs:
        m = cc.Mesh.from_raw(meshgen.bump_channel(c.nx, c.ny, c.seed, h=c.h))
        s = cc.Solver(m, cc.Config(mach=c.mach, aoa_deg=c.aoa_deg, flux="roe", grad="lsqr",
                                   ilimit=1, precision=precision, cfl_type="ramp", cfl_min=0.8,
                                   cfl_max=0.8, cfl_ramp_steps=1))
        s.sync()
        t0 = time.perf_counter()
        s.iterate(1, hist=False)
        s.sync()
        t1 = time.perf_counter()
        s.iterate(window, hist=False)
        s.sync()
        t2 = time.perf_counter()
        total += (t1 - t0) + (t2 - t1) * (maxiter - 1) / window
        s.close()
        m.close()
    return {"solves_per_s": len(cs) / total, "wall_s_scaled": total, "window": window}


def main_batch(args) -> int:
    """Many independent bump channels at once (seeded sizes, Mach numbers,
    bump heights, angles of attack), each to its density-residual tolerance
    or MaxIter, on one GPU."""
    from unsflow2d_amd import batch
    cs = batch.cases(args.cases)
    print(f"# batch: {len(cs)} cases, {sum(c.cells for c in cs)} cells, precision "
          f"{args.precision}, MaxIter {args.maxiter}, chunk {args.chunk}", file=sys.stderr,
          flush=True)
    r = batch.run(cs, precision=args.precision, maxiter=args.maxiter, minres=1e-8,
                  chunk=args.chunk, fused=not args.per_case, tune=not args.no_tune,
                  log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    achieved = r["cell_stages_per_s"] * MODEL_BYTES_PER_CELL_STAGE[args.precision] / 1e9
    seq = None
    if args.seq_window > 0:
        t0 = time.time()
        seq = sequential_baseline(cs, args.precision, args.maxiter, args.seq_window)
        print(f"# sequential per-case eager launches: {seq['solves_per_s']:.3f} solves/s "
              f"({seq['wall_s_scaled']:.1f} s for all cases, measured in {time.time() - t0:.1f} s)",
              file=sys.stderr, flush=True)
    line = {
        "metric": "batched solves/s (bump channels, 5k-200k cells each)",
        "value": round(r["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.maxiter, "warmup": 0, "ms_per_step": round(r["wall"] * 1e3 / args.maxiter, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": f"{len(cs)} seeded bump channels (cells 6 ny^2, log-uniform in "
                               "[5e3, 2e5]; M in [0.3, 0.8], h in [0.02, 0.1], AoA in [-2, 2] "
                               "deg), LSQ+Venkatakrishnan, Roe, RK3, CFL 0.8, each to a density "
                               "residual drop of 1e-8 or MaxIter",
                   "cases": len(cs), "cells": sum(c.cells for c in cs), "max_iter": args.maxiter,
                   "precision": args.precision,
                   "launch": "one context per case, each from its own graph" if args.per_case
                             else "one launch per kernel for all live cases"},
        "cell_stages_per_s": round(r["cell_stages_per_s"]),
        "sequential_baseline": None if seq is None else {
            "value": round(seq["solves_per_s"], 4), "unit": "solves/s",
            "speedup": round(r["solves_per_s"] / seq["solves_per_s"], 3),
            "method": f"each case alone, one after another, eager launches (ufx_cc_iterate): "
                      f"iteration 1 plus {seq['window']} timed iterations per case, scaled to "
                      f"MaxIter; setup excluded"},
        "converged": r["converged"], "wall_s": round(r["wall"], 3), "setup_s": round(r["setup"], 1),
        "autotune": r["tuned"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "basis": "model bytes per cell-stage x aggregate cell-stages/s"},
        "cpu_baseline": None if args.no_cpu else batch_cpu_baseline(cs, args.maxiter,
                                                                     args.cpu_budget),
    }
    print(json.dumps(line), flush=True)
    return 0


def viscous_solver(precision: str, mesh=None):
    """The laminar channel: 800 x 400 quads, first wall layer 1e-3 growing
    by 1.12 (capped), streamwise-only jitter (640 000 triangles); M = 0.5,
    Re = 5000 per unit length, Sutherland viscosity, Pr 0.72, adiabatic
    no-slip bump wall, GG-FC gradients with the limiter, Roe, RK3 with
    residual smoothing at CFL 0.8.  (With a first layer of 1e-5 or 1e-4 the
    reference scheme itself stops at iteration 1: its ghost primitives are
    still zero then, CCFVM/solution.F90:78-81, and across so thin a wall
    layer the reconstruction makes the Roe sonic speed negative.)"""
    from unsflow2d_amd import cc, meshgen
    if mesh is None:
        mesh = cc.Mesh.from_raw(meshgen.bump_channel(800, 400, 4, h=0.05, first=1e-3,
                                                     growth=1.12, jitter_y=False))
    cfg = cc.Config(mach=0.5, flow_type="laminar", reynolds=5000.0, iwall=1, flux="roe",
                    grad="ggfc", ilimit=1, irs=1, precision=precision, cfl_type="ramp",
                    cfl_min=0.8, cfl_max=0.8, cfl_ramp_steps=1)
    return mesh, cc.Solver(mesh, cfg)


def main_viscous(args) -> int:
    t0 = time.time()
    mesh, solver = viscous_solver(args.precision)
    counts = mesh.counts
    print(f"# setup {time.time() - t0:.1f}s: {counts['noc']} cells", file=sys.stderr, flush=True)
    solver.iterate(args.warmup, hist=True)
    solver.sync()
    t1 = time.perf_counter()
    solver.iterate(args.steps, hist=False)
    solver.sync()
    wall = time.perf_counter() - t1
    # a device error (the reference's STOPs) would have cut the timed work
    # short: the next whole iteration reports it
    solver.iterate(1, hist=True)
    solver.profile(True)
    solver.iterate(min(args.steps, 10), hist=False)
    prof = solver.profile_read()
    solver.profile(False)
    n = min(args.steps, 10)
    kern = {k: round(v[0] / n, 4) for k, v in prof.items() if v[1]}
    print(f"# kernel ms per iteration: {kern}", file=sys.stderr, flush=True)
    cf_err = None
    if args.cf_check:
        # wall skin friction of both precisions after the same iterations
        # from one developed FP64 state (100 FP64 iterations; the state,
        # primitives with the ghosts', and limiter values handed over): from
        # the free stream the mixed form cannot start, since the reference's
        # iteration-1 vertex averages (ghost primitives still zero) hold
        # negative densities and pressures at wall vertices, and FP32
        # rounding flips the sign of p/rho on some faces
        _, s64 = viscous_solver("fp64", mesh)
        s64.iterate(100, hist=False)
        _, s32 = viscous_solver("mixed", mesh)
        for name in ("qc", "qp", "phi"):
            s32.set(name, s64.get(name))
        s64.iterate(args.cf_check, hist=False)
        s32.iterate(args.cf_check, hist=False)
        ref = s64.forces(wall=True)["cf"]
        cf32 = s32.forces(wall=True)["cf"]
        cf_err = float(np.max(np.abs(cf32 - ref)) / np.max(np.abs(ref)))
        s64.close()
        s32.close()
    noc = counts["noc"]
    line = {
        "metric": "M cell-RK3-stages/s @640k cells, laminar Navier-Stokes channel",
        "value": round(3.0 * noc * args.steps / wall / 1e6, 2), "unit": "M cell-RK3-stages/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": "CCFVM laminar NS, bump channel 800x400 quads (640000 triangles), "
                               "first wall layer 1e-3 x1.12 capped, M=0.5, Re=5000, Sutherland, "
                               "Pr 0.72, adiabatic no-slip, GG-FC + limiter, Roe, RK3 + residual "
                               f"smoothing, CFL 0.8, {args.precision}",
                   "cells": noc, "precision": args.precision},
        "kernel_ms_per_iteration": kern,
        "cf_max_rel_error_mixed_vs_fp64": cf_err,
        "cf_check_iterations": args.cf_check or None,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def time_iterations(solver, args, category: str = "stage") -> tuple[float, float]:
    """Warm-up, then K iterations timed with the host clock around them
    (synchronised on both sides); HIP events on the library's stream bracket
    the launches of one kernel category (an event pair around every launch
    would itself cost about a tenth of the iteration).  Returns (wall
    seconds, mean launch ms of that category)."""
    solver.iterate(args.warmup, hist=True)
    solver.sync()
    solver.profile_only(category)
    solver.sync()
    t1 = time.perf_counter()
    solver.iterate(args.steps, hist=False)
    solver.sync()
    wall = time.perf_counter() - t1
    st_ms, st_n = solver.profile_read()[category]
    solver.profile(False)
    return wall, st_ms / max(1, st_n)


def precision_errors(mesh, niter: int) -> dict:
    """The mixed form against the FP64 form after the same `niter`
    iterations from the free stream: force coefficients (the pressure
    integral's and the flux-based pair the reference prints, and Cm), wall
    -Cp, the normalised residual history and the conserved field."""
    from unsflow2d_amd import cc
    out, sol = {}, {}
    for prec in ("fp64", "mixed"):
        s = cc.Solver(mesh, cc.Config(mach=0.75, flux="roe", grad="lsqr", ilimit=1,
                                      precision=prec))
        h = s.iterate(niter, hist=True)
        sol[prec] = (h, s.forces(wall=True), s.get("qc"))
        s.close()
    h64, f64, q64 = sol["fp64"]
    h32, f32, q32 = sol["mixed"]
    for k in ("cl", "cd", "cl1", "cd1", "cm"):
        out[k] = float(abs(f32[k] - f64[k]))
    out["cp_max"] = float(np.max(np.abs(f32["mcp"] - f64["mcp"])))
    out["history_max_rel"] = float(np.max(np.abs(h32 - h64) / np.abs(h64)))
    scale = np.max(np.abs(q64), axis=0)
    out["field_linf_rel"] = float(np.max(np.max(np.abs(q32 - q64), axis=0) / scale))
    out["residual_drop_fp64"] = float(h64[-1])
    out["cl_fp64"], out["cd_fp64"] = float(f64["cl1"]), float(f64["cd1"])
    out["iterations"] = niter
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: mixed for the batched cases (their scheme), else fp64")
    ap.add_argument("--quads", type=int, default=1000, help="quads per side of the channel")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters)")
    ap.add_argument("--workload", choices=["cc", "nc", "batch", "viscous"], default="cc",
                    help="cc: the cell-centred headline workload; nc: the vertex-centred one; "
                         "batch: many independent cell-centred cases; viscous: the laminar "
                         "channel")
    ap.add_argument("--cf-check", type=int, default=0, metavar="N",
                    help="viscous: also compare wall Cf of mixed vs FP64 after N iterations")
    ap.add_argument("--cases", type=int, default=256, help="batch: number of cases")
    ap.add_argument("--maxiter", type=int, default=2000, help="batch: MaxIter per case")
    ap.add_argument("--chunk", type=int, default=50, help="batch: iterations per graph replay")
    ap.add_argument("--per-case", action="store_true",
                    help="batch: launch every case's kernels separately (one graph per case)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default residual-stage form instead of tuning it on this GPU")
    ap.add_argument("--seq-window", type=int, default=100, metavar="N",
                    help="batch: also time every case alone, one after another, eager, over N "
                         "iterations scaled to MaxIter (0: skip)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="cc: skip the mixed-precision leg and its errors against FP64")
    ap.add_argument("--error-iters", type=int, default=2000,
                    help="cc: iterations from the free stream before comparing mixed with FP64")
    ap.add_argument("--nc-flux", choices=["vanleer", "ausm"], default="vanleer",
                    help="nc: dual-face flux (AUSM+-up is build-defined for the vertex-centred path)")
    ap.add_argument("--nc-limiter", choices=["vkn", "bj"], default="vkn",
                    help="nc: Venkatakrishnan or Barth-Jespersen (build-defined)")
    ap.add_argument("--lusgs-order", choices=["reference", "multicolour"], default="multicolour",
                    help="nc: LU-SGS sweep order (reference: level schedule of the reference's "
                         "numbering; multicolour: one launch per colour)")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "mixed" if args.workload == "batch" else "fp64"
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.workload == "viscous":
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            return 2
        return main_viscous(args)
    if args.workload == "batch":
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            return 2
        return main_batch(args)
    if args.workload == "nc":
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            return 2
        return main_nc(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path "
              "shards by mesh partition, which is not implemented)", file=sys.stderr)
        return 2

    from unsflow2d_amd import cc, meshgen
    t0 = time.time()
    raw = meshgen.bump_channel(args.quads, args.quads, 2, h=0.08)
    mesh = cc.Mesh.from_raw(raw)
    counts = mesh.counts
    nv2c = int(mesh.get("pt_nv2c").sum())
    cfg = cc.Config(mach=0.75, flux="roe", grad="lsqr", ilimit=1, precision=args.precision)
    solver = cc.Solver(mesh, cfg)
    print(f"# setup {time.time() - t0:.1f}s: {counts['noc']} cells, {counts['nof']} faces, "
          f"{counts['nop']} vertices", file=sys.stderr, flush=True)

    noc, nof, nop = counts["noc"], counts["nof"], counts["nop"]
    w = 8 if args.precision == "fp64" else 4
    pbytes = pass_bytes(noc, nof, nop, nv2c, w)
    say = lambda m: print(f"# {m}", file=sys.stderr, flush=True)
    # the residual-stage form that runs fastest on this GPU (the tuning
    # iterations are real iterations; they precede the warm-up)
    tune = None if args.no_tune else solver.tune_stage_form(log=say)
    form = solver.stage_form
    # the per-kernel breakdown, from a pass of K iterations with every launch
    # bracketed; the pass with the largest share is the roofline's kernel
    solver.iterate(args.warmup, hist=False)
    solver.profile(True)
    solver.iterate(args.steps, hist=False)
    prof_all = solver.profile_read()
    solver.profile(False)
    dominant = max(pbytes, key=lambda k: prof_all[k][0])
    # timed region: K iterations; HIP events on the library's stream bracket
    # the dominant pass's launches only
    wall, k_avg = time_iterations(solver, args, dominant)
    hist = solver.iterate(1, hist=True)
    if not np.all(np.isfinite(hist)):
        print("# residual is not finite", file=sys.stderr)
        return 1

    ms_step = wall * 1000.0 / args.steps
    value = 3.0 * noc * args.steps / wall / 1e6
    kb = pbytes[dominant]
    achieved = kb / (k_avg / 1000.0) / 1e9
    it_bytes = iteration_bytes(noc, nof, nop, nv2c, w)
    kern = {k: round(v[0] / max(1, v[1]), 4) for k, v in prof_all.items() if v[1]}
    per_it = {k: round(v[0] / args.steps, 4) for k, v in prof_all.items() if v[1]}
    say(f"stage form {form}; kernel avg ms: {kern}; per iteration: {per_it}; iteration "
        f"{ms_step:.3f} ms, {it_bytes / (ms_step / 1000.0) / 1e9:.0f} GB/s algorithmic over the "
        f"whole iteration")
    passes = {k: {"ms_per_iteration": per_it[k], "avg_launch_ms": kern[k],
                  "bytes_per_launch": int(pbytes[k]),
                  "frac": round(pbytes[k] / (kern[k] / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)}
              for k in pbytes if k in kern}

    solver.close()
    # the mixed form (FP32 records, gradients, limiter and fluxes; FP64 state,
    # residual sums, time steps and norms) on the same mesh, timed the same
    # way, and its errors against FP64 after the same iterations
    mixed = {}
    if args.precision == "fp64" and not args.no_mixed:
        s32 = cc.Solver(mesh, cc.Config(mach=0.75, flux="roe", grad="lsqr", ilimit=1,
                                        precision="mixed"))
        tune32 = None if args.no_tune else s32.tune_stage_form(log=say)
        wall32, st32 = time_iterations(s32, args, "stage")
        form32 = s32.stage_form
        s32.close()
        v32 = 3.0 * noc * args.steps / wall32 / 1e6
        sb32 = stage_bytes(noc, nof, 4)
        t0 = time.time()
        err = precision_errors(mesh, args.error_iters)
        say(f"mixed ({form32}): {wall32 * 1000.0 / args.steps:.3f} ms/iteration, {v32:.0f} M "
            f"cell-stages/s ({v32 / value:.3f}x); errors vs FP64 after {args.error_iters} "
            f"iterations ({time.time() - t0:.1f}s): {err}")
        mixed = {"mixed_value": round(v32, 2), "mixed_speedup": round(v32 / value, 4),
                 "mixed_ms_per_step": round(wall32 * 1000.0 / args.steps, 4),
                 "mixed_stage_form": form32,
                 "mixed_stage_roofline_frac": round(sb32 / (st32 / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                 "mixed_errors_vs_fp64": err}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_budget)
        except Exception as e:  # the baseline is reported, never required
            print(f"# cpu baseline failed: {e}", file=sys.stderr)

    line = {
        "metric": "M cell-RK3-stages/s @2M cells, FP64 vs mixed at CL/CD<=1e-6; batched solves/s",
        "value": round(value, 2),
        "unit": "M cell-RK3-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"CCFVM 2nd-order Euler, seeded bump channel {args.quads}x{args.quads} "
                               f"quads ({noc} triangles), h=0.08, M=0.75, LSQ+Venkatakrishnan, Roe, "
                               f"RK3, CFL 0.8 local time step, {args.precision}",
                   "cells": noc, "faces": nof, "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.precision,
                                                 kernel_name(dominant, args.precision, form)),
                     "kernel": kernel_name(dominant, args.precision, form),
                     "bytes_per_launch": int(kb),
                     "avg_launch_ms": round(k_avg, 5)},
        "iteration_roofline_frac": round(it_bytes / (ms_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
        "passes": passes,
        "stage_form": form,
        "stage_form_tuning": tune,
        **mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
