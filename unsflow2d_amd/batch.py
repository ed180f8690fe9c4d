"""Many independent cell-centred cases on one GPU.

The batched workload: N bump channels of random size, Mach number, bump
height and angle of attack, each solved with the reference's scheme (LSQ
gradient, Venkatakrishnan limiter, Roe flux, RK3 at CFL 0.8) until its
density residual has dropped by ``minres`` or it has run ``maxiter``
iterations (the reference's own loop condition, CCFVM/main.F90:57).

Every case has its own device context.  Iteration 1 runs eagerly (it fixes
the residual normalisation); after it, all cases advance together in chunks
of whole iterations with one launch per kernel for every live case
(``ufx_cc_batch_iterate``, replayed from a HIP graph), or, with
``fused=False``, each case from its own HIP graph on its own stream
(``ufx_cc_iterate_graph``), launched round-robin so that the GPU overlaps
small cases' kernels.  Between chunks the density-only residual is read back
per case; a case that reaches the tolerance or the iteration cap is retired
and its context freed.

Case k draws its parameters from ``numpy.random.default_rng(seed0 + k)``:
cells C = 6 ny^2 with ny = round(sqrt(C / 6)), C log-uniform in
[cmin, cmax]; M uniform in [0.3, 0.8]; bump height h uniform in
[0.02, 0.1]; angle of attack uniform in [-2, 2] degrees; the mesh seed is
seed0 + k.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import cc, meshgen


@dataclasses.dataclass
class Case:
    k: int
    ny: int
    mach: float
    h: float
    aoa_deg: float
    seed: int

    @property
    def nx(self) -> int:
        return 3 * self.ny

    @property
    def cells(self) -> int:
        return 2 * self.nx * self.ny


def cases(n: int = 256, seed0: int = 5000, cmin: float = 5e3, cmax: float = 2e5) -> list[Case]:
    out = []
    for k in range(n):
        rng = np.random.default_rng(seed0 + k)
        c = math.exp(rng.uniform(math.log(cmin), math.log(cmax)))
        ny = max(2, int(round(math.sqrt(c / 6.0))))
        out.append(Case(k=k, ny=ny, mach=float(rng.uniform(0.3, 0.8)),
                        h=float(rng.uniform(0.02, 0.1)), aoa_deg=float(rng.uniform(-2.0, 2.0)),
                        seed=seed0 + k))
    return out


FORMS = ("gather_lds", "gather", "shared", "shared_lds")


def autotune(fb: cc.Batch, solvers: list, iters: int = 10, reps: int = 2,
             forms=FORMS, log=None) -> dict:
    """Pick the residual-stage form (per-cell gather, or per-block edge
    batches of fluxes evaluated once in LDS; each with or without the
    block's records staged in LDS) that runs this batch fastest on this GPU.

    Each candidate is set on every case, one chunk of `iters` iterations is
    captured and replayed `reps` times, and the replays after the first are
    timed; the fastest form is left set.  The tuning iterations are real
    iterations of every case (they count towards MaxIter), so tuning costs
    only the difference between the forms.  Returns the choice and the
    measured milliseconds per iteration of every candidate (None where the
    form does not fit in LDS on some case's mesh)."""
    times = {}
    for form in forms:
        try:
            for s in solvers:
                s.set_stage_form(form)
        except cc.UfxError:
            times[form] = None
            continue
        fb.iterate(iters)  # capture (and run) the chunk
        t0 = time.perf_counter()
        for _ in range(reps - 1):
            fb.iterate(iters)
        times[form] = (time.perf_counter() - t0) * 1e3 / (iters * max(1, reps - 1))
        if log:
            log(f"autotune: {form} {times[form]:.3f} ms/iteration")
    best = min((f for f in forms if times[f] is not None), key=lambda f: times[f])
    for s in solvers:
        s.set_stage_form(best)
    return {"form": best, "ms_per_iteration": times, "iterations": iters * reps * len(forms)}


def density_norm(solver: cc.Solver) -> float:
    """sqrt(sum res_rho^2 / noc) of the last stage's residual, reduced on
    the device."""
    return solver.residual_norm(1)


def run(case_list: list[Case], *, precision: str = "mixed", maxiter: int = 2000,
        minres: float = 1e-8, chunk: int = 50, fused: bool = True, tune: bool = False,
        log=None) -> dict:
    """Solve every case; returns per-case iterations and final density drop
    and the aggregate timing (setup excluded).  With ``tune`` the residual
    stage's form is picked by :func:`autotune` on the live batch first (its
    iterations are part of the solve and of the timing)."""
    t_setup = time.time()
    live = []
    for c in case_list:
        raw = meshgen.bump_channel(c.nx, c.ny, c.seed, h=c.h)
        m = cc.Mesh.from_raw(raw)
        cfg = cc.Config(mach=c.mach, aoa_deg=c.aoa_deg, flux="roe", grad="lsqr", ilimit=1,
                        precision=precision, cfl_type="ramp", cfl_min=0.8, cfl_max=0.8,
                        cfl_ramp_steps=1)
        s = cc.Solver(m, cfg)
        s.keep_residual(True)
        live.append(dict(case=c, mesh=m, solver=s, r1=None, drop=1.0))
    t_setup = time.time() - t_setup
    for e in live:
        e["solver"].sync()
    t0 = time.time()
    done = []
    # iteration 1, eagerly, then the normalisation of each case's density residual
    for e in live:
        e["solver"].iterate(1, hist=False)
    for e in live:
        e["r1"] = density_norm(e["solver"])
    for i, e in enumerate(live):
        e["slot"] = i
    fb = cc.Batch([e["solver"] for e in live]) if fused else None
    tuned = None
    if fb is not None and tune:
        tuned = autotune(fb, [e["solver"] for e in live], log=log)
    rounds = 0
    while live:
        if fb is not None:
            # every live case is at the same iteration
            fb.iterate(min(chunk, maxiter - live[0]["solver"].iter))
        else:
            for e in live:
                n = min(chunk, maxiter - e["solver"].iter)
                e["solver"].iterate_graph(n, hist=False)
        rounds += 1
        # density norms: one launch for all live cases, or one per case
        norms = fb.residual_norms(1) if fb is not None else None
        still = []
        for e in live:
            s = e["solver"]
            rn = norms[e["slot"]] if norms is not None else density_norm(s)
            e["drop"] = rn / e["r1"] if e["r1"] > 0 else 0.0
            if e["drop"] <= minres or s.iter >= maxiter:
                done.append(dict(k=e["case"].k, cells=e["case"].cells, iters=s.iter,
                                 drop=e["drop"]))
                if fb is not None:
                    fb.retire(e["slot"])
                s.close()
                e["mesh"].close()
            else:
                still.append(e)
        live = still
        if log:
            log(f"round {rounds}: {len(live)} cases running, {len(done)} done, "
                f"{time.time() - t0:.1f}s")
    if fb is not None:
        fb.close()
    wall = time.time() - t0
    iters = sum(d["iters"] * d["cells"] for d in done)
    return dict(wall=wall, setup=t_setup, solves=len(done), solves_per_s=len(done) / wall,
                cell_iterations=iters, cell_stages_per_s=3.0 * iters / wall,
                converged=sum(1 for d in done if d["drop"] <= minres), cases=done, tuned=tuned)
