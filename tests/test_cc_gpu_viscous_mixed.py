"""The laminar channel in mixed precision against the FP64 device form.

The mixed form cannot start this case from the free stream: the
reference's iteration-1 vertex averages include ghost primitives that are
still zero (CCFVM/solution.F90:78-81), which puts negative densities and
pressures at wall vertices, and FP32 rounding then flips the sign of p/rho on
a few faces.  So both forms start from one developed FP64 state (state,
primitives with the ghosts', limiter values handed over) and run the same
iterations; the mixed form's wall skin friction, forces and residual history
are then compared with FP64's.  Mixed precision is build-defined (the
reference is FP64 only), so FP64 on the device, itself checked against the
oracle routine by routine, is the yardstick.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(mesh, precision):
    from unsflow2d_amd import cc
    cfg = cc.Config(mach=0.5, flow_type="laminar", reynolds=5000.0, iwall=1, flux="roe",
                    grad="ggfc", ilimit=1, irs=1, precision=precision, cfl_type="ramp",
                    cfl_min=0.8, cfl_max=0.8, cfl_ramp_steps=1)
    return cc.Solver(mesh, cfg)


def test_laminar_channel_mixed_tracks_fp64():
    """160 x 80 quads (25 600 triangles) with the laminar workload's wall
    layer (first cell 1e-3, growth 1.12, capped), M 0.5, Re 5000, adiabatic
    no-slip, GG-FC + limiter, Roe, RK3 with residual smoothing: 150 FP64
    iterations, then 300 more in each precision.  Wall Cf of the mixed form
    within 2e-3 of max |Cf|, CD within 1e-3 relative, and the residual
    history (relative to its first compared iteration) within 1e-3 of
    FP64's."""
    from unsflow2d_amd import cc, meshgen
    print("case: laminar channel 160x80 seed 4, FP64 150 it, then 300 it FP64 vs mixed", flush=True)
    mesh = cc.Mesh.from_raw(meshgen.bump_channel(160, 80, 4, h=0.05, first=1e-3, growth=1.12,
                                                 jitter_y=False))
    s64 = _solver(mesh, "fp64")
    s64.iterate(150, hist=False)
    s32 = _solver(mesh, "mixed")
    for name in ("qc", "qp", "phi"):
        s32.set(name, s64.get(name))
    s32.iter, s32.cfl = s64.iter, s64.cfl
    h64 = s64.iterate(300, hist=True)
    h32 = s32.iterate(300, hist=True)
    f64, f32 = s64.forces(wall=True), s32.forces(wall=True)
    cf_err = float(np.max(np.abs(f32["cf"] - f64["cf"])) / np.max(np.abs(f64["cf"])))
    cd_err = abs(f32["cd"] - f64["cd"]) / abs(f64["cd"])
    # the mixed context never ran iteration 1, which fixes the history's
    # normalisation (its entries are raw norms), so both histories are
    # compared relative to their first compared iteration
    r64, r32 = h64 / h64[0], h32 / h32[0]
    h_err = float(np.max(np.abs(r32 - r64) / np.abs(r64)))
    print(f"  Cf max err / max|Cf| {cf_err:.2e}, CD rel {cd_err:.2e}, history rel {h_err:.2e}, "
          f"residual over the 300 iterations x{r64[-1]:.3e}", flush=True)
    assert np.all(np.isfinite(h32)) and np.all(np.isfinite(f32["cf"]))
    assert cf_err <= 2e-3
    assert cd_err <= 1e-3
    assert h_err <= 1e-3
    s64.close()
    s32.close()
