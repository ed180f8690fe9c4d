"""The vertex-centred path's mixed-precision form against its FP64 form.

The mixed form keeps the vertex records (primitives as deviations from the
free stream, gradients, limiter values, coordinates), the edge geometry and
the least-squares weights in FP32 and runs the gradient and limiter in FP32;
reconstruction, fluxes, residual sums, spectral radii, time steps, the state,
the norms and the (multicolour) LU-SGS sweeps stay FP64.  It is build-defined
(the reference is FP64 only), so it is checked against the FP64 form on the
20 000-vertex O-grid of the vertex-centred workload's generator, LU-SGS at
CFL 50:
* with the limiter held at 0 (first order: nothing depends on the path) both
  converge to the same forces, CL and CD within 1e-6;
* second order, the reference's never-reset limiter ratchet makes the
  converged state depend on the path taken (the FP64 form itself lands on
  another CL when the CFL changes), so there the converged FP64 state must be
  a fixed point of the mixed iteration, and the early residual history must
  follow FP64's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mesh():
    from unsflow2d_amd import nc, nc_mesh
    return nc.NCMesh.from_raw(nc_mesh.o_grid(200, 100, 3, stretch=1.005 ** (999 / 99)))


def _solver(m, precision):
    from unsflow2d_amd import nc
    return nc.NCSolver(m, mach=0.4, cfl_max=50.0, grad="gg", timemode="lusgs",
                       lusgs_order="multicolour", precision=precision)


def test_nc_mixed_first_order_converges_to_fp64():
    print("case: nc mixed vs fp64, first order, o-grid 200x100 seed 3, 10000 iterations",
          flush=True)
    m = _mesh()
    nop = m.counts["nop"]
    out = {}
    for prec in ("fp64", "mixed"):
        s = _solver(m, prec)
        s.set("phi", np.zeros((nop, 4)))
        h = s.iterate(10000)
        f, cp = s.forces()
        out[prec] = (h[:, 0], f, cp)
        print(f"  {prec}: residual {h[-1, 0]:.3e} cl {f[0]:.12f} cd {f[1]:.12f}", flush=True)
        s.close()
    (h64, f64, cp64), (h32, f32, cp32) = out["fp64"], out["mixed"]
    assert h64[-1] < 1e-8 and h32[-1] < 1e-5
    assert abs(f32[0] - f64[0]) <= 1e-6
    assert abs(f32[1] - f64[1]) <= 1e-6
    assert np.max(np.abs(cp32 - cp64)) <= 1e-4


def test_nc_mixed_second_order_fixed_point_and_history():
    print("case: nc mixed vs fp64, second order, o-grid 200x100 seed 3", flush=True)
    m = _mesh()
    s64, s32 = _solver(m, "fp64"), _solver(m, "mixed")
    h64 = s64.iterate(200)[:, 0]
    h32 = s32.iterate(200)[:, 0]
    e = float(np.max(np.abs(h32 - h64) / h64))
    print(f"  residual history, 200 iterations: max rel {e:.2e}", flush=True)
    assert e <= 1e-3
    s64.iterate(7800, hist=False)
    f0, _ = s64.forces()
    s32.set("qc", s64.get("qc"))
    s32.set("phi", s64.get("phi"))
    s32.iter = s64.iter
    h = s32.iterate(500)[:, 0]
    f1, _ = s32.forces()
    print(f"  fp64 converged cl {f0[0]:.10f} cd {f0[1]:.10f}; 500 mixed iterations on: "
          f"cl {f1[0]:.10f} cd {f1[1]:.10f}, residual {h[-1]:.3e}", flush=True)
    assert abs(f1[0] - f0[0]) <= 1e-4 and abs(f1[1] - f0[1]) <= 1e-6
    assert h[-1] <= 1e-6
