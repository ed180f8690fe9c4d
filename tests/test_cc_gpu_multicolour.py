"""Device multicolour LU-SGS of the cell-centred path.

With lusgs_order="multicolour" the forward and backward sweeps of lusgs2
(CCFVM/lusgs.F90:207-379) take the cells colour by colour (ufx_mesh_get
"cell_colour"), one launch each.  That is lusgs2 with the cells numbered by
colour: the oracle's ordered sweep in that order (tests/test_multicolour.py
pins it against the reference-order sweep) must agree to the cell-centred
path's per-routine tolerance, 1e-12 per iteration, over a run.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grad,precision", [("gg", "fp64"), ("lsqr", "fp64")])
def test_cc_multicolour_equals_oracle_in_colour_order(grad, precision):
    from unsflow2d_amd import cc, meshgen
    from oracle.oracle import CCOracle
    m = cc.Mesh.from_raw(meshgen.bump_channel(40, 20, 8, h=0.08))
    col = m.get("cell_colour")
    order = np.lexsort((np.arange(len(col)), col)).astype(np.int32)
    print(f"case: cc multicolour lusgs {grad} bump 40x20 seed 8, {col.max() + 1} colours",
          flush=True)
    cfg = cc.Config(mach=0.5, grad=grad, timemode="lusgs", precision=precision,
                    cfl_type="ramp", cfl_min=1.0, cfl_max=5.0, cfl_ramp_steps=10,
                    lusgs_order="multicolour")
    s = cc.Solver(m, cfg)
    orc = CCOracle(m.arrays(), grad=grad)
    c = float("nan")
    for it in range(1, 16):
        c = cfg.cfl(it, c)
        h = s.iterate(1)[0]
        f = orc.iteration_ordered(it, c, order)
        e = relerr(s.get("qc"), orc.qc)
        print(f"  it {it:2d} qc rel {e:.2e} hist {h:.6e} {f:.6e}", flush=True)
        assert e <= 1e-12 * max(1, it) and abs(h - f) <= 1e-11 * abs(f)
