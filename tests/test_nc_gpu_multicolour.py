"""Device multicolour LU-SGS sweeps of the vertex-centred path.

With lusgs_order="multicolour" each colour of the mesh's greedy colouring
(ufx_mesh_get "pt_colour") is one launch, forward in ascending and reverse in
descending colour order.  That is the reference's sequential LU-SGS
(NCFVM/lusgs.f90:4-131) with the vertices numbered by colour: the oracle's
ordered sweep in that order (tests/test_multicolour.py pins it against the
reference-order sweep) must give the same iterates bit for bit -- the kernels
keep the reference's expression order without contraction.  Only the
residual norm is a parallel reduction in another order: 1e-12.  And one
operation is not exactly reproducible: pow in the farfield closure
(farfield_flux1), where the device's and the host C library's pow can differ
in the last bit; so the states must be equal bit for bit until the first
difference, which must involve farfield vertices and be a rounding
difference, after which they are held to 1e-12.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


class _Follow:
    """Bit-for-bit comparison of successive states until the first
    difference, which must involve farfield vertices (the pow of
    farfield_flux1) and be at rounding level; 1e-12 from then on."""

    def __init__(self, mesh):
        bc = mesh.get("fc_bc")
        far = np.zeros(len(mesh.get("pt_x")), bool)
        for k in ("fc_p1", "fc_p2"):
            far[mesh.get(k)[bc == 2001] - 1] = True
        self.far = far
        self.exact = True

    def check(self, name, got, want):
        g, w = np.asarray(got, np.float64), np.asarray(want, np.float64)
        e = relerr(g, w)
        bad = np.any(g.reshape(len(self.far), -1).view(np.int64)
                     != w.reshape(len(self.far), -1).view(np.int64), axis=1)
        print(f"  {name:12s} rel {e:.2e} differing vertices {int(bad.sum())}", flush=True)
        if self.exact and bad.any():
            # the sweep of the same iteration can carry it on to neighbours
            assert np.any(self.far[bad]), f"{name}: first difference not at the farfield"
            assert e <= 1e-15, f"{name}: first difference {e:.3e} beyond rounding"
            self.exact = False
        assert e <= 1e-12, f"{name}: {e:.3e}"


def _pair(ni, nj, seed, stretch, **kw):
    from unsflow2d_amd import nc, nc_mesh
    from oracle.oracle import NCOracle
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(ni, nj, seed, stretch=stretch))
    dev = nc.NCSolver(m, timemode="lusgs", lusgs_order="multicolour", **kw)
    orc = NCOracle(m.arrays(), **kw)
    col = m.get("pt_colour")
    order = np.lexsort((np.arange(len(col)), col)).astype(np.int32)
    return dev, orc, order, int(col.max()) + 1, _Follow(m)


@pytest.mark.parametrize("grad,flux,limiter", [("gg", "vanleer", "vkn"),
                                               ("lsqr", "vanleer", "vkn"),
                                               ("gg", "ausm", "vkn"), ("lsqr", "vanleer", "bj")])
def test_multicolour_equals_oracle_in_colour_order(grad, flux, limiter):
    dev, orc, order, nc_, fol = _pair(30, 12, 4, 1.2, mach=0.5, aoa_deg=1.5, cfl_max=50.0,
                                      grad=grad, flux=flux, limiter=limiter)
    print(f"case: nc multicolour {grad} {flux} {limiter} o-grid 30x12 seed 4, {nc_} colours",
          flush=True)
    n = 8 if limiter == "bj" else 15
    for it in range(1, n + 1):
        h = dev.iterate(1)[0, 0]
        f = orc.iteration_ordered(it, order)
        fol.check(f"qc it{it}", dev.get("qc"), orc.qc)
        assert abs(h - f) <= 1e-12 * abs(f)
    fol.check("phi", dev.get("phi"), orc.phi)


def test_multicolour_20k_vertices_equals_oracle():
    """The vertex-centred workload's generator at 200 x 100 points."""
    dev, orc, order, nc_, fol = _pair(200, 100, 3, 1.005 ** (999 / 99), mach=0.4, cfl_max=50.0,
                                      grad="gg")
    print(f"case: nc multicolour o-grid 200x100 seed 3, {nc_} colours", flush=True)
    for it in range(1, 11):
        dev.iterate(1)
        orc.iteration_ordered(it, order)
    fol.check("qc it10", dev.get("qc"), orc.qc)
    fol.check("res it10", dev.get("res"), orc.res)


def test_multicolour_iterations_keep_qold_for_readers():
    """Whole multicolour iterations leave the old state by vertex to be
    formed when something reads it: get("qold") is the state the last
    iteration started from, and a routine-by-routine RK stage after them
    (which reads qold) equals the same stage after an explicit save_old."""
    from unsflow2d_amd import nc, nc_mesh
    print("case: nc multicolour qold after whole iterations, o-grid 30x12 seed 4", flush=True)
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(30, 12, 4, stretch=1.2))
    kw = dict(mach=0.5, aoa_deg=1.5, cfl_max=50.0, grad="gg")
    a = nc.NCSolver(m, timemode="lusgs", lusgs_order="multicolour", **kw)
    b = nc.NCSolver(m, timemode="lusgs", lusgs_order="multicolour", **kw)
    a.iterate(4)
    b.iterate(4)
    start = a.get("qc")
    a.iterate(1)
    b.iterate(1)
    assert np.array_equal(a.get("qold"), start)
    # b: the stage straight after the iterations; a: after save_old of the
    # same state
    a.set("qc", start)
    b.set("qc", start)
    a.save_old()
    b.rk_stage(1, 1.0)
    a.rk_stage(1, 1.0)
    assert np.array_equal(a.get("qc"), b.get("qc"))


def test_fused_first_and_last_colour_equal_separate_launches(monkeypatch):
    """The first colour's forward update formed by the residual pass and the
    last colour's reverse update by its forward launch (neither has
    neighbours to sum) give the same iterates bit for bit as a launch per
    colour and direction (UFX_NC_FUSE_COLOURS=0)."""
    from unsflow2d_amd import nc, nc_mesh
    print("case: nc multicolour fused colour launches, o-grid 40x16 seed 5", flush=True)
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(40, 16, 5, stretch=1.15))
    kw = dict(timemode="lusgs", lusgs_order="multicolour", mach=0.45, cfl_max=50.0, grad="gg")
    a = nc.NCSolver(m, **kw)
    monkeypatch.setenv("UFX_NC_FUSE_COLOURS", "0")
    b = nc.NCSolver(m, **kw)
    ha, hb = a.iterate(12), b.iterate(12)
    assert np.array_equal(ha, hb)
    for name in ("qc", "res", "qold"):
        assert np.array_equal(a.get(name), b.get(name)), name
