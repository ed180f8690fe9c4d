"""The cell-centred preprocessing with its per-face, per-cell and per-vertex
passes on the GPU (ufx_cc_geometric_device) against the host form
(ufx_cc_geometric, which equals the reference's `geometric` bit for bit on
every fixture): every array of the preprocessed mesh, bit for bit, on
triangles, quadrilaterals and mixed cells, on a mesh whose face orientations
the reference has to repair (CCFVM/geometric.F90:62-198), and on the
2 000 000-cell workload mesh."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _arrays(mesh):
    from unsflow2d_amd import cc
    out = {}
    for name in cc.MESH_I32 + cc.MESH_I32_WALK + cc.MESH_F64:
        out[name] = mesh.get(name)
    return out


def _same(raw, label):
    from unsflow2d_amd import cc
    t0 = time.perf_counter()
    host = cc.Mesh.from_raw(raw)
    t1 = time.perf_counter()
    dev = cc.Mesh.from_raw(raw, device=True)
    t2 = time.perf_counter()
    print(f"  {label}: host {t1 - t0:.3f}s, device passes {t2 - t1:.3f}s", flush=True)
    a, b = _arrays(host), _arrays(dev)
    for name in a:
        assert a[name].dtype == b[name].dtype, name
        assert np.array_equal(a[name], b[name]), f"{label}: {name} differs"


@pytest.mark.parametrize("cells", ["tri", "quad", "mixed"])
def test_device_preprocessing_equals_host(cells):
    from unsflow2d_amd import meshgen
    print(f"case: device preprocessing, {cells} cells, bump 40x16 seed 9", flush=True)
    _same(meshgen.bump_channel(40, 16, 9, h=0.08, cells=cells), cells)


def test_device_preprocessing_repairs_orientation_as_host():
    """Every fifth boundary face given with its end points swapped: the
    reference's cell_face_norm swaps them back and recomputes the normal, on
    the device as on the host (the ghost centres, mirrored in file order
    before the repair, differ from the unflipped mesh's in the last bits)."""
    from unsflow2d_amd import cc, meshgen
    print("case: device preprocessing, flipped boundary faces, bump 20x10 seed 3", flush=True)
    raw = meshgen.bump_channel(20, 10, 3, h=0.08)
    raw.fbnd[::5, [0, 1]] = raw.fbnd[::5, [1, 0]]
    _same(raw, "flipped")
    m = cc.Mesh.from_raw(raw, device=True)
    ref = cc.Mesh.from_raw(meshgen.bump_channel(20, 10, 3, h=0.08), device=True)
    for name in ("fc_p1", "fc_p2", "fc_nx", "fc_ny", "fc_area"):
        assert np.array_equal(m.get(name), ref.get(name)), name


def test_device_preprocessing_workload_mesh():
    from unsflow2d_amd import meshgen
    print("case: device preprocessing, bump 1000x1000 seed 2 (2M cells)", flush=True)
    _same(meshgen.bump_channel(1000, 1000, 2, h=0.08), "2M cells")


def test_device_preprocessing_reports_reference_errors():
    """Meshes the reference rejects, with its messages, on the device too: a
    cell centroid outside its diamond triangles, and interior faces given
    the wrong way round (the reference's repair swaps sides and end points,
    which leaves the normal pointing into `in`, and the co-volume check
    fails)."""
    from unsflow2d_amd import cc, meshgen
    bad_centre = meshgen.bump_channel(8, 4, 1, h=0.08)
    bad_centre.cen[5] = bad_centre.cen[5] + 10.0
    flipped = meshgen.bump_channel(20, 10, 3, h=0.08)
    flipped.fint[::3, [0, 1]] = flipped.fint[::3, [1, 0]]
    for raw in (bad_centre, flipped):
        msgs = []
        for device in (False, True):
            with pytest.raises(cc.UfxError) as ei:
                cc.Mesh.from_raw(raw, device=device)
            msgs.append(str(ei.value))
        print(f"  {msgs}", flush=True)
        assert msgs[0] == msgs[1] == "[mesh] geom,triarea,  -ve area"


def _nc_same(raw, label):
    from unsflow2d_amd import nc
    t0 = time.perf_counter()
    host = nc.NCMesh.from_raw(raw)
    t1 = time.perf_counter()
    dev = nc.NCMesh.from_raw(raw, device=True)
    t2 = time.perf_counter()
    print(f"  {label}: host {t1 - t0:.3f}s, device passes {t2 - t1:.3f}s", flush=True)
    for name in nc.NC_MESH_I32 + nc.NC_MESH_F64:
        a, b = host.get(name), dev.get(name)
        assert a.dtype == b.dtype, name
        assert np.array_equal(a, b), f"{label}: {name} differs"


@pytest.mark.parametrize("shape", [(24, 10, 3, 1.2), (200, 100, 5, 1.02), (2000, 1000, 3, 1.005)])
def test_nc_device_preprocessing_equals_host(shape):
    """The vertex-centred preprocessing (ufx_nc_geometric_device: dual-face
    normals, median-dual volumes, co-volume check, least-squares weights on
    the GPU) against the host form, up to the 2 000 000-vertex O-grid."""
    from unsflow2d_amd import nc_mesh
    ni, nj, seed, stretch = shape
    print(f"case: nc device preprocessing, o-grid {ni}x{nj} seed {seed}", flush=True)
    _nc_same(nc_mesh.o_grid(ni, nj, seed, stretch=stretch), f"o-grid {ni}x{nj}")


def test_nc_device_preprocessing_reports_reference_errors():
    """A cell centroid moved far out of place makes an edge's co-volume
    non-positive: the reference's STOP message names the same edge on the
    device as on the host."""
    from unsflow2d_amd import nc, nc_mesh
    raw = nc_mesh.o_grid(24, 10, 3, stretch=1.2)
    raw.cen[7] = raw.cen[7] + 50.0
    msgs = []
    for device in (False, True):
        with pytest.raises(nc.UfxError) as ei:
            nc.NCMesh.from_raw(raw, device=device)
        msgs.append(str(ei.value))
    print(f"  {msgs}", flush=True)
    assert msgs[0] == msgs[1] and "-ve face co-volume" in msgs[0]
