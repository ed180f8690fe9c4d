"""The C-ABI library: loads without a GPU, exports every declared entry point,
and reports the reference's STOP conditions of the host preprocessing as
error codes."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "ufx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ufx_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    from unsflow2d_amd import cc, nc
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(cc.lib, n)]
    assert not missing, missing
    exported = set(cc.EXPORTED) | set(nc.NC_EXPORTED)
    assert sorted(cc.EXPORTED) == sorted(n for n in names if n in cc.EXPORTED)
    assert set(names) <= exported | {"ufx_mesh_size"}


def test_mesh_errors_map_to_reference_stops():
    from unsflow2d_amd import cc, meshgen
    raw = meshgen.bump_channel(6, 4, 1)
    # an interior face with a missing side (CCFVM/geometric.F90:13-16)
    bad = meshgen.RawMesh(**{k: getattr(raw, k).copy() for k in raw.__dataclass_fields__})
    bad.fint[3, 3] = 0
    with pytest.raises(cc.UfxError, match="in/out==0"):
        cc.Mesh.from_raw(bad)
    # `in` on the wrong side of an interior face: the reference's repair
    # leaves the face inconsistent and face_co_volume stops
    # (CCFVM/geometric.F90:90-102, :692)
    bad = meshgen.RawMesh(**{k: getattr(raw, k).copy() for k in raw.__dataclass_fields__})
    bad.fint[0, [2, 3]] = bad.fint[0, [3, 2]]
    with pytest.raises(cc.UfxError, match="-ve area"):
        cc.Mesh.from_raw(bad)


def test_mesh_import_roundtrip():
    """Arrays handed over as the reference holds them after its geometric
    give back the same mesh."""
    from unsflow2d_amd import cc, meshgen
    raw = meshgen.bump_channel(8, 5, 3)
    m = cc.Mesh.from_raw(raw)
    a = m.arrays()
    m2 = cc.Mesh.import_arrays(a, raw.nfi)
    b = m2.arrays()
    for n in cc.MESH_I32 + cc.MESH_F64:
        if n == "oldnum":
            continue
        assert np.array_equal(a[n], b[n]), n


def test_renumbering_is_breadth_first():
    """The cell numbering is the reference's breadth-first order from cell 1:
    every cell's lowest-numbered neighbour is numbered before it, and levels
    never decrease (CCFVM/geometric.F90:266-300)."""
    from unsflow2d_amd import cc, meshgen
    raw = meshgen.bump_channel(10, 7, 5)
    m = cc.Mesh.from_raw(raw)
    fin, fout = m.get("fc_in") - 1, m.get("fc_out") - 1
    noc = raw.noc
    nfi = raw.nfi
    nbr = [[] for _ in range(noc)]
    for f in range(nfi):
        nbr[fin[f]].append(fout[f])
        nbr[fout[f]].append(fin[f])
    level = np.full(noc, -1)
    level[0] = 0
    order = [0]
    for c in order:
        for n in nbr[c]:
            if level[n] < 0:
                level[n] = level[c] + 1
                order.append(n)
    assert sorted(order) == list(range(noc))
    assert np.all(np.diff(level) >= 0)


def test_batch_entry_points_reject_bad_arguments():
    """The batch entry points check their arguments before touching the
    device: no cases, a null batch, a bad chunk length or variable."""
    from unsflow2d_amd import cc
    lib = cc.lib
    out = ctypes.c_void_p()
    assert lib.ufx_cc_batch_create(None, 0, ctypes.byref(out)) == 1
    assert "n >= 1" in lib.ufx_last_error().decode()
    assert lib.ufx_cc_batch_iterate(None, 5, 0.8) == 1
    assert "null batch" in lib.ufx_last_error().decode()
    assert lib.ufx_cc_batch_set_active(None, None) == 1
    norms = np.zeros(2)
    assert lib.ufx_cc_batch_residual_norms(None, 1, norms.ctypes.data_as(cc._dp)) == 1
    assert lib.ufx_cc_batch_destroy(None) == 0
