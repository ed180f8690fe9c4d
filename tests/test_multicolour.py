"""Multicolour LU-SGS: the colouring and the ordered oracle sweep.

The multicolour form runs the reference's LU-SGS (NCFVM/lusgs.f90:4-131,
CCFVM/lusgs.F90:207-379) with the unknowns ordered by colour instead of by
number.  Checked here without a GPU:
* the colourings the library hands out are proper (no edge or interior face
  joins two vertices / cells of one colour), use few colours, and follow the
  greedy rule in smallest-last order (vertices taken out one by one at least
  remaining degree, lowest number on ties; coloured in reverse, each taking
  the smallest colour its already coloured neighbours leave free);
* the oracle's ordered sweep with the identity order is the reference-order
  sweep bit for bit (that one is pinned by the reference's fixtures), and with
  a colour order it is invariant to the order inside each colour -- the
  property that lets one launch update a whole colour.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

NC_ARRAYS = ["pt_x", "pt_y", "pt_cv", "fc_p1", "fc_p2", "fc_bc", "fc_nx", "fc_ny", "fc_area",
             "pt_nv2v", "pt_v2v", "pt_wx", "pt_wy", "pt_nv2f", "pt_v2f"]


def nc_case(ni=24, nj=10, seed=3, stretch=1.2):
    from unsflow2d_amd import nc, nc_mesh
    mesh = nc.NCMesh.from_raw(nc_mesh.o_grid(ni, nj, seed, stretch=stretch))
    return mesh, {k: mesh.get(k) for k in NC_ARRAYS}


def greedy_reference(n, nbrs):
    """Smallest-last order with a heap of (remaining degree, vertex), then
    greedy colouring in the reverse of that order."""
    import heapq
    deg = [len(nbrs[v]) for v in range(n)]
    heap = [(deg[v], v) for v in range(n)]
    heapq.heapify(heap)
    out, order = [False] * n, []
    while heap:
        d, v = heapq.heappop(heap)
        if out[v] or d != deg[v]:
            continue
        out[v] = True
        order.append(v)
        for u in nbrs[v]:
            if not out[u]:
                deg[u] -= 1
                heapq.heappush(heap, (deg[u], u))
    col = [-1] * n
    for v in reversed(order):
        taken = {col[o] for o in nbrs[v] if col[o] >= 0}
        c = 0
        while c in taken:
            c += 1
        col[v] = c
    return np.array(col)


@pytest.mark.parametrize("shape", [(24, 10, 3, 1.2), (60, 30, 5, 1.05)])
def test_vertex_colouring_is_proper_and_greedy(shape):
    mesh, a = nc_case(*shape)
    col = mesh.get("pt_colour")
    p1, p2 = a["fc_p1"] - 1, a["fc_p2"] - 1
    assert np.all(col[p1] != col[p2])
    assert col.min() == 0 and col.max() + 1 <= 5
    ptr = np.r_[0, np.cumsum(a["pt_nv2v"])]
    nbrs = [a["pt_v2v"][ptr[v]:ptr[v + 1]] - 1 for v in range(len(col))]
    np.testing.assert_array_equal(col, greedy_reference(len(col), nbrs))


def test_cell_colouring_is_proper_and_greedy():
    from unsflow2d_amd import cc, meshgen
    m = cc.Mesh.from_raw(meshgen.bump_channel(20, 10, 8, h=0.08))
    a = m.arrays()
    col = m.get("cell_colour")
    noc = len(a["cell_cv"])
    nfi = len(a["fc_p1"]) - (len(a["cell_cenx"]) - noc)
    fin, fout = a["fc_in"][:nfi] - 1, a["fc_out"][:nfi] - 1
    assert len(col) == noc
    assert np.all(col[fin] != col[fout])
    assert col.max() + 1 <= 4  # triangles: at most three neighbours each
    nbrs = [[] for _ in range(noc)]
    for i, o in zip(fin, fout):
        nbrs[i].append(o)
        nbrs[o].append(i)
    np.testing.assert_array_equal(col, greedy_reference(noc, nbrs))


def _run(orc, order, niter):
    hist = []
    for it in range(1, niter + 1):
        hist.append(orc.iteration(it) if order is None else orc.iteration_ordered(it, order))
    return np.array(hist)


def test_ordered_sweep_identity_is_the_reference_sweep():
    from oracle.oracle import NCOracle
    _, a = nc_case()
    x, y = NCOracle(a), NCOracle(a)
    hx = _run(x, None, 30)
    hy = _run(y, np.arange(len(a["pt_x"]), dtype=np.int32), 30)
    np.testing.assert_array_equal(hx, hy)
    np.testing.assert_array_equal(x.qc, y.qc)


def test_colour_order_is_independent_of_order_inside_a_colour():
    """Reversing the vertices inside every colour leaves the iterate
    unchanged bit for bit (each colour's vertices read only other colours),
    while the colour order itself gives another iterate than the numbering."""
    from oracle.oracle import NCOracle
    mesh, a = nc_case()
    col = mesh.get("pt_colour")
    n = len(col)
    up = np.lexsort((np.arange(n), col)).astype(np.int32)
    down = np.lexsort((-np.arange(n), col)).astype(np.int32)
    x, y, r = NCOracle(a), NCOracle(a), NCOracle(a)
    hx, hy, hr = _run(x, up, 20), _run(y, down, 20), _run(r, None, 20)
    np.testing.assert_array_equal(hx, hy)
    np.testing.assert_array_equal(x.qc, y.qc)
    assert not np.array_equal(x.qc, r.qc)
    # the same scheme all the same: the first iteration's residual agrees
    assert hx[0] == hr[0] == 1.0
    assert abs(hx[1] - hr[1]) < 1e-3 * hr[1]


def _cc_case():
    from unsflow2d_amd import cc, meshgen
    m = cc.Mesh.from_raw(meshgen.bump_channel(20, 10, 8, h=0.08))
    return m, m.arrays()


def test_cc_ordered_sweep_identity_is_the_reference_sweep():
    """lusgs2 in the identity order is the reference-order sweep (pinned by
    the cc_lusgs_gg fixture) bit for bit."""
    from oracle.oracle import CCOracle
    m, a = _cc_case()
    x, y = CCOracle(a, grad="gg"), CCOracle(a, grad="gg")
    ident = np.arange(x.noc, dtype=np.int32)
    for it in range(1, 11):
        hx = x.iteration(it, 5.0, "lusgs")
        hy = y.iteration_ordered(it, 5.0, ident)
        assert hx == hy
    np.testing.assert_array_equal(x.qc, y.qc)


def test_cc_colour_order_is_independent_of_order_inside_a_colour():
    from oracle.oracle import CCOracle
    m, a = _cc_case()
    col = m.get("cell_colour")
    n = len(col)
    up = np.lexsort((np.arange(n), col)).astype(np.int32)
    down = np.lexsort((-np.arange(n), col)).astype(np.int32)
    x, y, r = CCOracle(a, grad="gg"), CCOracle(a, grad="gg"), CCOracle(a, grad="gg")
    for it in range(1, 11):
        x.iteration_ordered(it, 5.0, up)
        y.iteration_ordered(it, 5.0, down)
        r.iteration(it, 5.0, "lusgs")
    np.testing.assert_array_equal(x.qc, y.qc)
    assert not np.array_equal(x.qc, r.qc)
