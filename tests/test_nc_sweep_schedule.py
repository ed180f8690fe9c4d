"""Host replay of the reference-order LU-SGS level schedule (no GPU).

The reference sweeps the vertices in number order (NCFVM/lusgs.f90:35-124):
forward, vertex p uses the new dq of its neighbours o < p; backward, of its
neighbours o > p.  The library groups the vertices into dependency levels
(level(p) = 1 + the largest level of its swept neighbours), orders the
positions level by level (inside a level by swept-neighbour count, most
first, then by the sweep's vertex order) and lets one workgroup walk the
levels, keeping the dq of the last SW_RING positions in an LDS ring.  A
neighbour at position o is read from the ring when o >= lim = (end of the
reader's level) - SW_RING, otherwise from HBM.

The GPU test that compares the streamed sweep with the level launches on the
600 x 300 O-grid only covers the HBM side of that read if some neighbours do
lie beyond the ring; this replay counts them, so a larger ring that made the
test ring-only would fail here.
"""
import os
import re

import numpy as np

from conftest import ROOT


def _ring_size():
    with open(os.path.join(ROOT, "unsflow2d_amd", "csrc", "nc_kernels.h")) as f:
        text = f.read()
    return int(re.search(r"constexpr int SW_RING = (\d+);", text).group(1))


def schedule(nop, ptr, nbr):
    """Levels and positions of both sweeps, as the library builds them.

    ptr/nbr: CSR of the 0-based vertex neighbours.  Returns, per sweep, the
    position of each vertex, the end position of each vertex's level and the
    swept-neighbour pairs (vertex, neighbour)."""
    out = []
    for fwd in (True, False):
        lev = [0] * nop
        order = range(nop) if fwd else range(nop - 1, -1, -1)
        for p in order:
            lv = 0
            for o in nbr[ptr[p]:ptr[p + 1]]:
                if (o < p) if fwd else (o > p):
                    lv = max(lv, lev[o] + 1)
            lev[p] = lv
        lev = np.asarray(lev)
        src = np.repeat(np.arange(nop), np.diff(ptr))
        dst = np.asarray(nbr)
        swept = dst < src if fwd else dst > src
        cnt = np.bincount(src[swept], minlength=nop)
        # level, then count (most first), then the sweep's vertex order
        tie = np.arange(nop) if fwd else -np.arange(nop)
        verts = np.lexsort((tie, -cnt, lev))
        pos = np.empty(nop, dtype=np.int64)
        pos[verts] = np.arange(nop)
        ends = np.cumsum(np.bincount(lev))
        out.append((pos, ends[lev], src[swept], dst[swept], lev))
    return out


def _mesh_csr(ni, nj, seed, stretch):
    from unsflow2d_amd import nc, nc_mesh
    m = nc.NCMesh.from_raw(nc_mesh.o_grid(ni, nj, seed, stretch=stretch))
    nop = m.counts["nop"]
    ptr = np.concatenate(([0], np.cumsum(m.get("pt_nv2v"))))
    return nop, ptr, (m.get("pt_v2v") - 1).tolist()


def test_levels_respect_the_sweep_dependencies():
    nop, ptr, nbr = _mesh_csr(40, 16, 5, 1.1)
    for pos, end, src, dst, lev in schedule(nop, ptr, nbr):
        assert np.all(lev[dst] < lev[src])
        assert np.all(pos[dst] < pos[src])
        assert np.all(pos < end)


def test_large_o_grid_reads_beyond_the_ring_in_both_sweeps():
    ring = _ring_size()
    nop, ptr, nbr = _mesh_csr(600, 300, 7, 1.005 ** (999 / 299))
    for name, (pos, end, src, dst, _) in zip(("forward", "backward"),
                                             schedule(nop, ptr, nbr)):
        beyond = int(np.count_nonzero(pos[dst] < end[src] - ring))
        frac = beyond / len(src)
        print(f"{name}: {beyond} of {len(src)} neighbour reads beyond the "
              f"{ring}-position ring ({100 * frac:.2f} %)")
        assert beyond > 0
        assert frac < 0.05
