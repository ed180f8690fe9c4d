"""Seeded O-grid meshes for the vertex-centred (median-dual) solver and its
text input formats.

The O-grid surrounds the unit circle (a slip wall, bc 1001) with ``ni``
points around and ``nj`` rings out to the square [-R, R]^2 (characteristic
farfield, bc 2001).  Ring ``j`` blends the circle and the square along each
ray, ``t_j = (q^j - 1) / (q^(nj-1) - 1)`` with stretching ``q``; interior
points are jittered by up to ``jitter`` of the local radial and angular
spacing, and every quad is split along a seeded diagonal.

The file follows NCFVM/read_input.f90:101-112 (header ``nop noc nof``, all
primal edges ``id p1 p2 in out nx ny bc``) and the orientation contracts the
reference's median-dual construction relies on:

* every edge has ``in`` on the LEFT of p1->p2 and ``out`` on the right, 0 for
  a missing side (the dual-volume shoelace, NCFVM/geometric.f90:279-313);
* wall edges keep the fluid on the LEFT (``in = c, out = 0``): solid_flux uses
  the right half-normal (NCFVM/solid_flux.f90:31-32);
* farfield edges keep the fluid on the RIGHT (``in = 0, out = c``):
  farfield_flux1 uses the left half-normal (NCFVM/farfield_flux.f90:13-14).

``nx ny`` are written as the edge's left normal; the reader ignores them.
"""
from __future__ import annotations

import dataclasses

import numpy as np

WALL = 1001
FARFIELD = 2001


@dataclasses.dataclass
class NCRawMesh:
    """A mesh in the form of the vertex-centred ``geometry.inp`` (1-based)."""
    x: np.ndarray        # (nop,) float64
    y: np.ndarray        # (nop,) float64
    pbc: np.ndarray      # (nop,) int32 vertex tag: 0, 1001, 2001
    cen: np.ndarray      # (noc, 2) float64 triangle centroids
    area: np.ndarray     # (noc,) float64 triangle areas
    edge: np.ndarray     # (nof, 4) int32: p1 p2 in out (0 = no cell)
    ebc: np.ndarray      # (nof,) int32: 0 interior, 1001 wall, 2001 farfield
    tri: np.ndarray      # (noc, 3) int32 counter-clockwise vertex ids

    @property
    def nop(self) -> int:
        return int(self.x.shape[0])

    @property
    def noc(self) -> int:
        return int(self.cen.shape[0])

    @property
    def nof(self) -> int:
        return int(self.edge.shape[0])


def o_grid(ni: int, nj: int, seed: int, *, stretch: float = 1.005, outer: float = 20.0,
           jitter: float = 0.2) -> NCRawMesh:
    """O-grid around the unit circle, ``ni`` x ``nj`` points (see module doc)."""
    if ni < 3 or nj < 2:
        raise ValueError("o_grid needs ni >= 3 and nj >= 2")
    rng = np.random.default_rng(seed)
    th = 2.0 * np.pi * np.arange(ni) / ni
    c, s = np.cos(th), np.sin(th)
    sq = outer / np.maximum(np.abs(c), np.abs(s))          # ray to the square
    j = np.arange(nj, dtype=np.float64)
    t = (stretch ** j - 1.0) / (stretch ** (nj - 1) - 1.0) if stretch != 1.0 else j / (nj - 1)
    r = (1.0 - t)[:, None] * 1.0 + t[:, None] * sq[None, :]  # (nj, ni)
    # jitter interior rings along the ray and along the ring
    if nj > 2 and jitter > 0.0:
        dr = np.minimum(r[1:-1] - r[:-2], r[2:] - r[1:-1])
        rt = r[1:-1] * (2.0 * np.pi / ni)
        r[1:-1] = r[1:-1] + rng.uniform(-jitter, jitter, dr.shape) * dr
        dth = rng.uniform(-jitter, jitter, dr.shape) * rt / np.maximum(r[1:-1], 1e-300)
    else:
        dth = np.zeros((max(nj - 2, 0), ni))
    ang = np.broadcast_to(th, (nj, ni)).copy()
    ang[1:-1] = ang[1:-1] + dth
    xs = (r * np.cos(ang)).ravel()
    ys = (r * np.sin(ang)).ravel()
    # exact circle and square on the boundary rings
    xs[:ni], ys[:ni] = c, s
    xs[(nj - 1) * ni:], ys[(nj - 1) * ni:] = sq * c, sq * s
    diag = rng.integers(0, 2, size=(nj - 1, ni))
    jj, ii = np.meshgrid(np.arange(nj - 1), np.arange(ni), indexing="ij")
    a = jj * ni + ii
    b = jj * ni + (ii + 1) % ni
    cc = (jj + 1) * ni + (ii + 1) % ni
    d = (jj + 1) * ni + ii
    # (a, b, cc, d) runs clockwise (b follows a counter-clockwise around the
    # ring, d lies outward of a): triangles reversed; quads in ring-major
    # order, each split along its seeded diagonal
    t1 = np.where(diag[..., None] == 1, np.stack([a, cc, b], -1), np.stack([a, d, b], -1))
    t2 = np.where(diag[..., None] == 1, np.stack([a, d, cc], -1), np.stack([b, d, cc], -1))
    tri = np.stack([t1, t2], axis=2).reshape(-1, 3).astype(np.int64)
    pbc = np.zeros(nj * ni, dtype=np.int32)
    pbc[:ni] = WALL
    pbc[(nj - 1) * ni:] = FARFIELD
    return _from_triangles(xs, ys, pbc, tri)


def _from_triangles(xs, ys, pbc, tri) -> NCRawMesh:
    """Edges of counter-clockwise triangles (0-based ids), numbered by first
    appearance walking the triangles and their edges (v0v1, v1v2, v2v0),
    oriented as that first triangle traverses them (so it lies on the left).
    A boundary edge whose vertices are both on the wall ring is a wall edge;
    other boundary edges are farfield edges and are reversed so the fluid is
    on their right."""
    nc = tri.shape[0]
    x0, y0 = xs[tri[:, 0]], ys[tri[:, 0]]
    x1, y1 = xs[tri[:, 1]], ys[tri[:, 1]]
    x2, y2 = xs[tri[:, 2]], ys[tri[:, 2]]
    cen = np.stack([(x0 + x1 + x2) / 3.0, (y0 + y1 + y2) / 3.0], axis=1)
    area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    if np.any(area <= 0):
        raise ValueError("generator produced a non-positive triangle area")
    ha = np.stack([tri[:, 0], tri[:, 1], tri[:, 2]], 1).ravel()
    hb = np.stack([tri[:, 1], tri[:, 2], tri[:, 0]], 1).ravel()
    hc = np.repeat(np.arange(nc), 3)
    nop = xs.shape[0]
    key = np.minimum(ha, hb) * nop + np.maximum(ha, hb)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    cnt = np.diff(np.r_[start, ks.size])
    if np.any(cnt > 2):
        raise ValueError("non-manifold edge")
    first = order[start]
    second = np.where(cnt == 2, order[np.minimum(start + 1, ks.size - 1)], -1)
    o = np.argsort(first, kind="stable")
    first, second, cnt = first[o], second[o], cnt[o]
    p1, p2 = ha[first].copy(), hb[first].copy()
    cin = hc[first] + 1
    cout = np.where(cnt == 2, hc[np.maximum(second, 0)] + 1, 0)
    ebc = np.zeros(first.size, dtype=np.int32)
    bnd = cnt == 1
    wall = bnd & (pbc[p1] == WALL) & (pbc[p2] == WALL)
    far = bnd & ~wall
    ebc[wall] = WALL
    ebc[far] = FARFIELD
    # farfield: fluid on the right
    p1[far], p2[far] = hb[first[far]], ha[first[far]]
    cout[far] = cin[far]
    cin[far] = 0
    edge = np.stack([p1 + 1, p2 + 1, cin, cout], 1).astype(np.int32)
    return NCRawMesh(x=np.ascontiguousarray(xs, dtype=np.float64),
                     y=np.ascontiguousarray(ys, dtype=np.float64), pbc=pbc.astype(np.int32),
                     cen=np.ascontiguousarray(cen), area=np.ascontiguousarray(area),
                     edge=edge, ebc=ebc, tri=(tri + 1).astype(np.int32))


def config3_mesh(scale: float = 1.0, seed: int = 3) -> NCRawMesh:
    """The vertex-centred workload: 2000 x 1000 points around the unit circle,
    stretching 1.005, farfield square [-20, 20]^2 (``scale`` shrinks both
    counts, keeping the stretching over the whole radial range)."""
    ni = max(8, int(round(2000 * scale)))
    nj = max(3, int(round(1000 * scale)))
    q = 1.005 ** (999.0 / max(nj - 1, 1))
    return o_grid(ni, nj, seed, stretch=q)


def _fmt_rows(fmt: str, cols) -> str:
    rows = np.rec.fromarrays(cols)
    return "".join(fmt % tuple(r) for r in rows.tolist())


def write_geometry_inp(mesh: NCRawMesh, path: str) -> None:
    """Write the vertex-centred ``geometry.inp`` (NCFVM/read_input.f90:101-112)."""
    g = "%.17g"
    with open(path, "w") as f:
        f.write(f"{mesh.nop} {mesh.noc} {mesh.nof}\n")
        f.write(_fmt_rows(f"%d {g} {g} %d\n", [np.arange(1, mesh.nop + 1), mesh.x, mesh.y, mesh.pbc]))
        f.write(_fmt_rows(f"%d {g} {g} {g}\n", [np.arange(1, mesh.noc + 1), mesh.cen[:, 0],
                                                mesh.cen[:, 1], mesh.area]))
        e = mesh.edge
        xs, ys = mesh.x, mesh.y
        nx = ys[e[:, 1] - 1] - ys[e[:, 0] - 1]
        ny = -(xs[e[:, 1] - 1] - xs[e[:, 0] - 1])
        f.write(_fmt_rows(f"%d %d %d %d %d {g} {g} %d\n",
                          [np.arange(1, mesh.nof + 1), e[:, 0], e[:, 1], e[:, 2], e[:, 3], nx, ny,
                           mesh.ebc]))


def read_geometry_inp(path: str) -> NCRawMesh:
    """Parse a vertex-centred ``geometry.inp`` (NCFVM/read_input.f90:101-126)
    back into an :class:`NCRawMesh` (the triangles' vertex lists are not part
    of the file and are left zero)."""
    with open(path) as f:
        nop, noc, nof = (int(t) for t in f.readline().split()[:3])
        pts = np.loadtxt(f, max_rows=nop, ndmin=2)
        cel = np.loadtxt(f, max_rows=noc, ndmin=2)
        edg = np.loadtxt(f, max_rows=nof, ndmin=2)
    return NCRawMesh(x=pts[:, 1].copy(), y=pts[:, 2].copy(), pbc=pts[:, 3].astype(np.int32),
                     cen=cel[:, 1:3].copy(), area=cel[:, 3].copy(),
                     edge=edg[:, 1:5].astype(np.int32), ebc=edg[:, 7].astype(np.int32),
                     tri=np.zeros((noc, 3), dtype=np.int32))


@dataclasses.dataclass
class NCDeck:
    """The vertex-centred input deck (NCFVM/read_input.f90:23-40)."""
    mach: float = 0.4
    aoa_deg: float = 0.0
    cfl_max: float = 50.0
    timemode: str = "lusgs"
    maxiter: int = 500
    minres: float = 0.0
    ilimit: int = 1

    def text(self) -> str:
        g = "%.17g"
        return "\n".join([
            "Start     1",
            "FlowType  inviscid",
            f"Mach      {g % self.mach}",
            f"AOA       {g % self.aoa_deg}",
            "Reynolds  5000.0",
            f"CFL       {g % self.cfl_max}",
            f"Timemode  {self.timemode}",
            "GMRES     20 1 0.1",
            "LastIter  0",
            f"MaxIter   {self.maxiter}",
            f"MinRes    {g % self.minres}",
            "SaveInt   1000000",
            "ScrInt    1000000",
            "NoIso     30",
            "Flux      vanleer",
            f"Limiter   {self.ilimit}",
            "FarBC     0  0.0  0.0",
            'GridFile  "geometry.inp"',
            "",
        ])
