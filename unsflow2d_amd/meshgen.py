"""Seeded synthetic meshes and the reference's text input formats.

The bump-channel generator produces the triangulated channel meshes the
cell-centred solver is exercised on: domain [0, L] x [0, H], ``nx`` x ``ny``
quads each split into two triangles along a seeded diagonal, interior vertices
jittered, and a Gaussian bump ``y_b = h * exp(-((x - xb)/w)^2)`` on the lower
wall.  The lower boundary is a slip wall (bc 1001), the other three sides are
characteristic farfield (bc 2001).

The output obeys every orientation contract of the reference's
``geometry.inp`` reader (CCFVM/read_input.F90:138-173, CCFVM/geometric.F90:62-198):

* interior face ``p1 p2 in out``: ``in`` lies to the LEFT of p1->p2, so the
  reference never "repairs" the face (its repair leaves the normal pointing into
  ``in``, CCFVM/geometric.F90:90-102);
* boundary face ``p1 p2 in nx ny bc``: ``in`` lies to the LEFT of p1->p2 and
  ``(nx, ny) = (y2 - y1, -(x2 - x1))`` exactly, which is the outward normal the
  reference keeps without recomputation (CCFVM/geometric.F90:116-152,185-195);
* cell centroid and area are written and used as given (area > 0).

Real numbers are written with 17 significant digits so that parsing the file
reproduces the in-memory doubles bit for bit; the solver input built from the
arrays and the reference run on the file therefore see identical values.
"""
from __future__ import annotations

import dataclasses
import numpy as np

WALL = 1001
FARFIELD = 2001


@dataclasses.dataclass
class RawMesh:
    """A mesh in the form of the reference's ``geometry.inp`` (1-based ids).

    Arrays follow CCFVM/read_input.F90:160-173: vertices, cells (centroid and
    area as given), interior faces ``p1 p2 in out`` and boundary faces
    ``p1 p2 in nx ny bc``.
    """
    x: np.ndarray            # (nop,) float64
    y: np.ndarray            # (nop,) float64
    pbc: np.ndarray          # (nop,) int32, vertex boundary tag (informational)
    cen: np.ndarray          # (noc, 2) float64
    area: np.ndarray         # (noc,) float64
    fint: np.ndarray         # (nfi, 4) int32: p1 p2 in out, 1-based
    fbnd: np.ndarray         # (nbf, 3) int32: p1 p2 in, 1-based
    fbn: np.ndarray          # (nbf, 2) float64: nx ny (unnormalised)
    fbc: np.ndarray          # (nbf,) int32: 1001 wall / 2001 farfield
    tri: np.ndarray          # (noc, k) int32, counter-clockwise vertex ids, 1-based (k = 3,
                             # or 4 with -1 padding for triangles among quadrilaterals)

    @property
    def nop(self) -> int:
        return int(self.x.shape[0])

    @property
    def noc(self) -> int:
        return int(self.cen.shape[0])

    @property
    def nfi(self) -> int:
        return int(self.fint.shape[0])

    @property
    def nbf(self) -> int:
        return int(self.fbnd.shape[0])


def _graded_eta(ny: int, first: float | None, growth: float | None) -> np.ndarray:
    """Wall-normal coordinate in [0, 1] for ``ny`` layers.

    Uniform when ``first`` is None.  Otherwise the first layer is ``first`` thick
    and each layer grows by ``growth`` until it reaches the spacing that fills
    the remaining height uniformly; from there the layers stay uniform (the
    growth is capped so the column closes exactly at 1).
    """
    if first is None:
        return np.linspace(0.0, 1.0, ny + 1)
    eta = [0.0]
    h = first
    k = 0
    while k < ny:
        rem = 1.0 - eta[-1]
        left = ny - k
        uni = rem / left
        if h >= uni:
            # uniform for the rest of the column
            for m in range(1, left + 1):
                eta.append(eta[-1] + uni if m < left else 1.0)
            break
        eta.append(eta[-1] + h)
        h *= growth
        k += 1
    eta = np.asarray(eta, dtype=np.float64)
    eta[-1] = 1.0
    return eta


def bump_channel(nx: int, ny: int, seed: int, *, h: float = 0.05,
                 length: float = 3.0, xb: float = 1.5, width: float = 0.2,
                 jitter: float = 0.2, jitter_y: bool = True,
                 first: float | None = None, growth: float | None = None,
                 cells: str = "tri") -> RawMesh:
    """Seeded jittered-triangle bump channel (see module docstring).

    ``first``/``growth`` grade the wall-normal spacing (anisotropic boundary
    layer); ``jitter_y=False`` jitters only in the streamwise direction.
    ``cells="quad"`` keeps the quadrilaterals whole, ``cells="mixed"`` splits
    a seeded half of them (the reference's solver takes any polygon; the
    triangulated channel is the default and the workloads').
    """
    rng = np.random.default_rng(seed)
    nxp, nyp = nx + 1, ny + 1
    dx = length / nx
    eta1 = _graded_eta(ny, first, growth)
    ii, jj = np.meshgrid(np.arange(nxp), np.arange(nyp), indexing="xy")
    ii = ii.ravel()
    jj = jj.ravel()
    xs = ii * dx
    eta = eta1[jj].copy()
    interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)
    jx = rng.uniform(-jitter, jitter, size=xs.shape)
    jy = rng.uniform(-jitter, jitter, size=xs.shape)
    if jitter_y:
        xs = np.where(interior, xs + jx * dx, xs)
        # local spacing = the smaller of the two adjacent layer thicknesses
        lo = np.diff(eta1)
        dloc = np.minimum(lo[np.clip(jj - 1, 0, ny - 1)], lo[np.clip(jj, 0, ny - 1)])
        eta = np.where(interior, eta + jy * dloc, eta)
    else:
        # Streamwise-only jitter moves each interior column as a whole (its
        # wall vertex slides along the bump), so columns stay straight and
        # layers far thinner than the jitter never invert.
        col = rng.uniform(-jitter, jitter, size=nxp)
        col[0] = col[nx] = 0.0
        xs = xs + col[ii] * dx
    yb = h * np.exp(-((xs - xb) / width) ** 2)
    ys = yb + eta * (1.0 - yb)
    # boundary tags for vertices (informational only; the reference ignores them)
    pbc = np.zeros(xs.shape, dtype=np.int32)
    pbc[jj == 0] = WALL
    pbc[(jj == ny) | (ii == 0) | (ii == nx)] = FARFIELD
    pbc[jj == 0] = WALL

    # triangles, counter-clockwise; quad corners a,b,c,d (0-based vertex ids)
    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    qi = qi.ravel()
    qj = qj.ravel()
    a = qj * nxp + qi
    b = a + 1
    c = a + nxp + 1
    d = a + nxp
    diag = rng.integers(0, 2, size=a.shape)  # 0: a-c diagonal, 1: b-d diagonal
    t0 = np.where(diag[:, None] == 0, np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t1 = np.where(diag[:, None] == 0, np.stack([a, c, d], 1), np.stack([b, c, d], 1))
    wall_of = lambda p, q: (jj[p] == 0) & (jj[q] == 0)
    if cells == "tri":
        tri = np.empty((2 * a.size, 3), dtype=np.int64)
        tri[0::2] = t0
        tri[1::2] = t1
        return _from_triangles(xs, ys, pbc, tri, wall_of=wall_of)
    if cells not in ("quad", "mixed"):
        raise ValueError(f"cells must be tri, quad or mixed, not {cells!r}")
    quad = np.stack([a, b, c, d], 1)
    split = (rng.integers(0, 2, size=a.shape) == 1) if cells == "mixed" else np.zeros(a.shape, bool)
    polys = []
    for k in range(a.size):
        if split[k]:
            polys.append(list(t0[k]))
            polys.append(list(t1[k]))
        else:
            polys.append(list(quad[k]))
    return _from_polygons(xs, ys, pbc, polys, wall_of=wall_of)


def _from_triangles(xs, ys, pbc, tri, wall_of) -> RawMesh:
    """Build faces from counter-clockwise triangles (0-based vertex ids).

    Faces are numbered in order of first appearance walking the triangles in
    order and their edges (v0v1, v1v2, v2v0); the first triangle to meet an
    edge is its ``in`` cell, with the edge oriented as that triangle traverses
    it, so ``in`` lies to the left.
    """
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
    first = order[start]                 # half-edge id of first appearance
    second = np.where(cnt == 2, order[np.minimum(start + 1, ks.size - 1)], -1)
    is_int = cnt == 2
    # interior faces in first-appearance order
    fi_first = first[is_int]
    fi_second = second[is_int]
    o = np.argsort(fi_first, kind="stable")
    fi_first, fi_second = fi_first[o], fi_second[o]
    fint = np.stack([ha[fi_first], hb[fi_first], hc[fi_first], hc[fi_second]], 1)
    fb_first = np.sort(first[~is_int])
    p1, p2 = ha[fb_first], hb[fb_first]
    fbnd = np.stack([p1, p2, hc[fb_first]], 1)
    fbn = np.stack([ys[p2] - ys[p1], -(xs[p2] - xs[p1])], 1)
    fbc = np.where(wall_of(p1, p2), WALL, FARFIELD).astype(np.int32)
    return RawMesh(x=np.ascontiguousarray(xs, dtype=np.float64),
                   y=np.ascontiguousarray(ys, dtype=np.float64),
                   pbc=pbc.astype(np.int32),
                   cen=np.ascontiguousarray(cen), area=np.ascontiguousarray(area),
                   fint=(fint + 1).astype(np.int32), fbnd=(fbnd + 1).astype(np.int32),
                   fbn=np.ascontiguousarray(fbn), fbc=fbc, tri=(tri + 1).astype(np.int32))


def _from_polygons(xs, ys, pbc, polys, wall_of) -> RawMesh:
    """Build faces from counter-clockwise polygons (lists of 0-based vertex
    ids), numbered as :func:`_from_triangles` numbers them: in order of
    first appearance walking the cells in order and their edges (v0v1, v1v2,
    ..., v(k-1)v0); the first cell to meet an edge is its ``in`` cell.  Area
    and centroid by the shoelace formula."""
    nc = len(polys)
    ha, hb, hc = [], [], []
    cen = np.empty((nc, 2))
    area = np.empty(nc)
    for c, pv in enumerate(polys):
        k = len(pv)
        px, py = xs[pv], ys[pv]
        qx, qy = np.roll(px, -1), np.roll(py, -1)
        cr = px * qy - qx * py
        a2 = cr.sum()
        area[c] = 0.5 * a2
        cen[c, 0] = ((px + qx) * cr).sum() / (3.0 * a2)
        cen[c, 1] = ((py + qy) * cr).sum() / (3.0 * a2)
        for i in range(k):
            ha.append(pv[i])
            hb.append(pv[(i + 1) % k])
            hc.append(c)
    if np.any(area <= 0):
        raise ValueError("generator produced a non-positive cell area")
    ha, hb, hc = np.array(ha), np.array(hb), np.array(hc)
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
    is_int = cnt == 2
    fi_first = first[is_int]
    fi_second = second[is_int]
    o = np.argsort(fi_first, kind="stable")
    fi_first, fi_second = fi_first[o], fi_second[o]
    fint = np.stack([ha[fi_first], hb[fi_first], hc[fi_first], hc[fi_second]], 1)
    fb_first = np.sort(first[~is_int])
    p1, p2 = ha[fb_first], hb[fb_first]
    fbnd = np.stack([p1, p2, hc[fb_first]], 1)
    fbn = np.stack([ys[p2] - ys[p1], -(xs[p2] - xs[p1])], 1)
    fbc = np.where(wall_of(p1, p2), WALL, FARFIELD).astype(np.int32)
    kmax = max(len(pv) for pv in polys)
    tri = np.full((nc, kmax), -1, dtype=np.int64)
    for c, pv in enumerate(polys):
        tri[c, :len(pv)] = pv
    return RawMesh(x=np.ascontiguousarray(xs, dtype=np.float64),
                   y=np.ascontiguousarray(ys, dtype=np.float64),
                   pbc=pbc.astype(np.int32),
                   cen=np.ascontiguousarray(cen), area=np.ascontiguousarray(area),
                   fint=(fint + 1).astype(np.int32), fbnd=(fbnd + 1).astype(np.int32),
                   fbn=np.ascontiguousarray(fbn), fbc=fbc,
                   tri=np.where(tri >= 0, tri + 1, -1).astype(np.int32))


def config_mesh(config: int, scale: float = 1.0, seed: int | None = None) -> RawMesh:
    """The bump-channel meshes of the benchmark workloads.

    config 1: 100x100 quads, h=0.05 (20 000 triangles);
    config 2: 1000x1000 quads, h=0.08 (2 000 000 triangles);
    config 4: 800x400 quads, first wall layer 1e-5 growing by 1.12 (capped),
              streamwise-only jitter (640 000 triangles).
    ``scale`` shrinks the quad counts for quick tests; the seed defaults to the
    config number.
    """
    seed = config if seed is None else seed
    if config == 1:
        n = max(2, int(round(100 * scale)))
        return bump_channel(n, n, seed, h=0.05)
    if config == 2:
        n = max(2, int(round(1000 * scale)))
        return bump_channel(n, n, seed, h=0.08)
    if config == 4:
        nxq = max(2, int(round(800 * scale)))
        nyq = max(2, int(round(400 * scale)))
        return bump_channel(nxq, nyq, seed, h=0.05, first=1e-5, growth=1.12, jitter_y=False)
    raise ValueError(f"no bump-channel mesh for config {config}")


def _fmt_rows(fmt: str, cols) -> str:
    rows = np.rec.fromarrays(cols)
    return "".join(fmt % tuple(r) for r in rows.tolist())


def write_geometry_inp(mesh: RawMesh, path: str) -> None:
    """Write the CCFVM ``geometry.inp`` (CCFVM/read_input.F90:138-173)."""
    g = "%.17g"
    with open(path, "w") as f:
        f.write(f"{mesh.nop} {mesh.noc} {mesh.nfi} {mesh.nbf}\n")
        ids = np.arange(1, mesh.nop + 1)
        f.write(_fmt_rows(f"%d {g} {g} %d\n", [ids, mesh.x, mesh.y, mesh.pbc]))
        ids = np.arange(1, mesh.noc + 1)
        f.write(_fmt_rows(f"%d {g} {g} {g}\n", [ids, mesh.cen[:, 0], mesh.cen[:, 1], mesh.area]))
        fi = mesh.fint
        xs, ys = mesh.x, mesh.y
        nxi = ys[fi[:, 1] - 1] - ys[fi[:, 0] - 1]
        nyi = -(xs[fi[:, 1] - 1] - xs[fi[:, 0] - 1])
        ids = np.arange(1, mesh.nfi + 1)
        f.write(_fmt_rows(f"%d %d %d %d %d {g} {g} %d\n",
                          [ids, fi[:, 0], fi[:, 1], fi[:, 2], fi[:, 3], nxi, nyi,
                           np.zeros(mesh.nfi, dtype=np.int64)]))
        fb = mesh.fbnd
        ids = np.arange(1, mesh.nbf + 1)
        f.write(_fmt_rows(f"%d %d %d %d {g} {g} %d\n",
                          [ids, fb[:, 0], fb[:, 1], fb[:, 2], mesh.fbn[:, 0], mesh.fbn[:, 1], mesh.fbc]))


def read_geometry_inp(path: str) -> RawMesh:
    """Parse a CCFVM ``geometry.inp`` back into a :class:`RawMesh`."""
    with open(path) as f:
        nop, noc, nfi, nbf = (int(t) for t in f.readline().split())
        pts = np.loadtxt(f, max_rows=nop, ndmin=2)
        cel = np.loadtxt(f, max_rows=noc, ndmin=2)
        fi = np.loadtxt(f, max_rows=nfi, ndmin=2)
        fb = np.loadtxt(f, max_rows=nbf, ndmin=2)
    return RawMesh(x=pts[:, 1].copy(), y=pts[:, 2].copy(), pbc=pts[:, 3].astype(np.int32),
                   cen=cel[:, 1:3].copy(), area=cel[:, 3].copy(),
                   fint=fi[:, 1:5].astype(np.int32), fbnd=fb[:, 1:4].astype(np.int32),
                   fbn=fb[:, 4:6].copy(), fbc=fb[:, 6].astype(np.int32),
                   tri=np.zeros((noc, 3), dtype=np.int32))


@dataclasses.dataclass
class Deck:
    """The CCFVM input deck (CCFVM/read_input.F90:31-50), label/value lines."""
    flow_type: str = "inviscid"
    iwall: int = 1
    t_wall: float = 300.0
    mach: float = 0.5
    aoa_deg: float = 0.0
    reynolds: float = 1000.0
    cfl_type: str = "ramp"
    cfl_min: float = 0.8
    cfl_max: float = 0.8
    cfl_ramp_steps: int = 1
    timemode: str = "rk3"
    grad_type: str = "lsqr"
    maxiter: int = 1000
    minres: float = 0.0
    flux_type: str = "roe"
    ilimit: int = 1
    irs: int = 0
    xref: float = 0.0
    yref: float = 0.0
    scrint: int = 100000000

    def text(self) -> str:
        g = "%.17g"
        return "\n".join([
            "Start     1",
            f"FlowType  {self.flow_type}  {self.iwall} {g % self.t_wall}",
            f"Mach      {g % self.mach}",
            f"AOA       {g % self.aoa_deg}",
            f"Reynolds  {g % self.reynolds}",
            f"CFL       {self.cfl_type}  {g % self.cfl_min}  {g % self.cfl_max} {self.cfl_ramp_steps}",
            f"Timemode  {self.timemode}",
            f"GradType  {self.grad_type}",
            "LastIter  0",
            f"MaxIter   {self.maxiter}",
            f"MinRes    {g % self.minres}",
            "SaveInt   100000000",
            f"ScrInt    {self.scrint}",
            "NoIso     30",
            f"Flux      {self.flux_type}",
            f"Limiter   {self.ilimit}",
            f"irs       {self.irs}",
            f"FarBC     0  {g % self.xref}  {g % self.yref}",
            'GridFile  "geometry.inp"',
            "",
        ])
