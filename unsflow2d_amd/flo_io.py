"""The files around the solvers, in the reference's formats.

Host-side I/O only (nothing here touches the GPU): the input decks of both
solvers, ``geometry.inp`` (see :mod:`meshgen` and :mod:`nc_mesh`), the
residual history ``FLO.RES``, the restart file ``FLO.DAT``, and the reports
``XYCpCf.dat`` and ``tecplt.dat``.

Fixed formats (``(i6,2e16.6)``, ``(4e20.10)``, ``(1x,7(f15.6,2x))``) are
written character for character as a Fortran compiler writes them.
List-directed records (``write(u,*)``) have a compiler-defined layout; they
are written here as blank-separated fields with 17 significant digits, which
any list-directed read (and the reference's own ``read(u,*)``) accepts.
"""
from __future__ import annotations

import math
import shlex

import numpy as np

from . import meshgen, nc_mesh

GAMMA = 1.4


# ------------------------------------------------------------ number formats
def fortran_e(v: float, w: int, d: int) -> str:
    """Fortran ``Ew.d``: ``0.ddddddE+xx`` (``+xxx`` without the E beyond
    two exponent digits), right-justified in ``w`` columns, ``*`` fill when it
    does not fit."""
    v = float(v)
    if math.isnan(v) or math.isinf(v):
        s = "NaN" if math.isnan(v) else ("-Infinity" if v < 0 else "Infinity")
        return s.rjust(w) if len(s) <= w else "*" * w
    if v == 0.0:
        body, sign = "0." + "0" * d + "E+00", "-" if math.copysign(1.0, v) < 0 else ""
    else:
        m, e = f"{abs(v):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        sign = "-" if v < 0 else ""
        if abs(ex) <= 99:
            body = f"0.{digits}E{'+' if ex >= 0 else '-'}{abs(ex):02d}"
        else:
            body = f"0.{digits}{'+' if ex >= 0 else '-'}{abs(ex):03d}"
    s = sign + body
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_f(v: float, w: int, d: int) -> str:
    """Fortran ``Fw.d``, right-justified, ``*`` fill when it does not fit."""
    s = f"{float(v):.{d}f}"
    return s.rjust(w) if len(s) <= w else "*" * w


def _ld(v) -> str:
    """A list-directed real field."""
    return f"{float(v):25.16E}"


# ------------------------------------------------------------------- decks
def _records(path: str) -> list[list[str]]:
    """One list-directed record per line: blank/comma separated values,
    quoted strings unquoted (the reference reads each line with
    ``read(inp,*) label, values...``)."""
    out = []
    with open(path) as f:
        for line in f:
            if not line.strip():
                continue
            lex = shlex.shlex(line.replace(",", " "), posix=True)
            lex.whitespace_split = True
            lex.commenters = ""
            out.append(list(lex))
    return out


def read_cc_deck(path: str) -> tuple[meshgen.Deck, dict]:
    """The cell-centred deck, CCFVM/read_input.F90:31-50 (nineteen records,
    each a label then its values), with the option checks of :52-120.
    Returns the deck and the fields the device path does not use (start
    option, iteration counters and intervals, vortex flag, grid file)."""
    r = _records(path)
    if len(r) < 19:
        raise ValueError(f"{path}: expected 19 records, found {len(r)}")
    d = meshgen.Deck(
        flow_type=r[1][1], iwall=int(r[1][2]), t_wall=float(r[1][3]),
        mach=float(r[2][1]), aoa_deg=float(r[3][1]), reynolds=float(r[4][1]),
        cfl_type=r[5][1], cfl_min=float(r[5][2]), cfl_max=float(r[5][3]),
        cfl_ramp_steps=int(r[5][4]), timemode=r[6][1], grad_type=r[7][1],
        maxiter=int(r[9][1]), minres=float(r[10][1]), flux_type=r[14][1],
        ilimit=int(r[15][1]), irs=int(r[16][1]), xref=float(r[17][2]), yref=float(r[17][3]))
    extra = dict(istart=int(r[0][1]), iterlast=int(r[8][1]), saveinterval=int(r[11][1]),
                 scrinterval=int(r[12][1]), niso=int(r[13][1]), vortex=int(r[17][1]),
                 gridfile=r[18][1])
    bad = []
    if extra["istart"] not in (1, 2):
        bad.append(f"start option {extra['istart']} (1 scratch, 2 restart)")
    if d.flow_type not in ("inviscid", "laminar", "rans"):
        bad.append(f"flow type {d.flow_type}")
    if d.iwall not in (0, 1):
        bad.append(f"wall condition {d.iwall}")
    if d.timemode not in ("rk3", "lusgs", "gmres"):
        bad.append(f"time scheme {d.timemode}")
    if d.cfl_type not in ("tanh", "ramp"):
        bad.append(f"CFL type {d.cfl_type}")
    if d.grad_type not in ("gg", "lsqr", "ggfc"):
        bad.append(f"gradient {d.grad_type}")
    if d.flux_type not in ("roe", "ausm", "vanleer", "rusanov"):
        bad.append(f"flux {d.flux_type}")
    if d.irs not in (0, 1):
        bad.append(f"residual smoothing option {d.irs}")
    if bad:
        raise ValueError(f"{path}: unknown " + "; ".join(bad))
    return d, extra


def read_nc_deck(path: str) -> tuple[nc_mesh.NCDeck, dict]:
    """The vertex-centred deck, NCFVM/read_input.f90:23-40 (eighteen records),
    with the checks of :44-86."""
    r = _records(path)
    if len(r) < 18:
        raise ValueError(f"{path}: expected 18 records, found {len(r)}")
    d = nc_mesh.NCDeck(mach=float(r[2][1]), aoa_deg=float(r[3][1]), cfl_max=float(r[5][1]),
                       timemode=r[6][1], maxiter=int(r[9][1]), minres=float(r[10][1]),
                       ilimit=int(r[15][1]))
    extra = dict(istart=int(r[0][1]), flow_type=r[1][1], reynolds=float(r[4][1]),
                 gmres=(int(r[7][1]), int(r[7][2]), float(r[7][3])), iterlast=int(r[8][1]),
                 saveinterval=int(r[11][1]), scrinterval=int(r[12][1]), niso=int(r[13][1]),
                 flux_type=r[14][1], vortex=int(r[16][1]), gridfile=r[17][1])
    bad = []
    if extra["istart"] not in (1, 2):
        bad.append(f"start option {extra['istart']}")
    if extra["flow_type"] not in ("inviscid", "laminar", "rans"):
        bad.append(f"flow type {extra['flow_type']}")
    if d.timemode not in ("rk3", "lusgs", "gmres"):
        bad.append(f"time scheme {d.timemode}")
    if d.ilimit not in (0, 1):
        bad.append(f"limiter option {d.ilimit}")
    if extra["vortex"] not in (0, 1):
        bad.append(f"vortex option {extra['vortex']}")
    if bad:
        raise ValueError(f"{path}: unknown " + "; ".join(bad))
    if d.timemode == "rk3" and d.cfl_max > 1.0:
        # NCFVM/main.f90:21-23 stops here
        raise ValueError(f"{path}: CFL should be less than 1 for rk3")
    return d, extra


# ------------------------------------------------------ history and restart
def flo_res_lines(iters, log_fres, cfl, log_fresi=None) -> str:
    """``FLO.RES``: ``(i6,2e16.6)`` iter, log10 L2, CFL
    (CCFVM/main.F90:114), or with ``log_fresi`` the vertex-centred
    ``(i6,3e16.6)`` iter, log10 L2, log10 Linf, CFL (NCFVM/main.f90:86)."""
    out = []
    for k, it in enumerate(iters):
        cols = [log_fres[k]] + ([log_fresi[k]] if log_fresi is not None else []) + [cfl[k]]
        out.append(f"{int(it):6d}" + "".join(fortran_e(c, 16, 6) for c in cols) + "\n")
    return "".join(out)


def write_flo_res(path: str, iters, log_fres, cfl, log_fresi=None) -> None:
    with open(path, "w") as f:
        f.write(flo_res_lines(iters, log_fres, cfl, log_fresi))


def save_flow(path: str, qc) -> None:
    """``FLO.DAT``: the conserved state, one ``(4e20.10)`` record per cell
    (CCFVM/solution.F90:88-100) or vertex (NCFVM/common.f90:58-70)."""
    q = np.asarray(qc, dtype=np.float64).reshape(-1, 4)
    with open(path, "w") as f:
        f.write("".join("".join(fortran_e(v, 20, 10) for v in row) + "\n" for row in q.tolist()))


def read_flow(path: str, n: int) -> np.ndarray:
    """Read ``n`` records of ``FLO.DAT`` (CCFVM/solution.F90:104-114)."""
    with open(path) as f:
        tok = f.read().replace("D", "E").replace("d", "e").split()
    if len(tok) < 4 * n:
        raise ValueError(f"{path}: expected {n} x 4 values, found {len(tok)}")
    return np.array(tok[:4 * n], dtype=np.float64).reshape(n, 4)


# ----------------------------------------------------------------- reports
def write_xycpcf(path: str, x, y, mcp, cf) -> None:
    """``XYCpCf.dat``: per wall face x, y, -Cp, Cf (CCFVM/main.F90:290-342,
    list-directed)."""
    with open(path, "w") as f:
        for a, b, c, d in zip(np.ravel(x), np.ravel(y), np.ravel(mcp), np.ravel(cf)):
            f.write(" " + _ld(a) + _ld(b) + _ld(c) + _ld(d) + "\n")


def _f15_rows(vals) -> str:
    """``(1x,7(f15.6,2x))`` with format reversion: seven fields on the
    first line, the rest on continuation lines without the leading 1x."""
    out = []
    for k in range(0, len(vals), 7):
        lead = " " if k == 0 else ""
        out.append(lead + "".join(fortran_f(v, 15, 6) + "  " for v in vals[k:k + 7]) + "\n")
    return "".join(out)


def write_tecplot_cc(path: str, x, y, prim, mu, nc2v, c2v, ent_inf: float) -> None:
    """``tecplt.dat`` of the cell-centred solver, CCFVM/main.F90:417-480: the
    vertex averages (x, y, density, u, v, Mach, pressure, temperature,
    entropy, |V|, mu) and the cells' vertex walks (triangles written with
    their third vertex repeated, ET=quadrilateral)."""
    prim = np.asarray(prim, dtype=np.float64)
    nop, noc = len(x), len(nc2v)
    with open(path, "w") as f:
        f.write(' TITLE = "flo2d output" \n')
        f.write(' VARIABLES="X","Y","density","u-velocity","v-velocity","Mach","Pressure",'
                '"Temprature","Entropy","Velocity" ,"mu"\n')
        f.write(" ZONE F=FEPOINT,ET=quadrilateral\n")
        f.write(f" N={nop:12d} ,E={noc:12d}\n")
        for i in range(nop):
            ro, uo, vo, po, to = prim[i, :5]
            q2 = uo ** 2 + vo ** 2
            mach = math.sqrt(q2 * ro / (GAMMA * po))
            entropy = math.log10(po / ro ** GAMMA / ent_inf)
            vmag = math.sqrt(uo * uo + vo * vo)
            f.write(_f15_rows([x[i], y[i], ro, uo, vo, mach, po, to, entropy, vmag, mu[i]]))
        ptr = np.r_[0, np.cumsum(nc2v)]
        for i in range(noc):
            w = c2v[ptr[i]:ptr[i + 1]]
            nf = len(w) - 1  # the walk is closed: faces = vertices - 1
            ids = list(w[:3]) + [w[2]] if nf == 3 else list(w[:nf])
            f.write(" " + "".join(f"{int(v):12d}" for v in ids) + "\n")


def write_tecplot_nc(path: str, x, y, qp, tri) -> None:
    """``tecplt.dat`` of the vertex-centred solver, NCFVM/main.f90:190-262:
    x, y, rho, u, v, p, Mach per vertex and the triangles' vertex walks."""
    qp = np.asarray(qp, dtype=np.float64)
    nop, noc = len(x), len(tri)
    with open(path, "w") as f:
        f.write(' TITLE = "flo2d output" \n')
        f.write(' VARIABLES="X","Y","rho","u","v","p","M"\n')
        f.write(" ZONE F=FEPOINT,ET=TRIANGLE\n")
        f.write(f" N={nop:12d} ,E={noc:12d}\n")
        for i in range(nop):
            ro, uo, vo, po = qp[i, :4]
            mach = math.sqrt((uo ** 2 + vo ** 2) * ro / (GAMMA * po))
            f.write(_f15_rows([x[i], y[i], ro, uo, vo, po, mach]))
        for t in np.asarray(tri):
            f.write(" " + "".join(f"{int(v):12d}" for v in t) + "\n")


def nc_cell_walks(a: dict) -> np.ndarray:
    """The vertex walk of every triangle of a vertex-centred mesh (1-based),
    NCFVM/geometric.f90:587-645: start at the cell's lowest-numbered edge
    (cell_2_face appends edges in ascending order, :393-410), oriented so the
    centroid lies to its left (dotprod < 0, NCFVM/common.f90:337-358), then
    the remaining vertex.  ``a`` holds the mesh arrays of :class:`nc.NCMesh`
    (post-numbering) and its cell centroids."""
    p1 = np.asarray(a["fc_p1"], dtype=np.int64)
    p2 = np.asarray(a["fc_p2"], dtype=np.int64)
    fin = np.asarray(a["fc_in"], dtype=np.int64)
    fout = np.asarray(a["fc_out"], dtype=np.int64)
    noc = len(a["cell_cenx"])
    e = np.arange(len(p1))
    cells = np.concatenate([fin, fout])
    edges = np.concatenate([e, e])
    keep = cells > 0
    cells, edges = cells[keep] - 1, edges[keep]
    first = np.full(noc, np.iinfo(np.int64).max)
    np.minimum.at(first, cells, edges)
    vsum = np.zeros(noc, dtype=np.int64)
    np.add.at(vsum, cells, p1[edges] + p2[edges])
    a0, b0 = p1[first], p2[first]
    third = vsum // 2 - a0 - b0
    x, y = np.asarray(a["pt_x"]), np.asarray(a["pt_y"])
    xc, yc = np.asarray(a["cell_cenx"]), np.asarray(a["cell_ceny"])
    x1, y1, x2, y2 = x[a0 - 1], y[a0 - 1], x[b0 - 1], y[b0 - 1]
    dx, dy = x2 - x1, y2 - y1
    ds = np.sqrt(dx * dx + dy * dy)
    nx, ny = dy / ds, -dx / ds
    rx, ry = xc - x1, yc - y1
    rs = np.sqrt(rx * rx + ry * ry)
    dot = (rx / rs) * nx + (ry / rs) * ny
    s1 = np.where(dot < 0.0, a0, b0)
    s2 = np.where(dot < 0.0, b0, a0)
    return np.stack([s1, s2, third], axis=1)
