// Host-side preprocessing of a cell-centred mesh into the state the hot path
// consumes: the restatement of the reference's read_grid ghost mirror
// (CCFVM/read_input.F90:179-203) and of geometric (CCFVM/geometric.F90:1-57).
//
// The result fixes everything the per-iteration kernels depend on for parity:
// the breadth-first cell numbering (renumber, :247-359), the face order and
// orientation (cell_face_norm, :62-198), the faces-of-a-cell order used by
// LU-SGS (cell_2_face + the reordering walk in cell_2_vertex, :425-616), the
// vertex->cell lists and pseudo-Laplacian weights (vertex_2_cell :459-531,
// pseudo_Laplacian :962-1013), least-squares factors (LSQR_Coeff :898-960) and
// diamond-path co-volumes (face_co_volume :619-674).  Floating-point
// expressions keep the reference's operation order so the arrays agree with
// the reference bit for bit (compiled without FMA contraction).
#include "ufx_internal.h"

#include <cmath>
#include <cstring>
#include <string>

namespace ufx {

namespace {

// Sign test of CCFVM/solution.F90:363-384 (dotprod): unit right-normal of
// n1->n2 dotted with the unit vector from n1 to (xc, yc).
double dotprod(const Mesh& m, int n1, int n2, double xc, double yc) {
  double x1 = m.px[n1], y1 = m.py[n1];
  double x2 = m.px[n2], y2 = m.py[n2];
  double dx = x2 - x1, dy = y2 - y1;
  double ds = std::sqrt(dx * dx + dy * dy);
  double nx = dy / ds, ny = -dx / ds;
  dx = xc - x1;
  dy = yc - y1;
  ds = std::sqrt(dx * dx + dy * dy);
  double rx = dx / ds, ry = dy / ds;
  return rx * nx + ry * ny;
}

// CCFVM/geometric.F90:676-693 (triarea), shoelace with the reference's order.
double triarea(const double* x, const double* y) {
  double area = 0.0;
  for (int i = 0; i < 3; ++i) {
    int ip1 = (i == 2) ? 0 : i + 1;
    double dx = 0.5 * (x[ip1] + x[i]);
    double dy = y[ip1] - y[i];
    area = area + dx * dy;
  }
  return area;
}

void push_unique(std::vector<int>& v, int val) {
  for (int e : v)
    if (e == val) return;
  v.push_back(val);
}

}  // namespace

// The input arrays, checked, 0-based, ghosts attached (read_grid,
// CCFVM/read_input.F90:179-203).
int cc_geom_read(const RawInput& in, Mesh& m, std::string& err) {
  const int nop = in.nop, noc = in.noc, nfi = in.nfi, nbf = in.nbf;
  if (nop <= 0 || noc <= 0 || nfi < 0 || nbf < 0) {
    err = "invalid mesh counts";
    return UFX_ERR_ARG;
  }
  const int nof = nfi + nbf, ncell = noc + nbf;
  m = Mesh();
  m.kind = UFX_KIND_CC;
  m.nop = nop; m.noc = noc; m.nfi = nfi; m.nbf = nbf;
  m.px.assign(in.px, in.px + nop);
  m.py.assign(in.py, in.py + nop);
  m.ccx.assign(ncell, 0.0);
  m.ccy.assign(ncell, 0.0);
  m.ccv.assign(noc, 0.0);
  for (int i = 0; i < noc; ++i) {
    m.ccx[i] = in.ccx[i];
    m.ccy[i] = in.ccy[i];
    m.ccv[i] = in.cv[i];
  }
  m.fp1.resize(nof); m.fp2.resize(nof); m.fin.resize(nof); m.fout.resize(nof);
  m.fbc.resize(nof);
  m.fnx.assign(nof, 0.0); m.fny.assign(nof, 0.0); m.farea.assign(nof, 0.0);
  m.fcx.assign(nof, 0.0); m.fcy.assign(nof, 0.0); m.fcov.assign(nof, 0.0);
  auto chk = [&](int v, int lo, int hi, const char* what) {
    if (v < lo || v > hi) {
      err = std::string("index out of range in ") + what;
      return false;
    }
    return true;
  };
  for (int i = 0; i < nfi; ++i) {
    int p1 = in.fi_p1[i] - 1, p2 = in.fi_p2[i] - 1;
    int ci = in.fi_in[i], co = in.fi_out[i];
    // CCFVM/geometric.F90:9-17: interior faces need both sides
    if (ci == 0 || co == 0) { err = "Inside faces in/out==0"; return UFX_ERR_MESH; }
    if (!chk(p1, 0, nop - 1, "interior face p1") || !chk(p2, 0, nop - 1, "interior face p2") ||
        !chk(ci, 1, noc, "interior face in") || !chk(co, 1, noc, "interior face out"))
      return UFX_ERR_MESH;
    m.fp1[i] = p1; m.fp2[i] = p2; m.fin[i] = ci - 1; m.fout[i] = co - 1;
    m.fbc[i] = in.fi_bc ? in.fi_bc[i] : 0;
  }
  for (int k = 0; k < nbf; ++k) {
    int i = nfi + k;
    int p1 = in.fb_p1[k] - 1, p2 = in.fb_p2[k] - 1, ci = in.fb_in[k];
    if (ci <= 0) { err = "BC faces in/out==0"; return UFX_ERR_MESH; }
    if (!chk(p1, 0, nop - 1, "boundary face p1") || !chk(p2, 0, nop - 1, "boundary face p2") ||
        !chk(ci, 1, noc, "boundary face in"))
      return UFX_ERR_MESH;
    m.fp1[i] = p1; m.fp2[i] = p2; m.fin[i] = ci - 1;
    m.fout[i] = noc + k;  // ghost cell, CCFVM/read_input.F90:183-186
    m.fnx[i] = in.fb_nx[k];
    m.fny[i] = in.fb_ny[k];
    m.fbc[i] = in.fb_bc[k];
  }
  return UFX_OK;
}

// Ghost centres (read_grid's mirror) and cell_face_norm, per face.
void cc_geom_faces(Mesh& m) {
  const int noc = m.noc, nfi = m.nfi, nbf = m.nbf, nof = m.nof();
  for (int k = 0; k < nbf; ++k) {
    // ghost centre = mirror of the inner centre about the face line
    // (pt2linemirror, CCFVM/types/tools.F90:30-46), file endpoint order
    const int i = nfi + k, p1 = m.fp1[i], p2 = m.fp2[i], ci = m.fin[i];
    double sx = m.px[p1], sy = m.py[p1], ex = m.px[p2], ey = m.py[p2];
    double px = m.ccx[ci], py = m.ccy[ci];
    double dx = ex - sx, dy = ey - sy;
    double x = sx - px, y = sy - py;
    double r = 1.0 / (dx * dx + dy * dy);
    m.ccx[noc + k] = px + 2.0 * (x - x * dx * dx * r - y * dx * dy * r);
    m.ccy[noc + k] = py + 2.0 * (y - y * dy * dy * r - x * dx * dy * r);
  }

  // ---- cell_face_norm (CCFVM/geometric.F90:62-198)
  for (int i = 0; i < nfi; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i];
    int out = m.fout[i];
    double xc = m.ccx[out], yc = m.ccy[out];
    double dx = m.px[p2] - m.px[p1], dy = m.py[p2] - m.py[p1];
    m.fnx[i] = dy;
    m.fny[i] = -dx;
    double check = dotprod(m, p1, p2, xc, yc);
    if (check < 0.0) {
      // the reference's repair swaps both the sides and the endpoints
      int t = m.fin[i]; m.fin[i] = m.fout[i]; m.fout[i] = t;
      m.fp1[i] = p2; m.fp2[i] = p1;
      dx = m.px[p1] - m.px[p2]; dy = m.py[p1] - m.py[p2];
      m.fnx[i] = dy;
      m.fny[i] = -dx;
    }
    m.fcx[i] = 0.5 * (m.px[p1] + m.px[p2]);
    m.fcy[i] = 0.5 * (m.py[p1] + m.py[p2]);
  }
  for (int i = nfi; i < nof; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i];
    int cin = m.fin[i];
    if (m.fbc[i] == 2001 || m.fbc[i] == 1001) {
      double check = dotprod(m, p1, p2, m.ccx[cin], m.ccy[cin]);
      if (check > 0.0) {
        m.fp1[i] = p2; m.fp2[i] = p1;
        double dx = m.px[p1] - m.px[p2], dy = m.py[p1] - m.py[p2];
        m.fnx[i] = dy;
        m.fny[i] = -dx;
      }
    }
    m.fcx[i] = 0.5 * (m.px[p1] + m.px[p2]);
    m.fcy[i] = 0.5 * (m.py[p1] + m.py[p2]);
  }
  for (int i = 0; i < nof; ++i) {
    double dx = m.fnx[i], dy = m.fny[i];
    m.farea[i] = std::sqrt(dx * dx + dy * dy);
    m.fnx[i] = m.fnx[i] / m.farea[i];
    m.fny[i] = m.fny[i] / m.farea[i];
  }
}

// Connectivity and the breadth-first numbering; the vertex -> cell lists go
// to m.v2c_ptr / m.v2c (the pseudo-Laplacian weights follow their order).
int cc_geom_connect(Mesh& m, std::string& err) {
  const int nop = m.nop, noc = m.noc, nfi = m.nfi, nof = m.nof();

  // ---- cell_2_cell (CCFVM/geometric.F90:200-245), interior faces only
  std::vector<std::vector<int>> c2c(noc);
  for (int i = 0; i < nfi; ++i) {
    int a = m.fin[i], b = m.fout[i];
    push_unique(c2c[a], b);
    push_unique(c2c[b], a);
  }

  // ---- renumber: breadth-first from cell 1 (CCFVM/geometric.F90:247-359)
  std::vector<int> oldnum(noc, -1), newnum(noc, -1);
  int c = 0;
  oldnum[0] = 0;
  newnum[0] = 0;
  for (int i = 0; i < noc; ++i) {
    int it = oldnum[i];
    if (it < 0) { err = "renumber: mesh is not connected"; return UFX_ERR_MESH; }
    for (int t1 : c2c[it]) {
      if (t1 >= 0 && t1 < noc && newnum[t1] < 0) {
        ++c;
        oldnum[c] = t1;
        newnum[t1] = c;
      }
    }
  }
  if (c + 1 != noc) { err = "renumber: count does not match no. of cells"; return UFX_ERR_MESH; }
  for (int i = 0; i < nof; ++i) {
    if (m.fin[i] < noc) m.fin[i] = newnum[m.fin[i]];
    if (m.fout[i] < noc) m.fout[i] = newnum[m.fout[i]];
  }
  {
    std::vector<double> cx(m.ccx), cy(m.ccy), cv(m.ccv);
    m.nc2c.resize(noc);
    for (int i = 0; i < noc; ++i) {
      int it = oldnum[i];
      m.ccx[i] = cx[it];
      m.ccy[i] = cy[it];
      m.ccv[i] = cv[it];
      m.nc2c[i] = (int)c2c[it].size();
    }
  }
  m.oldnum = oldnum;

  // ---- cell_2_face (CCFVM/geometric.F90:425-457)
  std::vector<std::vector<int>> c2f(noc);
  for (int i = 0; i < nof; ++i) {
    c2f[m.fin[i]].push_back(i);
    if (m.fout[i] < noc) c2f[m.fout[i]].push_back(i);
  }

  // ---- vertex_2_cell (CCFVM/geometric.F90:459-531), ghosts included
  std::vector<std::vector<int>> v2c(nop);
  for (int i = 0; i < nof; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i];
    push_unique(v2c[p1], m.fin[i]);
    push_unique(v2c[p2], m.fin[i]);
    push_unique(v2c[p1], m.fout[i]);
    push_unique(v2c[p2], m.fout[i]);
  }

  // ---- cell_2_vertex (CCFVM/geometric.F90:533-616): its walk round each
  // cell reorders the cell's face list, which fixes the LU-SGS sweep order;
  // the walk, closed with its first vertex (:611-613), is kept for output.
  m.c2v_ptr.assign(noc + 1, 0);
  m.c2v.clear();
  for (int i = 0; i < noc; ++i) {
    std::vector<int>& f = c2f[i];
    int nf = (int)f.size();
    if (nf < 2) { err = "cell with fewer than two faces"; return UFX_ERR_MESH; }
    double xc = m.ccx[i], yc = m.ccy[i];
    std::vector<int> c2v;
    int p1 = m.fp1[f[0]], p2 = m.fp2[f[0]];
    if (dotprod(m, p1, p2, xc, yc) < 0.0) { c2v.push_back(p1); c2v.push_back(p2); }
    else { c2v.push_back(p2); c2v.push_back(p1); }
    for (int k = 1; k < nf - 1; ++k) {
      int last = c2v.back();
      for (int j = k; j < nf; ++j) {
        int m1 = m.fp1[f[j]], m2 = m.fp2[f[j]];
        if (m1 == last || m2 == last) {
          if (dotprod(m, m1, m2, xc, yc) < 0.0) c2v.push_back(m2);
          else c2v.push_back(m1);
          int t = f[j]; f[j] = f[k]; f[k] = t;
        }
      }
    }
    c2v.push_back(c2v.front());
    m.c2v.insert(m.c2v.end(), c2v.begin(), c2v.end());
    m.c2v_ptr[i + 1] = (int)m.c2v.size();
  }
  m.c2f_ptr.assign(noc + 1, 0);
  for (int i = 0; i < noc; ++i) m.c2f_ptr[i + 1] = m.c2f_ptr[i] + (int)c2f[i].size();
  m.c2f.resize(m.c2f_ptr[noc]);
  for (int i = 0; i < noc; ++i)
    std::memcpy(&m.c2f[m.c2f_ptr[i]], c2f[i].data(), c2f[i].size() * sizeof(int));
  m.v2c_ptr.assign(nop + 1, 0);
  for (int i = 0; i < nop; ++i) m.v2c_ptr[i + 1] = m.v2c_ptr[i] + (int)v2c[i].size();
  m.v2c.resize(m.v2c_ptr[nop]);
  for (int i = 0; i < nop; ++i)
    if (!v2c[i].empty())
      std::memcpy(&m.v2c[m.v2c_ptr[i]], v2c[i].data(), v2c[i].size() * sizeof(int));
  return UFX_OK;
}

// Face co-volumes, least-squares factors and pseudo-Laplacian weights.
int cc_geom_cells(Mesh& m, std::string& err) {
  const int nop = m.nop, noc = m.noc, nfi = m.nfi, nof = m.nof(), ncell = m.ncell();

  // ---- face_co_volume (CCFVM/geometric.F90:619-674)
  m.ccov.assign(ncell, 0.0);
  for (int i = 0; i < nof; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i], ci = m.fin[i], co = m.fout[i];
    double x[3], y[3];
    x[0] = m.px[p1]; y[0] = m.py[p1];
    x[1] = m.px[p2]; y[1] = m.py[p2];
    x[2] = m.ccx[ci]; y[2] = m.ccy[ci];
    double a = triarea(x, y);
    if (a <= 0.0) { err = "geom,triarea,  -ve area"; return UFX_ERR_MESH; }
    m.fcov[i] = a;
    x[0] = m.px[p1]; y[0] = m.py[p1];
    x[1] = m.ccx[co]; y[1] = m.ccy[co];
    x[2] = m.px[p2]; y[2] = m.py[p2];
    a = triarea(x, y);
    if (a <= 0.0) { err = "geom,triarea,  -ve area"; return UFX_ERR_MESH; }
    m.fcov[i] = m.fcov[i] + a;
    m.ccov[ci] = m.ccov[ci] + m.fcov[i];
    m.ccov[co] = m.ccov[co] + m.fcov[i];
  }

  // ---- LSQR_Coeff (CCFVM/geometric.F90:898-960), unit weights
  m.r11.assign(noc, 0.0); m.r12.assign(noc, 0.0); m.r22.assign(noc, 0.0);
  for (int i = 0; i < nfi; ++i) {
    int a = m.fin[i], b = m.fout[i];
    double xi = m.ccx[a], yi = m.ccy[a], xo = m.ccx[b], yo = m.ccy[b];
    double dx = xo - xi, dy = yo - yi;
    m.r11[a] += dx * dx; m.r12[a] += dx * dy; m.r22[a] += dy * dy;
    dx = xi - xo; dy = yi - yo;
    m.r11[b] += dx * dx; m.r12[b] += dx * dy; m.r22[b] += dy * dy;
  }
  for (int i = nfi; i < nof; ++i) {
    int a = m.fin[i], b = m.fout[i];
    double dx = m.ccx[b] - m.ccx[a], dy = m.ccy[b] - m.ccy[a];
    m.r11[a] += dx * dx; m.r12[a] += dx * dy; m.r22[a] += dy * dy;
  }
  for (int i = 0; i < noc; ++i) {
    double r11 = std::sqrt(m.r11[i]);
    m.r11[i] = r11;
    double r12 = m.r12[i] / r11;
    m.r12[i] = r12;
    m.r22[i] = std::sqrt(m.r22[i] - r12 * r12);
  }

  // ---- pseudo_Laplacian (CCFVM/geometric.F90:962-1013)
  m.wt.resize(m.v2c_ptr[nop]);
  for (int i = 0; i < nop; ++i) {
    double ixx = 0.0, iyy = 0.0, ixy = 0.0, rx = 0.0, ry = 0.0;
    for (int j = m.v2c_ptr[i]; j < m.v2c_ptr[i + 1]; ++j) {
      const int cc = m.v2c[j];
      double dx = m.ccx[cc] - m.px[i], dy = m.ccy[cc] - m.py[i];
      rx = rx + dx; ry = ry + dy;
      ixx = ixx + dx * dx;
      iyy = iyy + dy * dy;
      ixy = ixy + dx * dy;
    }
    double det = ixx * iyy - ixy * ixy;
    if (det < 0.0) { err = "Determinant -ve @ pseudo_Laplacian"; return UFX_ERR_MESH; }
    double gx = (ixy * ry - iyy * rx) / det;
    double gy = (ixy * rx - ixx * ry) / det;
    for (int j = m.v2c_ptr[i]; j < m.v2c_ptr[i + 1]; ++j) {
      const int cc = m.v2c[j];
      double dx = m.ccx[cc] - m.px[i], dy = m.ccy[cc] - m.py[i];
      m.wt[j] = 1.0 + gx * dx + gy * dy;
    }
  }
  return UFX_OK;
}

int cc_geometric(const RawInput& in, Mesh& m, std::string& err) {
  if (int rc = cc_geom_read(in, m, err)) return rc;
  cc_geom_faces(m);
  if (int rc = cc_geom_connect(m, err)) return rc;
  return cc_geom_cells(m, err);
}

}  // namespace ufx
