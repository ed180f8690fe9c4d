// Host-side preprocessing of a vertex-centred (median-dual) mesh: the
// restatement of the reference's geometric (NCFVM/geometric.f90:1-46) for
// everything the hot path consumes.
//
//   vertex_2_vertex  :52-81     neighbour lists in edge order
//   renumber         :85-180    breadth-first vertex numbering from vertex 1
//                               (the reference leaves newnum/oldnum
//                               unallocated-uninitialised; zero is meant)
//   cell_face_norm   :184-345   edge midpoints, dual-face normals and lengths,
//                               median-dual volumes
//   vertex_2_face    :347-379   edges of each vertex in edge order
//   face_co_volume   :657-758   only its validity check (STOP on a
//                               non-positive co-volume)
//   LSQR_Coeff       :760-806   least-squares weights per neighbour
// The connectivity of cells (cell_2_face, vertex_2_cell, cell_2_vertex) and
// the averaged face lengths feed no routine of the iteration and are not
// built.  Floating-point expressions keep the reference's operation order
// (compiled without FMA contraction) so the arrays agree bit for bit.
#include "ufx_internal.h"

#include <cmath>
#include <string>

namespace ufx {

// The checked input, the neighbour lists, the breadth-first numbering and
// the edges of each vertex (connectivity).
int nc_geom_connect(const NCRawInput& in, Mesh& m, std::string& err) {
  const int nop = in.nop, noc = in.noc, nof = in.nof;
  if (nop <= 0 || noc <= 0 || nof <= 0) {
    err = "nc_geometric: empty mesh";
    return UFX_ERR_MESH;
  }
  for (int i = 0; i < nof; ++i) {
    if (in.e_p1[i] < 1 || in.e_p1[i] > nop || in.e_p2[i] < 1 || in.e_p2[i] > nop ||
        in.e_in[i] < 0 || in.e_in[i] > noc || in.e_out[i] < 0 || in.e_out[i] > noc ||
        (in.e_in[i] == 0 && in.e_out[i] == 0)) {
      err = "edge " + std::to_string(i + 1) + " has an invalid vertex or cell index";
      return UFX_ERR_MESH;
    }
  }
  m = Mesh();
  m.kind = UFX_KIND_NC;
  m.nop = nop;
  m.noc = noc;
  m.nfi = nof;
  m.nbf = 0;
  m.ccx.assign(in.ccx, in.ccx + noc);
  m.ccy.assign(in.ccy, in.ccy + noc);
  m.ccv.assign(in.cv, in.cv + noc);
  m.fin.resize(nof);
  m.fout.resize(nof);
  m.fbc.resize(nof);
  for (int i = 0; i < nof; ++i) {
    m.fin[i] = in.e_in[i] - 1;
    m.fout[i] = in.e_out[i] - 1;
    m.fbc[i] = in.e_bc[i];
  }

  // vertex_2_vertex, :61-76 (input numbering, 0-based)
  std::vector<std::vector<int>> nbr(nop);
  for (int i = 0; i < nof; ++i) {
    int p1 = in.e_p1[i] - 1, p2 = in.e_p2[i] - 1;
    nbr[p1].push_back(p2);
    nbr[p2].push_back(p1);
  }

  // renumber, :100-131: breadth-first from vertex 1 over the neighbour lists
  std::vector<int> oldnum(nop, -1), newnum(nop, -1);
  int c = 0;
  oldnum[0] = 0;
  newnum[0] = 0;
  for (int i = 0; i < nop; ++i) {
    int it = oldnum[i];
    if (it < 0) {
      err = "renumber: Fatal error. it is zero for i=" + std::to_string(i + 1);
      return UFX_ERR_MESH;
    }
    for (int t1 : nbr[it])
      if (newnum[t1] < 0) {
        ++c;
        oldnum[c] = t1;
        newnum[t1] = c;
      }
  }
  if (c + 1 != nop) {
    err = "renumber: count does not match nt.";
    return UFX_ERR_MESH;
  }
  m.pt_oldnum = oldnum;
  m.px.resize(nop);
  m.py.resize(nop);
  m.pbc.resize(nop);
  m.v2v_ptr.assign(nop + 1, 0);
  for (int i = 0; i < nop; ++i) {
    int it = oldnum[i];
    m.px[i] = in.px[it];
    m.py[i] = in.py[it];
    m.pbc[i] = in.pbc ? in.pbc[it] : 0;
    for (int t1 : nbr[it]) m.v2v.push_back(newnum[t1]);
    m.v2v_ptr[i + 1] = (int)m.v2v.size();
  }
  m.fp1.resize(nof);
  m.fp2.resize(nof);
  for (int i = 0; i < nof; ++i) {
    m.fp1[i] = newnum[in.e_p1[i] - 1];
    m.fp2[i] = newnum[in.e_p2[i] - 1];
  }

  // vertex_2_face, :355-374
  {
    std::vector<std::vector<int>> vf(nop);
    for (int i = 0; i < nof; ++i) {
      vf[m.fp1[i]].push_back(i);
      vf[m.fp2[i]].push_back(i);
    }
    m.v2f_ptr.assign(nop + 1, 0);
    for (int i = 0; i < nop; ++i) {
      for (int f : vf[i]) m.v2f.push_back(f);
      m.v2f_ptr[i + 1] = (int)m.v2f.size();
    }
  }

  return UFX_OK;
}

// Dual-face normals, median-dual volumes, the co-volume check and the
// least-squares weights (per edge and per vertex).
int nc_geom_numeric(Mesh& m, std::string& err) {
  const int nop = m.nop, nof = m.nfi;

  // cell_face_norm_c2f_dist, :198-255: dual-face normal of each edge, the
  // sum of the half normals centroid(in)->midpoint and midpoint->centroid(out),
  // each turned to point along p1->p2
  m.fnx.assign(nof, 0.0);
  m.fny.assign(nof, 0.0);
  m.farea.assign(nof, 0.0);
  m.fcx.assign(nof, 0.0);
  m.fcy.assign(nof, 0.0);
  for (int i = 0; i < nof; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i];
    double xm = 0.5 * (m.px[p1] + m.px[p2]);
    double ym = 0.5 * (m.py[p1] + m.py[p2]);
    double dx = (m.px[p2] - m.px[p1]);
    double dy = (m.py[p2] - m.py[p1]);
    m.fcx[i] = xm;
    m.fcy[i] = ym;
    int ci = m.fin[i], co = m.fout[i];
    if (ci >= 0) {
      double xc = m.ccx[ci], yc = m.ccy[ci];
      double sx = yc - ym;
      double sy = -(xc - xm);
      double vprod = sx * dx + sy * dy;
      if (vprod < 0.0) { sx = -sx; sy = -sy; }
      m.fnx[i] = sx;
      m.fny[i] = sy;
    }
    if (co >= 0) {
      double xc = m.ccx[co], yc = m.ccy[co];
      double sx = (ym - yc);
      double sy = -(xm - xc);
      double vprod = sx * dx + sy * dy;
      if (vprod < 0.0) { sx = -sx; sy = -sy; }
      m.fnx[i] = m.fnx[i] + sx;
      m.fny[i] = m.fny[i] + sy;
    }
  }
  for (int i = 0; i < nof; ++i) {
    double sx = m.fnx[i], sy = m.fny[i];
    double ds = std::sqrt(sx * sx + sy * sy);
    m.fnx[i] = m.fnx[i] / ds;
    m.fny[i] = m.fny[i] / ds;
    m.farea[i] = ds;
  }
  // median-dual volumes, :258-313 (shoelace pieces of each edge's dual face;
  // a missing side closes along the half edges)
  m.pcv.assign(nop, 0.0);
  for (int i = 0; i < nof; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i], ci = m.fin[i], co = m.fout[i];
    double xm = m.fcx[i], ym = m.fcy[i];
    double dx, dy, xc, yc;
    if (ci >= 0) {
      xc = m.ccx[ci]; yc = m.ccy[ci];
      dy = yc - ym; dx = 0.5 * (xc + xm);
      m.pcv[p1] = m.pcv[p1] + dx * dy;
      m.pcv[p2] = m.pcv[p2] - dx * dy;
    } else {
      xc = m.px[p1]; yc = m.py[p1];
      dy = yc - ym; dx = 0.5 * (xc + xm);
      m.pcv[p1] = m.pcv[p1] + dx * dy;
      xc = m.px[p2]; yc = m.py[p2];
      dy = yc - ym; dx = 0.5 * (xc + xm);
      m.pcv[p2] = m.pcv[p2] - dx * dy;
    }
    if (co >= 0) {
      xc = m.ccx[co]; yc = m.ccy[co];
      dy = ym - yc; dx = 0.5 * (xc + xm);
      m.pcv[p1] = m.pcv[p1] + dx * dy;
      m.pcv[p2] = m.pcv[p2] - dx * dy;
    } else {
      xc = m.px[p1]; yc = m.py[p1];
      dy = ym - yc; dx = 0.5 * (xc + xm);
      m.pcv[p1] = m.pcv[p1] + dx * dy;
      xc = m.px[p2]; yc = m.py[p2];
      dy = ym - yc; dx = 0.5 * (xc + xm);
      m.pcv[p2] = m.pcv[p2] - dx * dy;
    }
  }

  // face_co_volume, :668-738: the co-volume itself feeds nothing; its sign
  // check stops the reference
  for (int i = 0; i < nof; ++i) {
    int p1 = m.fp1[i], p2 = m.fp2[i], ci = m.fin[i], co = m.fout[i];
    double cov = 0.0, x1, y1, x2, y2, dx, dy;
    auto piece = [&](double ax, double ay, double bx, double by) {
      x1 = ax; y1 = ay; x2 = bx; y2 = by;
      dx = 0.5 * (x2 + x1); dy = y2 - y1;
      cov = cov + dx * dy;
    };
    if (ci >= 0 && co >= 0) {
      piece(m.px[p1], m.py[p1], m.ccx[co], m.ccy[co]);
      piece(m.ccx[co], m.ccy[co], m.px[p2], m.py[p2]);
      piece(m.px[p2], m.py[p2], m.ccx[ci], m.ccy[ci]);
      piece(m.ccx[ci], m.ccy[ci], m.px[p1], m.py[p1]);
    }
    if (ci < 0 && co >= 0) {
      piece(m.px[p1], m.py[p1], m.ccx[co], m.ccy[co]);
      piece(m.ccx[co], m.ccy[co], m.px[p2], m.py[p2]);
      piece(m.px[p2], m.py[p2], m.px[p1], m.py[p1]);
    }
    if (co < 0 && ci >= 0) {
      piece(m.px[p1], m.py[p1], m.px[p2], m.py[p2]);
      piece(m.px[p2], m.py[p2], m.ccx[ci], m.ccy[ci]);
      piece(m.ccx[ci], m.ccy[ci], m.px[p1], m.py[p1]);
    }
    if (cov <= 0.0) {
      err = std::to_string(i + 1) + " -ve face co-volume";
      return UFX_ERR_MESH;
    }
  }

  // LSQR_Coeff, :767-804
  m.wx.resize(m.v2v.size());
  m.wy.resize(m.v2v.size());
  for (int i = 0; i < nop; ++i) {
    double xc = m.px[i], yc = m.py[i];
    double r11 = 0.0, r12 = 0.0, r22 = 0.0;
    for (int j = m.v2v_ptr[i]; j < m.v2v_ptr[i + 1]; ++j) {
      int n = m.v2v[j];
      double dx = m.px[n] - xc, dy = m.py[n] - yc;
      r11 = r11 + dx * dx;
      r12 = r12 + dx * dy;
      r22 = r22 + dy * dy;
    }
    r11 = std::sqrt(r11);
    r12 = r12 / r11;
    r22 = std::sqrt(r22 - r12 * r12);
    for (int j = m.v2v_ptr[i]; j < m.v2v_ptr[i + 1]; ++j) {
      int n = m.v2v[j];
      double dx = m.px[n] - xc, dy = m.py[n] - yc;
      double alfa1 = dx / r11 / r11;
      double alfa2 = (dy - dx * r12 / r11) / r22 / r22;
      m.wx[j] = alfa1 - alfa2 * r12 / r11;
      m.wy[j] = alfa2;
    }
  }
  return UFX_OK;
}

int nc_geometric(const NCRawInput& in, Mesh& m, std::string& err) {
  if (int rc = nc_geom_connect(in, m, err)) return rc;
  return nc_geom_numeric(m, err);
}

}  // namespace ufx
