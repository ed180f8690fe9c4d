// C-ABI of the host-side mesh handling (include/ufx.h): preprocessing of a
// geometry.inp mesh (the reference's geometric), import of a mesh the
// reference has already preprocessed, and read-back of the mesh arrays.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <queue>
#include <string>
#include <vector>

#include "ufx_internal.h"

namespace ufx {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace ufx

using namespace ufx;

namespace ufx {

std::vector<int> greedy_colour(int n, const std::vector<int>& ptr, const std::vector<int>& adj,
                               int* ncolours) {
  // smallest-last order: repeatedly take out a vertex of least remaining
  // degree (ties: the lowest number); colour in the reverse of that order
  std::vector<int> deg(n, 0), order;
  std::vector<char> out(n, 0);
  order.reserve(n);
  using Item = std::pair<int, int>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  for (int v = 0; v < n; ++v) {
    for (int j = ptr[v]; j < ptr[v + 1]; ++j)
      if (adj[j] >= 0 && adj[j] < n) ++deg[v];
    pq.push({deg[v], v});
  }
  while (!pq.empty()) {
    const Item it = pq.top();
    pq.pop();
    const int v = it.second;
    if (out[v] || it.first != deg[v]) continue;
    out[v] = 1;
    order.push_back(v);
    for (int j = ptr[v]; j < ptr[v + 1]; ++j) {
      const int u = adj[j];
      if (u >= 0 && u < n && !out[u]) pq.push({--deg[u], u});
    }
  }
  std::reverse(order.begin(), order.end());
  std::vector<int> col(n, -1);
  std::vector<int> seen;  // seen[c] == v: colour c is taken by a neighbour of v
  int nc = 0;
  for (int v : order) {
    for (int j = ptr[v]; j < ptr[v + 1]; ++j) {
      int o = adj[j];
      if (o >= 0 && o < n && col[o] >= 0) seen[col[o]] = v;
    }
    int c = 0;
    while (c < nc && seen[c] == v) ++c;
    if (c == nc) {
      ++nc;
      seen.push_back(-1);
    }
    col[v] = c;
  }
  if (ncolours) *ncolours = nc;
  return col;
}

// Position along a Hilbert curve of order 16 of the point (x, y) of the
// 65536 x 65536 grid.
static uint32_t hilbert_d(uint32_t x, uint32_t y) {
  uint32_t d = 0;
  for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
    uint32_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = 65535u - x;
        y = 65535u - y;
      }
      uint32_t t = x; x = y; y = t;
    }
  }
  return d;
}

// Points ordered along the Hilbert curve through them (ties keep their
// given order).
std::vector<int> hilbert_order(const double* cx, const double* cy, int n) {
  if (n <= 0) return {};
  double x0 = cx[0], x1 = cx[0], y0 = cy[0], y1 = cy[0];
  for (int i = 0; i < n; ++i) {
    x0 = std::min(x0, cx[i]); x1 = std::max(x1, cx[i]);
    y0 = std::min(y0, cy[i]); y1 = std::max(y1, cy[i]);
  }
  double sc = 65535.0 / std::max(std::max(x1 - x0, y1 - y0), 1e-300);
  std::vector<std::pair<uint32_t, int>> key(n);
  for (int i = 0; i < n; ++i) {
    uint32_t ix = (uint32_t)std::min(65535.0, std::max(0.0, (cx[i] - x0) * sc));
    uint32_t iy = (uint32_t)std::min(65535.0, std::max(0.0, (cy[i] - y0) * sc));
    key[i] = {hilbert_d(ix, iy), i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = key[i].second;
  return order;
}

std::vector<int> lusgs_colours(const Mesh& m, int* ncolours) {
  if (m.kind == UFX_KIND_NC) return greedy_colour(m.nop, m.v2v_ptr, m.v2v, ncolours);
  std::vector<int> ptr(m.noc + 1, 0), adj;
  for (int f = 0; f < m.nfi; ++f) {
    ++ptr[m.fin[f] + 1];
    ++ptr[m.fout[f] + 1];
  }
  for (int c = 0; c < m.noc; ++c) ptr[c + 1] += ptr[c];
  adj.resize(ptr[m.noc]);
  std::vector<int> at(ptr.begin(), ptr.end() - 1);
  for (int f = 0; f < m.nfi; ++f) {
    adj[at[m.fin[f]]++] = m.fout[f];
    adj[at[m.fout[f]]++] = m.fin[f];
  }
  return greedy_colour(m.noc, ptr, adj, ncolours);
}

}  // namespace ufx

namespace ufx {
namespace {

int cc_geometric_entry(const char* who, bool device, int nop, int noc, int nfi, int nbf,
                       const double* px, const double* py, const double* ccx, const double* ccy,
                       const double* cv, const int* fi_p1, const int* fi_p2, const int* fi_in,
                       const int* fi_out, const int* fi_bc, const int* fb_p1, const int* fb_p2,
                       const int* fb_in, const double* fb_nx, const double* fb_ny,
                       const int* fb_bc, ufx_mesh** mesh) {
  if (!mesh || !px || !py || !ccx || !ccy || !cv || (nfi > 0 && (!fi_p1 || !fi_p2 || !fi_in || !fi_out)) ||
      (nbf > 0 && (!fb_p1 || !fb_p2 || !fb_in || !fb_nx || !fb_ny || !fb_bc))) {
    set_error(std::string(who) + ": null argument");
    return UFX_ERR_ARG;
  }
  RawInput in;
  in.nop = nop; in.noc = noc; in.nfi = nfi; in.nbf = nbf;
  in.px = px; in.py = py; in.ccx = ccx; in.ccy = ccy; in.cv = cv;
  in.fi_p1 = fi_p1; in.fi_p2 = fi_p2; in.fi_in = fi_in; in.fi_out = fi_out; in.fi_bc = fi_bc;
  in.fb_p1 = fb_p1; in.fb_p2 = fb_p2; in.fb_in = fb_in; in.fb_nx = fb_nx; in.fb_ny = fb_ny;
  in.fb_bc = fb_bc;
  ufx_mesh* m = new ufx_mesh();
  std::string err;
  int rc = device ? cc_geometric_device(in, m->m, err) : cc_geometric(in, m->m, err);
  if (rc) {
    delete m;
    set_error(err);
    return rc;
  }
  *mesh = m;
  return UFX_OK;
}


int nc_geometric_entry(const char* who, bool device, int nop, int noc, int nof, const double* px,
                       const double* py, const int* pbc, const double* ccx, const double* ccy,
                       const double* cv, const int* e_p1, const int* e_p2, const int* e_in,
                       const int* e_out, const int* e_bc, ufx_mesh** mesh) {
  if (!mesh || !px || !py || !ccx || !ccy || !cv || !e_p1 || !e_p2 || !e_in || !e_out || !e_bc) {
    set_error(std::string(who) + ": null argument");
    return UFX_ERR_ARG;
  }
  NCRawInput in;
  in.nop = nop; in.noc = noc; in.nof = nof;
  in.px = px; in.py = py; in.pbc = pbc; in.ccx = ccx; in.ccy = ccy; in.cv = cv;
  in.e_p1 = e_p1; in.e_p2 = e_p2; in.e_in = e_in; in.e_out = e_out; in.e_bc = e_bc;
  ufx_mesh* m = new ufx_mesh();
  std::string err;
  int rc = device ? nc_geometric_device(in, m->m, err) : nc_geometric(in, m->m, err);
  if (rc) {
    delete m;
    set_error(err);
    return rc;
  }
  *mesh = m;
  return UFX_OK;
}

}  // namespace
}  // namespace ufx

extern "C" {

const char* ufx_last_error(void) { return g_err.c_str(); }

int ufx_cc_geometric(int nop, int noc, int nfi, int nbf, const double* px, const double* py,
                     const double* ccx, const double* ccy, const double* cv, const int* fi_p1,
                     const int* fi_p2, const int* fi_in, const int* fi_out, const int* fi_bc,
                     const int* fb_p1, const int* fb_p2, const int* fb_in, const double* fb_nx,
                     const double* fb_ny, const int* fb_bc, ufx_mesh** mesh) {
  return ufx::cc_geometric_entry("ufx_cc_geometric", false, nop, noc, nfi, nbf, px, py, ccx, ccy, cv,
                            fi_p1, fi_p2, fi_in, fi_out, fi_bc, fb_p1, fb_p2, fb_in, fb_nx, fb_ny,
                            fb_bc, mesh);
}

int ufx_cc_geometric_device(int nop, int noc, int nfi, int nbf, const double* px,
                            const double* py, const double* ccx, const double* ccy,
                            const double* cv, const int* fi_p1, const int* fi_p2,
                            const int* fi_in, const int* fi_out, const int* fi_bc,
                            const int* fb_p1, const int* fb_p2, const int* fb_in,
                            const double* fb_nx, const double* fb_ny, const int* fb_bc,
                            ufx_mesh** mesh) {
  return ufx::cc_geometric_entry("ufx_cc_geometric_device", true, nop, noc, nfi, nbf, px, py, ccx,
                            ccy, cv, fi_p1, fi_p2, fi_in, fi_out, fi_bc, fb_p1, fb_p2, fb_in,
                            fb_nx, fb_ny, fb_bc, mesh);
}

int ufx_nc_geometric(int nop, int noc, int nof, const double* px, const double* py,
                     const int* pbc, const double* ccx, const double* ccy, const double* cv,
                     const int* e_p1, const int* e_p2, const int* e_in, const int* e_out,
                     const int* e_bc, ufx_mesh** mesh) {
  return ufx::nc_geometric_entry("ufx_nc_geometric", false, nop, noc, nof, px, py, pbc, ccx, ccy,
                                 cv, e_p1, e_p2, e_in, e_out, e_bc, mesh);
}

int ufx_nc_geometric_device(int nop, int noc, int nof, const double* px, const double* py,
                            const int* pbc, const double* ccx, const double* ccy,
                            const double* cv, const int* e_p1, const int* e_p2,
                            const int* e_in, const int* e_out, const int* e_bc,
                            ufx_mesh** mesh) {
  return ufx::nc_geometric_entry("ufx_nc_geometric_device", true, nop, noc, nof, px, py, pbc,
                                 ccx, ccy, cv, e_p1, e_p2, e_in, e_out, e_bc, mesh);
}

int ufx_cc_mesh_import(int nop, int noc, int nfi, int nbf, const double* px, const double* py,
                       const int* fc_p1, const int* fc_p2, const int* fc_in, const int* fc_out,
                       const int* fc_bc, const double* fc_nx, const double* fc_ny,
                       const double* fc_area, const double* fc_cenx, const double* fc_ceny,
                       const double* fc_cov, const double* cell_cenx, const double* cell_ceny,
                       const double* cell_cv, const double* cell_cov, const double* cell_r11,
                       const double* cell_r12, const double* cell_r22, const int* cell_nc2c,
                       const int* c2f_ptr, const int* c2f, const int* v2c_ptr, const int* v2c,
                       const double* v2c_wt, ufx_mesh** mesh) {
  if (!mesh || nop <= 0 || noc <= 0 || nfi < 0 || nbf < 0) {
    set_error("ufx_cc_mesh_import: invalid argument");
    return UFX_ERR_ARG;
  }
  const int nof = nfi + nbf, ncell = noc + nbf;
  ufx_mesh* h = new ufx_mesh();
  Mesh& m = h->m;
  m.kind = UFX_KIND_CC;
  m.nop = nop; m.noc = noc; m.nfi = nfi; m.nbf = nbf;
  m.px.assign(px, px + nop);
  m.py.assign(py, py + nop);
  auto idx = [](const int* a, int n, int base) {
    std::vector<int> v(n);
    for (int i = 0; i < n; ++i) v[i] = a[i] - base;
    return v;
  };
  m.fp1 = idx(fc_p1, nof, 1); m.fp2 = idx(fc_p2, nof, 1);
  m.fin = idx(fc_in, nof, 1); m.fout = idx(fc_out, nof, 1);
  m.fbc.assign(fc_bc, fc_bc + nof);
  m.fnx.assign(fc_nx, fc_nx + nof); m.fny.assign(fc_ny, fc_ny + nof);
  m.farea.assign(fc_area, fc_area + nof);
  m.fcx.assign(fc_cenx, fc_cenx + nof); m.fcy.assign(fc_ceny, fc_ceny + nof);
  m.fcov.assign(fc_cov, fc_cov + nof);
  m.ccx.assign(cell_cenx, cell_cenx + ncell); m.ccy.assign(cell_ceny, cell_ceny + ncell);
  m.ccov.assign(cell_cov, cell_cov + ncell);
  m.ccv.assign(cell_cv, cell_cv + noc);
  m.r11.assign(cell_r11, cell_r11 + noc); m.r12.assign(cell_r12, cell_r12 + noc);
  m.r22.assign(cell_r22, cell_r22 + noc);
  m.nc2c.assign(cell_nc2c, cell_nc2c + noc);
  m.c2f_ptr.assign(c2f_ptr, c2f_ptr + noc + 1);
  m.c2f = idx(c2f, c2f_ptr[noc], 1);
  m.v2c_ptr.assign(v2c_ptr, v2c_ptr + nop + 1);
  m.v2c = idx(v2c, v2c_ptr[nop], 1);
  m.wt.assign(v2c_wt, v2c_wt + v2c_ptr[nop]);
  m.oldnum.resize(noc);
  for (int i = 0; i < noc; ++i) m.oldnum[i] = i;
  for (int i = 0; i < nof; ++i) {
    if (m.fin[i] < 0 || m.fin[i] >= noc || m.fout[i] < 0 || m.fout[i] >= ncell ||
        m.fp1[i] < 0 || m.fp1[i] >= nop || m.fp2[i] < 0 || m.fp2[i] >= nop ||
        (i < nfi && m.fout[i] >= noc) || (i >= nfi && m.fout[i] != noc + (i - nfi))) {
      delete h;
      set_error("ufx_cc_mesh_import: face " + std::to_string(i + 1) + " has an invalid index");
      return UFX_ERR_MESH;
    }
  }
  for (int c : m.c2f)
    if (c < 0 || c >= nof) { delete h; set_error("ufx_cc_mesh_import: bad c2f"); return UFX_ERR_MESH; }
  for (int c : m.v2c)
    if (c < 0 || c >= ncell) { delete h; set_error("ufx_cc_mesh_import: bad v2c"); return UFX_ERR_MESH; }
  *mesh = h;
  return UFX_OK;
}

int ufx_mesh_size(const ufx_mesh* mesh, const char* name, int64_t* count) {
  if (!mesh || !name || !count) { set_error("ufx_mesh_size: null argument"); return UFX_ERR_ARG; }
  const Mesh& m = mesh->m;
  std::string n(name);
  int64_t v = -1;
  if (n.rfind("fc_", 0) == 0) v = m.nof();
  if (n == "cell_cenx" || n == "cell_ceny" || n == "cell_cov") v = m.ncell();
  if (n == "cell_cv" || n == "cell_r11" || n == "cell_r12" || n == "cell_r22" ||
      n == "cell_nc2c" || n == "cell_nc2f" || n == "oldnum") v = m.noc;
  if (n == "cell_c2f") v = (int64_t)m.c2f.size();
  if (n == "cell_nc2v") v = m.c2v_ptr.empty() ? 0 : m.noc;
  if (n == "cell_c2v") v = (int64_t)m.c2v.size();
  if (n == "pt_x" || n == "pt_y" || n == "pt_nv2c") v = m.nop;
  if (n == "pt_v2c" || n == "pt_wt") v = (int64_t)m.v2c.size();
  if (n == "counts") v = 4;
  if (n == "cell_colour") v = m.noc;
  if (m.kind == UFX_KIND_NC) {
    v = -1;
    if (n.rfind("fc_", 0) == 0) v = m.nof();
    if (n == "pt_x" || n == "pt_y" || n == "pt_cv" || n == "pt_bc" || n == "pt_nv2v" ||
        n == "pt_nv2f" || n == "oldnum")
      v = m.nop;
    if (n == "pt_v2v" || n == "pt_wx" || n == "pt_wy") v = (int64_t)m.v2v.size();
    if (n == "pt_v2f") v = (int64_t)m.v2f.size();
    if (n == "cell_cenx" || n == "cell_ceny" || n == "cell_cv") v = m.noc;
    if (n == "pt_colour") v = m.nop;
    if (n == "counts") v = 3;
  }
  if (v < 0) { set_error("ufx_mesh_size: unknown array " + n); return UFX_ERR_ARG; }
  *count = v;
  return UFX_OK;
}

int ufx_mesh_get(const ufx_mesh* mesh, const char* name, void* dst) {
  if (!mesh || !name || !dst) { set_error("ufx_mesh_get: null argument"); return UFX_ERR_ARG; }
  const Mesh& m = mesh->m;
  std::string n(name);
  int* di = (int*)dst;
  double* dd = (double*)dst;
  auto put1 = [&](const std::vector<int>& v) { for (size_t i = 0; i < v.size(); ++i) di[i] = v[i] + 1; };
  auto put0 = [&](const std::vector<int>& v) { for (size_t i = 0; i < v.size(); ++i) di[i] = v[i]; };
  auto putd = [&](const std::vector<double>& v) { if (!v.empty()) std::memcpy(dd, v.data(), v.size() * 8); };
  if (m.kind == UFX_KIND_NC) {
    if (n == "fc_p1") put1(m.fp1);
    else if (n == "fc_p2") put1(m.fp2);
    else if (n == "fc_in") put1(m.fin);
    else if (n == "fc_out") put1(m.fout);
    else if (n == "fc_bc") put0(m.fbc);
    else if (n == "fc_nx") putd(m.fnx);
    else if (n == "fc_ny") putd(m.fny);
    else if (n == "fc_area") putd(m.farea);
    else if (n == "fc_cenx") putd(m.fcx);
    else if (n == "fc_ceny") putd(m.fcy);
    else if (n == "pt_x") putd(m.px);
    else if (n == "pt_y") putd(m.py);
    else if (n == "pt_cv") putd(m.pcv);
    else if (n == "pt_bc") put0(m.pbc);
    else if (n == "pt_nv2v") { for (int i = 0; i < m.nop; ++i) di[i] = m.v2v_ptr[i + 1] - m.v2v_ptr[i]; }
    else if (n == "pt_v2v") put1(m.v2v);
    else if (n == "pt_wx") putd(m.wx);
    else if (n == "pt_wy") putd(m.wy);
    else if (n == "pt_nv2f") { for (int i = 0; i < m.nop; ++i) di[i] = m.v2f_ptr[i + 1] - m.v2f_ptr[i]; }
    else if (n == "pt_v2f") put1(m.v2f);
    else if (n == "oldnum") put1(m.pt_oldnum);
    else if (n == "cell_cenx") putd(m.ccx);
    else if (n == "cell_ceny") putd(m.ccy);
    else if (n == "cell_cv") putd(m.ccv);
    else if (n == "counts") { di[0] = m.nop; di[1] = m.noc; di[2] = m.nof(); }
    else if (n == "pt_colour") put0(lusgs_colours(m, nullptr));
    else { set_error("ufx_mesh_get: unknown array " + n); return UFX_ERR_ARG; }
    return UFX_OK;
  }
  if (n == "fc_p1") put1(m.fp1);
  else if (n == "fc_p2") put1(m.fp2);
  else if (n == "fc_in") put1(m.fin);
  else if (n == "fc_out") put1(m.fout);
  else if (n == "fc_bc") put0(m.fbc);
  else if (n == "fc_nx") putd(m.fnx);
  else if (n == "fc_ny") putd(m.fny);
  else if (n == "fc_area") putd(m.farea);
  else if (n == "fc_cenx") putd(m.fcx);
  else if (n == "fc_ceny") putd(m.fcy);
  else if (n == "fc_cov") putd(m.fcov);
  else if (n == "cell_cenx") putd(m.ccx);
  else if (n == "cell_ceny") putd(m.ccy);
  else if (n == "cell_cov") putd(m.ccov);
  else if (n == "cell_cv") putd(m.ccv);
  else if (n == "cell_r11") putd(m.r11);
  else if (n == "cell_r12") putd(m.r12);
  else if (n == "cell_r22") putd(m.r22);
  else if (n == "cell_nc2c") put0(m.nc2c);
  else if (n == "cell_nc2f") { for (int i = 0; i < m.noc; ++i) di[i] = m.c2f_ptr[i + 1] - m.c2f_ptr[i]; }
  else if (n == "cell_c2f") put1(m.c2f);
  else if (n == "cell_nc2v") { for (int i = 0; i < m.noc; ++i) di[i] = m.c2v_ptr[i + 1] - m.c2v_ptr[i]; }
  else if (n == "cell_c2v") put1(m.c2v);
  else if (n == "pt_x") putd(m.px);
  else if (n == "pt_y") putd(m.py);
  else if (n == "pt_nv2c") { for (int i = 0; i < m.nop; ++i) di[i] = m.v2c_ptr[i + 1] - m.v2c_ptr[i]; }
  else if (n == "pt_v2c") put1(m.v2c);
  else if (n == "pt_wt") putd(m.wt);
  else if (n == "oldnum") put1(m.oldnum);
  else if (n == "counts") { di[0] = m.nop; di[1] = m.noc; di[2] = m.nof(); di[3] = m.nbf; }
  else if (n == "cell_colour") put0(lusgs_colours(m, nullptr));
  else { set_error("ufx_mesh_get: unknown array " + n); return UFX_ERR_ARG; }
  return UFX_OK;
}

void ufx_mesh_destroy(ufx_mesh* mesh) { delete mesh; }

}  // extern "C"
