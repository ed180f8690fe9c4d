// The vertex-centred preprocessing (NCFVM/geometric.f90) with its per-edge
// and per-vertex passes on the GPU: dual-face normals and lengths
// (cell_face_norm_c2f_dist, :198-255) and the co-volume check (:668-738)
// one thread per edge; median-dual volumes (:258-313) and least-squares
// weights (LSQR_Coeff, :767-804) one thread per vertex, over its edges or
// neighbours in the reference's loop order.  The neighbour lists and the
// breadth-first numbering stay on the host (nc_geom_connect).  Compiled
// without floating-point contraction, with the reference's operation order
// and IEEE division and square root, so every array equals the host form's
// bit for bit.
#include <hip/hip_runtime.h>

#include <climits>
#include <string>
#include <vector>

#include "ufx_internal.h"

namespace ufx {

namespace {

constexpr int GB = 256;

struct NCGeomDev {
  int nop, nof;
  const double *px, *py, *ccx, *ccy;
  const int *fp1, *fp2, *fin, *fout;
  const int *v2f_ptr, *v2f, *v2v_ptr, *v2v;
  double *fnx, *fny, *farea, *fcx, *fcy, *pcv, *wx, *wy;
  int* err;  // lowest edge with a non-positive co-volume
};

// Dual-face normal of each edge: the half normals centroid(in)->midpoint
// and midpoint->centroid(out), each turned along p1->p2, summed; its length
// and the unit normal.
__global__ __launch_bounds__(GB) void k_ncg_edges(NCGeomDev g) {
  const int i = blockIdx.x * GB + threadIdx.x;
  if (i >= g.nof) return;
  const int p1 = g.fp1[i], p2 = g.fp2[i];
  double xm = 0.5 * (g.px[p1] + g.px[p2]);
  double ym = 0.5 * (g.py[p1] + g.py[p2]);
  double dx = (g.px[p2] - g.px[p1]);
  double dy = (g.py[p2] - g.py[p1]);
  g.fcx[i] = xm;
  g.fcy[i] = ym;
  double nx = 0.0, ny = 0.0;
  const int ci = g.fin[i], co = g.fout[i];
  if (ci >= 0) {
    double xc = g.ccx[ci], yc = g.ccy[ci];
    double sx = yc - ym;
    double sy = -(xc - xm);
    double vprod = sx * dx + sy * dy;
    if (vprod < 0.0) { sx = -sx; sy = -sy; }
    nx = sx;
    ny = sy;
  }
  if (co >= 0) {
    double xc = g.ccx[co], yc = g.ccy[co];
    double sx = (ym - yc);
    double sy = -(xm - xc);
    double vprod = sx * dx + sy * dy;
    if (vprod < 0.0) { sx = -sx; sy = -sy; }
    nx = nx + sx;
    ny = ny + sy;
  }
  double ds = sqrt(nx * nx + ny * ny);
  g.fnx[i] = nx / ds;
  g.fny[i] = ny / ds;
  g.farea[i] = ds;
}

// face_co_volume's sign check: the shoelace of the edge's diamond (or
// half-diamond at the boundary).
__global__ __launch_bounds__(GB) void k_ncg_cov(NCGeomDev g) {
  const int i = blockIdx.x * GB + threadIdx.x;
  if (i >= g.nof) return;
  const int p1 = g.fp1[i], p2 = g.fp2[i], ci = g.fin[i], co = g.fout[i];
  double cov = 0.0;
  auto piece = [&](double ax, double ay, double bx, double by) {
    double dx = 0.5 * (bx + ax), dy = by - ay;
    cov = cov + dx * dy;
  };
  if (ci >= 0 && co >= 0) {
    piece(g.px[p1], g.py[p1], g.ccx[co], g.ccy[co]);
    piece(g.ccx[co], g.ccy[co], g.px[p2], g.py[p2]);
    piece(g.px[p2], g.py[p2], g.ccx[ci], g.ccy[ci]);
    piece(g.ccx[ci], g.ccy[ci], g.px[p1], g.py[p1]);
  }
  if (ci < 0 && co >= 0) {
    piece(g.px[p1], g.py[p1], g.ccx[co], g.ccy[co]);
    piece(g.ccx[co], g.ccy[co], g.px[p2], g.py[p2]);
    piece(g.px[p2], g.py[p2], g.px[p1], g.py[p1]);
  }
  if (co < 0 && ci >= 0) {
    piece(g.px[p1], g.py[p1], g.px[p2], g.py[p2]);
    piece(g.px[p2], g.py[p2], g.ccx[ci], g.ccy[ci]);
    piece(g.ccx[ci], g.ccy[ci], g.px[p1], g.py[p1]);
  }
  if (cov <= 0.0) atomicMin(g.err, i);
}

// Per vertex: the median-dual volume, over its edges in ascending order
// with the reference's per-edge terms (+ at p1, - at p2; a missing cell
// closes along the half edge), and the least-squares weights of its
// neighbours.
__global__ __launch_bounds__(GB) void k_ncg_vertices(NCGeomDev g) {
  const int p = blockIdx.x * GB + threadIdx.x;
  if (p >= g.nop) return;
  double cv = 0.0;
  for (int j = g.v2f_ptr[p]; j < g.v2f_ptr[p + 1]; ++j) {
    const int i = g.v2f[j];
    const int p1 = g.fp1[i], p2 = g.fp2[i], ci = g.fin[i], co = g.fout[i];
    const bool first = p == p1;
    const double xm = g.fcx[i], ym = g.fcy[i];
    double dx, dy, xc, yc;
    if (ci >= 0) {
      xc = g.ccx[ci]; yc = g.ccy[ci];
    } else {
      xc = first ? g.px[p1] : g.px[p2];
      yc = first ? g.py[p1] : g.py[p2];
    }
    dy = yc - ym; dx = 0.5 * (xc + xm);
    cv = first ? cv + dx * dy : cv - dx * dy;
    if (co >= 0) {
      xc = g.ccx[co]; yc = g.ccy[co];
    } else {
      xc = first ? g.px[p1] : g.px[p2];
      yc = first ? g.py[p1] : g.py[p2];
    }
    dy = ym - yc; dx = 0.5 * (xc + xm);
    cv = first ? cv + dx * dy : cv - dx * dy;
  }
  g.pcv[p] = cv;
  const double xc = g.px[p], yc = g.py[p];
  double r11 = 0.0, r12 = 0.0, r22 = 0.0;
  for (int j = g.v2v_ptr[p]; j < g.v2v_ptr[p + 1]; ++j) {
    const int n = g.v2v[j];
    double dx = g.px[n] - xc, dy = g.py[n] - yc;
    r11 = r11 + dx * dx;
    r12 = r12 + dx * dy;
    r22 = r22 + dy * dy;
  }
  r11 = sqrt(r11);
  r12 = r12 / r11;
  r22 = sqrt(r22 - r12 * r12);
  for (int j = g.v2v_ptr[p]; j < g.v2v_ptr[p + 1]; ++j) {
    const int n = g.v2v[j];
    double dx = g.px[n] - xc, dy = g.py[n] - yc;
    double alfa1 = dx / r11 / r11;
    double alfa2 = (dy - dx * r12 / r11) / r22 / r22;
    g.wx[j] = alfa1 - alfa2 * r12 / r11;
    g.wy[j] = alfa2;
  }
}

struct Buffers {
  std::vector<void*> ptrs;
  ~Buffers() {
    for (void* q : ptrs) (void)hipFree(q);
  }
  template <typename T>
  hipError_t put(T** d, const std::vector<T>& h) {
    hipError_t e = hipMalloc((void**)d, (h.empty() ? 1 : h.size()) * sizeof(T));
    if (e != hipSuccess) return e;
    ptrs.push_back((void*)*d);
    return h.empty() ? hipSuccess
                     : hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  template <typename T>
  hipError_t alloc(T** d, size_t n) {
    hipError_t e = hipMalloc((void**)d, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back((void*)*d);
    return e;
  }
};

template <typename T>
hipError_t down(std::vector<T>& h, const T* d) {
  return h.empty() ? hipSuccess
                   : hipMemcpy(h.data(), d, h.size() * sizeof(T), hipMemcpyDeviceToHost);
}

int grid(int n) { return n > 0 ? (n + GB - 1) / GB : 1; }

}  // namespace

int nc_geometric_device(const NCRawInput& in, Mesh& m, std::string& err) {
  if (int rc = nc_geom_connect(in, m, err)) return rc;
  const int nop = m.nop, nof = m.nfi;
  Buffers b;
  NCGeomDev g{};
  g.nop = nop;
  g.nof = nof;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t r) { if (e == hipSuccess) e = r; return e == hipSuccess; };
  double *px, *py, *ccx, *ccy;
  int *fp1, *fp2, *fin, *fout, *v2fp, *v2f, *v2vp, *v2v;
  ok(b.put(&px, m.px)); ok(b.put(&py, m.py)); ok(b.put(&ccx, m.ccx)); ok(b.put(&ccy, m.ccy));
  ok(b.put(&fp1, m.fp1)); ok(b.put(&fp2, m.fp2)); ok(b.put(&fin, m.fin)); ok(b.put(&fout, m.fout));
  ok(b.put(&v2fp, m.v2f_ptr)); ok(b.put(&v2f, m.v2f)); ok(b.put(&v2vp, m.v2v_ptr));
  ok(b.put(&v2v, m.v2v));
  g.px = px; g.py = py; g.ccx = ccx; g.ccy = ccy; g.fp1 = fp1; g.fp2 = fp2; g.fin = fin;
  g.fout = fout; g.v2f_ptr = v2fp; g.v2f = v2f; g.v2v_ptr = v2vp; g.v2v = v2v;
  ok(b.alloc(&g.fnx, nof)); ok(b.alloc(&g.fny, nof)); ok(b.alloc(&g.farea, nof));
  ok(b.alloc(&g.fcx, nof)); ok(b.alloc(&g.fcy, nof)); ok(b.alloc(&g.pcv, nop));
  ok(b.alloc(&g.wx, m.v2v.size())); ok(b.alloc(&g.wy, m.v2v.size()));
  const std::vector<int> big(1, INT_MAX);
  ok(b.put(&g.err, big));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  k_ncg_edges<<<grid(nof), GB>>>(g);
  k_ncg_cov<<<grid(nof), GB>>>(g);
  k_ncg_vertices<<<grid(nop), GB>>>(g);
  ok(hipGetLastError()); ok(hipDeviceSynchronize());
  std::vector<int> bad(1, INT_MAX);
  ok(down(bad, g.err));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  if (bad[0] != INT_MAX) {
    err = std::to_string(bad[0] + 1) + " -ve face co-volume";
    return UFX_ERR_MESH;
  }
  m.fnx.assign(nof, 0.0); m.fny.assign(nof, 0.0); m.farea.assign(nof, 0.0);
  m.fcx.assign(nof, 0.0); m.fcy.assign(nof, 0.0); m.pcv.assign(nop, 0.0);
  m.wx.assign(m.v2v.size(), 0.0); m.wy.assign(m.v2v.size(), 0.0);
  ok(down(m.fnx, g.fnx)); ok(down(m.fny, g.fny)); ok(down(m.farea, g.farea));
  ok(down(m.fcx, g.fcx)); ok(down(m.fcy, g.fcy)); ok(down(m.pcv, g.pcv));
  ok(down(m.wx, g.wx)); ok(down(m.wy, g.wy));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  return UFX_OK;
}

}  // namespace ufx
