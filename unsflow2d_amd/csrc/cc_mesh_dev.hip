// The cell-centred preprocessing (CCFVM/geometric.F90) with its per-face,
// per-cell and per-vertex passes on the GPU: ghost centres and
// cell_face_norm (one thread per face), face co-volumes (per face), their
// per-cell sums and the least-squares factors (one thread per cell, over
// the cell's faces in ascending face order, the order of the reference's
// face loops), and the pseudo-Laplacian weights (one thread per vertex).
// The connectivity and the breadth-first numbering, sequential in the
// reference, stay on the host (cc_geom_connect).  Compiled without
// floating-point contraction, with the reference's operation order and
// IEEE division and square root, so every array equals the host form's
// (and the reference's) bit for bit.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ufx_internal.h"

namespace ufx {

namespace {

constexpr int GB = 256;

struct GeomDev {
  int nop, noc, nfi, nbf;
  const double *px, *py;
  double *ccx, *ccy;
  int *fp1, *fp2, *fin, *fout;
  const int* fbc;
  double *fnx, *fny, *farea, *fcx, *fcy, *fcov;
  double *ccov, *r11, *r12, *r22;
  const int *cf_ptr, *cf;      // faces of each cell (ghosts included), ascending
  const int *v2c_ptr, *v2c;
  double* wt;
  int* err;                    // [0] triarea <= 0, [1] determinant < 0
};

// CCFVM/solution.F90:363-384 (dotprod), as the host form.
__device__ double d_dotprod(const GeomDev& g, int n1, int n2, double xc, double yc) {
  double x1 = g.px[n1], y1 = g.py[n1];
  double x2 = g.px[n2], y2 = g.py[n2];
  double dx = x2 - x1, dy = y2 - y1;
  double ds = sqrt(dx * dx + dy * dy);
  double nx = dy / ds, ny = -dx / ds;
  dx = xc - x1;
  dy = yc - y1;
  ds = sqrt(dx * dx + dy * dy);
  double rx = dx / ds, ry = dy / ds;
  return rx * nx + ry * ny;
}

// CCFVM/geometric.F90:676-693 (triarea).
__device__ double d_triarea(const double* x, const double* y) {
  double area = 0.0;
  for (int i = 0; i < 3; ++i) {
    int ip1 = (i == 2) ? 0 : i + 1;
    double dx = 0.5 * (x[ip1] + x[i]);
    double dy = y[ip1] - y[i];
    area = area + dx * dy;
  }
  return area;
}

// Ghost centre of a boundary face (pt2linemirror, CCFVM/types/tools.F90:
// 30-46, file endpoint order) and cell_face_norm (CCFVM/geometric.F90:
// 62-198): orientation repair, normal, centre, length, unit normal.
__global__ __launch_bounds__(GB) void k_geom_faces(GeomDev g) {
  const int i = blockIdx.x * GB + threadIdx.x;
  const int nof = g.nfi + g.nbf;
  if (i >= nof) return;
  int p1 = g.fp1[i], p2 = g.fp2[i];
  if (i < g.nfi) {
    const int out = g.fout[i];
    double dx = g.px[p2] - g.px[p1], dy = g.py[p2] - g.py[p1];
    double nx = dy, ny = -dx;
    if (d_dotprod(g, p1, p2, g.ccx[out], g.ccy[out]) < 0.0) {
      const int t = g.fin[i];
      g.fin[i] = g.fout[i];
      g.fout[i] = t;
      g.fp1[i] = p2;
      g.fp2[i] = p1;
      dx = g.px[p1] - g.px[p2];
      dy = g.py[p1] - g.py[p2];
      nx = dy;
      ny = -dx;
    }
    g.fnx[i] = nx;
    g.fny[i] = ny;
  } else {
    const int k = i - g.nfi, ci = g.fin[i];
    {
      double sx = g.px[p1], sy = g.py[p1], ex = g.px[p2], ey = g.py[p2];
      double px = g.ccx[ci], py = g.ccy[ci];
      double dx = ex - sx, dy = ey - sy;
      double x = sx - px, y = sy - py;
      double r = 1.0 / (dx * dx + dy * dy);
      g.ccx[g.noc + k] = px + 2.0 * (x - x * dx * dx * r - y * dx * dy * r);
      g.ccy[g.noc + k] = py + 2.0 * (y - y * dy * dy * r - x * dx * dy * r);
    }
    if ((g.fbc[i] == 2001 || g.fbc[i] == 1001) &&
        d_dotprod(g, p1, p2, g.ccx[ci], g.ccy[ci]) > 0.0) {
      g.fp1[i] = p2;
      g.fp2[i] = p1;
      double dx = g.px[p1] - g.px[p2], dy = g.py[p1] - g.py[p2];
      g.fnx[i] = dy;
      g.fny[i] = -dx;
    }
  }
  p1 = g.fp1[i];
  p2 = g.fp2[i];
  g.fcx[i] = 0.5 * (g.px[p1] + g.px[p2]);
  g.fcy[i] = 0.5 * (g.py[p1] + g.py[p2]);
  double dx = g.fnx[i], dy = g.fny[i];
  const double a = sqrt(dx * dx + dy * dy);
  g.farea[i] = a;
  g.fnx[i] = g.fnx[i] / a;
  g.fny[i] = g.fny[i] / a;
}

// face_co_volume (CCFVM/geometric.F90:619-674), the per-face part: the
// diamond's two triangles.
__global__ __launch_bounds__(GB) void k_geom_fcov(GeomDev g) {
  const int i = blockIdx.x * GB + threadIdx.x;
  if (i >= g.nfi + g.nbf) return;
  const int p1 = g.fp1[i], p2 = g.fp2[i], ci = g.fin[i], co = g.fout[i];
  double x[3], y[3];
  x[0] = g.px[p1]; y[0] = g.py[p1];
  x[1] = g.px[p2]; y[1] = g.py[p2];
  x[2] = g.ccx[ci]; y[2] = g.ccy[ci];
  double a = d_triarea(x, y);
  const bool bad1 = a <= 0.0;
  double cov = a;
  x[0] = g.px[p1]; y[0] = g.py[p1];
  x[1] = g.ccx[co]; y[1] = g.ccy[co];
  x[2] = g.px[p2]; y[2] = g.py[p2];
  a = d_triarea(x, y);
  if (bad1 || a <= 0.0) atomicExch(&g.err[0], 1);
  g.fcov[i] = cov + a;
}

// Per cell, over its faces in ascending order: the co-volume sum (all
// cells) and the least-squares factors LSQR_Coeff (CCFVM/geometric.F90:
// 898-960, interior cells; interior faces precede boundary faces in the
// face numbering, so this is the reference's loop order too).
__global__ __launch_bounds__(GB) void k_geom_cells(GeomDev g) {
  const int c = blockIdx.x * GB + threadIdx.x;
  if (c >= g.noc + g.nbf) return;
  double cov = 0.0, r11 = 0.0, r12 = 0.0, r22 = 0.0;
  for (int j = g.cf_ptr[c]; j < g.cf_ptr[c + 1]; ++j) {
    const int i = g.cf[j];
    cov = cov + g.fcov[i];
    if (c >= g.noc) continue;
    const int a = g.fin[i], b = g.fout[i];
    double dx, dy;
    if (i < g.nfi && c == b) {
      dx = g.ccx[a] - g.ccx[b];
      dy = g.ccy[a] - g.ccy[b];
    } else {
      dx = g.ccx[b] - g.ccx[a];
      dy = g.ccy[b] - g.ccy[a];
    }
    r11 += dx * dx;
    r12 += dx * dy;
    r22 += dy * dy;
  }
  g.ccov[c] = cov;
  if (c >= g.noc) return;
  r11 = sqrt(r11);
  r12 = r12 / r11;
  g.r11[c] = r11;
  g.r12[c] = r12;
  g.r22[c] = sqrt(r22 - r12 * r12);
}

// pseudo_Laplacian (CCFVM/geometric.F90:962-1013), one vertex per thread.
__global__ __launch_bounds__(GB) void k_geom_pl(GeomDev g) {
  const int i = blockIdx.x * GB + threadIdx.x;
  if (i >= g.nop) return;
  double ixx = 0.0, iyy = 0.0, ixy = 0.0, rx = 0.0, ry = 0.0;
  for (int j = g.v2c_ptr[i]; j < g.v2c_ptr[i + 1]; ++j) {
    const int cc = g.v2c[j];
    double dx = g.ccx[cc] - g.px[i], dy = g.ccy[cc] - g.py[i];
    rx = rx + dx;
    ry = ry + dy;
    ixx = ixx + dx * dx;
    iyy = iyy + dy * dy;
    ixy = ixy + dx * dy;
  }
  double det = ixx * iyy - ixy * ixy;
  if (det < 0.0) {
    atomicExch(&g.err[1], 1);
    return;
  }
  double gx = (ixy * ry - iyy * rx) / det;
  double gy = (ixy * rx - ixx * ry) / det;
  for (int j = g.v2c_ptr[i]; j < g.v2c_ptr[i + 1]; ++j) {
    const int cc = g.v2c[j];
    double dx = g.ccx[cc] - g.px[i], dy = g.ccy[cc] - g.py[i];
    g.wt[j] = 1.0 + gx * dx + gy * dy;
  }
}

// Device buffers of one preprocessing run, freed together.
struct Buffers {
  std::vector<void*> ptrs;
  ~Buffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back((void*)*p);
    return e;
  }
};

template <typename T>
hipError_t up(T* d, const std::vector<T>& h) {
  return h.empty() ? hipSuccess
                   : hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}
template <typename T>
hipError_t down(std::vector<T>& h, const T* d) {
  return h.empty() ? hipSuccess
                   : hipMemcpy(h.data(), d, h.size() * sizeof(T), hipMemcpyDeviceToHost);
}

int grid(int n) { return n > 0 ? (n + GB - 1) / GB : 1; }

}  // namespace

int cc_geometric_device(const RawInput& in, Mesh& m, std::string& err) {
  if (int rc = cc_geom_read(in, m, err)) return rc;
  const int nop = m.nop, noc = m.noc, nfi = m.nfi, nbf = m.nbf, nof = m.nof(),
            ncell = m.ncell();
  Buffers b;
  GeomDev g{};
  g.nop = nop; g.noc = noc; g.nfi = nfi; g.nbf = nbf;
  double *px, *py;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t r) { if (e == hipSuccess) e = r; return e == hipSuccess; };
  ok(b.alloc(&px, nop)); ok(b.alloc(&py, nop));
  ok(b.alloc(&g.ccx, ncell)); ok(b.alloc(&g.ccy, ncell));
  ok(b.alloc(&g.fp1, nof)); ok(b.alloc(&g.fp2, nof)); ok(b.alloc(&g.fin, nof));
  ok(b.alloc(&g.fout, nof));
  int* fbc;
  ok(b.alloc(&fbc, nof));
  ok(b.alloc(&g.fnx, nof)); ok(b.alloc(&g.fny, nof)); ok(b.alloc(&g.farea, nof));
  ok(b.alloc(&g.fcx, nof)); ok(b.alloc(&g.fcy, nof)); ok(b.alloc(&g.fcov, nof));
  ok(b.alloc(&g.ccov, ncell)); ok(b.alloc(&g.r11, noc)); ok(b.alloc(&g.r12, noc));
  ok(b.alloc(&g.r22, noc));
  ok(b.alloc(&g.err, 2));
  g.px = px; g.py = py; g.fbc = fbc;
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  ok(up(px, m.px)); ok(up(py, m.py)); ok(up(g.ccx, m.ccx)); ok(up(g.ccy, m.ccy));
  ok(up(g.fp1, m.fp1)); ok(up(g.fp2, m.fp2)); ok(up(g.fin, m.fin)); ok(up(g.fout, m.fout));
  ok(up(fbc, m.fbc)); ok(up(g.fnx, m.fnx)); ok(up(g.fny, m.fny));
  ok(hipMemset(g.err, 0, 2 * sizeof(int)));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  // per face: ghost centres, orientation, normals, centres, lengths
  k_geom_faces<<<grid(nof), GB>>>(g);
  ok(hipGetLastError()); ok(hipDeviceSynchronize());
  ok(down(m.fp1, g.fp1)); ok(down(m.fp2, g.fp2)); ok(down(m.fin, g.fin));
  ok(down(m.fout, g.fout)); ok(down(m.fnx, g.fnx)); ok(down(m.fny, g.fny));
  ok(down(m.farea, g.farea)); ok(down(m.fcx, g.fcx)); ok(down(m.fcy, g.fcy));
  ok(down(m.ccx, g.ccx)); ok(down(m.ccy, g.ccy));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  // connectivity and numbering on the host
  if (int rc = cc_geom_connect(m, err)) return rc;
  // the faces of every cell in ascending order (ghost: its boundary face)
  std::vector<int> cf_ptr(ncell + 1, 0), cf;
  for (int i = 0; i < nof; ++i) {
    ++cf_ptr[m.fin[i] + 1];
    ++cf_ptr[m.fout[i] + 1];
  }
  for (int c = 0; c < ncell; ++c) cf_ptr[c + 1] += cf_ptr[c];
  cf.resize(cf_ptr[ncell]);
  {
    std::vector<int> at(cf_ptr.begin(), cf_ptr.end() - 1);
    for (int i = 0; i < nof; ++i) {
      cf[at[m.fin[i]]++] = i;
      cf[at[m.fout[i]]++] = i;
    }
  }
  int *cfp, *cfl, *v2cp, *v2c;
  ok(b.alloc(&cfp, ncell + 1)); ok(b.alloc(&cfl, cf.size()));
  ok(b.alloc(&v2cp, nop + 1)); ok(b.alloc(&v2c, m.v2c.size()));
  ok(b.alloc(&g.wt, m.v2c.size()));
  g.cf_ptr = cfp; g.cf = cfl; g.v2c_ptr = v2cp; g.v2c = v2c;
  // renumbered cells and faces
  ok(up(g.ccx, m.ccx)); ok(up(g.ccy, m.ccy)); ok(up(g.fin, m.fin)); ok(up(g.fout, m.fout));
  ok(up(cfp, cf_ptr)); ok(up(cfl, cf)); ok(up(v2cp, m.v2c_ptr)); ok(up(v2c, m.v2c));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  k_geom_fcov<<<grid(nof), GB>>>(g);
  k_geom_cells<<<grid(ncell), GB>>>(g);
  k_geom_pl<<<grid(nop), GB>>>(g);
  ok(hipGetLastError()); ok(hipDeviceSynchronize());
  int ferr[2] = {0, 0};
  ok(hipMemcpy(ferr, g.err, sizeof(ferr), hipMemcpyDeviceToHost));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  // the reference stops at the first of these, in its routine order
  if (ferr[0]) { err = "geom,triarea,  -ve area"; return UFX_ERR_MESH; }
  if (ferr[1]) { err = "Determinant -ve @ pseudo_Laplacian"; return UFX_ERR_MESH; }
  m.fcov.assign(nof, 0.0);
  m.ccov.assign(ncell, 0.0);
  m.r11.assign(noc, 0.0); m.r12.assign(noc, 0.0); m.r22.assign(noc, 0.0);
  m.wt.assign(m.v2c.size(), 0.0);
  ok(down(m.fcov, g.fcov)); ok(down(m.ccov, g.ccov)); ok(down(m.r11, g.r11));
  ok(down(m.r12, g.r12)); ok(down(m.r22, g.r22)); ok(down(m.wt, g.wt));
  if (e != hipSuccess) { err = std::string("device preprocessing: ") + hipGetErrorString(e); return UFX_ERR_HIP; }
  return UFX_OK;
}

}  // namespace ufx
