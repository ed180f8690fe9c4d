// Device residual hot path of the cell-centred solver for gfx950, and its
// C-ABI (include/ufx.h).
//
// Layout in HBM (structure of arrays, one array per variable):
//   cells: qc[4][ncell], qp[5][ncell] (interior then ghost cells),
//          qold/res/phi[4][noc], grad[10][noc] (d/dx, d/dy per variable),
//          dt, la, dtA (= dt / area) [noc], centroids [2][ncell];
//   faces: p1, p2, in, out, bc, nx, ny, area, centre [nof];
//   per-cell face slots: sl_face[4][noc] (faces of the cell in ascending face
//          order) and sl_nbr[4][noc] (the cell across the face, bit-inverted
//          when the cell is the face's `out` side), plus cnf[noc] = number of
//          faces | number of boundary faces << 8.
// Every per-cell sum is a gather over the cell's slots in the order the
// reference accumulates it (ascending face index for gradients and spectral
// radii; boundary faces first, then interior faces, for the residual), so
// the sums are deterministic and reproduce the reference's rounding order.
// Interior face fluxes are evaluated by both adjacent cells with identical
// inputs, which trades a second flux evaluation for the atomics and scatter
// of a face loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cc_physics.h"
#include "ufx_internal.h"

namespace ufx {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

constexpr int MAXF = 4;       // faces per cell (triangles and quadrilaterals)
constexpr int BLK = 256;      // threads per block of the cell/point kernels
constexpr int WALL = 1001, FARF = 2001;

// error word codes (device)
constexpr int DERR_ROE = 1, DERR_POS = 2, DERR_AREA = 3;

#define HIPCHK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " +    \
                __FILE__ + ":" + std::to_string(__LINE__));                      \
      return UFX_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

// Device view of a context (passed by value to kernels).
struct Dev {
  int nop, noc, nfi, nbf, nof, ncell;
  // mesh
  const double *px, *py;
  const int *fp1, *fp2, *fin, *fout, *fbc;
  const double *fnx, *fny, *farea, *fcx, *fcy;
  const double *ccx, *ccy, *ccv, *ccov, *r11, *r12, *r22;
  const int *nc2c;
  const int *sl_face, *sl_nbr, *cnf;
  const int *lu_face, *lu_nf;
  const int *v2c_ptr, *v2c;
  const double *wt;
  // state
  double *qc, *qp, *qold, *res, *grad, *phi, *dt, *la, *dta;
  double *fla, *fmu, *cmul, *pprim, *pmu, *fgrad;
  double *bflux;     // [2][nbf] x/y momentum flux of each boundary face
  double *part;      // reduction partials
  double *scal;      // small scalars (see SC_*)
  int *err;          // [0] code, [1] cell, [2] first non-positive cell (INT_MAX if none)
};

// scal[] slots
constexpr int SC_QMAX = 0, SC_QMIN = 4, SC_NORM = 8, SC_FRES1 = 9, SC_DTMIN = 10, SC_NSCAL = 16;

struct Params {
  double gamma, gamma1, znd, r_inf, u_inf, v_inf, p_inf, mref, t_wall, m_inf, rey, vll;
  double aoa, xref, yref;
  int flow, iwall, flux, grad, ilimit, irs, timemode, precision;
  double cfl_max;
};

// ------------------------------------------------------------- reductions
// Deterministic block reduction: fixed shared-memory tree.
template <int N, typename Op>
__device__ __forceinline__ void block_reduce(double (*sh)[BLK], double* v, Op op) {
  int t = threadIdx.x;
  for (int k = 0; k < N; ++k) sh[k][t] = v[k];
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < N; ++k) sh[k][t] = op(sh[k][t], sh[k][t + s]);
    __syncthreads();
  }
  for (int k = 0; k < N; ++k) v[k] = sh[k][0];
}

struct OpMax { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
struct OpMin { __device__ double operator()(double a, double b) const { return a < b ? a : b; } };
struct OpAdd { __device__ double operator()(double a, double b) const { return a + b; } };

// ---------------------------------------------------------------- kernels

// Primitive refresh of interior cells (CCFVM/main.F90:62-67), optionally
// qold = qc (CCFVM/solution.F90:118-132), and per-block max/min of the four
// primitives over all cells including ghosts (CCFVM/limiter.F90:146-149).
__global__ void k_prim(Dev d, Gas<double> g, int do_prim, int do_old, int do_range) {
  __shared__ double sh[4][BLK];
  int c = blockIdx.x * BLK + threadIdx.x;
  double mx[4], mn[4];
  for (int k = 0; k < 4; ++k) { mx[k] = -INFINITY; mn[k] = INFINITY; }
  if (c < d.ncell && !d.err[0]) {
    if (c < d.noc && do_prim) {
      double q[4], p[5];
      for (int k = 0; k < 4; ++k) q[k] = d.qc[k * d.ncell + c];
      con2prim(g, q, p);
      for (int k = 0; k < 5; ++k) d.qp[k * d.ncell + c] = p[k];
      // positivity of the refreshed primitives, CCFVM/main.F90:384-411
      // (the reference checks them after the update; they are unchanged)
      if (do_old && (p[0] <= 0.0 || p[3] <= 0.0)) {
        atomicMin(&d.err[2], c + 1);
        atomicCAS(d.err, 0, DERR_POS);
      }
      if (do_old)
        for (int k = 0; k < 4; ++k) d.qold[k * d.noc + c] = q[k];
      for (int k = 0; k < 4; ++k) mx[k] = mn[k] = p[k];
    } else {
      if (c < d.noc && do_old)
        for (int k = 0; k < 4; ++k) d.qold[k * d.noc + c] = d.qc[k * d.ncell + c];
      for (int k = 0; k < 4; ++k) mx[k] = mn[k] = d.qp[k * d.ncell + c];
    }
  }
  if (!do_range) return;
  block_reduce<4>(sh, mx, OpMax());
  block_reduce<4>(sh, mn, OpMin());
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) {
      d.part[blockIdx.x * 8 + k] = mx[k];
      d.part[blockIdx.x * 8 + 4 + k] = mn[k];
    }
}

// Final reduction of k_prim's partials into scal[SC_QMAX..SC_QMIN+3].
__global__ void k_range_final(Dev d, int nblk) {
  __shared__ double sh[4][BLK];
  double mx[4], mn[4];
  for (int k = 0; k < 4; ++k) { mx[k] = -INFINITY; mn[k] = INFINITY; }
  for (int b = threadIdx.x; b < nblk; b += BLK)
    for (int k = 0; k < 4; ++k) {
      double a = d.part[b * 8 + k], m = d.part[b * 8 + 4 + k];
      mx[k] = a > mx[k] ? a : mx[k];
      mn[k] = m < mn[k] ? m : mn[k];
    }
  block_reduce<4>(sh, mx, OpMax());
  block_reduce<4>(sh, mn, OpMin());
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) {
      d.scal[SC_QMAX + k] = mx[k];
      d.scal[SC_QMIN + k] = mn[k];
    }
}

// Vertex averages with pseudo-Laplacian weights, CCFVM/average.F90:11-29.
__global__ void k_avg_c2v(Dev d) {
  int p = blockIdx.x * BLK + threadIdx.x;
  if (p >= d.nop || d.err[0]) return;
  double w = 0.0, con[6] = {0, 0, 0, 0, 0, 0};
  for (int j = d.v2c_ptr[p]; j < d.v2c_ptr[p + 1]; ++j) {
    int c = d.v2c[j];
    double cw = d.wt[j];
    w = w + cw;
    for (int k = 0; k < 5; ++k) con[k] = con[k] + d.qp[k * d.ncell + c] * cw;
    con[5] = con[5] + d.cmul[c] * cw;
  }
  for (int k = 0; k < 5; ++k) d.pprim[k * d.nop + p] = con[k] / w;
  d.pmu[p] = con[5] / w;
}

// Face-averaged primitive (CCFVM/average.F90:32-36) computed where needed.
__device__ __forceinline__ double face_prim(const Dev& d, int f, int k) {
  return 0.5 * (d.pprim[k * d.nop + d.fp1[f]] + d.pprim[k * d.nop + d.fp2[f]]);
}

// Spectral radii and local time step, CCFVM/timestep.F90:98-177: each cell
// sums its faces' radii in ascending face order; the `in` cell also stores
// the face radius used by LU-SGS.
__global__ void k_timestep(Dev d, Gas<double> g, double vll, double cfl) {
  int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= d.noc || d.err[0]) return;
  int nf = d.cnf[c] & 0xff;
  double la = 0.0;
  for (int s = 0; s < nf; ++s) {
    int f = d.sl_face[s * d.noc + c];
    double r0 = face_prim(d, f, 0), u0 = face_prim(d, f, 1), v0 = face_prim(d, f, 2),
           p0 = face_prim(d, f, 3);
    double a0 = sqrt(g.gamma * p0 / r0);
    double un = u0 * d.fnx[f] + v0 * d.fny[f];
    double ll = d.farea[f] * (fabs(un) + a0);
    ll = ll + vll * d.fmu[f] / r0;
    la = la + ll;
    if (d.sl_nbr[s * d.noc + c] >= 0) d.fla[f] = ll;
  }
  d.la[c] = la;
  double dt = cfl * d.ccv[c] / la;
  d.dt[c] = dt;
  d.dta[c] = dt / d.ccv[c];
}

__device__ __forceinline__ double vk_limit(double a, double b, double tol) {
  double nr = a * a + 2.0 * a * b + tol * tol;
  double dr = a * a + 2.0 * b * b + a * b + tol * tol;
  return nr / dr;
}

// Venkatakrishnan limiter of one cell with the ratchet,
// CCFVM/limiter.F90:123-296 (kappa 0.05 of the global range).
__device__ __forceinline__ void limit_cell(const Dev& d, int c, const double* qpc,
                                           const double* gc, double* phi) {
  int info = d.cnf[c], nf = info & 0xff, nb = info >> 8;
  double umax[4], umin[4];
  for (int j = 0; j < 4; ++j) umax[j] = umin[j] = qpc[j];
  for (int s = 0; s < nf; ++s) {
    int n = d.sl_nbr[s * d.noc + c];
    if (s < nf - nb) {
      n = n >= 0 ? n : ~n;
      for (int j = 0; j < 4; ++j) {
        double q = d.qp[j * d.ncell + n];
        umax[j] = umax[j] > q ? umax[j] : q;
        umin[j] = umin[j] < q ? umin[j] : q;
      }
    } else {
      int f = d.sl_face[s * d.noc + c];
      for (int j = 0; j < 4; ++j) {
        double v = face_prim(d, f, j);
        umax[j] = umax[j] > v ? umax[j] : v;
        umin[j] = umin[j] < v ? umin[j] : v;
      }
    }
  }
  const double kappa = 0.05;
  double cx = d.ccx[c], cy = d.ccy[c];
  for (int j = 0; j < 4; ++j) phi[j] = d.phi[j * d.noc + c];
  for (int s = 0; s < nf; ++s) {
    int f = d.sl_face[s * d.noc + c];
    double dx = d.fcx[f] - cx, dy = d.fcy[f] - cy;
    for (int j = 0; j < 4; ++j) {
      double tol = kappa * (d.scal[SC_QMAX + j] - d.scal[SC_QMIN + j]);
      double dl = gc[2 * j] * dx + gc[2 * j + 1] * dy;
      double p0 = 1.0;
      if (dl > 0.0) p0 = vk_limit(umax[j] - qpc[j], dl, tol);
      else if (dl < 0.0) p0 = vk_limit(umin[j] - qpc[j], dl, tol);
      phi[j] = phi[j] < p0 ? phi[j] : p0;
    }
  }
}

// Least-squares gradient (CCFVM/gradient.F90:2-93) of one cell, optionally
// followed by the limiter; ascending face order, boundary faces last.
__global__ void k_grad_lsq(Dev d, int do_grad, int do_limit) {
  int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= d.noc || d.err[0]) return;
  double qpc[5];
  for (int k = 0; k < 5; ++k) qpc[k] = d.qp[k * d.ncell + c];
  double gc[10];
  if (do_grad) {
    for (int k = 0; k < 10; ++k) gc[k] = 0.0;
    int nf = d.cnf[c] & 0xff;
    double cx = d.ccx[c], cy = d.ccy[c];
    double r11 = d.r11[c], r12 = d.r12[c], r22 = d.r22[c];
    for (int s = 0; s < nf; ++s) {
      int f = d.sl_face[s * d.noc + c];
      int n = d.sl_nbr[s * d.noc + c];
      bool in_side = n >= 0;
      n = in_side ? n : ~n;
      double dx = d.ccx[n] - cx, dy = d.ccy[n] - cy;
      double alfa1 = dx / r11 / r11;
      double alfa2 = f < d.nfi ? (dy - dx * r12 / r11) / (r22 * r22)
                               : (dy - dx * r12 / r11) / r22 / r22;
      double wx = alfa1 - alfa2 * r12 / r11;
      double wy = alfa2;
      if (in_side) {
        for (int k = 0; k < 5; ++k) {
          double var = d.qp[k * d.ncell + n] - qpc[k];
          gc[2 * k] = gc[2 * k] + var * wx;
          gc[2 * k + 1] = gc[2 * k + 1] + var * wy;
        }
      } else {
        for (int k = 0; k < 5; ++k) {
          double var = qpc[k] - d.qp[k * d.ncell + n];
          gc[2 * k] = gc[2 * k] - var * wx;
          gc[2 * k + 1] = gc[2 * k + 1] - var * wy;
        }
      }
    }
    for (int k = 0; k < 10; ++k) d.grad[k * d.noc + c] = gc[k];
  } else {
    for (int k = 0; k < 10; ++k) gc[k] = d.grad[k * d.noc + c];
  }
  if (do_limit) {
    double phi[4];
    limit_cell(d, c, qpc, gc, phi);
    for (int j = 0; j < 4; ++j) d.phi[j * d.noc + c] = phi[j];
  }
}

template <int FLUX>
__device__ __forceinline__ bool face_flux(const Gas<double>& g, double* ql, double* qr, double nx,
                                          double ny, double area, double* F) {
  if (FLUX == UFX_FLUX_ROE) return roe_flux(g, ql, qr, nx, ny, area, F);
  if (FLUX == UFX_FLUX_AUSM) ausm_flux(g, ql, qr, nx, ny, F);
  if (FLUX == UFX_FLUX_VANLEER) vanleer_flux(g, ql, qr, nx, ny, area, F);
  if (FLUX == UFX_FLUX_RUSANOV) rusanov_flux(g, ql, qr, nx, ny, area, F);
  return true;
}

__device__ __forceinline__ void set_err(const Dev& d, int code, int cell) {
  if (atomicCAS(d.err, 0, code) == 0) d.err[1] = cell;
}

// Residual of one cell (CCFVM/fvresidual.F90:16-171) as a gather over its
// faces: boundary faces first (farfield state or wall ghost, flux, ghost
// update, wall force), then interior faces with the MUSCL reconstruction of
// recon (:190-252) on both sides.  mode 0 stores the residual; mode 1 also
// applies the Runge-Kutta stage update of CCFVM/main.F90:92-101 in place
// (no other cell reads this cell's qc during a stage); mode 2 does both and
// adds the squared residual to a block partial for the norm.
template <int FLUX>
__global__ void k_residual(Dev d, Gas<double> g, int viscous_wall, int iwall, double t_wall,
                           int mode, double ak, double bk) {
  __shared__ double sh[1][BLK];
  int c = blockIdx.x * BLK + threadIdx.x;
  double sq = 0.0;
  if (c < d.noc && !d.err[0]) {
    int info = d.cnf[c], nf = info & 0xff, nb = info >> 8;
    double qpc[4], gc[8], phic[4];
    for (int k = 0; k < 4; ++k) qpc[k] = d.qp[k * d.ncell + c];
    for (int k = 0; k < 8; ++k) gc[k] = d.grad[k * d.noc + c];
    for (int k = 0; k < 4; ++k) phic[k] = d.phi[k * d.noc + c];
    double cx = d.ccx[c], cy = d.ccy[c];
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    bool ok = true;
    // boundary faces, CCFVM/fvresidual.F90:27-130
    for (int s = nf - nb; s < nf; ++s) {
      int f = d.sl_face[s * d.noc + c];
      int gh = d.sl_nbr[s * d.noc + c];
      int bc = d.fbc[f];
      if (bc != FARF && bc != WALL) continue;
      double nx = d.fnx[f], ny = d.fny[f], area = d.farea[f];
      double dx = d.fcx[f] - cx, dy = d.fcy[f] - cy;
      double ql[4], qr[4], F[4];
      for (int i = 0; i < 4; ++i) ql[i] = qpc[i] + phic[i] * (gc[2 * i] * dx + gc[2 * i + 1] * dy);
      double gq[4], gp[5];
      if (bc == FARF) {
        farfield_state(g, ql, nx, ny, qr);
        ok = face_flux<FLUX>(g, ql, qr, nx, ny, area, F) && ok;
        prim2con(g, qr, gq);
        con2prim(g, gq, gp);
      } else {
        double q[4];
        for (int k = 0; k < 4; ++k) q[k] = d.qc[k * d.ncell + c];
        if (!viscous_wall) {
          // slip_wall, CCFVM/boundCond.F90:46-81
          double rho = q[1], u = q[2] / q[1], v = q[3] / q[1];
          double un = u * nx + v * ny;
          gq[0] = q[0];
          gq[1] = q[1];
          gq[2] = rho * u - 2.0 * un * nx;
          gq[3] = rho * v - 2.0 * un * ny;
          con2prim(g, gq, gp);
        } else if (iwall == 1) {
          // adiabatic no-slip, CCFVM/boundCond.F90:31-37
          gq[0] = q[0];
          gq[1] = q[1];
          gq[2] = -q[2];
          gq[3] = -q[3];
          con2prim(g, gq, gp);
        } else {
          // isothermal no-slip, CCFVM/boundCond.F90:19-30
          double pin[5];
          con2prim(g, q, pin);
          double tg = 2.0 * t_wall - pin[4];
          if (tg <= 0.0) tg = 0.01 * t_wall;
          gp[0] = pin[3] / (pin[0] * g.znd);
          gp[1] = -pin[1];
          gp[2] = -pin[2];
          gp[3] = pin[3];
          gp[4] = tg;
          prim2con(g, gp, gq);
        }
        for (int i = 0; i < 4; ++i) qr[i] = gp[i];
        ok = face_flux<FLUX>(g, ql, qr, nx, ny, area, F) && ok;
        int k = f - d.nfi;
        d.bflux[k] = F[2];
        d.bflux[d.nbf + k] = F[3];
      }
      for (int i = 0; i < 4; ++i) d.qc[i * d.ncell + gh] = gq[i];
      for (int i = 0; i < 5; ++i) d.qp[i * d.ncell + gh] = gp[i];
      for (int i = 0; i < 4; ++i) r[i] = r[i] + F[i];
    }
    // interior faces, CCFVM/fvresidual.F90:133-171
    const double zhi = 1.0 / 3.0;
    for (int s = 0; s < nf - nb; ++s) {
      int f = d.sl_face[s * d.noc + c];
      int n = d.sl_nbr[s * d.noc + c];
      bool in_side = n >= 0;
      n = in_side ? n : ~n;
      double qpn[4], gn[8], phin[4];
      for (int k = 0; k < 4; ++k) qpn[k] = d.qp[k * d.ncell + n];
      for (int k = 0; k < 8; ++k) gn[k] = d.grad[k * d.noc + n];
      for (int k = 0; k < 4; ++k) phin[k] = d.phi[k * d.noc + n];
      double fx = d.fcx[f], fy = d.fcy[f];
      const double *qa = in_side ? qpc : qpn, *qb = in_side ? qpn : qpc;
      const double *ga = in_side ? gc : gn, *gb = in_side ? gn : gc;
      const double *pa = in_side ? phic : phin, *pb = in_side ? phin : phic;
      double ax = in_side ? cx : d.ccx[n], ay = in_side ? cy : d.ccy[n];
      double bx = in_side ? d.ccx[n] : cx, by = in_side ? d.ccy[n] : cy;
      double dlx = fx - ax, dly = fy - ay, drx = fx - bx, dry = fy - by;
      double ql[4], qr[4], F[4];
      for (int i = 0; i < 4; ++i) {
        double alfa = ga[2 * i] * dlx + ga[2 * i + 1] * dly;
        double beta = qb[i] - qa[i];
        ql[i] = qa[i] + 0.5 * pa[i] * (zhi * beta + (1.0 - zhi) * alfa);
        alfa = gb[2 * i] * drx + gb[2 * i + 1] * dry;
        beta = qa[i] - qb[i];
        qr[i] = qb[i] + 0.5 * pb[i] * (zhi * beta + (1.0 - zhi) * alfa);
      }
      ok = face_flux<FLUX>(g, ql, qr, d.fnx[f], d.fny[f], d.farea[f], F) && ok;
      if (in_side)
        for (int i = 0; i < 4; ++i) r[i] = r[i] + F[i];
      else
        for (int i = 0; i < 4; ++i) r[i] = r[i] - F[i];
    }
    if (!ok) set_err(d, DERR_ROE, c + 1);
    for (int i = 0; i < 4; ++i) d.res[i * d.noc + c] = r[i];
    if (mode >= 1) {
      double dta = d.dta[c];
      for (int i = 0; i < 4; ++i) {
        double c1 = d.qold[i * d.noc + c], c2 = d.qc[i * d.ncell + c];
        d.qc[i * d.ncell + c] = ak * c1 + bk * (c2 - dta * r[i]);
      }
    }
    if (mode == 2)
      for (int i = 0; i < 4; ++i) sq = sq + r[i] * r[i];
  }
  if (mode == 2) {
    block_reduce<1>(sh, &sq, OpAdd());
    if (threadIdx.x == 0) d.part[blockIdx.x] = sq;
  }
}

// Runge-Kutta stage update alone, CCFVM/main.F90:92-101.
__global__ void k_rk_update(Dev d, double ak, double bk) {
  int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= d.noc || d.err[0]) return;
  double dta = d.dta[c];
  for (int i = 0; i < 4; ++i) {
    double c1 = d.qold[i * d.noc + c], c2 = d.qc[i * d.ncell + c];
    d.qc[i * d.ncell + c] = ak * c1 + bk * (c2 - dta * d.res[i * d.noc + c]);
  }
}

// Sum of squared residuals per block (CCFVM/solution.F90:152-157).
__global__ void k_sumsq(Dev d) {
  __shared__ double sh[1][BLK];
  int c = blockIdx.x * BLK + threadIdx.x;
  double sq = 0.0;
  if (c < d.noc)
    for (int i = 0; i < 4; ++i) {
      double r = d.res[i * d.noc + c];
      sq = sq + r * r;
    }
  block_reduce<1>(sh, &sq, OpAdd());
  if (threadIdx.x == 0) d.part[blockIdx.x] = sq;
}

// L2 norm from the partials (CCFVM/solution.F90:159-166): stores the raw norm,
// the iteration-1 normalisation and the normalised value of iteration k.
__global__ void k_norm_final(Dev d, int nblk, int first, double* hist, int k) {
  __shared__ double sh[1][BLK];
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += BLK) s = s + d.part[b];
  block_reduce<1>(sh, &s, OpAdd());
  if (threadIdx.x == 0) {
    double fres = sqrt(s / d.noc / 4);
    d.scal[SC_NORM] = fres;
    if (first) d.scal[SC_FRES1] = fres;
    double f1 = d.scal[SC_FRES1];
    if (hist) hist[k] = f1 >= 2.220446049250313e-16 ? fres / f1 : fres;
  }
}

// Positivity of the interior primitives, CCFVM/main.F90:384-411: the first
// failing cell (lowest index) goes to err[2].
__global__ void k_positivity(Dev d) {
  int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= d.noc) return;
  if (d.qp[c] <= 0.0 || d.qp[3 * d.ncell + c] <= 0.0) atomicMin(&d.err[2], c + 1);
}

// Minimum local time step (the screen's global dt), CCFVM/timestep.F90:111-119.
__global__ void k_dtmin(Dev d) {
  __shared__ double sh[1][BLK];
  double m = INFINITY;
  for (int c = threadIdx.x; c < d.noc; c += BLK) m = d.dt[c] < m ? d.dt[c] : m;
  block_reduce<1>(sh, &m, OpMin());
  if (threadIdx.x == 0) d.scal[SC_DTMIN] = m;
}

}  // namespace ufx

using namespace ufx;

// --------------------------------------------------------------- context
struct ufx_cc {
  Dev d{};
  Params p{};
  Gas<double> g{};
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int nblk_cell = 0, nblk_all = 0;
  int fres1_set = 0;
  double* d_hist = nullptr;
  int hist_cap = 0;
  double fres1_host = 0.0;
  std::vector<int> wall_faces;     // boundary ordinals of wall faces
  // kernel timing: event pool and (category, begin, end) records
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Rec { int cat; size_t a, b; };
  std::vector<Rec> recs;
};

namespace {

hipEvent_t prof_event(ufx_cc* x) {
  if (x->ev_used == x->ev_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    x->ev_pool.push_back(e);
  }
  return x->ev_pool[x->ev_used++];
}

int prof_begin(ufx_cc* x, int cat) {
  if (!x->prof) return -1;
  size_t i = x->ev_used;
  hipEvent_t e = prof_event(x);
  if (!e || hipEventRecord(e, x->stream) != hipSuccess) return -1;
  x->recs.push_back({cat, i, i});
  return (int)x->recs.size() - 1;
}

void prof_end(ufx_cc* x, int r) {
  if (r < 0) return;
  size_t i = x->ev_used;
  hipEvent_t e = prof_event(x);
  if (!e || hipEventRecord(e, x->stream) != hipSuccess) return;
  x->recs[r].b = i;
}

}  // namespace

// Launch `kern` on the context's stream, timed under category `cat` when
// profiling is on.
#define LAUNCH(x, cat, kern, grid, ...)                                  \
  do {                                                                   \
    int r_ = prof_begin((x), (cat));                                     \
    kern<<<(grid), BLK, 0, (x)->stream>>>(__VA_ARGS__);                  \
    prof_end((x), r_);                                                   \
  } while (0)

namespace {

template <typename T>
int dalloc(ufx_cc* x, T** p, size_t n) {
  if (n == 0) n = 1;
  HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
  x->allocs.push_back((void*)*p);
  HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), x->stream));
  return UFX_OK;
}

template <typename T>
int dupload(ufx_cc* x, const T** dst, const std::vector<T>& src) {
  T* p = nullptr;
  int rc = dalloc(x, &p, src.size());
  if (rc) return rc;
  if (!src.empty())
    HIPCHK(hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, x->stream));
  *dst = p;
  return UFX_OK;
}

int nblocks(int n) { return (n + BLK - 1) / BLK; }

int launch_check() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return UFX_ERR_HIP;
  }
  return UFX_OK;
}

int check_ctx(ufx_cc* x) {
  if (!x) {
    set_error("null context");
    return UFX_ERR_ARG;
  }
  return UFX_OK;
}

int device_error(ufx_cc* x) {
  int e[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(e, x->d.err, sizeof(e), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (e[0] == DERR_POS) e[1] = e[2];
  if (e[0] == DERR_ROE) {
    set_error("Roe scheme: sonic speed is negative (cell " + std::to_string(e[1]) + ")");
    return UFX_ERR_STATE;
  }
  if (e[0] == DERR_POS) {
    set_error("Density/pressure is negative at cell " + std::to_string(e[1]));
    return UFX_ERR_STATE;
  }
  if (e[0] != 0) {
    set_error("device error " + std::to_string(e[0]) + " at cell " + std::to_string(e[1]));
    return UFX_ERR_STATE;
  }
  return UFX_OK;
}

// Per-name descriptor of a state array for set/get: device pointer, number of
// variables, stride between variables (entities), entity count.
struct Field {
  double* ptr;
  int nvar;
  int stride;
  int count;
  bool writable;
};

bool field(ufx_cc* x, const char* name, Field& f) {
  Dev& d = x->d;
  std::string n(name);
  if (n == "qc") f = {d.qc, 4, d.ncell, d.ncell, true};
  else if (n == "qp") f = {d.qp, 5, d.ncell, d.ncell, true};
  else if (n == "qold") f = {d.qold, 4, d.noc, d.noc, true};
  else if (n == "res") f = {d.res, 4, d.noc, d.noc, true};
  else if (n == "phi") f = {d.phi, 4, d.noc, d.noc, true};
  else if (n == "grad") f = {d.grad, 10, d.noc, d.noc, true};
  else if (n == "dt") f = {d.dt, 1, d.noc, d.noc, true};
  else if (n == "la") f = {d.la, 1, d.noc, d.noc, true};
  else if (n == "fc_la") f = {d.fla, 1, d.nof, d.nof, true};
  else if (n == "fc_mu") f = {d.fmu, 1, d.nof, d.nof, true};
  else if (n == "cell_mul") f = {d.cmul, 1, d.ncell, d.ncell, true};
  else if (n == "pt_prim") f = {d.pprim, 5, d.nop, d.nop, true};
  else if (n == "fc_grad") f = {d.fgrad, 10, d.nof, d.nof, true};
  else return false;
  return true;
}

}  // namespace

extern "C" {

const char* ufx_last_error(void) { return g_err.c_str(); }

int ufx_cc_geometric(int nop, int noc, int nfi, int nbf, const double* px, const double* py,
                     const double* ccx, const double* ccy, const double* cv, const int* fi_p1,
                     const int* fi_p2, const int* fi_in, const int* fi_out, const int* fi_bc,
                     const int* fb_p1, const int* fb_p2, const int* fb_in, const double* fb_nx,
                     const double* fb_ny, const int* fb_bc, ufx_mesh** mesh) {
  if (!mesh || !px || !py || !ccx || !ccy || !cv || (nfi > 0 && (!fi_p1 || !fi_p2 || !fi_in || !fi_out)) ||
      (nbf > 0 && (!fb_p1 || !fb_p2 || !fb_in || !fb_nx || !fb_ny || !fb_bc))) {
    set_error("ufx_cc_geometric: null argument");
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
  int rc = cc_geometric(in, m->m, err);
  if (rc) {
    delete m;
    set_error(err);
    return rc;
  }
  *mesh = m;
  return UFX_OK;
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
  if (n == "pt_x" || n == "pt_y" || n == "pt_nv2c") v = m.nop;
  if (n == "pt_v2c" || n == "pt_wt") v = (int64_t)m.v2c.size();
  if (n == "counts") v = 4;
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
  else if (n == "pt_x") putd(m.px);
  else if (n == "pt_y") putd(m.py);
  else if (n == "pt_nv2c") { for (int i = 0; i < m.nop; ++i) di[i] = m.v2c_ptr[i + 1] - m.v2c_ptr[i]; }
  else if (n == "pt_v2c") put1(m.v2c);
  else if (n == "pt_wt") putd(m.wt);
  else if (n == "oldnum") put1(m.oldnum);
  else if (n == "counts") { di[0] = m.nop; di[1] = m.noc; di[2] = m.nof(); di[3] = m.nbf; }
  else { set_error("ufx_mesh_get: unknown array " + n); return UFX_ERR_ARG; }
  return UFX_OK;
}

void ufx_mesh_destroy(ufx_mesh* mesh) { delete mesh; }

int ufx_cc_destroy(ufx_cc* x) {
  if (!x) return UFX_OK;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  for (void* p : x->allocs) (void)hipFree(p);
  if (x->d_hist) (void)hipFree(x->d_hist);
  for (hipEvent_t e : x->ev_pool) (void)hipEventDestroy(e);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return UFX_OK;
}

int ufx_cc_create(const ufx_mesh* mesh, const ufx_cc_params* prm, ufx_cc** out) {
  if (!mesh || !prm || !out) { set_error("ufx_cc_create: null argument"); return UFX_ERR_ARG; }
  const Mesh& m = mesh->m;
  if (m.kind != UFX_KIND_CC) { set_error("ufx_cc_create: not a cell-centred mesh"); return UFX_ERR_ARG; }
  if (prm->flux < UFX_FLUX_ROE || prm->flux > UFX_FLUX_RUSANOV) { set_error("unknown flux"); return UFX_ERR_ARG; }
  if (prm->grad != UFX_GRAD_LSQR) {
    set_error("gradient method not implemented on the device yet");
    return UFX_ERR_UNSUPPORTED;
  }
  if (prm->flow_type != UFX_FLOW_INVISCID || prm->irs || prm->timemode != UFX_TIME_RK3 ||
      prm->precision != UFX_PREC_FP64) {
    set_error("option not implemented on the device yet");
    return UFX_ERR_UNSUPPORTED;
  }
  // per-cell face slots (ascending face order), boundary-face count
  const int noc = m.noc, nof = m.nof(), ncell = m.ncell();
  std::vector<int> sl_face(MAXF * (size_t)noc, -1), sl_nbr(MAXF * (size_t)noc, 0), cnf(noc, 0);
  {
    std::vector<int> cnt(noc, 0), nb(noc, 0);
    for (int f = 0; f < nof; ++f) {
      int sides[2] = {m.fin[f], f < m.nfi ? m.fout[f] : -1};
      for (int s = 0; s < 2; ++s) {
        int c = sides[s];
        if (c < 0) continue;
        if (cnt[c] >= MAXF) {
          set_error("cell " + std::to_string(c + 1) + " has more than 4 faces");
          return UFX_ERR_UNSUPPORTED;
        }
        int other = s == 0 ? m.fout[f] : m.fin[f];
        sl_face[(size_t)cnt[c] * noc + c] = f;
        sl_nbr[(size_t)cnt[c] * noc + c] = s == 0 ? other : ~other;
        cnt[c]++;
        if (f >= m.nfi) nb[c]++;
      }
    }
    for (int c = 0; c < noc; ++c) cnf[c] = cnt[c] | (nb[c] << 8);
  }
  ufx_cc* x = new ufx_cc();
  x->p = Params{};
  hipError_t e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    delete x;
    return UFX_ERR_HIP;
  }
  Dev& d = x->d;
  d.nop = m.nop; d.noc = noc; d.nfi = m.nfi; d.nbf = m.nbf; d.nof = nof; d.ncell = ncell;
  int rc = 0;
#define UP(dst, src) if ((rc = dupload(x, &d.dst, src))) { ufx_cc_destroy(x); return rc; }
#define AL(dst, n) if ((rc = dalloc(x, &d.dst, (size_t)(n)))) { ufx_cc_destroy(x); return rc; }
  UP(px, m.px); UP(py, m.py);
  UP(fp1, m.fp1); UP(fp2, m.fp2); UP(fin, m.fin); UP(fout, m.fout); UP(fbc, m.fbc);
  UP(fnx, m.fnx); UP(fny, m.fny); UP(farea, m.farea); UP(fcx, m.fcx); UP(fcy, m.fcy);
  UP(ccx, m.ccx); UP(ccy, m.ccy); UP(ccv, m.ccv); UP(ccov, m.ccov);
  UP(r11, m.r11); UP(r12, m.r12); UP(r22, m.r22); UP(nc2c, m.nc2c);
  UP(sl_face, sl_face); UP(sl_nbr, sl_nbr); UP(cnf, cnf);
  UP(v2c_ptr, m.v2c_ptr); UP(v2c, m.v2c); UP(wt, m.wt);
  AL(qc, 4 * (size_t)ncell); AL(qp, 5 * (size_t)ncell); AL(qold, 4 * (size_t)noc);
  AL(res, 4 * (size_t)noc); AL(grad, 10 * (size_t)noc); AL(phi, 4 * (size_t)noc);
  AL(dt, noc); AL(la, noc); AL(dta, noc); AL(fla, nof); AL(fmu, nof); AL(cmul, ncell);
  AL(pprim, 5 * (size_t)m.nop); AL(pmu, m.nop); AL(fgrad, 10 * (size_t)nof);
  AL(bflux, 2 * (size_t)std::max(1, m.nbf));
  x->nblk_cell = nblocks(noc);
  x->nblk_all = nblocks(ncell);
  AL(part, 8 * (size_t)std::max(x->nblk_all, 1));
  AL(scal, SC_NSCAL);
  {
    int* er = nullptr;
    if ((rc = dalloc(x, &er, 4))) { ufx_cc_destroy(x); return rc; }
    d.err = er;
    int init[4] = {0, 0, 0x7fffffff, 0};
    if (hipMemcpyAsync(er, init, sizeof(init), hipMemcpyHostToDevice, x->stream) != hipSuccess ||
        hipStreamSynchronize(x->stream) != hipSuccess) {
      set_error("upload of the error word failed");
      ufx_cc_destroy(x);
      return UFX_ERR_HIP;
    }
  }
#undef UP
#undef AL
  // constants and free stream, CCFVM/solution.F90:4-49
  Params& p = x->p;
  p.gamma = 1.4;
  p.gamma1 = p.gamma - 1.0;
  p.m_inf = prm->mach;
  double pi = 4.0 * std::atan(1.0);
  double aoa = prm->aoa_deg * pi / 180.0;
  p.aoa = aoa;
  p.xref = prm->xref;
  p.yref = prm->yref;
  double qinf[4];
  qinf[1] = 1.0;
  qinf[2] = std::cos(aoa) * 1.0;
  qinf[3] = std::sin(aoa) * 1.0;
  p.p_inf = 1.0 / (p.gamma * p.m_inf * p.m_inf);
  qinf[0] = p.p_inf / (p.gamma - 1.0) + 0.5;
  p.r_inf = qinf[1];
  p.u_inf = qinf[2] / qinf[1];
  p.v_inf = qinf[3] / qinf[1];
  p.znd = 1.0 / (p.gamma * p.m_inf * p.m_inf);
  p.mref = std::max((double)0.32f, 0.5 * p.m_inf);
  p.t_wall = prm->t_wall / 300.0;
  p.rey = prm->reynolds;
  p.flow = prm->flow_type; p.iwall = prm->iwall; p.flux = prm->flux; p.grad = prm->grad;
  p.ilimit = prm->ilimit; p.irs = prm->irs; p.timemode = prm->timemode;
  p.precision = prm->precision; p.cfl_max = prm->cfl_max;
  p.vll = 0.0;
  if (p.flow != UFX_FLOW_INVISCID) p.vll = 2.0 * p.gamma / (0.72 * p.rey * 0.25);
  Gas<double>& g = x->g;
  g.gamma = p.gamma; g.gamma1 = p.gamma1; g.znd = p.znd; g.r_inf = p.r_inf; g.u_inf = p.u_inf;
  g.v_inf = p.v_inf; g.p_inf = p.p_inf; g.mref = p.mref;
  // initialize: free stream everywhere, limiter 1 (CCFVM/solution.F90:78-81)
  {
    std::vector<double> q(4 * (size_t)ncell), ph(4 * (size_t)noc, 1.0);
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < ncell; ++c) q[(size_t)k * ncell + c] = qinf[k];
    e = hipMemcpyAsync(d.qc, q.data(), q.size() * 8, hipMemcpyHostToDevice, x->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d.phi, ph.data(), ph.size() * 8, hipMemcpyHostToDevice, x->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
    if (e != hipSuccess) {
      set_error(std::string("upload: ") + hipGetErrorString(e));
      ufx_cc_destroy(x);
      return UFX_ERR_HIP;
    }
  }
  for (int k = 0; k < m.nbf; ++k)
    if (m.fbc[m.nfi + k] == WALL) x->wall_faces.push_back(k);
  *out = x;
  return UFX_OK;
}

int ufx_cc_set(ufx_cc* x, const char* name, const double* src) {
  if (int rc = check_ctx(x)) return rc;
  Field f;
  if (!name || !src || !field(x, name, f)) { set_error(std::string("ufx_cc_set: unknown field ") + (name ? name : "")); return UFX_ERR_ARG; }
  std::vector<double> t((size_t)f.nvar * f.stride, 0.0);
  for (int c = 0; c < f.count; ++c)
    for (int k = 0; k < f.nvar; ++k) t[(size_t)k * f.stride + c] = src[(size_t)c * f.nvar + k];
  HIPCHK(hipMemcpyAsync(f.ptr, t.data(), t.size() * 8, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return UFX_OK;
}

int ufx_cc_get(ufx_cc* x, const char* name, double* dst) {
  if (int rc = check_ctx(x)) return rc;
  Field f;
  if (!name || !dst || !field(x, name, f)) { set_error(std::string("ufx_cc_get: unknown field ") + (name ? name : "")); return UFX_ERR_ARG; }
  std::vector<double> t((size_t)f.nvar * f.stride);
  HIPCHK(hipMemcpyAsync(t.data(), f.ptr, t.size() * 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int c = 0; c < f.count; ++c)
    for (int k = 0; k < f.nvar; ++k) dst[(size_t)c * f.nvar + k] = t[(size_t)k * f.stride + c];
  return UFX_OK;
}

int ufx_cc_prim(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  LAUNCH(x, UFX_K_PRIM, k_prim, x->nblk_all, x->d, x->g, 1, 0, 0);
  return launch_check();
}

int ufx_cc_save_old(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  LAUNCH(x, UFX_K_PRIM, k_prim, x->nblk_all, x->d, x->g, 0, 1, 0);
  return launch_check();
}

int ufx_cc_avg_c2v(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  LAUNCH(x, UFX_K_AVG, k_avg_c2v, nblocks(x->d.nop), x->d);
  return launch_check();
}

int ufx_cc_time_step(ufx_cc* x, double cfl) {
  if (int rc = check_ctx(x)) return rc;
  LAUNCH(x, UFX_K_TSTEP, k_timestep, x->nblk_cell, x->d, x->g, x->p.vll, cfl);
  return launch_check();
}

static int range_pass(ufx_cc* x) {
  LAUNCH(x, UFX_K_PRIM, k_prim, x->nblk_all, x->d, x->g, 0, 0, 1);
  LAUNCH(x, UFX_K_RANGE, k_range_final, 1, x->d, x->nblk_all);
  return launch_check();
}

int ufx_cc_limit(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  if (int rc = range_pass(x)) return rc;
  LAUNCH(x, UFX_K_GRAD, k_grad_lsq, x->nblk_cell, x->d, 0, 1);
  return launch_check();
}

int ufx_cc_gradient(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  if (x->p.ilimit)
    if (int rc = range_pass(x)) return rc;
  LAUNCH(x, UFX_K_GRAD, k_grad_lsq, x->nblk_cell, x->d, 1, x->p.ilimit);
  return launch_check();
}

static int residual(ufx_cc* x, int mode, double ak, double bk) {
  const Dev& d = x->d;
  int viscous = x->p.flow != UFX_FLOW_INVISCID;
  switch (x->p.flux) {
    case UFX_FLUX_ROE:
      LAUNCH(x, UFX_K_STAGE, k_residual<UFX_FLUX_ROE>, x->nblk_cell, d, x->g, viscous, x->p.iwall, x->p.t_wall, mode, ak, bk);
      break;
    case UFX_FLUX_AUSM:
      LAUNCH(x, UFX_K_STAGE, k_residual<UFX_FLUX_AUSM>, x->nblk_cell, d, x->g, viscous, x->p.iwall, x->p.t_wall, mode, ak, bk);
      break;
    case UFX_FLUX_VANLEER:
      LAUNCH(x, UFX_K_STAGE, k_residual<UFX_FLUX_VANLEER>, x->nblk_cell, d, x->g, viscous, x->p.iwall, x->p.t_wall, mode, ak, bk);
      break;
    default:
      LAUNCH(x, UFX_K_STAGE, k_residual<UFX_FLUX_RUSANOV>, x->nblk_cell, d, x->g, viscous, x->p.iwall, x->p.t_wall, mode, ak, bk);
      break;
  }
  return launch_check();
}

static const double RK_A[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
static const double RK_B[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};

int ufx_cc_fvresidual(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  return residual(x, 0, 0.0, 0.0);
}

int ufx_cc_rk_update(ufx_cc* x, int stage) {
  if (int rc = check_ctx(x)) return rc;
  if (stage < 1 || stage > 3) { set_error("stage must be 1..3"); return UFX_ERR_ARG; }
  LAUNCH(x, UFX_K_STAGE, k_rk_update, x->nblk_cell, x->d, RK_A[stage - 1], RK_B[stage - 1]);
  return launch_check();
}

int ufx_cc_rk_stage(ufx_cc* x, int stage) {
  if (int rc = check_ctx(x)) return rc;
  if (stage < 1 || stage > 3) { set_error("stage must be 1..3"); return UFX_ERR_ARG; }
  return residual(x, 1, RK_A[stage - 1], RK_B[stage - 1]);
}

int ufx_cc_lusgs(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  set_error("LU-SGS is not implemented on the device yet");
  return UFX_ERR_UNSUPPORTED;
}

static int ensure_hist(ufx_cc* x, int n) {
  if (n <= x->hist_cap) return UFX_OK;
  if (x->d_hist) HIPCHK(hipFree(x->d_hist));
  x->d_hist = nullptr;
  HIPCHK(hipMalloc((void**)&x->d_hist, (size_t)n * sizeof(double)));
  x->hist_cap = n;
  return UFX_OK;
}

int ufx_cc_residue(ufx_cc* x, int iter, double* fres, double* fres_raw) {
  if (int rc = check_ctx(x)) return rc;
  if (int rc = ensure_hist(x, 1)) return rc;
  LAUNCH(x, UFX_K_NORM, k_sumsq, x->nblk_cell, x->d);
  LAUNCH(x, UFX_K_NORM, k_norm_final, 1, x->d, x->nblk_cell, iter == 1, x->d_hist, 0);
  if (int rc = launch_check()) return rc;
  double h = 0.0, raw = 0.0;
  HIPCHK(hipMemcpyAsync(&h, x->d_hist, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(&raw, x->d.scal + SC_NORM, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (fres) *fres = h;
  if (fres_raw) *fres_raw = raw;
  return device_error(x);
}

int ufx_cc_check_positivity(ufx_cc* x, int* bad) {
  if (int rc = check_ctx(x)) return rc;
  int big = 0x7fffffff;
  HIPCHK(hipMemcpyAsync(x->d.err + 2, &big, sizeof(int), hipMemcpyHostToDevice, x->stream));
  LAUNCH(x, UFX_K_OTHER, k_positivity, x->nblk_cell, x->d);
  if (int rc = launch_check()) return rc;
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, x->d.err + 2, sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (bad) *bad = e == 0x7fffffff ? 0 : e;
  if (e != 0x7fffffff) {
    set_error("Density/pressure is negative at point " + std::to_string(e));
    return UFX_ERR_STATE;
  }
  return UFX_OK;
}

int ufx_cc_forces(ufx_cc* x, double* out, double* cp, double* cf) {
  if (int rc = check_ctx(x)) return rc;
  if (!out) { set_error("ufx_cc_forces: null output"); return UFX_ERR_ARG; }
  const Dev& d = x->d;
  const Params& p = x->p;
  LAUNCH(x, UFX_K_OTHER, k_dtmin, 1, x->d);
  if (int rc = launch_check()) return rc;
  int nbf = d.nbf;
  std::vector<double> bf(2 * (size_t)std::max(1, nbf)), dtmin(1);
  std::vector<int> fin(nbf), fp1(nbf), fp2(nbf);
  std::vector<double> fnx(nbf), fny(nbf), fa(nbf), px(d.nop), py(d.nop);
  std::vector<double> qp4((size_t)d.ncell);
  if (nbf > 0) {
    HIPCHK(hipMemcpyAsync(bf.data(), d.bflux, 2 * (size_t)nbf * 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(fin.data(), d.fin + d.nfi, nbf * 4, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(fp1.data(), d.fp1 + d.nfi, nbf * 4, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(fp2.data(), d.fp2 + d.nfi, nbf * 4, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(fnx.data(), d.fnx + d.nfi, nbf * 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(fny.data(), d.fny + d.nfi, nbf * 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(fa.data(), d.farea + d.nfi, nbf * 8, hipMemcpyDeviceToHost, x->stream));
  }
  HIPCHK(hipMemcpyAsync(px.data(), d.px, d.nop * 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(py.data(), d.py, d.nop * 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(qp4.data(), d.qp + 3 * (size_t)d.ncell, d.ncell * 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(dtmin.data(), d.scal + SC_DTMIN, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  // find_cl_cd, CCFVM/main.F90:263-355: pressure of the adjacent cell for
  // the pressure integral, momentum fluxes of the last residual for the
  // flux-based coefficients the reference prints (:351-353)
  double dyna = 0.5 * p.gamma * p.p_inf * p.m_inf * p.m_inf;
  double fx = 0.0, fy = 0.0, cm = 0.0, fx1 = 0.0, fy1 = 0.0;
  int iw = 0;
  for (int k : x->wall_faces) {
    double pwall = qp4[fin[k]];
    double xi = 0.5 * (px[fp1[k]] + px[fp2[k]]);
    double yi = 0.5 * (py[fp1[k]] + py[fp2[k]]);
    double xa = xi - p.xref, ya = yi - p.yref;
    double area = fa[k], nx = fnx[k], ny = fny[k];
    double cpv = (pwall - p.p_inf) / dyna;
    double dcx = cpv * nx * area, dcy = cpv * ny * area;
    fx = fx + pwall * nx * area;
    fy = fy + pwall * ny * area;
    cm = cm + dcx * ya - dcy * xa;
    if (cp) cp[iw] = -cpv;
    if (cf) cf[iw] = 0.0;
    fx1 = fx1 + bf[k];
    fy1 = fy1 + bf[nbf + k];
    ++iw;
  }
  double aoa = p.aoa;
  out[0] = (fy * std::cos(aoa) - fx * std::sin(aoa)) / dyna;
  out[1] = (fy * std::sin(aoa) + fx * std::cos(aoa)) / dyna;
  out[2] = (fy1 * std::cos(aoa) - fx1 * std::sin(aoa)) / dyna;
  out[3] = (fy1 * std::sin(aoa) + fx1 * std::cos(aoa)) / dyna;
  out[4] = cm;
  out[5] = fx1;
  out[6] = fy1;
  out[7] = dtmin[0];
  return device_error(x);
}

int ufx_cc_iterate(ufx_cc* x, int iter0, int niter, const double* cfl, double* hist) {
  if (int rc = check_ctx(x)) return rc;
  if (niter < 0 || (niter > 0 && !cfl)) { set_error("ufx_cc_iterate: bad arguments"); return UFX_ERR_ARG; }
  if (int rc = ensure_hist(x, std::max(1, niter))) return rc;
  const Dev& d = x->d;
  for (int k = 0; k < niter; ++k) {
    int iter = iter0 + 1 + k;
    // prim refresh + qold (save_old commutes with avg_c2v/time_step) + range
    LAUNCH(x, UFX_K_PRIM, k_prim, x->nblk_all, d, x->g, 1, 1, x->p.ilimit);
    LAUNCH(x, UFX_K_AVG, k_avg_c2v, nblocks(d.nop), d);
    LAUNCH(x, UFX_K_TSTEP, k_timestep, x->nblk_cell, d, x->g, x->p.vll, cfl[k]);
    if (x->p.ilimit)
      LAUNCH(x, UFX_K_RANGE, k_range_final, 1, d, x->nblk_all);
    LAUNCH(x, UFX_K_GRAD, k_grad_lsq, x->nblk_cell, d, 1, x->p.ilimit);
    for (int s = 0; s < 3; ++s)
      if (int rc = residual(x, s == 2 ? 2 : 1, RK_A[s], RK_B[s])) return rc;
    LAUNCH(x, UFX_K_NORM, k_norm_final, 1, d, x->nblk_cell, iter == 1, x->d_hist, k);
    if (int rc = launch_check()) return rc;
  }
  if (hist && niter > 0) {
    HIPCHK(hipMemcpyAsync(hist, x->d_hist, (size_t)niter * 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return device_error(x);
  }
  return UFX_OK;
}

int ufx_cc_profile(ufx_cc* x, int enable) {
  if (int rc = check_ctx(x)) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  x->prof = enable != 0;
  x->recs.clear();
  x->ev_used = 0;
  return UFX_OK;
}

int ufx_cc_profile_read(ufx_cc* x, double* ms, int* count) {
  if (int rc = check_ctx(x)) return rc;
  if (!ms || !count) { set_error("ufx_cc_profile_read: null output"); return UFX_ERR_ARG; }
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int k = 0; k < UFX_K_NCAT; ++k) { ms[k] = 0.0; count[k] = 0; }
  for (const auto& r : x->recs) {
    if (r.a == r.b) continue;
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, x->ev_pool[r.a], x->ev_pool[r.b]));
    ms[r.cat] += t;
    count[r.cat] += 1;
  }
  x->recs.clear();
  x->ev_used = 0;
  return UFX_OK;
}

int ufx_cc_sync(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  return device_error(x);
}

void* ufx_cc_stream(ufx_cc* x) { return x ? (void*)x->stream : nullptr; }

}  // extern "C"
