// Kernels of the vertex-centred (median-dual) hot path, FP64 (gfx950).
// Included once, by nc_solver.hip.  Each kernel cites the reference routine
// it performs.  The reference scatters edge contributions into both end
// points in edge order; here every vertex gathers over its edges in the
// same order (its v2f list is in ascending edge order), so each per-vertex
// sum keeps the reference's order.
//
// HBM layout (vertices in storage order, a Hilbert curve through them; vref
// holds their reference numbers):
//   vrec [nop][18] f64  primitives rho,u,v,p | limiter values | gradient
//                       (d/dx, d/dy per variable) | x, y
//   qc, qold, res [nop][4] f64 (conserved order E, rho, rho u, rho v)
//   la, dt, cv [nop] f64;  fla [nof] f64
//   erec [nof][3] f64   dual-face nx, ny, length;  ep [nof] int2 (p1, p2);
//                       edges stored in first-touch vertex order
//   v2f [..] int        edges of a vertex, bit-inverted when it is p2
//   v2fo [..] int2      the same with the vertex at the other end
//   v2v [..] int, wxy [..][2] f64   least-squares neighbours and weights
#pragma once

#include <hip/hip_cooperative_groups.h>

#include "cc_device.h"
#include "cc_physics.h"

namespace ufx {

constexpr int NW = 18;  // vertex record width (same offsets as the cell record)
constexpr int NW32 = 20;  // the same in FP32, padded to a multiple of 16 bytes
constexpr double NC_EPS_SP = 1.1920928955078125e-07;  // epsilon(1.0), NCFVM/All_modules.f90:9

struct NCFree {
  double gamma, gamma1, r_inf, u_inf, v_inf, p_inf, a_inf;
};

struct DevNC {
  int nop, nof;
  double* vrec;
  // compact per-vertex arrays beside the records: the primitives save_old
  // writes, the limiter values, the coordinates (the gradient pass reads
  // these and copies primitives and limiter values into the records)
  double *vqp, *vphc;
  const double* vxy;
  float *vqp32, *vphc32;
  const float* vxy32;
  double *qc, *qold, *res, *la, *dt, *dd, *dqc;
  const double* cv;
  const double* erec;
  const double* egg;         // [nof][2] nx ds, ny ds: the Green-Gauss pass's edge terms
  const int2* ep;
  const int* ebc;            // per edge: its index in the boundary-edge list, or -1
  // block-shared residual: per block of BLK vertices the edges they touch
  // (ascending stored edge number), and per v2f entry its edge's position
  // in its vertex's block list (bit-inverted when the vertex is p2)
  const int *be_ptr, *be_edge, *sl_loc;
  int be_max;
  const int* bedge;          // wall and farfield edges (stored edge numbers)
  const int* bkind;          // their kind (WALL or FARF)
  double* bclos;             // [nbnd][2 ends][5] closure terms (momentum/flux 4, radius 1)
  int nbnd;
  double* fla;
  const int *v2f_ptr, *v2f;
  const int2* v2fo;          // per v2f entry: (edge, bit-inverted when the vertex is p2; other end)
  const int *v2v_ptr, *v2v;
  const double* wxy;
  const int* vref;           // reference number of each stored vertex
  // the mixed form's records (FP32), used instead of vrec / erec / wxy
  float* vrec32;          // [nop][NW32]
  const float* erec32;    // [nof][4] nx, ny, length, pad
  const float* egg32;     // [nof][2]
  const float* wxy32;     // [..][2]
  const float2* ggs32;    // [nop] sum over the vertex's dual faces of +-n ds (FP64 sum)
  const double2* pxy;     // vertex coordinates in FP64 (wall forces)
  double* part;  // reduction partials [nblk][4]
  double* scal;  // [0] fres1
  int* err;
  // level-ordered sweep data (positions: forward levels, then backward)
  // Streamed sweeps: position s (0 .. 2 nop - 1) is the s-th vertex of the
  // forward levels then of the backward levels; every array below is
  // position-major (a wave's lanes read consecutive addresses), with sw_q
  // entry slots per position for its swept neighbours in edge order.
  const int* sw_vert;   // [2 nop] vertex of each position
  const int* sw_cnt;    // [2 nop] swept neighbours of each position
  const int* sw_nb;     // [sw_q][2 nop] neighbour's position, bit-inverted when the vertex is p2
  const int2* sw_pe;    // [sw_q][2 nop] (edge, neighbour vertex)
  const int* sw_d0;     // [2 nop] forward position of the vertex (read by backward positions)
  double2* sw_ent;      // [sw_q][6][2 nop] (nx,ny) (ds,la_e) F(qold_o) (2) qold_o (2)
  double2* sw_own;      // [4][2 nop] res (2), qold (2)
  double* sw_dd;        // [2 nop] diagonal
  double* sw_dq;        // [2 nop][4] dq of each position
  int sw_q;             // entry slots per position
  // multicolour sweeps: positions 0 .. nop - 1 are the vertices grouped by
  // colour (ascending vertex number inside a colour); per sweep direction and
  // position, mc_q entry slots (position-major) hold the neighbours of lower
  // (forward) or higher (backward) colour in edge order, as positions, with
  // the edge's geometry.  The sweeps keep their state by position: a block's
  // neighbours in another colour then lie in a narrow window of that
  // colour's positions instead of all over the vertex numbering
  const int* mc_vert;     // [nop] vertex of each position
  const int* mc_cnt;      // [2][nop] entries of each position per direction
  const int2* mc_eo;      // [2][mc_q][nop] (edge, bit-inverted when the vertex is p2; other end)
  const double2* mc_nxy;  // [2][mc_q][nop] edge nx, ny
  const double* mc_ds;    // [2][mc_q][nop] edge length
  const int* mc_pos;      // [nop] colour position of each vertex (null: no multicolour sweeps)
  double* mc_qold;        // [nop][4] old state by position (save_old writes it)
  double* mc_res;         // [nop][4] residual by position (the residual pass writes it)
  double* mc_dq;          // [nop][4] dq by position
  double* mc_dd;          // [nop] diagonal by position (the residual pass writes it)
  int mc_q;
  int mc_n0;              // positions of the first colour, whose forward dq the residual
                          // pass forms (0: the sweep launch does)
};

__device__ __forceinline__ void ld2d(const double* p, double& a, double& b) {
  double2 v = *reinterpret_cast<const double2*>(p);
  a = v.x;
  b = v.y;
}

// Copy n values (a multiple of 16 bytes, 16-byte aligned on both sides)
// from global memory into LDS, the block's threads striding over them.
template <typename R>
__device__ __forceinline__ void stage_block(R* dst, const R* src, int n) {
  constexpr int V = 16 / (int)sizeof(R);
  const int n4 = n / V;
  if constexpr (sizeof(R) == 8) {
    const double2* s2 = reinterpret_cast<const double2*>(src);
    double2* d2 = reinterpret_cast<double2*>(dst);
    for (int i = threadIdx.x; i < n4; i += BLK) d2[i] = s2[i];
  } else {
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    for (int i = threadIdx.x; i < n4; i += BLK) d4[i] = s4[i];
  }
}

// Vertex records of the FP64 form (vrec, double, NW wide) and of the mixed
// form (vrec32, float, padded to NW32 for 16-byte loads); edge records and
// least-squares weights likewise.  The FP64 instantiation of every kernel
// below keeps the reference's expression order (bit-exact); the mixed one
// evaluates the same expressions in FP32 and accumulates the residual, the
// spectral radii and the time step in FP64.
// The mixed form stores its primitives as deviations from the free stream:
// the limiter compares neighbour differences with single precision's epsilon
// (NCFVM/limiter.f90:45-46, 1.19e-7), which absolute FP32 values of the
// pressure (about 4.5 at M 0.4) cannot resolve; as deviations, differences,
// gradients and limiter inputs keep FP32's relative precision where the flow
// is nearly uniform.  The free stream is added back for the fluxes and the
// boundary closures; the Green-Gauss sum, which the reference leaves without
// a boundary closure, gets its free-stream part back from the vertex's FP64
// sum of +-n ds (zero but for rounding inside the mesh).
template <typename R> struct NCRec;
template <> struct NCRec<double> {
  static constexpr int W = NW;
  __device__ static double* rec(const DevNC& d) { return d.vrec; }
  __device__ static double* qp(const DevNC& d) { return d.vqp; }
  __device__ static double* phc(const DevNC& d) { return d.vphc; }
  __device__ static const double* xy(const DevNC& d) { return d.vxy; }
  __device__ static const double* edge(const DevNC& d) { return d.erec; }
  __device__ static const double* gg(const DevNC& d) { return d.egg; }
  __device__ static const double* lsq(const DevNC& d) { return d.wxy; }
};
template <> struct NCRec<float> {
  static constexpr int W = NW32;
  __device__ static float* rec(const DevNC& d) { return d.vrec32; }
  __device__ static float* qp(const DevNC& d) { return d.vqp32; }
  __device__ static float* phc(const DevNC& d) { return d.vphc32; }
  __device__ static const float* xy(const DevNC& d) { return d.vxy32; }
  __device__ static const float* edge(const DevNC& d) { return d.erec32; }
  __device__ static const float* gg(const DevNC& d) { return d.egg32; }
  __device__ static const float* lsq(const DevNC& d) { return d.wxy32; }
};

// An edge's nx, ny, length (and 0): three doubles per edge in FP64, four
// floats (one 16-byte load) in FP32.
template <typename R> __device__ __forceinline__ void ld_edge(const R* erec, int e, R* er);
template <> __device__ __forceinline__ void ld_edge<double>(const double* erec, int e, double* er) {
  const double* p = erec + (size_t)e * 3;
  er[0] = p[0]; er[1] = p[1]; er[2] = p[2]; er[3] = 0.0;
}
template <> __device__ __forceinline__ void ld_edge<float>(const float* erec, int e, float* er) {
  const float4 v = reinterpret_cast<const float4*>(erec)[e];
  er[0] = v.x; er[1] = v.y; er[2] = v.z; er[3] = v.w;
}

template <typename R> __device__ __forceinline__ R nc_abs(R x);
template <> __device__ __forceinline__ double nc_abs<double>(double x) { return fabs(x); }
template <> __device__ __forceinline__ float nc_abs<float>(float x) { return fabsf(x); }

// save_old, NCFVM/common.f90:88-104: primitives of every vertex (the only
// primitive refresh of an iteration, from the FP64 state) and qold = qc.
// wq = 0 (whole multicolour LU-SGS iterations, whose sweeps read the old
// state by colour position only): qold by vertex is left to
// k_nc_qold_from_mc, run when something else reads it.
template <typename R>
__global__ __launch_bounds__(BLK) void k_nc_save_old(DevNC d, NCFree g, int wq) {
  int p = cell_index();
  if (p >= d.nop) return;
  double q[4];
  ld4d(d.qc + (size_t)p * 4, q);
  if (wq) st4d(d.qold + (size_t)p * 4, q);
  // the multicolour sweeps read the old state by colour position
  if (d.mc_pos) st4d(d.mc_qold + (size_t)d.mc_pos[p] * 4, q);
  double rho = q[1], u = q[2] / q[1], v = q[3] / q[1], et = q[0];
  double qq = u * u + v * v;
  double pr = (g.gamma - 1.0) * (et - 0.5 * rho * qq);
  R o[4];
  if constexpr (sizeof(R) == 8) {
    o[0] = rho; o[1] = u; o[2] = v; o[3] = pr;
  } else {
    o[0] = (R)(rho - g.r_inf); o[1] = (R)(u - g.u_inf); o[2] = (R)(v - g.v_inf);
    o[3] = (R)(pr - g.p_inf);
  }
  stv<R, 4>(NCRec<R>::qp(d) + (size_t)p * 4, o);
}

// The residual and the LU-SGS diagonal by vertex from their copies by colour
// position.
__global__ __launch_bounds__(BLK) void k_nc_res_from_mc(DevNC d) {
  int p = cell_index();
  if (p >= d.nop) return;
  const size_t s = (size_t)d.mc_pos[p];
  double r[4];
  ld4d(d.mc_res + s * 4, r);
  st4d(d.res + (size_t)p * 4, r);
  d.dd[p] = d.mc_dd[s];
}

// qold by vertex from its copy by colour position.
__global__ __launch_bounds__(BLK) void k_nc_qold_from_mc(DevNC d) {
  int p = cell_index();
  if (p >= d.nop) return;
  double q[4];
  ld4d(d.mc_qold + (size_t)d.mc_pos[p] * 4, q);
  st4d(d.qold + (size_t)p * 4, q);
}

// The boundary closures of one wall or farfield edge, for both of its ends
// (side 0: p1, side 1: p2): the terms the residual adds right after the
// edge's van Leer flux, each already rounded as the reference's expression
// rounds it: for the wall (solid_flux, NCFVM/solid_flux.f90:1-43) the
// momentum terms pw n of its end and the spectral radius term; for the
// farfield (farfield_flux1, NCFVM/farfield_flux.f90:1-95) the flux times
// the half-edge length, the same into both ends, and the spectral radius
// term of each end.  Evaluated in FP64 in both forms.
template <typename R>
__device__ __forceinline__ void nc_closure(const DevNC& d, const NCFree& g, int k) {
  const R* qps = NCRec<R>::qp(d);
  const R* xys = NCRec<R>::xy(d);
  const int e = d.bedge[k], bc = d.bkind[k];
  const int2 pp = d.ep[e];
  const R* a = qps + (size_t)pp.x * 4;
  const R* b = qps + (size_t)pp.y * 4;
  double ax = (double)xys[(size_t)pp.x * 2], ay = (double)xys[(size_t)pp.x * 2 + 1];
  double bx = (double)xys[(size_t)pp.y * 2], by = (double)xys[(size_t)pp.y * 2 + 1];
  double ap[4] = {(double)a[0], (double)a[1], (double)a[2], (double)a[3]};
  double bp[4] = {(double)b[0], (double)b[1], (double)b[2], (double)b[3]};
  if constexpr (sizeof(R) == 4) {
    const double qi[4] = {g.r_inf, g.u_inf, g.v_inf, g.p_inf};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ap[i] = ap[i] + qi[i];
      bp[i] = bp[i] + qi[i];
    }
  }
  double* c0 = d.bclos + (size_t)k * 10;
  double* c1 = c0 + 5;
  if (bc == WALL) {
    // 3/4-1/4 end-point pressures on the half-edge right normals
    double p1w = 0.75 * ap[3] + 0.25 * bp[3];
    double p2w = 0.25 * ap[3] + 0.75 * bp[3];
    double wnx = 0.5 * (by - ay);
    double wny = -0.5 * (bx - ax);
    double area = sqrt(wnx * wnx + wny * wny);
    c0[0] = 0.0; c0[1] = 0.0; c0[2] = p1w * wnx; c0[3] = p1w * wny;
    c1[0] = 0.0; c1[1] = 0.0; c1[2] = p2w * wnx; c1[3] = p2w * wny;
    c0[4] = sqrt(g.gamma * ap[3] / ap[0]) * area;
    c1[4] = sqrt(g.gamma * bp[3] / bp[0]) * area;
  } else {
    // Riemann invariants on the half-edge left normal from the
    // edge-averaged primitives
    double fnx = -0.5 * (by - ay);
    double fny = 0.5 * (bx - ax);
    double dr = sqrt(fnx * fnx + fny * fny);
    fnx = fnx / dr;
    fny = fny / dr;
    double un_inf = g.u_inf * fnx + g.v_inf * fny;
    double rho = 0.5 * (ap[0] + bp[0]), u = 0.5 * (ap[1] + bp[1]);
    double v = 0.5 * (ap[2] + bp[2]), pr = 0.5 * (ap[3] + bp[3]);
    double as = sqrt(g.gamma * pr / rho);
    double un = u * fnx + v * fny;
    double rp = un + 2.0 * as / g.gamma1;
    double rm = un_inf - 2.0 * g.a_inf / g.gamma1;
    double unf = 0.5 * (rp + rm);
    double af = 0.25 * (rp - rm) * g.gamma1;
    double uf, vf, sf;
    if (unf > 0.0) {
      uf = u + fnx * (unf - un);
      vf = v + fny * (unf - un);
      sf = pr / pow(rho, g.gamma);
    } else {
      uf = g.u_inf + fnx * (unf - un_inf);
      vf = g.v_inf + fny * (unf - un_inf);
      sf = g.p_inf / pow(g.r_inf, g.gamma);
    }
    double rhof = pow(af * af / g.gamma / sf, 1.0 / g.gamma1);
    double pf = rhof * af * af / g.gamma;
    double ef = pf / g.gamma1 + 0.5 * rhof * (uf * uf + vf * vf);
    double fl[4];
    fl[0] = (ef + pf) * unf;
    fl[1] = rhof * unf;
    fl[2] = fl[1] * uf + pf * fnx;
    fl[3] = fl[1] * vf + pf * fny;
    double area = sqrt(fnx * fnx + fny * fny);
#pragma unroll
    for (int i = 0; i < 4; ++i) c0[i] = c1[i] = dr * fl[i];
    c0[4] = (fabs(unf) + sqrt(g.gamma * ap[3] / ap[0])) * area;
    c1[4] = (fabs(unf) + sqrt(g.gamma * bp[3] / bp[0])) * area;
  }
}

// Gradient and limiter of one vertex:
//   GRAD 1  Gradient_GG, NCFVM/gradient.f90:35-75 (edge midpoint values on
//           the dual faces, no boundary closure), / dual volume;
//   GRAD 2  Gradient_LSQR, NCFVM/gradient.f90:2-30;
// then limiter_VKN, NCFVM/limiter.f90:15-98: neighbour difference range
// (starting at -+eps of single precision), the Venkatakrishnan function
// with eps^2 = (kappa sqrt(V))^3, kappa = 0.01 in single precision, per edge
// end, ratcheted into the stored limiter value.
// Up to NCU edges of a vertex are fetched ahead of their gathers (the rest,
// rare, in a tail loop): the gathers then do not wait on the index chain.
constexpr int NCU = 8;
template <typename F>
__device__ __forceinline__ void for_edges(const DevNC& d, int j0, int n, const int2* nb, F&& f) {
#pragma unroll
  for (int k = 0; k < NCU; ++k)
    if (k < n) f(nb[k]);
  for (int k = NCU; k < n; ++k) f(d.v2fo[j0 + k]);
}

template <typename R, int GRAD, int LIM>
__global__ __launch_bounds__(BLK) void k_nc_grad_lim(DevNC d, NCFree g, int rev) {
  constexpr int W = NCRec<R>::W;
  // the primitives and coordinates of the block's vertices, staged in LDS:
  // the vertices are stored along a Hilbert curve, so most neighbours lie in
  // the same block
  __shared__ R s_qp[BLK * 4];
  __shared__ R s_xy[BLK * 2];
  // blocks past the vertices' form the boundary closures of this residual
  // evaluation (they read only primitives and coordinates, which this
  // kernel does not write)
  const int nvb = (d.nop + BLK - 1) / BLK;
  if ((int)blockIdx.x >= nvb) {
    const int k = ((int)blockIdx.x - nvb) * BLK + (int)threadIdx.x;
    if (k < d.nbnd) nc_closure<R>(d, g, k);
    return;
  }
  const int base = xcd_block_dir(blockIdx.x, nvb, rev != 0) * BLK;
  const int p = base + (int)threadIdx.x;
  const int cnt = min(BLK, d.nop - base);
  const R* qps = NCRec<R>::qp(d);
  const R* xys = NCRec<R>::xy(d);
  const R* egg = NCRec<R>::gg(d);
  R q[4], xy[2];
  if (p < d.nop) {
    ldv<R, 4>(qps + (size_t)p * 4, q);
    ldv<R, 2>(xys + (size_t)p * 2, xy);
    stv<R, 4>(s_qp + threadIdx.x * 4, q);
    stv<R, 2>(s_xy + threadIdx.x * 2, xy);
  }
  __syncthreads();
  if (p >= d.nop) return;
  auto nbr_qp = [&](int o, R* qo) {
    const unsigned l = (unsigned)(o - base);
    if (l < (unsigned)cnt) ldv<R, 4>(s_qp + l * 4, qo);
    else ldv<R, 4>(qps + (size_t)o * 4, qo);
  };
  const R x = xy[0], y = xy[1];
  R gr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int j0 = d.v2f_ptr[p], n = d.v2f_ptr[p + 1] - j0;
  int2 nb[NCU];
#pragma unroll
  for (int k = 0; k < NCU; ++k)
    if (k < n) nb[k] = d.v2fo[j0 + k];
  // the limiter's difference range over the edges, in edge order: in the
  // Green-Gauss loop (the same neighbour values), else a loop of its own
  const R eps = (R)NC_EPS_SP;
  R dumax[4], dumin[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { dumin[k] = eps; dumax[k] = -eps; }
  auto range = [&](bool first, const R* qo) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      R du = first ? qo[k] - q[k] : -(q[k] - qo[k]);
      dumax[k] = dumax[k] > du ? dumax[k] : du;
      dumin[k] = dumin[k] < du ? dumin[k] : du;
    }
  };
  if (GRAD == 1) {
    for_edges(d, j0, n, nb, [&](int2 eo) {
      bool first = eo.x >= 0;
      int e = first ? eo.x : ~eo.x, o = eo.y;
      R qo[4], dd[2];
      nbr_qp(o, qo);
      // nx ds and ny ds, the products formed once when the edges were stored
      ldv<R, 2>(egg + (size_t)e * 2, dd);
      const R dx = dd[0], dy = dd[1];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        R var = first ? R(0.5) * (q[k] + qo[k]) : R(0.5) * (qo[k] + q[k]);
        if (first) {
          gr[2 * k] = gr[2 * k] + var * dx;
          gr[2 * k + 1] = gr[2 * k + 1] + var * dy;
        } else {
          gr[2 * k] = gr[2 * k] - var * dx;
          gr[2 * k + 1] = gr[2 * k + 1] - var * dy;
        }
      }
      range(first, qo);
    });
    if constexpr (sizeof(R) == 4) {
      // the free-stream part of the open Green-Gauss sum
      const float2 sv = d.ggs32[p];
      const R qi[4] = {(R)g.r_inf, (R)g.u_inf, (R)g.v_inf, (R)g.p_inf};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gr[2 * k] = gr[2 * k] + qi[k] * sv.x;
        gr[2 * k + 1] = gr[2 * k + 1] + qi[k] * sv.y;
      }
    }
    R v = (R)d.cv[p];
#pragma unroll
    for (int k = 0; k < 8; ++k) gr[k] = gr[k] / v;
  } else {
    const R* wxy = NCRec<R>::lsq(d);
    for (int j = d.v2v_ptr[p]; j < d.v2v_ptr[p + 1]; ++j) {
      int c = d.v2v[j];
      R w[2], qo[4];
      ldv<R, 2>(wxy + (size_t)j * 2, w);
      nbr_qp(c, qo);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gr[2 * k] = gr[2 * k] + (qo[k] - q[k]) * w[0];
        gr[2 * k + 1] = gr[2 * k + 1] + (qo[k] - q[k]) * w[1];
      }
    }
  }
  if (GRAD != 1)
    for_edges(d, j0, n, nb, [&](int2 eo) {
      R qo[4];
      nbr_qp(eo.y, qo);
      range(eo.x >= 0, qo);
    });
  R phi[4];
  ldv<R, 4>(NCRec<R>::phc(d) + (size_t)p * 4, phi);
  const R kappa = (R)0.01f;
  R t = kappa * fsqrt<R>((R)d.cv[p]);
  const R tol = t * t * t;
  for_edges(d, j0, n, nb, [&](int2 eo) {
    bool first = eo.x >= 0;
    R xo[2];
    const unsigned l = (unsigned)(eo.y - base);
    if (l < (unsigned)cnt) ldv<R, 2>(s_xy + l * 2, xo);
    else ldv<R, 2>(xys + (size_t)eo.y * 2, xo);
    R d0 = first ? xo[0] - x : x - xo[0], d1 = first ? xo[1] - y : y - xo[1];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      R sum = (R(0) + gr[2 * k] * d0) + gr[2 * k + 1] * d1;
      R ph;
      if (LIM == 0) {
        R dl = first ? R(0.5) * sum : R(-0.5) * sum;
        R alfa = dl > R(0) ? dumax[k] : dumin[k];
        if (nc_abs<R>(alfa) < eps) alfa = R(0);
        R nr = alfa * alfa + R(2) * dl * alfa + tol;
        R dr = alfa * alfa + R(2) * dl * dl + dl * alfa + tol;
        ph = nr / dr;
        ph = R(1) < ph ? R(1) : ph;
      } else {
        // limiter_min_max (Barth-Jespersen), NCFVM/limiter.f90:192-260:
        // both edge ends take +1/2 of the projected gradient (:227, :241)
        R dl = R(0.5) * sum;
        if (dl > R(0)) {
          ph = dumax[k] / dl;
          ph = R(1) < ph ? R(1) : ph;
        } else if (dl < R(0)) {
          ph = dumin[k] / dl;
          ph = R(1) < ph ? R(1) : ph;
        } else {
          ph = R(1);
        }
      }
      phi[k] = phi[k] < ph ? phi[k] : ph;
    }
  });
  // the record the residual gathers: primitives, limiter values and
  // gradient in one piece; and the limiter values' compact copy
  static_assert(R_PHI == R_QP + 4 && R_G == R_PHI + 4, "record fields are adjacent");
  R o16[16];
  ldv<R, 4>(s_qp + threadIdx.x * 4, o16);  // (read back: q need not stay live)
#pragma unroll
  for (int k = 0; k < 4; ++k) o16[4 + k] = phi[k];
#pragma unroll
  for (int k = 0; k < 8; ++k) o16[8 + k] = gr[k];
  stv<R, 16>(NCRec<R>::rec(d) + (size_t)p * W + R_QP, o16);
  stv<R, 4>(NCRec<R>::phc(d) + (size_t)p * 4, phi);
}

// Spectral radius of a dual face from the Roe average of its two states,
// NCFVM/vanleer_flux.f90:103-115, times the face length.
template <typename R>
__device__ __forceinline__ R nc_roe_lambda(R gm1, R rl, R ul, R vl, R hl, R rr, R ur, R vr, R hr,
                                           R nx, R ny, R area) {
  R rt = fsqrt<R>(rr / rl);
  R u = (ul + rt * ur) / (R(1) + rt), v = (vl + rt * vr) / (R(1) + rt);
  R h = (hl + rt * hr) / (R(1) + rt);
  R a = fsqrt<R>(gm1 * (h - R(0.5) * (u * u + v * v)));
  return (nc_abs<R>(u * nx + v * ny) + a) * area;
}

// AUSM+-up on a dual face (build-defined: the vertex-centred reference has
// no compilable AUSM): the cell-centred routine's flux per unit length
// (CCFVM/ausm_up_flux.F90:2-162, unit normal), with the van Leer routine's
// Roe-averaged spectral radius.
template <typename R>
__device__ __forceinline__ void nc_ausm(const Gas<R>& ga, const R* ql, const R* qr, R nx, R ny,
                                        R area, R* F, R& ll) {
  R ql_[4] = {ql[0], ql[1], ql[2], ql[3]}, qr_[4] = {qr[0], qr[1], qr[2], qr[3]};
  ausm_flux<R>(ga, ql_, qr_, nx, ny, F);
  R hl = (ga.gamma * ql[3] / ql[0]) / ga.gamma1 + R(0.5) * (ql[1] * ql[1] + ql[2] * ql[2]);
  R hr = (ga.gamma * qr[3] / qr[0]) / ga.gamma1 + R(0.5) * (qr[1] * qr[1] + qr[2] * qr[2]);
  ll = nc_roe_lambda<R>(ga.gamma1, ql[0], ql[1], ql[2], hl, qr[0], qr[1], qr[2], hr, nx, ny, area);
}

// van Leer flux of NCFVM/vanleer_flux.f90:2-118 on one dual face between
// the reconstructed states ql, qr (primitives): F (per unit length) and the
// Roe-averaged spectral radius times the face length.
template <typename R>
__device__ __forceinline__ void nc_vanleer(R gm, R gm1, const R* ql, const R* qr, R nx, R ny,
                                           R area, R* F, R& ll) {
  R rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  R unl = ul * nx + vl * ny, ql2 = ul * ul + vl * vl;
  R al2 = gm * pl / rl, hl = al2 / gm1 + R(0.5) * ql2, cl = fsqrt<R>(al2), ml = unl / cl;
  R rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  R unr = ur * nx + vr * ny, qr2 = ur * ur + vr * vr;
  R ar2 = gm * pr / rr, hr = ar2 / gm1 + R(0.5) * qr2, cr = fsqrt<R>(ar2), mr = unr / cr;
  R plp = p5p<R>(ml, R(0)), prm = p5m<R>(mr, R(0));
  R mlp = m4p<R>(ml, R(0)), mlm = m4m<R>(ml, R(0));
  R mrp = m4p<R>(mr, R(0)), mrm = m4m<R>(mr, R(0));
  R p12 = plp * pl + prm * pr;
  R fmp = rl * cl * mlp, fmm = rr * cr * mrm;
  R fm12 = fmp + fmm;
  R delta = -(gm * (cr * mrp * mrm * pr - cl * mlp * mlm * pl) / (gm - R(1)));
  R dissi = fmp - fmm;
  R fl[4] = {hl, R(1), ul, vl}, fr[4] = {hr, R(1), ur, vr};
  R fp[4] = {delta, R(0), nx * p12, ny * p12};
#pragma unroll
  for (int i = 0; i < 4; ++i)
    F[i] = R(0.5) * fm12 * (fl[i] + fr[i]) - R(0.5) * dissi * (fr[i] - fl[i]) + fp[i];
  ll = nc_roe_lambda<R>(gm1, rl, ul, vl, hl, rr, ur, vr, hr, nx, ny, area);
}

// residue, NCFVM/common.f90:108-154: the partials of one block of vertices
// (`blk`: its index in vertex order) -- sum |res|^2, the largest per-vertex
// |res| with its (first, by reference number) vertex, and min dt.  Every
// thread of the block calls it; `have`: the thread has a vertex, whose
// residual is r and time step dt.  sh and si: LDS of 2 BLK doubles and BLK
// ints.
__device__ __forceinline__ void nc_norm_partial(const DevNC& d, double (*sh)[BLK], int* si, int blk,
                                                bool have, const double* r, double dt, int vref) {
  double s = 0.0, mx = 0.0, dmn = 1.0e20;
  int imx = 0x7fffffff;
  if (have) {
    double fr = 0.0;
    for (int j = 0; j < 4; ++j) fr = fr + r[j] * r[j];
    s = fr;
    mx = sqrt(fr);
    imx = vref;
    dmn = dt;
  }
  double v[1] = {s};
  block_reduce<1>(sh, v, OpAdd());
  // argmax, first index on ties (the reference's strict '>' scan)
  __syncthreads();
  sh[0][threadIdx.x] = mx;
  si[threadIdx.x] = imx;
  sh[1][threadIdx.x] = dmn;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double m2 = sh[0][threadIdx.x + w];
      int i2 = si[threadIdx.x + w];
      if (m2 > sh[0][threadIdx.x] || (m2 == sh[0][threadIdx.x] && i2 < si[threadIdx.x])) {
        sh[0][threadIdx.x] = m2;
        si[threadIdx.x] = i2;
      }
      double d2 = sh[1][threadIdx.x + w];
      sh[1][threadIdx.x] = d2 < sh[1][threadIdx.x] ? d2 : sh[1][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.part[blk * 4 + 0] = v[0];
    d.part[blk * 4 + 1] = sh[0][0];
    d.part[blk * 4 + 2] = (double)si[0];
    d.part[blk * 4 + 3] = sh[1][0];
  }
}

// A wall or farfield edge's closure (solid_flux, farfield_flux1), added
// right after the edge's flux as the reference does (bi: the edge's index
// in the boundary-edge list, -1 for other edges); its terms were formed per
// boundary edge by the gradient launch.
__device__ __forceinline__ void nc_add_closure(const DevNC& d, int bi, bool first, double* r,
                                               double& la) {
  if (bi < 0) return;
  const double* c = d.bclos + ((size_t)bi * 2 + (first ? 0 : 1)) * 5;
  if (d.bkind[bi] == WALL) {
    r[2] = r[2] + c[2];
    r[3] = r[3] + c[3];
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = r[i] + c[i];
  }
  la = la + c[4];
}

// The end of a vertex's residual: time_step2 (NCFVM/timestep.f90:5-37), the
// LU-SGS diagonal D = V/dt + omega/2 lambda (omega = 1.5,
// NCFVM/lusgs.f90:4-131; the multicolour sweeps read it once per vertex
// instead of gathering its three inputs), the residual, and for mode 1 the
// RK stage update of NCFVM/main.f90:59-70 in place.  Mode 2 (multicolour
// LU-SGS) stores residual and diagonal by colour position only
// (k_nc_res_from_mc forms them by vertex when something reads them).
// Returns dt.
__device__ __forceinline__ double nc_vertex_finish(const DevNC& d, int p, const double* r,
                                                   double la, double cfl, int mode, double ak,
                                                   double bk) {
  const double cvp = d.cv[p];
  const double dt = cfl * cvp / la;
  d.la[p] = la;
  d.dt[p] = dt;
  const double ddv = cvp / dt + 0.5 * 1.5 * la;
  if (mode == 0) d.dd[p] = ddv;
  if (mode != 2) st4d(d.res + (size_t)p * 4, r);
  if (d.mc_pos && mode != 1) {
    // the multicolour sweeps read residual and diagonal by colour position
    const size_t s = (size_t)d.mc_pos[p];
    st4d(d.mc_res + s * 4, r);
    d.mc_dd[s] = ddv;
    if (mode == 2 && s < (size_t)d.mc_n0) {
      // the first colour has no earlier neighbours: its forward dq is that
      // of k_nc_lusgs_mc with an empty sum, formed here
      double dq[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) dq[i] = (-r[i] - 0.0) / ddv;
      st4d(d.mc_dq + s * 4, dq);
    }
  }
  if (mode == 1) {
    double dtbyarea = dt / cvp, q[4], o[4];
    ld4d(d.qc + (size_t)p * 4, q);
    ld4d(d.qold + (size_t)p * 4, o);
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = ak * o[i] + bk * (q[i] - dtbyarea * r[i]);
    st4d(d.qc + (size_t)p * 4, q);
  }
  return dt;
}

// Residual of one vertex, NCFVM/fvresidual.f90:34-64 as a gather over its
// edges in edge order: reconstruction of both ends (:40-54), van Leer on the
// dual face (+ into p1, - into p2, spectral radius into both), then for a
// wall edge solid_flux (NCFVM/solid_flux.f90:1-43) and for a farfield edge
// farfield_flux1 (NCFVM/farfield_flux.f90:1-95); then time_step2
// (NCFVM/timestep.f90:5-37).  The two boundary closures are evaluated in
// FP64 in both forms (a few thousand edges).
//   mode 0: store res, la, dt;  mode 1: also the RK stage update of
//   NCFVM/main.f90:59-70 in place (no vertex reads another's qc in a stage).
// Built for 4 waves per SIMD (128 registers, a few values spilled): the
// third wave's latency hiding pays more than the spills cost.
template <typename R, int FLUX>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4)))
void k_nc_residual(DevNC d, NCFree g, Gas<double> ga64, double cfl, int mode, double ak,
                   double bk) {
  constexpr int W = NCRec<R>::W;
  // the block's vertex records, staged in LDS with one coalesced copy: the
  // vertices are stored along a Hilbert curve, so most neighbours a vertex
  // gathers lie in its own block, and its own record is read from here on
  // every edge instead of being held in registers
  __shared__ __align__(16) R s_rec[BLK * W];
  static_assert(BLK * W * sizeof(R) >= BLK * (2 * sizeof(double) + sizeof(int)),
                "the norm's reduction reuses the staged records' LDS");
  const int base = xcd_block(blockIdx.x, gridDim.x) * BLK;
  const int p = base + (int)threadIdx.x;
  const int cnt = min(BLK, d.nop - base);
  const R* recs = NCRec<R>::rec(d);
  stage_block<R>(s_rec, recs + (size_t)base * W, cnt * W);
  __syncthreads();
  const bool have = p < d.nop;
  const R* erec = NCRec<R>::edge(d);
  const R* me = s_rec + threadIdx.x * W;
  double r[4] = {0.0, 0.0, 0.0, 0.0}, la = 0.0, dt = 0.0;
  if (have) {
  const int j0 = d.v2f_ptr[p], n = d.v2f_ptr[p + 1] - j0;
  auto edge = [&](int2 eo) {
    const bool first = eo.x >= 0;
    const int e = first ? eo.x : ~eo.x;
    const unsigned l = (unsigned)(eo.y - base);
    const R* op = l < (unsigned)cnt ? s_rec + l * W : recs + (size_t)eo.y * W;
    R ot[W];
    ldv<R, W>(op, ot);
    R er[4];
    ld_edge<R>(erec, e, er);
    // reconstruction and flux in FP64 from the records (FP32 in the mixed
    // form: rounding absolute states to FP32 would put a noise floor of a
    // few 1e-7 under the residual).  a = p1's record, b = p2's: the edge
    // vector b - a, then each end's extrapolation to the dual face, p1 with
    // +, p2 with -, exactly as the reference writes them
    const double nx = er[0], ny = er[1], ds = er[2];
    const double mx = (double)me[R_CEN], my = (double)me[R_CEN + 1];
    const double ox = (double)ot[R_CEN], oy = (double)ot[R_CEN + 1];
    const double d0 = first ? ox - mx : mx - ox, d1 = first ? oy - my : my - oy;
    double ql[4], qr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double s = (0.0 + (double)me[R_G + 2 * k] * d0) + (double)me[R_G + 2 * k + 1] * d1;
      double t = 0.5 * s * (double)me[R_PHI + k];
      const double so = first ? (double)me[R_QP + k] + t : (double)me[R_QP + k] - t;
      s = (0.0 + (double)ot[R_G + 2 * k] * d0) + (double)ot[R_G + 2 * k + 1] * d1;
      t = 0.5 * s * (double)ot[R_PHI + k];
      const double sn = first ? (double)ot[R_QP + k] - t : (double)ot[R_QP + k] + t;
      ql[k] = first ? so : sn;
      qr[k] = first ? sn : so;
    }
    if constexpr (sizeof(R) == 4) {
      // deviations -> states
      const double qi[4] = {g.r_inf, g.u_inf, g.v_inf, g.p_inf};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ql[k] = ql[k] + qi[k];
        qr[k] = qr[k] + qi[k];
      }
    }
    double F[4], ll;
    if (FLUX == 2) nc_ausm<double>(ga64, ql, qr, nx, ny, ds, F, ll);
    else nc_vanleer<double>(g.gamma, g.gamma1, ql, qr, nx, ny, ds, F, ll);
    if (first) {
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = r[i] + F[i] * ds;
      d.fla[e] = ll;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = r[i] - F[i] * ds;
    }
    la = la + ll;
    nc_add_closure(d, d.ebc[e], first, r, la);
  };
  int2 nb[NCU];
#pragma unroll
  for (int k = 0; k < NCU; ++k)
    if (k < n) nb[k] = d.v2fo[j0 + k];
  for_edges(d, j0, n, nb, edge);
  dt = nc_vertex_finish(d, p, r, la, cfl, mode, ak, bk);
  }
  // the residual norm's partials of this block from the registers (the
  // staged records' LDS reused), as k_nc_norm forms them from memory
  __syncthreads();
  double(*sh)[BLK] = reinterpret_cast<double(*)[BLK]>(s_rec);
  int* si = reinterpret_cast<int*>(s_rec) + 4 * BLK;
  nc_norm_partial(d, sh, si, base / BLK, have, r, dt, have ? d.vref[p] : 0);
}

// The norm partials from the stored residual and time steps (the residual
// pass forms the same partials itself; this is for ufx_nc_residue).
__global__ __launch_bounds__(BLK) void k_nc_norm(DevNC d) {
  __shared__ double sh[2][BLK];
  __shared__ int si[BLK];
  const int p = blockIdx.x * BLK + threadIdx.x;
  const bool have = p < d.nop;
  double r[4] = {0.0, 0.0, 0.0, 0.0}, dt = 0.0;
  if (have) {
    ld4d(d.res + (size_t)p * 4, r);
    dt = d.dt[p];
  }
  nc_norm_partial(d, sh, si, blockIdx.x, have, r, dt, have ? d.vref[p] : 0);
}

// Final reduction: hist[k] = {fres (normalised by the iteration-1 value),
// fres raw, fresi, iresi (1-based), dtglobal}.
__global__ __launch_bounds__(BLK) void k_nc_norm_final(DevNC d, int nblk, int first,
                                                       double* hist, int k) {
  __shared__ double sh[2][BLK];
  __shared__ int si[BLK];
  double s = 0.0, mx = 0.0, dmn = 1.0e20;
  int imx = 0x7fffffff;
  // a thread's partials in block order, NU of them loaded together
  constexpr int NU = 8;
  for (int b0 = threadIdx.x; b0 < nblk; b0 += NU * BLK) {
    double2 pa[NU], pb[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int b = b0 + u * BLK;
      if (b < nblk) {
        pa[u] = reinterpret_cast<const double2*>(d.part)[2 * b];
        pb[u] = reinterpret_cast<const double2*>(d.part)[2 * b + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      if (b0 + u * BLK >= nblk) break;
      s = s + pa[u].x;
      double m2 = pa[u].y;
      int i2 = (int)pb[u].x;
      if (m2 > mx || (m2 == mx && i2 < imx)) { mx = m2; imx = i2; }
      double d2 = pb[u].y;
      dmn = d2 < dmn ? d2 : dmn;
    }
  }
  double v[1] = {s};
  block_reduce<1>(sh, v, OpAdd());
  __syncthreads();
  sh[0][threadIdx.x] = mx;
  si[threadIdx.x] = imx;
  sh[1][threadIdx.x] = dmn;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double m2 = sh[0][threadIdx.x + w];
      int i2 = si[threadIdx.x + w];
      if (m2 > sh[0][threadIdx.x] || (m2 == sh[0][threadIdx.x] && i2 < si[threadIdx.x])) {
        sh[0][threadIdx.x] = m2;
        si[threadIdx.x] = i2;
      }
      double d2 = sh[1][threadIdx.x + w];
      sh[1][threadIdx.x] = d2 < sh[1][threadIdx.x] ? d2 : sh[1][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double fres = sqrt(v[0] / (double)d.nop);
    if (first) d.scal[0] = fres;
    double f1 = d.scal[0];
    double* h = hist + (size_t)k * 5;
    h[0] = f1 != 0.0 ? fres / f1 : fres;
    h[1] = fres;
    h[2] = sh[0][0];
    h[3] = sh[0][0] > 0.0 ? (double)(si[0] + 1) : 0.0;
    h[4] = sh[1][0];
  }
}

// The same residual with each edge's flux evaluated once per block instead
// of once per end: the edges the block's vertices touch (a list per block,
// `be_ptr`/`be_edge`; vertices are stored along a Hilbert curve, so most
// edges have both ends in the block) are evaluated by the block's threads
// into LDS -- F ds and the spectral radius, from p1's and p2's records
// exactly as the reference reconstructs them -- then every vertex sums its
// edges from LDS in its own edge order (`sl_loc`: the entry's position in
// the block's list, bit-inverted when the vertex is p2), with the
// boundary closures after their edges.  The same sums in the same order as
// the per-vertex gather: the results are identical.
template <typename R, int FLUX>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4)))
void k_nc_residual_blk(DevNC d, NCFree g, Gas<double> ga64, double cfl, int mode, double ak,
                       double bk) {
  constexpr int W = NCRec<R>::W;
  extern __shared__ __align__(16) double s_f[];  // [be_max][5]: F ds (4), radius
  int* s_bi = reinterpret_cast<int*>(s_f + (size_t)d.be_max * 5);  // [be_max] boundary index
  const int nvb = (d.nop + BLK - 1) / BLK;
  const int blk = xcd_block(blockIdx.x, nvb);
  const int base = blk * BLK;
  const int p = base + (int)threadIdx.x;
  const R* recs = NCRec<R>::rec(d);
  const R* erec = NCRec<R>::edge(d);
  const int e0 = d.be_ptr[blk], ne = d.be_ptr[blk + 1] - e0;
  // each thread's next edge and its end points are fetched while the
  // current edge's flux is evaluated
  int e_nx = 0, bi_nx = 0;
  int2 pp_nx = make_int2(0, 0);
  R er_nx[4] = {R(0), R(0), R(0), R(0)};
  if ((int)threadIdx.x < ne) {
    e_nx = d.be_edge[e0 + threadIdx.x];
    pp_nx = d.ep[e_nx];
    ld_edge<R>(erec, e_nx, er_nx);
    bi_nx = d.ebc[e_nx];
  }
  for (int k = threadIdx.x; k < ne; k += BLK) {
    const int e = e_nx, bi = bi_nx;
    const int2 pp = pp_nx;
    R a[W], b[W], er[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) er[i] = er_nx[i];
    ldv<R, W>(recs + (size_t)pp.x * W, a);
    ldv<R, W>(recs + (size_t)pp.y * W, b);
    if (k + BLK < ne) {
      e_nx = d.be_edge[e0 + k + BLK];
      pp_nx = d.ep[e_nx];
      ld_edge<R>(erec, e_nx, er_nx);
      bi_nx = d.ebc[e_nx];
    }
    const double nx = er[0], ny = er[1], ds = er[2];
    const double d0 = (double)b[R_CEN] - (double)a[R_CEN];
    const double d1 = (double)b[R_CEN + 1] - (double)a[R_CEN + 1];
    double ql[4], qr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double t = (0.0 + (double)a[R_G + 2 * i] * d0) + (double)a[R_G + 2 * i + 1] * d1;
      ql[i] = (double)a[R_QP + i] + 0.5 * t * (double)a[R_PHI + i];
      t = (0.0 + (double)b[R_G + 2 * i] * d0) + (double)b[R_G + 2 * i + 1] * d1;
      qr[i] = (double)b[R_QP + i] - 0.5 * t * (double)b[R_PHI + i];
    }
    if constexpr (sizeof(R) == 4) {
      const double qi[4] = {g.r_inf, g.u_inf, g.v_inf, g.p_inf};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ql[i] = ql[i] + qi[i];
        qr[i] = qr[i] + qi[i];
      }
    }
    double F[4], ll;
    if (FLUX == 2) nc_ausm<double>(ga64, ql, qr, nx, ny, ds, F, ll);
    else nc_vanleer<double>(g.gamma, g.gamma1, ql, qr, nx, ny, ds, F, ll);
    double* o = s_f + (size_t)k * 5;
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = F[i] * ds;
    o[4] = ll;
    s_bi[k] = bi;
    d.fla[e] = ll;
  }
  __syncthreads();
  const bool have = p < d.nop;
  double r[4] = {0.0, 0.0, 0.0, 0.0}, la = 0.0, dt = 0.0;
  if (have) {
    auto add = [&](int loc) {
      const bool first = loc >= 0;
      const int k = first ? loc : ~loc;
      const double* o = s_f + (size_t)k * 5;
      if (first) {
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = r[i] + o[i];
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = r[i] - o[i];
      }
      la = la + o[4];
      nc_add_closure(d, s_bi[k], first, r, la);
    };
    // the first NCU list positions loaded together, then the sums in order
    const int j0 = d.v2f_ptr[p], n = d.v2f_ptr[p + 1] - j0;
    int locs[NCU];
#pragma unroll
    for (int u = 0; u < NCU; ++u)
      if (u < n) locs[u] = d.sl_loc[j0 + u];
#pragma unroll
    for (int u = 0; u < NCU; ++u)
      if (u < n) add(locs[u]);
    for (int j = j0 + NCU; j < j0 + n; ++j) add(d.sl_loc[j]);
    dt = nc_vertex_finish(d, p, r, la, cfl, mode, ak, bk);
  }
  __syncthreads();
  double(*sh)[BLK] = reinterpret_cast<double(*)[BLK]>(s_f);
  int* si = reinterpret_cast<int*>(s_f + 2 * BLK);
  nc_norm_partial(d, sh, si, blk, have, r, dt, have ? d.vref[p] : 0);
}

// LU-SGS of NCFVM/lusgs.f90:4-131 (omega = 1.5): diagonal D = V/dt +
// omega/2 lambda and dq = 0.
__global__ __launch_bounds__(BLK) void k_nc_lusgs_diag(DevNC d) {
  int p = cell_index();
  if (p >= d.nop) return;
  d.dd[p] = d.cv[p] / d.dt[p] + 0.5 * 1.5 * d.la[p];
  double z[4] = {0.0, 0.0, 0.0, 0.0};
  st4d(d.dqc + (size_t)p * 4, z);
}

// normalflux of NCFVM/lusgs.f90:135-155 (energy as (E/rho) rho).
__device__ __forceinline__ void nc_normalflux(const NCFree& g, const double* q, double nx,
                                              double ny, double ds, double* f) {
  double rho = q[1], u = q[2] / q[1], v = q[3] / q[1], et = q[0];
  double e = et / rho;
  double qq = u * u + v * v;
  double p = (g.gamma - 1.0) * (et - 0.5 * rho * qq);
  double ee = e * rho;
  double un = u * nx + v * ny;
  f[0] = (ee + p) * un;
  f[1] = rho * un;
  f[2] = f[1] * u + p * nx;
  f[3] = f[1] * v + p * ny;
#pragma unroll
  for (int i = 0; i < 4; ++i) f[i] = f[i] * ds;
}

// The update of one vertex in the forward (dir 0) or backward (dir 1)
// sweep, NCFVM/lusgs.f90:46-122: the off-diagonal sum over the neighbours
// already swept (lower reference number forward, higher backward), then
// dq and qc = qold + dq.
__device__ __forceinline__ void nc_lusgs_vertex(const DevNC& d, const NCFree& g, int it, int dir) {
  double cres[4] = {0.0, 0.0, 0.0, 0.0};
  const int j0 = d.v2f_ptr[it], n = d.v2f_ptr[it + 1] - j0;
  int2 nb[NCU];
#pragma unroll
  for (int k = 0; k < NCU; ++k)
    if (k < n) nb[k] = d.v2fo[j0 + k];
  for_edges(d, j0, n, nb, [&](int2 eo) {
    bool first = eo.x >= 0;
    int e = first ? eo.x : ~eo.x, o = eo.y;
    double sign = first ? 1.0 : -1.0;
    if (dir == 0 ? d.vref[o] < d.vref[it] : d.vref[o] > d.vref[it]) {
      double f1[4], f2[4], qo[4], qn[4], dq[4];
      const double nx = d.erec[(size_t)e * 3], ny = d.erec[(size_t)e * 3 + 1];
      const double ds = d.erec[(size_t)e * 3 + 2];
      ld4d(d.qold + (size_t)o * 4, qo);
      ld4d(d.qc + (size_t)o * 4, qn);
      ld4d(d.dqc + (size_t)o * 4, dq);
      nc_normalflux(g, qo, nx, ny, ds, f1);
      nc_normalflux(g, qn, nx, ny, ds, f2);
      double lam = d.fla[e];
#pragma unroll
      for (int i = 0; i < 4; ++i) cres[i] = cres[i] + 0.5 * (sign * (f2[i] - f1[i]) - 1.5 * lam * dq[i]);
    }
  });
  double dd = d.dd[it], res[4], dq[4], qo[4], q[4];
  ld4d(d.res + (size_t)it * 4, res);
  ld4d(d.qold + (size_t)it * 4, qo);
  if (dir == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) dq[i] = (-res[i] - cres[i]) / dd;
  } else {
    ld4d(d.dqc + (size_t)it * 4, dq);
#pragma unroll
    for (int i = 0; i < 4; ++i) dq[i] = dq[i] - cres[i] / dd;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] = qo[i] + dq[i];
  st4d(d.dqc + (size_t)it * 4, dq);
  st4d(d.qc + (size_t)it * 4, q);
}

// One level of a sweep as its own launch: the vertices of a level depend
// only on vertices of earlier levels, so they update concurrently and the
// sweep equals the reference's sequential one.
__global__ __launch_bounds__(BLK) void k_nc_lusgs_level(DevNC d, NCFree g, const int* verts,
                                                        int n, int dir) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n) return;
  nc_lusgs_vertex(d, g, verts[t], dir);
}

// ---- multicolour sweeps
// The sweeps of NCFVM/lusgs.f90:46-122 with the vertices ordered by colour
// instead of by number: no edge joins two vertices of one colour, so the
// vertices of a colour depend only on vertices of other colours and one
// launch updates them all.  Forward (dir 0, colours ascending): the
// already-swept neighbours are those of lower colour, whose dq this sweep
// has written; the diagonal is formed here and kept for the reverse sweep.
// Reverse (dir 1, colours descending): the neighbours of higher colour.  A
// neighbour's state is rebuilt as qold + dq, the sum that defines it, so the
// arithmetic is the reference loop's, term for term, in that vertex order.
// fin (forward sweep of the last colour, which has no later neighbours):
// also the reverse sweep's update of these vertices, whose sum is empty,
// qc = qold + dq.
constexpr int MCU = 8;
__global__ __launch_bounds__(BLK) void k_nc_lusgs_mc(DevNC d, NCFree g, int s0, int n, int dir,
                                                     int rev, int fin) {
  const int t = xcd_block_dir(blockIdx.x, gridDim.x, rev != 0) * BLK + threadIdx.x;
  if (t >= n) return;
  const int np = d.nop, s = s0 + t, it = d.mc_vert[s];
  const int cnt = d.mc_cnt[(size_t)dir * np + s];
  const size_t base = (size_t)dir * d.mc_q * np + s;
  double cres[4] = {0.0, 0.0, 0.0, 0.0};
  auto term = [&](int2 eo, double2 nxy, double ds) {
    const bool first = eo.x >= 0;
    const int e = first ? eo.x : ~eo.x, o = eo.y;
    const double sign = first ? 1.0 : -1.0;
    double f1[4], f2[4], qo[4], qn[4], dq[4];
    ld4d(d.mc_qold + (size_t)o * 4, qo);
    ld4d(d.mc_dq + (size_t)o * 4, dq);
#pragma unroll
    for (int i = 0; i < 4; ++i) qn[i] = qo[i] + dq[i];
    nc_normalflux(g, qo, nxy.x, nxy.y, ds, f1);
    nc_normalflux(g, qn, nxy.x, nxy.y, ds, f2);
    const double lam = d.fla[e];
#pragma unroll
    for (int i = 0; i < 4; ++i) cres[i] = cres[i] + 0.5 * (sign * (f2[i] - f1[i]) - 1.5 * lam * dq[i]);
  };
  // the first MCU entries' slot data are loaded together, ahead of the
  // gathers they select; the sum runs over the entries in order
  {
    int2 eo[MCU];
    double2 nxy[MCU];
    double dsv[MCU];
#pragma unroll
    for (int k = 0; k < MCU; ++k)
      if (k < cnt) {
        const size_t q = base + (size_t)k * np;
        eo[k] = d.mc_eo[q];
        nxy[k] = d.mc_nxy[q];
        dsv[k] = d.mc_ds[q];
      }
#pragma unroll
    for (int k = 0; k < MCU; ++k)
      if (k < cnt) term(eo[k], nxy[k], dsv[k]);
  }
  for (int k = MCU; k < cnt; ++k) {
    const size_t q = base + (size_t)k * np;
    term(d.mc_eo[q], d.mc_nxy[q], d.mc_ds[q]);
  }
  double res[4], dq[4], dd;
  if (dir == 0) {
    // the vertex's own residual and diagonal, in position order (the
    // residual pass wrote them there, as save_old the old state)
    dd = d.mc_dd[s];
    ld4d(d.mc_res + (size_t)s * 4, res);
#pragma unroll
    for (int i = 0; i < 4; ++i) dq[i] = (-res[i] - cres[i]) / dd;
    if (fin) {
      double qo[4], q[4];
      ld4d(d.mc_qold + (size_t)s * 4, qo);
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = qo[i] + dq[i];
      st4d(d.qc + (size_t)it * 4, q);
    }
  } else {
    dd = d.mc_dd[s];
    ld4d(d.mc_dq + (size_t)s * 4, dq);
#pragma unroll
    for (int i = 0; i < 4; ++i) dq[i] = dq[i] - cres[i] / dd;
    double qo[4], q[4];
    ld4d(d.mc_qold + (size_t)s * 4, qo);
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = qo[i] + dq[i];
    st4d(d.qc + (size_t)it * 4, q);
  }
  st4d(d.mc_dq + (size_t)s * 4, dq);
}

// ---- level-ordered one-workgroup sweeps
// A sweep level touches a few hundred vertices scattered over the mesh, and
// in one workgroup everything it reads enters through one CU, whose address
// unit handles about one distinct cache line per lane-access of a scattered
// load.  So the data that does not depend on the sweep is first written in
// level order by a parallel pass, position-major, and inside the sweep only
// the swept neighbours' dq (one 32-byte record per position) is gathered;
// their state is rebuilt as qold_o + dq_o, the same sum that produced it.
constexpr int SWE = 6;

// Per position: the vertex's residual, old state and diagonal; per swept
// entry: the edge's spectral radius, normalflux of the neighbour's old state
// (NCFVM/lusgs.f90:62-72) and that old state.  The geometry of the entries
// (nx, ny, ds) is written once at set-up.
__global__ __launch_bounds__(BLK) void k_nc_sweep_prep(DevNC d, NCFree g) {
  const int np = 2 * d.nop;
  int s = blockIdx.x * BLK + threadIdx.x;
  if (s >= np) return;
  const int v = d.sw_vert[s];
  const double2* r = reinterpret_cast<const double2*>(d.res + (size_t)v * 4);
  const double2* qo = reinterpret_cast<const double2*>(d.qold + (size_t)v * 4);
  d.sw_own[s] = r[0];
  d.sw_own[(size_t)np + s] = r[1];
  d.sw_own[2 * (size_t)np + s] = qo[0];
  d.sw_own[3 * (size_t)np + s] = qo[1];
  d.sw_dd[s] = d.dd[v];
  const int cnt = d.sw_cnt[s];
  for (int k = 0; k < cnt; ++k) {
    const int2 pe = d.sw_pe[(size_t)k * np + s];
    double2* en = d.sw_ent + (size_t)k * SWE * np + s;
    const double2 nxy = en[0];
    double2 dsl = en[np];
    double q[4], f1[4];
    ld4d(d.qold + (size_t)pe.y * 4, q);
    nc_normalflux(g, q, nxy.x, nxy.y, dsl.x, f1);
    dsl.y = d.fla[pe.x];
    en[np] = dsl;
    en[2 * (size_t)np] = make_double2(f1[0], f1[1]);
    en[3 * (size_t)np] = make_double2(f1[2], f1[3]);
    en[4 * (size_t)np] = make_double2(q[0], q[1]);
    en[5 * (size_t)np] = make_double2(q[2], q[3]);
  }
}

// One sweep in one workgroup over the level-ordered data.  The loads of the
// first SWP entry slots of a position are issued together, each under its
// slot's own k < cnt test.
constexpr int SWP = 5;
constexpr int STREAM_THREADS = 512;

struct SwSlot {
  double2 nxy, dsl, f01, f23, q01, q23;
  int nb;
  double dq[4];
};

// The dq of the last SW_RING positions of a sweep is also kept in LDS (a
// ring indexed by position mod SW_RING, 160 KB, all of a CU's LDS): a
// neighbour at position o is read from there when o >= lim = level start +
// level width - SW_RING, which no write of this level or an earlier one has
// overwritten; the rest (about 3 % on the 2M-vertex O-grid) from HBM.
constexpr int SW_RING = 5120;

__device__ __forceinline__ void nc_stream_load(const DevNC& d, int np, int s, int k, SwSlot& e,
                                               const double2* ring, int lim) {
  e.nb = d.sw_nb[(size_t)k * np + s];
  const double2* en = d.sw_ent + (size_t)k * SWE * np + s;
  e.nxy = en[0];
  e.dsl = en[np];
  e.f01 = en[2 * (size_t)np];
  e.f23 = en[3 * (size_t)np];
  e.q01 = en[4 * (size_t)np];
  e.q23 = en[5 * (size_t)np];
  const int o = e.nb >= 0 ? e.nb : ~e.nb;
  if (o >= lim) {
    const double2 a = ring[2 * (o % SW_RING)], b = ring[2 * (o % SW_RING) + 1];
    e.dq[0] = a.x;
    e.dq[1] = a.y;
    e.dq[2] = b.x;
    e.dq[3] = b.y;
  } else {
    ld4d(d.sw_dq + (size_t)o * 4, e.dq);
  }
}

__device__ __forceinline__ void nc_stream_term(const NCFree& g, const SwSlot& e, double* cres) {
  const double q[4] = {e.q01.x + e.dq[0], e.q01.y + e.dq[1], e.q23.x + e.dq[2], e.q23.y + e.dq[3]};
  const double f1[4] = {e.f01.x, e.f01.y, e.f23.x, e.f23.y};
  double f2[4];
  nc_normalflux(g, q, e.nxy.x, e.nxy.y, e.dsl.x, f2);
  double sign = e.nb >= 0 ? 1.0 : -1.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    cres[i] = cres[i] + 0.5 * (sign * (f2[i] - f1[i]) - 1.5 * e.dsl.y * e.dq[i]);
}

__device__ __forceinline__ void nc_stream_pos(const DevNC& d, const NCFree& g, int s, int dir,
                                              double2* ring, int lim) {
  const int np = 2 * d.nop;
  const int cnt = d.sw_cnt[s];
  const double2 r01 = d.sw_own[s], r23 = d.sw_own[(size_t)np + s];
  const double dd = d.sw_dd[s];
  double dq0[4];
  if (dir == 1) ld4d(d.sw_dq + (size_t)d.sw_d0[s] * 4, dq0);
  SwSlot e[SWP];
#pragma unroll
  for (int k = 0; k < SWP; ++k)
    if (k < cnt) nc_stream_load(d, np, s, k, e[k], ring, lim);
  double cres[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < SWP; ++k)
    if (k < cnt) nc_stream_term(g, e[k], cres);
  for (int k = SWP; k < cnt; ++k) {
    SwSlot e2;
    nc_stream_load(d, np, s, k, e2, ring, lim);
    nc_stream_term(g, e2, cres);
  }
  const double res[4] = {r01.x, r01.y, r23.x, r23.y};
  double out[4];
  if (dir == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = (-res[i] - cres[i]) / dd;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = dq0[i] - cres[i] / dd;
  }
  st4d(d.sw_dq + (size_t)s * 4, out);
  ring[2 * (s % SW_RING)] = make_double2(out[0], out[1]);
  ring[2 * (s % SW_RING) + 1] = make_double2(out[2], out[3]);
}

// A level's writes go to the slots of positions [b - SW_RING, b + n -
// SW_RING), below lim, so no read of the level meets them; a level wider
// than the ring reads only from HBM (lim > b).
__global__ __launch_bounds__(STREAM_THREADS) void k_nc_lusgs_stream(DevNC d, NCFree g,
                                                                  const int* off, int nlev,
                                                                  int dir) {
  __shared__ double2 ring[2 * SW_RING];
  for (int l = 0; l < nlev; ++l) {
    const int b = off[l], n = off[l + 1] - b;
    const int lim = b + n - SW_RING;
    for (int t = threadIdx.x; t < n; t += STREAM_THREADS)
      nc_stream_pos(d, g, b + t, dir, ring, lim);
    __syncthreads();
  }
}

// The sweeps' result: qc = qold + dq of the backward sweep, per backward
// position (the same sum the sweep used for the neighbours' state).
__global__ __launch_bounds__(BLK) void k_nc_sweep_final(DevNC d) {
  const int np = 2 * d.nop;
  int t = cell_index();
  if (t >= d.nop) return;
  const int s = d.nop + t;
  const int v = d.sw_vert[s];
  const double2 q01 = d.sw_own[2 * (size_t)np + s], q23 = d.sw_own[3 * (size_t)np + s];
  double dq[4];
  ld4d(d.sw_dq + (size_t)s * 4, dq);
  const double q[4] = {q01.x + dq[0], q01.y + dq[1], q23.x + dq[2], q23.y + dq[3]};
  st4d(d.qc + (size_t)v * 4, q);
}

// A whole sweep in one cooperative launch of a few workgroups: the vertices
// of a level are spread over all of them (so a level's scattered records
// come in through several CUs), and a grid barrier (HIP cooperative groups:
// the launch fails unless every workgroup is resident, and the barrier
// makes each level's stores visible device-wide) separates the levels.
// Every loop is bounded by the level lists.
constexpr int GRID_SWEEP_THREADS = 256;

__global__ __launch_bounds__(GRID_SWEEP_THREADS) void k_nc_lusgs_sweep_grid(DevNC d, NCFree g,
                                                                          const int* verts,
                                                                          const int* off,
                                                                          int nlev, int dir) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int nthr = gridDim.x * GRID_SWEEP_THREADS;
  const int t0 = blockIdx.x * GRID_SWEEP_THREADS + threadIdx.x;
  for (int l = 0; l < nlev; ++l) {
    const int b = off[l], n = off[l + 1] - b;
    for (int t = t0; t < n; t += nthr) nc_lusgs_vertex(d, g, verts[b + t], dir);
    grid.sync();
  }
}

// A whole sweep in one workgroup: the levels in order, a workgroup barrier
// between consecutive levels (which also orders the global-memory updates
// among the workgroup's waves).  Every loop is bounded by the level lists.
// Each level slot is packed: its vertex, the count of neighbours already
// swept and up to SWQ (edge, neighbour) pairs in edge order, so a level costs
// one dependent load; the next level's slots are fetched before the barrier.
constexpr int SWEEP_THREADS = 512;
constexpr int SWQ = 8;
struct SweepSlot {
  int vert, cnt;
  int2 eo[SWQ];
};

__device__ __forceinline__ void nc_sweep_slot(const DevNC& d, const NCFree& g, const SweepSlot& sl,
                                              int dir) {
  const int it = sl.vert;
  double cres[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < SWQ; ++k) {
    if (k < sl.cnt) {
      int2 eo = sl.eo[k];
      bool first = eo.x >= 0;
      int e = first ? eo.x : ~eo.x, o = eo.y;
      double sign = first ? 1.0 : -1.0;
      double f1[4], f2[4], qo[4], qn[4], dq[4];
      const double nx = d.erec[(size_t)e * 3], ny = d.erec[(size_t)e * 3 + 1];
      const double ds = d.erec[(size_t)e * 3 + 2];
      ld4d(d.qold + (size_t)o * 4, qo);
      ld4d(d.qc + (size_t)o * 4, qn);
      ld4d(d.dqc + (size_t)o * 4, dq);
      nc_normalflux(g, qo, nx, ny, ds, f1);
      nc_normalflux(g, qn, nx, ny, ds, f2);
      double lam = d.fla[e];
#pragma unroll
      for (int i = 0; i < 4; ++i) cres[i] = cres[i] + 0.5 * (sign * (f2[i] - f1[i]) - 1.5 * lam * dq[i]);
    }
  }
  double dd = d.dd[it], res[4], dq[4], qo[4], q[4];
  ld4d(d.res + (size_t)it * 4, res);
  ld4d(d.qold + (size_t)it * 4, qo);
  if (dir == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) dq[i] = (-res[i] - cres[i]) / dd;
  } else {
    ld4d(d.dqc + (size_t)it * 4, dq);
#pragma unroll
    for (int i = 0; i < 4; ++i) dq[i] = dq[i] - cres[i] / dd;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] = qo[i] + dq[i];
  st4d(d.dqc + (size_t)it * 4, dq);
  st4d(d.qc + (size_t)it * 4, q);
}

__global__ __launch_bounds__(SWEEP_THREADS) void k_nc_lusgs_sweep(DevNC d, NCFree g,
                                                                  const SweepSlot* slots,
                                                                  const int* off, int nlev,
                                                                  int dir) {
  // levels wider than the workgroup take several passes; the first pass of
  // the next level is prefetched
  SweepSlot cur;
  int b = off[0], n = off[1] - b;
  if ((int)threadIdx.x < n) cur = slots[b + threadIdx.x];
  for (int l = 0; l < nlev; ++l) {
    for (int t = threadIdx.x + SWEEP_THREADS; t < n; t += SWEEP_THREADS) {
      SweepSlot sl = slots[b + t];
      nc_sweep_slot(d, g, sl, dir);
    }
    if ((int)threadIdx.x < n) nc_sweep_slot(d, g, cur, dir);
    if (l + 1 < nlev) {
      b = off[l + 1];
      n = off[l + 2] - b;
      if ((int)threadIdx.x < n) cur = slots[b + threadIdx.x];
    }
    __syncthreads();
  }
}

// Wall forces of the vertex-centred path.  The reference has no force
// routine for it; here the force is the wall closure of solid_flux
// (NCFVM/solid_flux.f90:20-40) summed over the wall edges: per wall edge the
// end-point pressures weighted 3/4-1/4 on the half-edge right normal, i.e.
// the pressure terms the wall adds to the momentum residuals of its two end
// points.  Pressures come from the current conserved state (the expressions
// of save_old).
//   k_nc_wall_terms: per wall edge (edge order) its force contribution and
//                    -Cp at its first end point;
//   k_nc_forces_sum: one wave adds the contributions in edge order.
__device__ __forceinline__ double nc_pressure(const NCFree& g, const double* q) {
  double rho = q[1], u = q[2] / q[1], v = q[3] / q[1], et = q[0];
  double qq = u * u + v * v;
  return (g.gamma - 1.0) * (et - 0.5 * rho * qq);
}

__global__ __launch_bounds__(BLK) void k_nc_wall_terms(DevNC d, NCFree g, const int* wedge,
                                                       int nw, double dyna, double* fe,
                                                       double* cp) {
  int k = blockIdx.x * BLK + threadIdx.x;
  if (k >= nw) return;
  int2 pp = d.ep[wedge[k]];
  double q1[4], q2[4], x1, y1, x2, y2;
  ld4d(d.qc + (size_t)pp.x * 4, q1);
  ld4d(d.qc + (size_t)pp.y * 4, q2);
  const double2 c1 = d.pxy[pp.x], c2 = d.pxy[pp.y];
  x1 = c1.x; y1 = c1.y;
  x2 = c2.x; y2 = c2.y;
  double pr1 = nc_pressure(g, q1), pr2 = nc_pressure(g, q2);
  double p1w = 0.75 * pr1 + 0.25 * pr2, p2w = 0.25 * pr1 + 0.75 * pr2;
  double nx = 0.5 * (y2 - y1), ny = -0.5 * (x2 - x1);
  fe[2 * (size_t)k] = p1w * nx + p2w * nx;
  fe[2 * (size_t)k + 1] = p1w * ny + p2w * ny;
  cp[k] = -((pr1 - g.p_inf) / dyna);
}

__global__ __launch_bounds__(64) void k_nc_forces_sum(const double* fe, int nw, double* out) {
  double sx = 0.0, sy = 0.0;
  for (int b = 0; b < nw; b += 64) {
    int k = b + (int)threadIdx.x;
    double ex = 0.0, ey = 0.0;
    if (k < nw) {
      ex = fe[2 * (size_t)k];
      ey = fe[2 * (size_t)k + 1];
    }
    const int m = min(64, nw - b);
    for (int j = 0; j < m; ++j) {
      sx = sx + __shfl(ex, j);
      sy = sy + __shfl(ey, j);
    }
  }
  if (threadIdx.x == 0) {
    out[0] = sx;
    out[1] = sy;
  }
}

}  // namespace ufx
