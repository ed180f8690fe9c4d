// Kernels of the cell-centred residual hot path (gfx950).  Included once, by
// cc_solver.hip.  Each kernel cites the reference routine it performs; every
// per-cell sum runs over the cell's face slots in the reference's
// accumulation order.
#pragma once

#include "cc_device.h"
#include "cc_physics.h"

namespace ufx {

// The stage update of CCFVM/main.F90:92-101, q = a qold + b (q - dt/A R),
// rounded after every operation (nothing contracted into fused
// multiply-adds), so that the fused residual stage and the separate update
// kernel give the same bits however the compiler inlines them.
__device__ __forceinline__ double rk_update1(double ak, double o, double bk, double q, double dta,
                                            double r) {
#pragma clang fp contract(off)
  return ak * o + bk * (q - dta * r);
}

template <typename R>
__device__ __forceinline__ void face_rec(const DevT<R>& d, int f, R* fr) {
  ldv<R, Layout<R>::FW>(d.frec + (size_t)f * Layout<R>::FW, fr);
}

// Face average of the vertex primitives (CCFVM/average.F90:32-36), computed
// where it is used instead of being stored.
template <typename R, int NV>
__device__ __forceinline__ void face_prims(const DevT<R>& d, int f, R* fq) {
  constexpr int PW = Layout<R>::PW;
  int2 pp = d.fpt[f];
  R a[PW], b[PW];
  ldv<R, PW>(d.prec + (size_t)pp.x * PW, a);
  ldv<R, PW>(d.prec + (size_t)pp.y * PW, b);
#pragma unroll
  for (int k = 0; k < NV; ++k) fq[k] = R(0.5) * (a[k] + b[k]);
}

// Spectral radius of face f, CCFVM/timestep.F90:147-171 (spectral02):
// inviscid |u.n| + a from the face-averaged primitives fq, times the face
// length, plus the viscous term (mu: the face viscosity when vll != 0).
template <typename R>
__device__ __forceinline__ double face_lambda_val(const Gas<double>& g, const R* fq, const R* fr,
                                                  double vll, double mu) {
#pragma clang fp contract(off)
  double r0 = fq[0], u0 = fq[1], v0 = fq[2], p0 = fq[3];
  double a0 = sqrt(g.gamma * p0 / r0);
  double un = u0 * (double)fr[F_NX] + v0 * (double)fr[F_NY];
  double ll = (double)fr[F_AREA] * (fabs(un) + a0);
  if (vll != 0.0) ll = ll + vll * mu / r0;
  return ll;
}

// The spectral radius of every face, stored (for the time step of a routine
// call, LU-SGS and the residual smoothing).
template <typename R>
__device__ __forceinline__ void face_lambda_body(const DevT<R>& d, VBlock vb, Gas<double> g,
                                                 double vll) {
  int f = vb.cell();
  if (f >= d.nof || d.err[0]) return;
  // the normal and length from their compact copy (F_NX, F_NY, F_AREA are
  // its first three values)
  static_assert(F_NX == 0 && F_NY == 1 && F_AREA == 2, "compact face data layout");
  R fq[4], fr[4];
  face_prims<R, 4>(d, f, fq);
  ldv<R, 4>(d.fnl + (size_t)f * 4, fr);
  d.fla[f] = face_lambda_val<R>(g, fq, fr, vll, vll != 0.0 ? d.fmu[f] : 0.0);
}
template <typename R>
__global__ __launch_bounds__(BLK) void k_face_lambda(DevT<R> d, Gas<double> g, double vll) {
  face_lambda_body<R>(d, hw_block(), g, vll);
}

// ------------------------------------------------- primitive refresh + range
// CCFVM/main.F90:62-67 (con2prim of interior cells), qold = qc
// (CCFVM/solution.F90:118-132), positivity (CCFVM/main.F90:384-411) and the
// global range of the four primitives over all cells including ghosts
// (CCFVM/limiter.F90:146-149), accumulated atomically.
template <typename R>
__device__ __forceinline__ void prim_body(const DevT<R>& d, VBlock vb, Gas<double> g, int do_prim,
                                          int do_old, int do_range) {
  constexpr int RW = Layout<R>::RW;
  __shared__ double sh[4][BLK];
  int c = vb.cell();
  double mx[4], mn[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { mx[k] = -INFINITY; mn[k] = INFINITY; }
  if (c < d.ncell && !d.err[0]) {
    R q4[4];
    if (c < d.noc && (do_prim || do_old)) {
      double q[4];
      ld4d(d.qc + (size_t)c * 4, q);
      if (do_old == 1) st4d(d.qold + (size_t)c * 4, q);
      if (do_prim) {
        double p[5];
        con2prim(g, q, p);
#pragma unroll
        for (int k = 0; k < 4; ++k) q4[k] = (R)p[k];
        stv<R, 4>(d.rec + (size_t)c * RW + R_QP, q4);
        stv<R, 4>(d.qpc + (size_t)c * 4, q4);
        d.qT[c] = (R)p[4];
        d.qTc[c] = (R)p[4];
        // check_positivity (CCFVM/main.F90:384-411) runs after the
        // iteration's update on these start-of-iteration values; the
        // failing cell is noted here and becomes the error at the
        // iteration's end (norm_final_body), after any error of the stages
        if (do_old && (p[0] <= 0.0 || p[3] <= 0.0)) atomicMin(&d.err[2], d.cref[c] + 1);
      } else {
        ldv<R, 4>(d.rec + (size_t)c * RW + R_QP, q4);
      }
    } else {
      ldv<R, 4>(d.rec + (size_t)c * RW + R_QP, q4);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) mx[k] = mn[k] = (double)q4[k];
  }
  if (!do_range) return;
  // wave-level max/min by lane shuffles, then across the block's waves
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double a = __shfl_xor(mx[k], off), b = __shfl_xor(mn[k], off);
      mx[k] = a > mx[k] ? a : mx[k];
      mn[k] = b < mn[k] ? b : mn[k];
    }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int k = 0; k < 4; ++k) {
      sh[k][wave] = mx[k];
      sh[k][BLK / 64 + wave] = mn[k];
    }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) {
      double a = sh[k][0], b = sh[k][BLK / 64];
      for (int w = 1; w < BLK / 64; ++w) {
        a = sh[k][w] > a ? sh[k][w] : a;
        b = sh[k][BLK / 64 + w] < b ? sh[k][BLK / 64 + w] : b;
      }
      d.rpart[(size_t)vb.b * 8 + k] = a;
      d.rpart[(size_t)vb.b * 8 + 4 + k] = b;
    }
}
template <typename R>
__global__ __launch_bounds__(BLK) void k_prim(DevT<R> d, Gas<double> g, int do_prim, int do_old,
                                              int do_range) {
  prim_body<R>(d, hw_block(), g, do_prim, do_old, do_range);
}

// Final reduction of k_prim's block partials into the global range.
template <typename R>
__device__ __forceinline__ void range_final_body(const DevT<R>& d, VBlock vb, int nblk) {
  __shared__ double sh[4][BLK];
  double mx[4], mn[4];
  for (int k = 0; k < 4; ++k) { mx[k] = -INFINITY; mn[k] = INFINITY; }
  // 64 bytes per block partial, four of them in flight per thread (max and
  // min do not depend on the order)
  const double2* p2 = reinterpret_cast<const double2*>(d.rpart);
#pragma unroll 4
  for (int b = threadIdx.x; b < nblk; b += BLK) {
    const double2 v0 = p2[(size_t)b * 4], v1 = p2[(size_t)b * 4 + 1];
    const double2 v2 = p2[(size_t)b * 4 + 2], v3 = p2[(size_t)b * 4 + 3];
    const double a[4] = {v0.x, v0.y, v1.x, v1.y}, m[4] = {v2.x, v2.y, v3.x, v3.y};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      mx[k] = a[k] > mx[k] ? a[k] : mx[k];
      mn[k] = m[k] < mn[k] ? m[k] : mn[k];
    }
  }
  block_reduce<4>(sh, mx, OpMax());
  block_reduce<4>(sh, mn, OpMin());
  if (threadIdx.x < 4) {
    d.range[threadIdx.x] = ord_of(mx[threadIdx.x]);
    d.range[4 + threadIdx.x] = ord_of(mn[threadIdx.x]);
  }
}
template <typename R>
__global__ __launch_bounds__(BLK) void k_range_final(DevT<R> d, int nblk) {
  range_final_body<R>(d, hw_block(), nblk);
}

// ------------------------------------------------------- vertex averages
// CCFVM/average.F90:11-29: pseudo-Laplacian weighted average over the cells
// around a vertex (ghosts included), accumulated in double.  The cells
// around a block of vertices (numbered by first touch in cell order) lie
// mostly in a window of AVG_WIN cells starting at the block's lowest cell
// (avg_w0); their primitives are staged in LDS, the rest read from HBM.
// Interior primitives come from the compact copy (32 B per cell in FP64,
// where the same field inside the 144 B record would fetch whole lines);
// ghost primitives from the records.
constexpr int AVG_WIN = 768;
template <typename R, bool VISC>
__device__ __forceinline__ void avg_c2v_body(const DevT<R>& d, VBlock vb) {
  constexpr int RW = Layout<R>::RW, PW = Layout<R>::PW;
  __shared__ R s_q[AVG_WIN * 4];
  __shared__ R s_t[AVG_WIN];
  __shared__ R s_mu[VISC ? AVG_WIN : 1];
  const int pb = xcd_block(vb.b, vb.n);
  const int p = pb * BLK + threadIdx.x;
  const int w0 = d.avg_w0[pb], wn = min(AVG_WIN, d.noc - w0);
  for (int i = threadIdx.x; i < wn; i += BLK) {
    R q[4];
    ldv<R, 4>(d.qpc + (size_t)(w0 + i) * 4, q);
    stv<R, 4>(s_q + i * 4, q);
    s_t[i] = d.qTc[w0 + i];
    if (VISC) s_mu[i] = d.cmul[w0 + i];
  }
  __syncthreads();
  if (p >= d.nop || d.err[0]) return;
  double w = 0.0, con[6] = {0, 0, 0, 0, 0, 0};
  int j0 = d.v2c_ptr[p], j1 = d.v2c_ptr[p + 1];
  // the cell list first (independent loads), then the gathers, so the
  // latency of the list is paid once; a tail loop takes vertices with more
  // than AVG_UNR cells
  constexpr int AVG_UNR = 8;
  int cs[AVG_UNR];
  double ws[AVG_UNR];
  const int n = j1 - j0;
#pragma unroll
  for (int k = 0; k < AVG_UNR; ++k)
    if (k < n) {
      cs[k] = ld_nt(d.v2c + j0 + k);
      ws[k] = ld_nt(d.wt + j0 + k);
    }
  auto add = [&](int c, double cw) {
    R q[4], t, mu = R(0);
    unsigned l = (unsigned)(c - w0);
    if (l < (unsigned)wn) {
      ldv<R, 4>(s_q + l * 4, q);
      t = s_t[l];
      if (VISC) mu = s_mu[l];
    } else {
      if (c < d.noc) ldv<R, 4>(d.qpc + (size_t)c * 4, q);
      else ldv<R, 4>(d.rec + (size_t)c * RW + R_QP, q);
      t = c < d.noc ? d.qTc[c] : d.qT[c];
      if (VISC) mu = d.cmul[c];
    }
    w = w + cw;
#pragma unroll
    for (int k = 0; k < 4; ++k) con[k] = con[k] + (double)q[k] * cw;
    con[4] = con[4] + (double)t * cw;
    if (VISC) con[5] = con[5] + (double)mu * cw;
  };
#pragma unroll
  for (int k = 0; k < AVG_UNR; ++k)
    if (k < n) add(cs[k], ws[k]);
  for (int j = j0 + AVG_UNR; j < j1; ++j) add(ld_nt(d.v2c + j), ld_nt(d.wt + j));
  R o[PW];
#pragma unroll
  for (int k = 0; k < 5; ++k) o[k] = (R)(con[k] / w);
  o[P_MU] = (R)(con[5] / w);
#pragma unroll
  for (int k = 6; k < PW; ++k) o[k] = R(0);
  stv<R, PW>(d.prec + (size_t)p * PW, o);
}
template <typename R, bool VISC>
__global__ __launch_bounds__(BLK) void k_avg_c2v(DevT<R> d) {
  avg_c2v_body<R, VISC>(d, hw_block());
}

// The global range of k_range_final with wave shuffles and 256 bytes of
// LDS (max and min do not depend on the order), for a block riding along
// with another kernel's launch.
template <typename R>
__device__ __forceinline__ void range_final_small(const DevT<R>& d, int nblk) {
  __shared__ double shw[8][BLK / 64];
  double mx[4], mn[4];
  for (int k = 0; k < 4; ++k) { mx[k] = -INFINITY; mn[k] = INFINITY; }
  const double2* p2 = reinterpret_cast<const double2*>(d.rpart);
#pragma unroll 1
  for (int b = threadIdx.x; b < nblk; b += BLK) {
    const double2 v0 = p2[(size_t)b * 4], v1 = p2[(size_t)b * 4 + 1];
    const double2 v2 = p2[(size_t)b * 4 + 2], v3 = p2[(size_t)b * 4 + 3];
    const double a[4] = {v0.x, v0.y, v1.x, v1.y}, m[4] = {v2.x, v2.y, v3.x, v3.y};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      mx[k] = a[k] > mx[k] ? a[k] : mx[k];
      mn[k] = m[k] < mn[k] ? m[k] : mn[k];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double a = __shfl_xor(mx[k], off), b = __shfl_xor(mn[k], off);
      mx[k] = a > mx[k] ? a : mx[k];
      mn[k] = b < mn[k] ? b : mn[k];
    }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int k = 0; k < 4; ++k) {
      shw[k][wave] = mx[k];
      shw[4 + k][wave] = mn[k];
    }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    double a = shw[k][0], b = shw[4 + k][0];
    for (int w = 1; w < BLK / 64; ++w) {
      a = shw[k][w] > a ? shw[k][w] : a;
      b = shw[4 + k][w] < b ? shw[4 + k][w] : b;
    }
    d.range[k] = ord_of(a);
    d.range[4 + k] = ord_of(b);
  }
}

// Vertex averages with the limiter's range reduction as one extra block:
// neither needs the other, and the range no longer costs a one-block
// launch of its own.  The range block is block 0, so it starts first and
// its serial loop overlaps the averages.
template <typename R, bool VISC>
__global__ __launch_bounds__(BLK) void k_avg_c2v_range(DevT<R> d, int nblk_all) {
  if (blockIdx.x == 0) {
    range_final_small<R>(d, nblk_all);
    return;
  }
  avg_c2v_body<R, VISC>(d, VBlock{(int)blockIdx.x - 1, (int)gridDim.x - 1});
}

template <typename R>
__device__ __forceinline__ R vk_limit(R a, R b, R tol) {
  R nr = a * a + R(2) * a * b + tol * tol;
  R dr = a * a + R(2) * b * b + a * b + tol * tol;
  return nr / dr;
}

enum { G_NONE = 0, G_LSQR = 1, G_GG = 2, G_GGFC = 3 };

// ------------------------------------------------------ fused cell kernel
// Per cell, over its faces in ascending reference order:
//   gradient  LSQR: CCFVM/gradient.F90:2-93 (unit weights; the interior and
//             boundary loops round alfa2 differently, kept);
//             GG:   :95-156 (boundary faces with -n, /area);
//             GGFC: :159-323 cell part, sum of face gradients x co-area / cov;
//   limiter   Venkatakrishnan with the ratchet, CCFVM/limiter.F90:123-296;
//   time step spectral02 + time_step02, CCFVM/timestep.F90:98-177.
// wdt = 0 (whole RK3 iterations, whose stages read only dt/A): the
// spectral radius and dt are not stored; the same pass with GRAD = G_NONE,
// TSTEP forms them from the face radii when something reads them.
template <typename R, int GRAD, bool WITH_T, bool LIMIT, bool TSTEP>
__device__ __forceinline__ void cell_body(const DevT<R>& d, VBlock vb, Gas<double> g, double vll,
                                          double cfl, int wdt = 1) {
  constexpr int RW = Layout<R>::RW, FW = Layout<R>::FW;
  // the primitives, centroids (and temperatures) of the block's cells are
  // staged in LDS: most neighbours lie in the same block
  constexpr bool SL = GRAD == G_LSQR || LIMIT;
  __shared__ R s_qp[SL ? BLK * 4 : 1];
  __shared__ R s_cen[SL ? BLK * 2 : 1];
  __shared__ R s_t[SL && WITH_T ? BLK : 1];
  const int base = xcd_block(vb.b, vb.n) * BLK;
  const int c = base + threadIdx.x;
  const int cnt = min(BLK, d.noc - base);
  R own[RW];
  R tc = R(0);
  if (c < d.noc) {
    // the fields this kernel reads, from compact arrays (whole lines of
    // them, where the records would give a few values per line): the
    // primitives (which the last stage of a whole iteration refreshes), the
    // limiter values, the centroid; the old gradient only for a
    // limiter-only call
    ldv<R, 4>(d.qpc + (size_t)c * 4, own + R_QP);
    ldv<R, 4>(d.phc + (size_t)c * 4, own + R_PHI);
    ldv<R, 2>(d.cenc + (size_t)c * 2, own + R_CEN);
    if (GRAD == G_NONE) ldv<R, 8>(d.rec + (size_t)c * RW + R_G, own + R_G);
    tc = d.qTc[c];
  }
  if (SL) {
    if (c < d.noc) {
      stv<R, 4>(s_qp + threadIdx.x * 4, own + R_QP);
      stv<R, 2>(s_cen + threadIdx.x * 2, own + R_CEN);
      if (WITH_T) s_t[threadIdx.x] = tc;
    }
    __syncthreads();
  }
  if (c >= d.noc || d.err[0]) return;
  const int noc = d.noc;
  int info = ld_nt(d.cnf + c), nf = info & 0xff, nb = info >> 8;
  R cx = own[R_CEN], cy = own[R_CEN + 1];
  R gr[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) gr[k] = R(0);
  R umax[4], umin[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) umax[j] = umin[j] = own[R_QP + j];
  R fcx[MAXF], fcy[MAXF];
  double la = 0.0;
  R r11 = 0, r12 = 0, r22 = 0;
  if (GRAD == G_LSQR) { r11 = (R)d.r11[c]; r12 = (R)d.r12[c]; r22 = (R)d.r22[c]; }
  // slot lists first (independent loads), then the gathers
  int sf[MAXF], sn[MAXF];
#pragma unroll
  for (int s = 0; s < MAXF; ++s)
    if (s < nf) {
      sf[s] = ld_nt(d.sl_face + s * noc + c);
      sn[s] = ld_nt(d.sl_nbr + s * noc + c);
    }
#pragma unroll
  for (int s = 0; s < MAXF; ++s) {
    if (s >= nf) break;
    int f = sf[s];
    int n = sn[s];
    bool in_side = n >= 0;
    n = in_side ? n : ~n;
    bool bnd = n >= noc;
    R fr[FW];
    if (GRAD == G_GG) {
      face_rec(d, f, fr);
      fcx[s] = fr[F_CX];
      fcy[s] = fr[F_CY];
    } else {
      R fc[2];
      ldv<R, 2>(d.fcen + (size_t)f * 2, fc);
      fcx[s] = fc[0];
      fcy[s] = fc[1];
    }
    R fq[5];
    if (GRAD == G_GG || (LIMIT && bnd)) {
      if (GRAD == G_GG) face_prims<R, 5>(d, f, fq);
      else face_prims<R, 4>(d, f, fq);
    }
    // cell spectral radius: the face radii of k_face_lambda summed in
    // ascending face order, CCFVM/timestep.F90:171-173
    if (TSTEP) la = la + d.fla[f];
    R nq[4];
    const unsigned l = (unsigned)(n - base);
    const bool loc = SL && l < (unsigned)cnt;
    if (GRAD == G_LSQR || LIMIT) {
      if (loc) ldv<R, 4>(s_qp + l * 4, nq);
      else if (!bnd) ldv<R, 4>(d.qpc + (size_t)n * 4, nq);
      else ldv<R, 4>(d.rec + (size_t)n * RW + R_QP, nq);
    }
    if (GRAD == G_LSQR) {
      R ncen[2];
      if (loc) ldv<R, 2>(s_cen + l * 2, ncen);
      else ldv<R, 2>(d.cenc + (size_t)n * 2, ncen);
      R dx = ncen[0] - cx, dy = ncen[1] - cy;
      R alfa1 = dx / r11 / r11;
      R alfa2 = bnd ? (dy - dx * r12 / r11) / r22 / r22 : (dy - dx * r12 / r11) / (r22 * r22);
      R wx = alfa1 - alfa2 * r12 / r11, wy = alfa2;
      R nt = WITH_T ? (loc ? s_t[l] : (bnd ? d.qT[n] : d.qTc[n])) : R(0);
      if (in_side) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          R var = nq[k] - own[R_QP + k];
          gr[2 * k] = gr[2 * k] + var * wx;
          gr[2 * k + 1] = gr[2 * k + 1] + var * wy;
        }
        if (WITH_T) {
          R var = nt - tc;
          gr[8] = gr[8] + var * wx;
          gr[9] = gr[9] + var * wy;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          R var = own[R_QP + k] - nq[k];
          gr[2 * k] = gr[2 * k] - var * wx;
          gr[2 * k + 1] = gr[2 * k + 1] - var * wy;
        }
        if (WITH_T) {
          R var = tc - nt;
          gr[8] = gr[8] - var * wx;
          gr[9] = gr[9] - var * wy;
        }
      }
    } else if (GRAD == G_GG) {
      R dx = fr[F_NX] * fr[F_AREA], dy = fr[F_NY] * fr[F_AREA];
      if (bnd) { dx = -fr[F_NX] * fr[F_AREA]; dy = -fr[F_NY] * fr[F_AREA]; }
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        if (in_side) {
          gr[2 * k] = gr[2 * k] + fq[k] * dx;
          gr[2 * k + 1] = gr[2 * k + 1] + fq[k] * dy;
        } else {
          gr[2 * k] = gr[2 * k] - fq[k] * dx;
          gr[2 * k + 1] = gr[2 * k + 1] - fq[k] * dy;
        }
      }
    } else if (GRAD == G_GGFC) {
      R fg[10];
      ldv<R, 10>(d.fgrad + (size_t)f * 10, fg);
      R w = (R)d.fw[f];
#pragma unroll
      for (int k = 0; k < 10; ++k) gr[k] = gr[k] + fg[k] * w;
    }
    if (LIMIT) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        R v = bnd ? fq[j] : nq[j];
        umax[j] = umax[j] > v ? umax[j] : v;
        umin[j] = umin[j] < v ? umin[j] : v;
      }
    }
  }
  (void)nb;
  if (GRAD == G_GG) {
    R v = (R)d.cv[c];
#pragma unroll
    for (int k = 0; k < 10; ++k) gr[k] = gr[k] / v;
  } else if (GRAD == G_GGFC) {
    R v = (R)d.cov[c];
#pragma unroll
    for (int k = 0; k < 10; ++k) gr[k] = gr[k] / v;
  }
  if (GRAD != G_NONE) {
#pragma unroll
    for (int k = 0; k < 8; ++k) own[R_G + k] = gr[k];
    if (WITH_T) {
      R t2[2] = {gr[8], gr[9]};
      stv<R, 2>(d.gT + (size_t)c * 2, t2);
    }
  }
  if (LIMIT) {
    const R kappa = R(0.05);
    R tol[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) tol[j] = kappa * (R)(of_ord(d.range[j]) - of_ord(d.range[4 + j]));
#pragma unroll
    for (int s = 0; s < MAXF; ++s) {
      if (s >= nf) break;
      R dx = fcx[s] - cx, dy = fcy[s] - cy;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        R dl = own[R_G + 2 * j] * dx + own[R_G + 2 * j + 1] * dy;
        R p0 = R(1);
        if (dl > R(0)) p0 = vk_limit(umax[j] - own[R_QP + j], dl, tol[j]);
        else if (dl < R(0)) p0 = vk_limit(umin[j] - own[R_QP + j], dl, tol[j]);
        R ph = own[R_PHI + j];
        own[R_PHI + j] = ph < p0 ? ph : p0;
      }
    }
  }
  // the record's primitives are refreshed here, with the limiter values
  // and the gradient, for the stages' neighbour gathers
  static_assert(R_PHI == R_QP + 4 && R_G == R_PHI + 4, "record fields are adjacent");
  if (GRAD != G_NONE || LIMIT) {
    stv<R, 16>(d.rec + (size_t)c * RW + R_QP, own + R_QP);
    d.qT[c] = tc;
  }
  if (LIMIT) stv<R, 4>(d.phc + (size_t)c * 4, own + R_PHI);
  if (TSTEP) {
    double dt = cfl * d.cv[c] / la;
    if (wdt) {
      d.la[c] = la;
      d.dt[c] = dt;
    }
    d.dta[c] = dt / d.cv[c];
  }
}
template <typename R, int GRAD, bool WITH_T, bool LIMIT, bool TSTEP>
__global__ __launch_bounds__(BLK) void k_cell(DevT<R> d, Gas<double> g, double vll, double cfl,
                                              int wdt) {
  cell_body<R, GRAD, WITH_T, LIMIT, TSTEP>(d, hw_block(), g, vll, cfl, wdt);
}

// ------------------------------------------------ diamond-path face gradient
// gradtrixy over the triangles (p1, out, p2) and (p1, p2, in) and their
// area-weighted mean, CCFVM/gradient.F90:183-265,281-321.
template <typename R>
__device__ __forceinline__ bool gradtri(const double* x, const double* y, const R (*pr)[5],
                                        R* gr, double& ta) {
  double t = 0.0;
  for (int i = 0; i < 3; ++i) {
    int ip1 = i == 2 ? 0 : i + 1;
    t = t + (0.5 * (x[ip1] + x[i])) * (y[ip1] - y[i]);
  }
  ta = t;
  if (!(t > 0.0)) return false;
  for (int k = 0; k < 10; ++k) gr[k] = R(0);
  for (int i = 0; i < 3; ++i) {
    int ip1 = i == 2 ? 0 : i + 1;
    R dx = (R)(y[ip1] - y[i]);
    R dy = (R)(-(x[ip1] - x[i]));
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      R var = R(0.5) * (pr[i][k] + pr[ip1][k]);
      gr[2 * k] = gr[2 * k] + var * dx;
      gr[2 * k + 1] = gr[2 * k + 1] + var * dy;
    }
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) gr[k] = gr[k] / (R)t;
  return true;
}

template <typename R>
__global__ __launch_bounds__(BLK) void k_ggfc_face(DevT<R> d) {
  constexpr int RW = Layout<R>::RW, PW = Layout<R>::PW;
  int f = cell_index();
  if (f >= d.nof || d.err[0]) return;
  int2 pp = d.fpt[f];
  int ci = d.fin[f], co = d.fout[f];
  R a[PW], b[PW], qi[4], qo[4], ci2[2], co2[2];
  ldv<R, PW>(d.prec + (size_t)pp.x * PW, a);
  ldv<R, PW>(d.prec + (size_t)pp.y * PW, b);
  ldv<R, 4>(d.rec + (size_t)ci * RW + R_QP, qi);
  ldv<R, 4>(d.rec + (size_t)co * RW + R_QP, qo);
  ldv<R, 2>(d.rec + (size_t)ci * RW + R_CEN, ci2);
  ldv<R, 2>(d.rec + (size_t)co * RW + R_CEN, co2);
  R pr[3][5];
  double x[3], y[3];
  R g1[10], g2[10];
  double a1, a2;
  // triangle (p1, out, p2)
  x[0] = d.px[pp.x]; y[0] = d.py[pp.x];
  x[1] = (double)co2[0]; y[1] = (double)co2[1];
  x[2] = d.px[pp.y]; y[2] = d.py[pp.y];
  for (int k = 0; k < 5; ++k) { pr[0][k] = a[k]; pr[2][k] = b[k]; }
  for (int k = 0; k < 4; ++k) pr[1][k] = qo[k];
  pr[1][4] = d.qT[co];
  bool ok = gradtri(x, y, pr, g1, a1);
  // triangle (p1, p2, in)
  x[1] = d.px[pp.y]; y[1] = d.py[pp.y];
  x[2] = (double)ci2[0]; y[2] = (double)ci2[1];
  for (int k = 0; k < 5; ++k) { pr[0][k] = a[k]; pr[1][k] = b[k]; }
  for (int k = 0; k < 4; ++k) pr[2][k] = qi[k];
  pr[2][4] = d.qT[ci];
  ok = gradtri(x, y, pr, g2, a2) && ok;
  if (!ok) { set_err(d, DERR_AREA, d.fref[f]); return; }
  R o[10];
  R A1 = (R)a1, A2 = (R)a2;
#pragma unroll
  for (int k = 0; k < 10; ++k) o[k] = (A1 * g1[k] + A2 * g2[k]) / (A1 + A2);
  stv<R, 10>(d.fgrad + (size_t)f * 10, o);
  d.fw[f] = a1 + a2;
}

// ------------------------------------------------------------- residual
template <typename R, int FLUX>
__device__ __forceinline__ bool face_flux(const Gas<R>& g, R* ql, R* qr, R nx, R ny, R area, R* F) {
  if constexpr (FLUX == 1) return roe_flux(g, ql, qr, nx, ny, area, F);
  if constexpr (FLUX == 2) ausm_flux(g, ql, qr, nx, ny, F);
  if constexpr (FLUX == 3) vanleer_flux(g, ql, qr, nx, ny, area, F);
  if constexpr (FLUX == 4) rusanov_flux(g, ql, qr, nx, ny, area, F);
  return true;
}

// Where a stage update reads and writes its states.  src 0 (routine by
// routine): q = a qold + b (qc - dt/A R) in qc.  Whole RK3 iterations keep
// the iteration's start state in qc and the stage states in qs, so no copy
// to qold is needed: src 1 (stage 1) reads qc once and writes qs; src 2
// (stage 2) reads qc and qs and writes qs; src 3 (stage 3) reads qc and qs
// and writes qc (each cell reads only its own qc entry).
template <typename R>
__device__ __forceinline__ const double* stage_old(const DevT<R>& d, int src) {
  return src ? d.qc : d.qold;
}
template <typename R>
__device__ __forceinline__ const double* stage_in(const DevT<R>& d, int src) {
  return src >= 2 ? d.qs : d.qc;
}
template <typename R>
__device__ __forceinline__ double* stage_out(const DevT<R>& d, int src) {
  return src == 1 || src == 2 ? d.qs : d.qc;
}

// The boundary conditions of a residual evaluation.
struct BndOpts {
  Gas<double> g;
  int viscous_wall, iwall;
  double t_wall;
};

// Boundary face k of a residual evaluation, CCFVM/fvresidual.F90:27-130, in
// double: left state by the full-step reconstruction of the inner cell
// (record `own`), farfield state or wall ghost (built from the inner cell's
// current conserved state), flux F, ghost update and, for walls, the
// momentum flux of the force sums.  F = 0 for other boundary kinds.
template <typename R, int FLUX>
__device__ __forceinline__ void bface_one(const DevT<R>& d, const BndOpts& bo, int k,
                                          const R* own, int src, double* F) {
  constexpr int RW = Layout<R>::RW, FW = Layout<R>::FW;
  const Gas<double>& g = bo.g;
  int bc = d.fbc[k];
#pragma unroll
  for (int i = 0; i < 4; ++i) F[i] = 0.0;
  if (bc != FARF && bc != WALL) return;
  int f = d.bface[k], c = d.bcell[k], gh = d.noc + k;
  R fr[FW];
  face_rec(d, f, fr);
  double nx = fr[F_NX], ny = fr[F_NY], area = fr[F_AREA];
  double dx = (double)fr[F_CX] - (double)own[R_CEN], dy = (double)fr[F_CY] - (double)own[R_CEN + 1];
  double ql[4], qr[4], gq[4], gp[5];
  for (int i = 0; i < 4; ++i)
    ql[i] = (double)own[R_QP + i] +
            (double)own[R_PHI + i] * ((double)own[R_G + 2 * i] * dx + (double)own[R_G + 2 * i + 1] * dy);
  bool ok;
  if (bc == FARF) {
    farfield_state(g, ql, nx, ny, qr);
    ok = face_flux<double, FLUX>(g, ql, qr, nx, ny, area, F);
    prim2con(g, qr, gq);
    con2prim(g, gq, gp);
  } else {
    double q[4];
    ld4d(stage_in(d, src) + (size_t)c * 4, q);
    if (!bo.viscous_wall) {
      // slip_wall, CCFVM/boundCond.F90:46-81
      double rho = q[1], u = q[2] / q[1], v = q[3] / q[1];
      double un = u * nx + v * ny;
      gq[0] = q[0];
      gq[1] = q[1];
      gq[2] = rho * u - 2.0 * un * nx;
      gq[3] = rho * v - 2.0 * un * ny;
      con2prim(g, gq, gp);
    } else if (bo.iwall == 1) {
      // adiabatic no-slip, CCFVM/boundCond.F90:31-37
      gq[0] = q[0];
      gq[1] = q[1];
      gq[2] = -q[2];
      gq[3] = -q[3];
      con2prim(g, gq, gp);
    } else {
      // isothermal no-slip, CCFVM/boundCond.F90:19-30 (the ghost density
      // slot receives p/(rho znd), as in the reference)
      double pin[5];
      con2prim(g, q, pin);
      double tg = 2.0 * bo.t_wall - pin[4];
      if (tg <= 0.0) tg = 0.01 * bo.t_wall;
      gp[0] = pin[3] / (pin[0] * g.znd);
      gp[1] = -pin[1];
      gp[2] = -pin[2];
      gp[3] = pin[3];
      gp[4] = tg;
      prim2con(g, gp, gq);
    }
    for (int i = 0; i < 4; ++i) qr[i] = gp[i];
    ok = face_flux<double, FLUX>(g, ql, qr, nx, ny, area, F);
    d.bflux[k] = F[2];
    d.bflux[d.nbf + k] = F[3];
  }
  if (!ok) set_err(d, DERR_ROE, k);
  st4d(d.qc + (size_t)gh * 4, gq);
  R gp4[4] = {(R)gp[0], (R)gp[1], (R)gp[2], (R)gp[3]};
  stv<R, 4>(d.rec + (size_t)gh * RW + R_QP, gp4);
  d.qT[gh] = (R)gp[4];
}

// The boundary faces alone, one thread per face; the flux is stored for the
// stage kernel, which adds it first, as the reference does.
template <typename R, int FLUX>
__device__ __forceinline__ void bface_body(const DevT<R>& d, VBlock vb, Gas<double> g,
                                           int viscous_wall, int iwall, double t_wall, int src) {
  constexpr int RW = Layout<R>::RW;
  int k = vb.b * BLK + threadIdx.x;
  if (k >= d.nbf || d.err[0]) return;
  const BndOpts bo{g, viscous_wall, iwall, t_wall};
  R own[RW];
  ldv<R, RW>(d.rec + (size_t)d.bcell[k] * RW, own);
  double F[4];
  bface_one<R, FLUX>(d, bo, k, own, src, F);
  st4d(d.bres + (size_t)k * 4, F);
}
template <typename R, int FLUX>
__global__ __launch_bounds__(BLK) void k_bface(DevT<R> d, Gas<double> g, int viscous_wall,
                                               int iwall, double t_wall, int src) {
  bface_body<R, FLUX>(d, hw_block(), g, viscous_wall, iwall, t_wall, src);
}

// Copy the records of the block's cells [base, base + cnt) into LDS with
// coalesced 16-byte loads (records are 16-byte multiples).
template <typename R>
__device__ __forceinline__ void stage_block_recs(const DevT<R>& d, R* srec, int base, int cnt) {
  constexpr int RW = Layout<R>::RW;
  static_assert(RW * sizeof(R) % 16 == 0, "records are 16-byte multiples");
  const int4* src = reinterpret_cast<const int4*>(d.rec + (size_t)base * RW);
  int4* dst = reinterpret_cast<int4*>(srec);
  const int n16 = cnt * RW * (int)sizeof(R) / 16;
  for (int i = threadIdx.x; i < n16; i += BLK) dst[i] = src[i];
}

// Record of cell n: from the block's LDS copy when n is one of its cells
// (SL), else from HBM.
template <typename R, bool SL>
__device__ __forceinline__ void load_rec(const DevT<R>& d, const R* srec, int base, int n, R* out) {
  constexpr int RW = Layout<R>::RW;
  unsigned l = (unsigned)(n - base);
  if (SL && l < (unsigned)BLK) ldv<R, RW>(srec + l * RW, out);
  else ldv<R, RW>(d.rec + (size_t)n * RW, out);
}

// Residual of cell c, CCFVM/fvresidual.F90:16-171, as a gather over its
// faces: the boundary-face fluxes from k_bface first, then the interior
// faces with recon (:190-252) on both sides, each flux evaluated from the
// same inputs by both adjacent cells.  Returns the lowest failing face key
// (set_err), INT_MAX when every flux succeeded.  SL: records of the block's
// cells [base, base + BLK) are read from srec.
template <typename R, int FLUX, bool SL>
__device__ __forceinline__ int gather_residual(const DevT<R>& d, const R* srec, int base, int c,
                                               const Gas<R>& gr, double* r) {
  constexpr int RW = Layout<R>::RW, FW = Layout<R>::FW;
  const int noc = d.noc;
  int info = ld_nt(d.cnf + c), nf = info & 0xff, nb = info >> 8;
  R own[RW];
  load_rec<R, SL>(d, srec, base, c, own);
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = 0.0;
  int ekey = INT_MAX;
  for (int s = nf - nb; s < nf; ++s) {
    int k = ld_nt(d.sl_nbr + s * noc + c) - noc;
    double F[4];
    ld4d(d.bres + (size_t)k * 4, F);
    for (int i = 0; i < 4; ++i) r[i] = r[i] + F[i];
  }
  // interior faces, CCFVM/fvresidual.F90:133-171
  const R zhi = R(1) / R(3);
  int sf[MAXF], sn[MAXF];
#pragma unroll
  for (int s = 0; s < MAXF; ++s)
    if (s < nf - nb) {
      sf[s] = ld_nt(d.sl_face + s * noc + c);
      sn[s] = ld_nt(d.sl_nbr + s * noc + c);
    }
#pragma unroll
  for (int s = 0; s < MAXF; ++s) {
    if (s >= nf - nb) break;
    int f = sf[s];
    int n = sn[s];
    bool in_side = n >= 0;
    n = in_side ? n : ~n;
    R nr[RW], fr[FW];
    load_rec<R, SL>(d, srec, base, n, nr);
    face_rec(d, f, fr);
    R ql[4], qr[4], F[4];
    // recon of this cell's side and of the neighbour's side, then the
    // face's `in` side as ql and its `out` side as qr: the same arithmetic
    // as reconstructing the in and out sides directly, with only the
    // results selected
    R dox = fr[F_CX] - own[R_CEN], doy = fr[F_CY] - own[R_CEN + 1];
    R dnx = fr[F_CX] - nr[R_CEN], dny = fr[F_CY] - nr[R_CEN + 1];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      R qo = own[R_QP + i], qn = nr[R_QP + i];
      R alfa = own[R_G + 2 * i] * dox + own[R_G + 2 * i + 1] * doy;
      R beta = qn - qo;
      R so = qo + R(0.5) * own[R_PHI + i] * (zhi * beta + (R(1) - zhi) * alfa);
      alfa = nr[R_G + 2 * i] * dnx + nr[R_G + 2 * i + 1] * dny;
      beta = qo - qn;
      R sn = qn + R(0.5) * nr[R_PHI + i] * (zhi * beta + (R(1) - zhi) * alfa);
      ql[i] = in_side ? so : sn;
      qr[i] = in_side ? sn : so;
    }
    if (!face_flux<R, FLUX>(gr, ql, qr, fr[F_NX], fr[F_NY], fr[F_AREA], F))
      ekey = min(ekey, d.nbf + d.fref[f]);
    if (in_side) {
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = r[i] + (double)F[i];
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = r[i] - (double)F[i];
    }
  }
  return ekey;
}

// The primitive refresh of the next iteration (CCFVM/main.F90:62-67), done
// by the last stage of a whole RK3 iteration from the state it has just
// written, so that the next iteration needs no pass of its own for it: the
// compact primitives and temperature qpc, qTc (the gradient pass copies them
// into the records and qT, which this stage's other cells are still
// gathering, and which stay the reference's qp until then), the
// positivity note for the next iteration (err[1], moved up by the norm) and
// the block's partial of the limiter's global range over its cells and the
// ghosts of their boundary faces (written by this stage's k_bface).  Called
// by every thread of the block; `have`: the thread's cell is interior and
// its new state is q.
template <typename R>
__device__ __forceinline__ void prim_epilogue(const DevT<R>& d, VBlock vb, const Gas<double>& g,
                                              int c, bool have, const double* q) {
  constexpr int RW = Layout<R>::RW;
  __shared__ double shr[8][BLK / 64];
  double mx[4], mn[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { mx[k] = -INFINITY; mn[k] = INFINITY; }
  if (have) {
    double p[5];
    con2prim(g, q, p);
    R q4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q4[k] = (R)p[k];
    stv<R, 4>(d.qpc + (size_t)c * 4, q4);
    d.qTc[c] = (R)p[4];
    if (p[0] <= 0.0 || p[3] <= 0.0) atomicMin(&d.err[1], d.cref[c] + 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) mx[k] = mn[k] = (double)q4[k];
    const int info = ld_nt(d.cnf + c), nf = info & 0xff, nb = info >> 8;
    for (int s = nf - nb; s < nf; ++s) {
      const int gh = ld_nt(d.sl_nbr + s * d.noc + c);
      R gq[4];
      ldv<R, 4>(d.rec + (size_t)gh * RW + R_QP, gq);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double v = (double)gq[k];
        mx[k] = v > mx[k] ? v : mx[k];
        mn[k] = v < mn[k] ? v : mn[k];
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double a = __shfl_xor(mx[k], off), b = __shfl_xor(mn[k], off);
      mx[k] = a > mx[k] ? a : mx[k];
      mn[k] = b < mn[k] ? b : mn[k];
    }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      shr[k][wave] = mx[k];
      shr[4 + k][wave] = mn[k];
    }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    double a = shr[k][0];
    for (int w = 1; w < BLK / 64; ++w)
      a = k < 4 ? (shr[k][w] > a ? shr[k][w] : a) : (shr[k][w] < a ? shr[k][w] : a);
    d.rpart[(size_t)vb.b * 8 + k] = a;
  }
}

// Residual stage over the block's cells (gather_residual), then
//   mode & 3 = 0: store the residual only;
//              1: apply the stage update of CCFVM/main.F90:92-101 (no other
//                 cell reads this cell's state during a stage);
//              2: 1 and a block partial of the squared residual;
//   (mode >> 2) & 3 (src): where the update's states live, stage_old/in/out;
//   mode & 16: the XCDs walk their slices of blocks from the far end.
// SL: the block's records are first copied to LDS, so the neighbours inside
// the block (most of them, the cells being Hilbert-ordered) are read from
// there instead of through L2.  epi: the last stage of a whole iteration
// also refreshes the primitives (prim_epilogue).
template <typename R, int FLUX, bool SL>
__device__ __forceinline__ void stage_body(const DevT<R>& d, VBlock vb, Gas<R> gr, int mode,
                                           double ak, double bk, int write_res,
                                           const Gas<double>& g, int epi) {
  const int op = mode & 3, src = (mode >> 2) & 3;
  const bool rev = (mode & 16) != 0;
  extern __shared__ double lds_raw[];
  R* srec = reinterpret_cast<R*>(lds_raw);  // [BLK][RW] when SL
  __shared__ double sh[1][BLK];
  const int base = xcd_block_dir(vb.b, vb.n, rev) * BLK;
  const int c = base + threadIdx.x;
  if (SL) {
    stage_block_recs(d, srec, base, min(BLK, d.noc - base));
    __syncthreads();
  }
  double sq = 0.0, q[4] = {0.0, 0.0, 0.0, 0.0};
  bool have = false;
  if (c < d.noc && !d.err[0]) {
    double r[4];
    const int ekey = gather_residual<R, FLUX, SL>(d, srec, base, c, gr, r);
    if (ekey != INT_MAX) set_err(d, DERR_ROE, ekey);
    if (op == 0 || write_res) st4d_nt(d.res + (size_t)c * 4, r);
    if (op >= 1) {
      double dta = ld_nt(d.dta + c), o[4];
      ld4d_nt(stage_old(d, src) + (size_t)c * 4, o);
      if (src == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = o[i];
      } else {
        ld4d_nt(stage_in(d, src) + (size_t)c * 4, q);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = rk_update1(ak, o[i], bk, q[i], dta, r[i]);
      st4d_nt(stage_out(d, src) + (size_t)c * 4, q);
      have = true;
    }
    if (op == 2)
      for (int i = 0; i < 4; ++i) sq = sq + r[i] * r[i];
  }
  if (op == 2) {
    block_reduce<1>(sh, &sq, OpAdd());
    if (threadIdx.x == 0) d.part[vb.b] = sq;
  }
  if (epi) prim_epilogue<R>(d, vb, g, c, have, q);
}
template <typename R, int FLUX, bool SL>
__global__ __launch_bounds__(BLK) void k_stage(DevT<R> d, Gas<R> gr, int mode, double ak, double bk,
                                               int write_res, Gas<double> g, int epi) {
  stage_body<R, FLUX, SL>(d, hw_block(), gr, mode, ak, bk, write_res, g, epi);
}

// Flux through interior face `f` from the records of its `in` cell (a) and
// `out` cell (b): recon of both sides, CCFVM/fvresidual.F90:190-252
// (zeta = 1/3), then the flux function.  Same arithmetic as k_stage.
template <typename R, int FLUX>
__device__ __forceinline__ bool face_flux_ab(const Gas<R>& gr, const R* ra, const R* rb,
                                             const R* fr, R* F) {
  const R zhi = R(1) / R(3);
  R ql[4], qr[4];
  R dlx = fr[F_CX] - ra[R_CEN], dly = fr[F_CY] - ra[R_CEN + 1];
  R drx = fr[F_CX] - rb[R_CEN], dry = fr[F_CY] - rb[R_CEN + 1];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R qa = ra[R_QP + i], qb = rb[R_QP + i];
    R alfa = ra[R_G + 2 * i] * dlx + ra[R_G + 2 * i + 1] * dly;
    R beta = qb - qa;
    ql[i] = qa + R(0.5) * ra[R_PHI + i] * (zhi * beta + (R(1) - zhi) * alfa);
    alfa = rb[R_G + 2 * i] * drx + rb[R_G + 2 * i + 1] * dry;
    beta = qa - qb;
    qr[i] = qb + R(0.5) * rb[R_PHI + i] * (zhi * beta + (R(1) - zhi) * alfa);
  }
  return face_flux<R, FLUX>(gr, ql, qr, fr[F_NX], fr[F_NY], fr[F_AREA], F);
}

// The residual stage of k_stage with every interior flux evaluated once per
// block instead of once per adjacent cell.  Phase 1: the block's threads
// walk the interior faces touched by its BLK cells (Hilbert-ordered, so
// about 1.65 faces per cell instead of 3), evaluate each flux once and keep
// it in LDS.  Phase 2: each cell sums the boundary fluxes and then its
// interior fluxes from LDS in its slot order (the reference's accumulation
// order), then applies the stage update as k_stage does.  SL as in k_stage.
// LDS: [BLK][RW] records when SL, then [bf_max][4] fluxes; when SL and the
// list has at most 2 BLK faces, the fluxes are held in registers until every
// thread is done with the records and then written over them, so the block
// needs no more LDS than its records.
template <typename R, int FLUX, bool SL>
__device__ __forceinline__ void stage_blk_body(const DevT<R>& d, VBlock vb, Gas<R> gr, int mode,
                                               double ak, double bk, int write_res,
                                               const Gas<double>& g, int epi) {
  const int op = mode & 3, src = (mode >> 2) & 3;
  const bool rev = (mode & 16) != 0;
  constexpr int RW = Layout<R>::RW, FW = Layout<R>::FW;
  extern __shared__ double lds_raw[];
  R* srec = reinterpret_cast<R*>(lds_raw);
  R* fl = srec + (SL ? BLK * RW : 0);
  __shared__ double sh[1][BLK];
  __shared__ int s_err;
  if (threadIdx.x == 0) s_err = d.err[0];
  const int b = xcd_block_dir(vb.b, vb.n, rev), base = b * BLK;
  if (SL) stage_block_recs(d, srec, base, min(BLK, d.noc - base));
  __syncthreads();
  const bool run = s_err == 0;  // uniform over the block
  const int j0 = d.bf_ptr[b], nj = run ? d.bf_ptr[b + 1] - j0 : 0;
  if (SL && d.bf_max <= 2 * BLK) {
    R F2[2][4];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int j = threadIdx.x + k * BLK;
      if (j < nj) {
        const int2 ab = d.bf_ab[j0 + j];
        R ra[RW], rb[RW], fr[FW];
        ldv<R, FW>(d.bf_rec + (size_t)(j0 + j) * FW, fr);
        load_rec<R, SL>(d, srec, base, ab.x, ra);
        load_rec<R, SL>(d, srec, base, ab.y, rb);
        if (!face_flux_ab<R, FLUX>(gr, ra, rb, fr, F2[k]))
          set_err(d, DERR_ROE, d.nbf + d.bf_fref[j0 + j]);
      }
    }
    __syncthreads();
    fl = srec;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int j = threadIdx.x + k * BLK;
      if (j < nj)
#pragma unroll
        for (int i = 0; i < 4; ++i) fl[j * 4 + i] = F2[k][i];
    }
  } else
  for (int j = threadIdx.x; j < nj; j += BLK) {
    const int2 ab = d.bf_ab[j0 + j];
    const int ca = ab.x, cb = ab.y;
    R ra[RW], rb[RW], fr[FW], F[4];
    ldv<R, FW>(d.bf_rec + (size_t)(j0 + j) * FW, fr);
    load_rec<R, SL>(d, srec, base, ca, ra);
    load_rec<R, SL>(d, srec, base, cb, rb);
    if (!face_flux_ab<R, FLUX>(gr, ra, rb, fr, F)) set_err(d, DERR_ROE, d.nbf + d.bf_fref[j0 + j]);
#pragma unroll
    for (int i = 0; i < 4; ++i) fl[j * 4 + i] = F[i];
  }
  __syncthreads();
  const int c = base + threadIdx.x;
  double sq = 0.0, q[4] = {0.0, 0.0, 0.0, 0.0};
  bool have = false;
  if (run && c < d.noc) {
    const int noc = d.noc;
    int info = ld_nt(d.cnf + c), nf = info & 0xff, nb = info >> 8;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = nf - nb; s < nf; ++s) {
      int k = ld_nt(d.sl_nbr + s * noc + c) - noc;
      double F[4];
      ld4d(d.bres + (size_t)k * 4, F);
      for (int i = 0; i < 4; ++i) r[i] = r[i] + F[i];
    }
    int sl[MAXF];
#pragma unroll
    for (int s = 0; s < MAXF; ++s)
      if (s < nf - nb) sl[s] = ld_nt(d.sl_loc + s * noc + c);
#pragma unroll
    for (int s = 0; s < MAXF; ++s) {
      if (s >= nf - nb) break;
      int l = sl[s];
      bool in_side = l >= 0;
      l = in_side ? l : ~l;
      if (in_side) {
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = r[i] + (double)fl[l * 4 + i];
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = r[i] - (double)fl[l * 4 + i];
      }
    }
    if (op == 0 || write_res) st4d_nt(d.res + (size_t)c * 4, r);
    if (op >= 1) {
      double dta = ld_nt(d.dta + c), o[4];
      ld4d_nt(stage_old(d, src) + (size_t)c * 4, o);
      if (src == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = o[i];
      } else {
        ld4d_nt(stage_in(d, src) + (size_t)c * 4, q);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = rk_update1(ak, o[i], bk, q[i], dta, r[i]);
      st4d_nt(stage_out(d, src) + (size_t)c * 4, q);
      have = true;
    }
    if (op == 2)
      for (int i = 0; i < 4; ++i) sq = sq + r[i] * r[i];
  }
  if (op == 2) {
    block_reduce<1>(sh, &sq, OpAdd());
    if (threadIdx.x == 0) d.part[vb.b] = sq;
  }
  if (epi) prim_epilogue<R>(d, vb, g, c, have, q);
}
template <typename R, int FLUX, bool SL>
__global__ __launch_bounds__(BLK) void k_stage_blk(DevT<R> d, Gas<R> gr, int mode, double ak,
                                                   double bk, int write_res, Gas<double> g,
                                                   int epi) {
  stage_blk_body<R, FLUX, SL>(d, hw_block(), gr, mode, ak, bk, write_res, g, epi);
}

// Sutherland viscosity of the interior cells from the current state,
// CCFVM/viscous_flux.F90:20-22,127-139.
template <typename R>
__global__ __launch_bounds__(BLK) void k_visc_mul(DevT<R> d, Gas<double> g) {
  int c = cell_index();
  if (c >= d.noc || d.err[0]) return;
  double q[4], p[5];
  ld4d(d.qc + (size_t)c * 4, q);
  con2prim(g, q, p);
  double num = 1.0 + 300.0 / 110.4;
  double den = p[4] + 300.0 / 110.4;
  d.cmul[c] = (R)((num / den) * pow(p[4], 1.5));
}

// The viscous flux of every face once (the two cells of an interior face
// see the same face viscosity, velocities and gradients,
// CCFVM/viscous_flux.F90:24-76): the flux times the face length into fvis
// (the residual adds it at `out`, subtracts it at `in`), the face viscosity,
// and for boundary faces the wall force terms.
template <typename R>
__global__ __launch_bounds__(BLK) void k_visc_face(DevT<R> d, Gas<double> g, double rey,
                                                   double m_inf, double* bfluxv) {
  const int f = cell_index();
  if (f >= d.nof || d.err[0]) return;
  const int noc = d.noc, ci = d.fin[f], co = d.fout[f];
  const double mu = co < noc ? 0.5 * ((double)d.cmul[ci] + (double)d.cmul[co]) : (double)d.cmul[ci];
  R fr[Layout<R>::FW];
  face_rec(d, f, fr);
  const double nx = fr[F_NX], ny = fr[F_NY], area = fr[F_AREA];
  R fq[5];
  face_prims<R, 3>(d, f, fq);
  R fg[10];
  ldv<R, 10>(d.fgrad + (size_t)f * 10, fg);
  const double two = 2.0, two_third = two / 3.0;
  double uvel = fq[1], vvel = fq[2];
  double ux = fg[2], uy = fg[3], vx = fg[4], vy = fg[5], tx = fg[8], ty = fg[9];
  double tauxx = two_third * (two * ux - vy);
  double tauxy = uy + vx;
  double tauyy = two_third * (two * vy - ux);
  double factor = (g.gamma - 1.0) * m_inf * m_inf * 0.72;
  double qx = -(tx / factor), qy = -(ty / factor);
  double mr = mu / rey;
  double F[4];
  F[0] = mr * ((uvel * tauxx + vvel * tauxy - qx) * nx + (uvel * tauxy + vvel * tauyy - qy) * ny);
  F[1] = 0.0;
  F[2] = mr * (tauxx * nx + tauxy * ny);
  F[3] = mr * (tauxy * nx + tauyy * ny);
  double o[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = F[i] * area;
  st4d(d.fvis + (size_t)f * 4, o);
  d.fmu[f] = mu;
  if (co >= noc) {
    const int k = co - noc;
    bfluxv[k] = -o[2];
    bfluxv[d.nbf + k] = -o[3];
  }
}

// The viscous fluxes summed into each cell's residual over its faces in
// ascending order, after the inviscid ones.
template <typename R>
__global__ __launch_bounds__(BLK) void k_visc_sum(DevT<R> d) {
  const int c = cell_index();
  if (c >= d.noc || d.err[0]) return;
  const int noc = d.noc, nf = d.cnf[c] & 0xff;
  int sf[MAXF], sn[MAXF];
#pragma unroll
  for (int s = 0; s < MAXF; ++s)
    if (s < nf) {
      sf[s] = d.sl_face[s * noc + c];
      sn[s] = d.sl_nbr[s * noc + c];
    }
  double r[4];
  ld4d(d.res + (size_t)c * 4, r);
#pragma unroll
  for (int s = 0; s < MAXF; ++s) {
    if (s >= nf) break;
    double o[4];
    ld4d(d.fvis + (size_t)sf[s] * 4, o);
    if (sn[s] >= 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = r[i] - o[i];
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = r[i] + o[i];
    }
  }
  st4d(d.res + (size_t)c * 4, r);
}

// One Jacobi sweep of the central implicit residual smoothing,
// CCFVM/ImpResiSmoothening.F90:27-44.
// upd: also the stage update of CCFVM/main.F90:92-101 with the smoothed
// residual (the last sweep), as k_rk_update does it.
template <typename R>
__global__ __launch_bounds__(BLK) void k_cirs(DevT<R> d, double eps, const double* src,
                                              const double* old, double* dst, int upd,
                                              double ak, double bk) {
  int c = cell_index();
  if (c >= d.noc || d.err[0]) return;
  const int noc = d.noc;
  int info = d.cnf[c], nf = info & 0xff, nb = info >> 8;
  double it[4] = {0, 0, 0, 0};
  for (int s = 0; s < nf - nb; ++s) {
    int n = d.sl_nbr[s * noc + c];
    n = n >= 0 ? n : ~n;
    double v[4];
    ld4d(src + (size_t)n * 4, v);
    for (int i = 0; i < 4; ++i) it[i] = it[i] + v[i];
  }
  double den = 1.0 / (1.0 + eps * (double)d.nc2c[c]);
  double o[4], out[4];
  ld4d(old + (size_t)c * 4, o);
  for (int i = 0; i < 4; ++i) out[i] = (it[i] * eps + o[i]) * den;
  st4d(dst + (size_t)c * 4, out);
  if (upd) {
    double q[4], qo[4], dta = d.dta[c];
    ld4d(d.qc + (size_t)c * 4, q);
    ld4d(d.qold + (size_t)c * 4, qo);
    for (int i = 0; i < 4; ++i) q[i] = rk_update1(ak, qo[i], bk, q[i], dta, out[i]);
    st4d(d.qc + (size_t)c * 4, q);
  }
}

// Stage update alone, CCFVM/main.F90:92-101.
template <typename R>
__global__ __launch_bounds__(BLK) void k_rk_update(DevT<R> d, double ak, double bk) {
  int c = cell_index();
  if (c >= d.noc || d.err[0]) return;
  double q[4], o[4], r[4], dta = d.dta[c];
  ld4d(d.qc + (size_t)c * 4, q);
  ld4d(d.qold + (size_t)c * 4, o);
  ld4d(d.res + (size_t)c * 4, r);
  for (int i = 0; i < 4; ++i) q[i] = rk_update1(ak, o[i], bk, q[i], dta, r[i]);
  st4d(d.qc + (size_t)c * 4, q);
}

// Block partials of sum(res^2), CCFVM/solution.F90:152-157.
template <typename R>
__global__ __launch_bounds__(BLK) void k_sumsq(DevT<R> d) {
  __shared__ double sh[1][BLK];
  int c = cell_index();
  double sq = 0.0;
  if (c < d.noc) {
    double r[4];
    ld4d(d.res + (size_t)c * 4, r);
    for (int i = 0; i < 4; ++i) sq = sq + r[i] * r[i];
  }
  block_reduce<1>(sh, &sq, OpAdd());
  if (threadIdx.x == 0) d.part[blockIdx.x] = sq;
}

// Block partials of the squares of one stored residual variable (var 0..3)
// or of all four (var -1).
template <typename R>
__device__ __forceinline__ void sumsq_var_body(const DevT<R>& d, VBlock vb, int var,
                                               double* part) {
  __shared__ double sh[1][BLK];
  int c = vb.cell();
  double sq = 0.0;
  if (c < d.noc) {
    double r[4];
    ld4d(d.res + (size_t)c * 4, r);
    if (var < 0)
      for (int i = 0; i < 4; ++i) sq = sq + r[i] * r[i];
    else
      sq = r[var] * r[var];
  }
  block_reduce<1>(sh, &sq, OpAdd());
  if (threadIdx.x == 0) part[vb.b] = sq;
}
// block partials of the per-variable sum of squares into `part` (nblk_cell
// doubles: the context's reduction buffer, or a caller's scratch)
template <typename R>
__global__ __launch_bounds__(BLK) void k_sumsq_var(DevT<R> d, int var, double* part) {
  sumsq_var_body<R>(d, hw_block(), var, part);
}

template <typename R>
__device__ __forceinline__ void norm_var_final_body(const DevT<R>& d, int nblk, int var,
                                                    const double* part, double* out) {
  __shared__ double sh[1][BLK];
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += BLK) s = s + part[b];
  block_reduce<1>(sh, &s, OpAdd());
  if (threadIdx.x == 0) out[0] = var < 0 ? sqrt(s / d.noc / 4) : sqrt(s / d.noc);
}
template <typename R>
__global__ void k_norm_var_final(DevT<R> d, int nblk, int var, const double* part,
                                 double* out) {
  norm_var_final_body<R>(d, nblk, var, part, out);
}

// L2 norm (CCFVM/solution.F90:159-166): raw norm, iteration-1 normalisation,
// normalised history entry.
template <typename R>
__device__ __forceinline__ void norm_final_body(const DevT<R>& d, VBlock vb, int nblk, int first,
                                                double* hist, int k) {
  __shared__ double sh[1][BLK];
  double s = 0.0;
  // each thread's partials in a fixed order; unrolled so that their loads
  // are issued together
#pragma unroll 8
  for (int b = threadIdx.x; b < nblk; b += BLK) s = s + d.part[b];
  block_reduce<1>(sh, &s, OpAdd());
  if (threadIdx.x == 0) {
    // the positivity check of the iteration, CCFVM/main.F90:113, on the
    // primitives it started from (err[2], kept for the report); the note
    // taken on the primitives the next iteration starts from (err[1], when
    // the last stage refreshed them) moves up
    if (d.err[0] == 0 && d.err[2] != INT_MAX) d.err[0] = DERR_POS;
    if (d.err[0] == 0) {
      d.err[2] = d.err[1];
      d.err[1] = INT_MAX;
    }
    double fres = sqrt(s / d.noc / 4);
    d.scal[SC_NORM] = fres;
    if (first) d.scal[SC_FRES1] = fres;
    double f1 = d.scal[SC_FRES1];
    if (hist) hist[k] = f1 >= 2.220446049250313e-16 ? fres / f1 : fres;
  }
}
template <typename R>
__global__ void k_norm_final(DevT<R> d, int nblk, int first, double* hist, int k) {
  norm_final_body<R>(d, hw_block(), nblk, first, hist, k);
}

// The working primitives set back to the records' (the reference's current
// qp and temperature), and the positivity note taken ahead on the ones the
// next iteration would have started from dropped (a primitive refresh notes
// it again), before a routine call after a whole iteration that refreshed
// them ahead.
template <typename R>
__global__ __launch_bounds__(BLK) void k_prims_current(DevT<R> d) {
  constexpr int RW = Layout<R>::RW;
  int c = blockIdx.x * BLK + threadIdx.x;
  if (c == 0) d.err[2] = INT_MAX;
  if (c >= d.noc) return;
  R q[4];
  ldv<R, 4>(d.rec + (size_t)c * RW + R_QP, q);
  stv<R, 4>(d.qpc + (size_t)c * 4, q);
  d.qTc[c] = d.qT[c];
}

template <typename R>
__global__ __launch_bounds__(BLK) void k_positivity(DevT<R> d) {
  constexpr int RW = Layout<R>::RW;
  int c = cell_index();
  if (c >= d.noc) return;
  R q[4];
  ldv<R, 4>(d.rec + (size_t)c * RW + R_QP, q);
  if (q[0] <= R(0) || q[3] <= R(0)) atomicMin(&d.err[2], d.cref[c] + 1);
}

template <typename R>
__global__ void k_dtmin(DevT<R> d) {
  __shared__ double sh[1][BLK];
  double m = INFINITY;
  for (int c = threadIdx.x; c < d.noc; c += BLK) m = d.dt[c] < m ? d.dt[c] : m;
  block_reduce<1>(sh, &m, OpMin());
  if (threadIdx.x == 0) d.scal[SC_DTMIN] = m;
}

// ------------------------------------------------------------------ LU-SGS
// Normal flux difference of lusgs2's normal_flux, CCFVM/lusgs.F90:328-359.
__device__ __forceinline__ void normal_flux(const Gas<double>& g, const double* qold,
                                            const double* dq, double nx, double ny,
                                            double area, double* flux) {
  double q[4], p[5], f1[4], f2[4];
  for (int k = 0; k < 4; ++k) q[k] = qold[k];
  con2prim(g, q, p);
  double e = q[0] / p[0], et = e * p[0], un = p[1] * nx + p[2] * ny;
  f1[0] = (et + p[3]) * un;
  f1[1] = p[0] * un;
  f1[2] = f1[1] * p[1] + p[3] * nx;
  f1[3] = f1[1] * p[2] + p[3] * ny;
  for (int k = 0; k < 4; ++k) q[k] = qold[k] + dq[k];
  con2prim(g, q, p);
  e = q[0] / p[0];
  et = e * p[0];
  un = p[1] * nx + p[2] * ny;
  f2[0] = (et + p[3]) * un;
  f2[1] = p[0] * un;
  f2[2] = f2[1] * p[1] + p[3] * nx;
  f2[3] = f2[1] * p[2] + p[3] * ny;
  for (int k = 0; k < 4; ++k) flux[k] = area * (f2[k] - f1[k]);
}

// One level of the forward (dir 0) or backward (dir 1) sweep of lusgs2,
// CCFVM/lusgs.F90:244-310: every cell of a level depends only on cells of
// earlier levels, so a level runs in parallel and the sweep reproduces the
// reference's sequential order exactly.  With MC the cells are ordered by
// colour instead of by number (multicolour form): `cells` is one colour,
// whose cells have no face in common, and the already-swept neighbours are
// those of lower (forward) or higher (backward) colour.
template <typename R, bool MC>
__global__ __launch_bounds__(BLK) void k_lusgs_level(DevT<R> d, Gas<double> g, const int* cells,
                                                     int n, int dir) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n || d.err[0]) return;
  const double omega = 1.5;
  int it = cells[t];
  const int noc = d.noc;
  double cres[4] = {0, 0, 0, 0};
  int nf = d.lu_cnt[it];
  double* w = dir == 0 ? d.lower : d.upper;
  for (int j = 0; j < nf; ++j) {
    int f = d.lu_face[j * noc + it];
    R fr[Layout<R>::FW];
    face_rec(d, f, fr);
    double nx = fr[F_NX], ny = fr[F_NY];
    int in = d.fin[f], out = d.fout[f];
    if (out == it) { out = in; nx = -(double)fr[F_NX]; ny = -(double)fr[F_NY]; }
    // lower/upper neighbour in the reference's cell numbering (or colours)
    bool use;
    if constexpr (MC)
      use = out < noc && (dir == 0 ? d.ccol[out] < d.ccol[it] : d.ccol[out] > d.ccol[it]);
    else
      use = out < noc && (dir == 0 ? d.cref[out] < d.cref[it] : d.cref[out] > d.cref[it]);
    if (use) {
      double qo[4], dq[4], df[4];
      ld4d(d.qold + (size_t)out * 4, qo);
      ld4d(w + (size_t)out * 4, dq);
      normal_flux(g, qo, dq, nx, ny, (double)fr[F_AREA], df);
      double lam = d.fla[f];
      for (int k = 0; k < 4; ++k) cres[k] = cres[k] + (df[k] - omega * lam * dq[k]);
    }
  }
  double diag = d.cv[it] / d.dt[it] + 0.5 * omega * d.la[it];
  double o[4];
  if (dir == 0) {
    double r[4];
    ld4d(d.res + (size_t)it * 4, r);
    for (int k = 0; k < 4; ++k) o[k] = (-r[k] - 0.5 * cres[k]) / diag;
    st4d(d.lower + (size_t)it * 4, o);
  } else {
    double lo[4];
    ld4d(d.lower + (size_t)it * 4, lo);
    for (int k = 0; k < 4; ++k) o[k] = lo[k] - 0.5 * cres[k] / diag;
    st4d(d.upper + (size_t)it * 4, o);
  }
}

// qc = qold + upper, CCFVM/lusgs.F90:314-318.
template <typename R>
__global__ __launch_bounds__(BLK) void k_lusgs_final(DevT<R> d) {
  int c = cell_index();
  if (c >= d.noc || d.err[0]) return;
  double o[4], u[4];
  ld4d(d.qold + (size_t)c * 4, o);
  ld4d(d.upper + (size_t)c * 4, u);
  for (int k = 0; k < 4; ++k) o[k] = o[k] + u[k];
  st4d(d.qc + (size_t)c * 4, o);
}

}  // namespace ufx
