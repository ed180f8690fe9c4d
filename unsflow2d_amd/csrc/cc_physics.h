// Per-face and per-cell physics of the cell-centred solver, device side.
//
// Each function follows the reference routine cited beside it, keeping its
// operation order (left-to-right products, the same divisions) so the FP64
// form agrees with the reference to rounding.  Templated on the arithmetic
// type R: double for the FP64 form, float for the bulk work of the mixed form.
// Variable orders follow the reference: conserved qc = [E, rho, rho u, rho v],
// primitives qp = [rho, u, v, p, T], fluxes in the conserved order.
#pragma once

#include <hip/hip_runtime.h>

namespace ufx {

// Gas constants and free stream, as set by CCFVM/solution.F90:4-13,18-49.
template <typename R>
struct Gas {
  R gamma, gamma1;       // 1.4, gamma - 1
  R znd;                 // 1 / (gamma M^2), CCFVM/types/pri.F90:29
  R r_inf, u_inf, v_inf, p_inf;
  R mref;                // AUSM reference Mach, CCFVM/ausm_up_flux.F90:79
  R igamma1;             // 1 / (gamma - 1)
};

template <typename R> __device__ __forceinline__ R fsqrt(R x);
template <> __device__ __forceinline__ double fsqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float fsqrt<float>(float x) { return sqrtf(x); }
template <typename R> __device__ __forceinline__ R rabs(R x) { return x < R(0) ? -x : x; }
template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }
template <typename R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
template <typename R> __device__ __forceinline__ R rpow(R a, R b);
template <> __device__ __forceinline__ double rpow<double>(double a, double b) { return pow(a, b); }
template <> __device__ __forceinline__ float rpow<float>(float a, float b) { return powf(a, b); }

// con2prim, CCFVM/types/pri.F90:13-37 (also yields the sound speed).
template <typename R>
__device__ __forceinline__ void con2prim(const Gas<R>& g, const R* qc, R* qp) {
  R rho = qc[1];
  R u = qc[2] / qc[1];
  R v = qc[3] / qc[1];
  R et = qc[0];
  R q = u * u + v * v;
  R p = g.gamma1 * (et - R(0.5) * rho * q);
  qp[0] = rho;
  qp[1] = u;
  qp[2] = v;
  qp[3] = p;
  qp[4] = p / (rho * g.znd);
}

// prim2con, CCFVM/types/pri.F90:39-57.
template <typename R>
__device__ __forceinline__ void prim2con(const Gas<R>& g, const R* qp, R* qc) {
  R rho = qp[0], u = qp[1], v = qp[2], p = qp[3];
  qc[0] = p / g.gamma1 + R(0.5) * rho * (u * u + v * v);
  qc[1] = rho;
  qc[2] = rho * u;
  qc[3] = rho * v;
}

// Characteristic farfield boundary state, CCFVM/boundCond.F90:112-209.
template <typename R>
__device__ __forceinline__ void farfield_state(const Gas<R>& g, const R* ql, R nx, R ny, R* qr) {
  const R gam = g.gamma, gm1 = g.gamma1;
  R tx = -ny, ty = nx;
  R rinf = g.r_inf, uinf = g.u_inf, vinf = g.v_inf, pinf = g.p_inf;
  R ainf = fsqrt(gam * pinf / rinf);
  R rho = ql[0], u = ql[1], v = ql[2], p = ql[3];
  R a = fsqrt(gam * p / rho);
  R un = u * nx + v * ny;
  R tn = u * tx + v * ty;
  R mach = un / a;
  R rhob = rinf, ub = uinf, vb = vinf, pb = pinf;
  if (mach >= R(1)) {
    rhob = rho;
    ub = un * nx + tn * tx;
    vb = un * ny + tn * ty;
    pb = p;
  } else if (R(0) <= mach && mach < R(1)) {
    R vninf = uinf * nx + vinf * ny;
    R rp = un + R(2) * a / gm1;
    R rm = vninf - R(2) * ainf / gm1;
    R sb = p / rpow(rho, gam);
    R vnrm = R(0.5) * (rp + rm);
    ub = vnrm * nx + tn * tx;
    vb = vnrm * ny + tn * ty;
    R ab = R(0.25) * (rp - rm) * gm1;
    rhob = rpow(ab * ab / gam / sb, R(1) / gm1);
    pb = rhob * ab * ab / gam;
  } else if (R(-1) < mach && mach < R(0)) {
    R vninf = uinf * nx + vinf * ny;
    R vtinf = uinf * tx + vinf * ty;
    R rp = un + R(2) * a / gm1;
    R rm = vninf - R(2) * ainf / gm1;
    R sb = pinf / rpow(rinf, gam);
    R vnrm = R(0.5) * (rp + rm);
    ub = vnrm * nx + vtinf * tx;
    vb = vnrm * ny + vtinf * ty;
    R ab = R(0.25) * (rp - rm) * gm1;
    rhob = rpow(ab * ab / gam / sb, R(1) / gm1);
    pb = rhob * ab * ab / gam;
  }
  // mach <= -1 (and NaN) keeps the free stream, as the reference's last
  // branch and its defaults do
  qr[0] = rhob;
  qr[1] = ub;
  qr[2] = vb;
  qr[3] = pb;
}

// Roe flux with Harten's entropy fix, CCFVM/roe_flux.F90:1-155.  Returns false
// when the Roe-averaged sound speed squared is not positive (the reference
// STOPs there, :73-87).  The reference's divisions by gamma - 1, by the
// entropy-fix width and by the averaged sound speed (its square) are taken
// as products with one reciprocal each: 5 divisions per face instead of 11,
// a last-bit difference from the reference's order.
template <typename R>
__device__ __forceinline__ bool roe_flux(const Gas<R>& g, const R* ql, const R* qr, R nx, R ny,
                                         R area, R* flux) {
  const R gam = g.gamma, gm1 = g.gamma1, igm1 = g.igamma1;
  const R etol = R(0.1);
  R rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  R rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  R ql2 = ul * ul + vl * vl;
  R al2 = gam * pl / rl;
  R hl = al2 * igm1 + R(0.5) * ql2;
  R qr2 = ur * ur + vr * vr;
  R ar2 = gam * pr / rr;
  R hr = ar2 * igm1 + R(0.5) * qr2;
  R unl = ul * nx + vl * ny;
  R unr = ur * nx + vr * ny;
  R fc0 = rl * hl * unl + rr * hr * unr;
  R fc1 = rl * unl + rr * unr;
  R fc2 = pl * nx + rl * ul * unl + pr * nx + rr * ur * unr;
  R fc3 = pl * ny + rl * vl * unl + pr * ny + rr * vr * unr;
  R rl12 = fsqrt(rl);
  R rr12 = fsqrt(rr);
  R rd = R(1) / (rl12 + rr12);
  R ua = (ul * rl12 + ur * rr12) * rd;
  R va = (vl * rl12 + vr * rr12) * rd;
  R ha = (hl * rl12 + hr * rr12) * rd;
  R qa2 = ua * ua + va * va;
  R aa2 = gm1 * (ha - R(0.5) * qa2);
  bool ok = aa2 > R(0);
  R aa = fsqrt(aa2);
  R iaa = R(1) / aa;
  R una = ua * nx + va * ny;
  R vna = -ua * ny + va * nx;
  R e1 = rabs(una - aa);
  R e2 = rabs(una);
  R e4 = rabs(una + aa);
  R del = etol * aa;
  R idel = R(1) / del;
  R l1 = e1 < del ? R(0.5) * (del + e1 * e1 * idel) : e1;
  R l2 = e2 < del ? R(0.5) * (del + e2 * e2 * idel) : e2;
  R l3 = l2;
  R l4 = e4 < del ? R(0.5) * (del + e4 * e4 * idel) : e4;
  R du1 = rr - rl;
  R du2 = rr * ur - rl * ul;
  R du3 = rr * vr - rl * vl;
  R du4 = (rr * hr - pr) - (rl * hl - pl);
  R m1 = (nx * du2 + ny * du3 - una * du1) * iaa;
  R m2 = gm1 * (du4 - ua * du2 - va * du3 + qa2 * du1) * (iaa * iaa);
  R a1 = R(0.5) * (m2 - m1);
  R a2 = (ny * du2 - nx * du3 + vna * du1) * iaa;
  R a4 = R(0.5) * (m1 + m2);
  R a3 = du1 - a1 - a4;
  R a1l1 = a1 * l1, a2l2 = a2 * l2, a3l3 = a3 * l3, a4l4 = a4 * l4;
  R aact = aa * nx, aast = aa * ny;
  R fd0 = a1l1 * (ha - una * aa) + a2l2 * aa * vna + a3l3 * R(0.5) * qa2 + a4l4 * (ha + una * aa);
  R fd1 = a1l1 + a3l3 + a4l4;
  R fd2 = a1l1 * (ua - aact) + a2l2 * aa * ny + a3l3 * ua + a4l4 * (ua + aact);
  R fd3 = a1l1 * (va - aast) - a2l2 * aa * nx + a3l3 * va + a4l4 * (va + aast);
  R h = R(0.5) * area;
  flux[0] = h * (fc0 - fd0);
  flux[1] = h * (fc1 - fd1);
  flux[2] = h * (fc2 - fd2);
  flux[3] = h * (fc3 - fd3);
  return ok;
}

// Split Mach and pressure polynomials, CCFVM/solution.F90:265-359.
template <typename R> __device__ __forceinline__ R m1p(R m) { return R(0.5) * (m + rabs(m)); }
template <typename R> __device__ __forceinline__ R m1m(R m) { return R(0.5) * (m - rabs(m)); }
template <typename R> __device__ __forceinline__ R m2p(R m) { return R(0.25) * (m + R(1)) * (m + R(1)); }
template <typename R> __device__ __forceinline__ R m2m(R m) { return -(R(0.25) * (m - R(1)) * (m - R(1))); }
template <typename R> __device__ __forceinline__ R m4p(R m, R beta) {
  return rabs(m) >= R(1) ? R(0.5) * (m + rabs(m)) : m2p(m) * (R(1) - R(16) * beta * m2m(m));
}
template <typename R> __device__ __forceinline__ R m4m(R m, R beta) {
  return rabs(m) >= R(1) ? R(0.5) * (m - rabs(m)) : m2m(m) * (R(1) + R(16) * beta * m2p(m));
}
template <typename R> __device__ __forceinline__ R p5p(R m, R alpha) {
  return rabs(m) >= R(1) ? R(0.5) * (m + rabs(m)) / m
                         : m2p(m) * ((R(2) - m) - R(16) * alpha * m * m2m(m));
}
template <typename R> __device__ __forceinline__ R p5m(R m, R alpha) {
  return rabs(m) >= R(1) ? R(0.5) * (m - rabs(m)) / m
                         : m2m(m) * ((R(-2) - m) + R(16) * alpha * m * m2p(m));
}

// AUSM+-up, CCFVM/ausm_up_flux.F90:2-162.  As in the reference the face
// length multiplier is |n| of the unit normal (about 1), not the face length.
template <typename R>
__device__ __forceinline__ void ausm_flux(const Gas<R>& g, const R* ql, const R* qr, R nx, R ny,
                                          R* flux) {
  const R gam = g.gamma, gm1 = g.gamma1;
  const R sigma = R(1), ku = R(0.75), kp = R(0.25);
  R area = fsqrt(nx * nx + ny * ny);
  R rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  R rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  R unl = ul * nx + vl * ny;
  R hl = (gam * pl / rl) / gm1 + R(0.5) * (ul * ul + vl * vl);
  R unr = ur * nx + vr * ny;
  R hr = (gam * pr / rr) / gm1 + R(0.5) * (ur * ur + vr * vr);
  R als = fsqrt(R(2) * (gam - R(1)) * hl / (gam + R(1)));
  R ars = fsqrt(R(2) * (gam - R(1)) * hr / (gam + R(1)));
  R alh = als * als / rmax(als, rabs(unl));
  R arh = ars * ars / rmax(ars, R(-1) * rabs(unr));
  R c12 = rmin(alh, arh);
  R ml = unl / c12;
  R mr = unr / c12;
  R mbar2 = R(0.5) * (unl * unl + unr * unr) / (c12 * c12);
  R mref = g.mref;
  R fa;
  if (mbar2 >= R(1)) {
    fa = R(1);
  } else {
    R mstar2 = rmin(R(1), rmax(mbar2, mref * mref));
    R mo = mstar2;
    fa = mo * (R(2) - mo);
  }
  R beta = R(1) / R(8);
  R alpha = R(3) / R(16) * (R(-4) + R(5) * fa * fa);
  R plp = p5p(ml, alpha);
  R prm = p5m(mr, alpha);
  R mlp = m4p(ml, beta);
  R mrm = m4m(mr, beta);
  R ro12 = R(0.5) * (rl + rr);
  R mp = -((kp / fa) * rmax(R(1) - sigma * mbar2, R(0)) * (pr - pl) / (ro12 * c12 * c12));
  R m12 = mlp + mrm + mp;
  R pu = -(ku * plp * prm * R(2) * ro12 * fa * c12 * (unr - unl));
  R p12 = plp * pl + prm * pr + pu;
  R wl = m1p(m12) * rl, wr = m1m(m12) * rr;
  flux[0] = c12 * (wl * hl + wr * hr);
  flux[1] = c12 * (wl * R(1) + wr * R(1));
  flux[2] = c12 * (wl * ul + wr * ur);
  flux[3] = c12 * (wl * vl + wr * vr);
  flux[2] = flux[2] + nx * p12;
  flux[3] = flux[3] + ny * p12;
  for (int i = 0; i < 4; ++i) flux[i] = flux[i] * area;
}

// Van Leer flux-vector splitting, CCFVM/vanleer_flux.F90:1-101.
template <typename R>
__device__ __forceinline__ void vanleer_flux(const Gas<R>& g, const R* ql, const R* qr, R nx,
                                             R ny, R area, R* flux) {
  const R gam = g.gamma, gm1 = g.gamma1;
  R rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  R rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  R unl = ul * nx + vl * ny;
  R al2 = gam * pl / rl;
  R hl = al2 / gm1 + R(0.5) * (ul * ul + vl * vl);
  R cl = fsqrt(al2);
  R ml = unl / cl;
  R unr = ur * nx + vr * ny;
  R ar2 = gam * pr / rr;
  R hr = ar2 / gm1 + R(0.5) * (ur * ur + vr * vr);
  R cr = fsqrt(ar2);
  R mr = unr / cr;
  const R zero = R(0);
  R plp = p5p(ml, zero), prm = p5m(mr, zero);
  R mlp = m4p(ml, zero), mlm = m4m(ml, zero);
  R mrp = m4p(mr, zero), mrm = m4m(mr, zero);
  R p12 = plp * pl + prm * pr;
  R fmp = rl * cl * mlp;
  R fmm = rr * cr * mrm;
  R fm12 = fmp + fmm;
  R delta = -(gam * (cr * mrp * mrm * pr - cl * mlp * mlm * pl) / (gam - R(1)));
  R dissi = fmp - fmm;
  R fl[4] = {hl, R(1), ul, vl};
  R fr[4] = {hr, R(1), ur, vr};
  R fp[4] = {delta, R(0), nx * p12, ny * p12};
  for (int i = 0; i < 4; ++i)
    flux[i] = (R(0.5) * fm12 * (fl[i] + fr[i]) - R(0.5) * dissi * (fr[i] - fl[i]) + fp[i]) * area;
}

// Rusanov flux, CCFVM/rusanov_flux.F90:1-77.  The reference converts its
// argument states to conserved variables in place (:205-208); callers that
// reuse the right state afterwards see the conserved values (ql/qr are
// overwritten here the same way).
template <typename R>
__device__ __forceinline__ void rusanov_flux(const Gas<R>& g, R* ql, R* qr, R nx, R ny, R area,
                                             R* flux) {
  const R gam = g.gamma, gm1 = g.gamma1;
  R rol = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  R hl = (gam * pl / rol) / gm1 + R(0.5) * (ul * ul + vl * vl);
  R cl = fsqrt(gam * pl / rol);
  R ror = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  R hr = (gam * pr / ror) / gm1 + R(0.5) * (ur * ur + vr * vr);
  R cr = fsqrt(gam * pr / ror);
  R unl = ul * nx + vl * ny;
  R unr = ur * nx + vr * ny;
  R fl[4] = {rol * hl * unl, rol * unl, rol * unl * ul + nx * pl, rol * unl * vl + ny * pl};
  R fr[4] = {ror * hr * unr, ror * unr, ror * unr * ur + nx * pr, ror * unr * vr + ny * pr};
  R cq[4];
  prim2con(g, ql, cq);
  for (int i = 0; i < 4; ++i) ql[i] = cq[i];
  prim2con(g, qr, cq);
  for (int i = 0; i < 4; ++i) qr[i] = cq[i];
  R eig = rmax(rabs(unl) + cl, rabs(unr) + cr);
  for (int i = 0; i < 4; ++i) {
    R f = fl[i] + fr[i];
    f = f + eig * (ql[i] - qr[i]);
    flux[i] = R(0.5) * area * f;
  }
}

}  // namespace ufx
