/*
 * CPU oracle of the cell-centred hot path -- TEST INFRASTRUCTURE ONLY.
 * See cc_oracle.h.  Every routine restates the reference routine cited in
 * its comment, in the reference's loop order and expression order; built
 * without floating-point contraction so it reproduces the reference (built
 * the same way) bit for bit.
 */
#include "cc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double gamma, gamma1, znd, m_inf, r_inf, u_inf, v_inf, p_inf, mref, t_wall, aoa;
  double qinf[4];
} gas_t;

/* CCFVM/solution.F90:4-13 (math) and :18-49 (initialize) */
static gas_t gas_of(const orc_cc* s) {
  gas_t g;
  g.gamma = 1.4;
  g.gamma1 = g.gamma - 1.0;
  g.m_inf = s->mach;
  double pi = 4.0 * atan(1.0);
  g.aoa = s->aoa_deg * pi / 180.0;
  g.qinf[1] = 1.0;
  g.qinf[2] = cos(g.aoa) * 1.0;
  g.qinf[3] = sin(g.aoa) * 1.0;
  g.p_inf = 1.0 / (g.gamma * g.m_inf * g.m_inf);
  g.qinf[0] = g.p_inf / (g.gamma - 1.0) + 0.5;
  g.r_inf = g.qinf[1];
  g.u_inf = g.qinf[2] / g.qinf[1];
  g.v_inf = g.qinf[3] / g.qinf[1];
  g.znd = 1.0 / (g.gamma * g.m_inf * g.m_inf);
  double m32 = (double)0.32f, half = 0.5 * g.m_inf;
  g.mref = m32 > half ? m32 : half;
  g.t_wall = s->t_wall / 300.0;
  return g;
}

/* con2prim, CCFVM/types/pri.F90:13-37 */
static void con2prim(const gas_t* g, const double* q, double* p) {
  double rho = q[1];
  double u = q[2] / q[1];
  double v = q[3] / q[1];
  double et = q[0];
  double qq = u * u + v * v;
  double pr = (g->gamma - 1.0) * (et - 0.5 * rho * qq);
  p[0] = rho;
  p[1] = u;
  p[2] = v;
  p[3] = pr;
  p[4] = pr / (rho * g->znd);
}

/* prim2con, CCFVM/types/pri.F90:39-57 */
static void prim2con(const gas_t* g, const double* p, double* q) {
  double rho = p[0], u = p[1], v = p[2], pr = p[3];
  q[0] = pr / (g->gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  q[1] = rho;
  q[2] = rho * u;
  q[3] = rho * v;
}

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* farfield_flux, CCFVM/boundCond.F90:112-209 */
static void farfield(const gas_t* g, const double* ql, double nx, double ny, double* qr) {
  double tx = -ny, ty = nx;
  double rinf = g->r_inf, uinf = g->u_inf, vinf = g->v_inf, pinf = g->p_inf;
  double ainf = sqrt(g->gamma * pinf / rinf);
  double rho = ql[0], u = ql[1], v = ql[2], p = ql[3];
  double a = sqrt(g->gamma * p / rho);
  double un = u * nx + v * ny, tn = u * tx + v * ty;
  double rhob = rinf, ub = uinf, vb = vinf, pb = pinf;
  double mach = un / a;
  if (mach >= 1.0) {
    double vnrm = un, vtan = tn;
    rhob = rho;
    ub = vnrm * nx + vtan * tx;
    vb = vnrm * ny + vtan * ty;
    pb = p;
  } else if (0.0 <= mach && mach < 1.0) {
    double vninf = uinf * nx + vinf * ny;
    double rp = un + 2.0 * a / g->gamma1;
    double rm = vninf - 2.0 * ainf / g->gamma1;
    double sb = p / pow(rho, g->gamma);
    double vnrm = 0.5 * (rp + rm), vtan = tn;
    ub = vnrm * nx + vtan * tx;
    vb = vnrm * ny + vtan * ty;
    double ab = 0.25 * (rp - rm) * g->gamma1;
    rhob = pow(ab * ab / g->gamma / sb, 1.0 / g->gamma1);
    pb = rhob * ab * ab / g->gamma;
  } else if (-1.0 < mach && mach < 0.0) {
    double vninf = uinf * nx + vinf * ny;
    double vtinf = uinf * tx + vinf * ty;
    double rp = un + 2.0 * a / g->gamma1;
    double rm = vninf - 2.0 * ainf / g->gamma1;
    double sb = pinf / pow(rinf, g->gamma);
    double vnrm = 0.5 * (rp + rm), vtan = vtinf;
    ub = vnrm * nx + vtan * tx;
    vb = vnrm * ny + vtan * ty;
    double ab = 0.25 * (rp - rm) * g->gamma1;
    rhob = pow(ab * ab / g->gamma / sb, 1.0 / g->gamma1);
    pb = rhob * ab * ab / g->gamma;
  } else if (-1.0 <= mach) {
    rhob = rinf; ub = uinf; vb = vinf; pb = pinf;
  }
  qr[0] = rhob;
  qr[1] = ub;
  qr[2] = vb;
  qr[3] = pb;
}

/* roe_flux, CCFVM/roe_flux.F90:1-155 */
static int roe(const gas_t* g, const double* ql, const double* qr, double nx, double ny,
               double area, double* flux) {
  const double ETOL = 0.1, G = g->gamma, G1 = g->gamma1;
  double rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  double rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  double ql2 = ul * ul + vl * vl, al2 = G * pl / rl, hl = al2 / G1 + 0.5 * ql2;
  double qr2 = ur * ur + vr * vr, ar2 = G * pr / rr, hr = ar2 / G1 + 0.5 * qr2;
  double unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
  double f_c[4];
  f_c[0] = rl * hl * unl + rr * hr * unr;
  f_c[1] = rl * unl + rr * unr;
  f_c[2] = pl * nx + rl * ul * unl + pr * nx + rr * ur * unr;
  f_c[3] = pl * ny + rl * vl * unl + pr * ny + rr * vr * unr;
  double rl12 = sqrt(rl), rr12 = sqrt(rr), rd = 1.0 / (rl12 + rr12);
  double ua = (ul * rl12 + ur * rr12) * rd;
  double va = (vl * rl12 + vr * rr12) * rd;
  double ha = (hl * rl12 + hr * rr12) * rd;
  double qa2 = ua * ua + va * va;
  double aa2 = G1 * (ha - 0.5 * qa2);
  int bad = !(aa2 > 0.0);
  double aa = sqrt(aa2);
  double una = ua * nx + va * ny;
  double vna = -(ua * ny) + va * nx;
  double e1 = fabs(una - aa), e2 = fabs(una), e4 = fabs(una + aa);
  double del = ETOL * aa, l1, l2, l3, l4;
  if (e1 < del) l1 = 0.5 * (del + e1 * e1 / del); else l1 = e1;
  if (e2 < del) { l2 = 0.5 * (del + e2 * e2 / del); l3 = l2; } else { l2 = e2; l3 = l2; }
  if (e4 < del) l4 = 0.5 * (del + e4 * e4 / del); else l4 = e4;
  double du1 = rr - rl, du2 = rr * ur - rl * ul, du3 = rr * vr - rl * vl;
  double du4 = (rr * hr - pr) - (rl * hl - pl);
  double m1 = (nx * du2 + ny * du3 - una * du1) / aa;
  double m2 = G1 * (du4 - ua * du2 - va * du3 + qa2 * du1) / (aa * aa);
  double a1 = 0.5 * (m2 - m1);
  double a2 = (ny * du2 - nx * du3 + vna * du1) / aa;
  double a4 = 0.5 * (m1 + m2);
  double a3 = du1 - a1 - a4;
  double a1l1 = a1 * l1, a2l2 = a2 * l2, a3l3 = a3 * l3, a4l4 = a4 * l4;
  double aact = aa * nx, aast = aa * ny, fd[4];
  fd[0] = a1l1 * (ha - una * aa) + a2l2 * aa * vna + a3l3 * 0.5 * qa2 + a4l4 * (ha + una * aa);
  fd[1] = a1l1 + a3l3 + a4l4;
  fd[2] = a1l1 * (ua - aact) + a2l2 * aa * ny + a3l3 * ua + a4l4 * (ua + aact);
  fd[3] = a1l1 * (va - aast) - a2l2 * aa * nx + a3l3 * va + a4l4 * (va + aast);
  for (int i = 0; i < 4; ++i) flux[i] = 0.5 * area * (f_c[i] - fd[i]);
  return bad;
}

/* m1p..p5m, CCFVM/solution.F90:265-359 */
static double m1p(double m) { return 0.5 * (m + fabs(m)); }
static double m1m(double m) { return 0.5 * (m - fabs(m)); }
static double m2p(double m) { return 0.25 * (m + 1.0) * (m + 1.0); }
static double m2m(double m) { return -(0.25 * (m - 1.0) * (m - 1.0)); }
static double m4p(double m, double b) {
  if (fabs(m) >= 1.0) return 0.5 * (m + fabs(m));
  return m2p(m) * (1.0 - 16.0 * b * m2m(m));
}
static double m4m(double m, double b) {
  if (fabs(m) >= 1.0) return 0.5 * (m - fabs(m));
  return m2m(m) * (1.0 + 16.0 * b * m2p(m));
}
static double p5p(double m, double a) {
  if (fabs(m) >= 1.0) return 0.5 * (m + fabs(m)) / m;
  return m2p(m) * ((2.0 - m) - 16.0 * a * m * m2m(m));
}
static double p5m(double m, double a) {
  if (fabs(m) >= 1.0) return 0.5 * (m - fabs(m)) / m;
  return m2m(m) * ((-2.0 - m) + 16.0 * a * m * m2p(m));
}

/* ausmPlus_flux, CCFVM/ausm_up_flux.F90:2-162 */
static void ausm(const gas_t* g, const double* ql, const double* qr, double nx, double ny,
                 double* flux) {
  const double G = g->gamma, G1 = g->gamma1, sigma = 1.0, ku = 0.75, kp = 0.25;
  double area = sqrt(nx * nx + ny * ny);
  double rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  double unl = ul * nx + vl * ny, ql2 = ul * ul + vl * vl, al2 = G * pl / rl;
  double hl = al2 / G1 + 0.5 * ql2;
  double rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  double unr = ur * nx + vr * ny, qr2 = ur * ur + vr * vr, ar2 = G * pr / rr;
  double hr = ar2 / G1 + 0.5 * qr2;
  double als = sqrt(2.0 * (G - 1.0) * hl / (G + 1.0));
  double ars = sqrt(2.0 * (G - 1.0) * hr / (G + 1.0));
  double alh = als * als / dmax(als, fabs(unl));
  double arh = ars * ars / dmax(ars, -1.0 * fabs(unr));
  double c12 = dmin(alh, arh);
  double ml = unl / c12, mr = unr / c12;
  double mbar2 = 0.5 * (unl * unl + unr * unr) / (c12 * c12);
  double mref = g->mref, fa;
  if (mbar2 >= 1.0) {
    fa = 1.0;
  } else {
    double mstar2 = dmin(1.0, dmax(mbar2, mref * mref));
    double mo = mstar2;
    fa = mo * (2.0 - mo);
  }
  double beta = 1.0 / 8.0;
  double alpha = 3.0 / 16.0 * (-4.0 + 5.0 * fa * fa);
  double plp = p5p(ml, alpha), prm = p5m(mr, alpha);
  double mlp = m4p(ml, beta), mrm = m4m(mr, beta);
  double ro12 = 0.5 * (rl + rr);
  double mp = -((kp / fa) * dmax(1.0 - sigma * mbar2, 0.0) * (pr - pl) / (ro12 * c12 * c12));
  double m12 = mlp + mrm + mp;
  double pu = -(ku * plp * prm * 2.0 * ro12 * fa * c12 * (unr - unl));
  double p12 = plp * pl + prm * pr + pu;
  double fl[4] = {hl, 1.0, ul, vl}, fr[4] = {hr, 1.0, ur, vr};
  for (int i = 0; i < 4; ++i) flux[i] = c12 * (m1p(m12) * rl * fl[i] + m1m(m12) * rr * fr[i]);
  flux[2] = flux[2] + nx * p12;
  flux[3] = flux[3] + ny * p12;
  for (int i = 0; i < 4; ++i) flux[i] = flux[i] * area;
}

/* vanleer_flux, CCFVM/vanleer_flux.F90:1-101 */
static void vanleer(const gas_t* g, const double* ql, const double* qr, double nx, double ny,
                    double area, double* flux) {
  const double G = g->gamma, G1 = g->gamma1;
  double rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  double unl = ul * nx + vl * ny, ql2 = ul * ul + vl * vl, al2 = G * pl / rl;
  double hl = al2 / G1 + 0.5 * ql2, cl = sqrt(al2), ml = unl / cl;
  double rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  double unr = ur * nx + vr * ny, qr2 = ur * ur + vr * vr, ar2 = G * pr / rr;
  double hr = ar2 / G1 + 0.5 * qr2, cr = sqrt(ar2), mr = unr / cr;
  double beta = 0.0, alpha = 0.0;
  double plp = p5p(ml, alpha), prm = p5m(mr, alpha);
  double mlp = m4p(ml, beta), mlm = m4m(ml, beta), mrp = m4p(mr, beta), mrm = m4m(mr, beta);
  double p12 = plp * pl + prm * pr;
  double fmp = rl * cl * mlp, fmm = rr * cr * mrm, fm12 = fmp + fmm;
  double delta = -(G * (cr * mrp * mrm * pr - cl * mlp * mlm * pl) / (G - 1.0));
  double dissi = fmp - fmm;
  double fl[4] = {hl, 1.0, ul, vl}, fr[4] = {hr, 1.0, ur, vr};
  double fp[4] = {delta, 0.0, nx * p12, ny * p12};
  for (int i = 0; i < 4; ++i)
    flux[i] = 0.5 * fm12 * (fl[i] + fr[i]) - 0.5 * dissi * (fr[i] - fl[i]) + fp[i];
  for (int i = 0; i < 4; ++i) flux[i] = flux[i] * area;
}

/* rusanov_flux, CCFVM/rusanov_flux.F90:1-77 (converts ql, qr in place) */
static void rusanov(const gas_t* g, double* ql, double* qr, double nx, double ny, double area,
                    double* flux) {
  const double G = g->gamma, G1 = g->gamma1;
  double rol = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  double ql2 = ul * ul + vl * vl, al2 = G * pl / rol, hl = al2 / G1 + 0.5 * ql2;
  double cl = sqrt(G * pl / rol);
  double ror = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  double qr2 = ur * ur + vr * vr, ar2 = G * pr / ror, hr = ar2 / G1 + 0.5 * qr2;
  double cr = sqrt(G * pr / ror);
  double unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
  double fl[4] = {rol * hl * unl, rol * unl, rol * unl * ul + nx * pl, rol * unl * vl + ny * pl};
  double fr[4] = {ror * hr * unr, ror * unr, ror * unr * ur + nx * pr, ror * unr * vr + ny * pr};
  double t[4];
  prim2con(g, ql, t);
  memcpy(ql, t, sizeof t);
  prim2con(g, qr, t);
  memcpy(qr, t, sizeof t);
  double eigen = dmax(fabs(unl) + cl, fabs(unr) + cr);
  for (int i = 0; i < 4; ++i) flux[i] = fl[i] + fr[i];
  for (int i = 0; i < 4; ++i) flux[i] = flux[i] + eigen * (ql[i] - qr[i]);
  for (int i = 0; i < 4; ++i) flux[i] = 0.5 * area * flux[i];
}

static int num_flux(const orc_cc* s, const gas_t* g, double* ql, double* qr, int f, double* F) {
  double nx = s->fnx[f], ny = s->fny[f], area = s->farea[f];
  switch (s->flux) {
    case 1: return roe(g, ql, qr, nx, ny, area, F);
    case 2: ausm(g, ql, qr, nx, ny, F); return 0;
    case 3: vanleer(g, ql, qr, nx, ny, area, F); return 0;
    default: rusanov(g, ql, qr, nx, ny, area, F); return 0;
  }
}

int orc_flux(int kind, double mach, double* ql, double* qr, double nx, double ny, double area,
             double* flux) {
  orc_cc s;
  memset(&s, 0, sizeof s);
  s.mach = mach;
  gas_t g = gas_of(&s);
  switch (kind) {
    case 1: return roe(&g, ql, qr, nx, ny, area, flux);
    case 2: ausm(&g, ql, qr, nx, ny, flux); return 0;
    case 3: vanleer(&g, ql, qr, nx, ny, area, flux); return 0;
    default: rusanov(&g, ql, qr, nx, ny, area, flux); return 0;
  }
}

void orc_farfield_state(double mach, double aoa_deg, const double* ql, double nx, double ny,
                        double* qr) {
  orc_cc s;
  memset(&s, 0, sizeof s);
  s.mach = mach;
  s.aoa_deg = aoa_deg;
  gas_t g = gas_of(&s);
  farfield(&g, ql, nx, ny, qr);
}

static int ncell(const orc_cc* s) { return s->noc + s->nbf; }
static int nof(const orc_cc* s) { return s->nfi + s->nbf; }

/* initialize, CCFVM/solution.F90:78-81 (ghost primitives stay 0) */
void orc_cc_initialize(orc_cc* s) {
  gas_t g = gas_of(s);
  for (int c = 0; c < ncell(s); ++c)
    for (int k = 0; k < 4; ++k) s->qc[c * 4 + k] = g.qinf[k];
  for (int c = 0; c < s->noc; ++c)
    for (int k = 0; k < 4; ++k) s->phi[c * 4 + k] = 1.0;
  s->error = 0;
  s->error_cell = 0;
  s->error_face = 0;
}

/* CCFVM/main.F90:62-67 */
void orc_cc_prim(orc_cc* s) {
  gas_t g = gas_of(s);
  for (int c = 0; c < s->noc; ++c) con2prim(&g, &s->qc[c * 4], &s->qp[c * 5]);
}

/* avg_c2v, CCFVM/average.F90:1-39 */
void orc_cc_avg_c2v(orc_cc* s) {
  for (int i = 0; i < s->nop; ++i) {
    double wt = 0.0, con[6] = {0, 0, 0, 0, 0, 0};
    for (int j = s->v2c_ptr[i]; j < s->v2c_ptr[i + 1]; ++j) {
      int c = s->v2c[j];
      double cwt = s->wt[j];
      wt = wt + cwt;
      for (int k = 0; k < 5; ++k) con[k] = con[k] + s->qp[c * 5 + k] * cwt;
      con[5] = con[5] + s->cmul[c] * cwt;
    }
    for (int k = 0; k < 5; ++k) s->pprim[i * 5 + k] = con[k] / wt;
    s->pmu[i] = con[5] / wt;
  }
  for (int f = 0; f < nof(s); ++f)
    for (int k = 0; k < 5; ++k)
      s->fqp[f * 5 + k] = 0.5 * (s->pprim[s->fp1[f] * 5 + k] + s->pprim[s->fp2[f] * 5 + k]);
}

/* time_step02 + spectral02, CCFVM/timestep.F90:98-177 */
void orc_cc_time_step(orc_cc* s) {
  gas_t g = gas_of(s);
  for (int c = 0; c < s->noc; ++c) s->la[c] = 0.0;
  double vll = 0.0;
  if (s->flow_type != 1) vll = 2.0 * g.gamma / (0.72 * s->reynolds * 0.25);
  for (int f = 0; f < nof(s); ++f) {
    int in = s->fin[f], out = s->fout[f];
    double r0 = s->fqp[f * 5], u0 = s->fqp[f * 5 + 1], v0 = s->fqp[f * 5 + 2],
           p0 = s->fqp[f * 5 + 3];
    double a0 = sqrt(g.gamma * p0 / r0);
    double un = u0 * s->fnx[f] + v0 * s->fny[f];
    double ll = s->farea[f] * (fabs(un) + a0);
    ll = ll + vll * s->fmu[f] / r0;
    s->fla[f] = ll;
    s->la[in] = s->la[in] + ll;
    if (out < s->noc) s->la[out] = s->la[out] + ll;
  }
  s->dtglobal = 1.0e20;
  for (int c = 0; c < s->noc; ++c) {
    s->dt[c] = s->cfl * s->ccv[c] / s->la[c];
    s->dtglobal = dmin(s->dtglobal, s->dt[c]);
  }
}

/* save_old, CCFVM/solution.F90:118-132 */
void orc_cc_save_old(orc_cc* s) { memcpy(s->qold, s->qc, sizeof(double) * 4 * s->noc); }

static double vk_limit(double a, double b, double tol) {
  double nr = a * a + 2.0 * a * b + tol * tol;
  double dr = a * a + 2.0 * b * b + a * b + tol * tol;
  return nr / dr;
}

/* limiter_VKN1, CCFVM/limiter.F90:123-296 */
void orc_cc_limit(orc_cc* s) {
  const double kappa = 0.05;
  int noc = s->noc, nc = ncell(s);
  double qmax[4], qmin[4];
  for (int j = 0; j < 4; ++j) {
    qmax[j] = -INFINITY;
    qmin[j] = INFINITY;
    for (int c = 0; c < nc; ++c) {
      double v = s->qp[c * 5 + j];
      if (v > qmax[j]) qmax[j] = v;
      if (v < qmin[j]) qmin[j] = v;
    }
  }
  double* umax = malloc(sizeof(double) * 4 * noc);
  double* umin = malloc(sizeof(double) * 4 * noc);
  double* phi = malloc(sizeof(double) * 4 * noc);
  for (int i = 0; i < noc; ++i)
    for (int j = 0; j < 4; ++j) {
      phi[i * 4 + j] = s->phi[i * 4 + j];
      umax[i * 4 + j] = s->qp[i * 5 + j];
      umin[i * 4 + j] = s->qp[i * 5 + j];
    }
  for (int f = 0; f < s->nfi; ++f) {
    int in = s->fin[f], out = s->fout[f];
    for (int j = 0; j < 4; ++j) {
      umax[in * 4 + j] = dmax(umax[in * 4 + j], s->qp[out * 5 + j]);
      umin[in * 4 + j] = dmin(umin[in * 4 + j], s->qp[out * 5 + j]);
      umax[out * 4 + j] = dmax(umax[out * 4 + j], s->qp[in * 5 + j]);
      umin[out * 4 + j] = dmin(umin[out * 4 + j], s->qp[in * 5 + j]);
    }
  }
  for (int f = s->nfi; f < nof(s); ++f) {
    int in = s->fin[f], p1 = s->fp1[f], p2 = s->fp2[f];
    for (int j = 0; j < 4; ++j) {
      double var = 0.5 * (s->pprim[p1 * 5 + j] + s->pprim[p2 * 5 + j]);
      umax[in * 4 + j] = dmax(umax[in * 4 + j], var);
      umin[in * 4 + j] = dmin(umin[in * 4 + j], var);
    }
  }
  for (int f = 0; f < s->nfi; ++f) {
    int in = s->fin[f], out = s->fout[f];
    for (int j = 0; j < 4; ++j) {
      double tol = kappa * (qmax[j] - qmin[j]);
      int sides[2] = {in, out};
      for (int k = 0; k < 2; ++k) {
        int c = sides[k];
        double dx = s->fcx[f] - s->ccx[c], dy = s->fcy[f] - s->ccy[c];
        double dl = s->grad[c * 10 + 2 * j] * dx + s->grad[c * 10 + 2 * j + 1] * dy;
        double phi0 = 1.0;
        if (dl > 0.0) phi0 = vk_limit(umax[c * 4 + j] - s->qp[c * 5 + j], dl, tol);
        else if (dl < 0.0) phi0 = vk_limit(umin[c * 4 + j] - s->qp[c * 5 + j], dl, tol);
        phi[c * 4 + j] = dmin(phi[c * 4 + j], phi0);
      }
    }
  }
  for (int f = s->nfi; f < nof(s); ++f) {
    int c = s->fin[f];
    double dx = s->fcx[f] - s->ccx[c], dy = s->fcy[f] - s->ccy[c];
    for (int j = 0; j < 4; ++j) {
      double dl = s->grad[c * 10 + 2 * j] * dx + s->grad[c * 10 + 2 * j + 1] * dy;
      double tol = kappa * (qmax[j] - qmin[j]);
      double phi0 = 1.0;
      if (dl > 0.0) phi0 = vk_limit(umax[c * 4 + j] - s->qp[c * 5 + j], dl, tol);
      else if (dl < 0.0) phi0 = vk_limit(umin[c * 4 + j] - s->qp[c * 5 + j], dl, tol);
      phi[c * 4 + j] = dmin(phi[c * 4 + j], phi0);
    }
  }
  for (int i = 0; i < noc; ++i)
    for (int j = 0; j < 4; ++j) s->phi[i * 4 + j] = phi[i * 4 + j];
  free(umax);
  free(umin);
  free(phi);
}

/* Gradient_LSQR, CCFVM/gradient.F90:2-93 */
static void grad_lsqr(orc_cc* s) {
  for (int i = 0; i < 10 * s->noc; ++i) s->grad[i] = 0.0;
  for (int f = 0; f < s->nfi; ++f) {
    int in = s->fin[f], out = s->fout[f];
    double xc = s->ccx[in], yc = s->ccy[in], xn = s->ccx[out], yn = s->ccy[out];
    double dx = xn - xc, dy = yn - yc;
    double r11 = s->r11[in], r12 = s->r12[in], r22 = s->r22[in];
    double alfa1 = dx / r11 / r11;
    double alfa2 = (dy - dx * r12 / r11) / (r22 * r22);
    double wx0 = alfa1 - alfa2 * r12 / r11, wy0 = alfa2;
    dx = xc - xn;
    dy = yc - yn;
    r11 = s->r11[out]; r12 = s->r12[out]; r22 = s->r22[out];
    alfa1 = dx / r11 / r11;
    alfa2 = (dy - dx * r12 / r11) / (r22 * r22);
    double wx1 = alfa1 - alfa2 * r12 / r11, wy1 = alfa2;
    for (int k = 0; k < 5; ++k) {
      double var = s->qp[out * 5 + k] - s->qp[in * 5 + k];
      s->grad[in * 10 + 2 * k] = s->grad[in * 10 + 2 * k] + var * wx0;
      s->grad[in * 10 + 2 * k + 1] = s->grad[in * 10 + 2 * k + 1] + var * wy0;
      s->grad[out * 10 + 2 * k] = s->grad[out * 10 + 2 * k] - var * wx1;
      s->grad[out * 10 + 2 * k + 1] = s->grad[out * 10 + 2 * k + 1] - var * wy1;
    }
  }
  for (int f = s->nfi; f < nof(s); ++f) {
    int in = s->fin[f], out = s->fout[f];
    double dx = s->ccx[out] - s->ccx[in], dy = s->ccy[out] - s->ccy[in];
    double r11 = s->r11[in], r12 = s->r12[in], r22 = s->r22[in];
    double alfa1 = dx / r11 / r11;
    double alfa2 = (dy - dx * r12 / r11) / r22 / r22;
    double wx0 = alfa1 - alfa2 * r12 / r11, wy0 = alfa2;
    for (int k = 0; k < 5; ++k) {
      double var = s->qp[out * 5 + k] - s->qp[in * 5 + k];
      s->grad[in * 10 + 2 * k] = s->grad[in * 10 + 2 * k] + var * wx0;
      s->grad[in * 10 + 2 * k + 1] = s->grad[in * 10 + 2 * k + 1] + var * wy0;
    }
  }
}

/* Gradient_GG, CCFVM/gradient.F90:95-156 (boundary faces with -n) */
static void grad_gg(orc_cc* s) {
  for (int i = 0; i < 10 * s->noc; ++i) s->grad[i] = 0.0;
  for (int f = 0; f < s->nfi; ++f) {
    int in = s->fin[f], out = s->fout[f];
    double area = s->farea[f], dx = s->fnx[f] * area, dy = s->fny[f] * area;
    for (int j = 0; j < 5; ++j) {
      double var = s->fqp[f * 5 + j];
      s->grad[in * 10 + 2 * j] = s->grad[in * 10 + 2 * j] + var * dx;
      s->grad[in * 10 + 2 * j + 1] = s->grad[in * 10 + 2 * j + 1] + var * dy;
      s->grad[out * 10 + 2 * j] = s->grad[out * 10 + 2 * j] - var * dx;
      s->grad[out * 10 + 2 * j + 1] = s->grad[out * 10 + 2 * j + 1] - var * dy;
    }
  }
  for (int f = s->nfi; f < nof(s); ++f) {
    int in = s->fin[f];
    double area = s->farea[f], dx = -s->fnx[f] * area, dy = -s->fny[f] * area;
    for (int j = 0; j < 5; ++j) {
      double var = s->fqp[f * 5 + j];
      s->grad[in * 10 + 2 * j] = s->grad[in * 10 + 2 * j] + var * dx;
      s->grad[in * 10 + 2 * j + 1] = s->grad[in * 10 + 2 * j + 1] + var * dy;
    }
  }
  for (int c = 0; c < s->noc; ++c) {
    double var = s->ccv[c];
    for (int j = 0; j < 10; ++j) s->grad[c * 10 + j] = s->grad[c * 10 + j] / var;
  }
}

/* gradtrixy, CCFVM/gradient.F90:281-321 */
static int gradtrixy(const double prim[3][5], const double xy[3][2], double grad[10],
                     double* tarea) {
  double t = 0.0;
  for (int i = 0; i < 3; ++i) {
    int ip1 = i == 2 ? 0 : i + 1;
    double dx = 0.5 * (xy[ip1][0] + xy[i][0]);
    double dy = xy[ip1][1] - xy[i][1];
    t = t + dx * dy;
  }
  *tarea = t;
  if (t <= 0.0) return 1;
  for (int i = 0; i < 3; ++i) {
    int ip1 = i == 2 ? 0 : i + 1;
    double dx = xy[ip1][1] - xy[i][1];
    double dy = -(xy[ip1][0] - xy[i][0]);
    for (int k = 0; k < 5; ++k) {
      double var = 0.5 * (prim[i][k] + prim[ip1][k]);
      grad[2 * k] = grad[2 * k] + var * dx;
      grad[2 * k + 1] = grad[2 * k + 1] + var * dy;
    }
  }
  for (int k = 0; k < 10; ++k) grad[k] = grad[k] / t;
  return 0;
}

/* Gradient_GG_FC, CCFVM/gradient.F90:159-323 (diamond path) */
static void grad_ggfc(orc_cc* s) {
  for (int i = 0; i < 10 * s->noc; ++i) s->grad[i] = 0.0;
  for (int i = 0; i < 10 * nof(s); ++i) s->fgrad[i] = 0.0;
  for (int f = 0; f < nof(s); ++f) {
    int in = s->fin[f], out = s->fout[f], p1 = s->fp1[f], p2 = s->fp2[f];
    double prim[3][5], xy[3][2], g1[10] = {0}, g2[10] = {0}, a1, a2;
    xy[0][0] = s->px[p1]; xy[0][1] = s->py[p1];
    xy[1][0] = s->ccx[out]; xy[1][1] = s->ccy[out];
    xy[2][0] = s->px[p2]; xy[2][1] = s->py[p2];
    for (int k = 0; k < 5; ++k) {
      prim[0][k] = s->pprim[p1 * 5 + k];
      prim[1][k] = s->qp[out * 5 + k];
      prim[2][k] = s->pprim[p2 * 5 + k];
    }
    if (gradtrixy(prim, xy, g1, &a1)) { s->error = 2; s->error_cell = s->error_face = f + 1; return; }
    xy[0][0] = s->px[p1]; xy[0][1] = s->py[p1];
    xy[1][0] = s->px[p2]; xy[1][1] = s->py[p2];
    xy[2][0] = s->ccx[in]; xy[2][1] = s->ccy[in];
    for (int k = 0; k < 5; ++k) {
      prim[0][k] = s->pprim[p1 * 5 + k];
      prim[1][k] = s->pprim[p2 * 5 + k];
      prim[2][k] = s->qp[in * 5 + k];
    }
    if (gradtrixy(prim, xy, g2, &a2)) { s->error = 2; s->error_cell = s->error_face = f + 1; return; }
    for (int k = 0; k < 10; ++k) s->fgrad[f * 10 + k] = (a1 * g1[k] + a2 * g2[k]) / (a1 + a2);
    for (int k = 0; k < 10; ++k)
      s->grad[in * 10 + k] = s->grad[in * 10 + k] + s->fgrad[f * 10 + k] * (a1 + a2);
    if (f < s->nfi)
      for (int k = 0; k < 10; ++k)
        s->grad[out * 10 + k] = s->grad[out * 10 + k] + s->fgrad[f * 10 + k] * (a1 + a2);
  }
  for (int c = 0; c < s->noc; ++c)
    for (int k = 0; k < 10; ++k) s->grad[c * 10 + k] = s->grad[c * 10 + k] / s->ccov[c];
}

void orc_cc_gradient(orc_cc* s) {
  if (s->grad_type == 2) grad_gg(s);
  else if (s->grad_type == 3) grad_ggfc(s);
  else grad_lsqr(s);
  if (s->error) return;
  if (s->ilimit == 1) orc_cc_limit(s);
}

/* recon, CCFVM/fvresidual.F90:190-252 */
static void recon(const orc_cc* s, int f, int in, int out, double* qcl, double* qcr) {
  const double zhi = 1.0 / 3.0;
  double dlx = s->fcx[f] - s->ccx[in], dly = s->fcy[f] - s->ccy[in];
  double drx = s->fcx[f] - s->ccx[out], dry = s->fcy[f] - s->ccy[out];
  for (int i = 0; i < 4; ++i) {
    double phi = s->phi[in * 4 + i];
    double alfa = s->grad[in * 10 + 2 * i] * dlx + s->grad[in * 10 + 2 * i + 1] * dly;
    double beta = s->qp[out * 5 + i] - s->qp[in * 5 + i];
    qcl[i] = s->qp[in * 5 + i] + 0.5 * phi * (zhi * beta + (1.0 - zhi) * alfa);
    phi = s->phi[out * 4 + i];
    alfa = s->grad[out * 10 + 2 * i] * drx + s->grad[out * 10 + 2 * i + 1] * dry;
    beta = s->qp[in * 5 + i] - s->qp[out * 5 + i];
    qcr[i] = s->qp[out * 5 + i] + 0.5 * phi * (zhi * beta + (1.0 - zhi) * alfa);
  }
}

/* sutherland_viscosity, CCFVM/viscous_flux.F90:127-139 */
static double sutherland(const gas_t* g, const double* qc) {
  double p[5];
  con2prim(g, qc, p);
  double num = 1.0 + 300.0 / 110.4;
  double den = p[4] + 300.0 / 110.4;
  return (num / den) * pow(p[4], 1.5);
}

/* viscous_flux, CCFVM/viscous_flux.F90:1-82 */
static void viscous(orc_cc* s) {
  gas_t g = gas_of(s);
  if (s->grad_type != 3) {
    grad_ggfc(s);
    if (s->error) return;
    if (s->ilimit == 1) orc_cc_limit(s);
  }
  for (int c = 0; c < s->noc; ++c) s->cmul[c] = sutherland(&g, &s->qc[c * 4]);
  const double two = 2.0, two_third = two / 3.0;
  for (int f = 0; f < nof(s); ++f) {
    double nx = s->fnx[f], ny = s->fny[f], area = s->farea[f];
    int in = s->fin[f], out = s->fout[f], bc = s->fbc[f];
    double mu = out < s->noc ? 0.5 * (s->cmul[in] + s->cmul[out]) : s->cmul[in];
    s->fmu[f] = mu;
    double uvel = s->fqp[f * 5 + 1], vvel = s->fqp[f * 5 + 2];
    const double* fg = &s->fgrad[f * 10];
    double ux = fg[2], uy = fg[3], vx = fg[4], vy = fg[5], tx = fg[8], ty = fg[9];
    double tauxx = two_third * (two * ux - vy);
    double tauxy = uy + vx;
    double tauyy = two_third * (two * vy - ux);
    double factor = (g.gamma - 1.0) * g.m_inf * g.m_inf * 0.72;
    double qx = -(tx / factor), qy = -(ty / factor);
    double mr = mu / s->reynolds;
    double F[4];
    F[0] = mr * ((uvel * tauxx + vvel * tauxy - qx) * nx + (uvel * tauxy + vvel * tauyy - qy) * ny);
    F[1] = 0.0;
    F[2] = mr * (tauxx * nx + tauxy * ny);
    F[3] = mr * (tauxy * nx + tauyy * ny);
    for (int i = 0; i < 4; ++i) s->res[in * 4 + i] = s->res[in * 4 + i] - F[i] * area;
    if (out < s->noc)
      for (int i = 0; i < 4; ++i) s->res[out * 4 + i] = s->res[out * 4 + i] + F[i] * area;
    if (bc == 1001) {
      s->fx1 = s->fx1 - F[2] * area;
      s->fy1 = s->fy1 - F[3] * area;
    }
  }
}

/* cirs, CCFVM/ImpResiSmoothening.F90:1-47 */
static void cirs(orc_cc* s) {
  int noc = s->noc;
  double r = s->cfl / s->cfl_max;
  double eps = dmax(0.25 * (r * r - 1.0), 0.0);
  double* old = malloc(sizeof(double) * 4 * noc);
  double* itr = malloc(sizeof(double) * 4 * noc);
  memcpy(old, s->res, sizeof(double) * 4 * noc);
  for (int it = 0; it < 2; ++it) {
    memset(itr, 0, sizeof(double) * 4 * noc);
    for (int f = 0; f < s->nfi; ++f) {
      int in = s->fin[f], out = s->fout[f];
      for (int i = 0; i < 4; ++i) {
        itr[in * 4 + i] = itr[in * 4 + i] + s->res[out * 4 + i];
        itr[out * 4 + i] = itr[out * 4 + i] + s->res[in * 4 + i];
      }
    }
    for (int c = 0; c < noc; ++c) {
      double den = 1.0 / (1.0 + eps * (double)s->nc2c[c]);
      for (int i = 0; i < 4; ++i) s->res[c * 4 + i] = (itr[c * 4 + i] * eps + old[c * 4 + i]) * den;
    }
  }
  free(old);
  free(itr);
}

/* One face flux of fvresidual.  The reference STOPs inside roe_flux when the
 * Roe-averaged sound speed squared is not positive, printing the two states it
 * was given (CCFVM/roe_flux.F90:73-87); the first such face in fvresidual's
 * loop order (boundary faces, then interior faces) is recorded with those
 * states, 1-based. */
static void face_flux(orc_cc* s, const gas_t* g, double* qcl, double* qcr, int f, int in,
                      double* F) {
  double l[4], r[4];
  memcpy(l, qcl, sizeof l);
  memcpy(r, qcr, sizeof r);
  if (num_flux(s, g, qcl, qcr, f, F) && !s->error) {
    s->error = 1;
    s->error_cell = in + 1;
    s->error_face = f + 1;
    memcpy(s->error_ql, l, sizeof l);
    memcpy(s->error_qr, r, sizeof r);
  }
}

/* fvresidual, CCFVM/fvresidual.F90:4-187 */
void orc_cc_fvresidual(orc_cc* s) {
  gas_t g = gas_of(s);
  for (int i = 0; i < 4 * s->noc; ++i) s->res[i] = 0.0;
  s->fx1 = 0.0;
  s->fy1 = 0.0;
  for (int f = s->nfi; f < nof(s); ++f) {
    int in = s->fin[f], out = s->fout[f], bc = s->fbc[f];
    double qcl[4] = {0, 0, 0, 0}, qcr[4] = {0, 0, 0, 0}, F[4];
    if (bc != 2001 && bc != 1001) continue;
    double dx = s->fcx[f] - s->ccx[in], dy = s->fcy[f] - s->ccy[in];
    for (int i = 0; i < 4; ++i)
      qcl[i] = s->qp[in * 5 + i] +
               s->phi[in * 4 + i] * (s->grad[in * 10 + 2 * i] * dx + s->grad[in * 10 + 2 * i + 1] * dy);
    if (bc == 2001) {
      farfield(&g, qcl, s->fnx[f], s->fny[f], qcr);
      face_flux(s, &g, qcl, qcr, f, in, F);
      for (int i = 0; i < 4; ++i) s->res[in * 4 + i] = s->res[in * 4 + i] + F[i];
      double conv[4];
      prim2con(&g, qcr, conv);
      memcpy(&s->qc[out * 4], conv, sizeof conv);
      con2prim(&g, conv, &s->qp[out * 5]);
    } else {
      const double* q = &s->qc[in * 4];
      double* gq = &s->qc[out * 4];
      double* gp = &s->qp[out * 5];
      if (s->flow_type != 1) {
        if (s->iwall == 0) {
          double p[5];
          con2prim(&g, q, p);
          double tg = 2.0 * g.t_wall - p[4];
          if (tg <= 0.0) tg = 0.01 * g.t_wall;
          gp[0] = p[3] / (p[0] * g.znd);
          gp[1] = -p[1];
          gp[2] = -p[2];
          gp[3] = p[3];
          gp[4] = tg;
          prim2con(&g, gp, gq);
        } else {
          gq[0] = q[0];
          gq[1] = q[1];
          gq[2] = -q[2];
          gq[3] = -q[3];
          con2prim(&g, gq, gp);
        }
      } else {
        double rho = q[1], u = q[2] / q[1], v = q[3] / q[1];
        double un = u * s->fnx[f] + v * s->fny[f];
        double t[4] = {q[0], q[1], rho * u - 2.0 * un * s->fnx[f], rho * v - 2.0 * un * s->fny[f]};
        memcpy(gq, t, sizeof t);
        con2prim(&g, gq, gp);
      }
      for (int i = 0; i < 4; ++i) qcr[i] = gp[i];
      face_flux(s, &g, qcl, qcr, f, in, F);
      for (int i = 0; i < 4; ++i) s->res[in * 4 + i] = s->res[in * 4 + i] + F[i];
      s->fx1 = s->fx1 + F[2];
      s->fy1 = s->fy1 + F[3];
    }
  }
  for (int f = 0; f < s->nfi; ++f) {
    int in = s->fin[f], out = s->fout[f];
    double qcl[4], qcr[4], F[4];
    recon(s, f, in, out, qcl, qcr);
    face_flux(s, &g, qcl, qcr, f, in, F);
    for (int i = 0; i < 4; ++i) {
      s->res[in * 4 + i] = s->res[in * 4 + i] + F[i];
      s->res[out * 4 + i] = s->res[out * 4 + i] - F[i];
    }
  }
  if (s->flow_type != 1) viscous(s);
  if (s->irs == 1) cirs(s);
}

/* CCFVM/main.F90:92-101 */
void orc_cc_rk_update(orc_cc* s, int stage) {
  static const double a[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
  static const double b[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  double ak = a[stage - 1], bk = b[stage - 1];
  for (int c = 0; c < s->noc; ++c) {
    double dtbyarea = s->dt[c] / s->ccv[c];
    for (int j = 0; j < 4; ++j)
      s->qc[c * 4 + j] = ak * s->qold[c * 4 + j] + bk * (s->qc[c * 4 + j] - dtbyarea * s->res[c * 4 + j]);
  }
}

/* normal_flux, CCFVM/lusgs.F90:328-359 */
static void normal_flux(const gas_t* g, const double* qold, const double* dq, double nx, double ny,
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

/* lusgs2, CCFVM/lusgs.F90:207-379 */
void orc_cc_lusgs(orc_cc* s) {
  gas_t g = gas_of(s);
  const double omega = 1.5;
  int noc = s->noc;
  orc_cc_fvresidual(s);
  double* diag = malloc(sizeof(double) * noc);
  double* lower = calloc((size_t)4 * noc, sizeof(double));
  double* upper = calloc((size_t)4 * noc, sizeof(double));
  for (int c = 0; c < noc; ++c) diag[c] = s->ccv[c] / s->dt[c] + 0.5 * omega * s->la[c];
  for (int it = 0; it < noc; ++it) {
    double cres[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) lower[it * 4 + k] = 0.0;
    for (int j = s->c2f_ptr[it]; j < s->c2f_ptr[it + 1]; ++j) {
      int f = s->c2f[j];
      double nx = s->fnx[f], ny = s->fny[f];
      int in = s->fin[f], out = s->fout[f];
      if (out == it) { out = in; nx = -s->fnx[f]; ny = -s->fny[f]; }
      if (out < it && out >= 0 && out < noc) {
        double dflux[4];
        normal_flux(&g, &s->qold[out * 4], &lower[out * 4], nx, ny, s->farea[f], dflux);
        double lam = s->fla[f];
        for (int k = 0; k < 4; ++k) cres[k] = cres[k] + (dflux[k] - omega * lam * lower[out * 4 + k]);
      }
    }
    for (int k = 0; k < 4; ++k) lower[it * 4 + k] = (-s->res[it * 4 + k] - 0.5 * cres[k]) / diag[it];
  }
  for (int it = noc - 1; it >= 0; --it) {
    double cres[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) upper[it * 4 + k] = 0.0;
    for (int j = s->c2f_ptr[it]; j < s->c2f_ptr[it + 1]; ++j) {
      int f = s->c2f[j];
      double nx = s->fnx[f], ny = s->fny[f];
      int in = s->fin[f], out = s->fout[f];
      if (out == it) { out = in; nx = -s->fnx[f]; ny = -s->fny[f]; }
      if (out > it && out < noc) {
        double dflux[4];
        normal_flux(&g, &s->qold[out * 4], &upper[out * 4], nx, ny, s->farea[f], dflux);
        double lam = s->fla[f];
        for (int k = 0; k < 4; ++k) cres[k] = cres[k] + (dflux[k] - omega * lam * upper[out * 4 + k]);
      }
    }
    for (int k = 0; k < 4; ++k) upper[it * 4 + k] = lower[it * 4 + k] - 0.5 * cres[k] / diag[it];
  }
  for (int c = 0; c < noc; ++c)
    for (int k = 0; k < 4; ++k) s->qc[c * 4 + k] = s->qold[c * 4 + k] + upper[c * 4 + k];
  free(diag);
  free(lower);
  free(upper);
}

/* lusgs2 with the cells swept in the order `order` (order[k] = the k-th
 * cell) instead of their numbering: the forward loop visits them in that
 * order and takes as already swept the neighbours earlier in it, the reverse
 * loop the other way round.  The identity order gives orc_cc_lusgs; cells
 * grouped by colour give the multicolour sweep. */
void orc_cc_lusgs_ordered(orc_cc* s, const int* order) {
  gas_t g = gas_of(s);
  const double omega = 1.5;
  int noc = s->noc;
  orc_cc_fvresidual(s);
  double* diag = malloc(sizeof(double) * noc);
  double* lower = calloc((size_t)4 * noc, sizeof(double));
  double* upper = calloc((size_t)4 * noc, sizeof(double));
  int* rank = malloc(sizeof(int) * noc);
  for (int k = 0; k < noc; ++k) rank[order[k]] = k;
  for (int c = 0; c < noc; ++c) diag[c] = s->ccv[c] / s->dt[c] + 0.5 * omega * s->la[c];
  for (int pass = 0; pass < 2; ++pass) {
    double* w = pass == 0 ? lower : upper;
    for (int step = 0; step < noc; ++step) {
      int it = order[pass == 0 ? step : noc - 1 - step];
      double cres[4] = {0, 0, 0, 0};
      for (int k = 0; k < 4; ++k) w[it * 4 + k] = 0.0;
      for (int j = s->c2f_ptr[it]; j < s->c2f_ptr[it + 1]; ++j) {
        int f = s->c2f[j];
        double nx = s->fnx[f], ny = s->fny[f];
        int in = s->fin[f], out = s->fout[f];
        if (out == it) { out = in; nx = -s->fnx[f]; ny = -s->fny[f]; }
        if (out >= 0 && out < noc && (pass == 0 ? rank[out] < rank[it] : rank[out] > rank[it])) {
          double dflux[4];
          normal_flux(&g, &s->qold[out * 4], &w[out * 4], nx, ny, s->farea[f], dflux);
          double lam = s->fla[f];
          for (int k = 0; k < 4; ++k) cres[k] = cres[k] + (dflux[k] - omega * lam * w[out * 4 + k]);
        }
      }
      for (int k = 0; k < 4; ++k)
        w[it * 4 + k] = pass == 0 ? (-s->res[it * 4 + k] - 0.5 * cres[k]) / diag[it]
                                  : lower[it * 4 + k] - 0.5 * cres[k] / diag[it];
    }
  }
  for (int c = 0; c < noc; ++c)
    for (int k = 0; k < 4; ++k) s->qc[c * 4 + k] = s->qold[c * 4 + k] + upper[c * 4 + k];
  free(rank);
  free(diag);
  free(lower);
  free(upper);
}

/* residue, CCFVM/solution.F90:136-180 */
void orc_cc_residue(orc_cc* s, int iter) {
  double f = 0.0;
  for (int c = 0; c < s->noc; ++c)
    for (int j = 0; j < 4; ++j) f = f + s->res[c * 4 + j] * s->res[c * 4 + j];
  f = sqrt(f / s->noc / 4);
  if (iter == 1) s->fres1 = f;
  if (s->fres1 >= 2.220446049250313e-16) f = f / s->fres1;
  s->fres = f;
}

/* check_positivity, CCFVM/main.F90:358-414: 1-based first bad cell or 0 */
int orc_cc_check_positivity(const orc_cc* s) {
  for (int c = 0; c < s->noc; ++c)
    if (s->qp[c * 5] <= 0.0 || s->qp[c * 5 + 3] <= 0.0) return c + 1;
  return 0;
}

/* find_cl_cd, CCFVM/main.F90:263-355 (contained in screen).  Wall faces
 * (bc 1001) in boundary-face order: the wall pressure is the adjacent cell's
 * qp(4) (:302); Cp = (p - p_inf)/q_inf with q_inf = 0.5 gamma p_inf M^2
 * (:288); the force sums p n area (:316-317) and, for viscous flow, minus the
 * face stress with mu = fc%mu/Rey and fc%grad of the last viscous pass
 * (:319-332); Cf = tau_w area / q_inf (:334-337); Cm sums Cp n area's moment
 * about (xref, yref) at the face midpoint (:340).  out[0..6] = cl, cd of the
 * surface integral (:349-350, printed "Cl, Cd"), cl1, cd1 of the flux sums
 * Fx1, Fy1 of the last residual (:352-353, what the screen block then prints
 * as Cl, Cd), cm, fx, fy.  xw, yw, mcp, cf (any may be NULL) receive the rows
 * of XYCpCf.dat (:343): midpoint, -Cp, Cf.  Returns the wall face count. */
int orc_cc_find_cl_cd(const orc_cc* s, double xref, double yref, double* out, double* xw,
                      double* yw, double* mcp, double* cf) {
  gas_t g = gas_of(s);
  const double p_inf = g.p_inf, m_inf = s->mach;
  const double dyna = 0.5 * g.gamma * p_inf * m_inf * m_inf;
  const double two = 2.0, two_third = two / 3.0;
  const double aoa = g.aoa;
  double fx = 0.0, fy = 0.0, cm = 0.0;
  int n = 0;
  for (int f = s->nfi; f < nof(s); ++f) {
    if (s->fbc[f] != 1001) continue;
    double pwall = s->qp[s->fin[f] * 5 + 3];
    int p1 = s->fp1[f], p2 = s->fp2[f];
    double xi = 0.5 * (s->px[p1] + s->px[p2]);
    double yi = 0.5 * (s->py[p1] + s->py[p2]);
    double xa = xi - xref, ya = yi - yref;
    double area = s->farea[f], nx = s->fnx[f], ny = s->fny[f];
    double cp = (pwall - p_inf) / dyna;
    double dcx = cp * nx * area, dcy = cp * ny * area;
    fx = fx + pwall * nx * area;
    fy = fy + pwall * ny * area;
    double cfv = 0.0;
    if (s->flow_type != 1) {
      double mu = s->fmu[f] / s->reynolds;
      const double* fg = &s->fgrad[f * 10];
      double ux = fg[2], uy = fg[3], vx = fg[4], vy = fg[5];
      double tauxx = two_third * (two * ux - vy);
      double tauxy = uy + vx;
      double tauyy = two_third * (two * vy - ux);
      fx = fx - mu * (tauxx * nx + tauxy * ny) * area;
      fy = fy - mu * (tauxy * nx + tauyy * ny) * area;
      double tx = ny, ty = -nx;
      double tw = mu * ((ux * tx + vx * ty) * nx + (uy * tx + vy * ty) * ny);
      cfv = tw * area / dyna;
    }
    cm = cm + dcx * ya - dcy * xa;
    if (xw) xw[n] = xi;
    if (yw) yw[n] = yi;
    if (mcp) mcp[n] = -cp;
    if (cf) cf[n] = cfv;
    ++n;
  }
  out[0] = (fy * cos(aoa) - fx * sin(aoa)) / dyna;
  out[1] = (fy * sin(aoa) + fx * cos(aoa)) / dyna;
  out[2] = (s->fy1 * cos(aoa) - s->fx1 * sin(aoa)) / dyna;
  out[3] = (s->fy1 * sin(aoa) + s->fx1 * cos(aoa)) / dyna;
  out[4] = cm;
  out[5] = fx;
  out[6] = fy;
  return n;
}

/* one iteration of CCFVM/main.F90:59-115 */
void orc_cc_iteration(orc_cc* s, int iter, double cfl, int timemode) {
  s->cfl = cfl;
  orc_cc_prim(s);
  orc_cc_avg_c2v(s);
  orc_cc_time_step(s);
  orc_cc_save_old(s);
  orc_cc_gradient(s);
  if (s->error) return;
  if (timemode == 1) {
    for (int k = 1; k <= 3; ++k) {
      orc_cc_fvresidual(s);
      if (s->error) return;
      orc_cc_rk_update(s, k);
    }
  } else {
    orc_cc_lusgs(s);
  }
  orc_cc_residue(s, iter);
}
