/*
 * CPU oracle of the vertex-centred hot path -- TEST INFRASTRUCTURE ONLY.
 * See nc_oracle.h.  Expressions keep the reference's operation order
 * (Fortran left-to-right, compiled with -ffp-contract=off) so that the
 * results follow the reference bit for bit.
 */
#include "nc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double GAMMA = 1.4;
#define GAMMA1 (GAMMA - 1.0)
/* epsilon(1.0) of the single-precision kind, NCFVM/All_modules.f90:9 */
static const double EPS_SP = 1.1920928955078125e-07;

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

typedef struct { double r_inf, u_inf, v_inf, p_inf, a_inf; double qinf[4]; } freestream;

/* NCFVM/common.f90:21-58 (initialize) */
static freestream fs_of(const orc_nc* s) {
  freestream f;
  double pi = 4.0 * atan(1.0);
  double q_inf = 1.0;
  double aoa = s->aoa_deg * pi / 180.0;
  f.qinf[1] = 1.0;
  f.qinf[2] = cos(aoa) * q_inf;
  f.qinf[3] = sin(aoa) * q_inf;
  f.p_inf = 1.0 / (GAMMA * s->mach * s->mach);
  f.qinf[0] = f.p_inf / (GAMMA - 1.0) + 0.5;
  f.r_inf = f.qinf[1];
  f.a_inf = sqrt(GAMMA * f.p_inf / f.r_inf);
  f.u_inf = f.qinf[2] / f.qinf[1];
  f.v_inf = f.qinf[3] / f.qinf[1];
  return f;
}

void orc_nc_initialize(orc_nc* s) {
  freestream f = fs_of(s);
  const double phi0 = (double)1e20f;  /* default of pt%phi, NCFVM/All_modules.f90:200 */
  for (int p = 0; p < s->nop; ++p) {
    for (int k = 0; k < 4; ++k) {
      s->qc[p * 4 + k] = f.qinf[k];
      s->qp[p * 4 + k] = 0.0;
      s->qold[p * 4 + k] = 0.0;
      s->res[p * 4 + k] = 0.0;
      s->phi[p * 4 + k] = phi0;
      s->dumin[p * 4 + k] = 0.0;
      s->dumax[p * 4 + k] = 0.0;
    }
    for (int k = 0; k < 8; ++k) s->grad[p * 8 + k] = 0.0;
    s->la[p] = 0.0;
    s->dt[p] = 0.0;
  }
  for (int e = 0; e < s->nof; ++e) s->fla[e] = 0.0;
}

/* con2prim of NCFVM/All_modules.f90:146-169: [rho, u, v, p] and e = Et/rho */
static void con2prim(const double* q, double* prim, double* e) {
  double rho = q[1], u = q[2] / q[1], v = q[3] / q[1], et = q[0];
  double qq = u * u + v * v;
  double p = (GAMMA - 1.0) * (et - 0.5 * rho * qq);
  if (e) *e = et / rho;
  prim[0] = rho; prim[1] = u; prim[2] = v; prim[3] = p;
}

/* NCFVM/common.f90:88-104 */
void orc_nc_save_old(orc_nc* s) {
  for (int p = 0; p < s->nop; ++p) {
    con2prim(s->qc + p * 4, s->qp + p * 4, NULL);
    for (int k = 0; k < 4; ++k) s->qold[p * 4 + k] = s->qc[p * 4 + k];
  }
}

/* Gradient_GG, NCFVM/gradient.f90:35-75 (no boundary closure term) */
static void gradient_gg(orc_nc* s) {
  memset(s->grad, 0, sizeof(double) * 8 * (size_t)s->nop);
  for (int i = 0; i < s->nof; ++i) {
    int p1 = s->fp1[i], p2 = s->fp2[i];
    double ds = s->farea[i], dx = s->fnx[i] * ds, dy = s->fny[i] * ds;
    for (int k = 0; k < 4; ++k) {
      double var = 0.5 * (s->qp[p1 * 4 + k] + s->qp[p2 * 4 + k]);
      s->grad[p1 * 8 + 2 * k] = s->grad[p1 * 8 + 2 * k] + var * dx;
      s->grad[p1 * 8 + 2 * k + 1] = s->grad[p1 * 8 + 2 * k + 1] + var * dy;
      s->grad[p2 * 8 + 2 * k] = s->grad[p2 * 8 + 2 * k] - var * dx;
      s->grad[p2 * 8 + 2 * k + 1] = s->grad[p2 * 8 + 2 * k + 1] - var * dy;
    }
  }
  for (int i = 0; i < s->nop; ++i) {
    double var = s->pcv[i];
    for (int k = 0; k < 8; ++k) s->grad[i * 8 + k] = s->grad[i * 8 + k] / var;
  }
}

/* Gradient_LSQR, NCFVM/gradient.f90:2-30 */
static void gradient_lsqr(orc_nc* s) {
  memset(s->grad, 0, sizeof(double) * 8 * (size_t)s->nop);
  for (int i = 0; i < s->nop; ++i)
    for (int k = 0; k < 4; ++k)
      for (int j = s->v2v_ptr[i]; j < s->v2v_ptr[i + 1]; ++j) {
        int c = s->v2v[j];
        s->grad[i * 8 + 2 * k] =
            s->grad[i * 8 + 2 * k] + (s->qp[c * 4 + k] - s->qp[i * 4 + k]) * s->wx[j];
        s->grad[i * 8 + 2 * k + 1] =
            s->grad[i * 8 + 2 * k + 1] + (s->qp[c * 4 + k] - s->qp[i * 4 + k]) * s->wy[j];
      }
}

void orc_nc_gradient(orc_nc* s) {
  if (s->grad_kind == 2) gradient_lsqr(s);
  else gradient_gg(s);
}

/* limiter_VKN, NCFVM/limiter.f90:15-98: edge min/max of neighbour
 * differences (initial +-eps of single precision), then per edge end the
 * Venkatakrishnan function with eps^2 = (kappa sqrt(V))^3, kappa the single
 * precision 0.01, ratcheted into phi (never reset) */
void orc_nc_limiter(orc_nc* s) {
  for (int i = 0; i < s->nop; ++i)
    for (int j = 0; j < 4; ++j) {
      s->dumin[i * 4 + j] = EPS_SP;
      s->dumax[i * 4 + j] = -EPS_SP;
    }
  for (int i = 0; i < s->nof; ++i) {
    int p1 = s->fp1[i], p2 = s->fp2[i];
    for (int j = 0; j < 4; ++j) {
      double du = s->qp[p2 * 4 + j] - s->qp[p1 * 4 + j];
      s->dumax[p1 * 4 + j] = dmax(s->dumax[p1 * 4 + j], du);
      s->dumin[p1 * 4 + j] = dmin(s->dumin[p1 * 4 + j], du);
      s->dumax[p2 * 4 + j] = dmax(s->dumax[p2 * 4 + j], -du);
      s->dumin[p2 * 4 + j] = dmin(s->dumin[p2 * 4 + j], -du);
    }
  }
  const double kappa = (double)0.01f;
  for (int i = 0; i < s->nof; ++i) {
    int p1 = s->fp1[i], p2 = s->fp2[i];
    double d0 = s->px[p2] - s->px[p1], d1 = s->py[p2] - s->py[p1];
    for (int j = 0; j < 4; ++j) {
      for (int side = 0; side < 2; ++side) {
        int p = side == 0 ? p1 : p2;
        double sum = (0.0 + s->grad[p * 8 + 2 * j] * d0) + s->grad[p * 8 + 2 * j + 1] * d1;
        double phi;
        if (s->lim_kind == 1) {
          /* limiter_min_max, NCFVM/limiter.f90:221-257: +1/2 at both ends */
          double d_l = 0.5 * sum;
          if (d_l > 0.0) phi = dmin(1.0, s->dumax[p * 4 + j] / d_l);
          else if (d_l < 0.0) phi = dmin(1.0, s->dumin[p * 4 + j] / d_l);
          else phi = 1.0;
        } else {
          double d_l = side == 0 ? 0.5 * sum : -0.5 * sum;
          double t = kappa * sqrt(s->pcv[p]);
          double tol = t * t * t;
          double alfa = d_l > 0.0 ? s->dumax[p * 4 + j] : s->dumin[p * 4 + j];
          if (fabs(alfa) < EPS_SP) alfa = 0.0;
          double nr = alfa * alfa + 2.0 * d_l * alfa + tol;
          double dr = alfa * alfa + 2.0 * d_l * d_l + d_l * alfa + tol;
          phi = dmin(1.0, nr / dr);
        }
        s->phi[p * 4 + j] = dmin(s->phi[p * 4 + j], phi);
      }
    }
  }
}

/* split Mach and pressure polynomials, NCFVM/common.f90:247-330 */
static double m2p(double m) { return 0.25 * (m + 1.0) * (m + 1.0); }
static double m2m(double m) { return -0.25 * (m - 1.0) * (m - 1.0); }
static double m4p(double m, double beta) {
  return fabs(m) >= 1.0 ? 0.5 * (m + fabs(m)) : m2p(m) * (1.0 - 16.0 * beta * m2m(m));
}
static double m4m(double m, double beta) {
  return fabs(m) >= 1.0 ? 0.5 * (m - fabs(m)) : m2m(m) * (1.0 + 16.0 * beta * m2p(m));
}
static double p5p(double m, double alpha) {
  return fabs(m) >= 1.0 ? 0.5 * (m + fabs(m)) / m
                        : m2p(m) * ((2.0 - m) - 16.0 * alpha * m * m2m(m));
}
static double p5m(double m, double alpha) {
  return fabs(m) >= 1.0 ? 0.5 * (m - fabs(m)) / m
                        : m2m(m) * ((-2.0 - m) + 16.0 * alpha * m * m2p(m));
}

/* vanleer_flux, NCFVM/vanleer_flux.f90:2-118: van Leer splitting on the
 * dual face of edge ie (unit normal, length area) between the states ql, qr
 * (primitives); res(c1) += F area, res(c2) -= F area; Roe-averaged spectral
 * radius into la of both ends and fla of the edge */
static void vanleer(orc_nc* s, int ie, int c1, int c2, const double* ql, const double* qr) {
  double nx = s->fnx[ie], ny = s->fny[ie], area = s->farea[ie];
  double rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  double unl = ul * nx + vl * ny;
  double ql2 = ul * ul + vl * vl;
  double al2 = GAMMA * pl / rl;
  double hl = al2 / GAMMA1 + 0.5 * ql2;
  double cl = sqrt(al2);
  double ml = unl / cl;
  double rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  double unr = ur * nx + vr * ny;
  double qr2 = ur * ur + vr * vr;
  double ar2 = GAMMA * pr / rr;
  double hr = ar2 / GAMMA1 + 0.5 * qr2;
  double cr = sqrt(ar2);
  double mr = unr / cr;
  double beta = 0.0, alpha = 0.0;
  double plp = p5p(ml, alpha), prm = p5m(mr, alpha);
  double mlp = m4p(ml, beta), mlm = m4m(ml, beta);
  double mrp = m4p(mr, beta), mrm = m4m(mr, beta);
  double p12 = plp * pl + prm * pr;
  double fmass_p = rl * cl * mlp;
  double fmass_m = rr * cr * mrm;
  double fmass12 = fmass_p + fmass_m;
  double delta = -(GAMMA * (cr * mrp * mrm * pr - cl * mlp * mlm * pl) / (GAMMA - 1.0));
  double dissi = fmass_p - fmass_m;
  double fl[4] = {hl, 1.0, ul, vl}, fr[4] = {hr, 1.0, ur, vr};
  double fp[4] = {delta, 0.0, nx * p12, ny * p12};
  for (int i = 0; i < 4; ++i) {
    double flux = 0.5 * fmass12 * (fl[i] + fr[i]) - 0.5 * dissi * (fr[i] - fl[i]) + fp[i];
    s->res[c1 * 4 + i] = s->res[c1 * 4 + i] + flux * area;
    s->res[c2 * 4 + i] = s->res[c2 * 4 + i] - flux * area;
  }
  double rt = sqrt(rr / rl);
  double u = (ul + rt * ur) / (1.0 + rt);
  double v = (vl + rt * vr) / (1.0 + rt);
  double h = (hl + rt * hr) / (1.0 + rt);
  double a = sqrt(GAMMA1 * (h - 0.5 * (u * u + v * v)));
  double un = fabs(u * nx + v * ny);
  double ll = (un + a) * area;
  s->la[c1] = s->la[c1] + ll;
  s->la[c2] = s->la[c2] + ll;
  s->fla[ie] = ll;
}

/* AUSM+-up on a dual face (build-defined): the cell-centred routine
 * (orc_flux kind 2: CCFVM/ausm_up_flux.F90:2-162, unit normal) times the
 * face length, with van Leer's Roe-averaged spectral radius */
int orc_flux(int kind, double mach, double* ql, double* qr, double nx, double ny, double area,
             double* flux);
static void ausm_edge(orc_nc* s, int ie, int c1, int c2, const double* ql, const double* qr) {
  double nx = s->fnx[ie], ny = s->fny[ie], area = s->farea[ie];
  double a[4] = {ql[0], ql[1], ql[2], ql[3]}, b[4] = {qr[0], qr[1], qr[2], qr[3]}, F[4];
  orc_flux(2, s->mach, a, b, nx, ny, 1.0, F);
  for (int i = 0; i < 4; ++i) {
    s->res[c1 * 4 + i] = s->res[c1 * 4 + i] + F[i] * area;
    s->res[c2 * 4 + i] = s->res[c2 * 4 + i] - F[i] * area;
  }
  double rl = ql[0], ul = ql[1], vl = ql[2], pl = ql[3];
  double rr = qr[0], ur = qr[1], vr = qr[2], pr = qr[3];
  double hl = GAMMA * pl / rl / GAMMA1 + 0.5 * (ul * ul + vl * vl);
  double hr = GAMMA * pr / rr / GAMMA1 + 0.5 * (ur * ur + vr * vr);
  double rt = sqrt(rr / rl);
  double u = (ul + rt * ur) / (1.0 + rt);
  double v = (vl + rt * vr) / (1.0 + rt);
  double h = (hl + rt * hr) / (1.0 + rt);
  double aa = sqrt(GAMMA1 * (h - 0.5 * (u * u + v * v)));
  double ll = (fabs(u * nx + v * ny) + aa) * area;
  s->la[c1] = s->la[c1] + ll;
  s->la[c2] = s->la[c2] + ll;
  s->fla[ie] = ll;
}

/* solid_flux, NCFVM/solid_flux.f90:1-43: wall pressure on the half-edge
 * right normals with 3/4-1/4 weighting of the end-point pressures */
static void solid(orc_nc* s, int c1, int c2) {
  double p1 = 0.75 * s->qp[c1 * 4 + 3] + 0.25 * s->qp[c2 * 4 + 3];
  double p2 = 0.25 * s->qp[c1 * 4 + 3] + 0.75 * s->qp[c2 * 4 + 3];
  double nx = 0.5 * (s->py[c2] - s->py[c1]);
  double ny = -0.5 * (s->px[c2] - s->px[c1]);
  double area = sqrt(nx * nx + ny * ny);
  s->la[c1] = s->la[c1] + sqrt(GAMMA * s->qp[c1 * 4 + 3] / s->qp[c1 * 4]) * area;
  s->la[c2] = s->la[c2] + sqrt(GAMMA * s->qp[c2 * 4 + 3] / s->qp[c2 * 4]) * area;
  s->res[c1 * 4 + 2] = s->res[c1 * 4 + 2] + p1 * nx;
  s->res[c1 * 4 + 3] = s->res[c1 * 4 + 3] + p1 * ny;
  s->res[c2 * 4 + 2] = s->res[c2 * 4 + 2] + p2 * nx;
  s->res[c2 * 4 + 3] = s->res[c2 * 4 + 3] + p2 * ny;
}

/* farfield_flux1, NCFVM/farfield_flux.f90:1-95: Riemann-invariant state on
 * the half-edge left normal from the edge-averaged primitives; the flux is
 * added to both end points */
static void farfield(orc_nc* s, const freestream* f, int c1, int c2) {
  double nx = -0.5 * (s->py[c2] - s->py[c1]);
  double ny = 0.5 * (s->px[c2] - s->px[c1]);
  double dr = sqrt(nx * nx + ny * ny);
  nx = nx / dr;
  ny = ny / dr;
  double uinf = f->u_inf, vinf = f->v_inf, pinf = f->p_inf, rinf = f->r_inf, ainf = f->a_inf;
  double un_inf = uinf * nx + vinf * ny;
  double con[4];
  for (int k = 0; k < 4; ++k) con[k] = 0.5 * (s->qp[c1 * 4 + k] + s->qp[c2 * 4 + k]);
  double rho = con[0], u = con[1], v = con[2], p = con[3];
  double a = sqrt(GAMMA * p / rho);
  double un = u * nx + v * ny;
  double rp = un + 2.0 * a / GAMMA1;
  double rm = un_inf - 2.0 * ainf / GAMMA1;
  double unf = 0.5 * (rp + rm);
  double af = 0.25 * (rp - rm) * GAMMA1;
  double uf, vf, sf;
  if (unf > 0.0) {
    uf = u + nx * (unf - un);
    vf = v + ny * (unf - un);
    sf = p / pow(rho, GAMMA);
  } else {
    uf = uinf + nx * (unf - un_inf);
    vf = vinf + ny * (unf - un_inf);
    sf = pinf / pow(rinf, GAMMA);
  }
  double rhof = pow(af * af / GAMMA / sf, 1.0 / GAMMA1);
  double pf = rhof * af * af / GAMMA;
  double ef = pf / GAMMA1 + 0.5 * rhof * (uf * uf + vf * vf);
  double flux[4];
  flux[0] = (ef + pf) * unf;
  flux[1] = rhof * unf;
  flux[2] = flux[1] * uf + pf * nx;
  flux[3] = flux[1] * vf + pf * ny;
  double area = sqrt(nx * nx + ny * ny);
  s->la[c1] = s->la[c1] + (fabs(unf) + sqrt(GAMMA * s->qp[c1 * 4 + 3] / s->qp[c1 * 4])) * area;
  s->la[c2] = s->la[c2] + (fabs(unf) + sqrt(GAMMA * s->qp[c2 * 4 + 3] / s->qp[c2 * 4])) * area;
  for (int i = 0; i < 4; ++i) {
    s->res[c1 * 4 + i] = s->res[c1 * 4 + i] + dr * flux[i];
    s->res[c2 * 4 + i] = s->res[c2 * 4 + i] + dr * flux[i];
  }
}

/* fvresidual, NCFVM/fvresidual.f90:4-85 (Recon = 1), with time_step2
 * (NCFVM/timestep.f90:5-37) at its end */
void orc_nc_fvresidual(orc_nc* s) {
  freestream f = fs_of(s);
  orc_nc_gradient(s);
  memset(s->res, 0, sizeof(double) * 4 * (size_t)s->nop);
  memset(s->la, 0, sizeof(double) * (size_t)s->nop);
  memset(s->fla, 0, sizeof(double) * (size_t)s->nof);
  orc_nc_limiter(s);
  for (int ie = 0; ie < s->nof; ++ie) {
    int p1 = s->fp1[ie], p2 = s->fp2[ie];
    double d0 = s->px[p2] - s->px[p1], d1 = s->py[p2] - s->py[p1];
    double left[4], right[4];
    for (int k = 0; k < 4; ++k) {
      double g = (0.0 + s->grad[p1 * 8 + 2 * k] * d0) + s->grad[p1 * 8 + 2 * k + 1] * d1;
      left[k] = s->qp[p1 * 4 + k] + 0.5 * g * s->phi[p1 * 4 + k];
      g = (0.0 + s->grad[p2 * 8 + 2 * k] * d0) + s->grad[p2 * 8 + 2 * k + 1] * d1;
      right[k] = s->qp[p2 * 4 + k] - 0.5 * g * s->phi[p2 * 4 + k];
    }
    if (s->flux_kind == 2) ausm_edge(s, ie, p1, p2, left, right);
    else vanleer(s, ie, p1, p2, left, right);
    if (s->fbc[ie] == 1001) solid(s, p1, p2);
    if (s->fbc[ie] == 2001) farfield(s, &f, p1, p2);
  }
  s->dtglobal = 1.0e20;
  for (int i = 0; i < s->nop; ++i) {
    s->dt[i] = s->cfl * s->pcv[i] / s->la[i];
    s->dtglobal = dmin(s->dtglobal, s->dt[i]);
  }
}

/* inline RK3 stage update, NCFVM/main.f90:59-70 (coefficients :35-43) */
void orc_nc_rk_update(orc_nc* s, int irk) {
  static const double airk[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
  static const double birk[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  double a = airk[irk - 1], b = birk[irk - 1];
  for (int i = 0; i < s->nop; ++i) {
    double dtbyarea = s->dt[i] / s->pcv[i];
    for (int j = 0; j < 4; ++j)
      s->qc[i * 4 + j] = a * s->qold[i * 4 + j] + b * (s->qc[i * 4 + j] - dtbyarea * s->res[i * 4 + j]);
  }
}

/* normalflux of NCFVM/lusgs.f90:135-155 */
static void normalflux(const double* q, double nx, double ny, double ds, double* flux) {
  double pr[4], e;
  con2prim(q, pr, &e);
  double rho = pr[0], u = pr[1], v = pr[2], p = pr[3];
  double et = e * rho;
  double un = u * nx + v * ny;
  flux[0] = (et + p) * un;
  flux[1] = rho * un;
  flux[2] = flux[1] * u + p * nx;
  flux[3] = flux[1] * v + p * ny;
  for (int i = 0; i < 4; ++i) flux[i] = flux[i] * ds;
}

/* lusgs, NCFVM/lusgs.f90:4-131: fvresidual, then one symmetric
 * Gauss-Seidel pass (omega = 1.5) over the vertices in their numbering */
void orc_nc_lusgs(orc_nc* s) {
  const double omega = 1.5;
  orc_nc_fvresidual(s);
  int n = s->nop;
  double* dd = (double*)malloc(sizeof(double) * (size_t)n);
  double* dqc = (double*)calloc((size_t)n * 4, sizeof(double));
  for (int it = 0; it < n; ++it) dd[it] = s->pcv[it] / s->dt[it] + 0.5 * omega * s->la[it];
  for (int pass = 0; pass < 2; ++pass) {
    for (int step = 0; step < n; ++step) {
      int it = pass == 0 ? step : n - 1 - step;
      double cres[4] = {0.0, 0.0, 0.0, 0.0};
      for (int j = s->v2f_ptr[it]; j < s->v2f_ptr[it + 1]; ++j) {
        int c = s->v2f[j];
        double nx = s->fnx[c], ny = s->fny[c], ds = s->farea[c];
        int p1 = s->fp1[c], p2 = s->fp2[c];
        double sign = 1.0;
        if (p2 == it) { p2 = p1; sign = -1.0; }
        if (pass == 0 ? p2 < it : p2 > it) {
          double f1[4], f2[4];
          normalflux(s->qold + p2 * 4, nx, ny, ds, f1);
          normalflux(s->qc + p2 * 4, nx, ny, ds, f2);
          for (int iv = 0; iv < 4; ++iv)
            cres[iv] = cres[iv] + 0.5 * (sign * (f2[iv] - f1[iv]) - omega * s->fla[c] * dqc[p2 * 4 + iv]);
        }
      }
      for (int iv = 0; iv < 4; ++iv) {
        if (pass == 0) dqc[it * 4 + iv] = (-s->res[it * 4 + iv] - cres[iv]) / dd[it];
        else dqc[it * 4 + iv] = dqc[it * 4 + iv] - cres[iv] / dd[it];
        s->qc[it * 4 + iv] = s->qold[it * 4 + iv] + dqc[it * 4 + iv];
      }
    }
  }
  free(dd);
  free(dqc);
}

/* lusgs, NCFVM/lusgs.f90:4-131, with the vertices swept in the order
 * `order` (order[k] = the k-th vertex) instead of their numbering: the
 * forward loop visits them in that order and takes as already swept the
 * neighbours earlier in it, the reverse loop the other way round.  With the
 * identity order this is orc_nc_lusgs; with the vertices grouped by colour of
 * a colouring it is the multicolour sweep, whose vertices of one colour are
 * independent of each other. */
void orc_nc_lusgs_ordered(orc_nc* s, const int* order) {
  const double omega = 1.5;
  orc_nc_fvresidual(s);
  int n = s->nop;
  double* dd = (double*)malloc(sizeof(double) * (size_t)n);
  double* dqc = (double*)calloc((size_t)n * 4, sizeof(double));
  int* rank = (int*)malloc(sizeof(int) * (size_t)n);
  for (int k = 0; k < n; ++k) rank[order[k]] = k;
  for (int it = 0; it < n; ++it) dd[it] = s->pcv[it] / s->dt[it] + 0.5 * omega * s->la[it];
  for (int pass = 0; pass < 2; ++pass) {
    for (int step = 0; step < n; ++step) {
      int it = order[pass == 0 ? step : n - 1 - step];
      double cres[4] = {0.0, 0.0, 0.0, 0.0};
      for (int j = s->v2f_ptr[it]; j < s->v2f_ptr[it + 1]; ++j) {
        int c = s->v2f[j];
        double nx = s->fnx[c], ny = s->fny[c], ds = s->farea[c];
        int p1 = s->fp1[c], p2 = s->fp2[c];
        double sign = 1.0;
        if (p2 == it) { p2 = p1; sign = -1.0; }
        if (pass == 0 ? rank[p2] < rank[it] : rank[p2] > rank[it]) {
          double f1[4], f2[4];
          normalflux(s->qold + p2 * 4, nx, ny, ds, f1);
          normalflux(s->qc + p2 * 4, nx, ny, ds, f2);
          for (int iv = 0; iv < 4; ++iv)
            cres[iv] = cres[iv] + 0.5 * (sign * (f2[iv] - f1[iv]) - omega * s->fla[c] * dqc[p2 * 4 + iv]);
        }
      }
      for (int iv = 0; iv < 4; ++iv) {
        if (pass == 0) dqc[it * 4 + iv] = (-s->res[it * 4 + iv] - cres[iv]) / dd[it];
        else dqc[it * 4 + iv] = dqc[it * 4 + iv] - cres[iv] / dd[it];
        s->qc[it * 4 + iv] = s->qold[it * 4 + iv] + dqc[it * 4 + iv];
      }
    }
  }
  free(rank);
  free(dd);
  free(dqc);
}

/* residue, NCFVM/common.f90:108-154 */
void orc_nc_residue(orc_nc* s, int iter) {
  s->fres = 0.0;
  s->fresi = 0.0;
  s->iresi = 0;
  for (int i = 0; i < s->nop; ++i) {
    double fr = 0.0;
    for (int j = 0; j < 4; ++j) fr = fr + s->res[i * 4 + j] * s->res[i * 4 + j];
    s->fres = s->fres + fr;
    fr = sqrt(fr);
    if (fr > s->fresi) {
      s->fresi = fr;
      s->iresi = i + 1;
    }
  }
  s->fres = sqrt(s->fres / (double)s->nop);
  if (iter == 1) s->fres1 = s->fres;
  if (s->fres1 != 0.0) s->fres = s->fres / s->fres1;
}

/* NCFVM/main.f90:54 */
double orc_nc_cfl(double cfl_max, int iter) {
  return 0.5 * cfl_max * (1.0 + tanh((double)iter * 7.0 / 500.0 - 3.5));
}

void orc_nc_iteration(orc_nc* s, int iter, int timemode) {
  s->cfl = orc_nc_cfl(s->cfl_max, iter);
  orc_nc_save_old(s);
  if (timemode == 1) {
    for (int irk = 1; irk <= 3; ++irk) {
      orc_nc_fvresidual(s);
      orc_nc_rk_update(s, irk);
    }
  } else {
    orc_nc_lusgs(s);
  }
  orc_nc_residue(s, iter);
}

/* Wall forces (build-defined; the reference's NCFVM has no force routine):
 * the wall closure of solid_flux, NCFVM/solid_flux.f90:20-40, per wall edge
 * in edge order, from the current conserved state; cl/cd as
 * CCFVM/main.F90:347-348 with dyna = 0.5 gamma p_inf M^2. */
void orc_nc_forces(orc_nc* s, double* out, double* cp) {
  freestream f = fs_of(s);
  double dyna = 0.5 * GAMMA * f.p_inf * s->mach * s->mach;
  double fx = 0.0, fy = 0.0;
  int k = 0;
  for (int ie = 0; ie < s->nof; ++ie) {
    if (s->fbc[ie] != 1001) continue;
    int p1 = s->fp1[ie], p2 = s->fp2[ie];
    double a[4], b[4];
    con2prim(s->qc + p1 * 4, a, NULL);
    con2prim(s->qc + p2 * 4, b, NULL);
    double p1w = 0.75 * a[3] + 0.25 * b[3], p2w = 0.25 * a[3] + 0.75 * b[3];
    double nx = 0.5 * (s->py[p2] - s->py[p1]), ny = -0.5 * (s->px[p2] - s->px[p1]);
    fx = fx + (p1w * nx + p2w * nx);
    fy = fy + (p1w * ny + p2w * ny);
    if (cp) cp[k] = -((a[3] - f.p_inf) / dyna);
    ++k;
  }
  double pi = 4.0 * atan(1.0), aoa = s->aoa_deg * pi / 180.0;
  out[0] = (fy * cos(aoa) - fx * sin(aoa)) / dyna;
  out[1] = (fy * sin(aoa) + fx * cos(aoa)) / dyna;
  out[2] = fx;
  out[3] = fy;
}
