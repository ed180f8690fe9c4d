/*
 * CPU oracle of the vertex-centred (median-dual) hot path -- TEST
 * INFRASTRUCTURE ONLY.
 *
 * A plain C restatement of the NCFVM routines in the reference's own loop
 * structure (edge loops scattering into both end points, in edge order),
 * compiled without FMA contraction.  Only the tests, the smoke check and the
 * benchmark's CPU leg use it, to check results against.  Each function cites the
 * reference routine it restates.
 *
 * Per-vertex records keep the variables of one vertex contiguous:
 * qc[p*4+k] = [E, rho, rho u, rho v], qp[p*4+k] = [rho, u, v, p],
 * grad[p*8 + 2*k + d] (d = 0: x, 1: y), phi/dumin/dumax[p*4+k].
 * Indices are 0-based; vertices are in the reference's breadth-first order.
 */
#ifndef UFX_NC_ORACLE_H
#define UFX_NC_ORACLE_H

typedef struct {
  /* mesh (post-geometric) */
  int nop, nof;
  const double *px, *py, *pcv;
  const int *fp1, *fp2, *fbc;
  const double *fnx, *fny, *farea;
  const int *v2v_ptr, *v2v;
  const double *wx, *wy;
  const int *v2f_ptr, *v2f;
  /* parameters */
  double mach, aoa_deg, cfl_max;
  int grad_kind;               /* 1: Gradient_GG, 2: Gradient_LSQR */
  double cfl;                  /* CFL of the current iteration */
  /* state (caller-owned) */
  double *qc, *qp, *qold, *res;  /* 4*nop */
  double *grad;                  /* 8*nop */
  double *phi, *dumin, *dumax;   /* 4*nop */
  double *la, *dt;               /* nop */
  double *fla;                   /* nof */
  /* outputs */
  double dtglobal, fres, fres1, fresi;
  int iresi;
  /* options: flux 3 van Leer (the reference's), 2 AUSM+-up (build-defined:
   * the cell-centred routine on the dual face); limiter 0 limiter_VKN,
   * 1 limiter_min_max (Barth-Jespersen, dead code in the reference) */
  int flux_kind, lim_kind;
} orc_nc;

void orc_nc_initialize(orc_nc* s);
void orc_nc_save_old(orc_nc* s);
void orc_nc_gradient(orc_nc* s);
void orc_nc_limiter(orc_nc* s);
void orc_nc_fvresidual(orc_nc* s);
void orc_nc_rk_update(orc_nc* s, int irk);
void orc_nc_lusgs(orc_nc* s);
/* lusgs with the vertices swept in a given order (order[k] = k-th vertex);
 * "already swept" neighbours are those earlier in it (see nc_oracle.c). */
void orc_nc_lusgs_ordered(orc_nc* s, const int* order);
void orc_nc_residue(orc_nc* s, int iter);
double orc_nc_cfl(double cfl_max, int iter);
/* One iteration of NCFVM/main.f90:51-88 (timemode 1: rk3, 2: lusgs). */
void orc_nc_iteration(orc_nc* s, int iter, int timemode);
/* Wall forces, build-defined (see nc_oracle.c): out = cl, cd, fx, fy; cp
 * (may be NULL) -Cp at the first end point of each wall edge. */
void orc_nc_forces(orc_nc* s, double* out, double* cp);

#endif
