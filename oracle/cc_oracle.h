/*
 * CPU oracle of the cell-centred hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C restatement of the reference routines, kept in the reference's
 * own loop structure (face loops that scatter into cells, in face order) so
 * it is independent of the device code's per-cell gathers.  Only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark use it, as the
 * checker.  Each function cites the reference routine it restates.
 *
 * State arrays use the reference's per-cell record layout: variable index
 * fastest (qc[c*4+k], qp[c*5+k], grad[c*10 + 2*k + d]).  Indices are 0-based;
 * cells 0..noc-1 interior, noc..noc+nbf-1 ghosts; faces 0..nfi-1 interior.
 */
#ifndef UFX_CC_ORACLE_H
#define UFX_CC_ORACLE_H

typedef struct {
  /* mesh (post-geometric) */
  int nop, noc, nfi, nbf;
  const double *px, *py;
  const int *fp1, *fp2, *fin, *fout, *fbc;
  const double *fnx, *fny, *farea, *fcx, *fcy;
  const double *ccx, *ccy, *ccv, *ccov, *r11, *r12, *r22;
  const int *nc2c;
  const int *c2f_ptr, *c2f;          /* LU-SGS face order */
  const int *v2c_ptr, *v2c;
  const double *wt;
  /* parameters */
  double mach, aoa_deg, reynolds, t_wall; /* t_wall as in the deck (K) */
  int flow_type, iwall, flux, grad_type, ilimit, irs;
  double cfl, cfl_max;
  /* state (caller-owned, sizes in the header comment) */
  double *qc;      /* 4*ncell */
  double *qp;      /* 5*ncell */
  double *qold;    /* 4*noc */
  double *res;     /* 4*noc */
  double *grad;    /* 10*noc */
  double *phi;     /* 4*noc */
  double *dt, *la; /* noc */
  double *fla;     /* nof */
  double *fmu;     /* nof */
  double *fqp;     /* 5*nof */
  double *fgrad;   /* 10*nof */
  double *cmul;    /* ncell */
  double *pprim;   /* 5*nop */
  double *pmu;     /* nop */
  /* outputs */
  double fx1, fy1, dtglobal, fres, fres1;
  int error;       /* 1: Roe negative sound speed; 2: non-positive area (ggfc) */
  int error_cell;  /* 1-based: the failing face's `in` cell (Roe); the face (ggfc) */
  int error_face;  /* 1-based reference face of the first failure */
  double error_ql[4], error_qr[4]; /* Roe: the states roe_flux received */
} orc_cc;

void orc_cc_initialize(orc_cc* s);
void orc_cc_prim(orc_cc* s);
void orc_cc_avg_c2v(orc_cc* s);
void orc_cc_time_step(orc_cc* s);
void orc_cc_save_old(orc_cc* s);
void orc_cc_gradient(orc_cc* s);
void orc_cc_limit(orc_cc* s);
void orc_cc_fvresidual(orc_cc* s);
void orc_cc_rk_update(orc_cc* s, int stage);
void orc_cc_lusgs(orc_cc* s);
/* lusgs2 with the cells swept in a given order (order[k] = k-th cell);
 * "already swept" neighbours are those earlier in it (see cc_oracle.c). */
void orc_cc_lusgs_ordered(orc_cc* s, const int* order);
void orc_cc_residue(orc_cc* s, int iter);
int orc_cc_check_positivity(const orc_cc* s);
/* find_cl_cd, CCFVM/main.F90:263-355: out[0..6] = cl, cd (surface integral),
 * cl1, cd1 (flux sums), cm, fx, fy; per wall face the XYCpCf.dat row
 * (midpoint x, y, -Cp, Cf; any may be NULL).  Returns the wall face count. */
int orc_cc_find_cl_cd(const orc_cc* s, double xref, double yref, double* out, double* xw,
                      double* yw, double* mcp, double* cf);
/* One iteration of CCFVM/main.F90:59-115 with the given CFL (rk3 or lusgs
 * by s->... timemode argument: 1 rk3, 2 lusgs). */
void orc_cc_iteration(orc_cc* s, int iter, double cfl, int timemode);

/* Single-face kernels for known-answer tests: states are primitives
 * [rho,u,v,p], the flux comes back in conserved order [E,rho,rho u,rho v].
 * kind: 1 Roe, 2 AUSM+-up, 3 van Leer, 4 Rusanov.  Returns 0, or 1 when the
 * Roe average has a non-positive sound speed.  ql/qr are modified exactly as
 * the reference modifies them (Rusanov converts them to conserved). */
int orc_flux(int kind, double mach, double* ql, double* qr, double nx, double ny, double area,
             double* flux);
void orc_farfield_state(double mach, double aoa_deg, const double* ql, double nx, double ny,
                        double* qr);

#endif
