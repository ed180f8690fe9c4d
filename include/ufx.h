/*
 * ufx.h -- C-ABI of the MI355X (gfx950) residual hot path of UNSFlow2D's
 * cell-centred finite-volume solver.
 *
 * The reference's hot path is a set of parameterless Fortran subroutines that
 * read and write module globals (CCFVM/types/grid.F90:7-77 and friends).  This
 * interface replaces them one for one: a context owns the device copy of that
 * state, every entry point named after a reference routine performs that
 * routine on the device, and the state can be injected and read back between
 * calls (ufx_cc_set / ufx_cc_get) so a Fortran host can interleave reference
 * and device routines.  Plain pointers and sizes only; no torch types.
 *
 * Conventions
 *  - Return value: UFX_OK (0) or a positive error code; ufx_last_error() gives
 *    the message (thread-local).  Errors the reference reports with STOP are
 *    returned as codes; device-detected ones (negative Roe sound speed,
 *    non-positive density/pressure) surface at the next synchronising call,
 *    and ufx_cc_error names the face or cell.
 *  - Index arrays passed in are 1-based, in the reference's numbering; arrays
 *    read back with ufx_mesh_get are 1-based as well.
 *  - All calls on a context are asynchronous on the context's own HIP stream
 *    except those documented as synchronising.  A context is not thread-safe.
 *  - Host pointers are owned by the caller and copied; device memory is owned
 *    by the context.
 */
#ifndef UFX_H
#define UFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UFX_OK 0
#define UFX_ERR_ARG 1          /* invalid argument or unknown name */
#define UFX_ERR_MESH 2         /* mesh rejected (the reference STOPs here) */
#define UFX_ERR_HIP 3          /* HIP runtime failure */
#define UFX_ERR_STATE 4        /* device-detected invalid flow state */
#define UFX_ERR_UNSUPPORTED 5  /* option outside what this library implements */

#define UFX_KIND_CC 1
#define UFX_KIND_NC 2

/* flux_type, CCFVM/read_input.F90:93-98 */
#define UFX_FLUX_ROE 1
#define UFX_FLUX_AUSM 2
#define UFX_FLUX_VANLEER 3
#define UFX_FLUX_RUSANOV 4
/* grad_type, CCFVM/read_input.F90:87-91 */
#define UFX_GRAD_LSQR 1
#define UFX_GRAD_GG 2
#define UFX_GRAD_GGFC 3
/* flow_type, CCFVM/read_input.F90:61-66 */
#define UFX_FLOW_INVISCID 1
#define UFX_FLOW_LAMINAR 2
/* timemode, CCFVM/read_input.F90:74-79 */
#define UFX_TIME_RK3 1
#define UFX_TIME_LUSGS 2
/* arithmetic of the bulk flux/gradient/limiter work */
#define UFX_PREC_FP64 0
#define UFX_PREC_MIXED 1

/* Vertex (cell) order of the LU-SGS sweeps.  REFERENCE: the reference's own
 * numbering (NCFVM/lusgs.f90:46-122, CCFVM/lusgs.F90:245-310), as a level
 * schedule; the iterate is the reference's.  MULTICOLOUR: the vertices
 * (cells) grouped by the colours of a greedy distance-1 colouring in
 * smallest-last order (ufx_mesh_get "pt_colour" / "cell_colour"), each colour
 * one parallel step,
 * a few launches per sweep instead of thousands of dependent levels.  The
 * same LU-SGS under another ordering: another iterate with the same fixed
 * points (a converged reference-order state stays put), but the limiters'
 * never-reset ratchet makes the state a run converges to depend on its path,
 * as any change of CFL schedule does. */
#define UFX_LUSGS_REFERENCE 0
#define UFX_LUSGS_MULTICOLOUR 1

typedef struct ufx_mesh ufx_mesh;
typedef struct ufx_cc ufx_cc;

/* Message of the last failing call on this thread ("" if none). */
const char* ufx_last_error(void);

/* ------------------------------------------------------------------ mesh */

/* Preprocess a geometry.inp mesh into the reference's post-`geometric` state.
 * Replaces read_grid's ghost mirror (CCFVM/read_input.F90:179-203) and
 * geometric (CCFVM/geometric.F90:1-57): face normals and orientation,
 * breadth-first cell renumbering, faces-of-cell order, vertex->cell lists,
 * co-volumes, least-squares factors and pseudo-Laplacian weights.
 * Inputs follow CCFVM/read_input.F90:160-173: nop vertices (px, py), noc cells
 * (centroid ccx, ccy and area cv, used as given), nfi interior faces
 * (fi_p1, fi_p2, fi_in, fi_out; fi_bc may be NULL) and nbf boundary faces
 * (fb_p1, fb_p2, fb_in, fb_nx, fb_ny, fb_bc). */
int ufx_cc_geometric(int nop, int noc, int nfi, int nbf,
                     const double* px, const double* py,
                     const double* ccx, const double* ccy, const double* cv,
                     const int* fi_p1, const int* fi_p2, const int* fi_in,
                     const int* fi_out, const int* fi_bc,
                     const int* fb_p1, const int* fb_p2, const int* fb_in,
                     const double* fb_nx, const double* fb_ny, const int* fb_bc,
                     ufx_mesh** mesh);

/* The same preprocessing with its per-face, per-cell and per-vertex passes
 * on the current GPU (ghost centres and cell_face_norm, face co-volumes and
 * their cell sums, LSQR_Coeff, pseudo_Laplacian; CCFVM/geometric.F90:62-198,
 * 619-674, 898-1013); the connectivity and the breadth-first numbering
 * (:200-616) run on the host between them.  Every array equals
 * ufx_cc_geometric's bit for bit.  Allocates and frees its own device
 * memory; UFX_ERR_HIP when a HIP call fails. */
int ufx_cc_geometric_device(int nop, int noc, int nfi, int nbf,
                            const double* px, const double* py,
                            const double* ccx, const double* ccy, const double* cv,
                            const int* fi_p1, const int* fi_p2, const int* fi_in,
                            const int* fi_out, const int* fi_bc,
                            const int* fb_p1, const int* fb_p2, const int* fb_in,
                            const double* fb_nx, const double* fb_ny, const int* fb_bc,
                            ufx_mesh** mesh);

/* Build a mesh directly from arrays the reference holds after its own
 * `geometric` (CCFVM/types/grid.F90 fields fc%pt,in,out,bc,nx,ny,area,cen,cov;
 * cell%cen,cv,cov,r11,r12,r22,nc2c,c2f; pt%x,y,v2c,wt).  nof = nfi + nbf
 * faces, ncell = noc + nbf cells (ghosts last).  c2f_ptr has noc+1 entries
 * (0-based offsets), v2c_ptr nop+1.  Lets a Fortran host hand over the
 * reference's own preprocessing unchanged. */
int ufx_cc_mesh_import(int nop, int noc, int nfi, int nbf,
                       const double* px, const double* py,
                       const int* fc_p1, const int* fc_p2, const int* fc_in,
                       const int* fc_out, const int* fc_bc,
                       const double* fc_nx, const double* fc_ny, const double* fc_area,
                       const double* fc_cenx, const double* fc_ceny, const double* fc_cov,
                       const double* cell_cenx, const double* cell_ceny,
                       const double* cell_cv, const double* cell_cov,
                       const double* cell_r11, const double* cell_r12, const double* cell_r22,
                       const int* cell_nc2c, const int* c2f_ptr, const int* c2f,
                       const int* v2c_ptr, const int* v2c, const double* v2c_wt,
                       ufx_mesh** mesh);

/* Preprocess a vertex-centred geometry.inp mesh (NCFVM/read_input.f90:101-112:
 * nop vertices px, py with tag pbc, noc cells with centroid and area, nof
 * edges p1 p2 in out bc, 1-based, 0 for a missing cell) into the reference's
 * post-`geometric` state (NCFVM/geometric.f90:1-46): breadth-first vertex
 * numbering, dual-face normals and lengths, median-dual volumes, vertex
 * neighbour and edge lists, least-squares weights.  Fails like the
 * reference's STOPs (disconnected mesh, non-positive face co-volume).
 * Arrays read back with ufx_mesh_get: "pt_x","pt_y","pt_cv","pt_bc",
 * "pt_nv2v","pt_v2v","pt_wx","pt_wy","pt_nv2f","pt_v2f","oldnum","fc_p1",
 * "fc_p2","fc_in","fc_out","fc_bc","fc_nx","fc_ny","fc_area","fc_cenx",
 * "fc_ceny","cell_cenx","cell_ceny","cell_cv","counts" (nop,noc,nof),
 * "pt_colour". */
int ufx_nc_geometric(int nop, int noc, int nof, const double* px, const double* py,
                     const int* pbc, const double* ccx, const double* ccy, const double* cv,
                     const int* e_p1, const int* e_p2, const int* e_in, const int* e_out,
                     const int* e_bc, ufx_mesh** mesh);

/* The same preprocessing with its per-edge and per-vertex passes on the
 * current GPU (dual-face normals and lengths, the co-volume check, median-
 * dual volumes, least-squares weights; NCFVM/geometric.f90:198-313,
 * 668-804); the neighbour lists and the breadth-first numbering (:52-180,
 * 347-379) run on the host first.  Every array equals ufx_nc_geometric's bit
 * for bit.  Allocates and frees its own device memory; UFX_ERR_HIP when a
 * HIP call fails. */
int ufx_nc_geometric_device(int nop, int noc, int nof, const double* px, const double* py,
                            const int* pbc, const double* ccx, const double* ccy,
                            const double* cv, const int* e_p1, const int* e_p2,
                            const int* e_in, const int* e_out, const int* e_bc,
                            ufx_mesh** mesh);

/* Element count of a named mesh array (see ufx_mesh_get). */
int ufx_mesh_size(const ufx_mesh* mesh, const char* name, int64_t* count);

/* Copy a named mesh array out.  int32 arrays: "fc_p1","fc_p2","fc_in",
 * "fc_out","fc_bc","cell_nc2c","cell_nc2f","cell_c2f","pt_nv2c","pt_v2c",
 * "cell_nc2v","cell_c2v" (the vertex walk round each cell of
 * CCFVM/geometric.F90:546-613, closed with its first vertex; empty for an
 * imported mesh), "oldnum","counts" (nop,noc,nof,nbf); float64 arrays: "fc_nx","fc_ny",
 * "fc_area","fc_cenx","fc_ceny","fc_cov","cell_cenx","cell_ceny","cell_cv",
 * "cell_cov","cell_r11","cell_r12","cell_r22","pt_x","pt_y","pt_wt".
 * Indices are returned 1-based.  Also (int32, colour numbers from 0) the
 * colouring of the multicolour LU-SGS sweeps: "cell_colour" (interior cells
 * over the interior faces) of a cell-centred mesh, "pt_colour" (vertices over
 * the edges) of a vertex-centred one. */
int ufx_mesh_get(const ufx_mesh* mesh, const char* name, void* dst);

void ufx_mesh_destroy(ufx_mesh* mesh);

/* --------------------------------------------------------- solver context */

/* Run parameters of the input deck (CCFVM/read_input.F90:31-50). */
typedef struct {
  double mach;         /* Mach */
  double aoa_deg;      /* AOA in degrees */
  double reynolds;     /* Reynolds (laminar only) */
  double t_wall;       /* FlowType wall temperature (isothermal wall, K) */
  int flow_type;       /* UFX_FLOW_* */
  int iwall;           /* 0 isothermal, 1 adiabatic no-slip wall */
  int flux;            /* UFX_FLUX_* */
  int grad;            /* UFX_GRAD_* */
  int ilimit;          /* 1: Venkatakrishnan limiter after the gradient */
  int irs;             /* 1: central implicit residual smoothing */
  int timemode;        /* UFX_TIME_* */
  int precision;       /* UFX_PREC_* */
  double cfl_max;      /* needed by residual smoothing (CCFVM/ImpResiSmoothening.F90:17) */
  double xref, yref;   /* moment reference point (FarBC line) */
  int lusgs_order;     /* UFX_LUSGS_REFERENCE or UFX_LUSGS_MULTICOLOUR (timemode LU-SGS) */
} ufx_cc_params;

/* Create a context on the current HIP device: uploads the mesh in SoA layout
 * and allocates the solution state.  The state starts as the reference's
 * `initialize` leaves it (CCFVM/solution.F90:18-84): free stream in every
 * cell's qc including ghosts, limiter value 1, ghost primitives 0. */
int ufx_cc_create(const ufx_mesh* mesh, const ufx_cc_params* params, ufx_cc** ctx);
int ufx_cc_destroy(ufx_cc* ctx);

/* Inject / read back a named state array (host <-> device, synchronising).
 * Layout is cell-major with the variables of one cell contiguous, exactly as
 * the reference's per-cell records hold them:
 *   "qc"   4 x ncell  conserved [E, rho, rho u, rho v] (ghosts last)
 *   "qp"   5 x ncell  primitives [rho, u, v, p, T]
 *   "qold" 4 x noc    "res" 4 x noc   "phi" 4 x noc
 *   "grad" 10 x noc   (d/dx, d/dy) per variable, variable-major
 *   "dt","la" noc;  "fc_la" nof;  "pt_prim" 5 x nop;  "pt_mu" nop (get only);  "fc_qp" 5 x nof
 *   "fc_mu" nof;  "cell_mul" ncell;  "fc_grad" 10 x nof (viscous)
 * Not every name is writable ("fc_qp" is derived). */
int ufx_cc_set(ufx_cc* ctx, const char* name, const double* src);
int ufx_cc_get(ufx_cc* ctx, const char* name, double* dst);

/* ---- the routines of the hot path, one entry per reference routine ---- */

/* Interior primitive refresh, CCFVM/main.F90:62-67 (con2prim, types/pri.F90:13-37). */
int ufx_cc_prim(ufx_cc* ctx);
/* Vertex and face averages of the primitives, CCFVM/average.F90:1-39. */
int ufx_cc_avg_c2v(ufx_cc* ctx);
/* Face spectral radii and local time step, CCFVM/timestep.F90:98-177. */
int ufx_cc_time_step(ufx_cc* ctx, double cfl);
/* qold = qc, CCFVM/solution.F90:118-132. */
int ufx_cc_save_old(ufx_cc* ctx);
/* Gradient of the primitives by the context's method (CCFVM/gradient.F90:
 * LSQR :2-93, GG :95-156, GG_FC :159-323), followed by the limiter when
 * ilimit is set, as the reference does inside each gradient routine. */
int ufx_cc_gradient(ufx_cc* ctx);
/* Venkatakrishnan limiter with the ratchet, CCFVM/limiter.F90:3-17,123-296. */
int ufx_cc_limit(ufx_cc* ctx);
/* Finite-volume residual, CCFVM/fvresidual.F90:4-187 (boundary faces first,
 * then interior faces, then viscous flux and residual smoothing). */
int ufx_cc_fvresidual(ufx_cc* ctx);
/* Runge-Kutta stage update of CCFVM/main.F90:92-101 for stage 1..3.  Like
 * ufx_cc_rk_stage, ufx_cc_lusgs and ufx_cc_get("qold") it reads qold, and
 * returns UFX_ERR_STATE after whole inviscid RK3 iterations (which keep their
 * start state elsewhere and leave qold as it was) until ufx_cc_save_old or
 * ufx_cc_set("qold") writes it again. */
int ufx_cc_rk_update(ufx_cc* ctx, int stage);
/* fvresidual followed by the stage update in one pass (same results). */
int ufx_cc_rk_stage(ufx_cc* ctx, int stage);
/* Matrix-free LU-SGS step, CCFVM/lusgs.F90:5-10,207-379 (includes its
 * fvresidual call). */
int ufx_cc_lusgs(ufx_cc* ctx);
/* L2 residual norm, CCFVM/solution.F90:136-180; synchronising.  `iter` == 1
 * stores the normalisation as the reference does; *fres receives the
 * normalised value and *fres_raw the unnormalised one (either may be NULL). */
int ufx_cc_residue(ufx_cc* ctx, int iter, double* fres, double* fres_raw);
/* Positivity check of the interior primitives, CCFVM/main.F90:358-414;
 * synchronising.  *bad_cell is 0 when all cells pass, else the 1-based
 * index of the first failing cell (and UFX_ERR_STATE is returned). */
int ufx_cc_check_positivity(ufx_cc* ctx, int* bad_cell);
/* Force coefficients of CCFVM/main.F90:263-355 (find_cl_cd); synchronising.
 * out[0..7] = cl, cd (pressure integral), cl1, cd1 (flux-based, what the
 * reference prints), cm, fx, fy, dtglobal.  Optional cp (nwall) receives -Cp
 * per wall face in boundary order; cf likewise (NULL to skip). */
int ufx_cc_forces(ufx_cc* ctx, double* out, double* cp, double* cf);

/* One or more whole iterations of CCFVM/main.F90:57-128 on the device:
 * prim refresh, avg_c2v, time step, save_old, gradient (+limiter), then three
 * RK stages or one LU-SGS step, then the residual norm.  `iter0` is the
 * reference iteration counter before the first one; cfl[k] is the CFL of
 * iteration iter0+1+k (the host evaluates the reference's ramp).  hist, if
 * not NULL, receives the normalised residual of each iteration; that makes the
 * call synchronising.  Without hist no host synchronisation happens.
 * Inviscid RK3 iterations without residual smoothing keep the iteration's
 * start state in "qc" through the stages (the stage states live in a
 * separate buffer) instead of copying it to "qold": after them "qold" still
 * holds what the last save_old (or LU-SGS / viscous iteration) stored.
 * They store no "la" / "dt" either: those of the last time step are formed
 * from the face radii when get, set, forces or LU-SGS read them (the same
 * values bit for bit). */
int ufx_cc_iterate(ufx_cc* ctx, int iter0, int niter, const double* cfl, double* hist);

/* The same iterations replayed from a HIP graph (no reference counterpart;
 * for batched runs and serving loops, where the launch cost of small meshes
 * dominates).  Iterations iter0+1 .. iter0+niter with one constant CFL, from
 * iteration 2 on (iter0 >= 1: iteration 1 fixes the residual normalisation;
 * run it with ufx_cc_iterate).  The chunk is captured on the first call and
 * replayed while niter and cfl stay the same; a call with fewer iterations
 * than the captured chunk (the remainder of a run) launches them directly,
 * one with more captures anew.  hist as ufx_cc_iterate. */
int ufx_cc_iterate_graph(ufx_cc* ctx, int iter0, int niter, double cfl, double* hist);

/* Form of the residual-stage kernel (no reference counterpart; a tuning
 * knob for the autotuner, unsflow2d_amd.batch.autotune): shared = 1 evaluates
 * every interior face flux once per block of 256 cells into LDS (an edge
 * batch per block) and sums them per cell; shared = 0 lets every cell
 * evaluate its own faces (each interior flux twice).  lds = 1 stages the
 * block's cell records in LDS.  All forms give the same result to rounding.
 * Returns UFX_ERR_UNSUPPORTED when the form's LDS exceeds 64 KB per block on
 * this mesh.  The default at ufx_cc_create is the measured best for the
 * precision (UFX_STAGE_KERNEL / UFX_STAGE_LDS override it). */
int ufx_cc_set_stage_form(ufx_cc* ctx, int shared, int lds);
int ufx_cc_get_stage_form(ufx_cc* ctx, int* shared, int* lds);

/* The normalised residuals of the last whole-iteration call on the context
 * (ufx_cc_iterate, ufx_cc_iterate_graph, or the context's share of
 * ufx_cc_batch_iterate, which has no history argument): entries 0..n-1, one
 * per iteration of that call; synchronising.  n may not exceed the longest
 * call so far. */
int ufx_cc_history(ufx_cc* ctx, int n, double* out);

/* Keep the residual of the last stage of every whole iteration
 * (ufx_cc_iterate, ufx_cc_iterate_graph) in the context, for
 * ufx_cc_get("res") and ufx_cc_residual_norm.  No reference counterpart:
 * the fused stages otherwise reduce the residual without storing it.  Costs
 * one more write of the residual per iteration. */
int ufx_cc_keep_residual(ufx_cc* ctx, int on);
/* sqrt(sum_i res_i(var)^2 / noc) of the stored residual for one conserved
 * variable (0 E, 1 rho, 2 rho u, 3 rho v), or with var = -1 of all four (the
 * reference's raw norm, CCFVM/solution.F90:159-166); synchronising. */
int ufx_cc_residual_norm(ufx_cc* ctx, int var, double* out);

/* Many independent cases per launch.  No reference counterpart: the
 * reference solves one mesh per process; this is the scheduler of the
 * batched-cases workload (many small meshes on one GPU), where one launch
 * per kernel and case would leave the GPU mostly idle.  The cases are
 * existing contexts (inviscid, RK3, no residual smoothing, LSQ or
 * Green-Gauss gradient; equal precision, flux, gradient and limiter
 * options), each already past iteration 1.  ufx_cc_batch_iterate runs niter
 * whole iterations (constant CFL) of every active case with one launch per
 * kernel for all of them, replayed from a HIP graph captured on the first
 * call (a call with fewer iterations than the captured chunk, such as the
 * remainder of a run, launches them directly; one with more, or after a
 * retirement or a change of stage form, captures anew); per cell it computes
 * exactly what ufx_cc_iterate_graph computes for
 * each case alone.  It synchronises, and reports the first case whose
 * device error word is set ("case i: ...").  A case's own history buffer
 * receives its norms as with ufx_cc_iterate_graph.  Retire cases with
 * ufx_cc_batch_set_active (active[i] = 0; a retired case does not return);
 * a retired case may then be destroyed.  Destroying the batch leaves the
 * cases alone. */
typedef struct ufx_cc_batch ufx_cc_batch;
int ufx_cc_batch_create(ufx_cc* const* cases, int n, ufx_cc_batch** out);
int ufx_cc_batch_set_active(ufx_cc_batch* batch, const int* active);
int ufx_cc_batch_iterate(ufx_cc_batch* batch, int niter, double cfl);
/* ufx_cc_residual_norm(case, var, &out[i]) of every case i in one launch
 * (NaN for retired cases); synchronising. */
int ufx_cc_batch_residual_norms(ufx_cc_batch* batch, int var, double* out);
int ufx_cc_batch_destroy(ufx_cc_batch* batch);

/* ufx_cc_residual_norm without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving or batch-driver loops, harnesses
 * that forbid hipMalloc).  ufx_cc_residual_norm allocates its result slot on
 * first use; this variant calls no hipMalloc, hipFree or other allocating
 * routine: its block partials and result live in a static device array of
 * UFX_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), which
 * limits it to contexts of at most (UFX_NOALLOC_SCRATCH_BYTES / 8 - 32) * 256
 * interior cells (33.5 M); larger ones are refused with UFX_ERR_ARG.  Same
 * result as ufx_cc_residual_norm, bit for bit; synchronising.  All calls on
 * one GPU share the array: calls that overlap on one GPU (other threads,
 * other streams) race and must be serialised by the caller. */
#define UFX_NOALLOC_SCRATCH_BYTES (1 << 20)
int ufx_cc_residual_norm_noalloc(ufx_cc* ctx, int var, double* out);

/* Kernel timing (no reference counterpart; instrumentation for benchmarks).
 * When enabled, a pair of HIP events is recorded on the context's stream
 * around every kernel launch; ufx_cc_profile_read (synchronising) returns,
 * per category, the summed milliseconds and the number of launches since the
 * last read, then resets.  Categories: */
#define UFX_K_PRIM 0      /* primitive refresh + range partials */
#define UFX_K_RANGE 1     /* global range final reduction */
#define UFX_K_AVG 2       /* vertex averages */
#define UFX_K_TSTEP 3     /* spectral radius + time step */
#define UFX_K_GRAD 4      /* gradient (+ limiter) */
#define UFX_K_STAGE 5     /* residual (+ stage update) */
#define UFX_K_NORM 6      /* residual norm */
#define UFX_K_OTHER 7
#define UFX_K_BFACE 8     /* boundary faces of a residual */
#define UFX_K_NCAT 9
int ufx_cc_profile(ufx_cc* ctx, int enable);
/* Time only the categories whose bit is set in mask (1 << UFX_K_STAGE: the
 * residual stages alone); each timed launch adds an event pair to the
 * stream, so a narrow mask disturbs the timed work least.  0 stops. */
int ufx_cc_profile_mask(ufx_cc* ctx, unsigned mask);
int ufx_cc_profile_read(ufx_cc* ctx, double* ms, int* count);

/* Block until all queued work has finished; returns the first device-detected
 * error, if any. */
int ufx_cc_sync(ufx_cc* ctx);
/* The device-detected error of the context in reference terms; synchronising.
 * *kind: 0 none, 1 Roe's averaged sound speed squared not positive
 * (CCFVM/roe_flux.F90:73-87), 2 non-positive density or pressure at an
 * iteration's start (check_positivity, CCFVM/main.F90:358-414), 3
 * non-positive diamond-path area (CCFVM/gradient.F90:297).  *face: the
 * 1-based reference face at which the reference's sequential loop stops
 * (kinds 1 and 3; boundary faces before interior faces, as fvresidual loops
 * them), else 0; *cell: its `in` cell (kinds 1, 3) or the first failing cell
 * (kind 2), 1-based.  An error ends the context's useful life, as the
 * reference's STOP ends the run: later kernels skip their work. */
int ufx_cc_error(ufx_cc* ctx, int* kind, int* face, int* cell);
/* The context's HIP stream (hipStream_t), for timing with HIP events. */
void* ufx_cc_stream(ufx_cc* ctx);

/* ============================================ vertex-centred (median dual)
 * The NCFVM solver's hot path (NCFVM/main.f90:51-88).  State per vertex, in
 * the reference's breadth-first vertex order, variables of one vertex
 * contiguous, FP64 (the mixed form's records aside). */
typedef struct ufx_nc ufx_nc;

/* Run parameters (NCFVM/read_input.f90:23-40; the deck's Flux and Limiter
 * entries are ignored by the reference, which always runs van Leer with the
 * Venkatakrishnan limiter). */
typedef struct {
  double mach;       /* Mach */
  double aoa_deg;    /* AOA */
  double cfl_max;    /* CFL: the target of the tanh ramp, NCFVM/main.f90:54 */
  int grad;          /* UFX_GRAD_GG or UFX_GRAD_LSQR (the two sides of the
                        merge conflict at NCFVM/fvresidual.f90:11-17) */
  int timemode;      /* UFX_TIME_RK3 (needs cfl_max < 1) or UFX_TIME_LUSGS */
  int precision;     /* UFX_PREC_FP64, or UFX_PREC_MIXED: vertex records (primitives
                        as deviations from the free stream, gradients, limiter,
                        coordinates), edge geometry and least-squares weights in
                        FP32, gradient and limiter in FP32; reconstruction, fluxes,
                        residual sums, time steps, state, norms and LU-SGS in FP64.
                        Build-defined; LU-SGS then needs UFX_LUSGS_MULTICOLOUR */
  int flux;          /* UFX_FLUX_VANLEER (the reference's) or UFX_FLUX_AUSM (build-
                        defined: the cell-centred AUSM+-up on the dual faces) */
  int limiter;       /* UFX_NC_LIM_VKN (limiter_VKN, NCFVM/limiter.f90:15-98) or
                        UFX_NC_LIM_BJ (limiter_min_max, :192-260, dead code in the
                        reference, with its +1/2 at :241) */
  int lusgs_order;   /* UFX_LUSGS_REFERENCE or UFX_LUSGS_MULTICOLOUR */
} ufx_nc_params;

#define UFX_NC_LIM_VKN 0
#define UFX_NC_LIM_BJ 1

/* Context on the current device in the state `initialize` leaves
 * (NCFVM/common.f90:21-58): free stream everywhere, limiter value 1e20. */
int ufx_nc_create(const ufx_mesh* mesh, const ufx_nc_params* params, ufx_nc** ctx);
int ufx_nc_destroy(ufx_nc* ctx);
/* save_old, NCFVM/common.f90:88-104 (primitive refresh and qold = qc). */
int ufx_nc_save_old(ufx_nc* ctx);
/* fvresidual, NCFVM/fvresidual.f90:4-85: gradient, limiter, edge fluxes
 * with wall and farfield closure, then time_step2 with `cfl`. */
int ufx_nc_fvresidual(ufx_nc* ctx, double cfl);
/* fvresidual and the RK update of NCFVM/main.f90:59-70 for stage 1..3. */
int ufx_nc_rk_stage(ufx_nc* ctx, int stage, double cfl);
/* lusgs, NCFVM/lusgs.f90:4-131 (calls fvresidual itself). */
int ufx_nc_lusgs(ufx_nc* ctx, double cfl);
/* residue, NCFVM/common.f90:108-154; synchronising.  out[5] = {L2 / L2 of
 * iteration 1, L2, Linf, its vertex (1-based), global dt}. */
int ufx_nc_residue(ufx_nc* ctx, int iter, double* out);
/* Iterations iter0+1 .. iter0+niter of NCFVM/main.f90:51-88 (tanh CFL
 * ramp); hist (5 x niter, as ufx_nc_residue's out) may be NULL. */
int ufx_nc_iterate(ufx_nc* ctx, int iter0, int niter, double* hist);
/* Wall forces of the current state.  NCFVM has no force routine (its
 * find_cl_cd counterpart, CCFVM/main.F90:263-355, exists only in the
 * cell-centred solver); this one is consistent with the wall closure of
 * NCFVM/solid_flux.f90:20-40: per wall edge (bc 1001, ascending edge order)
 * the end-point pressures weighted 3/4-1/4 on the half-edge right normal,
 * summed in edge order.  out[0..3] = cl, cd, fx, fy with the dynamic
 * pressure 0.5 gamma p_inf M^2 and the angle of attack as
 * CCFVM/main.F90:347-348; cp (may be NULL) receives -Cp at the first end
 * point of each wall edge, one value per wall edge in edge order. */
int ufx_nc_forces(ufx_nc* ctx, double* out, double* cp);
/* "qc","qold" (4 x nop) writable, "phi" (4 x nop) writable; readable also
 * "qp" (4), "grad" (8: d/dx, d/dy per variable), "res" (4), "la", "dt"
 * (nop), "fc_la" (nof). */
int ufx_nc_set(ufx_nc* ctx, const char* name, const double* src);
int ufx_nc_get(ufx_nc* ctx, const char* name, double* dst);
int ufx_nc_sync(ufx_nc* ctx);
/* Kernel timing: categories save_old, gradient+limiter, residual, norm,
 * LU-SGS (5 entries). */
int ufx_nc_profile(ufx_nc* ctx, int enable);
int ufx_nc_profile_read(ufx_nc* ctx, double* ms, int* count);

#ifdef __cplusplus
}
#endif
#endif /* UFX_H */
