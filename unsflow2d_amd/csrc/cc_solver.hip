// Host side of the device residual hot path (gfx950) and its C-ABI
// (include/ufx.h): context creation (storage layout, slots, LU-SGS level
// schedule), one entry point per reference routine, whole iterations, and
// state injection / read-back in the reference's per-cell record layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "cc_batch.h"
#include "cc_kernels.h"
#include "ufx_internal.h"

using namespace ufx;

#define HIPCHK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " +    \
                __FILE__ + ":" + std::to_string(__LINE__));                      \
      return UFX_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

namespace {

struct Params {
  double gamma, gamma1, znd, r_inf, u_inf, v_inf, p_inf, mref, t_wall, m_inf, rey, vll;
  double aoa, xref, yref, cfl_max;
  double qinf[4];
  int flow, iwall, flux, grad, ilimit, irs, timemode, precision, lusgs_order;
};

// A captured chunk of whole iterations: the graph is kept with its
// executable and both are released together (executable first), and a new
// chunk is instantiated before the one it replaces is released.
struct Chunk {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int n = 0;            // iterations per replay
  double cfl = 0.0;
  int keep = -1;        // residual kept (batched chunks)
  bool stale = true;    // its kernel arguments no longer describe the cases
  bool prims_next = false;  // a replay leaves the next iteration's primitives in qpc
};

// Release a chunk; the first failing call's error is returned.
hipError_t release_chunk(Chunk& c) {
  hipError_t e = hipSuccess;
  if (c.exec) e = hipGraphExecDestroy(c.exec);
  if (c.graph) {
    hipError_t e2 = hipGraphDestroy(c.graph);
    if (e == hipSuccess) e = e2;
  }
  c = Chunk();
  return e;
}

// Host copies of the shared-flux residual's block face lists.
struct Slots {
  std::vector<int> bf_ptr, bf_face, sl_loc;
  int bf_max = 0;
};

}  // namespace

struct ufx_cc {
  Params p{};
  Gas<double> g{};
  Gas<float> gf{};
  DevT<double> d64{};
  DevT<float> d32{};
  int nop = 0, noc = 0, nfi = 0, nbf = 0, nof = 0, ncell = 0;
  int device = 0;                // the HIP device the context lives on
  double cfl = 0.0;              // CFL of the current iteration (param's CFL)
  bool stage_shared = true;      // residual stage with block-shared fluxes
  bool stage_local = true;       // block records staged in LDS
  size_t stage_lds = 0;          // LDS bytes per block of the stage kernel
  size_t lds_list = 0, lds_recs = 0;  // LDS of the block flux list / records
  bool lds_overlay = false;            // the flux list fits over the records (at most 2 BLK faces)
  // alternate traversal: in RK stages 1 and 3 the XCDs walk their slices of
  // blocks from the far end, so each stage starts on the data the pass
  // before it touched last, which the 256 MB memory-side cache still holds
  // (UFX_SERPENTINE=0: all passes forward)
  int stage_rev = 10;                  // bit s: stage s walks its blocks from the far end
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // a captured chunk of whole iterations (ufx_cc_iterate_graph)
  Chunk chunk;
  bool keep_res = false;           // whole iterations also store the last residual
  bool qold_stale = false;         // fused RK3 iterations ran since qold was last written
  bool dtla_pending = false;       // fused RK3 iterations left la and dt unstored (formed
  double dtla_cfl = 0.0;           // from the face radii at this CFL when read)
  bool prims_next = false;         // qpc/qTc hold the next iteration's primitives (written by
                                   // the last stage of a whole iteration; see prim_epilogue)
  bool iter_fusion = true;         // whole RK3 iterations refresh the primitives in their last stage
  unsigned prof_mask = ~0u;        // kernel categories timed while profiling
  double* d_out = nullptr;         // small device results (residual norms)
  int nblk_cell = 0, nblk_all = 0, nblk_pt = 0, nblk_face = 0;
  double* d_hist = nullptr;
  int hist_cap = 0;
  double* cirs_tmp = nullptr;    // [noc][4] the first smoothing sweep's result
  double* bfluxv = nullptr;      // viscous wall force terms [2][nbf]
  std::vector<int> sto_of_ref;   // reference cell -> storage cell (ncell)
  std::vector<int> ref_of_sto;   // storage cell -> reference cell
  std::vector<int> ref_of_face;  // storage face -> reference face
  std::vector<int> face_of_ref;  // reference face -> storage face
  std::vector<int> pt_of_ref;    // reference vertex -> storage vertex
  std::vector<int> wall_k;       // boundary ordinals of wall faces
  std::vector<int> fin_h, fp1_h, fp2_h;  // boundary faces (reference numbering)
  std::vector<int> fin_ref;        // `in` cell of every face (reference numbering)
  std::vector<double> fnx_h, fny_h, farea_h, px_h, py_h;
  // LU-SGS level schedule: cells grouped by level, forward then backward
  int* lv_cells = nullptr;
  std::vector<int> lv_fwd_off, lv_bwd_off;
  // multicolour LU-SGS: storage cells by colour, offsets of each colour
  int* mc_cells = nullptr;
  std::vector<int> mc_off;
  // kernel timing
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Rec { int cat; size_t a, b; };
  std::vector<Rec> recs;
};

namespace {
// LDS per block of a residual-stage form: the block's flux list (shared),
// its records (local), both -- or, when the list fits over the records, the
// records alone.
size_t stage_lds_need(const ufx_cc* x, bool shared, bool local) {
  if (shared && local && x->lds_overlay) return std::max(x->lds_list, x->lds_recs);
  return (shared ? x->lds_list : 0) + (local ? x->lds_recs : 0);
}
}  // namespace

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
  if (!x->prof || !((x->prof_mask >> cat) & 1u)) return -1;
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
// The same with `lds` bytes of dynamic shared memory per block.
#define LAUNCH_LDS(x, cat, kern, grid, lds, ...)                         \
  do {                                                                   \
    int r_ = prof_begin((x), (cat));                                     \
    kern<<<(grid), BLK, (lds), (x)->stream>>>(__VA_ARGS__);              \
    prof_end((x), r_);                                                   \
  } while (0)

namespace {

int nblocks(int n) { return std::max(1, (n + BLK - 1) / BLK); }

// LDS a residual-stage block may take (the kernels are launched with it as
// dynamic shared memory)
constexpr size_t STAGE_LDS_MAX = 64 * 1024;

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

// Whole RK3 iterations of the fused form keep their start state in qc and
// never write qold; a routine that reads qold afterwards would see the
// state of the last save_old.  Refuse until save_old (or set("qold")).
int check_qold(ufx_cc* x, const char* who) {
  if (!x->qold_stale) return UFX_OK;
  set_error(std::string(who) + ": qold is stale after whole RK3 iterations; call "
            "ufx_cc_save_old first");
  return UFX_ERR_STATE;
}

template <typename T>
int dalloc(ufx_cc* x, T** p, size_t n) {
  if (n == 0) n = 1;
  HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
  x->allocs.push_back((void*)*p);
  HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), x->stream));
  return UFX_OK;
}

template <typename T>
int dupload(ufx_cc* x, T** dst, const std::vector<T>& src) {
  T* p = nullptr;
  if (int rc = dalloc(x, &p, src.size())) return rc;
  if (!src.empty())
    HIPCHK(hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, x->stream));
  *dst = p;
  return UFX_OK;
}

template <typename T>
int dupload(ufx_cc* x, const T** dst, const std::vector<T>& src) {
  T* p = nullptr;
  int rc = dupload(x, &p, src);
  *dst = p;
  return rc;
}

int* err_word(ufx_cc* x) { return x->p.precision == UFX_PREC_FP64 ? x->d64.err : x->d32.err; }

// The device error word in reference terms: kind (DERR_*), 1-based
// reference face and cell (0 where they do not apply).
void decode_error(const ufx_cc* x, const int* e, int* kind, int* face, int* cell) {
  *kind = e[0];
  *face = 0;
  *cell = 0;
  if (e[0] == DERR_ROE && e[3] >= 0 && e[3] < x->nbf + x->nfi) {
    // boundary faces are keyed by their ordinal, interior ones after them
    int f = e[3] < x->nbf ? x->nfi + e[3] : e[3] - x->nbf;
    *face = f + 1;
    *cell = x->fin_ref[f] + 1;
  } else if (e[0] == DERR_AREA && e[3] >= 0 && e[3] < x->nof) {
    *face = e[3] + 1;
    *cell = x->fin_ref[e[3]] + 1;
  } else if (e[0] == DERR_POS) {
    *cell = e[2];
  }
}

int read_error(ufx_cc* x, int* e4) {
  HIPCHK(hipMemcpyAsync(e4, err_word(x), 4 * sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return UFX_OK;
}

int device_error(ufx_cc* x) {
  int e[4] = {0, 0, 0, 0};
  if (int rc = read_error(x, e)) return rc;
  if (e[0] == 0) return UFX_OK;
  int kind, face, cell;
  decode_error(x, e, &kind, &face, &cell);
  // the reference's STOP messages (CCFVM/roe_flux.F90:74-85,
  // CCFVM/main.F90:386, CCFVM/gradient.F90:297), with where it happened
  if (kind == DERR_ROE)
    set_error("Roe scheme: Sonic speed is negative (face " + std::to_string(face) + ", cell " +
              std::to_string(cell) + ")");
  else if (kind == DERR_POS)
    set_error("Density/pressure is negative at point " + std::to_string(cell));
  else if (kind == DERR_AREA)
    set_error("ggfc,gradtrixy,  -ve area (face " + std::to_string(face) + ")");
  else set_error("device error " + std::to_string(e[0]));
  return UFX_ERR_STATE;
}

// Run f(d) on the device view of the context's precision.
template <typename F>
int with_dev(ufx_cc* x, F&& f) {
  if (x->p.precision == UFX_PREC_FP64) return f(x->d64);
  return f(x->d32);
}

template <typename R>
const Gas<R>& gas_of(ufx_cc* x);
template <> const Gas<double>& gas_of<double>(ufx_cc* x) { return x->g; }
template <> const Gas<float>& gas_of<float>(ufx_cc* x) { return x->gf; }

// ---- per-routine launchers (templated on the bulk type)

template <typename R>
int run_prim(ufx_cc* x, DevT<R>& d, int do_prim, int do_old, int do_range) {
  LAUNCH(x, UFX_K_PRIM, k_prim<R>, x->nblk_all, d, x->g, do_prim, do_old, do_range);
  return launch_check();
}

// Before a routine call: the working primitives back to the current ones
// when a whole iteration refreshed them ahead (k_prims_current).
int prims_current(ufx_cc* x) {
  if (!x->prims_next) return UFX_OK;
  x->prims_next = false;
  return with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    LAUNCH(x, UFX_K_OTHER, k_prims_current<R>, x->nblk_cell, d);
    return launch_check();
  });
}

int check_routine(ufx_cc* x) {
  if (int rc = check_ctx(x)) return rc;
  return prims_current(x);
}

template <typename R>
int run_range_final(ufx_cc* x, DevT<R>& d) {
  LAUNCH(x, UFX_K_RANGE, k_range_final<R>, 1, d, x->nblk_all);
  return launch_check();
}

template <typename R>
int run_range(ufx_cc* x, DevT<R>& d) {
  if (int rc = run_prim(x, d, 0, 0, 1)) return rc;
  return run_range_final(x, d);
}

template <typename R>
int run_avg(ufx_cc* x, DevT<R>& d) {
  if (x->p.flow != UFX_FLOW_INVISCID) LAUNCH(x, UFX_K_AVG, (k_avg_c2v<R, true>), x->nblk_pt, d);
  else LAUNCH(x, UFX_K_AVG, (k_avg_c2v<R, false>), x->nblk_pt, d);
  return launch_check();
}

// avg_c2v with the limiter's range reduction riding along (k_prim's block
// partials must be complete, as for run_range_final).
template <typename R>
int run_avg_range(ufx_cc* x, DevT<R>& d, int nparts) {
  if (x->p.flow != UFX_FLOW_INVISCID)
    LAUNCH(x, UFX_K_AVG, (k_avg_c2v_range<R, true>), x->nblk_pt + 1, d, nparts);
  else
    LAUNCH(x, UFX_K_AVG, (k_avg_c2v_range<R, false>), x->nblk_pt + 1, d, nparts);
  return launch_check();
}

// k_cell with the options of one call; with_t when the temperature gradient
// has a consumer (the drop-in gradient routines always produce it); tstep:
// with the time step from the face radii of a k_face_lambda pass.  (Summing
// radii evaluated inside k_cell instead saves that pass but doubled the
// prec gathers' HBM traffic, the vertex averages no longer staying in L2
// beside k_cell's other streams: same time, 1.36 instead of 1.09 GB.)
template <typename R, int GR>
int run_cell_g(ufx_cc* x, DevT<R>& d, bool with_t, bool limit, bool tstep, double cfl,
               int wdt = 1) {
  double vll = x->p.vll;
  if (tstep) {
    x->dtla_pending = !wdt;
    x->dtla_cfl = cfl;
  }
  int cat = GR != G_NONE ? UFX_K_GRAD : (tstep ? UFX_K_TSTEP : UFX_K_GRAD);
  if (tstep) {
    LAUNCH(x, UFX_K_TSTEP, k_face_lambda<R>, x->nblk_face, d, x->g, vll);
    if (int rc = launch_check()) return rc;
  }
#define KC(WT, LM, TS) \
  LAUNCH(x, cat, (k_cell<R, GR, WT, LM, TS>), x->nblk_cell, d, x->g, vll, cfl, wdt)
  if (with_t) {
    if (limit) { if (tstep) KC(true, true, true); else KC(true, true, false); }
    else { if (tstep) KC(true, false, true); else KC(true, false, false); }
  } else {
    if (limit) { if (tstep) KC(false, true, true); else KC(false, true, false); }
    else { if (tstep) KC(false, false, true); else KC(false, false, false); }
  }
#undef KC
  return launch_check();
}

template <typename R>
int run_ggfc_face(ufx_cc* x, DevT<R>& d) {
  LAUNCH(x, UFX_K_GRAD, k_ggfc_face<R>, x->nblk_face, d);
  return launch_check();
}

// The gradient of the context's method, optionally fused with the limiter
// and the time step (the reference calls limit from inside each gradient
// routine when ILIMIT is set).
template <typename R>
int run_gradient(ufx_cc* x, DevT<R>& d, bool with_t, bool limit, bool tstep, double cfl,
                 int wdt = 1) {
  switch (x->p.grad) {
    case UFX_GRAD_GG: return run_cell_g<R, G_GG>(x, d, with_t, limit, tstep, cfl, wdt);
    case UFX_GRAD_GGFC:
      if (int rc = run_ggfc_face(x, d)) return rc;
      return run_cell_g<R, G_GGFC>(x, d, with_t, limit, tstep, cfl, wdt);
    default: return run_cell_g<R, G_LSQR>(x, d, with_t, limit, tstep, cfl, wdt);
  }
}

// la and dt of the last time step, when fused iterations left them unstored:
// the time-step pass alone, from the face radii (which nothing has changed
// since) at that step's CFL, gives them bit for bit.
template <typename R>
int restore_dtla(ufx_cc* x, DevT<R>& d) {
  if (!x->dtla_pending) return UFX_OK;
  x->dtla_pending = false;
  LAUNCH(x, UFX_K_TSTEP, (k_cell<R, G_NONE, false, false, true>), x->nblk_cell, d, x->g, x->p.vll,
         x->dtla_cfl, 1);
  return launch_check();
}

template <typename R>
int run_stage_kernel(ufx_cc* x, DevT<R>& d, int mode, double ak, double bk, int write_res,
                     int epi = 0) {
  const int src = (mode >> 2) & 3;
  int visc = x->p.flow != UFX_FLOW_INVISCID;
  const Gas<R>& gr = gas_of<R>(x);
  int nbb = nblocks(x->nbf);
#define KS(F)                                                                                   \
  do {                                                                                          \
    if (x->nbf > 0)                                                                             \
      LAUNCH(x, UFX_K_BFACE, (k_bface<R, F>), nbb, d, x->g, visc, x->p.iwall, x->p.t_wall, src); \
    if (x->stage_shared && x->stage_local)                                                      \
      LAUNCH_LDS(x, UFX_K_STAGE, (k_stage_blk<R, F, true>), x->nblk_cell, x->stage_lds, d, gr,  \
                 mode, ak, bk, write_res, x->g, epi);                                           \
    else if (x->stage_shared)                                                                   \
      LAUNCH_LDS(x, UFX_K_STAGE, (k_stage_blk<R, F, false>), x->nblk_cell, x->stage_lds, d, gr, \
                 mode, ak, bk, write_res, x->g, epi);                                           \
    else if (x->stage_local)                                                                    \
      LAUNCH_LDS(x, UFX_K_STAGE, (k_stage<R, F, true>), x->nblk_cell, x->stage_lds, d, gr,      \
                 mode, ak, bk, write_res, x->g, epi);                                           \
    else                                                                                        \
      LAUNCH(x, UFX_K_STAGE, (k_stage<R, F, false>), x->nblk_cell, d, gr, mode, ak, bk,         \
             write_res, x->g, epi);                                                             \
  } while (0)
  switch (x->p.flux) {
    case UFX_FLUX_ROE: KS(1); break;
    case UFX_FLUX_AUSM: KS(2); break;
    case UFX_FLUX_VANLEER: KS(3); break;
    default: KS(4); break;
  }
#undef KS
  return launch_check();
}

// Two Jacobi sweeps of the residual smoothing with the smoothing factor of
// the current CFL, CCFVM/ImpResiSmoothening.F90:13-44 (OldRes = residual on
// entry).  A sweep reads OldRes only at its own cell, before writing that
// cell, so the second sweep reads it from the residual it overwrites and
// no copy of it is kept.
// upd: the second sweep also applies the stage update (ak, bk).
template <typename R>
int run_cirs(ufx_cc* x, DevT<R>& d, bool upd = false, double ak = 0.0, double bk = 0.0) {
  double rr = x->cfl / x->p.cfl_max;
  double eps = std::max(0.25 * (rr * rr - 1.0), 0.0);
  double* tmp = x->cirs_tmp;
  LAUNCH(x, UFX_K_STAGE, k_cirs<R>, x->nblk_cell, d, eps, d.res, d.res, tmp, 0, 0.0, 0.0);
  LAUNCH(x, UFX_K_STAGE, k_cirs<R>, x->nblk_cell, d, eps, tmp, d.res, d.res, upd ? 1 : 0, ak, bk);
  return launch_check();
}

// fvresidual, CCFVM/fvresidual.F90:4-187: inviscid faces, then the viscous
// flux (which first recomputes the diamond-path gradient and the limiter
// unless that is the gradient method, CCFVM/viscous_flux.F90:17) and the
// residual smoothing.
// upd: a stage of a whole iteration with residual smoothing, whose last
// smoothing sweep also applies the stage update (ak, bk); returns with *upd_done set.
template <typename R>
int run_residual_full(ufx_cc* x, DevT<R>& d, bool upd = false, double ak = 0.0, double bk = 0.0,
                      bool* upd_done = nullptr) {
  if (int rc = run_stage_kernel(x, d, 0, 0.0, 0.0, 1)) return rc;
  if (x->p.flow != UFX_FLOW_INVISCID) {
    if (x->p.grad != UFX_GRAD_GGFC) {
      if (int rc = run_ggfc_face(x, d)) return rc;
      if (x->p.ilimit)
        if (int rc = run_range(x, d)) return rc;
      if (int rc = run_cell_g<R, G_GGFC>(x, d, true, x->p.ilimit != 0, false, 0.0)) return rc;
    }
    LAUNCH(x, UFX_K_STAGE, k_visc_mul<R>, x->nblk_cell, d, x->g);
    LAUNCH(x, UFX_K_STAGE, k_visc_face<R>, x->nblk_face, d, x->g, x->p.rey, x->p.m_inf,
           x->bfluxv);
    LAUNCH(x, UFX_K_STAGE, k_visc_sum<R>, x->nblk_cell, d);
    if (int rc = launch_check()) return rc;
  }
  if (x->p.irs) {
    if (int rc = run_cirs(x, d, upd, ak, bk)) return rc;
    if (upd_done) *upd_done = upd;
  }
  return UFX_OK;
}

template <typename R>
int run_rk_update(ufx_cc* x, DevT<R>& d, double ak, double bk) {
  LAUNCH(x, UFX_K_STAGE, k_rk_update<R>, x->nblk_cell, d, ak, bk);
  return launch_check();
}

const double RK_A[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
const double RK_B[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};

// One Runge-Kutta stage: residual and update fused in one pass when nothing
// else contributes to the residual, else the full residual then the update.
// whole: a stage of a whole iteration (its states in qc and qs, see
// stage_old/in/out), else routine by routine (qold and qc); epi: the stage
// also refreshes the next iteration's primitives (prim_epilogue).
template <typename R>
int run_stage(ufx_cc* x, DevT<R>& d, int stage, bool norm, bool write_res, bool whole,
              bool epi = false) {
  double ak = RK_A[stage - 1], bk = RK_B[stage - 1];
  if (x->p.flow == UFX_FLOW_INVISCID && !x->p.irs)
    return run_stage_kernel(x, d, (norm ? 2 : 1) | ((whole ? stage : 0) << 2) |
                                      ((x->stage_rev >> stage) & 1 ? 16 : 0),
                            ak, bk, write_res ? 1 : 0, epi ? 1 : 0);
  bool updated = false;
  if (int rc = run_residual_full(x, d, true, ak, bk, &updated)) return rc;
  if (!updated)
    if (int rc = run_rk_update(x, d, ak, bk)) return rc;
  if (norm) {
    LAUNCH(x, UFX_K_NORM, k_sumsq<R>, x->nblk_cell, d);
    return launch_check();
  }
  return UFX_OK;
}

// lusgs2, CCFVM/lusgs.F90:207-379: residual, level-scheduled forward and
// backward sweeps, update.
template <typename R>
int run_lusgs(ufx_cc* x, DevT<R>& d) {
  if (int rc = run_residual_full(x, d)) return rc;
  if (x->p.lusgs_order == UFX_LUSGS_MULTICOLOUR) {
    // one launch per colour: forward ascending, backward descending
    const int nc = (int)x->mc_off.size() - 1;
    for (int dir = 0; dir < 2; ++dir)
      for (int k = 0; k < nc; ++k) {
        int c = dir == 0 ? k : nc - 1 - k, a = x->mc_off[c], n = x->mc_off[c + 1] - a;
        if (n > 0)
          LAUNCH(x, UFX_K_OTHER, (k_lusgs_level<R, true>), nblocks(n), d, x->g, x->mc_cells + a,
                 n, dir);
      }
    LAUNCH(x, UFX_K_OTHER, k_lusgs_final<R>, x->nblk_cell, d);
    return launch_check();
  }
  int nlv = (int)x->lv_fwd_off.size() - 1;
  for (int l = 0; l < nlv; ++l) {
    int a = x->lv_fwd_off[l], n = x->lv_fwd_off[l + 1] - a;
    LAUNCH(x, UFX_K_OTHER, (k_lusgs_level<R, false>), nblocks(n), d, x->g, x->lv_cells + a, n, 0);
  }
  nlv = (int)x->lv_bwd_off.size() - 1;
  for (int l = 0; l < nlv; ++l) {
    int a = x->lv_bwd_off[l], n = x->lv_bwd_off[l + 1] - a;
    LAUNCH(x, UFX_K_OTHER, (k_lusgs_level<R, false>), nblocks(n), d, x->g, x->lv_cells + a, n, 1);
  }
  LAUNCH(x, UFX_K_OTHER, k_lusgs_final<R>, x->nblk_cell, d);
  return launch_check();
}

int ensure_hist(ufx_cc* x, int n) {
  if (n <= x->hist_cap) return UFX_OK;
  // a captured chunk writes the old buffer: it is never replayed again
  if (x->stream) HIPCHK(hipStreamSynchronize(x->stream));
  x->chunk.stale = true;
  if (x->d_hist) HIPCHK(hipFree(x->d_hist));
  x->d_hist = nullptr;
  HIPCHK(hipMalloc((void**)&x->d_hist, (size_t)n * sizeof(double)));
  x->hist_cap = n;
  return UFX_OK;
}

// One whole iteration of CCFVM/main.F90:59-115.
template <typename R>
int run_iteration(ufx_cc* x, DevT<R>& d, int iter, int k) {
  bool lim = x->p.ilimit != 0;
  bool visc = x->p.flow != UFX_FLOW_INVISCID;
  // prim refresh, qold (save_old commutes with avg_c2v and the time step),
  // positivity and the global range (re-armed by the previous norm)
  // the fused RK3 stages keep the start state in qc (no qold copy)
  const bool fused = x->p.timemode == UFX_TIME_RK3 && !visc && !x->p.irs;
  // the last stage of a fused iteration refreshes the next iteration's
  // primitives itself (not with the diamond-path gradient, whose face pass
  // reads the records' primitives before the gradient pass copies them in)
  const bool epi = fused && x->p.grad != UFX_GRAD_GGFC && x->iter_fusion;
  x->qold_stale = fused;
  const bool refresh = !(epi && x->prims_next);
  x->prims_next = false;
  if (refresh)
    if (int rc = run_prim(x, d, 1, fused ? 2 : 1, lim ? 1 : 0)) return rc;
  // vertex averages, with the range reduction as one more block
  if (int rc = lim ? run_avg_range(x, d, refresh ? x->nblk_all : x->nblk_cell) : run_avg(x, d))
    return rc;
  // gradient, limiter and time step in one pass over the cells; the
  // temperature gradient only when the viscous flux needs it
  if (int rc = run_gradient(x, d, visc, lim, true, x->cfl, fused ? 0 : 1)) return rc;
  if (x->p.timemode == UFX_TIME_RK3) {
    for (int s = 1; s <= 3; ++s)
      if (int rc = run_stage(x, d, s, s == 3, s == 3 && x->keep_res, fused, s == 3 && epi))
        return rc;
  } else {
    if (int rc = run_lusgs(x, d)) return rc;
    LAUNCH(x, UFX_K_NORM, k_sumsq<R>, x->nblk_cell, d);
  }
  LAUNCH(x, UFX_K_NORM, k_norm_final<R>, 1, d, x->nblk_cell, iter == 1, x->d_hist, k);
  x->prims_next = epi;
  return launch_check();
}

// ---- state transfer in the reference's per-cell record layout

template <typename R>
int get_field(ufx_cc* x, DevT<R>& d, const std::string& n, double* dst) {
  constexpr int RW = Layout<R>::RW, PW = Layout<R>::PW;
  auto copy = [&](void* h, const void* dv, size_t bytes) {
    return hipMemcpyAsync(h, dv, bytes, hipMemcpyDeviceToHost, x->stream);
  };
  const int noc = x->noc, ncell = x->ncell, nof = x->nof, nop = x->nop;
  const std::vector<int>& sto = x->sto_of_ref;
  if (n == "qc" || n == "qold" || n == "res") {
    int cnt = n == "qc" ? ncell : noc;
    const double* src = n == "qc" ? d.qc : (n == "qold" ? d.qold : d.res);
    std::vector<double> t((size_t)cnt * 4);
    HIPCHK(copy(t.data(), src, t.size() * 8));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int c = 0; c < cnt; ++c)
      for (int k = 0; k < 4; ++k) dst[(size_t)c * 4 + k] = t[(size_t)sto[c] * 4 + k];
    return UFX_OK;
  }
  if (n == "dt" || n == "la") {
    if (int rc = restore_dtla(x, d)) return rc;
    std::vector<double> t(noc);
    HIPCHK(copy(t.data(), n == "dt" ? d.dt : d.la, (size_t)noc * 8));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int c = 0; c < noc; ++c) dst[c] = t[sto[c]];
    return UFX_OK;
  }
  if (n == "qp" || n == "grad" || n == "phi") {
    std::vector<R> rec((size_t)ncell * RW), t((size_t)ncell), gt((size_t)noc * 2);
    HIPCHK(copy(rec.data(), d.rec, rec.size() * sizeof(R)));
    HIPCHK(copy(t.data(), d.qT, t.size() * sizeof(R)));
    HIPCHK(copy(gt.data(), d.gT, gt.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (n == "qp")
      for (int c = 0; c < ncell; ++c) {
        size_t s = sto[c];
        for (int k = 0; k < 4; ++k) dst[(size_t)c * 5 + k] = rec[s * RW + R_QP + k];
        dst[(size_t)c * 5 + 4] = t[s];
      }
    else if (n == "grad")
      for (int c = 0; c < noc; ++c) {
        size_t s = sto[c];
        for (int k = 0; k < 8; ++k) dst[(size_t)c * 10 + k] = rec[s * RW + R_G + k];
        dst[(size_t)c * 10 + 8] = gt[s * 2];
        dst[(size_t)c * 10 + 9] = gt[s * 2 + 1];
      }
    else
      for (int c = 0; c < noc; ++c)
        for (int k = 0; k < 4; ++k) dst[(size_t)c * 4 + k] = rec[(size_t)sto[c] * RW + R_PHI + k];
    return UFX_OK;
  }
  if (n == "fc_la" || n == "fc_mu") {
    std::vector<double> t(nof);
    HIPCHK(copy(t.data(), n == "fc_la" ? d.fla : d.fmu, (size_t)nof * 8));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int f = 0; f < nof; ++f) dst[f] = t[x->face_of_ref[f]];
    return UFX_OK;
  }
  if (n == "fc_grad") {
    std::vector<R> t((size_t)nof * 10);
    HIPCHK(copy(t.data(), d.fgrad, t.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int f = 0; f < nof; ++f)
      for (int k = 0; k < 10; ++k) dst[(size_t)f * 10 + k] = t[(size_t)x->face_of_ref[f] * 10 + k];
    return UFX_OK;
  }
  if (n == "pt_prim") {
    std::vector<R> t((size_t)nop * PW);
    HIPCHK(copy(t.data(), d.prec, t.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int i = 0; i < nop; ++i)
      for (int k = 0; k < 5; ++k) dst[(size_t)i * 5 + k] = t[(size_t)x->pt_of_ref[i] * PW + k];
    return UFX_OK;
  }
  if (n == "pt_mu") {
    std::vector<R> t((size_t)nop * PW);
    HIPCHK(copy(t.data(), d.prec, t.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int i = 0; i < nop; ++i) dst[i] = t[(size_t)x->pt_of_ref[i] * PW + P_MU];
    return UFX_OK;
  }
  if (n == "cell_mul") {
    std::vector<R> t((size_t)ncell);
    HIPCHK(copy(t.data(), d.cmul, t.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int c = 0; c < ncell; ++c) dst[c] = t[sto[c]];
    return UFX_OK;
  }
  set_error("ufx_cc_get: unknown field " + n);
  return UFX_ERR_ARG;
}

template <typename R>
int set_field(ufx_cc* x, DevT<R>& d, const std::string& n, const double* src) {
  constexpr int RW = Layout<R>::RW;
  auto up = [&](void* dv, const void* h, size_t bytes) {
    return hipMemcpyAsync(dv, h, bytes, hipMemcpyHostToDevice, x->stream);
  };
  const int noc = x->noc, ncell = x->ncell, nof = x->nof;
  const std::vector<int>& sto = x->sto_of_ref;
  if (n == "qc" || n == "qold" || n == "res") {
    int cnt = n == "qc" ? ncell : noc;
    double* dv = n == "qc" ? d.qc : (n == "qold" ? d.qold : d.res);
    std::vector<double> t((size_t)cnt * 4);
    for (int c = 0; c < cnt; ++c)
      for (int k = 0; k < 4; ++k) t[(size_t)sto[c] * 4 + k] = src[(size_t)c * 4 + k];
    HIPCHK(up(dv, t.data(), t.size() * 8));
    HIPCHK(hipStreamSynchronize(x->stream));
    return UFX_OK;
  }
  if (n == "dt" || n == "la") {
    // the other of the two stored first, if fused iterations left them
    if (int rc = restore_dtla(x, d)) return rc;
    std::vector<double> t(noc);
    for (int c = 0; c < noc; ++c) t[sto[c]] = src[c];
    HIPCHK(up(n == "dt" ? d.dt : d.la, t.data(), (size_t)noc * 8));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (n == "dt") {
      // dt / area, which the stage update uses
      std::vector<double> cv(noc), dta(noc);
      HIPCHK(hipMemcpyAsync(cv.data(), d.cv, (size_t)noc * 8, hipMemcpyDeviceToHost, x->stream));
      HIPCHK(hipStreamSynchronize(x->stream));
      for (int s = 0; s < noc; ++s) dta[s] = t[s] / cv[s];
      HIPCHK(up(d.dta, dta.data(), (size_t)noc * 8));
      HIPCHK(hipStreamSynchronize(x->stream));
    }
    return UFX_OK;
  }
  if (n == "qp" || n == "grad" || n == "phi") {
    std::vector<R> rec((size_t)ncell * RW), t((size_t)ncell), gt((size_t)noc * 2);
    HIPCHK(hipMemcpyAsync(rec.data(), d.rec, rec.size() * sizeof(R), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(t.data(), d.qT, t.size() * sizeof(R), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(gt.data(), d.gT, gt.size() * sizeof(R), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (n == "qp")
      for (int c = 0; c < ncell; ++c) {
        size_t s = sto[c];
        for (int k = 0; k < 4; ++k) rec[s * RW + R_QP + k] = (R)src[(size_t)c * 5 + k];
        t[s] = (R)src[(size_t)c * 5 + 4];
      }
    else if (n == "grad")
      for (int c = 0; c < noc; ++c) {
        size_t s = sto[c];
        for (int k = 0; k < 8; ++k) rec[s * RW + R_G + k] = (R)src[(size_t)c * 10 + k];
        gt[s * 2] = (R)src[(size_t)c * 10 + 8];
        gt[s * 2 + 1] = (R)src[(size_t)c * 10 + 9];
      }
    else
      for (int c = 0; c < noc; ++c)
        for (int k = 0; k < 4; ++k) rec[(size_t)sto[c] * RW + R_PHI + k] = (R)src[(size_t)c * 4 + k];
    HIPCHK(up(d.rec, rec.data(), rec.size() * sizeof(R)));
    HIPCHK(up(d.qT, t.data(), t.size() * sizeof(R)));
    HIPCHK(up(d.gT, gt.data(), gt.size() * sizeof(R)));
    std::vector<R> qpc((size_t)noc * 4);
    for (int s = 0; s < noc; ++s)
      for (int k = 0; k < 4; ++k) qpc[(size_t)s * 4 + k] = rec[(size_t)s * RW + R_QP + k];
    HIPCHK(up(d.qpc, qpc.data(), qpc.size() * sizeof(R)));
    HIPCHK(up(d.qTc, t.data(), (size_t)noc * sizeof(R)));
    std::vector<R> phc((size_t)noc * 4);
    for (int s = 0; s < noc; ++s)
      for (int k = 0; k < 4; ++k) phc[(size_t)s * 4 + k] = rec[(size_t)s * RW + R_PHI + k];
    HIPCHK(up(d.phc, phc.data(), phc.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    return UFX_OK;
  }
  if (n == "fc_la" || n == "fc_mu") {
    std::vector<double> t(nof);
    for (int f = 0; f < nof; ++f) t[x->face_of_ref[f]] = src[f];
    HIPCHK(up(n == "fc_la" ? d.fla : d.fmu, t.data(), (size_t)nof * 8));
    HIPCHK(hipStreamSynchronize(x->stream));
    return UFX_OK;
  }
  if (n == "cell_mul") {
    std::vector<R> t((size_t)ncell);
    for (int c = 0; c < ncell; ++c) t[sto[c]] = (R)src[c];
    HIPCHK(up(d.cmul, t.data(), t.size() * sizeof(R)));
    HIPCHK(hipStreamSynchronize(x->stream));
    return UFX_OK;
  }
  set_error("ufx_cc_set: unknown or read-only field " + n);
  return UFX_ERR_ARG;
}

// ---- context construction

template <typename R>
int build_dev(ufx_cc* x, DevT<R>& d, const Mesh& m, const std::vector<int>& sl_face,
              const std::vector<int>& sl_nbr, const std::vector<int>& cnf,
              const std::vector<int>& lu_face, const std::vector<int>& lu_cnt, const Slots& sl) {
  constexpr int RW = Layout<R>::RW, FW = Layout<R>::FW, PW = Layout<R>::PW;
  const int noc = m.noc, ncell = m.ncell(), nof = m.nof(), nop = m.nop, nbf = m.nbf;
  d.nop = nop; d.noc = noc; d.nfi = m.nfi; d.nbf = nbf; d.nof = nof; d.ncell = ncell;
  const std::vector<int>& sto = x->sto_of_ref;
  // faces in storage order
  std::vector<R> frec((size_t)nof * FW, R(0)), fcen((size_t)nof * 2), fnl((size_t)nof * 4, R(0));
  std::vector<int2> fpt(nof);
  std::vector<int> fin(nof), fout(nof);
  for (int f = 0; f < nof; ++f) {
    int s = x->face_of_ref[f];
    R* r = &frec[(size_t)s * FW];
    r[F_NX] = (R)m.fnx[f]; r[F_NY] = (R)m.fny[f]; r[F_AREA] = (R)m.farea[f];
    r[F_CX] = (R)m.fcx[f]; r[F_CY] = (R)m.fcy[f];
    fcen[(size_t)s * 2] = r[F_CX];
    fcen[(size_t)s * 2 + 1] = r[F_CY];
    fnl[(size_t)s * 4] = r[F_NX];
    fnl[(size_t)s * 4 + 1] = r[F_NY];
    fnl[(size_t)s * 4 + 2] = r[F_AREA];
    fpt[s] = make_int2(x->pt_of_ref[m.fp1[f]], x->pt_of_ref[m.fp2[f]]);
    fin[s] = sto[m.fin[f]];
    fout[s] = sto[m.fout[f]];
  }
  std::vector<int> fbc(std::max(1, nbf)), bface(std::max(1, nbf)), bcell(std::max(1, nbf));
  for (int k = 0; k < nbf; ++k) {
    fbc[k] = m.fbc[m.nfi + k];
    bface[k] = x->face_of_ref[m.nfi + k];
    bcell[k] = sto[m.fin[m.nfi + k]];
  }
  // per-cell mesh data in storage order
  auto perm = [&](const std::vector<double>& a, int n) {
    std::vector<double> o(n);
    for (int s = 0; s < n; ++s) o[s] = a[x->ref_of_sto[s]];
    return o;
  };
  // the least-squares factors and vertex weights in R: the gradient runs in
  // R, and the mixed form's vertex averages take FP32 weights (half the
  // bytes of the pass's largest stream)
  auto to_r = [](const std::vector<double>& a) { return std::vector<R>(a.begin(), a.end()); };
  std::vector<R> r11 = to_r(perm(m.r11, noc)), r12 = to_r(perm(m.r12, noc)),
                 r22 = to_r(perm(m.r22, noc));
  std::vector<double> cv = perm(m.ccv, noc), cov = perm(m.ccov, ncell);
  std::vector<int> nc2c(noc);
  for (int s = 0; s < noc; ++s) nc2c[s] = m.nc2c[x->ref_of_sto[s]];
  // vertices in storage order
  std::vector<int> v2c_ptr(nop + 1, 0), v2c;
  std::vector<R> wt;
  std::vector<double> px(nop), py(nop);
  std::vector<int> ref_of_pt(nop);
  for (int i = 0; i < nop; ++i) ref_of_pt[x->pt_of_ref[i]] = i;
  v2c.reserve(m.v2c.size());
  wt.reserve(m.wt.size());
  for (int s = 0; s < nop; ++s) {
    int i = ref_of_pt[s];
    for (int j = m.v2c_ptr[i]; j < m.v2c_ptr[i + 1]; ++j) {
      v2c.push_back(sto[m.v2c[j]]);
      wt.push_back((R)m.wt[j]);
    }
    v2c_ptr[s + 1] = (int)v2c.size();
    px[s] = m.px[i];
    py[s] = m.py[i];
  }
  // window start of the staged vertex averages: per block of BLK vertices,
  // the lowest interior cell around any of them (noc if none)
  std::vector<int> avg_w0(nblocks(nop), noc);
  for (int s = 0; s < nop; ++s)
    for (int j = v2c_ptr[s]; j < v2c_ptr[s + 1]; ++j)
      if (v2c[j] < noc) avg_w0[s / BLK] = std::min(avg_w0[s / BLK], v2c[j]);
  // cell records: primitives 0 (ghost primitives start at 0, as the
  // reference leaves them), gradients 0, limiter 1, centroid
  std::vector<R> rec((size_t)ncell * RW, R(0));
  for (int c = 0; c < ncell; ++c) {
    R* r = &rec[(size_t)c * RW];
    if (c < noc)
      for (int k = 0; k < 4; ++k) r[R_PHI + k] = R(1);
    r[R_CEN] = (R)m.ccx[x->ref_of_sto[c]];
    r[R_CEN + 1] = (R)m.ccy[x->ref_of_sto[c]];
  }
  // compact copies the gradient pass reads instead of the records
  std::vector<R> phc((size_t)noc * 4, R(1)), cenc((size_t)ncell * 2);
  for (int c = 0; c < ncell; ++c) {
    cenc[(size_t)c * 2] = rec[(size_t)c * RW + R_CEN];
    cenc[(size_t)c * 2 + 1] = rec[(size_t)c * RW + R_CEN + 1];
  }
  std::vector<double> qc((size_t)ncell * 4);
  for (int c = 0; c < ncell; ++c)
    for (int k = 0; k < 4; ++k) qc[(size_t)c * 4 + k] = x->p.qinf[k];
  int rc;
#define UP(dst, src) if ((rc = dupload(x, &d.dst, src))) return rc
#define AL(dst, n) if ((rc = dalloc(x, &d.dst, (size_t)(n)))) return rc
  UP(frec, frec); UP(fcen, fcen); UP(fnl, fnl); UP(phc, phc); UP(cenc, cenc);
  {
    int2* p = nullptr;
    if ((rc = dupload(x, &p, fpt))) return rc;
    d.fpt = p;
  }
  UP(fbc, fbc); UP(bface, bface); UP(bcell, bcell);
  UP(sl_face, sl_face); UP(sl_nbr, sl_nbr); UP(cnf, cnf);
  UP(bf_ptr, sl.bf_ptr); UP(sl_loc, sl.sl_loc);
  d.bf_max = sl.bf_max;
  {
    // the block face lists' cells and face records, in list order
    const size_t nbl = sl.bf_face.size();
    std::vector<int2> bf_ab(std::max<size_t>(1, nbl), make_int2(0, 0));
    std::vector<R> bf_rec(std::max<size_t>(1, nbl) * FW, R(0));
    for (size_t j = 0; j < nbl; ++j) {
      const int f = sl.bf_face[j];
      bf_ab[j] = make_int2(fin[f], fout[f]);
      for (int k = 0; k < FW; ++k) bf_rec[j * FW + k] = frec[(size_t)f * FW + k];
    }
    int2* p = nullptr;
    if ((rc = dupload(x, &p, bf_ab))) return rc;
    d.bf_ab = p;
    UP(bf_rec, bf_rec);
    std::vector<int> bf_fref(std::max<size_t>(1, nbl), 0);
    for (size_t j = 0; j < nbl; ++j) bf_fref[j] = x->ref_of_face[sl.bf_face[j]];
    UP(bf_fref, bf_fref);
  }
  UP(fref, x->ref_of_face);
  UP(r11, r11); UP(r12, r12); UP(r22, r22); UP(cv, cv); UP(cov, cov);
  UP(nc2c, nc2c); UP(cref, x->ref_of_sto);
  UP(v2c_ptr, v2c_ptr); UP(v2c, v2c); UP(wt, wt); UP(px, px); UP(py, py); UP(avg_w0, avg_w0);
  UP(fin, fin); UP(fout, fout); UP(lu_face, lu_face); UP(lu_cnt, lu_cnt);
  UP(rec, rec); UP(qc, qc);
  {
    std::vector<R> qpc((size_t)noc * 4, R(0));
    UP(qpc, qpc);
  }
  AL(fw, nof); AL(qT, ncell); AL(qTc, noc); AL(gT, 2 * (size_t)noc); AL(qold, 4 * (size_t)noc); AL(qs, 4 * (size_t)noc);
  AL(res, 4 * (size_t)noc); AL(dt, noc); AL(dta, noc); AL(la, noc); AL(fla, nof); AL(fmu, nof);
  if (x->p.flow != UFX_FLOW_INVISCID) AL(fvis, 4 * (size_t)nof);
  AL(cmul, ncell); AL(prec, (size_t)nop * PW); AL(fgrad, (size_t)nof * 10);
  AL(bflux, 2 * (size_t)std::max(1, nbf)); AL(bres, 4 * (size_t)std::max(1, nbf));
  AL(part, (size_t)std::max(x->nblk_all, 1) * 8); AL(rpart, (size_t)std::max(x->nblk_all, 1) * 8);
  AL(scal, SC_NSCAL); AL(range, 8); AL(err, 4); AL(lower, 4 * (size_t)noc); AL(upper, 4 * (size_t)noc);
#undef UP
#undef AL
  int init[4] = {0, INT_MAX, INT_MAX, INT_MAX};
  HIPCHK(hipMemcpyAsync(d.err, init, sizeof(init), hipMemcpyHostToDevice, x->stream));
  unsigned long long rg[8] = {0, 0, 0, 0, ~0ULL, ~0ULL, ~0ULL, ~0ULL};
  HIPCHK(hipMemcpyAsync(d.range, rg, sizeof(rg), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return UFX_OK;
}

}  // namespace

extern "C" {

int ufx_cc_destroy(ufx_cc* x) {
  if (!x) return UFX_OK;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  for (void* p : x->allocs) (void)hipFree(p);
  if (x->d_hist) (void)hipFree(x->d_hist);
  (void)release_chunk(x->chunk);
  if (x->d_out) (void)hipFree(x->d_out);
  for (hipEvent_t e : x->ev_pool) (void)hipEventDestroy(e);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return UFX_OK;
}

int ufx_cc_create(const ufx_mesh* mesh, const ufx_cc_params* prm, ufx_cc** out) {
  if (!mesh || !prm || !out) { set_error("ufx_cc_create: null argument"); return UFX_ERR_ARG; }
  const Mesh& m = mesh->m;
  if (m.kind != UFX_KIND_CC) { set_error("ufx_cc_create: not a cell-centred mesh"); return UFX_ERR_ARG; }
  if (prm->flux < UFX_FLUX_ROE || prm->flux > UFX_FLUX_RUSANOV) { set_error("Unknown flux"); return UFX_ERR_ARG; }
  if (prm->grad < UFX_GRAD_LSQR || prm->grad > UFX_GRAD_GGFC) { set_error("Unknown gradient type"); return UFX_ERR_ARG; }
  if (prm->flow_type != UFX_FLOW_INVISCID && prm->flow_type != UFX_FLOW_LAMINAR) {
    set_error("unsupported flow type (the reference's RANS is a stub)");
    return UFX_ERR_UNSUPPORTED;
  }
  if (prm->iwall != 0 && prm->iwall != 1) { set_error("Unknown no-slip wall BC"); return UFX_ERR_ARG; }
  if (prm->timemode != UFX_TIME_RK3 && prm->timemode != UFX_TIME_LUSGS) { set_error("Unknown time integration scheme"); return UFX_ERR_ARG; }
  if (prm->precision != UFX_PREC_FP64 && prm->precision != UFX_PREC_MIXED) { set_error("unknown precision"); return UFX_ERR_ARG; }
  if (prm->irs && !(prm->cfl_max > 0.0)) { set_error("residual smoothing needs cfl_max > 0"); return UFX_ERR_ARG; }
  if (prm->lusgs_order != UFX_LUSGS_REFERENCE && prm->lusgs_order != UFX_LUSGS_MULTICOLOUR) {
    set_error("ufx_cc_create: unknown LU-SGS sweep order");
    return UFX_ERR_ARG;
  }
  const int noc = m.noc, nof = m.nof(), nop = m.nop;
  // faces of each cell in ascending reference order (the order of every
  // per-cell sum)
  std::vector<std::vector<int>> cf(noc);
  for (int f = 0; f < nof; ++f) {
    cf[m.fin[f]].push_back(f);
    if (f < m.nfi) cf[m.fout[f]].push_back(f);
  }
  for (int c = 0; c < noc; ++c) {
    if (cf[c].size() > (size_t)MAXF || m.c2f_ptr[c + 1] - m.c2f_ptr[c] > MAXF) {
      set_error("cell " + std::to_string(c + 1) + " has more than 4 faces");
      return UFX_ERR_UNSUPPORTED;
    }
    std::sort(cf[c].begin(), cf[c].end());
  }
  ufx_cc* x = new ufx_cc();
  // storage order of the interior cells: Hilbert curve through the centroids
  // (compact thread blocks, local neighbour gathers); ghosts keep their
  // reference position after the interior cells
  x->ref_of_sto = hilbert_order(m.ccx.data(), m.ccy.data(), noc);
  x->sto_of_ref.assign(m.ncell(), 0);
  for (int s = 0; s < noc; ++s) x->sto_of_ref[x->ref_of_sto[s]] = s;
  for (int g = noc; g < m.ncell(); ++g) {
    x->ref_of_sto.push_back(g);
    x->sto_of_ref[g] = g;
  }
  const std::vector<int>& sto = x->sto_of_ref;
  // storage order of faces and vertices: first touch walking the cells in
  // storage order and their faces in ascending reference order
  x->face_of_ref.assign(nof, -1);
  x->pt_of_ref.assign(nop, -1);
  {
    int nf = 0, np = 0;
    for (int s = 0; s < noc; ++s)
      for (int f : cf[x->ref_of_sto[s]]) {
        if (x->face_of_ref[f] < 0) x->face_of_ref[f] = nf++;
        for (int p : {m.fp1[f], m.fp2[f]})
          if (x->pt_of_ref[p] < 0) x->pt_of_ref[p] = np++;
      }
    for (int f = 0; f < nof; ++f)
      if (x->face_of_ref[f] < 0) x->face_of_ref[f] = nf++;
    for (int p = 0; p < nop; ++p)
      if (x->pt_of_ref[p] < 0) x->pt_of_ref[p] = np++;
    x->ref_of_face.assign(nof, 0);
    for (int f = 0; f < nof; ++f) x->ref_of_face[x->face_of_ref[f]] = f;
  }
  std::vector<int> sl_face((size_t)MAXF * noc, 0), sl_nbr((size_t)MAXF * noc, 0), cnf(noc);
  for (int s = 0; s < noc; ++s) {
    int c = x->ref_of_sto[s];
    int nb = 0;
    for (size_t j = 0; j < cf[c].size(); ++j) {
      int f = cf[c][j];
      bool in_side = m.fin[f] == c;
      int other = sto[in_side ? m.fout[f] : m.fin[f]];
      sl_face[j * noc + s] = x->face_of_ref[f];
      sl_nbr[j * noc + s] = in_side ? other : ~other;
      if (f >= m.nfi) ++nb;
    }
    cnf[s] = (int)cf[c].size() | (nb << 8);
  }
  // shared-flux residual: per block of BLK storage cells, the interior faces
  // it touches (ascending storage order) and each slot's position among them
  Slots sl;
  {
    const int nblk = nblocks(noc);
    sl.bf_ptr.assign(nblk + 1, 0);
    sl.sl_loc.assign((size_t)MAXF * noc, 0);
    sl.bf_face.reserve((size_t)noc * 2);
    std::vector<int> loc(nof, 0), list;
    for (int b = 0; b < nblk; ++b) {
      int c0 = b * BLK, c1 = std::min(noc, c0 + BLK);
      list.clear();
      for (int s = c0; s < c1; ++s) {
        int ni = (cnf[s] & 0xff) - (cnf[s] >> 8);
        for (int j = 0; j < ni; ++j) list.push_back(sl_face[(size_t)j * noc + s]);
      }
      std::sort(list.begin(), list.end());
      list.erase(std::unique(list.begin(), list.end()), list.end());
      for (size_t i = 0; i < list.size(); ++i) {
        loc[list[i]] = (int)i;
        sl.bf_face.push_back(list[i]);
      }
      sl.bf_ptr[b + 1] = (int)sl.bf_face.size();
      sl.bf_max = std::max(sl.bf_max, (int)list.size());
      for (int s = c0; s < c1; ++s) {
        int ni = (cnf[s] & 0xff) - (cnf[s] >> 8);
        for (int j = 0; j < ni; ++j) {
          size_t q = (size_t)j * noc + s;
          sl.sl_loc[q] = sl_nbr[q] >= 0 ? loc[sl_face[q]] : ~loc[sl_face[q]];
        }
      }
    }
  }
  // LU-SGS: the reference's faces-of-cell order and the level schedule of
  // its sequential sweeps (a cell's level is one more than the highest level
  // among the neighbours it waits for), in the reference's cell numbering
  std::vector<int> lu_face((size_t)MAXF * noc, 0), lu_cnt(noc);
  std::vector<int> lvf(noc, 0), lvb(noc, 0);
  for (int c = 0; c < noc; ++c) {
    int n = m.c2f_ptr[c + 1] - m.c2f_ptr[c];
    lu_cnt[sto[c]] = n;
    for (int j = 0; j < n; ++j)
      lu_face[(size_t)j * noc + sto[c]] = x->face_of_ref[m.c2f[m.c2f_ptr[c] + j]];
  }
  auto nbr_of = [&](int c, int j) {
    int f = m.c2f[m.c2f_ptr[c] + j];
    return m.fout[f] == c ? m.fin[f] : m.fout[f];
  };
  int nlf = 0, nlb = 0;
  for (int c = 0; c < noc; ++c) {
    int l = 0;
    for (int j = 0; j < m.c2f_ptr[c + 1] - m.c2f_ptr[c]; ++j) {
      int o = nbr_of(c, j);
      if (o < c && o >= 0) l = std::max(l, lvf[o] + 1);
    }
    lvf[c] = l;
    nlf = std::max(nlf, l + 1);
  }
  for (int c = noc - 1; c >= 0; --c) {
    int l = 0;
    for (int j = 0; j < m.c2f_ptr[c + 1] - m.c2f_ptr[c]; ++j) {
      int o = nbr_of(c, j);
      if (o > c && o < noc) l = std::max(l, lvb[o] + 1);
    }
    lvb[c] = l;
    nlb = std::max(nlb, l + 1);
  }
  std::vector<int> lv_cells(2 * (size_t)noc);  // storage indices, by level
  {
    std::vector<int> cnt(nlf + 1, 0);
    for (int c = 0; c < noc; ++c) cnt[lvf[c] + 1]++;
    for (int l = 0; l < nlf; ++l) cnt[l + 1] += cnt[l];
    x->lv_fwd_off = cnt;
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int c = 0; c < noc; ++c) lv_cells[pos[lvf[c]]++] = sto[c];
    std::vector<int> cb(nlb + 1, 0);
    for (int c = 0; c < noc; ++c) cb[lvb[c] + 1]++;
    for (int l = 0; l < nlb; ++l) cb[l + 1] += cb[l];
    for (int l = 0; l <= nlb; ++l) cb[l] += noc;
    x->lv_bwd_off = cb;
    std::vector<int> pb(cb.begin(), cb.end() - 1);
    for (int c = noc - 1; c >= 0; --c) lv_cells[pb[lvb[c]]++] = sto[c];
  }
  hipError_t e = hipGetDevice(&x->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    delete x;
    return UFX_ERR_HIP;
  }
  x->nop = nop; x->noc = noc; x->nfi = m.nfi; x->nbf = m.nbf; x->nof = nof; x->ncell = m.ncell();
  x->nblk_cell = nblocks(noc);
  x->nblk_all = nblocks(m.ncell());
  x->nblk_pt = nblocks(nop);
  x->nblk_face = nblocks(nof);
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
  p.qinf[1] = 1.0;
  p.qinf[2] = std::cos(aoa) * 1.0;
  p.qinf[3] = std::sin(aoa) * 1.0;
  p.p_inf = 1.0 / (p.gamma * p.m_inf * p.m_inf);
  p.qinf[0] = p.p_inf / (p.gamma - 1.0) + 0.5;
  p.r_inf = p.qinf[1];
  p.u_inf = p.qinf[2] / p.qinf[1];
  p.v_inf = p.qinf[3] / p.qinf[1];
  p.znd = 1.0 / (p.gamma * p.m_inf * p.m_inf);
  p.mref = std::max((double)0.32f, 0.5 * p.m_inf);
  p.t_wall = prm->t_wall / 300.0;
  p.rey = prm->reynolds;
  p.flow = prm->flow_type; p.iwall = prm->iwall; p.flux = prm->flux; p.grad = prm->grad;
  p.ilimit = prm->ilimit; p.irs = prm->irs; p.timemode = prm->timemode;
  p.precision = prm->precision; p.cfl_max = prm->cfl_max;
  p.lusgs_order = prm->lusgs_order;
  p.vll = 0.0;
  if (p.flow != UFX_FLOW_INVISCID) p.vll = 2.0 * p.gamma / (0.72 * p.rey * 0.25);
  x->g = Gas<double>{p.gamma, p.gamma1, p.znd, p.r_inf, p.u_inf, p.v_inf, p.p_inf, p.mref,
                     1.0 / p.gamma1};
  x->gf = Gas<float>{(float)p.gamma, (float)p.gamma1, (float)p.znd, (float)p.r_inf,
                     (float)p.u_inf, (float)p.v_inf, (float)p.p_inf, (float)p.mref, (float)(1.0 / p.gamma1)};
  // residual stage form (measured on MI355X at 2M cells): FP64, the
  // per-cell gather with the block's records staged in LDS; mixed, the
  // shared-flux kernel reading records through L2 (its 80-byte records cache
  // well, and staging would cost it occupancy).  UFX_STAGE_KERNEL=
  // gather|shared and UFX_STAGE_LDS=0|1 override.  LDS per block: records
  // BLK x RW, and for the shared form the block's flux list.
  {
    size_t w = p.precision == UFX_PREC_FP64 ? sizeof(double) : sizeof(float);
    size_t rw = p.precision == UFX_PREC_FP64 ? Layout<double>::RW : Layout<float>::RW;
    const char* env = std::getenv("UFX_STAGE_KERNEL");
    bool shared = p.precision == UFX_PREC_MIXED;
    if (env && std::strcmp(env, "gather") == 0) shared = false;
    if (env && std::strcmp(env, "shared") == 0) shared = true;
    // UFX_ITER_FUSION=0: whole iterations refresh the primitives in a pass
    // of their own (same results; for comparison)
    if (const char* r = std::getenv("UFX_SERPENTINE"); r && std::strcmp(r, "0") == 0)
      x->stage_rev = 0;
    const char* enve = std::getenv("UFX_ITER_FUSION");
    x->iter_fusion = !(enve && std::strcmp(enve, "0") == 0);
    const char* envl = std::getenv("UFX_STAGE_LDS");
    bool local = p.precision == UFX_PREC_FP64;
    if (envl && std::strcmp(envl, "0") == 0) local = false;
    if (envl && std::strcmp(envl, "1") == 0) local = true;
    size_t lds_fl = (size_t)sl.bf_max * 4 * w, lds_rec = (size_t)BLK * rw * w;
    x->lds_list = lds_fl;
    x->lds_recs = lds_rec;
    x->lds_overlay = sl.bf_max <= 2 * BLK;
    if (shared && stage_lds_need(x, shared, local) > STAGE_LDS_MAX) shared = false;
    x->stage_shared = shared;
    x->stage_local = local;
    x->stage_lds = stage_lds_need(x, shared, local);
  }
  int rc = p.precision == UFX_PREC_FP64
               ? build_dev(x, x->d64, m, sl_face, sl_nbr, cnf, lu_face, lu_cnt, sl)
               : build_dev(x, x->d32, m, sl_face, sl_nbr, cnf, lu_face, lu_cnt, sl);
  if (!rc) rc = dupload(x, &x->lv_cells, lv_cells);
  if (!rc && p.timemode == UFX_TIME_LUSGS && p.lusgs_order == UFX_LUSGS_MULTICOLOUR) {
    // colours of the interior cells (reference numbering) -> storage order;
    // each colour's cells in ascending storage order (compact blocks)
    int ncol = 0;
    std::vector<int> col = lusgs_colours(m, &ncol), ccol(noc);
    for (int c = 0; c < noc; ++c) ccol[sto[c]] = col[c];
    x->mc_off.assign(ncol + 1, 0);
    for (int s2 = 0; s2 < noc; ++s2) x->mc_off[ccol[s2] + 1]++;
    for (int c = 0; c < ncol; ++c) x->mc_off[c + 1] += x->mc_off[c];
    std::vector<int> cells(std::max(1, noc)), at(x->mc_off.begin(), x->mc_off.end() - 1);
    for (int s2 = 0; s2 < noc; ++s2) cells[at[ccol[s2]]++] = s2;
    rc = dupload(x, &x->mc_cells, cells);
    const int* dc = nullptr;
    if (!rc) rc = dupload(x, &dc, ccol);
    x->d64.ccol = dc;
    x->d32.ccol = dc;
  }
  if (!rc) rc = dalloc(x, &x->cirs_tmp, 4 * (size_t)noc);
  if (!rc) rc = dalloc(x, &x->bfluxv, 2 * (size_t)std::max(1, m.nbf));
  if (!rc) {
    e = hipStreamSynchronize(x->stream);
    if (e != hipSuccess) { set_error(std::string("upload: ") + hipGetErrorString(e)); rc = UFX_ERR_HIP; }
  }
  if (rc) {
    ufx_cc_destroy(x);
    return rc;
  }
  for (int k = 0; k < m.nbf; ++k)
    if (m.fbc[m.nfi + k] == WALL) x->wall_k.push_back(k);
  x->fin_h.resize(m.nbf); x->fp1_h.resize(m.nbf); x->fp2_h.resize(m.nbf);
  x->fnx_h.resize(m.nbf); x->fny_h.resize(m.nbf); x->farea_h.resize(m.nbf);
  for (int k = 0; k < m.nbf; ++k) {
    int f = m.nfi + k;
    x->fin_h[k] = m.fin[f]; x->fp1_h[k] = m.fp1[f]; x->fp2_h[k] = m.fp2[f];
    x->fnx_h[k] = m.fnx[f]; x->fny_h[k] = m.fny[f]; x->farea_h[k] = m.farea[f];
  }
  x->px_h = m.px;
  x->py_h = m.py;
  x->fin_ref = m.fin;
  *out = x;
  return UFX_OK;
}

int ufx_cc_set(ufx_cc* x, const char* name, const double* src) {
  if (int rc = check_routine(x)) return rc;
  if (!name || !src) { set_error("ufx_cc_set: null argument"); return UFX_ERR_ARG; }
  std::string n(name);
  int rc = with_dev(x, [&](auto& d) { return set_field(x, d, n, src); });
  if (rc == UFX_OK && n == "qold") x->qold_stale = false;
  return rc;
}

int ufx_cc_get(ufx_cc* x, const char* name, double* dst) {
  if (int rc = check_ctx(x)) return rc;
  if (!name || !dst) { set_error("ufx_cc_get: null argument"); return UFX_ERR_ARG; }
  std::string n(name);
  if (n == "qold")
    if (int rc = check_qold(x, "ufx_cc_get(\"qold\")")) return rc;
  return with_dev(x, [&](auto& d) { return get_field(x, d, n, dst); });
}

int ufx_cc_prim(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  return with_dev(x, [&](auto& d) { return run_prim(x, d, 1, 0, 0); });
}

int ufx_cc_save_old(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  x->qold_stale = false;
  return with_dev(x, [&](auto& d) { return run_prim(x, d, 0, 1, 0); });
}

int ufx_cc_avg_c2v(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  return with_dev(x, [&](auto& d) { return run_avg(x, d); });
}

int ufx_cc_time_step(ufx_cc* x, double cfl) {
  if (int rc = check_routine(x)) return rc;
  x->cfl = cfl;
  return with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    return run_cell_g<R, G_NONE>(x, d, false, false, true, cfl);
  });
}

int ufx_cc_gradient(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  return with_dev(x, [&](auto& d) {
    if (x->p.ilimit)
      if (int rc = run_range(x, d)) return rc;
    return run_gradient(x, d, true, x->p.ilimit != 0, false, 0.0);
  });
}

int ufx_cc_limit(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  return with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    if (int rc = run_range(x, d)) return rc;
    return run_cell_g<R, G_NONE>(x, d, false, true, false, 0.0);
  });
}

int ufx_cc_fvresidual(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  return with_dev(x, [&](auto& d) { return run_residual_full(x, d); });
}

int ufx_cc_rk_update(ufx_cc* x, int stage) {
  if (int rc = check_routine(x)) return rc;
  if (stage < 1 || stage > 3) { set_error("stage must be 1..3"); return UFX_ERR_ARG; }
  if (int rc = check_qold(x, "ufx_cc_rk_update")) return rc;
  return with_dev(x, [&](auto& d) { return run_rk_update(x, d, RK_A[stage - 1], RK_B[stage - 1]); });
}

int ufx_cc_rk_stage(ufx_cc* x, int stage) {
  if (int rc = check_routine(x)) return rc;
  if (stage < 1 || stage > 3) { set_error("stage must be 1..3"); return UFX_ERR_ARG; }
  if (int rc = check_qold(x, "ufx_cc_rk_stage")) return rc;
  return with_dev(x, [&](auto& d) { return run_stage(x, d, stage, false, true, false); });
}

int ufx_cc_lusgs(ufx_cc* x) {
  if (int rc = check_routine(x)) return rc;
  if (int rc = check_qold(x, "ufx_cc_lusgs")) return rc;
  return with_dev(x, [&](auto& d) {
    if (int rc = restore_dtla(x, d)) return rc;
    return run_lusgs(x, d);
  });
}

int ufx_cc_residue(ufx_cc* x, int iter, double* fres, double* fres_raw) {
  if (int rc = check_routine(x)) return rc;
  if (int rc = ensure_hist(x, 1)) return rc;
  int rc = with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    LAUNCH(x, UFX_K_NORM, k_sumsq<R>, x->nblk_cell, d);
    LAUNCH(x, UFX_K_NORM, k_norm_final<R>, 1, d, x->nblk_cell, iter == 1, x->d_hist, 0);
    return launch_check();
  });
  if (rc) return rc;
  double h = 0.0, raw = 0.0;
  const double* scal = x->p.precision == UFX_PREC_FP64 ? x->d64.scal : x->d32.scal;
  HIPCHK(hipMemcpyAsync(&h, x->d_hist, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(&raw, scal + SC_NORM, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (fres) *fres = h;
  if (fres_raw) *fres_raw = raw;
  return device_error(x);
}

int ufx_cc_check_positivity(ufx_cc* x, int* bad) {
  if (int rc = check_ctx(x)) return rc;
  int* err = err_word(x);
  int big = INT_MAX, held = INT_MAX;
  HIPCHK(hipMemcpyAsync(&held, err + 2, sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpyAsync(err + 2, &big, sizeof(int), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  int rc = with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    LAUNCH(x, UFX_K_OTHER, k_positivity<R>, x->nblk_cell, d);
    return launch_check();
  });
  if (rc) return rc;
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, err + 2, sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (bad) *bad = e == INT_MAX ? 0 : e;
  // a report only: the cell noted here must not become an iteration's error
  // (the note an iteration may have taken ahead is put back)
  HIPCHK(hipMemcpyAsync(err + 2, &held, sizeof(int), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (e != INT_MAX) {
    set_error("Density/pressure is negative at point " + std::to_string(e));
    return UFX_ERR_STATE;
  }
  return UFX_OK;
}

int ufx_cc_forces(ufx_cc* x, double* out, double* cp, double* cf) {
  if (int rc = check_ctx(x)) return rc;
  if (!out) { set_error("ufx_cc_forces: null output"); return UFX_ERR_ARG; }
  const Params& p = x->p;
  const int nbf = x->nbf;
  std::vector<double> bf(2 * (size_t)std::max(1, nbf)), bv(2 * (size_t)std::max(1, nbf));
  std::vector<double> pin(std::max(1, nbf)), dtmin(1);
  int rc = with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    constexpr int RW = Layout<R>::RW;
    if (int r = restore_dtla(x, d)) return r;
    LAUNCH(x, UFX_K_OTHER, k_dtmin<R>, 1, d);
    if (int r = launch_check()) return r;
    std::vector<R> rec((size_t)x->ncell * RW);
    HIPCHK(hipMemcpyAsync(rec.data(), d.rec, rec.size() * sizeof(R), hipMemcpyDeviceToHost, x->stream));
    if (nbf > 0) {
      HIPCHK(hipMemcpyAsync(bf.data(), d.bflux, 2 * (size_t)nbf * 8, hipMemcpyDeviceToHost, x->stream));
      HIPCHK(hipMemcpyAsync(bv.data(), x->bfluxv, 2 * (size_t)nbf * 8, hipMemcpyDeviceToHost, x->stream));
    }
    HIPCHK(hipMemcpyAsync(dtmin.data(), d.scal + SC_DTMIN, 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int k = 0; k < nbf; ++k) pin[k] = rec[(size_t)x->sto_of_ref[x->fin_h[k]] * RW + R_QP + 3];
    return (int)UFX_OK;
  });
  if (rc) return rc;
  std::vector<double> fg, fmu;
  bool visc = p.flow != UFX_FLOW_INVISCID;
  if (visc) {
    fg.resize((size_t)x->nof * 10);
    fmu.resize(x->nof);
    if ((rc = ufx_cc_get(x, "fc_grad", fg.data()))) return rc;
    if ((rc = ufx_cc_get(x, "fc_mu", fmu.data()))) return rc;
  }
  // find_cl_cd, CCFVM/main.F90:263-355: pressure of the adjacent cell for
  // the surface integral; the momentum fluxes of the last residual for the
  // coefficients the reference prints (:351-353)
  double dyna = 0.5 * p.gamma * p.p_inf * p.m_inf * p.m_inf;
  double fx = 0.0, fy = 0.0, cm = 0.0, fx1 = 0.0, fy1 = 0.0;
  const double two = 2.0, two_third = two / 3.0;
  int iw = 0;
  for (int k : x->wall_k) {
    double pwall = pin[k];
    double xi = 0.5 * (x->px_h[x->fp1_h[k]] + x->px_h[x->fp2_h[k]]);
    double yi = 0.5 * (x->py_h[x->fp1_h[k]] + x->py_h[x->fp2_h[k]]);
    double xa = xi - p.xref, ya = yi - p.yref;
    double area = x->farea_h[k], nx = x->fnx_h[k], ny = x->fny_h[k];
    double cpv = (pwall - p.p_inf) / dyna;
    double dcx = cpv * nx * area, dcy = cpv * ny * area;
    fx = fx + pwall * nx * area;
    fy = fy + pwall * ny * area;
    double cfv = 0.0;
    if (visc) {
      int f = x->nfi + k;
      double mu = fmu[f] / p.rey;
      double ux = fg[(size_t)f * 10 + 2], uy = fg[(size_t)f * 10 + 3];
      double vx = fg[(size_t)f * 10 + 4], vy = fg[(size_t)f * 10 + 5];
      double tauxx = two_third * (two * ux - vy), tauxy = uy + vx, tauyy = two_third * (two * vy - ux);
      fx = fx - mu * (tauxx * nx + tauxy * ny) * area;
      fy = fy - mu * (tauxy * nx + tauyy * ny) * area;
      double tx = ny, ty = -nx;
      double tw = mu * ((ux * tx + vx * ty) * nx + (uy * tx + vy * ty) * ny);
      cfv = tw * area / dyna;
    }
    cm = cm + dcx * ya - dcy * xa;
    if (cp) cp[iw] = -cpv;
    if (cf) cf[iw] = cfv;
    ++iw;
  }
  for (int k : x->wall_k) { fx1 = fx1 + bf[k]; fy1 = fy1 + bf[nbf + k]; }
  if (visc)
    for (int k : x->wall_k) { fx1 = fx1 + bv[k]; fy1 = fy1 + bv[nbf + k]; }
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
  for (int k = 0; k < niter; ++k) {
    x->cfl = cfl[k];
    int rc = with_dev(x, [&](auto& d) { return run_iteration(x, d, iter0 + 1 + k, k); });
    if (rc) return rc;
  }
  if (hist && niter > 0) {
    HIPCHK(hipMemcpyAsync(hist, x->d_hist, (size_t)niter * 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return device_error(x);
  }
  return UFX_OK;
}

int ufx_cc_iterate_graph(ufx_cc* x, int iter0, int niter, double cfl, double* hist) {
  if (int rc = check_ctx(x)) return rc;
  if (niter <= 0 || iter0 < 1) {
    set_error("ufx_cc_iterate_graph: needs niter >= 1 and iter0 >= 1 (iteration 1 sets the "
              "residual normalisation)");
    return UFX_ERR_ARG;
  }
  if (x->prof) { set_error("ufx_cc_iterate_graph: not while profiling"); return UFX_ERR_ARG; }
  if (int rc = ensure_hist(x, niter)) return rc;
  // the launches of a chunk depend only on the CFL and the chunk length:
  // capture once, replay while both stay the same; a shorter chunk than the
  // captured one (the remainder of a run) is launched directly instead of
  // being captured anew
  x->cfl = cfl;
  const bool fresh = !x->chunk.exec || x->chunk.stale || x->chunk.cfl != cfl || niter > x->chunk.n;
  if (!fresh && niter < x->chunk.n) {
    for (int k = 0; k < niter; ++k) {
      int rc = with_dev(x, [&](auto& d) { return run_iteration(x, d, 2, k); });
      if (rc) return rc;
    }
  } else {
    if (fresh) {
      // the chunk's first iteration refreshes the primitives itself, so a
      // replay does not depend on what ran before it
      x->prims_next = false;
      HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
      int rc = 0;
      for (int k = 0; k < niter && !rc; ++k)
        rc = with_dev(x, [&](auto& d) { return run_iteration(x, d, 2, k); });
      Chunk c;
      hipError_t e = hipStreamEndCapture(x->stream, &c.graph);
      if (!rc && e == hipSuccess) e = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
      if (rc || e != hipSuccess) {
        (void)release_chunk(c);
        if (rc) return rc;
        HIPCHK(e);
      }
      c.n = niter;
      c.cfl = cfl;
      c.stale = false;
      c.prims_next = x->prims_next;
      HIPCHK(release_chunk(x->chunk));
      x->chunk = c;
    }
    HIPCHK(hipGraphLaunch(x->chunk.exec, x->stream));
    x->prims_next = x->chunk.prims_next;
  }
  (void)iter0;
  x->qold_stale = x->p.timemode == UFX_TIME_RK3 && x->p.flow == UFX_FLOW_INVISCID && !x->p.irs;
  // replayed iterations store no la / dt either (their captured time-step
  // pass ran at the chunk's CFL)
  x->dtla_pending = x->qold_stale;
  x->dtla_cfl = cfl;
  if (hist) {
    HIPCHK(hipMemcpyAsync(hist, x->d_hist, (size_t)niter * 8, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return device_error(x);
  }
  return UFX_OK;
}

int ufx_cc_history(ufx_cc* x, int n, double* out) {
  if (int rc = check_ctx(x)) return rc;
  if (n < 0 || n > x->hist_cap || (n > 0 && !out)) {
    set_error("ufx_cc_history: n must be 0.." + std::to_string(x->hist_cap));
    return UFX_ERR_ARG;
  }
  if (n > 0)
    HIPCHK(hipMemcpyAsync(out, x->d_hist, (size_t)n * 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return device_error(x);
}

int ufx_cc_set_stage_form(ufx_cc* x, int shared, int lds) {
  if (int rc = check_ctx(x)) return rc;
  size_t need = stage_lds_need(x, shared != 0, lds != 0);
  if (need > STAGE_LDS_MAX) {
    set_error("ufx_cc_set_stage_form: the block-shared form needs " + std::to_string(need) +
              " bytes of LDS per block on this mesh, more than " +
              std::to_string(STAGE_LDS_MAX));
    return UFX_ERR_UNSUPPORTED;
  }
  if ((x->stage_shared != (shared != 0)) || (x->stage_local != (lds != 0))) x->chunk.stale = true;
  x->stage_shared = shared != 0;
  x->stage_local = lds != 0;
  x->stage_lds = need;
  return UFX_OK;
}

int ufx_cc_get_stage_form(ufx_cc* x, int* shared, int* lds) {
  if (int rc = check_ctx(x)) return rc;
  if (!shared || !lds) { set_error("ufx_cc_get_stage_form: null output"); return UFX_ERR_ARG; }
  *shared = x->stage_shared ? 1 : 0;
  *lds = x->stage_local ? 1 : 0;
  return UFX_OK;
}

int ufx_cc_keep_residual(ufx_cc* x, int on) {
  if (int rc = check_ctx(x)) return rc;
  if (x->keep_res != (on != 0)) x->chunk.stale = true;
  x->keep_res = on != 0;
  return UFX_OK;
}

int ufx_cc_residual_norm(ufx_cc* x, int var, double* out) {
  if (int rc = check_ctx(x)) return rc;
  if (var < -1 || var > 3 || !out) { set_error("ufx_cc_residual_norm: var must be -1..3"); return UFX_ERR_ARG; }
  if (!x->d_out) HIPCHK(hipMalloc((void**)&x->d_out, 8 * sizeof(double)));
  int rc = with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    LAUNCH(x, UFX_K_NORM, k_sumsq_var<R>, x->nblk_cell, d, var, d.part);
    LAUNCH(x, UFX_K_NORM, k_norm_var_final<R>, 1, d, x->nblk_cell, var, d.part, x->d_out);
    return launch_check();
  });
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out, x->d_out, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return device_error(x);
}

// ---- residual norm without allocation
// Scratch of ufx_cc_residual_norm_noalloc: the result slot at [0], the
// block partials from [32] on.  Every GPU the library runs on holds its own
// copy until the process ends; in host code the array's name is only a
// handle, whose device address (per device) comes from hipGetSymbolAddress.
constexpr int NOALLOC_DOUBLES = UFX_NOALLOC_SCRATCH_BYTES / 8, NOALLOC_PART = 32;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on device dev, looked up once per device
// with that device current.
int scratch_of(int dev, double** out) {
  if (dev < 0 || dev >= MAX_DEVICES) {
    set_error("ufx_cc_residual_norm_noalloc: device number out of range");
    return UFX_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    HIPCHK(e);
    scratch_ptr[dev] = (double*)p;
  }
  *out = scratch_ptr[dev];
  return UFX_OK;
}
}  // namespace

int ufx_cc_residual_norm_noalloc(ufx_cc* x, int var, double* out) {
  if (int rc = check_ctx(x)) return rc;
  if (var < -1 || var > 3 || !out) {
    set_error("ufx_cc_residual_norm_noalloc: var must be -1..3 and out not NULL");
    return UFX_ERR_ARG;
  }
  if (x->nblk_cell > NOALLOC_DOUBLES - NOALLOC_PART) {
    set_error("ufx_cc_residual_norm_noalloc: " + std::to_string(x->noc) + " cells need " +
              std::to_string(x->nblk_cell) + " partial sums, more than the " +
              std::to_string(NOALLOC_DOUBLES - NOALLOC_PART) + " the scratch holds (" +
              std::to_string(UFX_NOALLOC_SCRATCH_BYTES) + " bytes)");
    return UFX_ERR_ARG;
  }
  double* ptr = nullptr;
  if (int rc = scratch_of(x->device, &ptr)) return rc;
  // every partial slot and the result are written before they are read: no
  // clearing needed
  int rc = with_dev(x, [&](auto& d) {
    using R = typename std::remove_reference_t<decltype(d)>::Real;
    // plain launches: the timing instrumentation would create events
    k_sumsq_var<R><<<x->nblk_cell, BLK, 0, x->stream>>>(d, var, ptr + NOALLOC_PART);
    k_norm_var_final<R><<<1, BLK, 0, x->stream>>>(d, x->nblk_cell, var, ptr + NOALLOC_PART, ptr);
    return launch_check();
  });
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out, ptr, 8, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return device_error(x);
}

int ufx_cc_profile_mask(ufx_cc* x, unsigned mask) {
  if (int rc = check_ctx(x)) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  x->prof = mask != 0;
  x->prof_mask = mask;
  x->recs.clear();
  x->ev_used = 0;
  return UFX_OK;
}

int ufx_cc_profile(ufx_cc* x, int enable) {
  if (int rc = check_ctx(x)) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  x->prof = enable != 0;
  x->prof_mask = ~0u;
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

int ufx_cc_error(ufx_cc* x, int* kind, int* face, int* cell) {
  if (int rc = check_ctx(x)) return rc;
  if (!kind || !face || !cell) { set_error("ufx_cc_error: null output"); return UFX_ERR_ARG; }
  int e[4] = {0, 0, 0, 0};
  if (int rc = read_error(x, e)) return rc;
  decode_error(x, e, kind, face, cell);
  return UFX_OK;
}

void* ufx_cc_stream(ufx_cc* x) { return x ? (void*)x->stream : nullptr; }

}  // extern "C"

// ---- many cases per launch (the batched-cases scheduler)

struct ufx_cc_batch {
  std::vector<ufx_cc*> cs;
  std::vector<char> on;
  std::vector<double*> hist_seen;  // each active case's history buffer in the descriptors
  bool f64 = true;
  hipStream_t stream = nullptr;
  void* d_desc = nullptr;          // CaseDesc<R>[n]
  int4* d_map = nullptr;           // the block tables, one after another
  size_t map_cap = 0;
  size_t off[6] = {0, 0, 0, 0, 0, 0}, cnt[6] = {0, 0, 0, 0, 0, 0};
  int* d_ids = nullptr;            // active cases
  int* d_err = nullptr;            // their error words
  double* d_norm = nullptr;        // per-case residual norms
  int nids = 0;
  size_t stage_lds = 0;
  Chunk chunk;                     // a captured chunk of batched iterations
  int chunk_form = -1;             // stage form the chunk was captured with
  bool dirty = true;               // descriptors or tables to rewrite
};

namespace {

// block tables: all cells (incl. ghosts), vertices, faces, interior cells,
// boundary faces, one block per case
enum { T_ALL, T_PT, T_FACE, T_CELL, T_BF, T_CASE, T_N };

template <typename R> DevT<R>& dev_of(ufx_cc* x);
template <> DevT<double>& dev_of<double>(ufx_cc* x) { return x->d64; }
template <> DevT<float>& dev_of<float>(ufx_cc* x) { return x->d32; }

// Descriptors of the active cases and the block tables over them.  Inactive
// cases are not touched (they may have been destroyed).
template <typename R>
int batch_write(ufx_cc_batch* b) {
  const int n = (int)b->cs.size();
  HIPCHK(hipStreamSynchronize(b->stream));
  // the captured chunk's launches describe the old tables: never replayed
  // again, released once its successor is instantiated
  b->chunk.stale = true;
  std::vector<CaseDesc<R>> desc(n);
  std::memset((void*)desc.data(), 0, desc.size() * sizeof(CaseDesc<R>));
  std::vector<int4> map;
  std::vector<int> ids;
  for (int i = 0; i < n; ++i) {
    if (!b->on[i]) continue;
    ufx_cc* x = b->cs[i];
    CaseDesc<R>& e = desc[i];
    e.d = dev_of<R>(x);
    e.g = x->g;
    e.gr = gas_of<R>(x);
    e.vll = x->p.vll;
    e.t_wall = x->p.t_wall;
    e.hist = x->d_hist;
    e.iwall = x->p.iwall;
    e.nblk_all = x->nblk_all;
    e.nblk_pt = x->nblk_pt;
    e.nblk_face = x->nblk_face;
    e.nblk_cell = x->nblk_cell;
    e.nblk_bf = x->nbf > 0 ? nblocks(x->nbf) : 0;
    b->hist_seen[i] = x->d_hist;
    ids.push_back(i);
  }
  for (int t = 0; t < T_N; ++t) {
    b->off[t] = map.size();
    for (int i : ids) {
      const CaseDesc<R>& e = desc[i];
      const int nb = t == T_ALL ? e.nblk_all : t == T_PT ? e.nblk_pt : t == T_FACE ? e.nblk_face
                   : t == T_CELL ? e.nblk_cell : t == T_BF ? e.nblk_bf : 1;
      for (int k = 0; k < nb; ++k) map.push_back(make_int4(i, k, nb, 0));
    }
    b->cnt[t] = map.size() - b->off[t];
  }
  if (!b->d_desc) HIPCHK(hipMalloc(&b->d_desc, (size_t)n * sizeof(CaseDesc<R>)));
  if (!b->d_ids) HIPCHK(hipMalloc((void**)&b->d_ids, (size_t)n * sizeof(int)));
  if (!b->d_err) HIPCHK(hipMalloc((void**)&b->d_err, (size_t)3 * n * sizeof(int)));
  if (!b->d_norm) HIPCHK(hipMalloc((void**)&b->d_norm, (size_t)n * sizeof(double)));
  if (map.size() > b->map_cap) {
    if (b->d_map) HIPCHK(hipFree(b->d_map));
    b->d_map = nullptr;
    HIPCHK(hipMalloc((void**)&b->d_map, map.size() * sizeof(int4)));
    b->map_cap = map.size();
  }
  HIPCHK(hipMemcpyAsync(b->d_desc, desc.data(), desc.size() * sizeof(CaseDesc<R>),
                        hipMemcpyHostToDevice, b->stream));
  if (!map.empty())
    HIPCHK(hipMemcpyAsync(b->d_map, map.data(), map.size() * sizeof(int4), hipMemcpyHostToDevice,
                          b->stream));
  if (!ids.empty())
    HIPCHK(hipMemcpyAsync(b->d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice,
                          b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->nids = (int)ids.size();
  b->dirty = false;
  return UFX_OK;
}

// One inviscid RK3 iteration of every active case, as run_iteration does it
// for one context (the first active case's options hold for all): the
// primitive refresh only when `refresh` (the first iteration of a call), the
// last stage refreshing them for the next iteration.
template <typename R>
int batch_iteration(ufx_cc_batch* b, ufx_cc* x0, int k, double cfl, int keep, bool refresh) {
  const CaseDesc<R>* cs = (const CaseDesc<R>*)b->d_desc;
  const hipStream_t s = b->stream;
  auto M = [&](int t) { return (const int4*)(b->d_map + b->off[t]); };
  auto G = [&](int t) { return dim3((unsigned)b->cnt[t]); };
  const bool lim = x0->p.ilimit != 0;
  if (refresh) km_prim<R><<<G(T_ALL), BLK, 0, s>>>(cs, M(T_ALL), 1, 2, lim ? 1 : 0);
  if (lim) km_range_final<R><<<G(T_CASE), BLK, 0, s>>>(cs, M(T_CASE), refresh ? 0 : 1);
  km_avg_c2v<R, false><<<G(T_PT), BLK, 0, s>>>(cs, M(T_PT));
  km_face_lambda<R><<<G(T_FACE), BLK, 0, s>>>(cs, M(T_FACE));
#define KMC(GR)                                                                            \
  do {                                                                                     \
    if (lim) km_cell<R, GR, false, true, true><<<G(T_CELL), BLK, 0, s>>>(cs, M(T_CELL), cfl); \
    else km_cell<R, GR, false, false, true><<<G(T_CELL), BLK, 0, s>>>(cs, M(T_CELL), cfl);   \
  } while (0)
  if (x0->p.grad == UFX_GRAD_GG) KMC(G_GG);
  else KMC(G_LSQR);
#undef KMC
  for (int st = 1; st <= 3; ++st) {
    const double ak = RK_A[st - 1], bk = RK_B[st - 1];
    const int mode = (st == 3 ? 2 : 1) | (st << 2) | ((x0->stage_rev >> st) & 1 ? 16 : 0),
              wr = st == 3 && keep ? 1 : 0;
    const int epi = st == 3 ? 1 : 0;
    const size_t lds = b->stage_lds;
#define KMS(F)                                                                                  \
  do {                                                                                          \
    if (b->cnt[T_BF]) km_bface<R, F><<<G(T_BF), BLK, 0, s>>>(cs, M(T_BF), 0, st);               \
    if (x0->stage_shared && x0->stage_local)                                                    \
      km_stage_blk<R, F, true><<<G(T_CELL), BLK, lds, s>>>(cs, M(T_CELL), mode, ak, bk, wr,     \
                                                           epi);                                \
    else if (x0->stage_shared)                                                                  \
      km_stage_blk<R, F, false><<<G(T_CELL), BLK, lds, s>>>(cs, M(T_CELL), mode, ak, bk, wr,    \
                                                            epi);                               \
    else if (x0->stage_local)                                                                   \
      km_stage<R, F, true><<<G(T_CELL), BLK, lds, s>>>(cs, M(T_CELL), mode, ak, bk, wr, epi);   \
    else                                                                                        \
      km_stage<R, F, false><<<G(T_CELL), BLK, 0, s>>>(cs, M(T_CELL), mode, ak, bk, wr, epi);    \
  } while (0)
    switch (x0->p.flux) {
      case UFX_FLUX_ROE: KMS(1); break;
      case UFX_FLUX_AUSM: KMS(2); break;
      case UFX_FLUX_VANLEER: KMS(3); break;
      default: KMS(4); break;
    }
#undef KMS
  }
  km_norm_final<R><<<G(T_CASE), BLK, 0, s>>>(cs, M(T_CASE), k);
  return launch_check();
}

template <typename R>
int batch_iterate(ufx_cc_batch* b, int niter, double cfl) {
  ufx_cc* x0 = nullptr;
  int keep = -1;
  for (size_t i = 0; i < b->cs.size(); ++i) {
    if (!b->on[i]) continue;
    ufx_cc* x = b->cs[i];
    if (!x0) x0 = x;
    if (keep < 0) keep = x->keep_res ? 1 : 0;
    if (keep != (x->keep_res ? 1 : 0)) {
      set_error("ufx_cc_batch_iterate: the active cases differ in ufx_cc_keep_residual");
      return UFX_ERR_ARG;
    }
    if (x->prof) { set_error("ufx_cc_batch_iterate: a case is profiling"); return UFX_ERR_ARG; }
    // the batch's first iteration refreshes the primitives from the state
    // (and notes their positivity again), whatever the context held
    x->prims_next = false;
    // the case's own queued work first
    HIPCHK(hipStreamSynchronize(x->stream));
    if (int rc = ensure_hist(x, niter)) return rc;
    if (x->d_hist != b->hist_seen[i]) b->dirty = true;
  }
  if (!x0) return UFX_OK;
  // the residual-stage form may have been changed (ufx_cc_set_stage_form):
  // all cases launch as one, so they must agree
  size_t lds = 0;
  for (size_t i = 0; i < b->cs.size(); ++i) {
    if (!b->on[i]) continue;
    ufx_cc* x = b->cs[i];
    if (x->stage_shared != x0->stage_shared || x->stage_local != x0->stage_local) {
      set_error("ufx_cc_batch_iterate: the active cases differ in their stage form");
      return UFX_ERR_ARG;
    }
    lds = std::max(lds, x->stage_lds);
  }
  const int form = (x0->stage_shared ? 2 : 0) | (x0->stage_local ? 1 : 0);
  if (form != b->chunk_form || lds != b->stage_lds) b->chunk.stale = true;
  b->stage_lds = lds;
  if (b->dirty)
    if (int rc = batch_write<R>(b)) return rc;
  // capture a chunk when there is none that fits (first call, new tables or
  // stage form, another CFL or residual option, a longer chunk); a shorter
  // chunk than the captured one (the remainder of a run) is launched directly
  const bool fresh = !b->chunk.exec || b->chunk.stale || b->chunk.cfl != cfl ||
                     b->chunk.keep != keep || niter > b->chunk.n;
  if (!fresh && niter < b->chunk.n) {
    for (int k = 0; k < niter; ++k)
      if (int rc = batch_iteration<R>(b, x0, k, cfl, keep, k == 0)) return rc;
  } else {
    if (fresh) {
      HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = 0;
      for (int k = 0; k < niter && !rc; ++k)
        rc = batch_iteration<R>(b, x0, k, cfl, keep, k == 0);
      Chunk c;
      hipError_t e = hipStreamEndCapture(b->stream, &c.graph);
      if (!rc && e == hipSuccess) e = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
      if (rc || e != hipSuccess) {
        (void)release_chunk(c);
        if (rc) return rc;
        HIPCHK(e);
      }
      c.n = niter;
      c.cfl = cfl;
      c.keep = keep;
      c.stale = false;
      HIPCHK(release_chunk(b->chunk));
      b->chunk = c;
      b->chunk_form = form;
    }
    HIPCHK(hipGraphLaunch(b->chunk.exec, b->stream));
  }
  // the error words of every active case: the first failing case is reported
  const CaseDesc<R>* cs = (const CaseDesc<R>*)b->d_desc;
  km_errors<R><<<nblocks(b->nids), BLK, 0, b->stream>>>(cs, b->d_ids, b->nids, b->d_err);
  if (int rc = launch_check()) return rc;
  std::vector<int> err((size_t)3 * b->nids);
  HIPCHK(hipMemcpyAsync(err.data(), b->d_err, err.size() * sizeof(int), hipMemcpyDeviceToHost,
                        b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  int j = 0;
  for (size_t i = 0; i < b->cs.size(); ++i) {
    if (!b->on[i]) continue;
    b->cs[i]->cfl = cfl;
    b->cs[i]->qold_stale = true;  // batched cases run the fused RK3 iteration
    b->cs[i]->prims_next = true;  // and its last stage refreshes the primitives
    b->cs[i]->dtla_pending = false;  // its time-step pass stores la and dt
    if (err[3 * j] != 0) {
      int rc = device_error(b->cs[i]);
      set_error("case " + std::to_string(i) + ": " + ufx_last_error());
      return rc ? rc : UFX_ERR_STATE;
    }
    ++j;
  }
  return UFX_OK;
}

template <typename R>
int batch_norms(ufx_cc_batch* b, int var, double* out) {
  const int n = (int)b->cs.size();
  for (int i = 0; i < n; ++i) out[i] = std::nan("");
  if (b->dirty)
    if (int rc = batch_write<R>(b)) return rc;
  if (b->nids == 0) return UFX_OK;
  const CaseDesc<R>* cs = (const CaseDesc<R>*)b->d_desc;
  km_sumsq_var<R><<<dim3((unsigned)b->cnt[T_CELL]), BLK, 0, b->stream>>>(
      cs, (const int4*)(b->d_map + b->off[T_CELL]), var);
  km_norm_var_final<R><<<dim3((unsigned)b->cnt[T_CASE]), BLK, 0, b->stream>>>(
      cs, (const int4*)(b->d_map + b->off[T_CASE]), var, b->d_norm);
  if (int rc = launch_check()) return rc;
  std::vector<double> v(n);
  HIPCHK(hipMemcpyAsync(v.data(), b->d_norm, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                        b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i)
    if (b->on[i]) out[i] = v[i];
  return UFX_OK;
}

}  // namespace

extern "C" {

int ufx_cc_batch_residual_norms(ufx_cc_batch* b, int var, double* out) {
  if (!b || !out) { set_error("ufx_cc_batch_residual_norms: null argument"); return UFX_ERR_ARG; }
  if (var < -1 || var > 3) { set_error("ufx_cc_batch_residual_norms: var must be -1..3"); return UFX_ERR_ARG; }
  for (size_t i = 0; i < b->cs.size(); ++i)
    if (b->on[i]) HIPCHK(hipStreamSynchronize(b->cs[i]->stream));
  return b->f64 ? batch_norms<double>(b, var, out) : batch_norms<float>(b, var, out);
}

int ufx_cc_batch_create(ufx_cc* const* cases, int n, ufx_cc_batch** out) {
  if (!out || !cases || n < 1) { set_error("ufx_cc_batch_create: needs n >= 1 cases"); return UFX_ERR_ARG; }
  *out = nullptr;
  ufx_cc* x0 = cases[0];
  for (int i = 0; i < n; ++i) {
    ufx_cc* x = cases[i];
    if (int rc = check_ctx(x)) return rc;
    for (int j = 0; j < i; ++j)
      if (cases[j] == x) { set_error("ufx_cc_batch_create: a case is listed twice"); return UFX_ERR_ARG; }
    const Params &p = x->p, &p0 = x0->p;
    if (p.flow != UFX_FLOW_INVISCID || p.irs || p.timemode != UFX_TIME_RK3 ||
        p.grad == UFX_GRAD_GGFC) {
      set_error("ufx_cc_batch_create: batched cases are inviscid RK3 without residual smoothing, "
                "with the LSQ or Green-Gauss gradient");
      return UFX_ERR_ARG;
    }
    if (p.precision != p0.precision || p.flux != p0.flux || p.grad != p0.grad ||
        (p.ilimit != 0) != (p0.ilimit != 0) || x->stage_shared != x0->stage_shared ||
        x->stage_local != x0->stage_local) {
      set_error("ufx_cc_batch_create: the cases differ in precision, flux, gradient, limiter or "
                "stage form");
      return UFX_ERR_ARG;
    }
  }
  auto* b = new ufx_cc_batch();
  b->cs.assign(cases, cases + n);
  b->on.assign(n, 1);
  b->hist_seen.assign(n, nullptr);
  b->f64 = x0->p.precision == UFX_PREC_FP64;
  for (int i = 0; i < n; ++i) b->stage_lds = std::max(b->stage_lds, cases[i]->stage_lds);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    set_error("ufx_cc_batch_create: stream creation failed");
    return UFX_ERR_HIP;
  }
  *out = b;
  return UFX_OK;
}

int ufx_cc_batch_set_active(ufx_cc_batch* b, const int* active) {
  if (!b || !active) { set_error("ufx_cc_batch_set_active: null argument"); return UFX_ERR_ARG; }
  bool changed = false;
  for (size_t i = 0; i < b->cs.size(); ++i) {
    char v = active[i] != 0;
    if (v && !b->on[i]) {
      set_error("ufx_cc_batch_set_active: a retired case cannot return");
      return UFX_ERR_ARG;
    }
    changed = changed || v != b->on[i];
    b->on[i] = v;
  }
  if (changed) b->dirty = true;
  return UFX_OK;
}

int ufx_cc_batch_iterate(ufx_cc_batch* b, int niter, double cfl) {
  if (!b) { set_error("ufx_cc_batch_iterate: null batch"); return UFX_ERR_ARG; }
  if (niter < 1) { set_error("ufx_cc_batch_iterate: needs niter >= 1"); return UFX_ERR_ARG; }
  return b->f64 ? batch_iterate<double>(b, niter, cfl) : batch_iterate<float>(b, niter, cfl);
}

int ufx_cc_batch_destroy(ufx_cc_batch* b) {
  if (!b) return UFX_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  (void)release_chunk(b->chunk);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_map) (void)hipFree(b->d_map);
  if (b->d_ids) (void)hipFree(b->d_ids);
  if (b->d_err) (void)hipFree(b->d_err);
  if (b->d_norm) (void)hipFree(b->d_norm);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return UFX_OK;
}

}  // extern "C"
