// Host side of the vertex-centred (median-dual) hot path and its C-ABI
// (include/ufx.h, ufx_nc_*): context creation from a preprocessed mesh, one
// entry per reference routine, whole iterations of NCFVM/main.f90:51-88, and
// state injection / read-back in the reference's per-vertex layout.  FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nc_kernels.h"
#include "ufx_internal.h"

using namespace ufx;

#define NCCHK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " +    \
                __FILE__ + ":" + std::to_string(__LINE__));                      \
      return UFX_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

struct ufx_nc {
  ufx_nc_params prm{};
  NCFree g{};
  Gas<double> ga{};  // for the AUSM+-up option (mref as CCFVM/ausm_up_flux.F90:79)
  Gas<float> gaf{};  // the same for the mixed form
  bool mixed = false;  // FP32 records, gradients, limiter and fluxes
  bool qold_pending = false;  // the last save_old (of a whole multicolour LU-SGS
                              // iteration) left qold by vertex to mc_qold
  bool res_pending = false;   // the last residual (of a multicolour LU-SGS step)
                              // left res and dd by vertex to mc_res, mc_dd
  DevNC d{};
  int nop = 0, nof = 0, nblk = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  double* d_hist = nullptr;
  int hist_cap = 0;
  int* lv_verts = nullptr;                 // forward levels then backward levels
  std::vector<int> lv_fwd_off, lv_bwd_off;  // offsets into lv_verts
  int *d_fwd_off = nullptr, *d_bwd_off = nullptr;  // the same on the device
  SweepSlot* d_slots = nullptr;             // packed sweep slots, forward then backward
  int nwall = 0;                            // wall edges (bc 1001)
  int* d_wedge = nullptr;                   // their stored edges, in ascending reference order
  std::vector<int> es, es_ref;              // storage slot of each reference edge, and back
  std::vector<int> rv, sv;                  // reference vertex of each storage slot, and back
  double *d_fe = nullptr, *d_cp = nullptr;  // [nwall][2] force terms, [nwall] -Cp
  int sweep_mode = 4;                       // 0 graph of level launches, 1 level launches, 2 one workgroup, 3 cooperative grid, 4 level-ordered one workgroup
  int grid_wgs = 8;                         // workgroups of the cooperative sweep
  size_t res_lds = 0;                       // LDS of the block-shared residual (0: per-vertex gather form)
  // alternate traversal: the gradient pass and every other colour launch of
  // the sweeps walk their blocks from the far end of each XCD's slice, so a
  // pass starts on the data the one before touched last, which the
  // memory-side cache still holds (UFX_SERPENTINE=0: all forward)
  bool serpentine = true;
  std::vector<int> mc_off;                  // multicolour sweeps: positions of each colour
  std::vector<int> mc_pos;                  // and the position of each stored vertex
  hipGraph_t sweep_graph = nullptr;         // diagonal + both sweeps, captured once
  hipGraphExec_t sweep_exec = nullptr;
  double cfl = 0.0;
  bool prof = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_cat;
  size_t ev_used = 0;
};

namespace {

int nblk_of(int n) { return std::max(1, (n + BLK - 1) / BLK); }

template <typename T>
int nc_alloc(ufx_nc* x, T** p, size_t n) {
  if (n == 0) n = 1;
  NCCHK(hipMalloc((void**)p, n * sizeof(T)));
  x->allocs.push_back((void*)*p);
  NCCHK(hipMemsetAsync(*p, 0, n * sizeof(T), x->stream));
  return UFX_OK;
}

template <typename T>
int nc_upload(ufx_nc* x, T** dst, const std::vector<T>& src) {
  T* p = nullptr;
  if (int rc = nc_alloc(x, &p, src.size())) return rc;
  if (!src.empty())
    NCCHK(hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, x->stream));
  *dst = p;
  return UFX_OK;
}

template <typename T>
int nc_upload(ufx_nc* x, const T** dst, const std::vector<T>& src) {
  T* p = nullptr;
  int rc = nc_upload(x, &p, src);
  *dst = p;
  return rc;
}

int launch_ok() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return UFX_ERR_HIP;
  }
  return UFX_OK;
}

// kernel timing categories (ufx_nc_profile_read)
enum { NK_SAVE = 0, NK_GRAD = 1, NK_RES = 2, NK_NORM = 3, NK_LUSGS = 4, NK_NCAT = 5 };

void prof_mark(ufx_nc* x, int cat) {
  if (!x->prof) return;
  if (x->ev_used == x->ev.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return;
    x->ev.push_back(e);
    x->ev_cat.push_back(-1);
  }
  x->ev_cat[x->ev_used] = cat;
  (void)hipEventRecord(x->ev[x->ev_used++], x->stream);
}

#define NLAUNCH(x, cat, kern, grid, ...)                    \
  do {                                                      \
    prof_mark((x), (cat));                                  \
    kern<<<(grid), BLK, 0, (x)->stream>>>(__VA_ARGS__);     \
    prof_mark((x), -1);                                     \
  } while (0)

template <typename R> const Gas<R>& nc_gas(const ufx_nc* x);
template <> const Gas<double>& nc_gas<double>(const ufx_nc* x) { return x->ga; }
template <> const Gas<float>& nc_gas<float>(const ufx_nc* x) { return x->gaf; }

template <typename R>
int run_save_old_r(ufx_nc* x, int wq) {
  NLAUNCH(x, NK_SAVE, k_nc_save_old<R>, x->nblk, x->d, x->g, wq);
  return launch_ok();
}

// gradient + limiter, with extra blocks for the boundary closures
template <typename R>
int run_grad_lim_r(ufx_nc* x) {
  bool lsq = x->prm.grad == UFX_GRAD_LSQR, bj = x->prm.limiter == UFX_NC_LIM_BJ;
  const int nb = x->nblk + (x->d.nbnd + BLK - 1) / BLK, rev = x->serpentine ? 1 : 0;
  if (lsq && bj) NLAUNCH(x, NK_GRAD, (k_nc_grad_lim<R, 2, 1>), nb, x->d, x->g, rev);
  else if (lsq) NLAUNCH(x, NK_GRAD, (k_nc_grad_lim<R, 2, 0>), nb, x->d, x->g, rev);
  else if (bj) NLAUNCH(x, NK_GRAD, (k_nc_grad_lim<R, 1, 1>), nb, x->d, x->g, rev);
  else NLAUNCH(x, NK_GRAD, (k_nc_grad_lim<R, 1, 0>), nb, x->d, x->g, rev);
  return launch_ok();
}

// fvresidual (gradient, limiter, edge fluxes, time step); mode 1 fuses the
// RK stage update with coefficients (ak, bk)
template <typename R>
int run_fvresidual_r(ufx_nc* x, double cfl, int mode, double ak, double bk) {
  if (int rc = run_grad_lim_r<R>(x)) return rc;
  x->res_pending = mode == 2;
  const Gas<double>& ga = x->ga;
  if (x->res_lds > 0) {
    const size_t lds = x->res_lds;
    prof_mark(x, NK_RES);
    if (x->prm.flux == UFX_FLUX_AUSM)
      k_nc_residual_blk<R, 2><<<x->nblk, BLK, lds, x->stream>>>(x->d, x->g, ga, cfl, mode, ak, bk);
    else
      k_nc_residual_blk<R, 3><<<x->nblk, BLK, lds, x->stream>>>(x->d, x->g, ga, cfl, mode, ak, bk);
    prof_mark(x, -1);
  } else if (x->prm.flux == UFX_FLUX_AUSM) {
    NLAUNCH(x, NK_RES, (k_nc_residual<R, 2>), x->nblk, x->d, x->g, ga, cfl, mode, ak, bk);
  } else {
    NLAUNCH(x, NK_RES, (k_nc_residual<R, 3>), x->nblk, x->d, x->g, ga, cfl, mode, ak, bk);
  }
  return launch_ok();
}

// whole: the save_old of a whole iteration; in multicolour LU-SGS the sweeps
// read the old state by colour position only, so qold by vertex is written
// when something reads it (restore_qold)
int run_save_old(ufx_nc* x, bool whole = false) {
  const bool skip = whole && x->prm.timemode != UFX_TIME_RK3 &&
                    x->prm.lusgs_order == UFX_LUSGS_MULTICOLOUR;
  x->qold_pending = skip;
  return x->mixed ? run_save_old_r<float>(x, skip ? 0 : 1) : run_save_old_r<double>(x, skip ? 0 : 1);
}

int restore_res(ufx_nc* x) {
  if (!x->res_pending) return UFX_OK;
  x->res_pending = false;
  NLAUNCH(x, NK_NORM, k_nc_res_from_mc, x->nblk, x->d);
  return launch_ok();
}

int restore_qold(ufx_nc* x) {
  if (!x->qold_pending) return UFX_OK;
  x->qold_pending = false;
  NLAUNCH(x, NK_SAVE, k_nc_qold_from_mc, x->nblk, x->d);
  return launch_ok();
}

int run_fvresidual(ufx_nc* x, double cfl, int mode, double ak, double bk) {
  return x->mixed ? run_fvresidual_r<float>(x, cfl, mode, ak, bk)
                  : run_fvresidual_r<double>(x, cfl, mode, ak, bk);
}

// The diagonal and both sweeps as level launches (each level one kernel).
int launch_sweeps_levels(ufx_nc* x) {
  k_nc_lusgs_diag<<<x->nblk, BLK, 0, x->stream>>>(x->d);
  for (int dir = 0; dir < 2; ++dir) {
    const std::vector<int>& off = dir == 0 ? x->lv_fwd_off : x->lv_bwd_off;
    for (size_t l = 0; l + 1 < off.size(); ++l) {
      int n = off[l + 1] - off[l];
      if (n <= 0) continue;
      k_nc_lusgs_level<<<nblk_of(n), BLK, 0, x->stream>>>(x->d, x->g, x->lv_verts + off[l], n, dir);
    }
  }
  return launch_ok();
}

// Capture the level launches once into a graph (their arguments never
// change: the sweeps read the residual, time steps and states through the
// context's fixed device arrays).  The graph holds kernel nodes only.
int build_sweep_graph(ufx_nc* x) {
  NCCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_sweeps_levels(x);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(x->stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  NCCHK(e);
  x->sweep_graph = gr;
  NCCHK(hipGraphInstantiate(&x->sweep_exec, gr, nullptr, nullptr, 0));
  return UFX_OK;
}

int run_lusgs(ufx_nc* x, double cfl) {
  const bool mc = x->prm.lusgs_order == UFX_LUSGS_MULTICOLOUR;
  if (int rc = run_fvresidual(x, cfl, mc ? 2 : 0, 0.0, 0.0)) return rc;
  if (mc) {
    // one launch per colour: forward in ascending, reverse in descending
    // order.  With two colours or more, the first colour's forward dq came
    // from the residual pass (mc_n0), and the last colour's forward launch
    // also does its reverse update (fin): neither has neighbours to sum.
    const int nc = (int)x->mc_off.size() - 1;
    const bool fuse = x->d.mc_n0 > 0;
    for (int dir = 0; dir < 2; ++dir)
      for (int k = 0; k < nc; ++k) {
        const int c = dir == 0 ? k : nc - 1 - k, s0 = x->mc_off[c], n = x->mc_off[c + 1] - s0;
        if (fuse && ((dir == 0 && c == 0) || (dir == 1 && c == nc - 1))) continue;
        const int fin = fuse && dir == 0 && c == nc - 1 ? 1 : 0;
        // consecutive colour launches walk their positions in opposite
        // directions (when serpentine), each starting where the last ended
        const int rev = x->serpentine ? (dir * nc + k) & 1 : 0;
        if (n > 0)
          NLAUNCH(x, NK_LUSGS, k_nc_lusgs_mc, nblk_of(n), x->d, x->g, s0, n, dir, rev, fin);
      }
    return launch_ok();
  }
  if (x->sweep_mode == 0) {
    if (!x->sweep_exec)
      if (int rc = build_sweep_graph(x)) return rc;
    prof_mark(x, NK_LUSGS);
    NCCHK(hipGraphLaunch(x->sweep_exec, x->stream));
    prof_mark(x, -1);
    return UFX_OK;
  }
  NLAUNCH(x, NK_LUSGS, k_nc_lusgs_diag, x->nblk, x->d);
  if (x->sweep_mode == 4) {
    NLAUNCH(x, NK_LUSGS, k_nc_sweep_prep, nblk_of(2 * x->nop), x->d, x->g);
    for (int dir = 0; dir < 2; ++dir) {
      const std::vector<int>& off = dir == 0 ? x->lv_fwd_off : x->lv_bwd_off;
      const int* doff = dir == 0 ? x->d_fwd_off : x->d_bwd_off;
      prof_mark(x, NK_LUSGS);
      k_nc_lusgs_stream<<<1, STREAM_THREADS, 0, x->stream>>>(x->d, x->g, doff,
                                                            (int)off.size() - 1, dir);
      prof_mark(x, -1);
    }
    NLAUNCH(x, NK_LUSGS, k_nc_sweep_final, x->nblk, x->d);
    return launch_ok();
  }
  for (int dir = 0; dir < 2; ++dir) {
    const std::vector<int>& off = dir == 0 ? x->lv_fwd_off : x->lv_bwd_off;
    if (x->sweep_mode == 3) {
      // one cooperative launch per sweep (offsets relative to lv_verts)
      const int* doff = dir == 0 ? x->d_fwd_off : x->d_bwd_off;
      int nlev = (int)off.size() - 1;
      DevNC dd = x->d;
      NCFree gg = x->g;
      const int* vv = x->lv_verts;
      void* args[] = {&dd, &gg, &vv, &doff, &nlev, &dir};
      prof_mark(x, NK_LUSGS);
      NCCHK(hipLaunchCooperativeKernel((const void*)k_nc_lusgs_sweep_grid, dim3(x->grid_wgs),
                                       dim3(GRID_SWEEP_THREADS), args, 0, x->stream));
      prof_mark(x, -1);
      continue;
    }
    if (x->sweep_mode == 2) {
      // one workgroup walks all levels (level offsets relative to lv_verts)
      const int* doff = dir == 0 ? x->d_fwd_off : x->d_bwd_off;
      prof_mark(x, NK_LUSGS);
      k_nc_lusgs_sweep<<<1, SWEEP_THREADS, 0, x->stream>>>(x->d, x->g, x->d_slots, doff,
                                                          (int)off.size() - 1, dir);
      prof_mark(x, -1);
      continue;
    }
    for (size_t l = 0; l + 1 < off.size(); ++l) {
      int n = off[l + 1] - off[l];
      if (n <= 0) continue;
      NLAUNCH(x, NK_LUSGS, k_nc_lusgs_level, nblk_of(n), x->d, x->g, x->lv_verts + off[l], n, dir);
    }
  }
  return launch_ok();
}

int ensure_hist(ufx_nc* x, int n) {
  if (n <= x->hist_cap) return UFX_OK;
  if (x->d_hist) (void)hipFree(x->d_hist);
  x->d_hist = nullptr;
  NCCHK(hipMalloc((void**)&x->d_hist, sizeof(double) * 5 * (size_t)n));
  x->hist_cap = n;
  return UFX_OK;
}

// residue: the block partials from the stored residual (`partials`), or
// those the last residual pass formed itself (whole iterations), then the
// final reduction into hist[k]
int run_norm(ufx_nc* x, int iter, int k, bool partials) {
  if (partials) NLAUNCH(x, NK_NORM, k_nc_norm, x->nblk, x->d);
  NLAUNCH(x, NK_NORM, k_nc_norm_final, 1, x->d, x->nblk, iter == 1 ? 1 : 0, x->d_hist, k);
  return launch_ok();
}

// NCFVM/main.f90:54
double nc_cfl(double cfl_max, int iter) {
  return 0.5 * cfl_max * (1.0 + std::tanh((double)iter * 7.0 / 500.0 - 3.5));
}

int check(ufx_nc* x) {
  if (!x) { set_error("null context"); return UFX_ERR_ARG; }
  return UFX_OK;
}

// The vertex records as FP64, NW per vertex, whichever form holds them;
// their primitives and limiter values from the compact copies, which are
// current (the records' copies are refreshed by the gradient pass).
int read_records(ufx_nc* x, std::vector<double>& v) {
  const size_t nop = (size_t)x->nop;
  v.assign(nop * NW, 0.0);
  std::vector<double> qp(nop * 4), ph(nop * 4);
  if (!x->mixed) {
    NCCHK(hipMemcpyAsync(v.data(), x->d.vrec, v.size() * 8, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipMemcpyAsync(qp.data(), x->d.vqp, qp.size() * 8, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipMemcpyAsync(ph.data(), x->d.vphc, ph.size() * 8, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
  } else {
    std::vector<float> f(nop * NW32), fq(nop * 4), fp(nop * 4);
    NCCHK(hipMemcpyAsync(f.data(), x->d.vrec32, f.size() * 4, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipMemcpyAsync(fq.data(), x->d.vqp32, fq.size() * 4, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipMemcpyAsync(fp.data(), x->d.vphc32, fp.size() * 4, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
    const double qi[4] = {x->g.r_inf, x->g.u_inf, x->g.v_inf, x->g.p_inf};
    for (size_t p = 0; p < nop; ++p) {
      for (int k = 0; k < NW; ++k) v[p * NW + k] = (double)f[p * NW32 + k];
      for (int k = 0; k < 4; ++k) {
        qp[p * 4 + k] = (double)fq[p * 4 + k] + qi[k];
        ph[p * 4 + k] = (double)fp[p * 4 + k];
      }
    }
  }
  for (size_t p = 0; p < nop; ++p)
    for (int k = 0; k < 4; ++k) {
      v[p * NW + R_QP + k] = qp[p * 4 + k];
      v[p * NW + R_PHI + k] = ph[p * 4 + k];
    }
  return UFX_OK;
}

// The records, and the compact primitives and limiter values, from FP64
// records.
int write_records(ufx_nc* x, const std::vector<double>& v) {
  const size_t nop = (size_t)x->nop;
  std::vector<double> qp(nop * 4), ph(nop * 4);
  for (size_t p = 0; p < nop; ++p)
    for (int k = 0; k < 4; ++k) {
      qp[p * 4 + k] = v[p * NW + R_QP + k];
      ph[p * 4 + k] = v[p * NW + R_PHI + k];
    }
  if (!x->mixed) {
    NCCHK(hipMemcpyAsync(x->d.vrec, v.data(), v.size() * 8, hipMemcpyHostToDevice, x->stream));
    NCCHK(hipMemcpyAsync(x->d.vqp, qp.data(), qp.size() * 8, hipMemcpyHostToDevice, x->stream));
    NCCHK(hipMemcpyAsync(x->d.vphc, ph.data(), ph.size() * 8, hipMemcpyHostToDevice, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
    return UFX_OK;
  }
  std::vector<float> f(nop * NW32, 0.0f), fq(nop * 4), fp(nop * 4);
  const double qi[4] = {x->g.r_inf, x->g.u_inf, x->g.v_inf, x->g.p_inf};
  for (size_t p = 0; p < nop; ++p) {
    for (int k = 0; k < NW; ++k)
      f[p * NW32 + k] = (float)(v[p * NW + k] - (k >= R_QP && k < R_QP + 4 ? qi[k - R_QP] : 0.0));
    for (int k = 0; k < 4; ++k) {
      fq[p * 4 + k] = (float)(qp[p * 4 + k] - qi[k]);
      fp[p * 4 + k] = (float)ph[p * 4 + k];
    }
  }
  NCCHK(hipMemcpyAsync(x->d.vrec32, f.data(), f.size() * 4, hipMemcpyHostToDevice, x->stream));
  NCCHK(hipMemcpyAsync(x->d.vqp32, fq.data(), fq.size() * 4, hipMemcpyHostToDevice, x->stream));
  NCCHK(hipMemcpyAsync(x->d.vphc32, fp.data(), fp.size() * 4, hipMemcpyHostToDevice, x->stream));
  NCCHK(hipStreamSynchronize(x->stream));
  return UFX_OK;
}

// The mesh with its vertices renumbered to storage order (rv: reference
// vertex of each storage slot, sv: the inverse): per-vertex arrays and
// lists moved, vertex numbers in the lists and edge end points mapped.
// Edges keep their reference numbers, and every list keeps its order.
Mesh nc_storage_mesh(const Mesh& m, const std::vector<int>& rv, const std::vector<int>& sv) {
  const int nop = m.nop, nof = m.nof();
  Mesh s;
  s.kind = m.kind;
  s.nop = nop;
  s.noc = m.noc;
  s.nfi = m.nfi;
  s.nbf = m.nbf;
  s.fnx = m.fnx;
  s.fny = m.fny;
  s.farea = m.farea;
  s.fbc = m.fbc;
  s.fp1.resize(nof);
  s.fp2.resize(nof);
  for (int e = 0; e < nof; ++e) {
    s.fp1[e] = sv[m.fp1[e]];
    s.fp2[e] = sv[m.fp2[e]];
  }
  s.px.resize(nop);
  s.py.resize(nop);
  s.pcv.resize(nop);
  s.v2f_ptr.assign(nop + 1, 0);
  s.v2v_ptr.assign(nop + 1, 0);
  s.v2f.reserve(m.v2f.size());
  s.v2v.reserve(m.v2v.size());
  s.wx.reserve(m.wx.size());
  s.wy.reserve(m.wy.size());
  for (int i = 0; i < nop; ++i) {
    const int r = rv[i];
    s.px[i] = m.px[r];
    s.py[i] = m.py[r];
    s.pcv[i] = m.pcv[r];
    for (int j = m.v2f_ptr[r]; j < m.v2f_ptr[r + 1]; ++j) s.v2f.push_back(m.v2f[j]);
    for (int j = m.v2v_ptr[r]; j < m.v2v_ptr[r + 1]; ++j) {
      s.v2v.push_back(sv[m.v2v[j]]);
      s.wx.push_back(m.wx[j]);
      s.wy.push_back(m.wy[j]);
    }
    s.v2f_ptr[i + 1] = (int)s.v2f.size();
    s.v2v_ptr[i + 1] = (int)s.v2v.size();
  }
  return s;
}

}  // namespace

extern "C" {

int ufx_nc_destroy(ufx_nc* x) {
  if (!x) return UFX_OK;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  for (void* p : x->allocs) (void)hipFree(p);
  if (x->d_hist) (void)hipFree(x->d_hist);
  for (hipEvent_t e : x->ev) (void)hipEventDestroy(e);
  if (x->sweep_exec) (void)hipGraphExecDestroy(x->sweep_exec);
  if (x->sweep_graph) (void)hipGraphDestroy(x->sweep_graph);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return UFX_OK;
}

int ufx_nc_create(const ufx_mesh* mesh, const ufx_nc_params* prm, ufx_nc** out) {
  if (!mesh || !prm || !out) { set_error("ufx_nc_create: null argument"); return UFX_ERR_ARG; }
  const Mesh& m0 = mesh->m;
  if (m0.kind != UFX_KIND_NC) { set_error("ufx_nc_create: not a vertex-centred mesh"); return UFX_ERR_ARG; }
  if (prm->grad != UFX_GRAD_GG && prm->grad != UFX_GRAD_LSQR) {
    set_error("ufx_nc_create: the vertex-centred path has the GG and LSQR gradients");
    return UFX_ERR_ARG;
  }
  if (prm->timemode != UFX_TIME_RK3 && prm->timemode != UFX_TIME_LUSGS) {
    set_error("Unknown time integration scheme");
    return UFX_ERR_ARG;
  }
  if (prm->precision != UFX_PREC_FP64 && prm->precision != UFX_PREC_MIXED) {
    set_error("ufx_nc_create: unknown precision");
    return UFX_ERR_ARG;
  }
  if (prm->precision == UFX_PREC_MIXED && prm->timemode == UFX_TIME_LUSGS &&
      prm->lusgs_order != UFX_LUSGS_MULTICOLOUR) {
    set_error("ufx_nc_create: the mixed form's LU-SGS runs the multicolour sweeps");
    return UFX_ERR_UNSUPPORTED;
  }
  if (prm->flux != UFX_FLUX_VANLEER && prm->flux != UFX_FLUX_AUSM) {
    set_error("ufx_nc_create: the vertex-centred path has the van Leer and AUSM+-up fluxes");
    return UFX_ERR_ARG;
  }
  if (prm->limiter != UFX_NC_LIM_VKN && prm->limiter != UFX_NC_LIM_BJ) {
    set_error("ufx_nc_create: unknown limiter");
    return UFX_ERR_ARG;
  }
  if (prm->lusgs_order != UFX_LUSGS_REFERENCE && prm->lusgs_order != UFX_LUSGS_MULTICOLOUR) {
    set_error("ufx_nc_create: unknown LU-SGS sweep order");
    return UFX_ERR_ARG;
  }
  if (prm->timemode == UFX_TIME_RK3 && !(prm->cfl_max < 1.0)) {
    // NCFVM/main.f90:31-44 sets the RK coefficients only for cfl_max < 1
    set_error("CFL should be less than 1 for rk3 !!!");
    return UFX_ERR_ARG;
  }
  ufx_nc* x = new ufx_nc();
  x->prm = *prm;
  const int nop = m0.nop, nof = m0.nof();
  x->nop = nop;
  x->nof = nof;
  x->nblk = nblk_of(nop);
  if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete x;
    return UFX_ERR_HIP;
  }
  // free stream, NCFVM/common.f90:21-58
  const double gamma = 1.4, pi = 4.0 * std::atan(1.0);
  double aoa = prm->aoa_deg * pi / 180.0;
  double qinf[4];
  qinf[1] = 1.0;
  qinf[2] = std::cos(aoa) * 1.0;
  qinf[3] = std::sin(aoa) * 1.0;
  double p_inf = 1.0 / (gamma * prm->mach * prm->mach);
  qinf[0] = p_inf / (gamma - 1.0) + 0.5;
  x->g.gamma = gamma;
  x->g.gamma1 = gamma - 1.0;
  x->g.r_inf = qinf[1];
  x->g.p_inf = p_inf;
  x->g.a_inf = std::sqrt(gamma * p_inf / qinf[1]);
  x->g.u_inf = qinf[2] / qinf[1];
  x->g.v_inf = qinf[3] / qinf[1];
  x->ga = Gas<double>{gamma, gamma - 1.0, 1.0 / (gamma * prm->mach * prm->mach), qinf[1],
                      x->g.u_inf, x->g.v_inf, p_inf, std::max((double)0.32f, 0.5 * prm->mach),
                      1.0 / (gamma - 1.0)};
  const Gas<double>& G = x->ga;
  x->gaf = Gas<float>{(float)G.gamma, (float)G.gamma1, (float)G.znd, (float)G.r_inf,
                      (float)G.u_inf, (float)G.v_inf, (float)G.p_inf, (float)G.mref,
                      (float)G.igamma1};
  x->mixed = prm->precision == UFX_PREC_MIXED;
  if (const char* r = std::getenv("UFX_SERPENTINE"); r && std::strcmp(r, "0") == 0)
    x->serpentine = false;
  // vertices are stored in Hilbert-curve order of their coordinates, so a
  // block of consecutive vertices is a compact patch of the mesh and most of
  // the neighbours it gathers lie in the same block (in the breadth-first
  // reference numbering a vertex's neighbours lie a whole front away);
  // UFX_NC_ORDER=reference keeps the reference numbering.  Every sum runs
  // in the reference's order whatever the storage order, and the sweeps of
  // the reference-order LU-SGS compare reference numbers.
  std::vector<int>& rv = x->rv;
  std::vector<int>& sv = x->sv;
  {
    const char* ord = std::getenv("UFX_NC_ORDER");
    if (ord && std::strcmp(ord, "reference") == 0) {
      rv.resize(nop);
      for (int i = 0; i < nop; ++i) rv[i] = i;
    } else {
      rv = hilbert_order(m0.px.data(), m0.py.data(), nop);
    }
    sv.assign(nop, 0);
    for (int i = 0; i < nop; ++i) sv[rv[i]] = i;
  }
  const Mesh ms = nc_storage_mesh(m0, rv, sv);
  const Mesh& m = ms;
  // edges are stored in first-touch order over the vertices (each vertex's
  // edges in its list order), so the edges a block of vertices gathers lie
  // in a few contiguous runs instead of all over the reference numbering;
  // every per-vertex sum still runs over the vertex's list, in ascending
  // reference edge order
  std::vector<int>& es = x->es;
  es.assign(nof, -1);
  {
    int ne = 0;
    for (int p = 0; p < nop; ++p)
      for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j)
        if (es[m.v2f[j]] < 0) es[m.v2f[j]] = ne++;
    for (int e = 0; e < nof; ++e)
      if (es[e] < 0) es[e] = ne++;
    x->es_ref.assign(nof, 0);
    for (int e = 0; e < nof; ++e) x->es_ref[es[e]] = e;
  }
  // device arrays
  std::vector<double> vrec((size_t)nop * NW, 0.0);
  const double phi0 = (double)1e20f;  // default limiter value, NCFVM/All_modules.f90:200
  for (int p = 0; p < nop; ++p) {
    for (int k = 0; k < 4; ++k) vrec[(size_t)p * NW + R_PHI + k] = phi0;
    vrec[(size_t)p * NW + R_CEN] = m.px[p];
    vrec[(size_t)p * NW + R_CEN + 1] = m.py[p];
  }
  std::vector<double> qc((size_t)nop * 4), erec((size_t)nof * 4, 0.0);
  for (int p = 0; p < nop; ++p)
    for (int k = 0; k < 4; ++k) qc[(size_t)p * 4 + k] = qinf[k];
  std::vector<int2> ep(nof);
  std::vector<int> ebc(std::max(1, nof), -1), bedge, bkind;
  for (int e = 0; e < nof; ++e) {
    const size_t s = es[e];
    erec[s * 4] = m.fnx[e];
    erec[s * 4 + 1] = m.fny[e];
    erec[s * 4 + 2] = m.farea[e];
    ep[s] = make_int2(m.fp1[e], m.fp2[e]);
  }
  // the wall and farfield edges, whose closures the residual adds
  for (int s = 0; s < nof; ++s) {
    const int e = x->es_ref[s];
    if (m.fbc[e] == WALL || m.fbc[e] == FARF) {
      ebc[s] = (int)bedge.size();
      bedge.push_back(s);
      bkind.push_back(m.fbc[e]);
    }
  }
  std::vector<int> v2f(m.v2f.size());
  std::vector<int2> v2fo(m.v2f.size());
  for (int p = 0; p < nop; ++p)
    for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
      int e = m.v2f[j];
      bool first = m.fp1[e] == p;
      v2f[j] = first ? es[e] : ~es[e];
      v2fo[j] = make_int2(v2f[j], first ? m.fp2[e] : m.fp1[e]);
    }
  // block-shared residual: per block of BLK vertices the stored edges its
  // vertices touch, ascending, and per v2f entry the position of its edge
  // in that list (bit-inverted when the vertex is p2)
  std::vector<int> be_ptr(x->nblk + 1, 0), be_edge, sl_loc(m.v2f.size());
  int be_max = 0;
  {
    std::vector<int> list;
    for (int b = 0; b < x->nblk; ++b) {
      const int p0 = b * BLK, p1 = std::min(nop, p0 + BLK);
      list.clear();
      for (int p = p0; p < p1; ++p)
        for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) list.push_back(es[m.v2f[j]]);
      std::sort(list.begin(), list.end());
      list.erase(std::unique(list.begin(), list.end()), list.end());
      for (int p = p0; p < p1; ++p)
        for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
          const int e = m.v2f[j];
          const int k = (int)(std::lower_bound(list.begin(), list.end(), es[e]) - list.begin());
          sl_loc[j] = m.fp1[e] == p ? k : ~k;
        }
      be_edge.insert(be_edge.end(), list.begin(), list.end());
      be_ptr[b + 1] = (int)be_edge.size();
      be_max = std::max(be_max, (int)list.size());
    }
  }
  {
    // per edge 5 doubles and a boundary index; at least what the norm's
    // block reduction reuses
    const size_t lds = std::max((size_t)be_max * (5 * sizeof(double) + sizeof(int)),
                                (size_t)BLK * (2 * sizeof(double) + sizeof(int)));
    const char* form = std::getenv("UFX_NC_RESIDUAL");
    const bool gather = form && std::strcmp(form, "gather") == 0;
    if (!gather && lds <= 64 * 1024) x->res_lds = lds;
  }
  std::vector<double> wxy(2 * m.v2v.size());
  for (size_t j = 0; j < m.v2v.size(); ++j) {
    wxy[2 * j] = m.wx[j];
    wxy[2 * j + 1] = m.wy[j];
  }
  // LU-SGS level schedule of the sequential sweeps in vertex order (only
  // the reference-order sweeps use it)
  if (prm->timemode == UFX_TIME_LUSGS && prm->lusgs_order == UFX_LUSGS_REFERENCE) {
    std::vector<int> lf(nop, 0), lb(nop, 0);
    int nlf = 0, nlb = 0;
    // (vertices visited in reference order; p is a storage slot, r its
    // reference number)
    for (int r = 0; r < nop; ++r) {
      const int p = sv[r];
      int l = 0;
      for (int j = m.v2v_ptr[p]; j < m.v2v_ptr[p + 1]; ++j)
        if (rv[m.v2v[j]] < r) l = std::max(l, lf[m.v2v[j]] + 1);
      lf[p] = l;
      nlf = std::max(nlf, l + 1);
    }
    for (int r = nop - 1; r >= 0; --r) {
      const int p = sv[r];
      int l = 0;
      for (int j = m.v2v_ptr[p]; j < m.v2v_ptr[p + 1]; ++j)
        if (rv[m.v2v[j]] > r) l = std::max(l, lb[m.v2v[j]] + 1);
      lb[p] = l;
      nlb = std::max(nlb, l + 1);
    }
    std::vector<int> verts(2 * (size_t)nop);
    std::vector<int> cf(nlf + 1, 0), cb(nlb + 1, 0);
    for (int p = 0; p < nop; ++p) { cf[lf[p] + 1]++; cb[lb[p] + 1]++; }
    for (int l = 0; l < nlf; ++l) cf[l + 1] += cf[l];
    for (int l = 0; l < nlb; ++l) cb[l + 1] += cb[l];
    for (int l = 0; l <= nlb; ++l) cb[l] += nop;
    x->lv_fwd_off = cf;
    x->lv_bwd_off = cb;
    std::vector<int> pf(cf.begin(), cf.end() - 1), pb(cb.begin(), cb.end() - 1);
    for (int r = 0; r < nop; ++r) verts[pf[lf[sv[r]]]++] = sv[r];
    for (int r = nop - 1; r >= 0; --r) verts[pb[lb[sv[r]]]++] = sv[r];
    // inside a level the vertices are independent, so any order gives the
    // same bits; ordering them by swept-neighbour count (most first) lets a
    // wave's lanes share their count and skip the unused entry slots
    {
      std::vector<int> nlo(nop, 0), nhi(nop, 0);
      for (int p = 0; p < nop; ++p)
        for (int j = m.v2v_ptr[p]; j < m.v2v_ptr[p + 1]; ++j)
          ++(rv[m.v2v[j]] < rv[p] ? nlo[p] : nhi[p]);
      auto sort_levels = [&](const std::vector<int>& off, const std::vector<int>& cnt) {
        for (size_t l = 0; l + 1 < off.size(); ++l)
          std::stable_sort(verts.begin() + off[l], verts.begin() + off[l + 1],
                           [&](int a, int b) { return cnt[a] > cnt[b]; });
      };
      sort_levels(cf, nlo);
      sort_levels(cb, nhi);
    }
    int rc = nc_upload(x, &x->lv_verts, verts);
    if (!rc) rc = nc_upload(x, &x->d_fwd_off, x->lv_fwd_off);
    if (!rc) rc = nc_upload(x, &x->d_bwd_off, x->lv_bwd_off);
    // packed slots of the one-workgroup sweeps: per level position its
    // vertex and the (edge, neighbour) pairs already swept, in edge order
    bool fits = true;
    std::vector<SweepSlot> slots(2 * (size_t)nop);
    for (size_t s = 0; s < slots.size(); ++s) {
      int p = verts[s], dir = s < (size_t)nop ? 0 : 1, cnt = 0;
      SweepSlot& sl = slots[s];
      sl.vert = p;
      for (int k = 0; k < SWQ; ++k) sl.eo[k] = make_int2(0, 0);  // unused slots: in bounds
      for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
        int e = m.v2f[j];
        bool first = m.fp1[e] == p;
        int o = first ? m.fp2[e] : m.fp1[e];
        if (dir == 0 ? rv[o] < rv[p] : rv[o] > rv[p]) {
          if (cnt == SWQ) { fits = false; break; }
          sl.eo[cnt++] = make_int2(first ? es[e] : ~es[e], o);
        }
      }
      sl.cnt = cnt;
    }
    if (!rc && fits) rc = nc_upload(x, &x->d_slots, slots);
    // position-major data of the streamed one-workgroup sweeps: per
    // position sw_q slots (the largest swept-neighbour count) holding its
    // swept neighbours in edge order, with the edge geometry (unused
    // slots: zeros, never read)
    {
      const size_t np = verts.size();
      std::vector<int> pos(np);  // position of each vertex in each sweep
      for (size_t s = 0; s < np; ++s) pos[(s < (size_t)nop ? 0 : nop) + verts[s]] = (int)s;
      std::vector<int> sw_cnt(np, 0), sw_d0(np, 0);
      std::vector<int2> eo_all;
      std::vector<size_t> eo_ptr(np + 1, 0);
      int q = 1;
      for (size_t s = 0; s < np; ++s) {
        int p = verts[s], dir = s < (size_t)nop ? 0 : 1, cnt = 0;
        for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
          int e = m.v2f[j];
          bool first = m.fp1[e] == p;
          int o = first ? m.fp2[e] : m.fp1[e];
          if (dir == 0 ? rv[o] < rv[p] : rv[o] > rv[p]) {
            eo_all.push_back(make_int2(first ? e : ~e, o));
            ++cnt;
          }
        }
        sw_cnt[s] = cnt;
        sw_d0[s] = pos[p];
        eo_ptr[s + 1] = eo_all.size();
        q = std::max(q, cnt);
      }
      const size_t nq = (size_t)q;
      std::vector<int> sw_nb(nq * np, 0);
      std::vector<int2> sw_pe(nq * np, make_int2(0, 0));
      std::vector<double2> sw_ent(nq * SWE * np, make_double2(0.0, 0.0));
      for (size_t s = 0; s < np; ++s)
        for (size_t j = eo_ptr[s]; j < eo_ptr[s + 1]; ++j) {
          const size_t k = j - eo_ptr[s];
          const int2 eo = eo_all[j];
          const int e = eo.x >= 0 ? eo.x : ~eo.x;
          const int on = pos[(s < (size_t)nop ? 0 : nop) + eo.y];
          sw_nb[k * np + s] = eo.x >= 0 ? on : ~on;
          sw_pe[k * np + s] = make_int2(es[e], eo.y);
          sw_ent[k * SWE * np + s] = make_double2(m.fnx[e], m.fny[e]);
          sw_ent[(k * SWE + 1) * np + s] = make_double2(m.farea[e], 0.0);
        }
      x->d.sw_q = q;
      x->d.sw_vert = x->lv_verts;
      if (!rc) rc = nc_upload(x, &x->d.sw_cnt, sw_cnt);
      if (!rc) rc = nc_upload(x, &x->d.sw_nb, sw_nb);
      if (!rc) rc = nc_upload(x, &x->d.sw_d0, sw_d0);
      if (!rc) rc = nc_upload(x, &x->d.sw_pe, sw_pe);
      if (!rc) rc = nc_upload(x, &x->d.sw_ent, sw_ent);
      if (!rc) rc = nc_alloc(x, &x->d.sw_own, 4 * np);
      if (!rc) rc = nc_alloc(x, &x->d.sw_dd, np);
      if (!rc) rc = nc_alloc(x, &x->d.sw_dq, 4 * np);
    }
    if (rc) { ufx_nc_destroy(x); return rc; }
    // sweep form: one workgroup over the level-ordered data (default,
    // UFX_NC_SWEEP=stream), a replayed graph of the level launches (graph),
    // the level launches themselves (levels), one workgroup walking the
    // packed slots (single; needs at most SWQ swept neighbours per vertex),
    // or one cooperative launch of several workgroups (grid)
    const char* env = std::getenv("UFX_NC_SWEEP");
    x->sweep_mode = 4;
    if (env && std::strcmp(env, "graph") == 0) x->sweep_mode = 0;
    if (env && std::strcmp(env, "levels") == 0) x->sweep_mode = 1;
    if (env && std::strcmp(env, "single") == 0 && fits) x->sweep_mode = 2;
    if (env && std::strcmp(env, "grid") == 0) x->sweep_mode = 3;
    if (env && std::strcmp(env, "stream") == 0) x->sweep_mode = 4;
    if (const char* w = std::getenv("UFX_NC_SWEEP_WGS")) x->grid_wgs = std::max(1, std::min(64, std::atoi(w)));
  }
  {
    int rc = 0;
    // multicolour sweeps: vertices grouped by colour, and per direction the
    // neighbours of lower (forward) / higher (reverse) colour in edge order
    if (prm->timemode == UFX_TIME_LUSGS && prm->lusgs_order == UFX_LUSGS_MULTICOLOUR) {
      // the colouring of the reference numbering (ufx_mesh_get "pt_colour"),
      // carried to storage slots; inside a colour the vertices are
      // independent and taken in storage order
      int ncol = 0;
      const std::vector<int> col0 = lusgs_colours(m0, &ncol);
      std::vector<int> col(nop);
      for (int p = 0; p < nop; ++p) col[p] = col0[rv[p]];
      x->mc_off.assign(ncol + 1, 0);
      for (int p = 0; p < nop; ++p) x->mc_off[col[p] + 1]++;
      for (int c = 0; c < ncol; ++c) x->mc_off[c + 1] += x->mc_off[c];
      std::vector<int> mv(nop), at(x->mc_off.begin(), x->mc_off.end() - 1), mpos(nop);
      for (int p = 0; p < nop; ++p) mv[at[col[p]]++] = p;
      for (int sp = 0; sp < nop; ++sp) mpos[mv[sp]] = sp;
      std::vector<int> cnt(2 * (size_t)nop, 0);
      int q = 1;
      for (int sp = 0; sp < nop; ++sp) {
        int p = mv[sp];
        for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
          int e = m.v2f[j], o = m.fp1[e] == p ? m.fp2[e] : m.fp1[e];
          if (col[o] != col[p]) ++cnt[(col[o] < col[p] ? 0 : (size_t)nop) + sp];
        }
        q = std::max(q, std::max(cnt[sp], cnt[(size_t)nop + sp]));
      }
      const size_t nq = (size_t)q * nop;
      std::vector<int2> eo(2 * nq, make_int2(0, 0));
      std::vector<double2> nxy(2 * nq, make_double2(0.0, 0.0));
      std::vector<double> dsv(2 * nq, 0.0);
      for (int sp = 0; sp < nop; ++sp) {
        int p = mv[sp], k[2] = {0, 0};
        for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
          int e = m.v2f[j];
          bool first = m.fp1[e] == p;
          int o = first ? m.fp2[e] : m.fp1[e];
          if (col[o] == col[p]) continue;  // not a proper colouring: never taken
          int dir = col[o] < col[p] ? 0 : 1;
          size_t at2 = (size_t)dir * nq + (size_t)k[dir]++ * nop + sp;
          eo[at2] = make_int2(first ? es[e] : ~es[e], mpos[o]);
          nxy[at2] = make_double2(m.fnx[e], m.fny[e]);
          dsv[at2] = m.farea[e];
        }
      }
      x->d.mc_q = q;
      // with two colours or more the residual pass forms the first colour's
      // forward dq (UFX_NC_FUSE_COLOURS=0: the sweep launches do)
      const char* fz = std::getenv("UFX_NC_FUSE_COLOURS");
      x->d.mc_n0 = (ncol >= 2 && !(fz && std::string(fz) == "0")) ? x->mc_off[1] : 0;
      if (!rc) rc = nc_upload(x, &x->d.mc_vert, mv);
      if (!rc) rc = nc_upload(x, &x->d.mc_cnt, cnt);
      if (!rc) rc = nc_upload(x, &x->d.mc_eo, eo);
      if (!rc) rc = nc_upload(x, &x->d.mc_nxy, nxy);
      if (!rc) rc = nc_upload(x, &x->d.mc_ds, dsv);
      if (!rc) rc = nc_alloc(x, &x->d.mc_qold, 4 * (size_t)nop);
      if (!rc) rc = nc_alloc(x, &x->d.mc_res, 4 * (size_t)nop);
      if (!rc) rc = nc_upload(x, &x->d.mc_pos, mpos);
      x->mc_pos = mpos;
      if (!rc) rc = nc_alloc(x, &x->d.mc_dq, 4 * (size_t)nop);
      if (!rc) rc = nc_alloc(x, &x->d.mc_dd, (size_t)nop);
    }
    if (rc) { ufx_nc_destroy(x); return rc; }
  }
  DevNC& d = x->d;
  d.nop = nop;
  d.nof = nof;
  int rc = 0;
#define NUP(dst, src) if (!rc) rc = nc_upload(x, &d.dst, src)
#define NAL(dst, n) if (!rc) rc = nc_alloc(x, &d.dst, (size_t)(n))
  if (x->mixed) {
    // the records in FP32 (the FP64 ones are not allocated)
    std::vector<float> vr32((size_t)nop * NW32, 0.0f), er32(erec.size()), w32(wxy.size());
    for (int p = 0; p < nop; ++p)
      for (int k = 0; k < NW; ++k) vr32[(size_t)p * NW32 + k] = (float)vrec[(size_t)p * NW + k];
    for (size_t i = 0; i < erec.size(); ++i) er32[i] = (float)erec[i];
    // the Green-Gauss terms nx ds, ny ds as the FP32 pass forms them (one
    // single-precision product each)
    std::vector<float> gg32((size_t)nof * 2);
    for (int s = 0; s < nof; ++s) {
      gg32[(size_t)s * 2] = er32[(size_t)s * 4] * er32[(size_t)s * 4 + 2];
      gg32[(size_t)s * 2 + 1] = er32[(size_t)s * 4 + 1] * er32[(size_t)s * 4 + 2];
    }
    NUP(egg32, gg32);
    for (size_t i = 0; i < wxy.size(); ++i) w32[i] = (float)wxy[i];
    // primitives as deviations from the free stream (the state starts
    // there: zero), and the open Green-Gauss sum's +-n ds per vertex
    for (int p = 0; p < nop; ++p)
      for (int k = 0; k < 4; ++k) vr32[(size_t)p * NW32 + R_QP + k] = 0.0f;
    std::vector<float2> ggs(nop);
    for (int p = 0; p < nop; ++p) {
      double sx = 0.0, sy = 0.0;
      for (int j = m.v2f_ptr[p]; j < m.v2f_ptr[p + 1]; ++j) {
        int e = m.v2f[j];
        double dx = m.fnx[e] * m.farea[e], dy = m.fny[e] * m.farea[e];
        if (m.fp1[e] == p) { sx = sx + dx; sy = sy + dy; }
        else { sx = sx - dx; sy = sy - dy; }
      }
      ggs[p] = make_float2((float)sx, (float)sy);
    }
    NUP(vrec32, vr32); NUP(erec32, er32); NUP(wxy32, w32); NUP(ggs32, ggs);
    std::vector<float> q32((size_t)nop * 4, 0.0f), p32((size_t)nop * 4, (float)phi0), x32((size_t)nop * 2);
    for (int p = 0; p < nop; ++p) {
      x32[(size_t)p * 2] = (float)m.px[p];
      x32[(size_t)p * 2 + 1] = (float)m.py[p];
    }
    NUP(vqp32, q32); NUP(vphc32, p32); NUP(vxy32, x32);
  } else {
    // edge records packed to their three values
    std::vector<double> er64((size_t)nof * 3);
    for (int s = 0; s < nof; ++s)
      for (int k = 0; k < 3; ++k) er64[(size_t)s * 3 + k] = erec[(size_t)s * 4 + k];
    NUP(vrec, vrec); NUP(erec, er64); NUP(wxy, wxy);
    std::vector<double> gg((size_t)nof * 2);
    for (int s = 0; s < nof; ++s) {
      gg[(size_t)s * 2] = erec[(size_t)s * 4] * erec[(size_t)s * 4 + 2];
      gg[(size_t)s * 2 + 1] = erec[(size_t)s * 4 + 1] * erec[(size_t)s * 4 + 2];
    }
    NUP(egg, gg);
    std::vector<double> q64((size_t)nop * 4, 0.0), p64((size_t)nop * 4, phi0), x64((size_t)nop * 2);
    for (int p = 0; p < nop; ++p) {
      x64[(size_t)p * 2] = m.px[p];
      x64[(size_t)p * 2 + 1] = m.py[p];
    }
    NUP(vqp, q64); NUP(vphc, p64); NUP(vxy, x64);
  }
  NUP(qc, qc); NUP(cv, m.pcv); NUP(ebc, ebc); NUP(vref, rv);
  NUP(bedge, bedge); NUP(bkind, bkind); NAL(bclos, 10 * bedge.size());
  NUP(be_ptr, be_ptr); NUP(be_edge, be_edge); NUP(sl_loc, sl_loc);
  d.be_max = be_max;
  d.nbnd = (int)bedge.size();
  NUP(v2f_ptr, m.v2f_ptr); NUP(v2f, v2f); NUP(v2v_ptr, m.v2v_ptr); NUP(v2v, m.v2v);
  if (!rc) {
    std::vector<double2> pxy(nop);
    for (int p = 0; p < nop; ++p) pxy[p] = make_double2(m.px[p], m.py[p]);
    double2* q = nullptr;
    rc = nc_upload(x, &q, pxy);
    d.pxy = q;
  }
  if (!rc) {
    int2* p = nullptr;
    rc = nc_upload(x, &p, ep);
    d.ep = p;
  }
  if (!rc) {
    int2* p = nullptr;
    rc = nc_upload(x, &p, v2fo);
    d.v2fo = p;
  }
  NAL(qold, 4 * (size_t)nop); NAL(res, 4 * (size_t)nop); NAL(la, nop); NAL(dt, nop);
  NAL(dd, nop); NAL(dqc, 4 * (size_t)nop); NAL(fla, nof); NAL(part, 4 * (size_t)x->nblk);
  NAL(scal, 4); NAL(err, 4);
  {
    std::vector<int> wedge;
    for (int e = 0; e < nof; ++e)
      if (m.fbc[e] == 1001) wedge.push_back(es[e]);
    x->nwall = (int)wedge.size();
    if (!rc) rc = nc_upload(x, &x->d_wedge, wedge);
    if (!rc) rc = nc_alloc(x, &x->d_fe, 2 * (size_t)x->nwall + 2);
    if (!rc) rc = nc_alloc(x, &x->d_cp, (size_t)x->nwall);
  }
#undef NUP
#undef NAL
  if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) {
    set_error("upload failed");
    rc = UFX_ERR_HIP;
  }
  if (rc) { ufx_nc_destroy(x); return rc; }
  *out = x;
  return UFX_OK;
}

int ufx_nc_save_old(ufx_nc* x) {
  if (int rc = check(x)) return rc;
  return run_save_old(x);
}

int ufx_nc_fvresidual(ufx_nc* x, double cfl) {
  if (int rc = check(x)) return rc;
  x->cfl = cfl;
  return run_fvresidual(x, cfl, 0, 0.0, 0.0);
}

int ufx_nc_rk_stage(ufx_nc* x, int stage, double cfl) {
  if (int rc = check(x)) return rc;
  static const double airk[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
  static const double birk[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  if (stage < 1 || stage > 3) { set_error("stage must be 1..3"); return UFX_ERR_ARG; }
  if (int rc = restore_qold(x)) return rc;
  x->cfl = cfl;
  return run_fvresidual(x, cfl, 1, airk[stage - 1], birk[stage - 1]);
}

int ufx_nc_lusgs(ufx_nc* x, double cfl) {
  if (int rc = check(x)) return rc;
  if (int rc = restore_qold(x)) return rc;
  x->cfl = cfl;
  return run_lusgs(x, cfl);
}

int ufx_nc_residue(ufx_nc* x, int iter, double* out) {
  if (int rc = check(x)) return rc;
  if (int rc = restore_res(x)) return rc;
  if (int rc = ensure_hist(x, 1)) return rc;
  if (int rc = run_norm(x, iter, 0, true)) return rc;
  double h[5];
  NCCHK(hipMemcpyAsync(h, x->d_hist, sizeof(h), hipMemcpyDeviceToHost, x->stream));
  NCCHK(hipStreamSynchronize(x->stream));
  if (out) std::memcpy(out, h, sizeof(h));
  return UFX_OK;
}

int ufx_nc_forces(ufx_nc* x, double* out, double* cp) {
  if (int rc = check(x)) return rc;
  if (!out) { set_error("ufx_nc_forces: out is NULL"); return UFX_ERR_ARG; }
  const NCFree& g = x->g;
  const double dyna = 0.5 * g.gamma * g.p_inf * x->prm.mach * x->prm.mach;
  if (x->nwall > 0) {
    k_nc_wall_terms<<<nblk_of(x->nwall), BLK, 0, x->stream>>>(x->d, g, x->d_wedge, x->nwall, dyna,
                                                              x->d_fe, x->d_cp);
    k_nc_forces_sum<<<1, 64, 0, x->stream>>>(x->d_fe, x->nwall, x->d_fe + 2 * (size_t)x->nwall);
    NCCHK(hipGetLastError());
  }
  double f[2] = {0.0, 0.0};
  if (x->nwall > 0)
    NCCHK(hipMemcpyAsync(f, x->d_fe + 2 * (size_t)x->nwall, sizeof(f), hipMemcpyDeviceToHost,
                         x->stream));
  if (cp && x->nwall > 0)
    NCCHK(hipMemcpyAsync(cp, x->d_cp, sizeof(double) * x->nwall, hipMemcpyDeviceToHost,
                         x->stream));
  NCCHK(hipStreamSynchronize(x->stream));
  const double pi = 4.0 * std::atan(1.0), aoa = x->prm.aoa_deg * pi / 180.0;
  out[0] = (f[1] * std::cos(aoa) - f[0] * std::sin(aoa)) / dyna;
  out[1] = (f[1] * std::sin(aoa) + f[0] * std::cos(aoa)) / dyna;
  out[2] = f[0];
  out[3] = f[1];
  return UFX_OK;
}

int ufx_nc_iterate(ufx_nc* x, int iter0, int niter, double* hist) {
  if (int rc = check(x)) return rc;
  if (niter <= 0) return UFX_OK;
  if (int rc = ensure_hist(x, niter)) return rc;
  for (int k = 0; k < niter; ++k) {
    int iter = iter0 + k + 1;
    double cfl = nc_cfl(x->prm.cfl_max, iter);
    x->cfl = cfl;
    if (int rc = run_save_old(x, true)) return rc;
    if (x->prm.timemode == UFX_TIME_RK3) {
      for (int s = 1; s <= 3; ++s)
        if (int rc = ufx_nc_rk_stage(x, s, cfl)) return rc;
    } else {
      if (int rc = run_lusgs(x, cfl)) return rc;
    }
    if (int rc = run_norm(x, iter, k, false)) return rc;
  }
  if (hist) {
    NCCHK(hipMemcpyAsync(hist, x->d_hist, sizeof(double) * 5 * (size_t)niter,
                         hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
  }
  return UFX_OK;
}

int ufx_nc_set(ufx_nc* x, const char* name, const double* src) {
  if (int rc = check(x)) return rc;
  if (!name || !src) { set_error("ufx_nc_set: null argument"); return UFX_ERR_ARG; }
  std::string n(name);
  const int nop = x->nop;
  // arrays come in the reference's vertex numbering; the device holds them
  // by storage slot
  if (n == "qc" || n == "qold") {
    std::vector<double> t(4 * (size_t)nop);
    for (int p = 0; p < nop; ++p)
      for (int k = 0; k < 4; ++k) t[(size_t)p * 4 + k] = src[(size_t)x->rv[p] * 4 + k];
    NCCHK(hipMemcpyAsync(n == "qc" ? x->d.qc : x->d.qold, t.data(), sizeof(double) * 4 * nop,
                         hipMemcpyHostToDevice, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
    if (n == "qold") x->qold_pending = false;
    if (n == "qold" && !x->mc_pos.empty()) {
      // the multicolour sweeps' copy by colour position
      std::vector<double> u(4 * (size_t)nop);
      for (int p = 0; p < nop; ++p)
        for (int k = 0; k < 4; ++k) u[(size_t)x->mc_pos[p] * 4 + k] = t[(size_t)p * 4 + k];
      NCCHK(hipMemcpyAsync(x->d.mc_qold, u.data(), sizeof(double) * 4 * nop, hipMemcpyHostToDevice,
                           x->stream));
      NCCHK(hipStreamSynchronize(x->stream));
    }
  } else if (n == "phi") {
    std::vector<double> v;
    if (int rc = read_records(x, v)) return rc;
    for (int p = 0; p < nop; ++p)
      for (int k = 0; k < 4; ++k) v[(size_t)p * NW + R_PHI + k] = src[(size_t)x->rv[p] * 4 + k];
    if (int rc = write_records(x, v)) return rc;
  } else {
    set_error("ufx_nc_set: unknown or read-only array " + n);
    return UFX_ERR_ARG;
  }
  NCCHK(hipStreamSynchronize(x->stream));
  return UFX_OK;
}

int ufx_nc_get(ufx_nc* x, const char* name, double* dst) {
  if (int rc = check(x)) return rc;
  if (!name || !dst) { set_error("ufx_nc_get: null argument"); return UFX_ERR_ARG; }
  std::string n(name);
  const int nop = x->nop;
  auto copy = [&](const double* src, size_t cnt) -> int {
    NCCHK(hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
    return UFX_OK;
  };
  // per-vertex arrays (w values each) from storage slots to the reference's
  // vertex numbering
  auto vertex_array = [&](const double* src, int w) -> int {
    std::vector<double> t((size_t)w * nop);
    NCCHK(hipMemcpyAsync(t.data(), src, t.size() * 8, hipMemcpyDeviceToHost, x->stream));
    NCCHK(hipStreamSynchronize(x->stream));
    for (int p = 0; p < nop; ++p)
      for (int k = 0; k < w; ++k) dst[(size_t)x->rv[p] * w + k] = t[(size_t)p * w + k];
    return UFX_OK;
  };
  if (n == "qc") return vertex_array(x->d.qc, 4);
  if (n == "qold") {
    if (int rc = restore_qold(x)) return rc;
    return vertex_array(x->d.qold, 4);
  }
  if (n == "res") {
    if (int rc = restore_res(x)) return rc;
    return vertex_array(x->d.res, 4);
  }
  if (n == "la") return vertex_array(x->d.la, 1);
  if (n == "dt") return vertex_array(x->d.dt, 1);
  if (n == "fc_la") {
    std::vector<double> t(x->nof);
    if (int rc = copy(x->d.fla, x->nof)) return rc;
    std::copy(dst, dst + x->nof, t.begin());
    for (int e = 0; e < x->nof; ++e) dst[e] = t[x->es[e]];
    return UFX_OK;
  }
  if (n == "qp" || n == "grad" || n == "phi") {
    std::vector<double> v;
    if (int rc = read_records(x, v)) return rc;
    int off = n == "qp" ? R_QP : n == "grad" ? R_G : R_PHI, w = n == "grad" ? 8 : 4;
    for (int p = 0; p < nop; ++p)
      for (int k = 0; k < w; ++k) dst[(size_t)x->rv[p] * w + k] = v[(size_t)p * NW + off + k];
    return UFX_OK;
  }
  set_error("ufx_nc_get: unknown array " + n);
  return UFX_ERR_ARG;
}

int ufx_nc_sync(ufx_nc* x) {
  if (int rc = check(x)) return rc;
  NCCHK(hipStreamSynchronize(x->stream));
  return UFX_OK;
}

int ufx_nc_profile(ufx_nc* x, int enable) {
  if (int rc = check(x)) return rc;
  NCCHK(hipStreamSynchronize(x->stream));
  x->prof = enable != 0;
  x->ev_used = 0;
  return UFX_OK;
}

int ufx_nc_profile_read(ufx_nc* x, double* ms, int* count) {
  if (int rc = check(x)) return rc;
  NCCHK(hipStreamSynchronize(x->stream));
  for (int c = 0; c < NK_NCAT; ++c) { ms[c] = 0.0; count[c] = 0; }
  for (size_t i = 0; i + 1 < x->ev_used; ++i) {
    int c = x->ev_cat[i];
    if (c < 0) continue;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, x->ev[i], x->ev[i + 1]) == hipSuccess) {
      ms[c] += t;
      count[c] += 1;
    }
  }
  x->ev_used = 0;
  return UFX_OK;
}

}  // extern "C"
