// Many independent cell-centred cases per launch (gfx950).  Included once,
// by cc_solver.hip, after cc_kernels.h.
//
// Each kernel of the inviscid RK3 iteration exists as a body over a virtual
// block (cc_kernels.h).  A many-case launch runs one hardware block per
// entry of a block table: entry i = (case, block among that case's blocks
// of this kernel, that case's block count).  The block loads the case's
// descriptor (its device arrays, free stream and sizes) and runs the same
// body the one-context kernel runs, so every per-cell value equals the
// one-context result bit for bit.  Hardware blocks are dealt to the 8 XCDs
// round-robin; within a case, blocks with equal index mod 8 land on one XCD,
// so each die still walks contiguous slices of each case's cells.
#pragma once

#include "cc_kernels.h"

namespace ufx {

// What a many-case launch needs of one context.
template <typename R>
struct CaseDesc {
  DevT<R> d;
  Gas<double> g;
  Gas<R> gr;
  double vll;
  double t_wall;
  double* hist;  // the context's residual history
  int iwall;
  int nblk_all, nblk_pt, nblk_face, nblk_cell, nblk_bf;
};

__device__ __forceinline__ VBlock case_block(const int4* map, int& c) {
  const int4 m = map[blockIdx.x];
  c = m.x;
  return VBlock{m.y, m.z};
}
// The same walking the table from its far end when `rev` (the residual
// stages of the serpentine traversal: each starts on the cases the stage
// before touched last).
__device__ __forceinline__ VBlock case_block_dir(const int4* map, int& c, bool rev) {
  const int4 m = map[rev ? (int)gridDim.x - 1 - (int)blockIdx.x : (int)blockIdx.x];
  c = m.x;
  return VBlock{m.y, m.z};
}

template <typename R>
__global__ __launch_bounds__(BLK) void km_prim(const CaseDesc<R>* cs, const int4* map,
                                               int do_prim, int do_old, int do_range) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  prim_body<R>(e.d, vb, e.g, do_prim, do_old, do_range);
}

// one block per case; the partials come from the primitive refresh (one
// per block of all cells) or from the previous iteration's last stage (one
// per block of interior cells)
template <typename R>
__global__ __launch_bounds__(BLK) void km_range_final(const CaseDesc<R>* cs, const int4* map,
                                                      int from_stage) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  range_final_body<R>(e.d, vb, from_stage ? e.nblk_cell : e.nblk_all);
}

template <typename R, bool VISC>
__global__ __launch_bounds__(BLK) void km_avg_c2v(const CaseDesc<R>* cs, const int4* map) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  avg_c2v_body<R, VISC>(e.d, vb);
}

template <typename R>
__global__ __launch_bounds__(BLK) void km_face_lambda(const CaseDesc<R>* cs, const int4* map) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  face_lambda_body<R>(e.d, vb, e.g, e.vll);
}

template <typename R, int GRAD, bool WITH_T, bool LIMIT, bool TSTEP>
__global__ __launch_bounds__(BLK) void km_cell(const CaseDesc<R>* cs, const int4* map,
                                               double cfl) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  cell_body<R, GRAD, WITH_T, LIMIT, TSTEP>(e.d, vb, e.g, e.vll, cfl);
}

template <typename R, int FLUX>
__global__ __launch_bounds__(BLK) void km_bface(const CaseDesc<R>* cs, const int4* map,
                                                int viscous_wall, int src) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  bface_body<R, FLUX>(e.d, vb, e.g, viscous_wall, e.iwall, e.t_wall, src);
}

template <typename R, int FLUX, bool SL>
__global__ __launch_bounds__(BLK) void km_stage(const CaseDesc<R>* cs, const int4* map, int mode,
                                                double ak, double bk, int write_res, int epi) {
  int c;
  const VBlock vb = case_block_dir(map, c, (mode & 16) != 0);
  const CaseDesc<R> e = cs[c];
  stage_body<R, FLUX, SL>(e.d, vb, e.gr, mode, ak, bk, write_res, e.g, epi);
}

template <typename R, int FLUX, bool SL>
__global__ __launch_bounds__(BLK) void km_stage_blk(const CaseDesc<R>* cs, const int4* map,
                                                    int mode, double ak, double bk,
                                                    int write_res, int epi) {
  int c;
  const VBlock vb = case_block_dir(map, c, (mode & 16) != 0);
  const CaseDesc<R> e = cs[c];
  stage_blk_body<R, FLUX, SL>(e.d, vb, e.gr, mode, ak, bk, write_res, e.g, epi);
}

// one block per case; history slot k of every case
template <typename R>
__global__ void km_norm_final(const CaseDesc<R>* cs, const int4* map, int k) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  norm_final_body<R>(e.d, vb, e.nblk_cell, 0, e.hist, k);
}

// Per-variable residual norm of every case (as ufx_cc_residual_norm): block
// partials over the interior-cell table, then one block per case writing
// out[case].
template <typename R>
__global__ __launch_bounds__(BLK) void km_sumsq_var(const CaseDesc<R>* cs, const int4* map,
                                                    int var) {
  int c;
  const VBlock vb = case_block(map, c);
  const CaseDesc<R> e = cs[c];
  sumsq_var_body<R>(e.d, vb, var, e.d.part);
}

template <typename R>
__global__ void km_norm_var_final(const CaseDesc<R>* cs, const int4* map, int var, double* out) {
  int c;
  case_block(map, c);
  const CaseDesc<R> e = cs[c];
  norm_var_final_body<R>(e.d, e.nblk_cell, var, e.d.part, out + c);
}

// The error words of the listed cases, three ints each, into out.
template <typename R>
__global__ void km_errors(const CaseDesc<R>* cs, const int* ids, int n, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int* w = cs[ids[i]].d.err;
  out[3 * i] = w[0];
  out[3 * i + 1] = w[1];
  out[3 * i + 2] = w[2];
}

}  // namespace ufx
