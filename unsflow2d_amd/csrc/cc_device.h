// Device data layout of the cell-centred solver and small device helpers.
//
// HBM layout (one context):
//   rec  [ncell][RW] R   per-cell record gathered by neighbours: primitives
//                        rho,u,v,p | limiter values | gradient d/dx,d/dy of
//                        each | centroid (interior cells then ghosts); the
//                        gradient/limiter pass reads the first 8 values and
//                        the centroid and writes values 4..15 in one piece
//   qT   [ncell] R       temperature (fifth primitive)
//   qpc, qTc [noc] R     the interior primitives an iteration works from:
//                        written by the primitive refresh or by the last
//                        stage of the previous whole iteration, copied into
//                        rec and qT by the gradient pass
//   gT   [noc][2] R      its gradient (only when a consumer needs it)
//   qc   [ncell][4] f64  conserved state [E, rho, rho u, rho v] (ghosts last)
//   qold [noc][4] f64, res [noc][4] f64, dt/dta/la [noc] f64
//   qs   [noc][4] f64    stage states of a whole RK3 iteration (qc keeps
//                        the iteration's start state until the last stage)
//   frec [nof][FW] R     face record nx, ny, area, centre x, y
//   fcen [nof][2] R      face centre alone (for the gradient/limiter pass)
//   fnl [nof][4] R       normal and length alone (for the face spectral radii)
//   fpt  [nof] int2      face endpoints (storage numbering of points)
//   prec [nop][PW] R     vertex record: averaged primitives rho,u,v,p,T, mu
//   sl_face/sl_nbr [MAXF][noc] int  the faces of each cell in ascending
//                        reference face order, and the cell across each
//                        (bit-inverted when the cell is the face's `out`)
//   bf_ptr/bf_ab/bf_rec  per block of BLK cells, the interior faces it touches
//                        (in/out cells and face record, in list order)
//   sl_loc [MAXF][noc]   per cell slot, the face's position in its block's list
// R is double for the FP64 form and float for the bulk data of the mixed
// form; states, residuals and their sums stay double in both.  Face and
// vertex storage follows the cell order (first touch) so gathers stay local;
// the reference order of every per-cell sum is carried by the slot order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ufx {

constexpr int MAXF = 4;     // faces per cell (triangles, quadrilaterals)
constexpr int BLK = 256;    // threads per block for cell/point/face kernels
constexpr int NXCD = 8;     // MI355X: 8 accelerator complex dies, own L2 each
constexpr int WALL = 1001, FARF = 2001;

template <typename R> struct Layout;
template <> struct Layout<double> { static constexpr int RW = 18, FW = 5, PW = 6; };
template <> struct Layout<float> { static constexpr int RW = 20, FW = 5, PW = 6; };
// record offsets
constexpr int R_QP = 0, R_PHI = 4, R_G = 8, R_CEN = 16;
constexpr int F_NX = 0, F_NY = 1, F_AREA = 2, F_CX = 3, F_CY = 4;
constexpr int P_MU = 5;

// error word codes
constexpr int DERR_ROE = 1, DERR_POS = 2, DERR_AREA = 3;

// scalar slots (double) and ordered-integer range accumulators
constexpr int SC_NORM = 0, SC_FRES1 = 1, SC_DTMIN = 2, SC_NSCAL = 8;

template <typename R>
struct DevT {
  using Real = R;
  int nop, noc, nfi, nbf, nof, ncell;
  // mesh
  const R* frec;
  const R* fcen;             // [nof][2] face centres (the limiter's only face data)
  const R* fnl;              // [nof][4] nx, ny, area, 0 (the face spectral radii's face data)
  const int2* fpt;
  const int* fbc;            // bc code per boundary ordinal (nbf)
  const int *bface, *bcell;  // storage face and inner cell per boundary ordinal
  double* bres;              // [nbf][4] boundary-face fluxes of the current stage
  const int* sl_face;        // [MAXF][noc]
  const int* sl_nbr;         // [MAXF][noc]
  const int* cnf;            // nf | nb << 8
  // shared-flux residual: the interior faces touched by each block of BLK
  // cells (ascending storage order) and, per cell slot, the position of its
  // face in that list (bit-inverted when the cell is the face's `out`)
  const int *bf_ptr, *sl_loc;
  const int2* bf_ab;  // per block-list entry: the face's in and out cell
  const R* bf_rec;    // per block-list entry: the face record
  const int* bf_fref; // per block-list entry: the reference face number (error reports)
  int bf_max;
  const R *r11, *r12, *r22;  // noc: least-squares factors, in R (the gradient runs in R)
  const double *cv, *cov;    // noc (cov: ncell)
  const int* nc2c;
  const int* cref;           // storage cell -> reference cell number (ncell)
  const int *v2c_ptr, *v2c;  // point storage order
  const int* avg_w0;         // per block of BLK vertices: its lowest interior cell
  const R* wt;                // pseudo-Laplacian weights per v2c entry, in R
  const double *px, *py;     // point storage order
  const int *fin, *fout;     // face sides (storage numbering of faces)
  const int* fref;           // storage face -> reference face number (error reports)
  double* fw;                // a1 + a2 of the diamond path per face (GG_FC)
  // LU-SGS
  const int *lu_face, *lu_cnt;   // [MAXF][noc] reference c2f order
  const int* ccol;               // [noc] colour of each cell (multicolour LU-SGS)
  // state
  R* rec;
  R* qpc;                    // [noc][4] the interior cells' record primitives, compact
  R* qT;
  R* qTc;                    // [noc] the interior cells' temperatures beside qpc
  double* fvis;              // [nof][4] viscous face fluxes times the face length
  R* phc;                    // [noc][4] the interior cells' limiter values, compact (the
                             // ratchet's previous values; the records hold a copy)
  const R* cenc;             // [ncell][2] centroids, compact
  R* gT;
  double* qc;
  double* qold;
  double* qs;                // [noc][4] stage states of whole RK3 iterations
  double* res;
  double* dt;
  double* dta;
  double* la;
  double* fla;               // nof (LU-SGS)
  double* fmu;               // nof (viscous)
  R* cmul;                   // ncell (viscous)
  R* prec;                   // nop * PW
  R* fgrad;                  // nof * 10 (GG_FC)
  double* bflux;             // [2][nbf]
  double* part;              // reduction partials
  double* rpart;             // [blocks][8] block partials of the limiter's range (max q0..3, min q0..3)
  double* scal;              // SC_* slots
  unsigned long long* range; // [8]: max of q0..3 (ordered), min of q0..3
  int* err;                  // [0] code, [1] and [2] lowest 1-based reference cell with
                             // rho <= 0 or p <= 0 among the primitives the next iteration
                             // starts from ([1], noted by the last stage) and the current
                             // one started from ([2]), [3] lowest failing face key (set_err)
  double* lower;             // LU-SGS work [noc][4]
  double* upper;
};

// ---------------------------------------------------------------- loads
template <typename R, int N>
__device__ __forceinline__ void ldv(const R* __restrict__ p, R* out) {
  constexpr int B = N * (int)sizeof(R);
  if constexpr (B % 16 == 0) {
    if constexpr (sizeof(R) == 8) {
      const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
      for (int i = 0; i < N / 2; ++i) { double2 v = q[i]; out[2 * i] = v.x; out[2 * i + 1] = v.y; }
    } else {
      const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
      for (int i = 0; i < N / 4; ++i) {
        float4 v = q[i];
        out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
      }
    }
  } else if constexpr (sizeof(R) == 8 || N % 2 != 0) {
    // records of an odd number of values (aligned to one value): one load each
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = p[i];
  } else {
    const float2* q = reinterpret_cast<const float2*>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) { float2 v = q[i]; out[2 * i] = v.x; out[2 * i + 1] = v.y; }
  }
}

template <typename R, int N>
__device__ __forceinline__ void stv(R* __restrict__ p, const R* v) {
  constexpr int B = N * (int)sizeof(R);
  if constexpr (B % 16 == 0) {
    if constexpr (sizeof(R) == 8) {
      double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
      for (int i = 0; i < N / 2; ++i) q[i] = make_double2(v[2 * i], v[2 * i + 1]);
    } else {
      float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
      for (int i = 0; i < N / 4; ++i) q[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
  } else {
    float2* q = reinterpret_cast<float2*>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) q[i] = make_float2(v[2 * i], v[2 * i + 1]);
  }
}

__device__ __forceinline__ void ld4d(const double* p, double* o) {
  const double2* q = reinterpret_cast<const double2*>(p);
  double2 a = q[0], b = q[1];
  o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4d(double* p, const double* v) {
  double2* q = reinterpret_cast<double2*>(p);
  q[0] = make_double2(v[0], v[1]);
  q[1] = make_double2(v[2], v[3]);
}

// Streaming (non-temporal) forms for data a kernel touches exactly once: they
// keep such data from displacing the neighbour records in L2.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld4d_nt(const double* p, double* o) {
  const nt_d2* q = reinterpret_cast<const nt_d2*>(p);
  nt_d2 a = __builtin_nontemporal_load(q), b = __builtin_nontemporal_load(q + 1);
  o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4d_nt(double* p, const double* v) {
  nt_d2* q = reinterpret_cast<nt_d2*>(p);
  nt_d2 a = {v[0], v[1]}, b = {v[2], v[3]};
  __builtin_nontemporal_store(a, q);
  __builtin_nontemporal_store(b, q + 1);
}
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }

// Block index remapped so that each XCD (blocks are dealt to the 8 dies
// round-robin) works on one contiguous slice of the cells: neighbour data
// then mostly sits in the die's own L2.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  int q = nblk / NXCD, r = nblk % NXCD;
  int x = b % NXCD, i = b / NXCD;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// The same with each XCD walking its slice from the far end when `rev`: a
// kernel that follows one which walked forward then starts on the data the
// other touched last, which the memory-side cache still holds.
__device__ __forceinline__ int xcd_block_dir(int b, int nblk, bool rev) {
  int q = nblk / NXCD, r = nblk % NXCD;
  int x = b % NXCD, i = b / NXCD;
  if (rev) i = (x < r ? q + 1 : q) - 1 - i;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

__device__ __forceinline__ int cell_index() {
  return xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
}

// The block a kernel body works on: the hardware block of a one-context
// launch, or a (case, block) entry of a many-case launch's block table.
// b: block index among the context's blocks of this kernel, n: their count.
struct VBlock {
  int b, n;
  __device__ int cell() const { return xcd_block(b, n) * BLK + (int)threadIdx.x; }
};
__device__ __forceinline__ VBlock hw_block() { return VBlock{(int)blockIdx.x, (int)gridDim.x}; }

// Monotone map of doubles onto unsigned integers, for atomic max/min.
__device__ __forceinline__ unsigned long long ord_of(double x) {
  long long i = __double_as_longlong(x);
  return i >= 0 ? (unsigned long long)i ^ 0x8000000000000000ULL : ~(unsigned long long)i;
}
__host__ __device__ inline double of_ord(unsigned long long u) {
  long long i = (u & 0x8000000000000000ULL) ? (long long)(u ^ 0x8000000000000000ULL)
                                            : (long long)~u;
  double d;
  static_assert(sizeof(d) == sizeof(i), "");
  __builtin_memcpy(&d, &i, sizeof d);
  return d;
}

// Deterministic block reduction (fixed shared-memory tree).
template <int N, typename Op>
__device__ __forceinline__ void block_reduce(double (*sh)[BLK], double* v, Op op) {
  int t = threadIdx.x;
  __syncthreads();
  for (int k = 0; k < N; ++k) sh[k][t] = v[k];
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < N; ++k) sh[k][t] = op(sh[k][t], sh[k][t + s]);
    __syncthreads();
  }
  for (int k = 0; k < N; ++k) v[k] = sh[k][0];
}
struct OpMax { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
struct OpMin { __device__ double operator()(double a, double b) const { return a < b ? a : b; } };
struct OpAdd { __device__ double operator()(double a, double b) const { return a + b; } };

// Record an error of kind `code` at face `key`.  The first kind to arrive
// wins; among failures of that kind the lowest key is kept, so the report
// names the face the reference's sequential loop would have stopped at:
// Roe, key = boundary ordinal for boundary faces (looped first,
// CCFVM/fvresidual.F90:27) and nbf + reference face for interior faces
// (:133); diamond-path area, key = reference face.
template <typename R>
__device__ __forceinline__ void set_err(const DevT<R>& d, int code, int key) {
  int old = atomicCAS(d.err, 0, code);
  if (old == 0 || old == code) atomicMin(&d.err[3], key);
}

}  // namespace ufx
