// Internal host-side types of the ufx library (not part of the C-ABI).
#pragma once

#include <string>
#include <vector>

#include "../../include/ufx.h"

namespace ufx {

// Raw mesh as the reference reads it from geometry.inp
// (CCFVM/read_input.F90:160-173); indices 1-based.
struct RawInput {
  int nop = 0, noc = 0, nfi = 0, nbf = 0;
  const double *px = nullptr, *py = nullptr;
  const double *ccx = nullptr, *ccy = nullptr, *cv = nullptr;
  const int *fi_p1 = nullptr, *fi_p2 = nullptr, *fi_in = nullptr, *fi_out = nullptr,
            *fi_bc = nullptr;
  const int *fb_p1 = nullptr, *fb_p2 = nullptr, *fb_in = nullptr, *fb_bc = nullptr;
  const double *fb_nx = nullptr, *fb_ny = nullptr;
};

// A preprocessed mesh, 0-based.  Cells 0..noc-1 are interior (reference
// numbering after its breadth-first renumbering), noc..noc+nbf-1 are ghost
// cells, ghost k belonging to boundary face nfi+k.  Faces 0..nfi-1 interior,
// nfi..nfi+nbf-1 boundary, in the reference's face order.
struct Mesh {
  int kind = 0;
  int nop = 0, noc = 0, nfi = 0, nbf = 0;
  std::vector<double> px, py;
  std::vector<int> fp1, fp2, fin, fout, fbc;
  std::vector<double> fnx, fny, farea, fcx, fcy, fcov;
  std::vector<double> ccx, ccy, ccov;   // ncell
  std::vector<double> ccv, r11, r12, r22;  // noc
  std::vector<int> nc2c;                // noc
  std::vector<int> c2f_ptr, c2f;        // LU-SGS face order per cell
  std::vector<int> c2v_ptr, c2v;        // vertex walk round each cell, closed (first repeated)
  std::vector<int> v2c_ptr, v2c;        // vertex -> cells (ghosts included)
  std::vector<double> wt;               // pseudo-Laplacian weights per v2c entry
  std::vector<int> oldnum;              // new cell -> input cell (0-based)
  // vertex-centred meshes (kind UFX_KIND_NC): every edge is a face (nfi =
  // nof, nbf = 0), fin/fout are the cells left/right of p1->p2 (-1 = none),
  // vertices are in the reference's breadth-first order
  std::vector<int> pbc;                 // vertex tag
  std::vector<double> pcv;              // median-dual volume
  std::vector<int> v2v_ptr, v2v;        // vertex neighbours (edge order)
  std::vector<double> wx, wy;           // least-squares weights per v2v entry
  std::vector<int> v2f_ptr, v2f;        // edges of a vertex (edge order)
  std::vector<int> pt_oldnum;           // new vertex -> input vertex (0-based)
  int nof() const { return nfi + nbf; }
  int ncell() const { return noc + nbf; }
};

int cc_geometric(const RawInput& in, Mesh& m, std::string& err);
// Its phases (cc_geometric = read, faces, connect, cells): the checked input
// with ghosts attached; per face the ghost centres and cell_face_norm; the
// connectivity and breadth-first numbering (vertex -> cell lists into
// m.v2c_ptr / m.v2c); per cell and vertex the co-volumes, least-squares
// factors and pseudo-Laplacian weights.  The per-face and per-cell phases
// also run on the GPU (cc_geometric_device, cc_mesh_dev.hip).
int cc_geom_read(const RawInput& in, Mesh& m, std::string& err);
void cc_geom_faces(Mesh& m);
int cc_geom_connect(Mesh& m, std::string& err);
int cc_geom_cells(Mesh& m, std::string& err);
int cc_geometric_device(const RawInput& in, Mesh& m, std::string& err);

// Raw vertex-centred mesh as NCFVM/read_input.f90:101-112 reads it (1-based,
// 0 = no cell on that side of an edge).
struct NCRawInput {
  int nop = 0, noc = 0, nof = 0;
  const double *px = nullptr, *py = nullptr;
  const int* pbc = nullptr;
  const double *ccx = nullptr, *ccy = nullptr, *cv = nullptr;
  const int *e_p1 = nullptr, *e_p2 = nullptr, *e_in = nullptr, *e_out = nullptr,
            *e_bc = nullptr;
};

int nc_geometric(const NCRawInput& in, Mesh& m, std::string& err);
// Its phases: connectivity and numbering, then the per-edge and per-vertex
// geometry (also on the GPU, nc_geometric_device, nc_mesh_dev.hip).
int nc_geom_connect(const NCRawInput& in, Mesh& m, std::string& err);
int nc_geom_numeric(Mesh& m, std::string& err);
int nc_geometric_device(const NCRawInput& in, Mesh& m, std::string& err);

// Greedy distance-1 colouring of a graph given as adjacency lists (ptr has
// n + 1 offsets) in smallest-last order (the vertices taken out one by one,
// each time one of least remaining degree, the lowest-numbered on ties; then
// coloured in the reverse of that order, each taking the smallest colour its
// already coloured neighbours leave free).  No edge joins two vertices of one
// colour, so the vertices of a colour can be updated at once in a
// Gauss-Seidel sweep ordered by colour; the order keeps the colours few and
// even (four on triangulated O-grids, where index order leaves six or seven,
// the last ones nearly empty).  Returns colours 0..*ncolours-1.
std::vector<int> greedy_colour(int n, const std::vector<int>& ptr, const std::vector<int>& adj,
                               int* ncolours);
// The colouring the LU-SGS sweeps of the multicolour form use: vertex-centred
// meshes colour their vertices over the edges; cell-centred meshes colour the
// interior cells over the interior faces.
std::vector<int> lusgs_colours(const Mesh& m, int* ncolours);

// Points (x[i], y[i]) ordered along a Hilbert curve of order 16 over their
// bounding square (ties keep the given order): storage orders whose blocks of
// consecutive points are compact patches of the mesh.
std::vector<int> hilbert_order(const double* x, const double* y, int n);

void set_error(const std::string& msg);

}  // namespace ufx

struct ufx_mesh {
  ufx::Mesh m;
};
