// Host setup of one instance: geometry, stencils, least-squares
// pseudo-inverses and the structure-of-arrays image the kernels read.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace zfv {

// Compile-time shape family of the kernels (see kernels.hip for the
// instantiations that exist).
struct Shape {
  int dim = 2;      // 2 or 3
  int nvars = 4;    // conserved variables stored: rho, m (dim), E
  int nfaces = 3;   // faces per cell
  int q = 2;        // quadrature points per face
  int qv = 6;       // quadrature points per cell (volume rule)
  int dofc = 6;     // dof of the highest-order (central) stencil's polynomial
  int rc = 10;      // rows of the central LSQ system (members - 1)
  int nb = 3;       // number of one-sided stencils
  int dofb = 3;     // dof of the one-sided polynomials
  int rb = 3;       // rows of each one-sided LSQ system
  int m = 68;       // pseudo-inverse entries per cell
  int nrows = 19;   // stencil rows per cell = rc + nb*rb
  int nmom = 3;     // stored moments (indices 3 .. dofc-1)
};

// Per-cell meta bits
enum : int32_t {
  META_KHIGH_MASK = 0x7,      // index of the stencil that receives the CWENO correction
  META_FIRST_ORDER = 1 << 3,  // single one-cell stencil: constant reconstruction
  META_GHOST = 1 << 4,        // frozen ghost cell
};

// cface bits: (face_id << 2) | (side << 0) | (exterior << 1)
constexpr int CFACE_SIDE_R = 1;
constexpr int CFACE_EXTERIOR = 2;
// a face between two ghost cells: no flux and no flux BC, geometry only
// (include/zisa/fvm_loops/flux_loop.hpp:71-83 leaves such faces out)
constexpr int CFACE_GHOST_PAIR = CFACE_EXTERIOR | CFACE_SIDE_R;

struct HostImage {
  Shape s;
  int64_t n = 0;   // cells
  int64_t ne = 0;  // interior faces
  std::vector<double> P;         // m * n
  std::vector<int32_t> rows;     // nrows * n   (product numbering)
  // one-sided row r (rc <= r < nrows) whose member (cell and translation) is
  // also central row twin[(r - rc) * n + i]: its equilibrium average is the
  // central row's; -1 if none
  std::vector<int8_t> twin;      // (nrows - rc) * n
  std::vector<double> crn;       // (nfaces*dim) * n, corner k coordinate d at (k*dim+d)
  std::vector<double> xc;        // dim * n
  std::vector<double> len, vol, inr;
  std::vector<double> mom;       // nmom * n
  std::vector<int32_t> meta;     // n
  std::vector<int32_t> cface;    // 4 * n: per cell its face words (2D: the fourth unused)
  // the cell-face slots (i * nfaces + k, product numbering) of each interior
  // face's left (side 0, first ne entries) and right (side 1) cell: where the
  // reconstruction leaves the face's point values
  std::vector<int32_t> fslot;    // ne pairs (left slot, right slot)
  std::vector<uint32_t> fperm;   // n: own-corner index of the face's left-cell vertex order
  std::vector<int32_t> perm;     // product index -> caller index
  std::vector<int32_t> iperm;    // caller index -> product index
  std::vector<uint8_t> ghost;    // product numbering
  // reference-layout bookkeeping for zfv_get_stencils (caller numbering)
  std::vector<int32_t> l2g_off, l2g;               // CSR of the combined stencil
  std::vector<int32_t> mem_off, members;           // CSR of [order, size, ids...]
  // normalized quadrature rules used by the kernels (device constants)
  std::vector<double> face_w, face_x;  // q weights; face_x: q * (dim) barycentric / GL coordinates
  std::vector<double> vol_w, vol_x;    // qv weights; vol_x: qv * (dim+1) barycentric
  // periodic translations of stencil members: row code c (bits 27..30 of a
  // row entry) adds shift[3c..3c+2] to the member's coordinates; code 0 = none
  std::vector<double> shift = std::vector<double>(48, 0.0);
  double max_cond = 0.0;               // largest estimated condition number of A^T A
  double linear_weights[ZFV_MAX_STENCILS] = {0};
};

// Hilbert index of integer coordinates with `bits` bits per axis (bits <= 32
// in 2D, <= 21 in 3D): the cell order of the instances and of the mesh
// renumbering tool.
uint64_t hilbert2(uint32_t x, uint32_t y, int bits);
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits);

// Builds everything on the host.  Throws ZfvError on unsupported input.
void build_image(const zfv_mesh &mesh, const zfv_params &p, HostImage &img);

// Shape implied by the parameters (throws if the GPU kernels have no
// instantiation for it).
Shape shape_for(const zfv_params &p);
// true when the volume rule and the cell moments both integrate the central
// polynomial's basis exactly: then the quadrature mean of a reconstruction
// over its cell is its constant coefficient (every other basis function has
// mean zero), which the light set's gravity volume term uses
bool volume_mean_is_constant_term(const zfv_params &p, const Shape &s);

// The caller's n x 5 row-major state (AllVariables::cvars) to and from the
// product's nvars x n planes in product order (perm: product index -> caller
// index), on the host's threads.  from_soa adds into aos when accumulate.
void to_soa(const HostImage &img, const double *aos, double *soa);
void from_soa(const HostImage &img, const double *soa, double *aos, bool accumulate);

}  // namespace zfv
