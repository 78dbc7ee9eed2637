// Multigrid preconditioner of the precision form's solve: host hierarchy
// (mg_setup.cpp, plain C++) and the device solve (k_mg.hip).
//
// The precision form solves A phi = b to the reference's stopping rule
// |r| < tol |b| (CG_solve.hpp:71-72,103-107) with conjugate gradients in FP64
// preconditioned by one FP32 multigrid V-cycle per iteration: damped Jacobi
// (one sweep before and one after the coarse correction), trilinear
// interpolation on nested node lattices, Galerkin coarse operators and a dense
// inverse on the coarsest level.  The V-cycle is a fixed symmetric linear
// operator (up to FP32 rounding), so CG keeps its short recurrences; the FP64
// residual decides convergence.  Used when the mesh is a brick lattice (its
// descriptor names the lattice) whose cell counts halve to a small coarsest
// level; otherwise the refinement CG runs.
#ifndef MPIC_MG_HPP_
#define MPIC_MG_HPP_

#include <stdint.h>

#include <string>
#include <vector>

namespace mpic {

constexpr int MG_DIRECT_MAX = 512;  // unknowns of a level solved by its dense inverse
constexpr float MG_OMEGA = 0.8f;    // Jacobi damping
constexpr int MG_MAX_LEVELS = 12;
constexpr int MG_MAX_RECHECKS = 3;  // restarts from the true residual per solve

// One coarse level: interior points of a lattice of (cx, cy, cz) cells,
// numbered ((i - 1) (cy - 1) + (j - 1)) (cz - 1) + (k - 1); its operator as a
// 27-point stencil, column-major ([27][n]), entry (di + 1) 9 + (dj + 1) 3 +
// (dk + 1) coupling the point to its neighbour at offset (di, dj, dk); zero
// where that neighbour is not an interior point.  The columns (col; the row
// itself where the value is zero) serve the host-side checks; the device
// kernels take a row's neighbours from the stencil offsets.
struct MGHostLevel {
  int cx = 0, cy = 0, cz = 0;
  int32_t n = 0;
  std::vector<int32_t> col;
  std::vector<float> val, dinv;
};

struct MGHost {
  int nx = 0, ny = 0, nz = 0;         // level 0 cells (the mesh)
  std::vector<float> dinv0;           // [nn] level 0 diagonal inverse
  // Level 0's matrix as a stencil over all nodes: slot s couples node n to
  // node n + (di (ny + 1) + dj) (nz + 1) + dk, where slot0[s] = (di + 1) 9 +
  // (dj + 1) 3 + (dk + 1); the slots any row uses, in ascending order (so a
  // row's entries run in ascending column order); values [S][nn], zero where
  // the row has no such entry (wall rows keep only their diagonal).
  std::vector<int> slot0;
  std::vector<double> val0;
  std::vector<MGHostLevel> levels;    // levels 1..L; the last is solved directly
  std::vector<float> coarse_inv;      // inverse of the coarsest operator, column-major ([j][i] = (i, j))
};

// Builds the hierarchy from the FP64 stiffness matrix in CSR (rows of the
// brick's nodes, numbered (i (ny + 1) + j) (nz + 1) + k; Dirichlet rows are
// identity rows), coarsening until a level has at most direct_max unknowns
// (<= MG_DIRECT_MAX).  false (with the reason) when multigrid does not apply.
bool mg_build(int nx, int ny, int nz, int32_t nn, const uint8_t* boundary, const int32_t* row_ptr, const int32_t* col,
              const double* val, MGHost& out, std::string& why, int direct_max = MG_DIRECT_MAX);

}  // namespace mpic

#endif  // MPIC_MG_HPP_
