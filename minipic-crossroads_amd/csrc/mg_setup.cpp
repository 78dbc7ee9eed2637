// Host setup of the multigrid preconditioner of the precision form's solve
// (mg.hpp).  Levels are nested node lattices of the brick mesh: level 0 is the
// mesh itself (nodes (i, j, k), 0 <= i <= nx, ..., numbered as the brick
// generator numbers them, mesh.cpp:52-54); level l + 1 keeps every second
// point of level l: a level of c cells per direction has interior points
// 1..c - 1, and its coarse level has floor((c + 1) / 2) cells, coarse point I
// sitting on fine point 2I (an odd c leaves a last coarse cell of one fine
// cell).  Only interior points are unknowns (Dirichlet walls are zero at every
// level).
//
// Interpolation P from level l + 1 to level l is linear per direction on the
// lattice: a fine point (i, j, k) takes, per direction, the coarse point i/2
// (i even, weight 1) or the two points (i -+ 1)/2 (i odd, weight 1/2 each);
// coarse points on the walls contribute nothing.  Restriction is P^T (full
// weighting: coarse point I gathers the interior fine points among 2I - 1, 2I,
// 2I + 1 per direction with weights 1/2, 1, 1/2).  The coarse operator is the Galerkin product
// A_{l+1} = P^T A_l P, formed here in FP64 over the interior unknowns; the
// stencil of a tet mesh on a brick lattice couples points at most one cell
// apart in each direction, so every coarse operator is a 27-point stencil.
#include "mg.hpp"

#include <cmath>
#include <cstring>
#include <string>

namespace mpic {
namespace {

// A level's points (interior lattice of cells (cx, cy, cz) at level >= 1, the
// full node lattice at level 0) and the index of a point.
struct Grid {
  int cx, cy, cz;
  bool full;  // level 0: all nodes, numbered (i * (ny + 1) + j) * (nz + 1) + k
  int64_t index(int i, int j, int k) const {
    if (full) return ((int64_t)i * (cy + 1) + j) * (cz + 1) + k;
    return ((int64_t)(i - 1) * (cy - 1) + (j - 1)) * (cz - 1) + (k - 1);
  }
  bool interior(int i, int j, int k) const { return i > 0 && i < cx && j > 0 && j < cy && k > 0 && k < cz; }
  int64_t count() const {
    return full ? (int64_t)(cx + 1) * (cy + 1) * (cz + 1) : (int64_t)(cx - 1) * (cy - 1) * (cz - 1);
  }
};

// A level's operator over its interior points as a 27-point stencil:
// st[(index) * 27 + (di + 1) * 9 + (dj + 1) * 3 + (dk + 1)].
using Stencil = std::vector<double>;

// Level 0's stencil from the CSR rows of the interior nodes (columns on the
// walls dropped: their values are zero in every preconditioner vector).
bool stencil_from_csr(const Grid& g, const int32_t* row_ptr, const int32_t* col, const double* val, Stencil& st,
                      std::string& why) {
  const int64_t n = g.count();
  st.assign((size_t)n * 27, 0.0);
  const int ny1 = g.cy + 1, nz1 = g.cz + 1;
  for (int i = 1; i < g.cx; ++i)
    for (int j = 1; j < g.cy; ++j)
      for (int k = 1; k < g.cz; ++k) {
        const int64_t r = g.index(i, j, k);
        for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
          const int c = col[e];
          const int ci = c / (ny1 * nz1), cj = (c / nz1) % ny1, ck = c % nz1;
          const int di = ci - i, dj = cj - j, dk = ck - k;
          if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) {
            why = "a matrix entry couples lattice points more than one cell apart";
            return false;
          }
          if (!g.interior(ci, cj, ck)) continue;
          st[(size_t)r * 27 + (di + 1) * 9 + (dj + 1) * 3 + (dk + 1)] += val[e];
        }
      }
  return true;
}

inline double w1(int d) { return d == 0 ? 1.0 : 0.5; }

// A_c = P^T A P for coarse grid gc of fine grid gf.
void galerkin(const Grid& gf, const Stencil& af, const Grid& gc, Stencil& ac) {
  ac.assign((size_t)gc.count() * 27, 0.0);
  for (int I = 1; I < gc.cx; ++I)
    for (int J = 1; J < gc.cy; ++J)
      for (int K = 1; K < gc.cz; ++K) {
        double acc[27] = {0};
        // fine points f = 2I + d (2I + 1 is the fine wall when the fine cell count is odd)
        for (int a = -1; a <= 1; ++a)
          for (int b = -1; b <= 1; ++b)
            for (int c = -1; c <= 1; ++c) {
              const int fi = 2 * I + a, fj = 2 * J + b, fk = 2 * K + c;
              if (!gf.interior(fi, fj, fk)) continue;
              const double wf = w1(a) * w1(b) * w1(c);
              const double* row = &af[(size_t)gf.index(fi, fj, fk) * 27];
              for (int s = 0; s < 27; ++s) {
                const double v = row[s];
                if (v == 0.0) continue;
                const int gi = fi + s / 9 - 1, gj = fj + (s / 3) % 3 - 1, gk = fk + s % 3 - 1;
                // coarse parents of g (interior ones only)
                const int i0 = gi >> 1, i1 = (gi + 1) >> 1, j0 = gj >> 1, j1 = (gj + 1) >> 1, k0 = gk >> 1,
                          k1 = (gk + 1) >> 1;
                for (int pi = i0; pi <= i1; ++pi)
                  for (int pj = j0; pj <= j1; ++pj)
                    for (int pk = k0; pk <= k1; ++pk) {
                      if (!gc.interior(pi, pj, pk)) continue;
                      const double wg = w1(gi - 2 * pi) * w1(gj - 2 * pj) * w1(gk - 2 * pk);
                      acc[(pi - I + 1) * 9 + (pj - J + 1) * 3 + (pk - K + 1)] += wf * v * wg;
                    }
              }
            }
        std::memcpy(&ac[(size_t)gc.index(I, J, K) * 27], acc, sizeof(acc));
      }
}

// Inverse of the SPD matrix a (n x n, row-major) by Cholesky; false if not SPD.
bool spd_inverse(int n, std::vector<double>& a) {
  std::vector<double> L((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = a[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(d > 0)) return false;
    const double ljj = std::sqrt(d);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = a[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / ljj;
    }
  }
  // inverse column by column: L L^T x = e_c
  std::vector<double> inv((size_t)n * n), y(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
      y[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * inv[(size_t)k * n + c];
      inv[(size_t)i * n + c] = s / L[(size_t)i * n + i];
    }
  }
  a.swap(inv);
  return true;
}

}  // namespace

bool mg_build(int nx, int ny, int nz, int32_t nn, const uint8_t* boundary, const int32_t* row_ptr, const int32_t* col,
              const double* val, MGHost& out, std::string& why, int direct_max) {
  out = MGHost{};
  if (direct_max < 1 || direct_max > MG_DIRECT_MAX) {
    why = "bad size of the directly solved level";
    return false;
  }
  if (nx < 3 || ny < 3 || nz < 3 || (int64_t)(nx + 1) * (ny + 1) * (nz + 1) != nn) {
    why = "the mesh is not a brick lattice of at least 3 cells per direction";
    return false;
  }
  const Grid g0{nx, ny, nz, true};
  // the walls of the lattice are exactly the Dirichlet nodes
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j <= ny; ++j)
      for (int k = 0; k <= nz; ++k) {
        const bool wall = !g0.interior(i, j, k);
        if ((boundary[g0.index(i, j, k)] != 0) != wall) {
          why = "the Dirichlet nodes are not the walls of the lattice";
          return false;
        }
      }
  // the V-cycle keeps the walls at zero: their rows must be identity rows
  for (int32_t r = 0; r < nn; ++r) {
    if (!boundary[r]) continue;
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
      if ((col[e] == r) != (val[e] != 0.0)) {
        why = "a Dirichlet row is not an identity row";
        return false;
      }
  }
  out.nx = nx;
  out.ny = ny;
  out.nz = nz;
  Grid gf = g0;
  Stencil af;
  if (!stencil_from_csr(g0, row_ptr, col, val, af, why)) return false;
  while (gf.cx >= 3 && gf.cy >= 3 && gf.cz >= 3 && (int)out.levels.size() < MG_MAX_LEVELS) {
    const Grid gc{(gf.cx + 1) / 2, (gf.cy + 1) / 2, (gf.cz + 1) / 2, false};
    Stencil ac;
    galerkin(gf, af, gc, ac);
    MGHostLevel L;
    L.cx = gc.cx;
    L.cy = gc.cy;
    L.cz = gc.cz;
    L.n = (int32_t)gc.count();
    // 27-point ELL (column-major), padding (col = row, val = 0); diagonal inverse
    L.col.assign((size_t)27 * L.n, 0);
    L.val.assign((size_t)27 * L.n, 0.0f);
    L.dinv.assign(L.n, 0.0f);
    for (int I = 1; I < gc.cx; ++I)
      for (int J = 1; J < gc.cy; ++J)
        for (int K = 1; K < gc.cz; ++K) {
          const int64_t r = gc.index(I, J, K);
          for (int s = 0; s < 27; ++s) {
            const int ci = I + s / 9 - 1, cj = J + (s / 3) % 3 - 1, ck = K + s % 3 - 1;
            const double v = ac[(size_t)r * 27 + s];
            const bool in = gc.interior(ci, cj, ck);
            L.col[(size_t)s * L.n + r] = in ? (int32_t)gc.index(ci, cj, ck) : (int32_t)r;
            L.val[(size_t)s * L.n + r] = in ? (float)v : 0.0f;
          }
          const double d = ac[(size_t)r * 27 + 13];
          if (!(d > 0)) {
            why = "a coarse operator has a non-positive diagonal";
            return false;
          }
          L.dinv[r] = (float)(1.0 / d);
        }
    out.levels.push_back(std::move(L));
    gf = gc;
    af.swap(ac);
    if (gf.count() <= direct_max) break;
  }
  const int64_t nc = gf.count();
  if (out.levels.empty() || nc > direct_max) {
    why = "the coarsest level has " + std::to_string(nc) +
          " unknowns (a direction with fewer than 3 cells stops the coarsening)";
    return false;
  }
  // dense inverse of the coarsest operator
  std::vector<double> dense((size_t)nc * nc, 0.0);
  for (int I = 1; I < gf.cx; ++I)
    for (int J = 1; J < gf.cy; ++J)
      for (int K = 1; K < gf.cz; ++K) {
        const int64_t r = gf.index(I, J, K);
        for (int s = 0; s < 27; ++s) {
          const int ci = I + s / 9 - 1, cj = J + (s / 3) % 3 - 1, ck = K + s % 3 - 1;
          if (gf.interior(ci, cj, ck)) dense[(size_t)r * nc + gf.index(ci, cj, ck)] += af[(size_t)r * 27 + s];
        }
      }
  if (!spd_inverse((int)nc, dense)) {
    why = "the coarsest operator is not positive definite";
    return false;
  }
  // column-major: the device's thread i reads element (i, j) at j * nc + i, so
  // the threads of a wavefront read consecutive words
  out.coarse_inv.resize(dense.size());
  for (int64_t i = 0; i < nc; ++i)
    for (int64_t j = 0; j < nc; ++j) out.coarse_inv[(size_t)j * nc + i] = (float)dense[(size_t)i * nc + j];
  // level 0's diagonal inverse (all nodes; the Dirichlet rows are identity rows)
  out.dinv0.assign(nn, 1.0f);
  for (int32_t r = 0; r < nn; ++r)
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
      if (col[e] == r && val[e] != 0.0) out.dinv0[r] = (float)(1.0 / val[e]);
  // level 0 as a stencil over all nodes (interior rows were checked to couple
  // neighbours at most one cell apart; wall rows are identity rows)
  {
    const int ny1 = ny + 1, nz1 = nz + 1;
    bool used[27] = {false};
    auto slot_of = [&](int32_t r, int32_t c) {
      const int ri = r / (ny1 * nz1), rj = (r / nz1) % ny1, rk = r % nz1;
      const int ci = c / (ny1 * nz1), cj = (c / nz1) % ny1, ck = c % nz1;
      return (ci - ri + 1) * 9 + (cj - rj + 1) * 3 + (ck - rk + 1);
    };
    for (int32_t r = 0; r < nn; ++r)
      for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
        if (val[e] != 0.0) used[slot_of(r, col[e])] = true;
    for (int s = 0; s < 27; ++s)
      if (used[s]) out.slot0.push_back(s);
    int idx[27];
    for (size_t k = 0; k < out.slot0.size(); ++k) idx[out.slot0[k]] = (int)k;
    out.val0.assign(out.slot0.size() * (size_t)nn, 0.0);
    for (int32_t r = 0; r < nn; ++r)
      for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
        if (val[e] != 0.0) out.val0[(size_t)idx[slot_of(r, col[e])] * nn + r] += val[e];
  }
  return true;
}

}  // namespace mpic
