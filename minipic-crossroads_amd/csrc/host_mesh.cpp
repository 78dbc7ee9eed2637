// Host-side mesh and discretisation setup, done once per run:
//   Mesh::Mesh      (mesh.cpp:78-714)  brick generator, face neighbours, normals, Jacobians
//   ES::ES          (ES.cpp:47-213)    P1 stiffness CSR with identity Dirichlet rows, weighted basis values
//   PIC::PIC        (PIC.cpp:203-242)  physical basis gradients
//   ES::set_analytic_RHS (ES.cpp:235-257) for a constant background charge.
// Floating-point expressions keep the reference's association; the library is
// built with -ffp-contract=off.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "minipic_c.h"
#include "p1_cubature.hpp"

namespace mpic {
std::string& global_error();
}

struct mpic_mesh {
  int nx, ny, nz, split;
  double lo[3], hi[3];
  int32_t nn = 0, ne = 0;
  std::vector<double> nodes, normals, jinv, detj, grad, wbv, val, background;
  std::vector<int32_t> elem_nodes, nbrs, row_ptr, col;
  std::vector<uint8_t> boundary;
  bool has_background = false;
  bool periodic = false;             // joined in x (the two-stream mesh)
  std::vector<int32_t> node_pair;    // periodic: image node -> partner on x = lo, else -1
  // The row that takes node n's charge and stiffness (Mesh::node_pairs).
  int32_t row_of(int32_t n) const { return periodic && node_pair[n] >= 0 ? node_pair[n] : n; }
};

namespace {

// Corner c of a hex (i,j,k): c = 4*di + 2*dj + dk, the numbering of mesh.cpp:200-208.
// Tet tables: the reference's alternating 5-tet split (mesh.cpp:211-223) and a
// 6-tet Kuhn split along the corner 0 -> corner 7 diagonal (odd axis
// permutations have their middle vertices swapped so det J > 0).
constexpr int kSplit5Odd[5][4] = {{3, 1, 0, 5}, {6, 4, 5, 0}, {6, 2, 0, 3}, {5, 7, 6, 3}, {5, 6, 0, 3}};
constexpr int kSplit5Even[5][4] = {{2, 3, 1, 7}, {7, 6, 4, 2}, {2, 0, 4, 1}, {4, 5, 7, 1}, {2, 4, 7, 1}};
constexpr int kKuhn[6][4] = {{0, 4, 6, 7}, {0, 5, 4, 7}, {0, 6, 2, 7}, {0, 2, 3, 7}, {0, 1, 5, 7}, {0, 3, 1, 7}};
// Local faces (mesh.cpp:238-250): face f holds these local nodes.
constexpr int kFaces[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
// Face normal f = (n[C]-n[B]) x (n[A]-n[B]) (mesh.cpp:654-666).
constexpr int kNa[4] = {0, 1, 0, 0}, kNb[4] = {1, 2, 3, 2}, kNc[4] = {3, 3, 2, 1};
// P1 gradients on the reference tet (the cubature rule: p1_cubature.hpp).
using mpic::kGradRef;
using mpic::kCub;
using mpic::kCubW;

// Mesh::apply_mapping, mesh.cpp:763-792: smooth maps of the brick applied to
// the node coordinates once topology and boundary sets exist (mesh.cpp:587-588).
void map_node(double* x, int mapping) {
  const double X = x[0], Y = x[1], Z = x[2];
  switch (mapping) {
    case 1: {
      const double r = (1.0 - std::fabs(X)) * (1.0 - std::fabs(Y));
      x[0] = X * std::cos(r) - Y * std::sin(r);
      x[1] = X * std::sin(r) + Y * std::cos(r);
    } break;
    case 2:
      x[0] = X * std::sqrt(1 - .5 * Y * Y - .5 * Z * Z + Y * Y * Z * Z / 3.);
      x[1] = Y * std::sqrt(1 - .5 * X * X - .5 * Z * Z + X * X * Z * Z / 3.);
      x[2] = Z * std::sqrt(1 - .5 * Y * Y - .5 * X * X + Y * Y * X * X / 3.);
      break;
    case 3:
      x[0] = X < 0 ? -(X * X) : X * X;
      break;
    default:
      break;
  }
}

uint64_t jitter_next(uint64_t& s) {
  s = 6364136223846793005ull * s + 1442695040888963407ull;
  return s;
}

class BrickBuilder {
 public:
  explicit BrickBuilder(mpic_mesh& m) : m_(m) {}

  bool run(double jitter, uint64_t jseed, int mapping) {
    make_nodes(jitter, jseed);
    if (m_.periodic) make_pairs();
    if (mapping != 0)
      for (int32_t g = 0; g < m_.nn; ++g) map_node(&m_.nodes[(size_t)g * 3], mapping);
    make_elements();
    match_faces();
    if (!element_geometry()) return false;
    assemble();
    return true;
  }

 private:
  mpic_mesh& m_;

  int64_t node_id(int i, int j, int k) const {
    return (int64_t)i * (m_.ny + 1) * (m_.nz + 1) + (int64_t)j * (m_.nz + 1) + k;  // mesh.cpp:52-54
  }

  void make_nodes(double jitter, uint64_t jseed) {
    const int n[3] = {m_.nx, m_.ny, m_.nz};
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = (m_.hi[a] - m_.lo[a]) / n[a];
    m_.nn = (int32_t)((int64_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1));
    m_.nodes.assign((size_t)m_.nn * 3, 0.0);
    m_.boundary.assign(m_.nn, 0);
    for (int i = 0; i <= n[0]; ++i)
      for (int j = 0; j <= n[1]; ++j)
        for (int k = 0; k <= n[2]; ++k) {
          const int64_t g = node_id(i, j, k);
          double* x = &m_.nodes[g * 3];
          const int idx[3] = {i, j, k};
          bool on_wall = false;
          for (int a = 0; a < 3; ++a) {
            x[a] = m_.lo[a] + idx[a] * h[a];  // -1 + i*dx in mesh.cpp:141
            // the x walls are joined, not Dirichlet, in a periodic mesh (mesh.cpp:410)
            on_wall = on_wall || ((a > 0 || !m_.periodic) && (idx[a] == 0 || idx[a] == n[a]));
          }
          m_.boundary[g] = on_wall;  // nodeset, mesh.cpp:408-412
          if (on_wall || jitter == 0.0) continue;
          uint64_t s = (uint64_t)(g + 1) * 0x9E3779B97F4A7C15ull + jseed;
          for (int t = 0; t < 4; ++t) jitter_next(s);
          for (int a = 0; a < 3; ++a) {
            const double u = (double)jitter_next(s) * (1.0 / 18446744073709551616.0);
            x[a] = x[a] + (jitter * h[a]) * (2.0 * u - 1.0);
          }
        }
  }

  // Mesh::node_pairs (mesh.cpp:178-189): a node on x = hi away from the y and z
  // walls is the image of the node on x = lo with the same (j, k).
  void make_pairs() {
    m_.node_pair.assign(m_.nn, -1);
    for (int j = 1; j < m_.ny; ++j)
      for (int k = 1; k < m_.nz; ++k) m_.node_pair[node_id(m_.nx, j, k)] = (int32_t)node_id(0, j, k);
  }

  void make_elements() {
    const int per = m_.split;
    m_.ne = (int32_t)((int64_t)m_.nx * m_.ny * m_.nz * per);
    m_.elem_nodes.assign((size_t)m_.ne * 4, 0);
    int64_t e = 0;
    for (int i = 0; i < m_.nx; ++i)
      for (int j = 0; j < m_.ny; ++j)
        for (int k = 0; k < m_.nz; ++k) {
          int32_t corner[8];
          for (int c = 0; c < 8; ++c) corner[c] = (int32_t)node_id(i + ((c >> 2) & 1), j + ((c >> 1) & 1), k + (c & 1));
          const int(*table)[4] = per == 6 ? kKuhn : ((i + j + k) % 2 != 0 ? kSplit5Odd : kSplit5Even);
          for (int t = 0; t < per; ++t, ++e)
            for (int v = 0; v < 4; ++v) m_.elem_nodes[e * 4 + v] = corner[table[t][v]];
        }
  }

  // Faces shared by two elements are neighbours; unmatched faces are walls (-1).
  // Periodic: a face with all three nodes on x = hi is keyed by its x = lo
  // images (Mesh::periodic_face, mesh.cpp:62-76), so it matches the face across
  // the period; such pairs are stored as -2 - neighbour (mesh.cpp:376-389).
  void match_faces() {
    const int64_t ne = m_.ne;
    const uint64_t nn = (uint64_t)m_.nn;
    const int64_t plane = (int64_t)(m_.ny + 1) * (m_.nz + 1);  // nodes per x plane; i = id / plane
    std::vector<std::pair<uint64_t, int64_t>> keys(ne * 4);
    for (int64_t e = 0; e < ne; ++e)
      for (int f = 0; f < 4; ++f) {
        uint64_t v[3];
        bool on_hi = m_.periodic;
        for (int c = 0; c < 3; ++c) {
          v[c] = (uint64_t)m_.elem_nodes[e * 4 + kFaces[f][c]];
          on_hi = on_hi && (int64_t)v[c] / plane == m_.nx;
        }
        if (on_hi)
          for (int c = 0; c < 3; ++c) v[c] -= (uint64_t)(m_.nx * plane);
        std::sort(v, v + 3);
        keys[e * 4 + f] = {(v[0] * nn + v[1]) * nn + v[2], e * 4 + f};
      }
    std::sort(keys.begin(), keys.end());
    m_.nbrs.assign(ne * 4, -1);
    for (size_t a = 0; a + 1 < keys.size(); ++a)
      if (keys[a].first == keys[a + 1].first) {
        const int32_t e0 = (int32_t)(keys[a].second / 4), e1 = (int32_t)(keys[a + 1].second / 4);
        // joined across the period: the largest node id of the key is on x = lo
        const bool across = m_.periodic && (int64_t)(keys[a].first % nn) < plane;
        m_.nbrs[keys[a].second] = across ? -2 - e1 : e1;
        m_.nbrs[keys[a + 1].second] = across ? -2 - e0 : e0;
        ++a;
      }
  }

  bool element_geometry() {
    const int64_t ne = m_.ne;
    m_.normals.assign(ne * 12, 0.0);
    m_.jinv.assign(ne * 9, 0.0);
    m_.detj.assign(ne, 0.0);
    m_.grad.assign(ne * 12, 0.0);
    m_.wbv.assign(ne * 16, 0.0);
    bool ok = true;
    for (int64_t e = 0; e < ne; ++e) {
      const double* X[4];
      for (int v = 0; v < 4; ++v) X[v] = &m_.nodes[(size_t)m_.elem_nodes[e * 4 + v] * 3];
      for (int f = 0; f < 4; ++f) {
        double a[3], b[3];
        for (int d = 0; d < 3; ++d) {
          a[d] = X[kNa[f]][d] - X[kNb[f]][d];
          b[d] = X[kNc[f]][d] - X[kNb[f]][d];
        }
        double n[3] = {b[1] * a[2] - b[2] * a[1], b[2] * a[0] - b[0] * a[2], b[0] * a[1] - b[1] * a[0]};
        const double scale = 1. / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);  // utils.hpp:75-80
        for (int d = 0; d < 3; ++d) m_.normals[e * 12 + f * 3 + d] = n[d] * scale;
      }
      double J[3][3];  // J[r][c] = x_{c+1,r} - x_{0,r}
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = X[c + 1][r] - X[0][r];
      const double det = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
                         J[0][2] * J[1][1] * J[2][0] - J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2];
      ok = ok && det > 0;  // mesh.cpp:692
      m_.detj[e] = det;
      double* inv = &m_.jinv[e * 9];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const int a0 = (c + 1) % 3, a1 = (c + 2) % 3, b0 = (r + 1) % 3, b1 = (r + 2) % 3;
          inv[r * 3 + c] = (J[a0][b0] * J[a1][b1] - J[a0][b1] * J[a1][b0]) / det;
        }
      for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) {
          double g = 0;
          for (int j = 0; j < 3; ++j) g += inv[j * 3 + d] * kGradRef[i][j];
          m_.grad[e * 12 + i * 3 + d] = g;
        }
      for (int c = 0; c < 4; ++c) {
        const double meas = det * kCubW;
        const double* p = kCub[c];
        const double N[4] = {1 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
        for (int i = 0; i < 4; ++i) m_.wbv[e * 16 + i * 4 + c] = meas * N[i];
      }
    }
    if (!ok) mpic::global_error() = "element with det J <= 0";
    return ok;
  }

  void element_matrix(int64_t e, double K[4][4]) const {
    const double* G = &m_.grad[e * 12];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double s = 0;
        for (int c = 0; c < 4; ++c) {
          const double meas = m_.detj[e] * kCubW;
          for (int d = 0; d < 3; ++d) s += G[i * 3 + d] * (G[j * 3 + d] * meas);
        }
        K[i][j] = s;
      }
  }

  void assemble() {
    const int64_t nn = m_.nn, ne = m_.ne;
    // Row patterns: interior rows take every node of every touching element;
    // Dirichlet rows only their diagonal (ES.cpp:91-101).  Periodic: rows and
    // columns of image nodes go to their partners; an image row holds itself and
    // its partner (ES.cpp:104-110).
    std::vector<std::vector<int32_t>> pattern(nn);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 4; ++a) {
        const int32_t row = m_.row_of(m_.elem_nodes[e * 4 + a]);
        if (m_.boundary[row]) continue;
        for (int b = 0; b < 4; ++b) pattern[row].push_back(m_.row_of(m_.elem_nodes[e * 4 + b]));
      }
    m_.row_ptr.assign(nn + 1, 0);
    for (int64_t r = 0; r < nn; ++r) {
      auto& p = pattern[r];
      if (m_.boundary[r]) p.assign(1, (int32_t)r);
      if (m_.row_of((int32_t)r) != r) p = {(int32_t)r, m_.row_of((int32_t)r)};
      std::sort(p.begin(), p.end());
      p.erase(std::unique(p.begin(), p.end()), p.end());
      m_.row_ptr[r + 1] = m_.row_ptr[r] + (int32_t)p.size();
    }
    m_.col.resize(m_.row_ptr[nn]);
    for (int64_t r = 0; r < nn; ++r) std::copy(pattern[r].begin(), pattern[r].end(), m_.col.begin() + m_.row_ptr[r]);
    pattern.clear();
    m_.val.assign(m_.col.size(), 0.0);
    // Element contributions summed in element order (sumIntoGlobalValues, ES.cpp:166-191).
    for (int64_t e = 0; e < ne; ++e) {
      double K[4][4];
      element_matrix(e, K);
      for (int a = 0; a < 4; ++a) {
        const int32_t row = m_.row_of(m_.elem_nodes[e * 4 + a]);
        if (m_.boundary[row]) continue;
        const int32_t* first = m_.col.data() + m_.row_ptr[row];
        const int32_t* last = m_.col.data() + m_.row_ptr[row + 1];
        for (int b = 0; b < 4; ++b) {
          const int64_t at = std::lower_bound(first, last, m_.row_of(m_.elem_nodes[e * 4 + b])) - m_.col.data();
          m_.val[at] += K[a][b];
        }
      }
    }
    for (int64_t r = 0; r < nn; ++r) {
      if (m_.boundary[r]) m_.val[m_.row_ptr[r]] = 1.0;  // ES.cpp:198-204
      const int32_t pr = m_.row_of((int32_t)r);
      if (pr != r)  // image row: -phi_r + phi_partner (ES.cpp:181-186)
        for (int32_t at = m_.row_ptr[r]; at < m_.row_ptr[r + 1]; ++at) m_.val[at] = m_.col[at] == r ? -1.0 : 1.0;
    }
  }
};

}  // namespace

namespace {
int make_brick(int nx, int ny, int nz, int split, const double* lo, const double* hi, double jitter,
               uint64_t jitter_seed, int mapping, bool periodic, mpic_mesh** out) {
  if (!out || !lo || !hi || nx < 1 || ny < 1 || nz < 1 || (split != 5 && split != 6) || mapping < 0 || mapping > 3) {
    mpic::global_error() = "mpic_mesh_brick: bad arguments";
    return MPIC_ERR_ARG;
  }
  if (periodic && nx % 2 != 0) {  // mesh.cpp:91-94
    mpic::global_error() = "mpic_mesh_brick_periodic: a periodic mesh needs an even number of hexes in x";
    return MPIC_ERR_ARG;
  }
  const double nodes = (double)(nx + 1) * (ny + 1) * (nz + 1);
  if (nodes > 2.0e6 || (double)nx * ny * nz * split > 2.0e9) {
    mpic::global_error() = "mpic_mesh_brick: mesh too large for 64-bit face keys";
    return MPIC_ERR_ARG;
  }
  mpic_mesh* m = new mpic_mesh();
  m->periodic = periodic;
  m->nx = nx;
  m->ny = ny;
  m->nz = nz;
  m->split = split;
  for (int a = 0; a < 3; ++a) {
    m->lo[a] = lo[a];
    m->hi[a] = hi[a];
  }
  BrickBuilder b(*m);
  if (!b.run(jitter, jitter_seed, mapping)) {
    delete m;
    return MPIC_ERR_ARG;
  }
  *out = m;
  return MPIC_OK;
}
}  // namespace

extern "C" int mpic_mesh_brick(int nx, int ny, int nz, int split, const double* lo, const double* hi, double jitter,
                               uint64_t jitter_seed, int mapping, mpic_mesh** out) {
  return make_brick(nx, ny, nz, split, lo, hi, jitter, jitter_seed, mapping, false, out);
}

extern "C" int mpic_mesh_brick_periodic(int nx, int ny, int nz, int split, const double* lo, const double* hi,
                                        int mapping, mpic_mesh** out) {
  return make_brick(nx, ny, nz, split, lo, hi, 0.0, 0, mapping, true, out);
}

extern "C" int mpic_mesh_set_background(mpic_mesh* m, double f) {
  if (!m) return MPIC_ERR_ARG;
  m->background.assign(m->nn, 0.0);
  for (int64_t e = 0; e < m->ne; ++e)
    for (int i = 0; i < 4; ++i) {
      double v = 0;
      for (int c = 0; c < 4; ++c) v += -f * m->wbv[e * 16 + i * 4 + c];
      const int32_t node = m->elem_nodes[e * 4 + i];
      if (m->boundary[node]) v = 0;
      m->background[m->row_of(node)] += v;  // ES.cpp:247-251
    }
  m->has_background = true;
  return MPIC_OK;
}

// ES::set_analytic_RHS (ES.cpp:235-257): f at each cubature point (refToPhysical,
// mesh.hpp:183-190), val_i = sum_c -f(x_c) wbv[e][i][c], zero at Dirichlet rows,
// summed into the node in element order.
extern "C" int mpic_mesh_load_vector(const mpic_mesh* m, double (*f)(const double* x, void* user), void* user,
                                     double* out) {
  if (!m || !f || !out) {
    mpic::global_error() = "mpic_mesh_load_vector: bad arguments";
    return MPIC_ERR_ARG;
  }
  std::fill(out, out + m->nn, 0.0);
  for (int64_t e = 0; e < m->ne; ++e) {
    double fc[4];
    for (int c = 0; c < 4; ++c) {
      const double* p = kCub[c];
      const double wts[4] = {1 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
      double x[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) x[d] += wts[i] * m->nodes[(size_t)m->elem_nodes[e * 4 + i] * 3 + d];
      fc[c] = f(x, user);
    }
    for (int i = 0; i < 4; ++i) {
      double v = 0;
      for (int c = 0; c < 4; ++c) v += -fc[c] * m->wbv[e * 16 + i * 4 + c];
      const int32_t node = m->elem_nodes[e * 4 + i];
      if (m->boundary[node]) v = 0;
      out[m->row_of(node)] += v;  // ES.cpp:247-251
    }
  }
  return MPIC_OK;
}

extern "C" int mpic_mesh_describe(const mpic_mesh* m, mpic_mesh_desc* d) {
  if (!m || !d) return MPIC_ERR_ARG;
  d->num_nodes = m->nn;
  d->num_elems = m->ne;
  d->nnz = (int64_t)m->col.size();
  d->nodes = m->nodes.data();
  d->elem_nodes = m->elem_nodes.data();
  d->elem_nbrs = m->nbrs.data();
  d->normals = m->normals.data();
  d->inv_jacobian = m->jinv.data();
  d->det_jacobian = m->detj.data();
  d->grad_basis = m->grad.data();
  d->wbv = m->wbv.data();
  d->boundary = m->boundary.data();
  d->row_ptr = m->row_ptr.data();
  d->col_idx = m->col.data();
  d->values = m->val.data();
  d->background = m->has_background ? m->background.data() : nullptr;
  d->node_pair = m->periodic ? m->node_pair.data() : nullptr;
  d->lattice[0] = m->periodic ? 0 : m->nx;
  d->lattice[1] = m->periodic ? 0 : m->ny;
  d->lattice[2] = m->periodic ? 0 : m->nz;
  return MPIC_OK;
}

extern "C" void mpic_mesh_free(mpic_mesh* m) { delete m; }
