// MiniPIC electrostatic timestep: FP64 CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the CPU
// baseline for the MI355X build.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; the product library never links,
// loads or calls it.
//
// It restates, operation for operation, the CPU (Kokkos OpenMP, NOSORT) path of
// the reference MiniPIC (npe9/miniPIC-crossroads).  File:line citations below
// point into that reference.  Every floating-point expression keeps the
// reference's association order; it is compiled with -ffp-contract=off so no
// multiply-add is fused.
//
// Conventions fixed by this restatement where the reference leaves them to an
// external library or to thread scheduling (each is pinned by a test in
// tests/test_oracle_*.py, or marked "parity unpinned" there):
//   * Jacobian, inverse (cofactor / det) and determinant (Sarrus) of the affine
//     P1 tet replace Intrepid CellTools (mesh.cpp:684-700).
//   * P1 basis values and the 4-point degree-2 tet cubature rule replace the
//     Intrepid basis / cubature factory (ES.cpp:51-81, basis.hpp:161-163).
//   * Node charge: the reference adds doubles with atomics in an unspecified
//     order (PIC.cpp:135).  Here each per-particle, per-node contribution is
//     converted to a 128-bit fixed-point integer and summed exactly, so the sum
//     is independent of order (see to_fixed/from_fixed).
//   * CG dot products use one fixed reduction tree (dot_tree) in place of the
//     Tpetra serial sum (CG_solve.hpp:58,66,87,100).
//   * Particle order: fill order is (fill call, element, ipart), one of the
//     orders the reference's team append can produce (particle_list.hpp:103-139).
//   * Periodic meshes (the two-stream problem): faces joined across x are
//     stored as -2 - neighbour; pinned by geometry (the joined faces are the
//     same triangle on opposite walls), the analytic periodic Poisson solution
//     and one-period free streaming (tests/test_oracle.py).

#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <numeric>
#include <omp.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

namespace {

// ---------------------------------------------------------------------------
// Random numbers: dRandom, random.hpp:43-87.
//   seed = A*seed + C (mod 2^64), A = 6364136223846793005, C = 1442695040888963407
//   (the "& M" with M = 2^64-1 is the identity), drand = seed * (1/(double)M);
//   (double)M rounds to 2^64, so drand = (double)seed * 2^-64.
// ---------------------------------------------------------------------------
struct LCG {
  uint64_t s;
  uint64_t next() { s = 6364136223846793005ull * s + 1442695040888963407ull; return s; }
  double drand() { const double d = 1.0 / 18446744073709551616.0; return (double)next() * d; }
};

// The sine of the fill (the beam perturbation of set_vel, particle_fill.hpp:106-111,
// and the configurations' weight perturbation).  The reference calls the
// platform's sin, whose last bit differs between the device and the C library;
// the build specifies this one: x = k pi/2 + r by Cody and Waite's reduction
// (pi/2 = fdlibm's 33-bit head 0x1.921fb544p+0 plus its tail), then Taylor
// polynomials of sin r (terms to r^21) and cos r (to r^20), Horner in z = r^2,
// coefficients (-1)^k/(2k+1)! and (-1)^k/(2k)! rounded to double.  Pinned
// against the C library's sin (within 1 ulp) in tests/test_oracle.py.
double fill_sin(double x) {
  const double k = std::rint(x * 0x1.45f306dc9c883p-1);
  const double r = (x - k * 0x1.921fb54400000p+0) - k * 0x1.0b4611a626331p-34;
  const double z = r * r;
  static const double S[10] = {1.9572941063391263e-20, -8.22063524662433e-18, 2.8114572543455206e-15,
                               -7.647163731819816e-13, 1.6059043836821613e-10, -2.505210838544172e-08,
                               2.7557319223985893e-06, -0.0001984126984126984, 0.008333333333333333,
                               -0.16666666666666666};
  static const double C[10] = {4.110317623312165e-19, -1.5619206968586225e-16, 4.779477332387385e-14,
                               -1.1470745597729725e-11, 2.08767569878681e-09, -2.755731922398589e-07,
                               2.48015873015873e-05, -0.001388888888888889, 0.041666666666666664, -0.5};
  double s = S[0], c = C[0];
  for (int i = 1; i < 10; ++i) {
    s = s * z + S[i];
    c = c * z + C[i];
  }
  const double sn = r + (r * z) * s;
  const double cs = 1.0 + z * c;
  switch ((long long)k & 3) {
    case 0: return sn;
    case 1: return cs;
    case 2: return -sn;
    default: return -cs;
  }
}

// 4-point degree-2 cubature on the reference tet (inferred Intrepid default,
// used through ES.cpp:55,70): a = (5 - sqrt 5)/20, b = 1 - 3a, weight 1/24.
const double CUB_A = 0.1381966011250105;
const double CUB_B = 0.5854101966249685;
const double CUB_W = 1.0 / 24.0;
const double CUB_PTS[4][3] = {{CUB_A, CUB_A, CUB_A}, {CUB_B, CUB_A, CUB_A},
                              {CUB_A, CUB_B, CUB_A}, {CUB_A, CUB_A, CUB_B}};
// Reference-element gradients of the P1 basis (Basis_HGRAD_TET_C1_FEM, PIC.cpp:215,227).
const double GREF[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

// Tet faces, mesh.cpp:238-250 and 654: face f = local nodes FACE[f].
const int FACE[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};

// Hex corner local index L = 4*di + 2*dj + dk, as mesh.cpp:200-208.
const int REF5_ODD[5][4] = {{3, 1, 0, 5}, {6, 4, 5, 0}, {6, 2, 0, 3}, {5, 7, 6, 3}, {5, 6, 0, 3}};
const int REF5_EVEN[5][4] = {{2, 3, 1, 7}, {7, 6, 4, 2}, {2, 0, 4, 1}, {4, 5, 7, 1}, {2, 4, 7, 1}};
// Kuhn split along the 0-7 diagonal, one tet per axis permutation, with the two
// middle vertices swapped for odd permutations so that det J > 0.
const int KUHN6[6][4] = {{0, 4, 6, 7}, {0, 5, 4, 7}, {0, 6, 2, 7}, {0, 2, 3, 7}, {0, 1, 5, 7}, {0, 3, 1, 7}};

struct Mesh {
  int nx, ny, nz, split;
  double lo[3], hi[3];
  int32_t nn = 0, ne = 0;
  std::vector<double> nodes;    // [nn][3]
  std::vector<int32_t> en;      // [ne][4] element -> node
  std::vector<int32_t> nb;      // [ne][4] element face -> neighbour element, -1 wall
  std::vector<double> normals;  // [ne][4][3]
  std::vector<double> jinv;     // [ne][3][3]
  std::vector<double> detj;     // [ne]
  std::vector<double> sixodet;  // [ne] 6/detJ
  std::vector<double> grad;     // [ne][4][3]
  std::vector<double> wbv;      // [ne][4][4]
  std::vector<uint8_t> bnd;     // [nn]
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;
  // Periodic in x (Mesh::Mesh with periodic = true): image node -> partner on
  // x = lo (Mesh::node_pairs, mesh.cpp:178-189), else -1; faces joined across
  // the period are stored in nb as -2 - neighbour; px = lo[0] + hi[0].
  bool periodic = false;
  std::vector<int32_t> pair;
  double px = 0.0;
  int32_t row_of(int32_t n) const { return periodic && pair[n] >= 0 ? pair[n] : n; }
};

inline double sq_sum3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Mesh::apply_mapping, mesh.cpp:763-792 (applied to node coordinates after the
// topology and boundary sets are built, mesh.cpp:587-588).
void apply_mapping(double& inx, double& iny, double& inz, int mapping) {
  double outx = inx, outy = iny, outz = inz;
  switch (mapping) {
    case 0:
      return;
    case 1: {
      double r = (1.0 - std::fabs(inx)) * (1.0 - std::fabs(iny));
      outx = inx * std::cos(r) - iny * std::sin(r);
      outy = inx * std::sin(r) + iny * std::cos(r);
      outz = inz;
    } break;
    case 2:
      outx = inx * std::sqrt(1 - .5 * iny * iny - .5 * inz * inz + iny * iny * inz * inz / 3.);
      outy = iny * std::sqrt(1 - .5 * inx * inx - .5 * inz * inz + inx * inx * inz * inz / 3.);
      outz = inz * std::sqrt(1 - .5 * iny * iny - .5 * inx * inx + iny * iny * inx * inx / 3.);
      break;
    case 3:
      outx = inx * inx;
      if (inx < 0) outx = -outx;
      outy = iny;
      outz = inz;
      break;
  }
  inx = outx;
  iny = outy;
  inz = outz;
}

void build_mesh(Mesh& m, double jitter, uint64_t jseed, int mapping) {
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int tph = m.split;  // tets per hex
  const int64_t ne = (int64_t)nx * ny * nz * tph;
  m.nn = (int32_t)nn;
  m.ne = (int32_t)ne;
  double d[3];
  const int n3[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) d[a] = (m.hi[a] - m.lo[a]) / n3[a];
  // Nodes: x = lo + i*d, mesh.cpp:135-142 (lo=-1, d=2/n there).
  m.nodes.assign(nn * 3, 0.0);
  m.bnd.assign(nn, 0);
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j <= ny; ++j)
      for (int k = 0; k <= nz; ++k) {
        int64_t g = (int64_t)i * (ny + 1) * (nz + 1) + (int64_t)j * (nz + 1) + k;  // mesh.cpp:52-54
        double* p = &m.nodes[g * 3];
        p[0] = m.lo[0] + i * d[0];
        p[1] = m.lo[1] + j * d[1];
        p[2] = m.lo[2] + k * d[2];
        // mesh.cpp:408-411; the x walls of a periodic mesh are joined, not Dirichlet (:410)
        bool b = ((!m.periodic && (i == 0 || i == nx)) || j == 0 || j == ny || k == 0 || k == nz);
        m.bnd[g] = b;
        if (!b && jitter != 0.0) {
          // Seeded interior jitter, uniform in [-jitter*d, +jitter*d] per axis.
          LCG r{(uint64_t)(g + 1) * 0x9E3779B97F4A7C15ull + jseed};
          for (int t = 0; t < 4; ++t) r.next();
          for (int a = 0; a < 3; ++a) {
            double u = r.drand();
            p[a] = p[a] + (jitter * d[a]) * (2.0 * u - 1.0);
          }
        }
      }
  for (int64_t g = 0; g < nn; ++g) apply_mapping(m.nodes[g * 3], m.nodes[g * 3 + 1], m.nodes[g * 3 + 2], mapping);
  const int64_t plane = (int64_t)(ny + 1) * (nz + 1);  // node id = i*plane + j*(nz+1) + k
  if (m.periodic) {  // mesh.cpp:178-189
    m.pair.assign(nn, -1);
    for (int j = 1; j < ny; ++j)
      for (int k = 1; k < nz; ++k) m.pair[(int64_t)nx * plane + (int64_t)j * (nz + 1) + k] = (int32_t)((int64_t)j * (nz + 1) + k);
    m.px = m.lo[0] + m.hi[0];
  }
  // Elements, mesh.cpp:196-224 (5-tet) or the Kuhn 6-tet split.
  m.en.assign(ne * 4, 0);
  int64_t cnt = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        int32_t hx[8];
        for (int L = 0; L < 8; ++L) {
          int di = (L >> 2) & 1, dj = (L >> 1) & 1, dk = L & 1;
          hx[L] = (int32_t)((int64_t)(i + di) * (ny + 1) * (nz + 1) + (int64_t)(j + dj) * (nz + 1) + (k + dk));
        }
        const int(*tab)[4] = (tph == 6) ? KUHN6 : (((i + j + k) % 2 != 0) ? REF5_ODD : REF5_EVEN);
        for (int t = 0; t < tph; ++t, ++cnt)
          for (int c = 0; c < 4; ++c) m.en[cnt * 4 + c] = hx[tab[t][c]];
      }
  // Face -> neighbour by matching sorted node triples, mesh.cpp:233-250,396-405.
  std::vector<std::pair<uint64_t, int64_t>> faces(ne * 4);
#pragma omp parallel for
  for (int64_t e = 0; e < ne; ++e)
    for (int f = 0; f < 4; ++f) {
      uint64_t a = m.en[e * 4 + FACE[f][0]], b = m.en[e * 4 + FACE[f][1]], c = m.en[e * 4 + FACE[f][2]];
      // Mesh::periodic_face (mesh.cpp:62-76): a face on x = hi is keyed by its x = lo nodes
      const uint64_t shift = (uint64_t)nx * (uint64_t)plane;
      if (m.periodic && (int64_t)a / plane == nx && (int64_t)b / plane == nx && (int64_t)c / plane == nx) {
        a -= shift;
        b -= shift;
        c -= shift;
      }
      if (a > b) std::swap(a, b);
      if (b > c) std::swap(b, c);
      if (a > b) std::swap(a, b);
      faces[e * 4 + f] = {(a * (uint64_t)nn + b) * (uint64_t)nn + c, e * 4 + f};
    }
  std::sort(faces.begin(), faces.end());
  m.nb.assign(ne * 4, -1);
  for (size_t q = 0; q < faces.size();) {
    size_t r = q + 1;
    while (r < faces.size() && faces[r].first == faces[q].first) ++r;
    if (r - q == 2) {
      int64_t f0 = faces[q].second, f1 = faces[q + 1].second;
      // a matched face whose nodes all lie on x = lo joins the two sides of the
      // period (mesh.cpp:376-389: elem_face_periodic)
      const bool across = m.periodic && (int64_t)(faces[q].first % (uint64_t)nn) < plane;
      m.nb[f0] = across ? -2 - (int32_t)(f1 / 4) : (int32_t)(f1 / 4);
      m.nb[f1] = across ? -2 - (int32_t)(f0 / 4) : (int32_t)(f0 / 4);
    }
    q = r;
  }
  // Normals, mesh.cpp:651-669.
  m.normals.assign(ne * 12, 0.0);
  m.jinv.assign(ne * 9, 0.0);
  m.detj.assign(ne, 0.0);
  m.sixodet.assign(ne, 0.0);
  m.grad.assign(ne * 12, 0.0);
  m.wbv.assign(ne * 16, 0.0);
  // For face f: v1 = n[A]-n[B], v2 = n[C]-n[B], normal = v2 x v1 with (A,B,C) per mesh.cpp:655-666.
  static const int NA[4] = {0, 1, 0, 0}, NB_[4] = {1, 2, 3, 2}, NC[4] = {3, 3, 2, 1};
  long bad = 0;
#pragma omp parallel for reduction(+ : bad)
  for (int64_t e = 0; e < ne; ++e) {
    const double* X[4];
    for (int c = 0; c < 4; ++c) X[c] = &m.nodes[(int64_t)m.en[e * 4 + c] * 3];
    for (int f = 0; f < 4; ++f) {
      double v1[3], v2[3], nrm[3];
      for (int a = 0; a < 3; ++a) {
        v1[a] = X[NA[f]][a] - X[NB_[f]][a];
        v2[a] = X[NC[f]][a] - X[NB_[f]][a];
      }
      nrm[0] = v2[1] * v1[2] - v2[2] * v1[1];
      nrm[1] = v2[2] * v1[0] - v2[0] * v1[2];
      nrm[2] = v2[0] * v1[1] - v2[1] * v1[0];
      double factor = 1. / std::sqrt(sq_sum3(nrm, nrm));  // utils.hpp:69-80
      for (int a = 0; a < 3; ++a) m.normals[e * 12 + f * 3 + a] = nrm[a] * factor;
    }
    // Affine Jacobian J[r][c] = x_{c+1,r} - x_{0,r}.
    double J[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) J[r][c] = X[c + 1][r] - X[0][r];
    double det = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
                 J[0][2] * J[1][1] * J[2][0] - J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2];
    if (!(det > 0)) ++bad;  // mesh.cpp:692
    m.detj[e] = det;
    m.sixodet[e] = 6 / det;  // the (6/ detJ) factor of PIC.cpp:125
    double* Ji = &m.jinv[e * 9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        // inverse[r][c] = cofactor(c, r) / det with cyclic indices
        int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
        double cof = J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1];
        Ji[r * 3 + c] = cof / det;
      }
    // Physical basis gradients, PIC.cpp:229-241 (HGRADtransformGRAD: sum_j Jinv[j][d]*gref[i][j]).
    for (int i = 0; i < 4; ++i)
      for (int dd = 0; dd < 3; ++dd) {
        double acc = 0;
        for (int j = 0; j < 3; ++j) acc += Ji[j * 3 + dd] * GREF[i][j];
        m.grad[e * 12 + i * 3 + dd] = acc;
      }
    // Weighted basis values, ES.cpp:193-195: (detJ * w_c) * N_i(xi_c).
    for (int c = 0; c < 4; ++c) {
      double meas = det * CUB_W;
      const double* p = CUB_PTS[c];
      double N[4] = {1 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
      for (int i = 0; i < 4; ++i) m.wbv[e * 16 + i * 4 + c] = meas * N[i];
    }
  }
  if (bad) m.ne = -1;  // signals a degenerate mesh to the caller
}

// Element stiffness, ES.cpp:141-162: K(i,j) = sum_c sum_d G[i][d] * (G[j][d] * meas_c).
void elem_stiffness(const Mesh& m, int64_t e, double K[4][4]) {
  const double* G = &m.grad[e * 12];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double acc = 0;
      for (int c = 0; c < 4; ++c) {
        double meas = m.detj[e] * CUB_W;
        for (int dd = 0; dd < 3; ++dd) acc += G[i * 3 + dd] * (G[j * 3 + dd] * meas);
      }
      K[i][j] = acc;
    }
}

// Stiffness CSR, ES.cpp:83-112,163-205: interior rows hold every node of every
// element touching the row (boundary columns kept); boundary rows are identity.
// Columns ascend within a row; each entry sums element contributions in element order.
void build_csr(Mesh& m) {
  // Periodic: rows and columns of image nodes go to their partners; an image
  // row is (-1 at itself, +1 at its partner), ES.cpp:91-110,168-189.
  const int64_t nn = m.nn, ne = m.ne;
  std::vector<int64_t> ecount(nn + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int c = 0; c < 4; ++c) ecount[m.row_of(m.en[e * 4 + c]) + 1]++;
  for (int64_t i = 0; i < nn; ++i) ecount[i + 1] += ecount[i];
  std::vector<int32_t> elist(ecount[nn]);
  {
    std::vector<int64_t> pos(ecount.begin(), ecount.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int c = 0; c < 4; ++c) elist[pos[m.row_of(m.en[e * 4 + c])]++] = (int32_t)e;
  }
  std::vector<std::vector<int32_t>> cols(nn);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < nn; ++i) {
    std::vector<int32_t>& cl = cols[i];
    if (m.bnd[i]) {
      cl.push_back((int32_t)i);
      continue;
    }
    if (m.row_of((int32_t)i) != i) {
      cl.push_back(m.row_of((int32_t)i));
      cl.push_back((int32_t)i);
      continue;
    }
    for (int64_t q = ecount[i]; q < ecount[i + 1]; ++q)
      for (int c = 0; c < 4; ++c) cl.push_back(m.row_of(m.en[(int64_t)elist[q] * 4 + c]));
    std::sort(cl.begin(), cl.end());
    cl.erase(std::unique(cl.begin(), cl.end()), cl.end());
  }
  m.rowptr.assign(nn + 1, 0);
  for (int64_t i = 0; i < nn; ++i) m.rowptr[i + 1] = m.rowptr[i] + (int32_t)cols[i].size();
  m.col.resize(m.rowptr[nn]);
  m.val.assign(m.rowptr[nn], 0.0);
  for (int64_t i = 0; i < nn; ++i) std::copy(cols[i].begin(), cols[i].end(), m.col.begin() + m.rowptr[i]);
  for (int64_t e = 0; e < ne; ++e) {
    double K[4][4];
    elem_stiffness(m, e, K);
    for (int a = 0; a < 4; ++a) {
      int32_t row = m.row_of(m.en[e * 4 + a]);
      if (m.bnd[row]) continue;  // ES.cpp:178-180
      for (int b = 0; b < 4; ++b) {
        int32_t cc = m.row_of(m.en[e * 4 + b]);
        int32_t* beg = &m.col[m.rowptr[row]];
        int32_t* end = &m.col[m.rowptr[row + 1]];
        int64_t p = std::lower_bound(beg, end, cc) - &m.col[0];
        m.val[p] += K[a][b];
      }
    }
  }
  for (int64_t i = 0; i < nn; ++i) {
    if (m.bnd[i]) m.val[m.rowptr[i]] = 1.0;  // ES.cpp:198-204
    if (m.row_of((int32_t)i) != i)          // image row, ES.cpp:181-186
      for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q) m.val[q] = m.col[q] == i ? -1.0 : 1.0;
  }
}

// ---------------------------------------------------------------------------
// Exact 128-bit fixed point.  v -> round-to-nearest-even(v * 2^S) as an integer;
// back: |a| split in two 64-bit halves, (double)hi * 2^64 + (double)lo, * 2^-S.
// ---------------------------------------------------------------------------
inline i128 to_fixed(double v, int S) { return (i128)std::rint(std::ldexp(v, S)); }
inline double from_fixed(i128 a, int S) {
  bool neg = a < 0;
  u128 mg = neg ? (u128)0 - (u128)a : (u128)a;
  uint64_t hi = (uint64_t)(mg >> 64), lo = (uint64_t)mg;
  double dv = (double)hi * 18446744073709551616.0 + (double)lo;
  dv = std::ldexp(dv, -S);
  return neg ? -dv : dv;
}
inline int scale_for_bound(double bound) {
  if (!(bound > 0) || !std::isfinite(bound)) return 100;
  int e = std::ilogb(bound) + 1;
  int S = 124 - e;
  if (S > 1000) S = 1000;
  if (S < -1000) S = -1000;
  return S;
}

// Per-particle, per-node charge, PIC.cpp:121-131:
//   tot_q = weight*q*(6/detJ);  val_i = sum_c (-tot_q)*wbv[e][i][c].
inline void particle_charge(const Mesh& m, int32_t e, double w, double q, double val[4]) {
  double tq = w * q * m.sixodet[e];
  const double* W = &m.wbv[(int64_t)e * 16];
  for (int i = 0; i < 4; ++i) {
    double acc = 0;
    for (int c = 0; c < 4; ++c) acc += -tq * W[i * 4 + c];
    val[i] = acc;
  }
}

// Fixed reduction tree used for every CG dot product (the GPU kernels follow the
// same tree): chunks of 256 items, item t of a chunk held by thread t as
// (0 + term); the 64 lanes of each wave combine by an xor butterfly (offsets
// 32,16,...,1); the 4 wave sums combine as (w0+w1)+(w2+w3).  Chunk partials are
// reduced the same way after thread t sums partials t, t+256, ... in order.
double tree256(const double* tv) {
  double ws[4];
  for (int w = 0; w < 4; ++w) {
    double a[64], b[64];
    for (int l = 0; l < 64; ++l) a[l] = tv[w * 64 + l];
    for (int off = 32; off >= 1; off >>= 1) {
      for (int l = 0; l < 64; ++l) b[l] = a[l] + a[l ^ off];
      std::memcpy(a, b, sizeof(a));
    }
    ws[w] = a[0];
  }
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}
double reduce_partials(const std::vector<double>& part) {
  double tv[256];
  const int64_t P = (int64_t)part.size();
  for (int t = 0; t < 256; ++t) {
    double acc = 0;
    for (int64_t i = t; i < P; i += 256) acc += part[i];
    tv[t] = acc;
  }
  return tree256(tv);
}
template <class F>
double tree_sum(int64_t n, F term) {
  const int64_t nchunk = (n + 255) / 256;
  std::vector<double> part(nchunk);
  for (int64_t b = 0; b < nchunk; ++b) {
    double tv[256];
    for (int t = 0; t < 256; ++t) {
      double acc = 0;
      int64_t i = b * 256 + t;
      if (i < n) acc += term(i);
      tv[t] = acc;
    }
    part[b] = tree256(tv);
  }
  return reduce_partials(part);
}

void spmv(const Mesh& m, const double* x, double* y) {
  for (int64_t i = 0; i < m.nn; ++i) {
    double acc = 0;
    for (int32_t p = m.rowptr[i]; p < m.rowptr[i + 1]; ++p) acc += m.val[p] * x[m.col[p]];
    y[i] = acc;
  }
}

// Gather (PIC.cpp:162-184): weights = getWeights(x_ref) (mesh.hpp:170-176),
// E_d = sum_i sum_j (phi_i*w_j)*G[e][i][j]_d; G[e][i][j] is the same for every j.
inline void gather_E(const Mesh& m, int32_t e, const double xr[3], const double* phi, double E[3]) {
  double wt[4] = {1 - xr[0] - xr[1] - xr[2], xr[0], xr[1], xr[2]};
  double dof[4];
  for (int i = 0; i < 4; ++i) dof[i] = phi[m.en[(int64_t)e * 4 + i]];
  const double* G = &m.grad[(int64_t)e * 12];
  E[0] = E[1] = E[2] = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int dd = 0; dd < 3; ++dd) E[dd] += dof[i] * wt[j] * G[i * 3 + dd];
}

// physToReferenceVector, mesh.hpp:245-250.
inline void p2r_vec(const double* Ji, const double v[3], double out[3]) {
  for (int r = 0; r < 3; ++r) {
    double acc = 0.0;
    for (int c = 0; c < 3; ++c) acc += Ji[r * 3 + c] * v[c];
    out[r] = acc;
  }
}

// Boris push with the reference's B = 0 (particle_move.hpp:364-392); E from gather.
inline void kick(const Mesh& m, int32_t e, double v[3], const double E[3], double qm, double dt, double scale, double vr[3]) {
  const double B[3] = {0.0, 0.0, 0.0};
  double sdt = dt * scale;
  double t[3];
  t[0] = qm * 0.5 * B[0];
  t[1] = qm * 0.5 * B[1];
  t[2] = qm * 0.5 * B[2];
  for (int d = 0; d < 3; ++d) v[d] += sdt * E[d] * qm * 0.5;
  double tmp[3] = {v[0], v[1], v[2]};
  tmp[0] += v[1] * t[2] - v[2] * t[1];  // utils.hpp:55-59
  tmp[1] += v[2] * t[0] - v[0] * t[2];
  tmp[2] += v[0] * t[1] - v[1] * t[0];
  double tsq = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
  for (int d = 0; d < 3; ++d) t[d] *= 2. / (1 + tsq);
  v[0] += tmp[1] * t[2] - tmp[2] * t[1];
  v[1] += tmp[2] * t[0] - tmp[0] * t[2];
  v[2] += tmp[0] * t[1] - tmp[1] * t[0];
  for (int d = 0; d < 3; ++d) v[d] += sdt * E[d] * qm * 0.5;
  p2r_vec(&m.jinv[(int64_t)e * 9], v, vr);
}

const int WALK_CAP = 1000;  // crossings per particle per step; the reference has no cap

struct WalkStats {
  int64_t crossings = 0, killed = 0, capped = 0, nonpos = 0, reflections = 0;
};

// MoveParticle::operator() and find_intersection, particle_move.hpp:103-124,136-225.
// Returns false if the particle was killed on a wall.
inline bool walk(const Mesh& m, int32_t& e, double x[3], double xr[3], double v[3], double vr[3], double dt, bool kill,
                 WalkStats& st) {
  double tr = 1.0;
  int iters = 0;
  while (tr > 0) {
    if (iters++ >= WALK_CAP) {
      st.capped++;
      break;
    }
    const double ldt = dt * tr;
    double s = 1.0;
    int face = -1;
    double dxr[3], xf[3];
    for (int d = 0; d < 3; ++d) {
      dxr[d] = ldt * vr[d];
      xf[d] = xr[d] + dxr[d];
    }
    if (vr[2] < 0 && xf[2] < 0) {
      face = 3;
      s = -xr[2] / (dxr[2]);
    }
    if (vr[1] < 0 && xf[1] < 0) {
      double sfy = -xr[1] / (dxr[1]);
      if (sfy < s) {
        face = 0;
        s = sfy;
      }
    }
    if (vr[0] < 0 && xf[0] < 0) {
      double sfx = -xr[0] / (dxr[0]);
      if (sfx < s) {
        face = 2;
        s = sfx;
      }
    }
    double xfs = xf[0] + xf[1] + xf[2];
    double vrs = vr[0] + vr[1] + vr[2];
    if (vrs > 0 && xfs > 1) {
      double sfs = (1 - xr[0] - xr[1] - xr[2]) / (dxr[0] + dxr[1] + dxr[2]);
      if (sfs < s) {
        face = 1;
        s = sfs;
      }
    }
    for (int d = 0; d < 3; ++d) {
      x[d] = x[d] + s * ldt * v[d];
      xr[d] = xr[d] + s * ldt * vr[d];
    }
    if (s < 1) {
      int32_t ne2 = m.nb[(int64_t)e * 4 + face];
      if (ne2 != -1) {
        if (ne2 < -1) {  // across the period: x[0] = -x[0] on [-1,1], particle_move.hpp:190-191
          ne2 = -2 - ne2;
          x[0] = m.px - x[0];
        }
        e = ne2;
        st.crossings++;
        // physToReference, mesh.hpp:197-211
        const double* X0 = &m.nodes[(int64_t)m.en[(int64_t)e * 4] * 3];
        double dl[3] = {x[0] - X0[0], x[1] - X0[1], x[2] - X0[2]};
        double dr[3];
        xr[0] = xr[1] = xr[2] = 0.0;
        p2r_vec(&m.jinv[(int64_t)e * 9], dl, dr);
        for (int d = 0; d < 3; ++d) xr[d] += dr[d];
        p2r_vec(&m.jinv[(int64_t)e * 9], v, vr);
      } else {
        if (kill) {
          st.killed++;
          return false;
        }
        st.reflections++;
        const double* nrm = &m.normals[(int64_t)e * 12 + face * 3];
        double vn = 0;
        for (int d = 0; d < 3; ++d) vn += v[d] * nrm[d];
        for (int d = 0; d < 3; ++d) v[d] -= 2 * vn * nrm[d];
        p2r_vec(&m.jinv[(int64_t)e * 9], v, vr);
      }
    }
    if (!(s > 0)) st.nonpos++;
    tr -= s * tr;
  }
  return true;
}

}  // namespace

// ===========================================================================
// C API (ctypes).  Particle arrays are component planes: x[d*stride + p].
// ===========================================================================
extern "C" {

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
void orc_fill_sin(int64_t n, const double* x, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = fill_sin(x[i]);
}
int orc_get_threads() { return omp_get_max_threads(); }

void* orc_mesh_create(int nx, int ny, int nz, int split, const double* lo, const double* hi, double jitter,
                      uint64_t jseed, int mapping, int periodic) {
  if (nx < 1 || ny < 1 || nz < 1 || (split != 5 && split != 6) || mapping < 0 || mapping > 3) return nullptr;
  if (periodic && (nx % 2 != 0 || jitter != 0.0)) return nullptr;  // mesh.cpp:91-94
  Mesh* m = new Mesh();
  m->periodic = periodic != 0;
  m->nx = nx;
  m->ny = ny;
  m->nz = nz;
  m->split = split;
  for (int a = 0; a < 3; ++a) {
    m->lo[a] = lo[a];
    m->hi[a] = hi[a];
  }
  build_mesh(*m, jitter, jseed, mapping);
  if (m->ne < 0) {
    delete m;
    return nullptr;
  }
  build_csr(*m);
  return m;
}

void orc_mesh_destroy(void* h) { delete (Mesh*)h; }

void orc_mesh_counts(void* h, int64_t* out) {
  Mesh* m = (Mesh*)h;
  out[0] = m->nn;
  out[1] = m->ne;
  out[2] = (int64_t)m->col.size();
  int64_t nb = 0;
  for (auto b : m->bnd) nb += b;
  out[3] = nb;
}

void orc_mesh_export(void* h, double* nodes, int32_t* en, int32_t* nb, double* normals, double* jinv, double* detj,
                     double* grad, double* wbv, uint8_t* bnd, int32_t* rowptr, int32_t* col, double* val) {
  Mesh* m = (Mesh*)h;
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(nodes, m->nodes);
  cp(en, m->en);
  cp(nb, m->nb);
  cp(normals, m->normals);
  cp(jinv, m->jinv);
  cp(detj, m->detj);
  cp(grad, m->grad);
  cp(wbv, m->wbv);
  cp(bnd, m->bnd);
  cp(rowptr, m->rowptr);
  cp(col, m->col);
  cp(val, m->val);
}

// Partner of each node (-1: none); all -1 for a mesh that is not periodic.
void orc_mesh_pairs(void* h, int32_t* out) {
  Mesh* m = (Mesh*)h;
  for (int32_t i = 0; i < m->nn; ++i) out[i] = m->periodic ? m->pair[i] : -1;
}

void orc_elem_stiffness(void* h, int64_t e, double* K) {
  double k[4][4];
  elem_stiffness(*(Mesh*)h, e, k);
  std::memcpy(K, k, sizeof(k));
}

// Uniform background charge density f as a load vector, the reference's
// ES::set_analytic_RHS with a constant source (ES.cpp:235-257).
void orc_background(void* h, double f, double* bg) {
  Mesh& m = *(Mesh*)h;
  for (int64_t i = 0; i < m.nn; ++i) bg[i] = 0.0;
  for (int64_t e = 0; e < m.ne; ++e)
    for (int i = 0; i < 4; ++i) {
      double v = 0;
      for (int c = 0; c < 4; ++c) v += -f * m.wbv[e * 16 + i * 4 + c];
      int32_t node = m.en[e * 4 + i];
      if (m.bnd[node]) v = 0;
      bg[m.row_of(node)] += v;  // ES.cpp:247-251
    }
}

// ParticleFill / FillFunctor, particle_fill.hpp:59-164 (one element per team,
// seed = elem + (34030+ipart)*Nglob + Nglob*12345*type, 10 burned draws,
// rejection-sampled reference position, velocity by set_vel).
// Extensions: seed_offset adds seed_offset * 0x9E3779B97F4A7C15 to the seed
// (0 = the reference), and pert_amp multiplies each weight by
// 1 + pert_amp*sin(2*pi*(x0 - lo0)/Lx).
int64_t orc_fill(void* h, int nppe, int type, double density, const double* vel, double temp, uint64_t seed_offset,
                 double pert_amp, const double* species_m, int64_t offset, int64_t stride, double* x, double* xr, double* v,
                 int32_t* elem, double* w, int32_t* typ, uint32_t* id) {
  Mesh& m = *(Mesh*)h;
  const int64_t ne = m.ne;
  const int64_t total = ne * nppe;
  if (offset + total > stride) return -1;
  const uint64_t N = (uint64_t)ne;
  const double mass = species_m[type];
  const double Lx = m.hi[0] - m.lo[0];
  const double kx = (2.0 * M_PI) / Lx;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < ne; ++e) {
    const double vol = m.detj[e] / 6.;
    const double wt = density * vol / nppe;
    for (int ip = 0; ip < nppe; ++ip) {
      int64_t p = offset + e * nppe + ip;
      LCG r{(uint64_t)e + (34030ull + (uint64_t)ip) * N + N * 12345ull * (uint64_t)type};
      r.s += seed_offset * 0x9E3779B97F4A7C15ull;
      for (int i = 0; i < 10; ++i) r.next();
      double ref[3];
      do {
        ref[0] = r.drand();
        ref[1] = r.drand();
        ref[2] = r.drand();
      } while (ref[0] + ref[1] + ref[2] > 1.0);
      // refToPhysical, mesh.hpp:183-190
      double wts[4] = {1 - ref[0] - ref[1] - ref[2], ref[0], ref[1], ref[2]};
      double ph[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) ph[d] += wts[i] * m.nodes[(int64_t)m.en[e * 4 + i] * 3 + d];
      double vv[3];
      if (temp == 0) {
        vv[0] = vel[0];
        vv[1] = vel[1];
        vv[2] = vel[2];
      } else if (temp < 0) {
        vv[0] = vel[0] * (1. + temp * fill_sin(1. * M_PI * ph[0]));
        vv[1] = vv[2] = 0.0;
      } else {
        while (true) {
          double rr = r.drand();
          double vmax = std::sqrt(6 * temp / mass);
          double u[3];
          u[0] = vmax * 2 * (-.5 + r.drand());
          u[1] = vmax * 2 * (-.5 + r.drand());
          u[2] = vmax * 2 * (-.5 + r.drand());
          double prob = std::exp(-(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / temp);
          if (rr < prob) {
            vv[0] = vel[0] + u[0];
            vv[1] = vel[1] + u[1];
            vv[2] = vel[2] + u[2];
            break;
          }
        }
      }
      double wp = wt;
      if (pert_amp != 0.0) wp = wt * (1.0 + pert_amp * fill_sin(kx * (ph[0] - m.lo[0])));
      for (int d = 0; d < 3; ++d) {
        x[d * stride + p] = ph[d];
        xr[d * stride + p] = ref[d];
        v[d * stride + p] = vv[d];
      }
      elem[p] = (int32_t)e;
      w[p] = wp;
      typ[p] = type;
      id[p] = (uint32_t)p;
    }
  }
  return offset + total;
}

// Fixed-point scale for the node-charge sum: bound = sum|bg| + n_live*max|w*q|.
int orc_charge_scale(void* h, int64_t n, const int32_t* elem, const double* w, const int32_t* typ,
                     const double* species_q, const double* bg) {
  Mesh& m = *(Mesh*)h;
  double sbg = 0;
  if (bg)
    for (int64_t i = 0; i < m.nn; ++i) sbg += std::fabs(bg[i]);
  double mx = 0;
  int64_t live = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (elem[p] < 0) continue;
    ++live;
    double a = std::fabs(w[p] * species_q[typ[p]]);
    if (a > mx) mx = a;
  }
  return scale_for_bound(sbg + (double)live * mx);
}

// weight_charge (PIC.cpp:139-160): rhs = background + sum of particle charges,
// boundary rows zeroed.  Sums are exact in 128-bit fixed point at scale S.
void orc_deposit(void* h, int64_t n, const int32_t* elem, const double* w, const int32_t* typ, const double* species_q,
                 const double* bg, int S, double* rhs) {
  Mesh& m = *(Mesh*)h;
  const int nt = omp_get_max_threads();
  std::vector<std::vector<i128>> acc(nt, std::vector<i128>(m.nn, 0));
#pragma omp parallel
  {
    std::vector<i128>& a = acc[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
      int32_t e = elem[p];
      if (e < 0) continue;
      double val[4];
      particle_charge(m, e, w[p], species_q[typ[p]], val);
      for (int i = 0; i < 4; ++i) a[m.row_of(m.en[(int64_t)e * 4 + i])] += to_fixed(val[i], S);  // PIC.cpp:132-135
    }
  }
  for (int64_t i = 0; i < m.nn; ++i) {
    i128 s = bg ? to_fixed(bg[i], S) : 0;
    for (int t = 0; t < nt; ++t) s += acc[t][i];
    rhs[i] = m.bnd[i] ? 0.0 : from_fixed(s, S);
  }
}

// Same charges summed in plain double, particle order (for pinning the exact sum).
void orc_deposit_naive(void* h, int64_t n, const int32_t* elem, const double* w, const int32_t* typ,
                       const double* species_q, double* rhs) {
  Mesh& m = *(Mesh*)h;
  for (int64_t i = 0; i < m.nn; ++i) rhs[i] = 0.0;
  for (int64_t p = 0; p < n; ++p) {
    int32_t e = elem[p];
    if (e < 0) continue;
    double val[4];
    particle_charge(m, e, w[p], species_q[typ[p]], val);
    for (int i = 0; i < 4; ++i) rhs[m.row_of(m.en[(int64_t)e * 4 + i])] += val[i];
  }
  for (int64_t i = 0; i < m.nn; ++i)
    if (m.bnd[i]) rhs[i] = 0.0;
}

double orc_dot(int64_t n, const double* a, const double* b) {
  return tree_sum(n, [&](int64_t i) { return a[i] * b[i]; });
}

void orc_spmv(void* h, const double* x, double* y) { spmv(*(Mesh*)h, x, y); }

// CG_solve, CG_solve.hpp:54-114: warm start from phi, stop when rho/rho0 < tol^2,
// at most max_iter iterations.  Returns 0 converged, 1 not converged.
int orc_cg(void* h, const double* rhs, double* x, double tol, int max_iter, int* iters_out, double* relres_out) {
  Mesh& m = *(Mesh*)h;
  const int64_t n = m.nn;
  std::vector<double> p(n), q(n), r(n);
  double rho0 = orc_dot(n, rhs, rhs);
  spmv(m, x, r.data());
  for (int64_t i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
  double rho = orc_dot(n, r.data(), r.data()), rho_old = 0, beta = 0, alpha = 0;
  if (rho == 0) {
    *iters_out = 0;
    *relres_out = 0;
    return 0;
  }
  int iter = 0;
  bool conv = false;
  for (; iter < max_iter; ++iter) {
    if (iter == 0)
      for (int64_t i = 0; i < n; ++i) p[i] = r[i];
    else {
      beta = rho / rho_old;
      for (int64_t i = 0; i < n; ++i) p[i] = beta * p[i] + r[i];
    }
    spmv(m, p.data(), q.data());
    alpha = orc_dot(n, p.data(), q.data());
    alpha = rho / alpha;
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
    for (int64_t i = 0; i < n; ++i) r[i] = r[i] + -alpha * q[i];
    rho_old = rho;
    rho = orc_dot(n, r.data(), r.data());
    if (rho / rho0 < tol * tol) {
      conv = true;
      ++iter;
      break;
    }
  }
  *iters_out = iter;
  *relres_out = std::sqrt(rho / rho0);
  return conv ? 0 : 1;
}

// BiCGStab_solve, CG_solve.hpp:115-206 (the solver PIC::time_step calls,
// PIC.cpp:251): x is scaled to 0, r = b - A x, r~ = r; each Tpetra
// update(alpha, A, beta) is this = beta*this + alpha*A; alpha uses rhs.v (r~ = b
// since x starts at 0).  Stops when r.r < rho0*tol*tol after the alpha step
// (x += alpha p first) or r.r/rho0 < tol*tol after the omega step; exits early
// on rho = 0, omega = 0 or t.t = 0 as the reference does.  iters_out = the
// reference's iter at the exit.  Returns 0 converged, 1 max_iter, 2 early exit.
int orc_bicgstab(void* h, const double* rhs, double* x, double tol, int max_iter, int* iters_out,
                 double* relres_out) {
  Mesh& m = *(Mesh*)h;
  const int64_t n = m.nn;
  std::vector<double> r(n), rt(n), p(n), v(n), t(n);
  const double rho0 = orc_dot(n, rhs, rhs);
  for (int64_t i = 0; i < n; ++i) x[i] = 0. * x[i];  // x->scale(0.)
  spmv(m, x, r.data());
  for (int64_t i = 0; i < n; ++i) r[i] = -1.0 * r[i] + 1.0 * rhs[i];
  rt = r;
  double rho = orc_dot(n, rt.data(), r.data()), rho_old = 0, beta = 0, alpha = 0, omega = 0, resid = 0;
  if (rho == 0) {
    *iters_out = 0;
    *relres_out = 0;
    return 0;
  }
  int iter = 0, status = 1;
  for (; iter < max_iter; ++iter) {
    if (iter == 0) {
      p = r;
    } else {
      rho = orc_dot(n, rt.data(), r.data());
      if (rho == 0) {
        status = 2;
        break;
      }
      beta = rho / rho_old * alpha / omega;
      for (int64_t i = 0; i < n; ++i) p[i] = 1.0 * p[i] + -omega * v[i];
      for (int64_t i = 0; i < n; ++i) p[i] = beta * p[i] + 1.0 * r[i];
    }
    spmv(m, p.data(), v.data());
    alpha = orc_dot(n, rhs, v.data());
    alpha = rho / alpha;
    for (int64_t i = 0; i < n; ++i) r[i] = 1.0 * r[i] + -alpha * v[i];
    if ((resid = orc_dot(n, r.data(), r.data())) < rho0 * tol * tol) {
      for (int64_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + alpha * p[i];
      status = 0;
      break;
    }
    spmv(m, r.data(), t.data());
    const double tdott = orc_dot(n, t.data(), t.data());
    omega = orc_dot(n, t.data(), r.data()) / tdott;
    if (omega == 0.0 || tdott == 0) {
      status = 2;
      break;
    }
    for (int64_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + alpha * p[i];
    for (int64_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + omega * r[i];
    for (int64_t i = 0; i < n; ++i) r[i] = 1.0 * r[i] + -omega * t[i];
    rho_old = rho;
    resid = orc_dot(n, r.data(), r.data());
    if (resid / rho0 < tol * tol) {
      status = 0;
      break;
    }
  }
  *iters_out = iter;
  *relres_out = std::sqrt(resid / rho0);
  return status;
}

// weight_Efield (PIC.cpp:162-189) for every live particle; E is [3][stride].
void orc_gather(void* h, int64_t n, int64_t stride, const double* xr, const int32_t* elem, const double* phi, double* E) {
  Mesh& m = *(Mesh*)h;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    if (elem[p] < 0) continue;
    double r[3] = {xr[p], xr[stride + p], xr[2 * stride + p]}, Ev[3];
    gather_E(m, elem[p], r, phi, Ev);
    for (int d = 0; d < 3; ++d) E[d * stride + p] = Ev[d];
  }
}

// ParticleMove::move (particle_move.hpp:398-404) after weight_Efield: gather,
// kick, walk.  stats = [crossings, killed, capped, nonpositive s, reflections].
void orc_push(void* h, int64_t n, int64_t stride, double* x, double* xr, double* v, int32_t* elem, const double* phi,
              const int32_t* typ, const double* species_qm, double dt, double scale, int kill, int64_t* stats) {
  Mesh& m = *(Mesh*)h;
  int64_t c = 0, k = 0, cp = 0, np = 0, rf = 0;
#pragma omp parallel for schedule(static) reduction(+ : c, k, cp, np, rf)
  for (int64_t p = 0; p < n; ++p) {
    int32_t e = elem[p];
    if (e < 0) continue;
    double X[3], R[3], V[3], VR[3], E[3];
    for (int d = 0; d < 3; ++d) {
      X[d] = x[d * stride + p];
      R[d] = xr[d * stride + p];
      V[d] = v[d * stride + p];
    }
    gather_E(m, e, R, phi, E);
    kick(m, e, V, E, species_qm[typ[p]], dt, scale, VR);
    WalkStats st;
    bool alive = walk(m, e, X, R, V, VR, dt, kill != 0, st);
    for (int d = 0; d < 3; ++d) {
      x[d * stride + p] = X[d];
      xr[d * stride + p] = R[d];
      v[d * stride + p] = V[d];
    }
    elem[p] = alive ? e : -1;
    c += st.crossings;
    k += st.killed;
    cp += st.capped;
    np += st.nonpos;
    rf += st.reflections;
  }
  stats[0] = c;
  stats[1] = k;
  stats[2] = cp;
  stats[3] = np;
  stats[4] = rf;
}

// Kinetic energy (PIC.cpp:191-200): sum of 0.5*(w*m)*|v|^2, summed exactly in
// fixed point at a scale chosen from n_live * max term.
double orc_ke(int64_t n, int64_t stride, const double* v, const int32_t* elem, const double* w, const int32_t* typ,
              const double* species_m) {
  std::vector<double> term(n, 0.0);
  double mx = 0;
  int64_t live = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (elem[p] < 0) continue;
    ++live;
    double tm = w[p] * species_m[typ[p]];
    double vs = 0;
    for (int d = 0; d < 3; ++d) vs += v[d * stride + p] * v[d * stride + p];
    term[p] = 0.5 * tm * vs;
    if (term[p] > mx) mx = term[p];
  }
  int S = scale_for_bound((double)live * mx);
  i128 s = 0;
  for (int64_t p = 0; p < n; ++p)
    if (elem[p] >= 0) s += to_fixed(term[p], S);
  return from_fixed(s, S);
}

// Fixed-point helpers exposed for tests.
void orc_to_fixed(double v, int S, int64_t* hi, uint64_t* lo) {
  i128 a = to_fixed(v, S);
  *hi = (int64_t)(a >> 64);
  *lo = (uint64_t)a;
}
double orc_from_fixed(int64_t hi, uint64_t lo, int S) {
  i128 a = ((i128)hi << 64) | (i128)lo;
  return from_fixed(a, S);
}
int orc_scale_for_bound(double b) { return scale_for_bound(b); }

// PIC::time_step (PIC.cpp:245-267) with CG at cg_tol in place of the hard-coded
// BiCGStab: deposit -> solve (warm start) -> gather+kick+walk -> KE.
// out = [cg_iters, relres, ke, crossings, killed, capped, nonpos, reflections, S]
int orc_step(void* h, int64_t n, int64_t stride, double* x, double* xr, double* v, int32_t* elem, const double* w,
             const int32_t* typ, const double* species_q, const double* species_m, const double* bg, int S, double dt,
             double cg_tol, int max_iter, int kill, double* phi, double* rhs, double* out) {
  Mesh& m = *(Mesh*)h;
  double species_qm[16];
  for (int t = 0; t < 16; ++t) species_qm[t] = species_m[t] != 0 ? species_q[t] / species_m[t] : 0;
  orc_deposit(h, n, elem, w, typ, species_q, bg, S, rhs);
  int it = 0;
  double rr = 0;
  int st = orc_cg(h, rhs, phi, cg_tol, max_iter, &it, &rr);
  int64_t ws[5];
  orc_push(h, n, stride, x, xr, v, elem, phi, typ, species_qm, dt, 1.0, kill, ws);
  double ke = orc_ke(n, stride, v, elem, w, typ, species_m);
  out[0] = it;
  out[1] = rr;
  out[2] = ke;
  for (int i = 0; i < 5; ++i) out[3 + i] = (double)ws[i];
  out[8] = S;
  (void)m;
  return st;
}

}  // extern "C"
