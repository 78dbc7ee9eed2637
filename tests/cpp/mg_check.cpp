// The multigrid hierarchy of the precision form's solve (csrc/mg_setup.cpp), on
// the CPU: host code only, no GPU is touched.  Usage: mg_check
//
// For brick meshes built by the library (Kuhn and 5-tet splits, a node
// mapping, interior jitter):
//   1. every coarse operator equals P^T A P formed densely from the
//      interpolation applied to unit vectors (meshes of 7 to 12 cells, even
//      and odd counts);
//   2. conjugate gradients in FP64 preconditioned by an FP32 V-cycle built
//      from the hierarchy -- the algorithm of the device solve: damped Jacobi
//      (omega = 0.8) once before and once after the coarse correction, full
//      weighting, trilinear interpolation, the dense coarsest inverse -- reach
//      |r| < 1e-10 |b| from zero in few iterations, independent of the mesh
//      size, where plain CG needs hundreds;
//   3. a direction of fewer than 3 cells and a periodic mesh are refused with
//      a reason;
//   4. level 0 as a stencil (slots of the lattice offsets) gives the CSR's
//      row sums bit for bit (same entries, ascending columns).
// Prints one line per case and "mg_check ok", exit 0; or the failure, exit 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "minipic_c.h"
#include "../../minipic-crossroads_amd/csrc/mg.hpp"

namespace {

int failures = 0;
void check(bool ok, const std::string& what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what.c_str());
    ++failures;
  }
}

struct Csr {
  int n;
  std::vector<int32_t> rp, col;
  std::vector<double> val;
};

// Level geometry: cells (cx, cy, cz); level 0 holds every node.
struct Lev {
  int cx, cy, cz;
  bool full;
  long index(int i, int j, int k) const {
    if (full) return ((long)i * (cy + 1) + j) * (cz + 1) + k;
    return ((long)(i - 1) * (cy - 1) + (j - 1)) * (cz - 1) + (k - 1);
  }
  bool interior(int i, int j, int k) const { return i > 0 && i < cx && j > 0 && j < cy && k > 0 && k < cz; }
  long count() const { return full ? (long)(cx + 1) * (cy + 1) * (cz + 1) : (long)(cx - 1) * (cy - 1) * (cz - 1); }
};
inline float w1(int d) { return d == 0 ? 1.0f : 0.5f; }

// coarse b = P^T r (fine level f, coarse level c; interior fine points only:
// 2I + 1 is the fine wall when the fine cell count is odd)
void restrict_(const Lev& f, const Lev& c, const std::vector<float>& r, std::vector<float>& b) {
  b.assign(c.count(), 0.0f);
  for (int I = 1; I < c.cx; ++I)
    for (int J = 1; J < c.cy; ++J)
      for (int K = 1; K < c.cz; ++K) {
        float s = 0.0f;
        for (int a = -1; a <= 1; ++a)
          for (int bb = -1; bb <= 1; ++bb)
            for (int cc = -1; cc <= 1; ++cc)
              if (f.interior(2 * I + a, 2 * J + bb, 2 * K + cc))
                s += w1(a) * w1(bb) * w1(cc) * r[f.index(2 * I + a, 2 * J + bb, 2 * K + cc)];
        b[c.index(I, J, K)] = s;
      }
}
// fine y += P x (interior fine points only)
void prolong_add(const Lev& f, const Lev& c, const std::vector<float>& x, std::vector<float>& y) {
  for (int i = 1; i < f.cx; ++i)
    for (int j = 1; j < f.cy; ++j)
      for (int k = 1; k < f.cz; ++k) {
        float s = 0.0f;
        for (int I = i >> 1; I <= (i + 1) >> 1; ++I)
          for (int J = j >> 1; J <= (j + 1) >> 1; ++J)
            for (int K = k >> 1; K <= (k + 1) >> 1; ++K)
              if (c.interior(I, J, K)) s += w1(i - 2 * I) * w1(j - 2 * J) * w1(k - 2 * K) * x[c.index(I, J, K)];
        y[f.index(i, j, k)] += s;
      }
}

struct VCycle {
  const mpic::MGHost& h;
  const Csr& A0;
  std::vector<float> a0;  // level 0 values in FP32
  std::vector<Lev> lev;
  explicit VCycle(const mpic::MGHost& host, const Csr& A) : h(host), A0(A) {
    a0.assign(A.val.begin(), A.val.end());
    lev.push_back(Lev{h.nx, h.ny, h.nz, true});
    for (const auto& L : h.levels) lev.push_back(Lev{L.cx, L.cy, L.cz, false});
  }
  // y = A_l x
  void apply(size_t l, const std::vector<float>& x, std::vector<float>& y) const {
    if (l == 0) {
      y.assign(A0.n, 0.0f);
      for (int r = 0; r < A0.n; ++r) {
        float s = 0.0f;
        for (int e = A0.rp[r]; e < A0.rp[r + 1]; ++e) s += a0[e] * x[A0.col[e]];
        y[r] = s;
      }
      return;
    }
    const auto& L = h.levels[l - 1];
    y.assign(L.n, 0.0f);
    for (int r = 0; r < L.n; ++r) {
      float s = 0.0f;
      for (int e = 0; e < 27; ++e) s += L.val[(size_t)e * L.n + r] * x[L.col[(size_t)e * L.n + r]];
      y[r] = s;
    }
  }
  const std::vector<float>& dinv(size_t l) const { return l == 0 ? h.dinv0 : h.levels[l - 1].dinv; }
  std::vector<float> run(size_t l, const std::vector<float>& b) const {
    const size_t n = b.size();
    if (l == lev.size() - 1) {
      std::vector<float> x(n, 0.0f);
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) x[i] += h.coarse_inv[j * n + i] * b[j];  // column-major
      return x;
    }
    const auto& D = dinv(l);
    std::vector<float> x(n), t, r(n);
    for (size_t i = 0; i < n; ++i) x[i] = mpic::MG_OMEGA * D[i] * b[i];
    apply(l, x, t);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
    if (l == 0)  // walls carry no residual
      for (int i = 0; i <= lev[0].cx; ++i)
        for (int j = 0; j <= lev[0].cy; ++j)
          for (int k = 0; k <= lev[0].cz; ++k)
            if (!lev[0].interior(i, j, k)) r[lev[0].index(i, j, k)] = 0.0f;
    std::vector<float> bc;
    restrict_(lev[l], lev[l + 1], r, bc);
    const std::vector<float> xc = run(l + 1, bc);
    prolong_add(lev[l], lev[l + 1], xc, x);
    apply(l, x, t);
    std::vector<float> z(n);
    for (size_t i = 0; i < n; ++i) z[i] = x[i] + mpic::MG_OMEGA * D[i] * (b[i] - t[i]);
    return z;
  }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
void spmv(const Csr& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(A.n, 0.0);
  for (int r = 0; r < A.n; ++r)
    for (int e = A.rp[r]; e < A.rp[r + 1]; ++e) y[r] += A.val[e] * x[A.col[e]];
}

// PCG from zero; returns iterations to |r| < tol |b| (or -1 past max_it).
int pcg(const Csr& A, const std::vector<double>& b, const std::function<std::vector<double>(const std::vector<double>&)>& M,
        double tol, int max_it) {
  std::vector<double> x(A.n, 0.0), r = b, q;
  const double bb = dot(b, b);
  std::vector<double> z = M(r), p = z;
  double rz = dot(r, z);
  for (int it = 0; it <= max_it; ++it) {
    if (dot(r, r) < tol * tol * bb) return it;
    spmv(A, p, q);
    const double alpha = rz / dot(p, q);
    for (int i = 0; i < A.n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    z = M(r);
    const double rz2 = dot(r, z);
    for (int i = 0; i < A.n; ++i) p[i] = z[i] + (rz2 / rz) * p[i];
    rz = rz2;
  }
  return -1;
}

struct Case {
  int nx, ny, nz, split, mapping;
  double jitter;
};

bool build(const Case& cs, mpic_mesh** m, mpic_mesh_desc& d) {
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  if (mpic_mesh_brick(cs.nx, cs.ny, cs.nz, cs.split, lo, hi, cs.jitter, 7, cs.mapping, m) != MPIC_OK) return false;
  std::memset(&d, 0, sizeof(d));
  return mpic_mesh_describe(*m, &d) == MPIC_OK;
}

Csr csr_of(const mpic_mesh_desc& d) {
  Csr A;
  A.n = d.num_nodes;
  A.rp.assign(d.row_ptr, d.row_ptr + d.num_nodes + 1);
  A.col.assign(d.col_idx, d.col_idx + d.nnz);
  A.val.assign(d.values, d.values + d.nnz);
  return A;
}

// 1. coarse operator of level 1 against a dense P^T A P
void galerkin_case(const Case& cs) {
  mpic_mesh* m = nullptr;
  mpic_mesh_desc d;
  if (!build(cs, &m, d)) {
    check(false, "mesh build");
    return;
  }
  mpic::MGHost h;
  std::string why;
  const bool ok = mpic::mg_build(d.lattice[0], d.lattice[1], d.lattice[2], d.num_nodes, d.boundary, d.row_ptr,
                                 d.col_idx, d.values, h, why);
  check(ok, "mg_build " + why);
  if (ok) {
    const Lev f{cs.nx, cs.ny, cs.nz, true}, c{h.levels[0].cx, h.levels[0].cy, h.levels[0].cz, false};
    const long nf = f.count(), nc = c.count();
    // P columns: prolongation of unit coarse vectors
    std::vector<std::vector<double>> P(nc);
    for (long j = 0; j < nc; ++j) {
      std::vector<float> e(nc, 0.0f), y(nf, 0.0f);
      e[j] = 1.0f;
      prolong_add(f, c, e, y);
      P[j].assign(y.begin(), y.end());
    }
    // A restricted to interior rows and columns
    Csr A = csr_of(d);
    std::vector<char> in(nf, 0);
    for (int i = 1; i < f.cx; ++i)
      for (int j = 1; j < f.cy; ++j)
        for (int k = 1; k < f.cz; ++k) in[f.index(i, j, k)] = 1;
    double maxdiff = 0, maxval = 0;
    for (long J = 0; J < nc; ++J) {
      std::vector<double> ap(nf, 0.0);
      for (long r = 0; r < nf; ++r) {
        if (!in[r]) continue;
        for (int e = A.rp[r]; e < A.rp[r + 1]; ++e)
          if (in[A.col[e]]) ap[r] += A.val[e] * P[J][A.col[e]];
      }
      for (long I = 0; I < nc; ++I) {
        const double want = dot(P[I], ap);
        double got = 0;
        const auto& L = h.levels[0];
        for (int e = 0; e < 27; ++e)  // padding entries (col = row, val = 0) add nothing
          if (L.col[(size_t)e * L.n + I] == J) got += L.val[(size_t)e * L.n + I];
        maxdiff = std::fmax(maxdiff, std::fabs(want - got));
        maxval = std::fmax(maxval, std::fabs(want));
      }
    }
    std::printf("galerkin %dx%dx%d split %d mapping %d: %ld coarse unknowns, max |P^T A P - A_1| %.3e of %.3e\n",
                cs.nx, cs.ny, cs.nz, cs.split, cs.mapping, nc, maxdiff, maxval);
    check(maxdiff <= 1e-6 * maxval, "galerkin operator");  // FP32 storage of the coarse operator
  }
  mpic_mesh_free(m);
}

// 2. V-cycle preconditioned CG
void pcg_case(const Case& cs, int max_pcg_its) {
  mpic_mesh* m = nullptr;
  mpic_mesh_desc d;
  if (!build(cs, &m, d)) {
    check(false, "mesh build");
    return;
  }
  mpic::MGHost h;
  std::string why;
  const bool ok = mpic::mg_build(d.lattice[0], d.lattice[1], d.lattice[2], d.num_nodes, d.boundary, d.row_ptr,
                                 d.col_idx, d.values, h, why);
  check(ok, "mg_build " + why);
  if (ok) {
    const Csr A = csr_of(d);
    std::mt19937_64 gen(11);
    std::normal_distribution<double> nd;
    std::vector<double> b(A.n);
    for (int i = 0; i < A.n; ++i) b[i] = d.boundary[i] ? 0.0 : nd(gen);
    const VCycle V(h, A);
    auto M = [&](const std::vector<double>& r) {
      std::vector<float> rf(r.begin(), r.end());
      const std::vector<float> z = V.run(0, rf);
      return std::vector<double>(z.begin(), z.end());
    };
    // 4. the level-0 stencil against the CSR rows, every node
    {
      std::vector<double> xr(A.n), y;
      for (int i = 0; i < A.n; ++i) xr[i] = nd(gen);
      spmv(A, xr, y);
      const int S = (int)h.slot0.size(), ny1 = cs.ny + 1, nz1 = cs.nz + 1;
      long bad = 0;
      for (int r = 0; r < A.n; ++r) {
        double acc = 0;
        for (int s = 0; s < S; ++s) {
          const int sl = h.slot0[s];
          const int c = r + ((sl / 9 - 1) * ny1 + (sl / 3) % 3 - 1) * nz1 + sl % 3 - 1;
          const double v = h.val0[(size_t)s * A.n + r];
          if (v != 0.0) acc += v * xr[c];
        }
        bad += acc != y[r];
      }
      std::printf("stencil %dx%dx%d: %d slots, %ld rows differ from the CSR\n", cs.nx, cs.ny, cs.nz, S, bad);
      check(bad == 0 && S <= 27, "level-0 stencil");
    }
    const int it_mg = pcg(A, b, M, 1e-10, 200);
    const int it_cg = pcg(A, b, [](const std::vector<double>& r) { return r; }, 1e-10, 5000);
    std::printf("pcg %dx%dx%d split %d mapping %d jitter %.2f: %zu levels, multigrid %d iterations, plain CG %d\n",
                cs.nx, cs.ny, cs.nz, cs.split, cs.mapping, cs.jitter, h.levels.size() + 1, it_mg, it_cg);
    check(it_mg > 0 && it_mg <= max_pcg_its, "multigrid iterations");
    check(it_cg > 2 * it_mg, "plain CG needs more iterations");
  }
  mpic_mesh_free(m);
}

void refuse_case(const Case& cs, bool periodic) {
  mpic_mesh* m = nullptr;
  mpic_mesh_desc d;
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  const int rc = periodic ? mpic_mesh_brick_periodic(cs.nx, cs.ny, cs.nz, cs.split, lo, hi, cs.mapping, &m)
                          : mpic_mesh_brick(cs.nx, cs.ny, cs.nz, cs.split, lo, hi, 0.0, 0, cs.mapping, &m);
  if (rc != MPIC_OK || mpic_mesh_describe(m, &d) != MPIC_OK) {
    check(false, "mesh build");
    return;
  }
  mpic::MGHost h;
  std::string why;
  const bool ok = periodic ? (d.lattice[0] != 0 &&
                              mpic::mg_build(d.lattice[0], d.lattice[1], d.lattice[2], d.num_nodes, d.boundary,
                                             d.row_ptr, d.col_idx, d.values, h, why))
                           : mpic::mg_build(d.lattice[0], d.lattice[1], d.lattice[2], d.num_nodes, d.boundary,
                                            d.row_ptr, d.col_idx, d.values, h, why);
  std::printf("refused %dx%dx%d%s: %s\n", cs.nx, cs.ny, cs.nz, periodic ? " periodic" : "",
              periodic && d.lattice[0] == 0 ? "no lattice in the descriptor" : why.c_str());
  check(!ok, "refusal");
  mpic_mesh_free(m);
}

}  // namespace

int main() {
  galerkin_case({8, 8, 8, 6, 0, 0.0});
  galerkin_case({12, 8, 8, 5, 2, 0.0});
  galerkin_case({9, 8, 7, 6, 0, 0.0});
  galerkin_case({7, 7, 7, 5, 1, 0.0});
  pcg_case({16, 16, 16, 6, 0, 0.0}, 20);
  pcg_case({32, 32, 32, 6, 0, 0.0}, 20);
  pcg_case({16, 16, 16, 5, 1, 0.0}, 25);
  pcg_case({16, 16, 8, 6, 2, 0.0}, 25);
  pcg_case({16, 16, 16, 6, 0, 0.2}, 25);
  pcg_case({15, 15, 15, 6, 0, 0.0}, 25);
  pcg_case({23, 23, 23, 6, 0, 0.0}, 25);
  pcg_case({17, 12, 9, 5, 0, 0.2}, 25);
  pcg_case({31, 31, 31, 6, 0, 0.0}, 25);
  refuse_case({16, 16, 2, 6, 0, 0.0}, false);
  refuse_case({8, 8, 8, 5, 0, 0.0}, true);
  if (failures) {
    std::printf("mg_check: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("mg_check ok\n");
  return 0;
}
