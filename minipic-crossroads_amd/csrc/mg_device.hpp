// Device helpers of the multigrid V-cycle shared by the single context's solve
// (k_mg.hip) and the batching form's (k_batch_mg.hip): coarse-level indexing,
// the full-weighting restriction, the linear prolongation and the 27-point
// stencil row of a coarse level.  The hierarchy is described in mg.hpp.
#pragma once

#include "mpic_internal.hpp"

namespace mpic {

__device__ __forceinline__ float mg_w1(int d) { return d == 0 ? 1.0f : 0.5f; }

// Coarse level points: interior lattice points (1..c-1 per direction),
// numbered ((i - 1) (cy - 1) + (j - 1)) (cz - 1) + (k - 1).
__device__ __forceinline__ int mg_cidx(const MGLevelDev& L, int i, int j, int k) {
  return ((i - 1) * (L.cy - 1) + (j - 1)) * (L.cz - 1) + (k - 1);
}
__device__ __forceinline__ void mg_cdecode(const MGLevelDev& L, int u, int& i, int& j, int& k) {
  const int ez = L.cz - 1, ey = L.cy - 1;
  k = u % ez + 1;
  j = (u / ez) % ey + 1;
  i = u / (ez * ey) + 1;
}
__device__ __forceinline__ bool mg_cinterior(const MGLevelDev& L, int i, int j, int k) {
  return i > 0 && i < L.cx && j > 0 && j < L.cy && k > 0 && k < L.cz;
}

// P^T t at coarse point (I, J, K): full weighting of the fine points
// 2I - 1..2I + 1 per direction, idx(i, j, k) giving the index of fine point
// (i, j, k) in t and (fx, fy, fz) the fine level's cells.  The fine point
// 2I + 1 is the fine wall when it equals the cell count (odd counts): it takes
// weight zero and its load goes to the clamped point 2I, so all 27 loads stay
// in bounds and are issued unconditionally.
template <class Idx>
__device__ __forceinline__ float mg_restrict(const float* __restrict__ t, Idx idx, int fx, int fy, int fz, int I, int J,
                                             int K) {
  const bool okx = 2 * I + 1 < fx, oky = 2 * J + 1 < fy, okz = 2 * K + 1 < fz;
  const float wx[3] = {0.5f, 1.0f, okx ? 0.5f : 0.0f}, wy[3] = {0.5f, 1.0f, oky ? 0.5f : 0.0f},
              wz[3] = {0.5f, 1.0f, okz ? 0.5f : 0.0f};
  const int ox[3] = {-1, 0, okx ? 1 : 0}, oy[3] = {-1, 0, oky ? 1 : 0}, oz[3] = {-1, 0, okz ? 1 : 0};
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) s += (wx[a] * wy[b] * wz[c]) * t[idx(2 * I + ox[a], 2 * J + oy[b], 2 * K + oz[c])];
  return s;
}

// P z of the coarser level C at point (i, j, k) of the finer one (z is in C.b):
// per direction the parents i >> 1 (weight 1 for even i, 1/2 for odd) and
// (i >> 1) + 1 (weight 1/2 for odd i, 0 for even); a parent on the wall takes
// weight 0 and its load goes to a clamped interior point, so all 8 loads are
// issued unconditionally (the zero-weight terms add nothing to the sum).
__device__ __forceinline__ float mg_prolong(const MGLevelDev& C, int i, int j, int k) {
  int I[2], J[2], K[2];
  float wi[2], wj[2], wk[2];
  auto parents = [](int f, int c, int(&P)[2], float(&w)[2]) {
    const int p0 = f >> 1;
    const bool odd = f & 1;
    w[0] = odd ? 0.5f : 1.0f;
    w[1] = odd ? 0.5f : 0.0f;
    P[0] = p0;
    P[1] = p0 + 1;
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (P[q] < 1 || P[q] > c - 1) {
        w[q] = 0.0f;
        P[q] = min(max(P[q], 1), c - 1);
      }
  };
  parents(i, C.cx, I, wi);
  parents(j, C.cy, J, wj);
  parents(k, C.cz, K, wk);
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) s += (wi[a] * wj[b] * wk[c]) * C.b[mg_cidx(C, I[a], J[b], K[c])];
  return s;
}

// Row u = (i, j, k) of a coarse level's 27-point stencil applied to v.  An
// entry whose neighbour is not an interior point is zero; its load goes to v[u].
__device__ __forceinline__ float mg_stencil_row(const MGLevelDev& L, int u, int i, int j, int k,
                                                const float* __restrict__ v) {
  float acc = 0.0f;
#pragma unroll
  for (int e = 0; e < 27; ++e) {
    const int ci = i + e / 9 - 1, cj = j + (e / 3) % 3 - 1, ck = k + e % 3 - 1;
    const int c = mg_cinterior(L, ci, cj, ck) ? mg_cidx(L, ci, cj, ck) : u;
    acc += L.val[e * L.n + u] * v[c];
  }
  return acc;
}

// Level 0 row sum over the FP32 sliced ELL (diagonal first), four entries per batch.
__device__ __forceinline__ float mg_row32(const DevMesh& m, const float* __restrict__ val, int row,
                                          const float* __restrict__ v) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const float* __restrict__ vp = val + m.ell_off[slice] + lane;
  float acc = 0.0f;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const int c0 = cp[(k + 0) * 64], c1 = cp[(k + 1) * 64], c2 = cp[(k + 2) * 64], c3 = cp[(k + 3) * 64];
    const float a0 = vp[(k + 0) * 64], a1 = vp[(k + 1) * 64], a2 = vp[(k + 2) * 64], a3 = vp[(k + 3) * 64];
    const float x0 = v[c0], x1 = v[c1], x2 = v[c2], x3 = v[c3];
    acc += a0 * x0;
    acc += a1 * x1;
    acc += a2 * x2;
    acc += a3 * x3;
  }
  for (; k < len; ++k) acc += vp[k * 64] * v[cp[k * 64]];
  return acc;
}

// Level 0 FP64 row sum with the operand formed at each column by f(col).
template <class F>
__device__ __forceinline__ double mg_row64(const DevMesh& m, int row, F f) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  double acc = 0.0;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const int c0 = cp[(k + 0) * 64], c1 = cp[(k + 1) * 64], c2 = cp[(k + 2) * 64], c3 = cp[(k + 3) * 64];
    const double a0 = vp[(k + 0) * 64], a1 = vp[(k + 1) * 64], a2 = vp[(k + 2) * 64], a3 = vp[(k + 3) * 64];
    const double x0 = f(c0), x1 = f(c1), x2 = f(c2), x3 = f(c3);
    acc += a0 * x0;
    acc += a1 * x1;
    acc += a2 * x2;
    acc += a3 * x3;
  }
  for (; k < len; ++k) acc += vp[k * 64] * f(cp[k * 64]);
  return acc;
}

// Level 0 stencil rows (the mesh's matrix on its node lattice; mg.hpp): G rows
// row[g] of nodes lo..hi-1 at once, slot s of a row at val[s * ns + (row - lo)]
// with column row + off[s], S slots of at most SM.  A column outside lo..hi-1
// belongs to a slot the row does not use (value zero): its load goes to the
// clamped node.  All loads of the G rows are issued before the sums, which run
// in slot order (ascending columns).
template <int G, int SM>
__device__ __forceinline__ void mg_stencil_rows32(const float* __restrict__ val, int ns, int lo, int hi, int S,
                                                  const int* off, const int (&row)[G], const float* __restrict__ x,
                                                  float (&w)[G]) {
  float v[G][SM], xv[G][SM];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int s = 0; s < SM; ++s)
      if (s < S) {
        v[g][s] = val[s * ns + (row[g] - lo)];
        xv[g][s] = x[min(max(row[g] + off[s], lo), hi - 1)];
      }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    float acc = 0.0f;
#pragma unroll
    for (int s = 0; s < SM; ++s)
      if (s < S) acc = fmaf(v[g][s], xv[g][s], acc);
    w[g] = acc;
  }
}

template <int G, int SM, class F>
__device__ __forceinline__ void mg_stencil_rows64(const double* __restrict__ val, int ns, int lo, int hi, int S,
                                                  const int* off, const int (&row)[G], F f, double (&w)[G]) {
  double v[G][SM], xv[G][SM];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int s = 0; s < SM; ++s)
      if (s < S) {
        v[g][s] = val[s * ns + (row[g] - lo)];
        xv[g][s] = f(min(max(row[g] + off[s], lo), hi - 1));
      }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < SM; ++s)
      if (s < S) acc += v[g][s] * xv[g][s];
    w[g] = acc;
  }
}

}  // namespace mpic
