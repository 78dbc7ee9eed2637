// Conjugate gradients in FP64 preconditioned by an FP32 multigrid V-cycle: the
// precision form's solve of A phi = b (ES::solve, ES.cpp:216-231) to the
// reference's stopping rule |r| < tol |b| (CG_solve.hpp:71-72,103-107) when
// the mesh is a brick lattice (hierarchy built by mg_setup.cpp, mg.hpp).
//
// PCG iteration it (standard CG with z = M r; CG_solve.hpp:54-114 is the
// unpreconditioned recurrence):
//   k_mg_ap      rho = r.z (partials of the previous V-cycle); beta = rho/rho_prev;
//                p = z + beta p_old formed at every column of a row, q = A p
//                (FP64); partials of p.q
//   k_mg_update  alpha = rho / p.q; x += alpha p; r -= alpha q; the V-cycle's
//                right-hand side b0 = (float) r and first smoothing step
//                x0 = omega D^-1 b0; partials of r.r
//   V-cycle      k_mg_resid0 (convergence test, then t0 = b0 - A x0), restriction
//                to level 1, the large coarse levels (residual, restriction) one
//                launch each, k_mg_coarse (the small levels and the coarsest
//                dense solve in one workgroup), then back up: prolongation
//                t = x + P z_coarse and the second smoothing step
//                z = t + omega D^-1 (b - A t) per level; level 0's last kernel
//                writes z (into b0) and the partials of r.z.
// The V-cycle is one damped-Jacobi sweep before and one after the coarse
// correction, restriction = P^T of the trilinear interpolation P: a symmetric
// preconditioner (up to FP32 rounding).  Every launch after the stop returns at
// once.  Dot products use the fixed chunk tree of the other solves (chunk =
// 256 rows, partials reduced redundantly by every workgroup of the next
// launch), so results do not depend on scheduling.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

constexpr int CB_THREADS = 1024;  // the coarse-block kernel's single workgroup

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

// The chunk tree: xor butterfly in each wave, then (w0 + w1) + (w2 + w3).
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double ws[BLOCK / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__device__ __forceinline__ double reduce(const double* __restrict__ part, int P) {
  return block_sum(thread_partials(part, P));
}

// The stop flag, read once per workgroup (a launch whose workgroup 0 may set
// it must not let the waves of another workgroup disagree before a barrier).
__device__ __forceinline__ bool wg_done(const MGState* st) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = __hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return flag != 0;
}

__device__ __forceinline__ float w1(int d) { return d == 0 ? 1.0f : 0.5f; }

// Coarse level points: interior lattice points (1..c-1 per direction).
__device__ __forceinline__ int cidx(const MGLevelDev& L, int i, int j, int k) {
  return ((i - 1) * (L.cy - 1) + (j - 1)) * (L.cz - 1) + (k - 1);
}
__device__ __forceinline__ void cdecode(const MGLevelDev& L, int u, int& i, int& j, int& k) {
  const int ez = L.cz - 1, ey = L.cy - 1;
  k = u % ez + 1;
  j = (u / ez) % ey + 1;
  i = u / (ez * ey) + 1;
}
__device__ __forceinline__ bool cinterior(const MGLevelDev& L, int i, int j, int k) {
  return i > 0 && i < L.cx && j > 0 && j < L.cy && k > 0 && k < L.cz;
}
// Level 0 points: every node, numbered as the brick generator does.
__device__ __forceinline__ int nidx(const MGBuffers& g, int i, int j, int k) {
  return (i * (g.ny + 1) + j) * (g.nz + 1) + k;
}

// Row sums of the 27-entry ELL of a coarse level.
__device__ __forceinline__ float crow(const MGLevelDev& L, int u, const float* __restrict__ v) {
  float acc = 0.0f;
#pragma unroll 9
  for (int e = 0; e < 27; ++e) acc += L.val[e * L.n + u] * v[L.col[e * L.n + u]];
  return acc;
}

// Level 0 row sum over the FP32 sliced ELL (diagonal first), four entries per batch.
__device__ __forceinline__ float row32(const DevMesh& m, const float* __restrict__ val, int row,
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
__device__ __forceinline__ double row64(const DevMesh& m, int row, F f) {
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

struct MGArgs {
  DevMesh m;           // level 0 matrix: diagonal-first ELL, FP64 values
  const float* val32;  // its FP32 values
  MGBuffers g;
  double* x;           // phi
  const double* b;     // rhs
  int it;
  int restart;
};

// ---- restriction and prolongation (shared by the multi-workgroup and the
// single-workgroup kernels) ----

// Coarse point u of level lc from level lc - 1's t: b = P^T t, x = omega D^-1 b.
__device__ __forceinline__ void restrict_point(const MGArgs& a, int lc, int u) {
  const MGLevelDev& C = a.g.lev[lc];
  int I, J, K;
  cdecode(C, u, I, J, K);
  float s = 0.0f;
  if (lc == 1) {
    const float* __restrict__ t = a.g.t0;
#pragma unroll
    for (int di = -1; di <= 1; ++di)
#pragma unroll
      for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk)
          s += (w1(di) * w1(dj) * w1(dk)) * t[nidx(a.g, 2 * I + di, 2 * J + dj, 2 * K + dk)];
  } else {
    const MGLevelDev& F = a.g.lev[lc - 1];
#pragma unroll
    for (int di = -1; di <= 1; ++di)
#pragma unroll
      for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk)
          s += (w1(di) * w1(dj) * w1(dk)) * F.t[cidx(F, 2 * I + di, 2 * J + dj, 2 * K + dk)];
  }
  C.b[u] = s;
  C.x[u] = MG_OMEGA * C.dinv[u] * s;
}

// P z of level lc + 1 at point (i, j, k) of level lc (z is in that level's b).
__device__ __forceinline__ float prolong_at(const MGLevelDev& C, int i, int j, int k) {
  float s = 0.0f;
  for (int I = i >> 1; I <= (i + 1) >> 1; ++I)
    for (int J = j >> 1; J <= (j + 1) >> 1; ++J)
      for (int K = k >> 1; K <= (k + 1) >> 1; ++K)
        if (cinterior(C, I, J, K)) s += (w1(i - 2 * I) * w1(j - 2 * J) * w1(k - 2 * K)) * C.b[cidx(C, I, J, K)];
  return s;
}

// ---- kernels ----

__global__ void k_mg_setup(MGState* st, double tol2, int max_iter) {
  MGState h{};
  h.tol2 = tol2;
  h.max_iter = max_iter;
  *st = h;
}

__global__ void k_mg_reopen(MGState* st) {
  st->done = 0;
  st->converged = 0;
}

// The stop test of iteration it on the partials of k_mg_init, in one workgroup
// (the check of the true residual after the recursive one met the tolerance).
__global__ __launch_bounds__(BLOCK) void k_mg_test(MGArgs a) {
  MGState* st = a.g.st;
  const double rr = reduce(a.g.part_rr, a.g.nchunks);
  const double bb = reduce(a.g.part_bb, a.g.nchunks);
  if (threadIdx.x == 0) {
    const bool conv = rr == 0.0 || rr < st->tol2 * bb;
    st->rr = rr;
    st->bb = bb;
    st->iters = a.it;
    if (conv || a.it >= st->max_iter) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
    }
  }
}

// r = b - A x (FP64), b0 = (float) r, x0 = omega D^-1 b0; partials of b.b and r.r.
__global__ __launch_bounds__(BLOCK) void k_mg_init(MGArgs a) {
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0) return;
  const int i = chunk * BLOCK + threadIdx.x;
  double tb = 0.0, tr = 0.0;
  if (i < a.m.nn) {
    const double* __restrict__ x = a.x;
    const double bi = a.b[i];
    const double ri = bi - row64(a.m, i, [&](int c) { return x[c]; });
    a.g.r[i] = ri;
    const float bf = (float)ri;
    a.g.b0[i] = bf;
    a.g.x0[i] = MG_OMEGA * a.g.dinv0[i] * bf;
    tb = bi * bi;
    tr = ri * ri;
  }
  const double sb = block_sum(tb);
  const double sr = block_sum(tr);
  if (threadIdx.x == 0) {
    a.g.part_bb[chunk] = sb;
    a.g.part_rr[chunk] = sr;
  }
}

// Stop test of iteration it (|r|^2 < tol^2 |b|^2, or it >= max_iter), then
// level 0's residual after the first smoothing step: t0 = b0 - A x0.
__global__ __launch_bounds__(BLOCK) void k_mg_resid0(MGArgs a) {
  MGState* st = a.g.st;
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0 || wg_done(st)) return;
  const int i = chunk * BLOCK + threadIdx.x;
  const double rr = reduce(a.g.part_rr, a.g.nchunks);
  const double bb = reduce(a.g.part_bb, a.g.nchunks);
  const double tol2 = st->tol2;
  const bool conv = rr == 0.0 || rr < tol2 * bb;
  const bool stop = conv || a.it >= st->max_iter;
  if (chunk == 0 && threadIdx.x == 0) {
    st->rr = rr;
    st->bb = bb;
    st->iters = a.it;
    if (stop) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
    }
  }
  if (stop) return;
  if (i < a.m.nn) a.g.t0[i] = a.g.b0[i] - row32(a.m, a.val32, i, a.g.x0);
}

// Coarse level l (below the single-workgroup levels): t = b - A x.
__global__ __launch_bounds__(BLOCK) void k_mg_resid(MGArgs a, int l) {
  if (a.g.st->done) return;
  const MGLevelDev& L = a.g.lev[l];
  const int u = blockIdx.x * BLOCK + threadIdx.x;
  if (u < L.n) L.t[u] = L.b[u] - crow(L, u, L.x);
}

// Restriction to coarse level lc.
__global__ __launch_bounds__(BLOCK) void k_mg_restrict(MGArgs a, int lc) {
  if (a.g.st->done) return;
  const int u = blockIdx.x * BLOCK + threadIdx.x;
  if (u < a.g.lev[lc].n) restrict_point(a, lc, u);
}

// Prolongation to coarse level l (from l + 1) and its second smoothing step,
// as two launches: t = x + P z, then z = t + omega D^-1 (b - A t) into b.
__global__ __launch_bounds__(BLOCK) void k_mg_prolong(MGArgs a, int l) {
  if (a.g.st->done) return;
  const MGLevelDev& L = a.g.lev[l];
  const int u = blockIdx.x * BLOCK + threadIdx.x;
  if (u >= L.n) return;
  int i, j, k;
  cdecode(L, u, i, j, k);
  L.t[u] = L.x[u] + prolong_at(a.g.lev[l + 1], i, j, k);
}

__global__ __launch_bounds__(BLOCK) void k_mg_post(MGArgs a, int l) {
  if (a.g.st->done) return;
  const MGLevelDev& L = a.g.lev[l];
  const int u = blockIdx.x * BLOCK + threadIdx.x;
  if (u < L.n) L.b[u] = L.t[u] + MG_OMEGA * L.dinv[u] * (L.b[u] - crow(L, u, L.t));
}

// Levels cb..nlev in one workgroup: down (residual, restriction), the
// coarsest dense solve, up (prolongation, second smoothing step).
__global__ __launch_bounds__(CB_THREADS) void k_mg_coarse(MGArgs a) {
  if (wg_done(a.g.st)) return;
  const int tid = threadIdx.x;
  const int nlev = a.g.nlev;
  for (int l = a.g.cb; l < nlev; ++l) {
    const MGLevelDev& L = a.g.lev[l];
    for (int u = tid; u < L.n; u += CB_THREADS) L.t[u] = L.b[u] - crow(L, u, L.x);
    __syncthreads();
    const int nc = a.g.lev[l + 1].n;
    for (int u = tid; u < nc; u += CB_THREADS) restrict_point(a, l + 1, u);
    __syncthreads();
  }
  {
    const MGLevelDev& C = a.g.lev[nlev];
    const int n = C.n;  // <= MG_DIRECT_MAX <= CB_THREADS: one row of the dense inverse per thread
    float z = 0.0f;
    if (tid < n) {
      // the inverse is stored column-major: (tid, j) at j * n + tid
      const float* __restrict__ col = a.g.cinv + tid;
      const float* __restrict__ b = C.b;
      int j = 0;
      float z1 = 0.0f, z2 = 0.0f, z3 = 0.0f;
      for (; j + 4 <= n; j += 4) {
        z += col[(size_t)j * n] * b[j];
        z1 += col[(size_t)(j + 1) * n] * b[j + 1];
        z2 += col[(size_t)(j + 2) * n] * b[j + 2];
        z3 += col[(size_t)(j + 3) * n] * b[j + 3];
      }
      for (; j < n; ++j) z += col[(size_t)j * n] * b[j];
      z = (z + z1) + (z2 + z3);
    }
    __syncthreads();
    if (tid < n) C.b[tid] = z;
    __syncthreads();
  }
  for (int l = nlev - 1; l >= a.g.cb; --l) {
    const MGLevelDev& L = a.g.lev[l];
    for (int u = tid; u < L.n; u += CB_THREADS) {
      int i, j, k;
      cdecode(L, u, i, j, k);
      L.t[u] = L.x[u] + prolong_at(a.g.lev[l + 1], i, j, k);
    }
    __syncthreads();
    for (int u = tid; u < L.n; u += CB_THREADS) L.b[u] = L.t[u] + MG_OMEGA * L.dinv[u] * (L.b[u] - crow(L, u, L.t));
    __syncthreads();
  }
}

// Level 0 prolongation: t0 = x0 + P z1 on interior nodes, x0 (= 0) on the walls.
__global__ __launch_bounds__(BLOCK) void k_mg_prolong0(MGArgs a) {
  if (a.g.st->done) return;
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0) return;
  const int u = chunk * BLOCK + threadIdx.x;
  if (u >= a.m.nn) return;
  const int nz1 = a.g.nz + 1, ny1 = a.g.ny + 1;
  const int k = u % nz1, j = (u / nz1) % ny1, i = u / (nz1 * ny1);
  float y = a.g.x0[u];
  if (i > 0 && i < a.g.nx && j > 0 && j < a.g.ny && k > 0 && k < a.g.nz) y += prolong_at(a.g.lev[1], i, j, k);
  a.g.t0[u] = y;
}

// Level 0 second smoothing step: z = t0 + omega D^-1 (b0 - A t0) into b0;
// partials of r.z.
__global__ __launch_bounds__(BLOCK) void k_mg_post0(MGArgs a) {
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0 || wg_done(a.g.st)) return;
  const int i = chunk * BLOCK + threadIdx.x;
  double t = 0.0;
  if (i < a.m.nn) {
    const float z = a.g.t0[i] + MG_OMEGA * a.g.dinv0[i] * (a.g.b0[i] - row32(a.m, a.val32, i, a.g.t0));
    a.g.b0[i] = z;
    t = a.g.r[i] * (double)z;
  }
  const double s = block_sum(t);
  if (threadIdx.x == 0) a.g.part_rz[chunk] = s;
}

// rho = r.z; p = z + beta p_old (beta = 0 at the first iteration and after a
// restart) formed at each column; q = A p; partials of p.q.
__global__ __launch_bounds__(BLOCK) void k_mg_ap(MGArgs a) {
  MGState* st = a.g.st;
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0 || wg_done(st)) return;
  const int it = a.it;
  const double rho = reduce(a.g.part_rz, a.g.nchunks);
  const bool fresh = it == 0 || a.restart;
  const double beta = fresh ? 0.0 : rho / a.g.rz_hist[it - 1];
  if (chunk == 0 && threadIdx.x == 0) a.g.rz_hist[it] = rho;
  const float* __restrict__ z = a.g.b0;
  const double* __restrict__ po = a.g.p[(it + 1) & 1];
  double* __restrict__ pn = a.g.p[it & 1];
  const int i = chunk * BLOCK + threadIdx.x;
  double t = 0.0;
  if (i < a.m.nn) {
    double qi;
    if (fresh)
      qi = row64(a.m, i, [&](int c) { return (double)z[c]; });
    else
      qi = row64(a.m, i, [&](int c) { return (double)z[c] + beta * po[c]; });
    const double pi = fresh ? (double)z[i] : (double)z[i] + beta * po[i];
    pn[i] = pi;
    a.g.q[i] = qi;
    t = pi * qi;
  }
  const double s = block_sum(t);
  if (threadIdx.x == 0) a.g.part_pq[chunk] = s;
}

// alpha = rho / p.q; x += alpha p; r -= alpha q; b0 = (float) r, x0 = omega D^-1 b0;
// partials of r.r.  A non-positive p.q (breakdown) stops the solve unconverged.
__global__ __launch_bounds__(BLOCK) void k_mg_update(MGArgs a) {
  MGState* st = a.g.st;
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0 || wg_done(st)) return;
  const double pq = reduce(a.g.part_pq, a.g.nchunks);
  if (!(pq > 0.0)) {
    if (chunk == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->converged = 0;
      st->iters = a.it;
    }
    return;
  }
  const double alpha = a.g.rz_hist[a.it] / pq;
  const int i = chunk * BLOCK + threadIdx.x;
  double t = 0.0;
  if (i < a.m.nn) {
    a.x[i] = a.x[i] + alpha * a.g.p[a.it & 1][i];
    const double ri = a.g.r[i] - alpha * a.g.q[i];
    a.g.r[i] = ri;
    const float bf = (float)ri;
    a.g.b0[i] = bf;
    a.g.x0[i] = MG_OMEGA * a.g.dinv0[i] * bf;
    t = ri * ri;
  }
  const double s = block_sum(t);
  if (threadIdx.x == 0) a.g.part_rr[chunk] = s;
}

MGArgs make_args(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, int it, bool restart) {
  MGArgs a;
  a.m = diag_first(m, m32);
  a.val32 = m32.ell_val;
  a.g = g;
  a.x = x;
  a.b = b;
  a.it = it;
  a.restart = restart ? 1 : 0;
  return a;
}

inline unsigned blocks_for(int n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// The V-cycle of iteration a.it (its first kernel tests the stop).
void enqueue_vcycle(const MGArgs& a, hipStream_t s) {
  const MGBuffers& g = a.g;
  const dim3 grid0(xcd_grid(g.nchunks));
  hipLaunchKernelGGL(k_mg_resid0, grid0, dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mg_restrict, dim3(blocks_for(g.host_lev[1].n)), dim3(BLOCK), 0, s, a, 1);
  for (int l = 1; l < g.cb; ++l) {
    hipLaunchKernelGGL(k_mg_resid, dim3(blocks_for(g.host_lev[l].n)), dim3(BLOCK), 0, s, a, l);
    hipLaunchKernelGGL(k_mg_restrict, dim3(blocks_for(g.host_lev[l + 1].n)), dim3(BLOCK), 0, s, a, l + 1);
  }
  hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(CB_THREADS), 0, s, a);
  for (int l = g.cb - 1; l >= 1; --l) {
    hipLaunchKernelGGL(k_mg_prolong, dim3(blocks_for(g.host_lev[l].n)), dim3(BLOCK), 0, s, a, l);
    hipLaunchKernelGGL(k_mg_post, dim3(blocks_for(g.host_lev[l].n)), dim3(BLOCK), 0, s, a, l);
  }
  hipLaunchKernelGGL(k_mg_prolong0, grid0, dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mg_post0, grid0, dim3(BLOCK), 0, s, a);
}

bool valid(const MGBuffers& g) {
  return g.lev && g.nlev >= 1 && g.nlev <= MG_MAX_LEVELS && g.cb >= 1 && g.cb <= g.nlev && g.dinv0 && g.cinv &&
         g.b0 && g.x0 && g.t0 && g.r && g.q && g.p[0] && g.p[1] && g.part_bb && g.part_rr && g.part_rz &&
         g.part_pq && g.rz_hist && g.st && g.nchunks > 0 && g.host_lev[g.nlev].n <= MG_DIRECT_MAX;
}

}  // namespace

hipError_t launch_mg_start(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, double tol,
                           int max_iter, hipStream_t s) {
  if (!valid(g) || !m32.ell_val || !m32.ell_col || !m32.ell_val64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mg_setup, dim3(1), dim3(1), 0, s, g.st, tol * tol, max_iter);
  const MGArgs a = make_args(m, m32, g, x, b, 0, false);
  hipLaunchKernelGGL(k_mg_init, dim3(xcd_grid(g.nchunks)), dim3(BLOCK), 0, s, a);
  enqueue_vcycle(a, s);
  return hipGetLastError();
}

hipError_t launch_mg_iter(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, int it,
                          bool restart, hipStream_t s) {
  if (!valid(g)) return hipErrorInvalidValue;
  const MGArgs a = make_args(m, m32, g, x, b, it, restart);
  const dim3 grid0(xcd_grid(g.nchunks));
  hipLaunchKernelGGL(k_mg_ap, grid0, dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mg_update, grid0, dim3(BLOCK), 0, s, a);
  enqueue_vcycle(make_args(m, m32, g, x, b, it + 1, false), s);
  return hipGetLastError();
}

hipError_t launch_mg_recheck(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, int it,
                             hipStream_t s) {
  if (!valid(g)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mg_reopen, dim3(1), dim3(1), 0, s, g.st);
  const MGArgs a = make_args(m, m32, g, x, b, it, false);
  hipLaunchKernelGGL(k_mg_init, dim3(xcd_grid(g.nchunks)), dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mg_test, dim3(1), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mg_vcycle(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, int it,
                            hipStream_t s) {
  if (!valid(g)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mg_reopen, dim3(1), dim3(1), 0, s, g.st);
  enqueue_vcycle(make_args(m, m32, g, x, b, it, false), s);
  return hipGetLastError();
}

}  // namespace mpic
