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
//   V-cycle      k_mg_resid0 (convergence test, then t0 = b0 - A x0); per coarse
//                level one k_mg_rr (restriction b = P^T t, first smoothing step
//                x = omega D^-1 b, residual t = b - A x, over 3-D tiles with
//                their halo in LDS); k_mg_cs (the coarsest level's restriction
//                and dense solve, rows spread over workgroups); back up, per
//                level one k_mg_pp (y = x + P z_coarse and the second smoothing
//                step z = y + omega D^-1 (b - A y), into b); level 0's
//                prolongation, then its smoothing step, which writes z (into b0)
//                and the partials of r.z.
// The V-cycle is one damped-Jacobi sweep before and one after the coarse
// correction, restriction = P^T of the trilinear interpolation P: a symmetric
// preconditioner (up to FP32 rounding).  Every launch after the stop returns at
// once.  Dot products use the fixed chunk tree of the other solves (chunk =
// 256 rows, partials reduced redundantly by every workgroup of the next
// launch), so results do not depend on scheduling.
#include "mg_device.hpp"
#include "mpic_device.hpp"
#include "mpic_internal.hpp"

namespace mpic {
namespace {

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

// A state flag, read once per workgroup (a launch whose workgroup 0 may set it
// must not let the waves of another workgroup disagree before a barrier).
__device__ __forceinline__ bool wg_flag(const int* f) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return flag != 0;
}
__device__ __forceinline__ bool wg_done(const MGState* st) { return wg_flag(&st->done); }

// Level 0 points: every node, numbered as the brick generator does.
__device__ __forceinline__ int nidx(const MGBuffers& g, int i, int j, int k) {
  return (i * (g.ny + 1) + j) * (g.nz + 1) + k;
}

struct MGArgs {
  DevMesh m;           // level 0 matrix: diagonal-first ELL, FP64 values
  const float* val32;  // its FP32 values
  MGBuffers g;
  double* x;           // phi
  const double* b;     // rhs
  int it;
  int restart;
  int cond;  // k_mg_init / k_mg_test as the check of a converged solve: only when armed
};

// ---- restriction, prolongation and the tiles of the coarse levels ----

// P^T t of the finer level at point (I, J, K) of a coarse level: level 0 (t0
// on the node lattice) when F.n == 0, else F.
__device__ __forceinline__ float restrict_at(const MGBuffers& g, const MGLevelDev& F, int I, int J, int K) {
  if (F.n == 0) return mg_restrict(g.t0, [&](int i, int j, int k) { return nidx(g, i, j, k); }, g.nx, g.ny, g.nz, I, J, K);
  return mg_restrict(F.t, [&](int i, int j, int k) { return mg_cidx(F, i, j, k); }, F.cx, F.cy, F.cz, I, J, K);
}

// P z of level lc + 1 at point (i, j, k) of level lc (z is in that level's b).
__device__ __forceinline__ float prolong_at(const MGLevelDev& C, int i, int j, int k) {
  return mg_prolong(C, i, j, k);
}

// A coarse level in tiles of TX x TY x TZ points, one per workgroup (one point
// per thread), each with a one-point halo in LDS: a row's 27 stencil columns are
// its halo neighbours.  Halo point (hi, hj, hk) is lattice point
// (i0 + hi, j0 + hj, k0 + hk); the tile's own points have h* in 1..T*.
constexpr int TX = 4, TY = 8, TZ = 8;
constexpr int HY = TY + 2, HZ = TZ + 2, HALO = (TX + 2) * HY * HZ;
static_assert(TX * TY * TZ == BLOCK, "one tile point per thread");

__host__ __device__ inline int tiles_along(int cells, int t) { return (cells - 1 + t - 1) / t; }
inline unsigned tiles_of(const MGLevelDev& L) {
  return (unsigned)(tiles_along(L.cx, TX) * tiles_along(L.cy, TY) * tiles_along(L.cz, TZ));
}

struct TileAt {
  int i0, j0, k0;  // lattice point of halo (0, 0, 0)
  int hc;          // this thread's own halo index
  int i, j, k;     // this thread's own lattice point
};
__device__ __forceinline__ TileAt tile_at(const MGLevelDev& L) {
  const int ty = tiles_along(L.cy, TY), tz = tiles_along(L.cz, TZ);
  const int b = blockIdx.x;
  TileAt T;
  T.i0 = (b / (ty * tz)) * TX;
  T.j0 = ((b / tz) % ty) * TY;
  T.k0 = (b % tz) * TZ;
  const int t = threadIdx.x;
  const int hi = t / (TY * TZ) + 1, hj = (t / TZ) % TY + 1, hk = t % TZ + 1;
  T.hc = (hi * HY + hj) * HZ + hk;
  T.i = T.i0 + hi;
  T.j = T.j0 + hj;
  T.k = T.k0 + hk;
  return T;
}
__device__ __forceinline__ void halo_point(const TileAt& T, int h, int& i, int& j, int& k) {
  i = T.i0 + h / (HY * HZ);
  j = T.j0 + (h / HZ) % HY;
  k = T.k0 + h % HZ;
}
// Halo offset of stencil entry e = (di + 1) 9 + (dj + 1) 3 + (dk + 1).
__device__ __forceinline__ int halo_off(int e) { return ((e / 9 - 1) * HY + (e / 3) % 3 - 1) * HZ + e % 3 - 1; }

// Row sum of the tile's own point over its 27 stencil entries, the operand
// from the halo (zero outside the level, where the entry is zero too).
__device__ __forceinline__ float tile_row(const MGLevelDev& L, int u, const float* hv, int hc) {
  float acc = 0.0f;
#pragma unroll
  for (int e = 0; e < 27; ++e) acc += L.val[e * L.n + u] * hv[hc + halo_off(e)];
  return acc;
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

// Opens the check of the true residual only when the recursive one has met the
// tolerance; armed says whether it did (the check's kernels return at once if not).
__global__ void k_mg_reopen_if(MGState* st) {
  const int arm = st->done && st->converged ? 1 : 0;
  st->armed = arm;
  if (arm) {
    st->done = 0;
    st->converged = 0;
  }
}

// The stop test on the partials of k_mg_init, in one workgroup (the check of the
// true residual after the recursive one met the tolerance, at the iteration that
// met it).
__global__ __launch_bounds__(BLOCK) void k_mg_test(MGArgs a) {
  MGState* st = a.g.st;
  if (a.cond && !wg_flag(&st->armed)) return;
  const int it = st->iters;
  const double rr = reduce(a.g.part_rr, a.g.nchunks);
  const double bb = reduce(a.g.part_bb, a.g.nchunks);
  if (threadIdx.x == 0) {
    const bool conv = rr == 0.0 || rr < st->tol2 * bb;
    st->rr = rr;
    st->bb = bb;
    if (conv || it >= st->max_iter) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
    }
  }
}

// r = b - A x (FP64), b0 = (float) r, x0 = omega D^-1 b0; partials of b.b and r.r
// (as a check, a.cond: only when armed).
__global__ __launch_bounds__(BLOCK) void k_mg_init(MGArgs a) {
  const int chunk = xcd_chunk(blockIdx.x, a.g.nchunks);
  if (chunk < 0) return;
  if (a.cond && !wg_flag(&a.g.st->armed)) return;
  const int i = chunk * BLOCK + threadIdx.x;
  double tb = 0.0, tr = 0.0;
  if (i < a.m.nn) {
    const double* __restrict__ x = a.x;
    const double bi = a.b[i];
    const double ri = bi - mg_row64(a.m, i, [&](int c) { return x[c]; });
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
  if (i < a.m.nn) a.g.t0[i] = a.g.b0[i] - mg_row32(a.m, a.val32, i, a.g.x0);
}

// The coarse-level kernels take their levels by value (the descriptors'
// device pointers): a launch is a short chain of dependent memory accesses.

// Coarse level L on the way down, one tile per workgroup: the restriction
// b = P^T t of the finer level F (level 0 when F.n == 0) and the first
// smoothing step x = omega D^-1 b over the tile and its halo (into LDS), then
// the residual t = b - A x of the tile's points.
__global__ __launch_bounds__(BLOCK) void k_mg_rr(MGArgs a, MGLevelDev L, MGLevelDev F) {
  __shared__ float xs[HALO], bs[HALO];
  if (a.g.st->done) return;  // set by an earlier launch: the same for every wave
  const TileAt T = tile_at(L);
  for (int h = threadIdx.x; h < HALO; h += BLOCK) {
    int i, j, k;
    halo_point(T, h, i, j, k);
    float b = 0.0f, x = 0.0f;
    if (mg_cinterior(L, i, j, k)) {
      b = restrict_at(a.g, F, i, j, k);
      x = MG_OMEGA * L.dinv[mg_cidx(L, i, j, k)] * b;
    }
    bs[h] = b;
    xs[h] = x;
  }
  __syncthreads();
  if (!mg_cinterior(L, T.i, T.j, T.k)) return;
  const int u = mg_cidx(L, T.i, T.j, T.k);
  float acc = 0.0f;
#pragma unroll
  for (int e = 0; e < 27; ++e) acc += L.val[e * L.n + u] * xs[T.hc + halo_off(e)];
  const float b = bs[T.hc];
  L.b[u] = b;
  L.x[u] = xs[T.hc];
  L.t[u] = b - acc;
}

// Coarse level L on the way up, one tile per workgroup: the corrected iterate
// y = x + P z of the coarser level C over the tile and its halo (into LDS),
// then the second smoothing step z = y + omega D^-1 (b - A y) of the tile's
// points, into b.
__global__ __launch_bounds__(BLOCK) void k_mg_pp(MGArgs a, MGLevelDev L, MGLevelDev C) {
  __shared__ float ys[HALO];
  if (a.g.st->done) return;
  const TileAt T = tile_at(L);
  const bool own = mg_cinterior(L, T.i, T.j, T.k);
  const int u = own ? mg_cidx(L, T.i, T.j, T.k) : 0;
  const float bu = own ? L.b[u] : 0.0f, du = own ? L.dinv[u] : 0.0f;
  for (int h = threadIdx.x; h < HALO; h += BLOCK) {
    int i, j, k;
    halo_point(T, h, i, j, k);
    ys[h] = mg_cinterior(L, i, j, k) ? L.x[mg_cidx(L, i, j, k)] + prolong_at(C, i, j, k) : 0.0f;
  }
  __syncthreads();
  if (!own) return;
  float acc = 0.0f;
#pragma unroll
  for (int e = 0; e < 27; ++e) acc += L.val[e * L.n + u] * ys[T.hc + halo_off(e)];
  const float y = ys[T.hc];
  L.b[u] = y + MG_OMEGA * du * (bu - acc);
}

// The coarsest level C: every workgroup restricts the whole level (at most
// MG_DIRECT_MAX points) from the finer level F into LDS, then forms CS_ROWS
// rows of z = A^-1 b from the dense inverse (stored column-major), the columns
// split in CS_PARTS contiguous parts summed in a fixed order.
constexpr int CS_ROWS = 16, CS_PARTS = BLOCK / CS_ROWS, CS_PER = MG_DIRECT_MAX / CS_PARTS;
static_assert(CS_PER * CS_PARTS == MG_DIRECT_MAX, "column parts");
__global__ __launch_bounds__(BLOCK) void k_mg_cs(MGArgs a, MGLevelDev C, MGLevelDev F) {
  __shared__ float bsh[MG_DIRECT_MAX];
  __shared__ float part[CS_PARTS][CS_ROWS];
  if (a.g.st->done) return;
  const int n = C.n;  // <= MG_DIRECT_MAX
  const int rr = threadIdx.x % CS_ROWS, p = threadIdx.x / CS_ROWS;
  const int row = blockIdx.x * CS_ROWS + rr;
  const int per = (n + CS_PARTS - 1) / CS_PARTS;  // <= CS_PER
  const int j0 = p * per, j1 = min(n, j0 + per);
  float m[CS_PER];
  const float* __restrict__ col = a.g.cinv + row;
#pragma unroll
  for (int q = 0; q < CS_PER; ++q) m[q] = (row < n && j0 + q < j1) ? col[(size_t)(j0 + q) * n] : 0.0f;
  for (int u = threadIdx.x; u < MG_DIRECT_MAX; u += BLOCK) {
    float b = 0.0f;  // zero past n: the clamped reads below multiply it by zero
    if (u < n) {
      int I, J, K;
      mg_cdecode(C, u, I, J, K);
      b = restrict_at(a.g, F, I, J, K);
    }
    bsh[u] = b;
  }
  __syncthreads();
  float z = 0.0f, z1 = 0.0f, z2 = 0.0f, z3 = 0.0f;
#pragma unroll
  for (int q = 0; q < CS_PER; q += 4) {
    // entries past j1 are zero and read a valid LDS word (clamped)
    z += m[q] * bsh[min(j0 + q, MG_DIRECT_MAX - 1)];
    z1 += m[q + 1] * bsh[min(j0 + q + 1, MG_DIRECT_MAX - 1)];
    z2 += m[q + 2] * bsh[min(j0 + q + 2, MG_DIRECT_MAX - 1)];
    z3 += m[q + 3] * bsh[min(j0 + q + 3, MG_DIRECT_MAX - 1)];
  }
  part[p][rr] = (z + z1) + (z2 + z3);
  __syncthreads();
  if (threadIdx.x < CS_ROWS && row < n) {
    float w[CS_PARTS];
#pragma unroll
    for (int q = 0; q < CS_PARTS; ++q) w[q] = part[q][rr];
#pragma unroll
    for (int step = 1; step < CS_PARTS; step *= 2)
#pragma unroll
      for (int q = 0; q + step < CS_PARTS; q += 2 * step) w[q] += w[q + step];
    C.b[row] = w[0];
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
  if (i > 0 && i < a.g.nx && j > 0 && j < a.g.ny && k > 0 && k < a.g.nz) y += prolong_at(a.g.host_lev[1], i, j, k);
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
    const float z = a.g.t0[i] + MG_OMEGA * a.g.dinv0[i] * (a.g.b0[i] - mg_row32(a.m, a.val32, i, a.g.t0));
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
      qi = mg_row64(a.m, i, [&](int c) { return (double)z[c]; });
    else
      qi = mg_row64(a.m, i, [&](int c) { return (double)z[c] + beta * po[c]; });
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
  a.cond = 0;
  return a;
}

// The V-cycle of iteration a.it (its first kernel tests the stop): level 0's
// residual; levels 1..nlev-1 restricted with their residuals; the coarsest
// level restricted and solved; back up, levels nlev-1..1 corrected with their
// second smoothing step; level 0's prolongation and second smoothing step.
void enqueue_vcycle(const MGArgs& a, hipStream_t s) {
  const MGBuffers& g = a.g;
  const dim3 grid0(xcd_grid(g.nchunks));
  const MGLevelDev none{};  // n == 0: the finer level is level 0
  hipLaunchKernelGGL(k_mg_resid0, grid0, dim3(BLOCK), 0, s, a);
  for (int l = 1; l < g.nlev; ++l)
    hipLaunchKernelGGL(k_mg_rr, dim3(tiles_of(g.host_lev[l])), dim3(BLOCK), 0, s, a, g.host_lev[l],
                       l == 1 ? none : g.host_lev[l - 1]);
  const unsigned ncs = (unsigned)((g.host_lev[g.nlev].n + CS_ROWS - 1) / CS_ROWS);
  hipLaunchKernelGGL(k_mg_cs, dim3(ncs), dim3(BLOCK), 0, s, a, g.host_lev[g.nlev],
                     g.nlev == 1 ? none : g.host_lev[g.nlev - 1]);
  for (int l = g.nlev - 1; l >= 1; --l)
    hipLaunchKernelGGL(k_mg_pp, dim3(tiles_of(g.host_lev[l])), dim3(BLOCK), 0, s, a, g.host_lev[l], g.host_lev[l + 1]);
  hipLaunchKernelGGL(k_mg_prolong0, grid0, dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mg_post0, grid0, dim3(BLOCK), 0, s, a);
}

bool valid(const MGBuffers& g) {
  return g.lev && g.nlev >= 1 && g.nlev <= MG_MAX_LEVELS && g.dinv0 && g.cinv &&
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

hipError_t launch_mg_recheck(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b,
                             hipStream_t s) {
  if (!valid(g)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mg_reopen_if, dim3(1), dim3(1), 0, s, g.st);
  MGArgs a = make_args(m, m32, g, x, b, 0, false);
  a.cond = 1;
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
