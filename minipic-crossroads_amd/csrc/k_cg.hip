// Conjugate gradient of ES::solve (ES.cpp:216-231) as CG_solve (CG_solve.hpp:54-114),
// on the device with no host round trip per iteration.
//
// One iteration = two launches:
//   k1: p = beta*p_old + r (p = r at iteration 0) fused into q = A p: every
//       row recomputes the p entries of its columns, so p is never re-read after
//       the update; partial sums of p.q.
//   k2: x += alpha p, r += (-alpha) q, partial sums of r.r.
// The last workgroup of each launch (found with an atomic counter) reduces the
// partials and updates the scalars (alpha; rho, beta, convergence) in CGState.
//
// Dot products use one fixed tree: chunk = 256 rows = one workgroup, item t held
// by thread t as (0 + term); xor butterfly over the 64 lanes of a wave
// (offsets 32..1); waves combine as (w0+w1)+(w2+w3); chunk partials reduce the
// same way after thread t sums partials t, t+256, ... in order.  Every row sum
// of A p is sequential in the CSR column order.  So every value is independent
// of the launch geometry and equals the CPU restatement bit for bit.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

__device__ __forceinline__ double wave_butterfly(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

// Tree of one value per thread over the workgroup; result valid in all threads.
__device__ __forceinline__ double block_tree(double v) {
  __shared__ double ws[BLOCK / 64];
  v = wave_butterfly(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__device__ __forceinline__ double reduce_partials(const double* part, int P) {
  double acc = 0;
  for (int i = threadIdx.x; i < P; i += BLOCK) acc += __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_tree(acc);
}

// Publishes this workgroup's partial and returns true in the last workgroup to finish.
__device__ __forceinline__ bool publish_and_check_last(double* part, double val, unsigned int* counter) {
  __shared__ int last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(part + blockIdx.x, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned int prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// Sequential ELL row sum: acc = 0; acc += val_k * x(col_k) in CSR order.
template <class F>
__device__ __forceinline__ double ell_row(const DevMesh& m, int row, F xval) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int64_t base = m.ell_off[slice] + lane;
  double acc = 0;
  for (int k = 0; k < len; ++k) {
    const int c = m.ell_col[base + (int64_t)k * 64];
    const double v = m.ell_val[base + (int64_t)k * 64];
    acc += v * xval(c);
  }
  return acc;
}

struct CGArgs {
  DevMesh m;
  double *x, *b, *r, *pold, *pnew, *q;
  double *part_bb, *part_rr, *part_pq;
  CGState* st;
  int it;
};

// r = b - A x; rho0 = b.b; rho = r.r (CG_solve.hpp:58-72).
__global__ __launch_bounds__(BLOCK) void k_cg_init(CGArgs a) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double tb = 0, tr = 0;
  if (i < a.m.nn) {
    const double bi = a.b[i];
    const double* x = a.x;
    const double ax = ell_row(a.m, i, [&](int c) { return x[c]; });
    const double ri = bi - ax;
    a.r[i] = ri;
    tb += bi * bi;
    tr += ri * ri;
  }
  const double sb = block_tree(tb);
  const double sr = block_tree(tr);
  __shared__ int last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(a.part_bb + blockIdx.x, sb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.part_rr + blockIdx.x, sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned int prev = __hip_atomic_fetch_add(&a.st->cnt_init, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  const double rho0 = reduce_partials(a.part_bb, gridDim.x);
  const double rho = reduce_partials(a.part_rr, gridDim.x);
  if (threadIdx.x == 0) {
    CGState* st = a.st;
    st->cnt_init = 0;
    st->rho0 = rho0;
    st->rho = rho;
    st->iters = 0;
    st->done = (rho == 0) ? 1 : 0;
    st->converged = (rho == 0) ? 1 : 0;
    __threadfence();
  }
}

// p update fused into q = A p; partial p.q; alpha = rho / p.q (CG_solve.hpp:77-88).
__global__ __launch_bounds__(BLOCK) void k_cg_k1(CGArgs a) {
  if (a.st->done) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const bool first = (a.it == 0);
  const double beta = a.st->beta;
  double t = 0;
  if (i < a.m.nn) {
    const double* r = a.r;
    const double* po = a.pold;
    double pi, qi;
    if (first) {
      pi = r[i];
      qi = ell_row(a.m, i, [&](int c) { return r[c]; });
    } else {
      pi = beta * po[i] + r[i];
      qi = ell_row(a.m, i, [&](int c) { return beta * po[c] + r[c]; });
    }
    a.pnew[i] = pi;
    a.q[i] = qi;
    t += pi * qi;
  }
  const double s = block_tree(t);
  if (!publish_and_check_last(a.part_pq, s, &a.st->cnt_k1)) return;
  const double pq = reduce_partials(a.part_pq, gridDim.x);
  if (threadIdx.x == 0) {
    CGState* st = a.st;
    st->cnt_k1 = 0;
    double alpha = pq;
    alpha = st->rho / alpha;
    st->alpha = alpha;
    __threadfence();
  }
}

// x += alpha p; r += (-alpha) q; rho = r.r; convergence test rho/rho0 < tol^2
// (CG_solve.hpp:91-107); beta = rho/rho_old for the next iteration.
__global__ __launch_bounds__(BLOCK) void k_cg_k2(CGArgs a) {
  if (a.st->done) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const double alpha = a.st->alpha;
  double t = 0;
  if (i < a.m.nn) {
    a.x[i] = a.x[i] + alpha * a.pnew[i];
    const double ri = a.r[i] + -alpha * a.q[i];
    a.r[i] = ri;
    t += ri * ri;
  }
  const double s = block_tree(t);
  if (!publish_and_check_last(a.part_rr, s, &a.st->cnt_k2)) return;
  const double rho = reduce_partials(a.part_rr, gridDim.x);
  if (threadIdx.x == 0) {
    CGState* st = a.st;
    st->cnt_k2 = 0;
    st->rho_old = st->rho;
    st->rho = rho;
    st->iters = a.it + 1;
    if (rho / st->rho0 < st->tol2) {
      st->done = 1;
      st->converged = 1;
    } else if (a.it + 1 >= st->max_iter) {
      st->done = 1;
      st->converged = 0;
    }
    st->beta = rho / st->rho_old;
    __threadfence();
  }
}

__global__ void k_cg_setup(CGState* st, double tol2, int max_iter) {
  CGState h{};
  h.tol2 = tol2;
  h.max_iter = max_iter;
  *st = h;
}

CGArgs make_args(const DevMesh& m, CGBuffers& c, int it) {
  CGArgs a;
  a.m = m;
  a.x = c.x;
  a.b = c.b;
  a.r = c.r;
  a.pold = (it & 1) ? c.p1 : c.p0;
  a.pnew = (it & 1) ? c.p0 : c.p1;
  a.q = c.q;
  a.part_bb = c.part_bb;
  a.part_rr = c.part_rr;
  a.part_pq = c.part_pq;
  a.st = c.st;
  a.it = it;
  return a;
}

}  // namespace

hipError_t launch_cg_init(const DevMesh& m, CGBuffers& c, double tol, int max_iter, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(1), 0, s, c.st, tol * tol, max_iter);
  hipLaunchKernelGGL(k_cg_init, dim3(c.nchunks), dim3(BLOCK), 0, s, make_args(m, c, 0));
  return hipGetLastError();
}

hipError_t launch_cg_iter(const DevMesh& m, CGBuffers& c, int it, hipStream_t s) {
  const CGArgs a = make_args(m, c, it);
  hipLaunchKernelGGL(k_cg_k1, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_cg_k2, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

}  // namespace mpic
