// Precision form of ES::solve (ES.cpp:216-231): CG (CG_solve.hpp:54-114) in FP32
// inside FP64 iterative refinement, to the reference's stopping rule
// |r|^2 < tol^2 |b|^2 on the FP64 residual.
//
// Outer pass (FP64): r = b - A x; stop when r.r < tol^2 b.b; otherwise an inner
// FP32 CG solves A d = r/|r| from d = 0 to the relative residual
// max(in_tol_min, 0.3 tol |b|/|r|), and x += |r| d.  The FP32 matrix holds the
// FP64 stiffness values rounded to FP32 in the same sliced ELL layout; inner dot
// products accumulate in FP64.  Kernels follow the FP64 solver's structure: each
// launch reduces the previous launch's per-workgroup partials redundantly in
// every workgroup, workgroup 0 records the scalars, and kernels after the stop
// return at once, so the host can queue iterations in batches.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

constexpr int IR_MAX_OUTER = 40;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double ws[BLOCK / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__device__ __forceinline__ double reduce(const double* __restrict__ part, int P) {
  double acc = 0;
  for (int i = threadIdx.x; i < P; i += BLOCK) acc += part[i];
  return block_sum(acc);
}

// FP64 row of A x (sliced ELL, entries in CSR order).
__device__ __forceinline__ double row64(const DevMesh& m, int row, const double* __restrict__ x) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  double acc = 0;
  for (int k = 0; k < len; ++k) acc += vp[k * 64] * x[cp[k * 64]];
  return acc;
}

#pragma clang fp contract(fast)

// FP32 row of A p with p = beta p_old + r formed per column.
template <class F>
__device__ __forceinline__ float row32(const DevMesh& m, const float* __restrict__ val, int row, F pval) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const float* __restrict__ vp = val + m.ell_off[slice] + lane;
  float acc = 0;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const int c0 = cp[(k + 0) * 64], c1 = cp[(k + 1) * 64], c2 = cp[(k + 2) * 64], c3 = cp[(k + 3) * 64];
    const float v0 = vp[(k + 0) * 64], v1 = vp[(k + 1) * 64], v2 = vp[(k + 2) * 64], v3 = vp[(k + 3) * 64];
    acc += v0 * pval(c0);
    acc += v1 * pval(c1);
    acc += v2 * pval(c2);
    acc += v3 * pval(c3);
  }
  for (; k < len; ++k) acc += vp[k * 64] * pval(cp[k * 64]);
  return acc;
}

struct IRArgs {
  DevMesh m;
  const float* val32;
  double *x, *b, *r;
  float *d, *r32, *pold, *pnew, *q;
  double *part_bb, *part_rr, *part_irr, *part_ipq;
  double* rho_hist;
  IRState* st;
  int it;
};

__global__ void k_ir_setup(IRState* st, double tol2, int max_iter) {
  IRState h{};
  h.tol2 = tol2;
  h.max_iter = max_iter;
  *st = h;
}

// r = b - A x; partials of b.b and r.r.
__global__ __launch_bounds__(BLOCK) void k_ir_resid(IRArgs a) {
  if (a.st->done) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double tb = 0, tr = 0;
  if (i < a.m.nn) {
    const double bi = a.b[i];
    const double ri = bi - row64(a.m, i, a.x);
    a.r[i] = ri;
    tb = bi * bi;
    tr = ri * ri;
  }
  const double sb = block_sum(tb);
  const double sr = block_sum(tr);
  if (threadIdx.x == 0) {
    a.part_bb[blockIdx.x] = sb;
    a.part_rr[blockIdx.x] = sr;
  }
}

// Outer convergence test; otherwise starts the inner solve on r/|r|.
__global__ __launch_bounds__(BLOCK) void k_ir_start(IRArgs a, double in_tol_min) {
  IRState* st = a.st;
  if (st->done) return;
  const double bb = reduce(a.part_bb, gridDim.x);
  const double rr = reduce(a.part_rr, gridDim.x);
  const int outer = st->outer;
  const bool conv = rr == 0 || rr < st->tol2 * bb;
  const bool stop = conv || outer >= IR_MAX_OUTER;
  const double scale = sqrt(rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->bb = bb;
    st->rr = rr;
    if (stop) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
    } else {
      st->outer = outer + 1;
      st->in_done = 0;
      st->in_iters = 0;
      st->scale = scale;
      const double want = 0.3 * sqrt(st->tol2 * bb / rr);
      const double eps = want > in_tol_min ? want : in_tol_min;
      st->in_tol2 = eps * eps;
    }
  }
  if (stop) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double t = 0;
  if (i < a.m.nn) {
    const float ri = (float)(a.r[i] / scale);
    a.r32[i] = ri;
    a.d[i] = 0.0f;
    t = (double)ri * (double)ri;
  }
  const double s = block_sum(t);
  if (threadIdx.x == 0) a.part_irr[blockIdx.x] = s;
}

// Inner iteration, first half: stopping test, p = beta p_old + r fused into
// q = A p, partials of p.q.
__global__ __launch_bounds__(BLOCK) void k_ir_k1(IRArgs a) {
  IRState* st = a.st;
  if (st->done || st->in_done) return;
  const int it = a.it;
  const double rho = reduce(a.part_irr, gridDim.x);
  const double rho0 = it == 0 ? rho : st->in_rho0;
  const bool conv = rho == 0 || (it > 0 && rho < st->in_tol2 * rho0);
  const bool stop = conv || it >= st->max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->in_rho = rho;
    if (it == 0) st->in_rho0 = rho0;
    a.rho_hist[it] = rho;
    if (stop) {
      st->in_done = 1;
      st->in_iters = it;
      st->total_inner += it;
    }
  }
  if (stop) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double t = 0;
  if (i < a.m.nn) {
    const float* __restrict__ r = a.r32;
    const float* __restrict__ po = a.pold;
    float pi, qi;
    if (it == 0) {
      pi = r[i];
      qi = row32(a.m, a.val32, i, [&](int c) { return r[c]; });
    } else {
      const float beta = (float)(rho / a.rho_hist[it - 1]);
      pi = beta * po[i] + r[i];
      qi = row32(a.m, a.val32, i, [&](int c) { return beta * po[c] + r[c]; });
    }
    a.pnew[i] = pi;
    a.q[i] = qi;
    t = (double)pi * (double)qi;
  }
  const double s = block_sum(t);
  if (threadIdx.x == 0) a.part_ipq[blockIdx.x] = s;
}

// Inner iteration, second half: alpha = rho / p.q; d += alpha p; r -= alpha q.
__global__ __launch_bounds__(BLOCK) void k_ir_k2(IRArgs a) {
  if (a.st->done || a.st->in_done) return;
  const float alpha = (float)(a.rho_hist[a.it] / reduce(a.part_ipq, gridDim.x));
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double t = 0;
  if (i < a.m.nn) {
    a.d[i] = a.d[i] + alpha * a.pnew[i];
    const float ri = a.r32[i] - alpha * a.q[i];
    a.r32[i] = ri;
    t = (double)ri * (double)ri;
  }
  const double s = block_sum(t);
  if (threadIdx.x == 0) a.part_irr[blockIdx.x] = s;
}

// x += |r| d in FP64.
__global__ __launch_bounds__(BLOCK) void k_ir_update(IRArgs a) {
  if (a.st->done) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < a.m.nn) a.x[i] = a.x[i] + a.st->scale * (double)a.d[i];
}

IRArgs make_args(const DevMesh& m, const Mesh32& m32, IRBuffers& c, int it) {
  IRArgs a;
  a.m = m;
  a.val32 = m32.ell_val;
  a.x = c.x;
  a.b = c.b;
  a.r = c.r;
  a.d = c.d;
  a.r32 = c.r32;
  a.pold = (it & 1) ? c.p1 : c.p0;
  a.pnew = (it & 1) ? c.p0 : c.p1;
  a.q = c.q;
  a.part_bb = c.part_a;
  a.part_rr = c.part_a + c.nchunks;
  a.part_irr = c.part_b;
  a.part_ipq = c.part_b + c.nchunks;
  a.rho_hist = c.rho_hist;
  a.st = c.st;
  a.it = it;
  return a;
}

}  // namespace

hipError_t launch_ir_outer(const DevMesh& m, const Mesh32& m32, IRBuffers& c, bool first, double tol, int max_iter,
                           double in_tol_min, hipStream_t s) {
  if (!c.x || !c.b || !c.r || !c.d || !c.r32 || !c.p0 || !c.p1 || !c.q || !c.part_a || !c.part_b || !c.rho_hist ||
      !c.st || !m32.ell_val || c.nchunks <= 0)
    return hipErrorInvalidValue;
  if (first) hipLaunchKernelGGL(k_ir_setup, dim3(1), dim3(1), 0, s, c.st, tol * tol, max_iter);
  const IRArgs a = make_args(m, m32, c, 0);
  hipLaunchKernelGGL(k_ir_resid, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_ir_start, dim3(c.nchunks), dim3(BLOCK), 0, s, a, in_tol_min);
  return hipGetLastError();
}

hipError_t launch_ir_inner(const DevMesh& m, const Mesh32& m32, IRBuffers& c, int it, hipStream_t s) {
  const IRArgs a = make_args(m, m32, c, it);
  hipLaunchKernelGGL(k_ir_k1, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_ir_k2, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ir_update(const DevMesh& m, IRBuffers& c, hipStream_t s) {
  Mesh32 none{};
  hipLaunchKernelGGL(k_ir_update, dim3(c.nchunks), dim3(BLOCK), 0, s, make_args(m, none, c, 0));
  return hipGetLastError();
}

}  // namespace mpic
