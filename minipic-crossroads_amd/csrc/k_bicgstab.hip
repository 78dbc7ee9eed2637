// BiCGStab_solve (CG_solve.hpp:115-206), the solver ES::solve and
// PIC::time_step call in the reference (ES.cpp:219, PIC.cpp:251), on the device.
//
// x is scaled to 0 and r = b - A x, r~ = r (= b); every Tpetra update(alpha, A,
// beta) is this = beta*this + alpha*A, and alpha uses rhs.v as the reference
// does.  One iteration is three launches:
//   KA: tests of the previous iteration (r.r/rho0 < tol^2, max_iter, rho = r~.r
//       = 0); p = beta*(p + (-omega) v) + r and v = A p (each row recomputes p
//       at its columns from the previous p, v and r); partials of b.v;
//   KB: alpha = rho / b.v; r' = r + (-alpha) v and t = A r' (r' recomputed at
//       the columns); partials of r'.r', t.t and t.r';
//   KC: r'.r' < rho0 tol tol: x += alpha p, converged; t.t = 0 or omega = 0:
//       early exit; otherwise x += alpha p, x += omega r', r = r' + (-omega) t;
//       partials of r.r and b.r.
// Dot products use the fixed tree of k_cg.hip (256-row chunks, butterfly, then
// chunk partials), so every value equals the FP64 restatement's.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

__device__ __forceinline__ double wave_butterfly(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double block_tree(double v) {
  __shared__ double ws[BLOCK / 64];
  v = wave_butterfly(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__device__ __forceinline__ double reduce_partials(const double* __restrict__ part, int P) {
  const double acc = thread_partials(part, P);
  return block_tree(acc);
}

// acc = 0; acc += v_k * f(col_k) in CSR order.
template <class F>
__device__ __forceinline__ double row_sum(const DevMesh& m, int row, F f) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  double acc = 0;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const int c0 = cp[(k + 0) * 64], c1 = cp[(k + 1) * 64], c2 = cp[(k + 2) * 64], c3 = cp[(k + 3) * 64];
    const double v0 = vp[(k + 0) * 64], v1 = vp[(k + 1) * 64], v2 = vp[(k + 2) * 64], v3 = vp[(k + 3) * 64];
    const double x0 = f(c0), x1 = f(c1), x2 = f(c2), x3 = f(c3);
    acc += v0 * x0;
    acc += v1 * x1;
    acc += v2 * x2;
    acc += v3 * x3;
  }
  for (; k < len; ++k) acc += vp[k * 64] * f(cp[k * 64]);
  return acc;
}

struct BiArgs {
  DevMesh m;
  BiBuffers c;
  int it;
};

__global__ void k_bi_setup(BiState* st, double tol, int max_iter) {
  BiState h{};
  h.tol = tol;
  h.max_iter = max_iter;
  *st = h;
}

// x = 0 x; r = b - A x; partials of b.b and r~.r (r~ = r = b).
__global__ __launch_bounds__(BLOCK) void k_bi_init(BiArgs a) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const BiBuffers& c = a.c;
  double tb = 0, tr = 0;
  if (i < a.m.nn) {
    const double xi = 0. * c.x[i];
    c.x[i] = xi;
    const double* __restrict__ x = c.x;
    const double ax = row_sum(a.m, i, [&](int col) { return 0. * x[col]; });
    const double ri = -1.0 * ax + 1.0 * c.b[i];
    c.r[i] = ri;
    tb += c.b[i] * c.b[i];
    tr += ri * ri;
  }
  const double sb = block_tree(tb);
  const double sr = block_tree(tr);
  if (threadIdx.x == 0) {
    c.part_bb[blockIdx.x] = sb;
    c.part_r[2 * blockIdx.x + 1] = sr;  // r~.r with r~ = r
  }
}

__device__ __forceinline__ void finish(BiState* st, int status, int iters) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->done = 1;
    st->status = status;
    st->iters = iters;
  }
}

__global__ __launch_bounds__(BLOCK) void k_bi_a(BiArgs a) {
  BiState* st = a.c.st;
  if (st->done) return;
  const BiBuffers& c = a.c;
  const int it = a.it;
  const int G = gridDim.x;
  // partial rows are interleaved: part_r[2g] = r.r, part_r[2g+1] = r~.r
  double rr_acc = 0, br_acc = 0;
  for (int g = threadIdx.x; g < G; g += BLOCK) {
    rr_acc += c.part_r[2 * g];
    br_acc += c.part_r[2 * g + 1];
  }
  const double resid = block_tree(rr_acc);
  const double rho = block_tree(br_acc);
  double rho0;
  if (it == 0) {
    rho0 = reduce_partials(c.part_bb, G);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rho0 = rho0;
    if (rho == 0) {
      finish(st, 0, 0);
      return;
    }
  } else {
    rho0 = st->rho0;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->resid = resid;
    if (resid / rho0 < st->tol * st->tol) {
      finish(st, 0, it - 1);
      return;
    }
  }
  if (it >= st->max_iter) {
    finish(st, 1, it);
    return;
  }
  if (it > 0 && rho == 0) {
    finish(st, 2, it);
    return;
  }
  const double* __restrict__ pold = (it & 1) ? c.p1 : c.p0;
  double* __restrict__ pnew = (it & 1) ? c.p0 : c.p1;
  const double* __restrict__ vold = (it & 1) ? c.v1 : c.v0;
  double* __restrict__ vnew = (it & 1) ? c.v0 : c.v1;
  const double* __restrict__ r = c.r;
  double beta = 0, omega = 0;
  if (it > 0) {
    const double rho_old = c.hist[3 * (it - 1)], alpha = c.hist[3 * (it - 1) + 1];
    omega = c.hist[3 * (it - 1) + 2];
    beta = rho / rho_old * alpha / omega;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.hist[3 * it] = rho;
    st->rho = rho;
  }
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double t = 0;
  if (i < a.m.nn) {
    double pi, vi;
    if (it == 0) {
      pi = r[i];
      vi = row_sum(a.m, i, [&](int col) { return r[col]; });
    } else {
      pi = beta * (1.0 * pold[i] + -omega * vold[i]) + 1.0 * r[i];
      vi = row_sum(a.m, i, [&](int col) { return beta * (1.0 * pold[col] + -omega * vold[col]) + 1.0 * r[col]; });
    }
    pnew[i] = pi;
    vnew[i] = vi;
    t += c.b[i] * vi;
  }
  const double s = block_tree(t);
  if (threadIdx.x == 0) c.part_bv[blockIdx.x] = s;
}

__global__ __launch_bounds__(BLOCK) void k_bi_b(BiArgs a) {
  BiState* st = a.c.st;
  if (st->done) return;
  const BiBuffers& c = a.c;
  const int it = a.it;
  double* __restrict__ vnew = (it & 1) ? c.v0 : c.v1;
  const double* __restrict__ r = c.r;
  const double alpha = c.hist[3 * it] / reduce_partials(c.part_bv, gridDim.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) c.hist[3 * it + 1] = alpha;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double t1 = 0, t2 = 0, t3 = 0;
  if (i < a.m.nn) {
    const double rn = 1.0 * r[i] + -alpha * vnew[i];
    const double ti = row_sum(a.m, i, [&](int col) { return 1.0 * r[col] + -alpha * vnew[col]; });
    c.rn[i] = rn;
    c.t[i] = ti;
    t1 += rn * rn;
    t2 += ti * ti;
    t3 += ti * rn;
  }
  const double s1 = block_tree(t1);
  const double s2 = block_tree(t2);
  const double s3 = block_tree(t3);
  if (threadIdx.x == 0) {
    c.part_rt[3 * blockIdx.x] = s1;
    c.part_rt[3 * blockIdx.x + 1] = s2;
    c.part_rt[3 * blockIdx.x + 2] = s3;
  }
}

__global__ __launch_bounds__(BLOCK) void k_bi_c(BiArgs a) {
  BiState* st = a.c.st;
  if (st->done) return;
  const BiBuffers& c = a.c;
  const int it = a.it;
  const int G = gridDim.x;
  double a1 = 0, a2 = 0, a3 = 0;
  for (int g = threadIdx.x; g < G; g += BLOCK) {
    a1 += c.part_rt[3 * g];
    a2 += c.part_rt[3 * g + 1];
    a3 += c.part_rt[3 * g + 2];
  }
  const double resid = block_tree(a1);
  const double tdott = block_tree(a2);
  const double tr = block_tree(a3);
  const double alpha = c.hist[3 * it + 1];
  const double* __restrict__ pnew = (it & 1) ? c.p0 : c.p1;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const bool row = i < a.m.nn;
  if (resid < st->rho0 * st->tol * st->tol) {
    if (row) c.x[i] = 1.0 * c.x[i] + alpha * pnew[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) st->resid = resid;
    finish(st, 0, it);
    return;
  }
  const double omega = tr / tdott;
  if (omega == 0.0 || tdott == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->resid = resid;
    finish(st, 2, it);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) c.hist[3 * it + 2] = omega;
  double t1 = 0, t2 = 0;
  if (row) {
    double xi = 1.0 * c.x[i] + alpha * pnew[i];
    xi = 1.0 * xi + omega * c.rn[i];
    c.x[i] = xi;
    const double ri = 1.0 * c.rn[i] + -omega * c.t[i];
    c.r[i] = ri;
    t1 += ri * ri;
    t2 += c.b[i] * ri;
  }
  const double s1 = block_tree(t1);
  const double s2 = block_tree(t2);
  if (threadIdx.x == 0) {
    c.part_r[2 * blockIdx.x] = s1;
    c.part_r[2 * blockIdx.x + 1] = s2;
  }
}

}  // namespace

hipError_t launch_bi_init(const DevMesh& m, BiBuffers& c, double tol, int max_iter, hipStream_t s) {
  if (!c.x || !c.b || !c.r || !c.rn || !c.t || !c.p0 || !c.p1 || !c.v0 || !c.v1 || !c.part_bb || !c.part_r ||
      !c.part_bv || !c.part_rt || !c.hist || !c.st || c.nchunks <= 0)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bi_setup, dim3(1), dim3(1), 0, s, c.st, tol, max_iter);
  hipLaunchKernelGGL(k_bi_init, dim3(c.nchunks), dim3(BLOCK), 0, s, BiArgs{m, c, 0});
  return hipGetLastError();
}

hipError_t launch_bi_iter(const DevMesh& m, BiBuffers& c, int it, hipStream_t s) {
  const BiArgs a{m, c, it};
  hipLaunchKernelGGL(k_bi_a, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_bi_b, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_bi_c, dim3(c.nchunks), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

}  // namespace mpic
