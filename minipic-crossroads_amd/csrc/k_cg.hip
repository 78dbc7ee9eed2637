// Conjugate gradient of ES::solve (ES.cpp:216-231) as CG_solve (CG_solve.hpp:54-114),
// on the device with no host round trip per iteration.
//
// One iteration = two launches:
//   k1: convergence test of the previous residual; p = beta*p_old + r (p = r at
//       iteration 0) fused into q = A p -- every row recomputes the p entries of
//       its columns, so p is never re-read after the update; partials of p.q.
//   k2: alpha = rho / p.q; x += alpha p, r += (-alpha) q; partials of r.r.
// Each launch reduces the previous launch's partials redundantly in every
// workgroup (they are visible at the kernel boundary), so no atomics, fences or
// extra reduction launches are needed; workgroup 0 records the scalars.
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

__device__ __forceinline__ double reduce_partials(const double* __restrict__ part, int P) {
  const double acc = thread_partials(part, P);
  return block_tree(acc);
}

// Sequential ELL row sum: acc = 0; acc += val_k * x(col_k) in CSR order.  Loads
// are issued four entries at a time; the additions keep their order.
template <class F>
__device__ __forceinline__ double ell_row(const DevMesh& m, int row, F xval) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  double acc = 0;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const int c0 = cp[(k + 0) * 64], c1 = cp[(k + 1) * 64], c2 = cp[(k + 2) * 64], c3 = cp[(k + 3) * 64];
    const double v0 = vp[(k + 0) * 64], v1 = vp[(k + 1) * 64], v2 = vp[(k + 2) * 64], v3 = vp[(k + 3) * 64];
    const double x0 = xval(c0), x1 = xval(c1), x2 = xval(c2), x3 = xval(c3);
    acc += v0 * x0;
    acc += v1 * x1;
    acc += v2 * x2;
    acc += v3 * x3;
  }
  for (; k < len; ++k) acc += vp[k * 64] * xval(cp[k * 64]);
  return acc;
}

// The iteration state lives in one 16-B word per row, (p_{i-1}, r_i), so the
// two vectors a row of A p needs at a column come with one load; the words
// alternate between two buffers by iteration parity.
// One ELL row staged in registers (columns, values and the gathered words),
// loaded before anything waits on a reduction.  The sums below keep the
// sequential CSR order of ell_row exactly; rows longer than ROW_REGS read their
// tail directly.  Ten staged entries keep k_cg_k1 within 96 VGPRs, i.e. five
// waves per SIMD: 1280 resident workgroups, so a 65^3-node mesh (1073 chunks)
// runs in one round instead of two.
constexpr int ROW_REGS = 10;

struct Row64 {
  int len;
  double v[ROW_REGS], a[ROW_REGS], b[ROW_REGS];
};

__device__ __forceinline__ void stage_row(const DevMesh& m, int row, const double2* __restrict__ pr, Row64& R) {
  const int slice = row >> 6, lane = row & 63;
  R.len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  int c[ROW_REGS];
#pragma unroll
  for (int k = 0; k < ROW_REGS; ++k) c[k] = k < R.len ? cp[k * 64] : row;
#pragma unroll
  for (int k = 0; k < ROW_REGS; ++k) {
    R.v[k] = k < R.len ? vp[k * 64] : 0.0;
    const double2 t = pr[c[k]];
    R.a[k] = t.x;
    R.b[k] = t.y;
  }
}

// acc = 0; acc += v_k * x_k in CSR order, x_k = r_k (first iteration) or beta p_k + r_k.
__device__ __forceinline__ double row_apply(const DevMesh& m, int row, const Row64& R, bool first, double beta,
                                            const double2* __restrict__ pr) {
  double acc = 0;
#pragma unroll
  for (int k = 0; k < ROW_REGS; ++k)
    if (k < R.len) acc += R.v[k] * (first ? R.b[k] : beta * R.a[k] + R.b[k]);
  if (R.len > ROW_REGS) {
    const int slice = row >> 6, lane = row & 63;
    const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
    const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
    for (int k = ROW_REGS; k < R.len; ++k) {
      const double2 t = pr[cp[k * 64]];
      acc += vp[k * 64] * (first ? t.y : beta * t.x + t.y);
    }
  }
  return acc;
}

struct CGArgs {
  DevMesh m;
  double *x, *b, *q;
  const double2* pr_cur;  // (p_{i-1}, r_i) (read)
  double2* pr_nxt;        // (p_i, r_{i+1}) (written: p by k1, both by k2)
  double *part_bb, *part_rr, *part_pq;
  double* rho_hist;  // [max_iter+1] rho of each iteration
  CGState* st;
  int it;
  int nchunks;  // row chunks of BLOCK rows; partials are indexed by chunk
};

// r = b - A x; partials of b.b and r.r (CG_solve.hpp:58-66).
__global__ __launch_bounds__(BLOCK) void k_cg_init(CGArgs a) {
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  if (chunk < 0) return;
  const int i = chunk * BLOCK + threadIdx.x;
  double tb = 0, tr = 0;
  if (i < a.m.nn) {
    const double bi = a.b[i];
    const double* __restrict__ x = a.x;
    const double ax = ell_row(a.m, i, [&](int c) { return x[c]; });
    const double ri = bi - ax;
    a.pr_nxt[i] = make_double2(0.0, ri);  // the state of iteration 0 (p_{-1} unused)
    tb += bi * bi;
    tr += ri * ri;
  }
  const double sb = block_tree(tb);
  const double sr = block_tree(tr);
  if (threadIdx.x == 0) {
    a.part_bb[chunk] = sb;
    a.part_rr[chunk] = sr;
  }
}

// Convergence test (CG_solve.hpp:71-72,103-107; max_iter :73,113), then
// p = beta p_old + r fused into q = A p; partials of p.q (CG_solve.hpp:77-87).
// The row's operands are loaded before the reduction of the previous partials.
__global__ __launch_bounds__(BLOCK, 5) void k_cg_k1(CGArgs a) {
  CGState* st = a.st;
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  if (chunk < 0 || st->done) return;
  const int it = a.it;
  const int i = chunk * BLOCK + threadIdx.x;
  const bool row = i < a.m.nn;
  Row64 R;
  double poi = 0, ri = 0;
  if (row) {
    stage_row(a.m, i, a.pr_cur, R);
    const double2 own = a.pr_cur[i];
    poi = own.x;
    ri = own.y;
  }
  const double rho = reduce_partials(a.part_rr, a.nchunks);
  const double rho0 = it == 0 ? reduce_partials(a.part_bb, a.nchunks) : st->rho0;
  bool stop = false, conv = false;
  if (it == 0) {
    if (rho == 0) stop = conv = true;
  } else if (rho / rho0 < st->tol2) {
    stop = conv = true;
  }
  if (!stop && it >= st->max_iter) stop = true;
  if (chunk == 0 && threadIdx.x == 0) {
    st->rho = rho;
    if (it == 0) st->rho0 = rho0;
    a.rho_hist[it] = rho;
    if (stop) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
      st->iters = it;
    }
  }
  if (stop) return;
  double t = 0;
  if (row) {
    double pi, qi;
    if (it == 0) {
      pi = ri;
      qi = row_apply(a.m, i, R, true, 0.0, a.pr_cur);
    } else {
      const double beta = rho / a.rho_hist[it - 1];
      pi = beta * poi + ri;
      qi = row_apply(a.m, i, R, false, beta, a.pr_cur);
    }
    a.pr_nxt[i].x = pi;
    a.q[i] = qi;
    t += pi * qi;
  }
  const double s = block_tree(t);
  if (threadIdx.x == 0) a.part_pq[chunk] = s;
}

// alpha = rho / p.q; x += alpha p; r += (-alpha) q; partials of r.r (CG_solve.hpp:87-100).
__global__ __launch_bounds__(BLOCK) void k_cg_k2(CGArgs a) {
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  if (chunk < 0 || a.st->done) return;
  const int i = chunk * BLOCK + threadIdx.x;
  const bool row = i < a.m.nn;
  double xi = 0, pi = 0, ri = 0, qi = 0;
  if (row) {
    xi = a.x[i];
    pi = a.pr_nxt[i].x;
    ri = a.pr_cur[i].y;
    qi = a.q[i];
  }
  double alpha = reduce_partials(a.part_pq, a.nchunks);
  alpha = a.rho_hist[a.it] / alpha;
  double t = 0;
  if (row) {
    a.x[i] = xi + alpha * pi;
    const double rn = ri + -alpha * qi;
    a.pr_nxt[i] = make_double2(pi, rn);
    t += rn * rn;
  }
  const double s = block_tree(t);
  if (threadIdx.x == 0) a.part_rr[chunk] = s;
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
  a.pr_cur = c.pr[it & 1];
  a.pr_nxt = c.pr[(it & 1) ^ 1];
  a.q = c.q;
  a.part_bb = c.part_bb;
  a.part_rr = c.part_rr;
  a.part_pq = c.part_pq;
  a.rho_hist = c.rho_hist;
  a.st = c.st;
  a.it = it;
  a.nchunks = c.nchunks;
  return a;
}

}  // namespace

hipError_t launch_cg_init(const DevMesh& m, CGBuffers& c, double tol, int max_iter, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(1), 0, s, c.st, tol * tol, max_iter);
  // k_cg_init writes the parity-0 words (the arguments of iteration -1).
  hipLaunchKernelGGL(k_cg_init, dim3(xcd_grid(c.nchunks)), dim3(BLOCK), 0, s, make_args(m, c, -1));
  return hipGetLastError();
}

hipError_t launch_cg_setup(CGBuffers& c, double tol, int max_iter, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(1), 0, s, c.st, tol * tol, max_iter);
  return hipGetLastError();
}

hipError_t launch_cg_iter(const DevMesh& m, CGBuffers& c, int it, int max_iter, hipStream_t s) {
  const CGArgs a = make_args(m, c, it);
  const dim3 grid(xcd_grid(c.nchunks));
  hipLaunchKernelGGL(k_cg_k1, grid, dim3(BLOCK), 0, s, a);
  if (it < max_iter) hipLaunchKernelGGL(k_cg_k2, grid, dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

}  // namespace mpic
