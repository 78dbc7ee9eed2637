// Persistent conjugate-gradient solves: one cooperative launch per solve.
//
// ES::solve's CG (CG_solve.hpp:54-114) runs as one launch when the whole grid
// fits on the device at once (hipLaunchCooperativeKernel refuses otherwise).
// Thread t of workgroup g owns row g*256+t -- the chunk of one workgroup of the
// per-iteration kernels (k_cg.hip) -- and keeps x, r, p and q of its row in
// registers for the whole solve; the matrix is read each iteration (it stays in
// the 256 MB last-level cache at this size).
//
// Workgroups exchange per-chunk partial sums through a grid-wide epoch array: a
// workgroup writes its partial, makes its stores visible at device scope, then
// writes its epoch; a reader waits until every epoch has reached the wanted
// value, then reads all partials.  Every workgroup reduces the partials itself
// in the fixed tree of k_cg.hip, so all of them take the same decisions and
// leave the loop together, and every value equals the per-iteration kernels'
// bit for bit.  A wait longer than WAIT_TICKS sets the error flag and the
// kernel returns, so a launch always ends; the host then solves again with the
// per-iteration kernels.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

constexpr unsigned long long WAIT_TICKS = 100000000ull;  // 1 s of the 100 MHz real-time counter

struct GridSync {
  unsigned* epoch;  // [gridDim.x], zero at launch
  int* err;         // set when a wait times out
};

// Stores of every thread of the workgroup become visible at device scope, then
// the workgroup's epoch is set to ep.
__device__ __forceinline__ void grid_publish(const GridSync& s, unsigned ep) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(s.epoch + blockIdx.x, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until every workgroup has published ep (true), or until the error flag
// is set or the wait exceeds WAIT_TICKS (false, flag set).  The decision is
// taken by thread 0 and shared, so the whole workgroup returns the same value.
__device__ __forceinline__ bool grid_wait(const GridSync& s, unsigned ep) {
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while (true) {
    int ok = 1;
    for (int g = threadIdx.x; g < (int)gridDim.x; g += BLOCK)
      if ((int)(__hip_atomic_load(s.epoch + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - ep) < 0) ok = 0;
    if (__syncthreads_and(ok)) break;
    int quit = 0;
    if (threadIdx.x == 0)
      quit = __hip_atomic_load(s.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
             __builtin_amdgcn_s_memrealtime() - t0 > WAIT_TICKS;
    if (__syncthreads_or(quit)) {
      if (threadIdx.x == 0) __hip_atomic_store(s.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

__device__ __forceinline__ double wave_butterfly(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

// The fixed tree of k_cg.hip: butterfly in each wave, then (w0+w1)+(w2+w3).
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

// Sequential ELL row sum: acc = 0; acc += val_k * f(col_k) in CSR order (the
// order of k_cg.hip).  Columns and values are read from the ELL arrays, four
// entries per batch of loads.
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

struct PCGArgs {
  DevMesh m;
  double *x, *b, *r, *p0, *p1;
  double *part_bb, *part_rr, *part_pq;
  double* rho_hist;
  CGState* st;
  GridSync sync;
  double tol2;
  int max_iter;
};

// The whole CG of CG_solve.hpp:56-113 (the steps of k_cg_init, k_cg_k1, k_cg_k2).
__global__ __launch_bounds__(BLOCK, 5) void k_cg_persist(PCGArgs a) {
  const int row = blockIdx.x * BLOCK + threadIdx.x;
  const bool has = row < a.m.nn;
  const int G = gridDim.x;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  double xi = 0, ri = 0, pi = 0, qi = 0;
  // r = b - A x; partials of b.b and r.r (CG_solve.hpp:58-66)
  double tb = 0, tr = 0;
  if (has) {
    const double* __restrict__ x = a.x;
    const double bi = a.b[row];
    xi = x[row];
    const double ax = row_sum(a.m, row, [&](int c) { return x[c]; });
    ri = bi - ax;
    a.r[row] = ri;
    tb += bi * bi;
    tr += ri * ri;
  }
  double s = block_tree(tb);
  const double s2 = block_tree(tr);
  if (threadIdx.x == 0) {
    a.part_bb[blockIdx.x] = s;
    a.part_rr[blockIdx.x] = s2;
  }
  unsigned ep = 1;
  grid_publish(a.sync, ep);
  if (!grid_wait(a.sync, ep)) return;
  double rho = reduce_partials(a.part_rr, G);
  const double rho0 = reduce_partials(a.part_bb, G);
  double rho_prev = 0;
  for (int it = 0;; ++it) {
    // convergence test (CG_solve.hpp:71-72,103-107), max_iter (:73,113)
    bool stop = false, conv = false;
    if (it == 0) {
      if (rho == 0) stop = conv = true;
    } else if (rho / rho0 < a.tol2) {
      stop = conv = true;
    }
    if (!stop && it >= a.max_iter) stop = true;
    if (lead) {
      a.st->rho = rho;
      if (it == 0) a.st->rho0 = rho0;
      a.rho_hist[it] = rho;
      if (stop) {
        a.st->done = 1;
        a.st->converged = conv ? 1 : 0;
        a.st->iters = it;
      }
    }
    if (stop) break;
    // p = beta p_old + r (p = r at iteration 0), q = A p (CG_solve.hpp:77-84)
    const double* __restrict__ pold = (it & 1) ? a.p1 : a.p0;
    double* __restrict__ pnew = (it & 1) ? a.p0 : a.p1;
    const double* __restrict__ r = a.r;
    double t = 0;
    if (has) {
      if (it == 0) {
        pi = ri;
        qi = row_sum(a.m, row, [&](int c) { return r[c]; });
      } else {
        const double beta = rho / rho_prev;
        pi = beta * pi + ri;
        qi = row_sum(a.m, row, [&](int c) { return beta * pold[c] + r[c]; });
      }
      pnew[row] = pi;
      t += pi * qi;
    }
    s = block_tree(t);
    if (threadIdx.x == 0) a.part_pq[blockIdx.x] = s;
    grid_publish(a.sync, ++ep);
    if (!grid_wait(a.sync, ep)) return;
    // alpha = rho / p.q; x += alpha p; r -= alpha q (CG_solve.hpp:86-100)
    const double alpha = rho / reduce_partials(a.part_pq, G);
    double t2 = 0;
    if (has) {
      xi = xi + alpha * pi;
      ri = ri + -alpha * qi;
      a.r[row] = ri;
      t2 += ri * ri;
    }
    s = block_tree(t2);
    if (threadIdx.x == 0) a.part_rr[blockIdx.x] = s;
    grid_publish(a.sync, ++ep);
    if (!grid_wait(a.sync, ep)) return;
    rho_prev = rho;
    rho = reduce_partials(a.part_rr, G);
  }
  if (has) a.x[row] = xi;
}


// ---- precision form: FP32 CG inside FP64 refinement (k_cg32.hip) -----------

__device__ __forceinline__ double block_sum_d(double v) { return block_tree(v); }

// FP64 row of A x (sequential, as k_cg32.hip's row64).
__device__ __forceinline__ double row64(const DevMesh& m, int row, const double* __restrict__ x) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  double acc = 0;
  for (int k = 0; k < len; ++k) acc += vp[k * 64] * x[cp[k * 64]];
  return acc;
}

constexpr int ROW_REGS32 = 16;  // entries of the inner row sum taken in one batch (as k_cg32.hip)

struct PIRArgs {
  DevMesh m;
  const float* val32;
  double *x, *b, *r;
  float *rv0, *rv1, *wv0, *wv1, *sv0, *sv1;
  double* part_a;  // [2G]: b.b, r.r of the outer residual
  double* part_b;  // [2][2G]: r.r, w.r of the inner residual by iteration parity
  double* part_e;  // [2G]: r.s', p.s' at the start of a continued pass
  IRState* st;
  GridSync sync;
  double tol2, in_tol_min;
  int max_iter;
  int pass_len;  // inner iterations after which a pass ends (k_ir_update's pass_len)
};

// The refinement solve of k_cg32.hip in one launch: per pass the FP64
// residual, the start of the inner pass on r/|r| (continuing the previous
// pass's direction when it stopped at its tolerance), the Chronopoulos-Gear
// iterations and x += |r| d, with the arithmetic of k_ir_resid, k_ir_start,
// k_ir_cg and k_ir_update.
__global__ __launch_bounds__(BLOCK, 5) void k_ir_persist(PIRArgs a) {
  const int row = blockIdx.x * BLOCK + threadIdx.x;
  const bool has = row < a.m.nn;
  const int G = gridDim.x;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int slice = row >> 6, lane = row & 63;
  double xi = has ? a.x[row] : 0.0;
  unsigned ep = 0;
  int total_inner = 0;
  float si = 0, pi = 0;  // the last pass's direction and A times it (this row)
  bool cont_ok = false;
  double scale_prev = 0, rho_last = 0;
  for (int outer = 0;; ++outer) {
    // r = b - A x (FP64); partials of b.b and r.r
    double tb = 0, tr = 0;
    if (has) {
      const double bi = a.b[row];
      const double ri = bi - row64(a.m, row, a.x);
      a.r[row] = ri;
      tb = bi * bi;
      tr = ri * ri;
    }
    double s1 = block_sum_d(tb), s2 = block_sum_d(tr);
    if (threadIdx.x == 0) {
      a.part_a[blockIdx.x] = s1;
      a.part_a[G + blockIdx.x] = s2;
    }
    grid_publish(a.sync, ++ep);
    if (!grid_wait(a.sync, ep)) return;
    // outer test (the reference's |r| < tol |b|); start of the inner solve
    const double bb = reduce_partials(a.part_a, G);
    const double rr = reduce_partials(a.part_a + G, G);
    const bool conv = rr == 0 || rr < a.tol2 * bb;
    const bool stop = conv || outer >= IR_MAX_OUTER;
    const double scale = sqrt(rr);
    const double want = 0.3 * sqrt(a.tol2 * bb / rr);
    const double eps = want > a.in_tol_min ? want : a.in_tol_min;
    const double in_tol2 = eps * eps;
    if (lead) {
      a.st->bb = bb;
      a.st->rr = rr;
      if (stop) {
        a.st->done = 1;
        a.st->converged = conv ? 1 : 0;
        a.st->total_inner = total_inner;
      } else {
        a.st->outer = outer + 1;
        a.st->scale = scale;
        a.st->in_tol2 = in_tol2;
      }
    }
    if (stop) break;
    const bool cont = outer > 0 && cont_ok;
    const double cs = cont ? scale_prev / scale : 0.0;
    const double rho_c = cs * cs * rho_last;
    float ri = 0, wi = 0, di = 0;
    double tg = 0, td = 0, te = 0, tf = 0;
    if (has) {
      const double inv = 1.0 / scale;
      const int len = a.m.ell_len[slice];
      const int* __restrict__ cp = a.m.ell_col + a.m.ell_off[slice] + lane;
      const float* __restrict__ vp = a.val32 + a.m.ell_off[slice] + lane;
      float w = 0;
      for (int k = 0; k < len; ++k) w = fmaf(vp[k * 64], (float)(a.r[cp[k * 64]] * inv), w);
      ri = (float)(a.r[row] * inv);
      wi = w;
      if (cont) {
        const float cf = (float)cs;
        si = cf * si;
        pi = cf * pi;
        te = (double)ri * (double)si;
        tf = (double)pi * (double)si;
      } else {
        si = 0.0f;
        pi = 0.0f;
      }
      a.rv0[row] = ri;
      a.wv0[row] = wi;
      a.sv1[row] = si;  // s' at the columns of iteration 0
      tg = (double)ri * (double)ri;
      td = (double)wi * (double)ri;
    }
    s1 = block_sum_d(tg);
    s2 = block_sum_d(td);
    if (threadIdx.x == 0) {
      a.part_b[blockIdx.x] = s1;
      a.part_b[G + blockIdx.x] = s2;
    }
    if (cont) {
      const double s3 = block_sum_d(te);
      const double s4 = block_sum_d(tf);
      if (threadIdx.x == 0) {
        a.part_e[blockIdx.x] = s3;
        a.part_e[G + blockIdx.x] = s4;
      }
    }
    grid_publish(a.sync, ++ep);
    if (!grid_wait(a.sync, ep)) return;
    double rho0 = 0, rho_prev = rho_c, alpha_prev = 1.0;
    cont_ok = false;
    for (int it = 0;; ++it) {
      const int cur = it & 1;
      const double* __restrict__ pc = a.part_b + (size_t)2 * G * cur;
      double pg = 0, pd = 0;
      for (int k = threadIdx.x; k < G; k += BLOCK) {
        pg += pc[k];
        pd += pc[G + k];
      }
      const double gamma = block_sum_d(pg);
      const double delta = block_sum_d(pd);
      if (it == 0) rho0 = gamma;
      const bool carry = it > 0 || cont;
      double beta = 0, alpha = 0, denom = delta;
      if (it > 0) {
        beta = gamma / rho_prev;
        denom = delta - beta * gamma / alpha_prev;
      } else if (cont) {
        const double rs = reduce_partials(a.part_e, G);
        const double ps = reduce_partials(a.part_e + G, G);
        beta = gamma / rho_prev;
        denom = fma(beta * beta, ps, fma(2.0 * beta, rs, delta));  // same rounding in every build
      }
      if (it >= a.pass_len) {  // the pass ends at its length, as k_ir_update ends it
        total_inner += it;
        cont_ok = it > 0;
        rho_last = it > 0 ? rho_prev : 0.0;
        break;
      }
      const bool iconv = gamma == 0 || (it > 0 && gamma < in_tol2 * rho0);
      const bool istop = iconv || it >= a.max_iter || !(denom > 0);
      if (!istop) alpha = gamma / denom;
      if (lead) {
        a.st->in_rho = gamma;
        if (it == 0) a.st->in_rho0 = rho0;
        if (istop) {
          a.st->in_done = 1;
          a.st->in_iters = it;
        }
      }
      if (istop) {
        total_inner += it;
        cont_ok = iconv && it > 0;
        rho_last = it > 0 ? rho_prev : 0.0;
        break;
      }
      rho_prev = gamma;
      alpha_prev = alpha;
      const float* __restrict__ r_cur = cur ? a.rv1 : a.rv0;
      const float* __restrict__ w_cur = cur ? a.wv1 : a.wv0;
      const float* __restrict__ s_prev = cur ? a.sv0 : a.sv1;
      float* __restrict__ r_nxt = cur ? a.rv0 : a.rv1;
      float* __restrict__ w_nxt = cur ? a.wv0 : a.wv1;
      float* __restrict__ s_cur = cur ? a.sv1 : a.sv0;
      const float bf = (float)beta, af = (float)alpha;
      double t1 = 0, t2 = 0;
      if (has) {
        const int len = a.m.ell_len[slice];
        const int* __restrict__ cp = a.m.ell_col + a.m.ell_off[slice] + lane;
        const float* __restrict__ vp = a.val32 + a.m.ell_off[slice] + lane;
        const float p = carry ? fmaf(bf, pi, ri) : ri;
        const float sv = carry ? fmaf(bf, si, wi) : wi;
        const float rn = fmaf(-af, sv, ri);
        s_cur[row] = sv;
        di = fmaf(af, p, di);
        r_nxt[row] = rn;
        // w = A r_new, r_new at each column recomputed as r - alpha (w + beta s_prev)
        float w = 0;
        // entries k < ROW_REGS32 past the row's length are (col = row, val = 0), as
        // k_ir_cg stages them; four entries per batch of loads
        const int kend = len > ROW_REGS32 ? len : ROW_REGS32;
        auto term = [&](int k, float wacc) {
          const int c = k < len ? cp[k * 64] : row;
          const float v = k < len ? vp[k * 64] : 0.0f;
          const float sk = carry ? fmaf(bf, s_prev[c], w_cur[c]) : w_cur[c];
          return fmaf(v, fmaf(-af, sk, r_cur[c]), wacc);
        };
        int k = 0;
#pragma unroll 1
        for (; k + 4 <= kend; k += 4) {
          const int c0 = k < len ? cp[k * 64] : row, c1 = k + 1 < len ? cp[(k + 1) * 64] : row;
          const int c2 = k + 2 < len ? cp[(k + 2) * 64] : row, c3 = k + 3 < len ? cp[(k + 3) * 64] : row;
          const float v0 = k < len ? vp[k * 64] : 0.0f, v1 = k + 1 < len ? vp[(k + 1) * 64] : 0.0f;
          const float v2 = k + 2 < len ? vp[(k + 2) * 64] : 0.0f, v3 = k + 3 < len ? vp[(k + 3) * 64] : 0.0f;
          const float s0 = carry ? fmaf(bf, s_prev[c0], w_cur[c0]) : w_cur[c0];
          const float s1v = carry ? fmaf(bf, s_prev[c1], w_cur[c1]) : w_cur[c1];
          const float s2v = carry ? fmaf(bf, s_prev[c2], w_cur[c2]) : w_cur[c2];
          const float s3 = carry ? fmaf(bf, s_prev[c3], w_cur[c3]) : w_cur[c3];
          const float r0 = r_cur[c0], r1 = r_cur[c1], r2 = r_cur[c2], r3 = r_cur[c3];
          w = fmaf(v0, fmaf(-af, s0, r0), w);
          w = fmaf(v1, fmaf(-af, s1v, r1), w);
          w = fmaf(v2, fmaf(-af, s2v, r2), w);
          w = fmaf(v3, fmaf(-af, s3, r3), w);
        }
        for (; k < kend; ++k) w = term(k, w);
        w_nxt[row] = w;
        pi = p;
        si = sv;
        ri = rn;
        wi = w;
        t1 = (double)rn * (double)rn;
        t2 = (double)w * (double)rn;
      }
      s1 = block_sum_d(t1);
      s2 = block_sum_d(t2);
      if (threadIdx.x == 0) {
        double* pn = a.part_b + (size_t)2 * G * (cur ^ 1);
        pn[blockIdx.x] = s1;
        pn[G + blockIdx.x] = s2;
      }
      grid_publish(a.sync, ++ep);
      if (!grid_wait(a.sync, ep)) return;
    }
    // x += |r| d (FP64)
    if (has) {
      xi = fma(scale, (double)di, xi);
      a.x[row] = xi;
    }
    scale_prev = scale;
    grid_publish(a.sync, ++ep);
    if (!grid_wait(a.sync, ep)) return;
  }
}

}  // namespace

// Largest grid of a persistent kernel that is resident at once (0 when
// cooperative launches are unsupported).
int persist_capacity(int which) {
  int dev = 0, coop = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const hipError_t e = which == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_persist, BLOCK, 0)
                                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ir_persist, BLOCK, 0);
  if (e != hipSuccess) return 0;
  return per_cu * cus;
}

hipError_t launch_cg_persist(const DevMesh& m, CGBuffers& c, double tol, int max_iter, unsigned* epoch, int* err,
                             hipStream_t s) {
  if (!c.x || !c.b || !c.r || !c.p0 || !c.p1 || !c.part_bb || !c.part_rr || !c.part_pq || !c.rho_hist || !c.st ||
      !epoch || !err || c.nchunks <= 0 || (int64_t)c.nchunks * BLOCK < m.nn)
    return hipErrorInvalidValue;
  PCGArgs a;
  a.m = m;
  a.x = c.x;
  a.b = c.b;
  a.r = c.r;
  a.p0 = c.p0;
  a.p1 = c.p1;
  a.part_bb = c.part_bb;
  a.part_rr = c.part_rr;
  a.part_pq = c.part_pq;
  a.rho_hist = c.rho_hist;
  a.st = c.st;
  a.sync = GridSync{epoch, err};
  a.tol2 = tol * tol;
  a.max_iter = max_iter;
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_cg_persist), dim3(c.nchunks), dim3(BLOCK), args,
                                    0, s);
}

}  // namespace mpic

namespace mpic {

hipError_t launch_ir_persist(const DevMesh& m, const Mesh32& m32, IRBuffers& c, double tol, int max_iter,
                             double in_tol_min, int pass_len, unsigned* epoch, int* err, hipStream_t s) {
  if (!c.x || !c.b || !c.r || !c.rv[0] || !c.rv[1] || !c.wv[0] || !c.wv[1] || !c.sv[0] || !c.sv[1] || !c.part_a ||
      !c.part_b || !c.part_e || !c.st || !m32.ell_val || !epoch || !err || c.nchunks <= 0 || (int64_t)c.nchunks * BLOCK < m.nn)
    return hipErrorInvalidValue;
  PIRArgs a;
  a.m = diag_first(m, m32);
  a.val32 = m32.ell_val;
  a.x = c.x;
  a.b = c.b;
  a.r = c.r;
  a.rv0 = c.rv[0];
  a.rv1 = c.rv[1];
  a.wv0 = c.wv[0];
  a.wv1 = c.wv[1];
  a.sv0 = c.sv[0];
  a.sv1 = c.sv[1];
  a.part_a = c.part_a;
  a.part_b = c.part_b;
  a.part_e = c.part_e;
  a.st = c.st;
  a.sync = GridSync{epoch, err};
  a.tol2 = tol * tol;
  a.in_tol_min = in_tol_min;
  a.max_iter = max_iter;
  a.pass_len = pass_len;
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_ir_persist), dim3(c.nchunks), dim3(BLOCK), args,
                                    0, s);
}

}  // namespace mpic
