// Kernels of the batching form: an ensemble of independent problems stored as
// one disjoint mesh (problem p owns nodes [node_off[p], node_off[p+1]) and
// elements [elem_off[p], elem_off[p+1])).
//
//   k_batch_ir<T>  the solve of every problem in one launch, one workgroup of T
//                  threads per problem: FP32 CG (Chronopoulos-Gear) inside FP64
//                  iterative refinement to the reference's |r| < tol |b|
//                  (CG_solve.hpp:71-72,103-107), each problem with its own
//                  scalars and stopping point; the problem's vectors stay in the
//                  device's caches (at most a few hundred KB per problem);
//   k_batch_stats  per-problem kinetic energy (PIC.cpp:191-200, exact
//                  fixed-point sum) and particle count over the problem's
//                  contiguous range of the sorted list.
// Sums reduce per thread over its rows in order, then by the fixed workgroup
// tree, so results are reproducible for a given T.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

template <int T>
__device__ __forceinline__ void tree2(double& a, double& b, double* ws) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    a = a + __shfl_xor(a, off);
    b = b + __shfl_xor(b, off);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    ws[threadIdx.x >> 6] = a;
    ws[T / 64 + (threadIdx.x >> 6)] = b;
  }
  __syncthreads();
  double sa = ws[0], sb = ws[T / 64];
#pragma unroll
  for (int w = 1; w < T / 64; ++w) {
    sa = sa + ws[w];
    sb = sb + ws[T / 64 + w];
  }
  a = sa;
  b = sb;
}

__device__ __forceinline__ double row64(const DevMesh& m, int row, const double* __restrict__ x) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  double acc = 0;
  for (int k = 0; k < len; ++k) acc += vp[k * 64] * x[cp[k * 64]];
  return acc;
}

__device__ __forceinline__ float row32(const DevMesh& m, const float* __restrict__ val, int row,
                                       const float* __restrict__ x) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const float* __restrict__ vp = val + m.ell_off[slice] + lane;
  float acc = 0;
  for (int k = 0; k < len; ++k) acc = fmaf(vp[k * 64], x[cp[k * 64]], acc);
  return acc;
}

template <int T>
__global__ __launch_bounds__(T) void k_batch_ir(BatchSolveArgs a) {
  __shared__ double ws[2 * (T / 64)];
  const int p = blockIdx.x;
  const int r0 = a.node_off[p], r1 = a.node_off[p + 1];
  const DevMesh& m = a.m;
  double bb = 0, rr = 0;
  int outer = 0, inner_total = 0;
  bool conv = false;
  for (;; ++outer) {
    // r = b - A x (FP64)
    double tb = 0, tr = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += T) {
      const double bi = a.b[i];
      const double ri = bi - row64(m, i, a.x);
      a.r64[i] = ri;
      tb += bi * bi;
      tr += ri * ri;
    }
    tree2<T>(tb, tr, ws);
    bb = tb;
    rr = tr;
    conv = rr == 0 || rr < a.tol2 * bb;
    if (conv || outer >= IR_MAX_OUTER) break;
    // inner FP32 CG on A d = r/|r| to max(in_tol_min, 0.3 tol |b|/|r|)
    const double scale = sqrt(rr);
    const double inv = 1.0 / scale;
    const double want = 0.3 * sqrt(a.tol2 * bb / rr);
    const double eps = want > a.in_tol_min ? want : a.in_tol_min;
    const double in_tol2 = eps * eps;
    for (int i = r0 + threadIdx.x; i < r1; i += T) {
      a.rf[i] = (float)(a.r64[i] * inv);
      a.df[i] = 0.0f;
    }
    __syncthreads();
    double g = 0, d = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += T) {
      const float w = row32(m, a.val32, i, a.rf);
      a.wf[i] = w;
      const float ri = a.rf[i];
      g += (double)ri * (double)ri;
      d += (double)w * (double)ri;
    }
    tree2<T>(g, d, ws);
    double rho0 = g, rho_prev = 1.0, alpha_prev = 1.0;
    int it = 0;
    for (;; ++it) {
      const double gamma = g, delta = d;
      double beta = 0, alpha = 0, denom = delta;
      if (it > 0) {
        beta = gamma / rho_prev;
        denom = delta - beta * gamma / alpha_prev;
      }
      const bool iconv = gamma == 0 || (it > 0 && gamma < in_tol2 * rho0);
      if (iconv || it >= a.max_iter || !(denom > 0)) break;
      alpha = gamma / denom;
      rho_prev = gamma;
      alpha_prev = alpha;
      const float bf = (float)beta, af = (float)alpha;
      // p = r + beta p; s = w + beta s; d += alpha p; r -= alpha s
      for (int i = r0 + threadIdx.x; i < r1; i += T) {
        const float ri = a.rf[i], wi = a.wf[i];
        const float pi = it > 0 ? fmaf(bf, a.pf[i], ri) : ri;
        const float si = it > 0 ? fmaf(bf, a.sf[i], wi) : wi;
        a.pf[i] = pi;
        a.sf[i] = si;
        a.df[i] = fmaf(af, pi, a.df[i]);
        a.rf[i] = fmaf(-af, si, ri);
      }
      __syncthreads();
      // w = A r; r.r and w.r
      g = 0;
      d = 0;
      for (int i = r0 + threadIdx.x; i < r1; i += T) {
        const float w = row32(m, a.val32, i, a.rf);
        a.wf[i] = w;
        const float ri = a.rf[i];
        g += (double)ri * (double)ri;
        d += (double)w * (double)ri;
      }
      tree2<T>(g, d, ws);
    }
    inner_total += it;
    // x += |r| d (FP64)
    for (int i = r0 + threadIdx.x; i < r1; i += T) a.x[i] = fma(scale, (double)a.df[i], a.x[i]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.iters[p] = inner_total;
    a.converged[p] = conv ? 1 : 0;
    a.relres[p] = bb > 0 ? sqrt(rr / bb) : 0.0;
  }
}

// Kinetic energy and particle count of each problem over its range of the
// element-sorted particle list.  The order inside an element is not fixed (the
// ranking's atomics), so the energy terms are summed exactly in 128-bit fixed
// point at scale 2^S (independent of order), then converted once.
__global__ __launch_bounds__(BLOCK) void k_batch_stats(const Rec32* __restrict__ r, const int64_t* __restrict__ poff,
                                                       const double* __restrict__ sp, int S, double* __restrict__ ke,
                                                       int64_t* __restrict__ count) {
  __shared__ fx128 wsum[BLOCK / 64];
  __shared__ long long wcnt[BLOCK / 64];
  const int p = blockIdx.x;
  fx128 s{0, 0};
  long long c = 0;
  for (int64_t i = poff[p] + threadIdx.x; i < poff[p + 1]; i += BLOCK) {
    const Rec32 q = load_rec(r, i);
    if (q.elem < 0) continue;
    const double V[3] = {(double)q.v[0], (double)q.v[1], (double)q.v[2]};
    s = fx_add(s, to_fixed(ke_term(q.w, sp[16 + q.type], V), S));
    c += 1;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long olo = __shfl_xor(s.lo, off);
    const long long ohi = __shfl_xor(s.hi, off);
    s = fx_add(s, fx128{olo, ohi});
    c += __shfl_xor(c, off);
  }
  if ((threadIdx.x & 63) == 0) {
    wsum[threadIdx.x >> 6] = s;
    wcnt[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    fx128 t = wsum[0];
    long long n = wcnt[0];
    for (int w = 1; w < BLOCK / 64; ++w) {
      t = fx_add(t, wsum[w]);
      n += wcnt[w];
    }
    ke[p] = from_fixed(t, S);
    count[p] = n;
  }
}

}  // namespace

hipError_t launch_batch_ir(const BatchSolveArgs& a, int nprob, int threads, hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  switch (threads) {
    case 256: hipLaunchKernelGGL(k_batch_ir<256>, dim3(nprob), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL(k_batch_ir<512>, dim3(nprob), dim3(512), 0, s, a); break;
    case 1024: hipLaunchKernelGGL(k_batch_ir<1024>, dim3(nprob), dim3(1024), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_stats(const Rec32* r, const int64_t* poff, const double* species, int S, double* ke,
                              int64_t* count, int nprob, hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_stats, dim3(nprob), dim3(BLOCK), 0, s, r, poff, species, S, ke, count);
  return hipGetLastError();
}

}  // namespace mpic
