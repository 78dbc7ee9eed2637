// Kernels of the batching form: an ensemble of independent problems stored as
// one disjoint mesh (problem p owns nodes [node_off[p], node_off[p+1]) and
// elements [elem_off[p], elem_off[p+1])).
//
//   k_batch_ir<T>  the solve of every problem in one launch, one workgroup of T
//                  threads per problem: FP32 CG (Chronopoulos-Gear) inside FP64
//                  iterative refinement to the reference's |r| < tol |b|
//                  (CG_solve.hpp:71-72,103-107), each problem with its own
//                  scalars and stopping point; the inner residual, the one
//                  vector the row sums gather, sits in LDS when the largest
//                  problem fits (up to 36864 nodes), the other vectors stay in
//                  the device's caches (at most a few hundred KB per problem);
//   k_batch_stats  per-problem kinetic energy (PIC.cpp:191-200, exact
//                  fixed-point sum) and particle count over the problem's
//                  contiguous range of the sorted list: partials of fixed-size
//                  chunks of the range, then k_batch_stats_finish adds each
//                  problem's partials.
// Sums reduce per thread over its rows in order, then by the fixed workgroup
// tree, so results are reproducible for a given T.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

constexpr int UROWS = 4;  // rows of a thread updated together

// LN > 0: the inner residual (the only vector the row sums gather) lives in
// LDS, LN entries, and every problem has at most LN nodes; LN == 0: in a.rf.
template <int T, int LN>
__global__ __launch_bounds__(T) void k_batch_ir(BatchSolveArgs a) {
  __shared__ double ws[2 * (T / 64)];
  __shared__ float rl[LN > 0 ? LN : 1];
  const int p = blockIdx.x;
  const int r0 = a.node_off[p], r1 = a.node_off[p + 1];
  const DevMesh& m = a.m;
  float* __restrict__ rv;
  if constexpr (LN > 0)
    rv = rl - r0;
  else
    rv = a.rf;
  // w = A r over the problem's rows, SROWS rows of the thread at a time (1024
  // rows of the workgroup in flight together); r.r and w.r summed over the
  // thread's rows in row order
  constexpr int SROWS = 1024 / T;
  auto spmv = [&](double& g, double& d) {
    g = 0;
    d = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += SROWS * T) {
      int row[SROWS];
      bool ok[SROWS];
      float w[SROWS];
#pragma unroll
      for (int r = 0; r < SROWS; ++r) {
        ok[r] = i + r * T < r1;
        row[r] = ok[r] ? i + r * T : i;
      }
      ell_rows32<SROWS>(m, a.val32, row, ok, rv, w);
#pragma unroll
      for (int r = 0; r < SROWS; ++r)
        if (ok[r]) {
          a.wf[row[r]] = w[r];
          const float ri = rv[row[r]];
          g += (double)ri * (double)ri;
          d += (double)w[r] * (double)ri;
        }
    }
  };
  double bb = 0, rr = 0;
  int outer = 0, inner_total = 0;
  bool conv = false;
  for (;; ++outer) {
    // r = b - A x (FP64)
    double tb = 0, tr = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += T) {
      const double bi = a.b[i];
      const double ri = bi - ell_row64(m, i, a.x);
      a.r64[i] = ri;
      tb += bi * bi;
      tr += ri * ri;
    }
    wg_sum2<T>(tb, tr, ws);
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
      rv[i] = (float)(a.r64[i] * inv);
      a.df[i] = 0.0f;
    }
    __syncthreads();
    double g = 0, d = 0;
    spmv(g, d);
    wg_sum2<T>(g, d, ws);
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
      // p = r + beta p; s = w + beta s; d += alpha p; r -= alpha s, UROWS rows
      // of the thread at a time with all their loads in flight
      for (int i = r0 + threadIdx.x; i < r1; i += UROWS * T) {
        float ri[UROWS], wi[UROWS], po[UROWS], so[UROWS], di[UROWS];
#pragma unroll
        for (int u = 0; u < UROWS; ++u) {
          const int j = i + u * T;
          if (j < r1) {
            ri[u] = rv[j];
            wi[u] = a.wf[j];
            po[u] = it > 0 ? a.pf[j] : 0.0f;
            so[u] = it > 0 ? a.sf[j] : 0.0f;
            di[u] = a.df[j];
          }
        }
#pragma unroll
        for (int u = 0; u < UROWS; ++u) {
          const int j = i + u * T;
          if (j < r1) {
            const float pi = it > 0 ? fmaf(bf, po[u], ri[u]) : ri[u];
            const float si = it > 0 ? fmaf(bf, so[u], wi[u]) : wi[u];
            a.pf[j] = pi;
            a.sf[j] = si;
            a.df[j] = fmaf(af, pi, di[u]);
            rv[j] = fmaf(-af, si, ri[u]);
          }
        }
      }
      __syncthreads();
      // w = A r; r.r and w.r
      spmv(g, d);
      wg_sum2<T>(g, d, ws);
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
// element-sorted particle list, in two launches.  Each problem's range is cut
// into chunks of STAT_CHUNK particles (chunk_off[p] is the first chunk of
// problem p); one workgroup per chunk writes its partial, and one wave per
// problem adds the problem's partials.  The order inside an element is not
// fixed (the ranking's atomics), so the energy terms are summed exactly in
// 128-bit fixed point at scale 2^S (independent of order), then converted once.
__global__ __launch_bounds__(BLOCK) void k_batch_stats(const Rec32* __restrict__ r, const int64_t* __restrict__ poff,
                                                       const int32_t* __restrict__ chunk_off, int nprob,
                                                       const double* __restrict__ sp, int S,
                                                       StatPartial* __restrict__ part) {
  __shared__ fx128 wsum[BLOCK / 64];
  __shared__ long long wcnt[BLOCK / 64];
  const int b = blockIdx.x;
  // the problem whose chunks hold b: the last p with chunk_off[p] <= b
  int lo = 0, hi = nprob;  // chunk_off[lo] <= b < chunk_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (chunk_off[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t i0 = poff[lo] + (int64_t)(b - chunk_off[lo]) * STAT_CHUNK;
  const int64_t i1 = min(i0 + STAT_CHUNK, poff[lo + 1]);
  fx128 s{0, 0};
  long long c = 0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += BLOCK) {
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
    part[b] = StatPartial{t.lo, t.hi, n, 0};
  }
}

__global__ __launch_bounds__(64) void k_batch_stats_finish(const StatPartial* __restrict__ part,
                                                           const int32_t* __restrict__ chunk_off, int S,
                                                           double* __restrict__ ke, int64_t* __restrict__ count) {
  const int p = blockIdx.x;
  fx128 s{0, 0};
  long long c = 0;
  for (int j = chunk_off[p] + threadIdx.x; j < chunk_off[p + 1]; j += 64) {
    const StatPartial q = part[j];
    s = fx_add(s, fx128{q.lo, q.hi});
    c += q.n;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long olo = __shfl_xor(s.lo, off);
    const long long ohi = __shfl_xor(s.hi, off);
    s = fx_add(s, fx128{olo, ohi});
    c += __shfl_xor(c, off);
  }
  if (threadIdx.x == 0) {
    ke[p] = from_fixed(s, S);
    count[p] = c;
  }
}

}  // namespace

hipError_t launch_batch_ir(const BatchSolveArgs& a, int nprob, int threads, hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
#define MPIC_BATCH_IR(T, LN) hipLaunchKernelGGL((k_batch_ir<T, LN>), dim3(nprob), dim3(T), 0, s, a)
  const int ln = a.max_nodes <= BATCH_LDS_SMALL ? 1 : a.max_nodes <= BATCH_LDS_LARGE ? 2 : 0;
  switch (threads * 4 + ln) {
    case 256 * 4 + 0: MPIC_BATCH_IR(256, 0); break;
    case 256 * 4 + 1: MPIC_BATCH_IR(256, BATCH_LDS_SMALL); break;
    case 256 * 4 + 2: MPIC_BATCH_IR(256, BATCH_LDS_LARGE); break;
    case 512 * 4 + 0: MPIC_BATCH_IR(512, 0); break;
    case 512 * 4 + 1: MPIC_BATCH_IR(512, BATCH_LDS_SMALL); break;
    case 512 * 4 + 2: MPIC_BATCH_IR(512, BATCH_LDS_LARGE); break;
    case 1024 * 4 + 0: MPIC_BATCH_IR(1024, 0); break;
    case 1024 * 4 + 1: MPIC_BATCH_IR(1024, BATCH_LDS_SMALL); break;
    case 1024 * 4 + 2: MPIC_BATCH_IR(1024, BATCH_LDS_LARGE); break;
    default: return hipErrorInvalidValue;
  }
#undef MPIC_BATCH_IR
  return hipGetLastError();
}

hipError_t launch_batch_stats(const Rec32* r, const int64_t* poff, const int32_t* chunk_off, int nchunks,
                              const double* species, int S, StatPartial* part, double* ke, int64_t* count, int nprob,
                              hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  if (nchunks > 0)
    hipLaunchKernelGGL(k_batch_stats, dim3(nchunks), dim3(BLOCK), 0, s, r, poff, chunk_off, nprob, species, S, part);
  hipLaunchKernelGGL(k_batch_stats_finish, dim3(nprob), dim3(64), 0, s, part, chunk_off, S, ke, count);
  return hipGetLastError();
}

}  // namespace mpic
