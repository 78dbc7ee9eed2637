// The batching form's multigrid solve: every problem of the ensemble solves
// A phi = b (ES::solve, ES.cpp:216-231) to the reference's stopping rule
// |r| < tol |b| (CG_solve.hpp:71-72,103-107) by conjugate gradients in FP64
// preconditioned by one FP32 multigrid V-cycle per iteration -- the single
// context's method (k_mg.hip) -- in one launch, one workgroup per problem, each
// problem with its own hierarchy, scalars and stopping point.
//
// A workgroup runs its problem's iterations with barriers between the phases
// (a phase reads what other threads of the workgroup wrote in the one before):
//   q = A p, p.q;  x += alpha p, r -= alpha q, r.r;  the stop test (recursive
//   residual, then the true residual b - A x, restarting from it when the
//   recursive one was wrong);  b0 = (float) r, x0 = omega D^-1 b0;  the V-cycle:
//   t0 = b0 - A x0;  per coarse level the restriction b = P^T t with
//   x = omega D^-1 b, then t = b - A x;  the coarsest level's restriction (into
//   LDS) and dense solve;  per level on the way up y = x + P z_coarse, then
//   z = y + omega D^-1 (b - A y);  at level 0 the same into b0, with r.z;
//   p = z + beta p.
// Sums reduce over each thread's rows in order, then by the fixed workgroup
// tree (wg_sum2), so results are reproducible for a given workgroup size.
#include "mg_device.hpp"
#include "mpic_device.hpp"
#include "mpic_internal.hpp"

namespace mpic {
namespace {

// Rows of the problem a thread takes together (loads of all of them in flight
// before the arithmetic): G64 for FP64 row sums, G32 for FP32 row sums, GV for
// vector updates; fewer at 1024 threads, whose registers are fewer.
template <int T, int SM>
struct Groups {
  static constexpr int G64 = T >= 1024 || SM > 16 ? 1 : 2;
  static constexpr int G32 = T >= 1024 ? (SM > 16 ? 1 : 2) : 4;
  static constexpr int GV = T >= 1024 ? 2 : 4;
};

// SM: the largest level-0 stencil of the ensemble (slots), 16 or 27.
template <int T, int SM>
__global__ __launch_bounds__(T) void k_batch_mg(BatchMGArgs a) {
  constexpr int G64 = Groups<T, SM>::G64, G32 = Groups<T, SM>::G32, GV = Groups<T, SM>::GV;
  __shared__ double ws[2 * (T / 64)];
  __shared__ float cb[BATCH_MG_DIRECT_MAX];
  __shared__ float part[T];
  const BatchMGProb P = a.prob[blockIdx.x];
  const int r0 = P.r0, r1 = P.r0 + P.nn;
  const int tid = threadIdx.x;
  const DevMesh& m = a.m;
  const MGLevelDev* __restrict__ lev = a.lev + P.lev0;  // lev[l - 1] is coarse level l
  const int ny1 = P.ny + 1, nz1 = P.nz + 1;
  double* __restrict__ x = a.x;
  double* __restrict__ r = a.r;
  double* __restrict__ p = a.p;
  double* __restrict__ q = a.q;
  float* __restrict__ b0 = a.b0;
  float* __restrict__ x0 = a.x0;
  float* __restrict__ t0 = a.t0;
  const float* __restrict__ dinv0 = a.dinv0;
  const double* __restrict__ v64 = a.val0_64 + P.v0_off;
  const float* __restrict__ v32 = a.val0_32 + P.v0_off;
  auto node = [&](int i, int j, int k) { return r0 + (i * ny1 + j) * nz1 + k; };
  // P^T t of the level finer than coarse level l, at its point (I, J, K)
  auto restrict_to = [&](int l, int I, int J, int K) -> float {
    if (l == 1) return mg_restrict(t0, node, P.nx, P.ny, P.nz, I, J, K);
    const MGLevelDev& F = lev[l - 2];
    return mg_restrict(F.t, [&](int i, int j, int k) { return mg_cidx(F, i, j, k); }, F.cx, F.cy, F.cz, I, J, K);
  };

  // One V-cycle on b0 (x0 = omega D^-1 b0 formed, a barrier passed): z into
  // b0; returns the thread's part of r.z.
  auto vcycle = [&]() -> double {
    for (int i = r0 + tid; i < r1; i += G32 * T) {
      int row[G32];
      bool ok[G32];
      float w[G32], bv[G32];
#pragma unroll
      for (int g = 0; g < G32; ++g) {
        ok[g] = i + g * T < r1;
        row[g] = ok[g] ? i + g * T : i;
        bv[g] = b0[row[g]];
      }
      mg_stencil_rows32<G32, SM>(v32, P.nn, r0, r1, P.s0, P.off0, row, x0, w);
#pragma unroll
      for (int g = 0; g < G32; ++g)
        if (ok[g]) t0[row[g]] = bv[g] - w[g];
    }
    __syncthreads();
    for (int l = 1; l < P.nlev; ++l) {
      const MGLevelDev& L = lev[l - 1];
      for (int u = tid; u < L.n; u += T) {
        int I, J, K;
        mg_cdecode(L, u, I, J, K);
        const float s = restrict_to(l, I, J, K);
        L.b[u] = s;
        L.x[u] = MG_OMEGA * L.dinv[u] * s;
      }
      __syncthreads();
      for (int u = tid; u < L.n; u += T) {
        int i, j, k;
        mg_cdecode(L, u, i, j, k);
        L.t[u] = L.b[u] - mg_stencil_row(L, u, i, j, k, L.x);
      }
      __syncthreads();
    }
    {
      // the coarsest level: its restriction into LDS, then z = A^-1 b with the
      // columns of each row split in Q parts (Q threads per row), the parts
      // added in order
      const MGLevelDev& C = lev[P.nlev - 1];
      const int n = C.n;  // <= BATCH_MG_DIRECT_MAX <= T
      for (int u = tid; u < n; u += T) {
        int I, J, K;
        mg_cdecode(C, u, I, J, K);
        cb[u] = restrict_to(P.nlev, I, J, K);
      }
      __syncthreads();
      const int Q = min(T / n, 8);
      const int u = tid % n, qp = tid / n;
      if (qp < Q) {
        const int per = (n + Q - 1) / Q;
        const int j0 = qp * per, j1 = min(n, j0 + per);
        const float* __restrict__ inv = a.cinv + P.cinv_off + u;
        float z0 = 0.0f, z1 = 0.0f, z2 = 0.0f, z3 = 0.0f;
        int j = j0;
        for (; j + 4 <= j1; j += 4) {
          z0 += inv[(size_t)j * n] * cb[j];
          z1 += inv[(size_t)(j + 1) * n] * cb[j + 1];
          z2 += inv[(size_t)(j + 2) * n] * cb[j + 2];
          z3 += inv[(size_t)(j + 3) * n] * cb[j + 3];
        }
        for (; j < j1; ++j) z0 += inv[(size_t)j * n] * cb[j];
        part[qp * n + u] = (z0 + z1) + (z2 + z3);
      }
      __syncthreads();
      if (tid < n) {
        float z = part[tid];
        for (int k = 1; k < Q; ++k) z += part[k * n + tid];
        C.b[tid] = z;
      }
      __syncthreads();
    }
    for (int l = P.nlev - 1; l >= 1; --l) {
      const MGLevelDev& L = lev[l - 1];
      const MGLevelDev& C = lev[l];
      for (int u = tid; u < L.n; u += T) {
        int i, j, k;
        mg_cdecode(L, u, i, j, k);
        L.t[u] = L.x[u] + mg_prolong(C, i, j, k);
      }
      __syncthreads();
      for (int u = tid; u < L.n; u += T) {
        int i, j, k;
        mg_cdecode(L, u, i, j, k);
        L.b[u] = L.t[u] + MG_OMEGA * L.dinv[u] * (L.b[u] - mg_stencil_row(L, u, i, j, k, L.t));
      }
      __syncthreads();
    }
    for (int i = r0 + tid; i < r1; i += GV * T) {
      float y[GV];
#pragma unroll
      for (int g = 0; g < GV; ++g) {
        const int n = i + g * T;
        if (n < r1) {
          const int li = n - r0;
          const int k = li % nz1, j = (li / nz1) % ny1, ii = li / (nz1 * ny1);
          y[g] = x0[n];
          if (ii > 0 && ii < P.nx && j > 0 && j < P.ny && k > 0 && k < P.nz) y[g] += mg_prolong(lev[0], ii, j, k);
        }
      }
#pragma unroll
      for (int g = 0; g < GV; ++g)
        if (i + g * T < r1) t0[i + g * T] = y[g];
    }
    __syncthreads();
    double rz = 0.0;
    for (int i = r0 + tid; i < r1; i += G32 * T) {
      int row[G32];
      bool ok[G32];
      float w[G32], tv[G32], bv[G32], dv[G32];
      double rv[G32];
#pragma unroll
      for (int g = 0; g < G32; ++g) {
        ok[g] = i + g * T < r1;
        row[g] = ok[g] ? i + g * T : i;
        tv[g] = t0[row[g]];
        bv[g] = b0[row[g]];
        dv[g] = dinv0[row[g]];
        rv[g] = r[row[g]];
      }
      mg_stencil_rows32<G32, SM>(v32, P.nn, r0, r1, P.s0, P.off0, row, t0, w);
#pragma unroll
      for (int g = 0; g < G32; ++g)
        if (ok[g]) {
          const float z = tv[g] + MG_OMEGA * dv[g] * (bv[g] - w[g]);
          b0[row[g]] = z;
          rz += rv[g] * (double)z;
        }
    }
    return rz;
  };

  double bb = 0.0, rr = 0.0;
  // r = b - A x (FP64), b.b and r.r
  auto residual = [&]() {
    double tb = 0.0, tr = 0.0;
    for (int i = r0 + tid; i < r1; i += G64 * T) {
      int row[G64];
      bool ok[G64];
      double w[G64], bv[G64];
#pragma unroll
      for (int g = 0; g < G64; ++g) {
        ok[g] = i + g * T < r1;
        row[g] = ok[g] ? i + g * T : i;
        bv[g] = a.b[row[g]];
      }
      mg_stencil_rows64<G64, SM>(v64, P.nn, r0, r1, P.s0, P.off0, row, [&](int c) { return x[c]; }, w);
#pragma unroll
      for (int g = 0; g < G64; ++g)
        if (ok[g]) {
          const double ri = bv[g] - w[g];
          r[row[g]] = ri;
          tb += bv[g] * bv[g];
          tr += ri * ri;
        }
    }
    wg_sum2<T>(tb, tr, ws);
    bb = tb;
    rr = tr;
  };
  // z = M r into b0; returns r.z
  auto precondition = [&]() -> double {
    for (int i = r0 + tid; i < r1; i += GV * T) {
      double rv[GV];
      float dv[GV];
#pragma unroll
      for (int g = 0; g < GV; ++g)
        if (i + g * T < r1) {
          rv[g] = r[i + g * T];
          dv[g] = dinv0[i + g * T];
        }
#pragma unroll
      for (int g = 0; g < GV; ++g)
        if (i + g * T < r1) {
          const float bf = (float)rv[g];
          b0[i + g * T] = bf;
          x0[i + g * T] = MG_OMEGA * dv[g] * bf;
        }
    }
    __syncthreads();
    double rz = vcycle(), none = 0.0;
    wg_sum2<T>(rz, none, ws);
    return rz;
  };
  // p = z + beta p (beta = 0: p = z)
  auto direction = [&](double beta, bool fresh) {
    for (int i = r0 + tid; i < r1; i += GV * T) {
      float zv[GV];
      double pv[GV];
#pragma unroll
      for (int g = 0; g < GV; ++g)
        if (i + g * T < r1) {
          zv[g] = b0[i + g * T];
          pv[g] = fresh ? 0.0 : p[i + g * T];
        }
#pragma unroll
      for (int g = 0; g < GV; ++g)
        if (i + g * T < r1) p[i + g * T] = fresh ? (double)zv[g] : (double)zv[g] + beta * pv[g];
    }
    __syncthreads();
  };

  int it = 0, rechecks = 0;
  bool conv = false;
  residual();
  for (;;) {  // a (re)start from the residual in r
    conv = rr == 0.0 || rr < a.tol2 * bb;
    if (conv || it >= a.max_iter) break;
    double rz = precondition();
    direction(0.0, true);
    bool restart = false;
    for (;;) {
      double pq = 0.0, none = 0.0;
      for (int i = r0 + tid; i < r1; i += G64 * T) {
        int row[G64];
        bool ok[G64];
        double w[G64], pv[G64];
#pragma unroll
        for (int g = 0; g < G64; ++g) {
          ok[g] = i + g * T < r1;
          row[g] = ok[g] ? i + g * T : i;
          pv[g] = p[row[g]];
        }
        mg_stencil_rows64<G64, SM>(v64, P.nn, r0, r1, P.s0, P.off0, row, [&](int c) { return p[c]; }, w);
#pragma unroll
        for (int g = 0; g < G64; ++g)
          if (ok[g]) {
            q[row[g]] = w[g];
            pq += pv[g] * w[g];
          }
      }
      wg_sum2<T>(pq, none, ws);
      if (!(pq > 0.0)) break;  // breakdown: stops unconverged
      const double alpha = rz / pq;
      double tr = 0.0, none2 = 0.0;
      for (int i = r0 + tid; i < r1; i += GV * T) {
        double xv[GV], pv[GV], rv[GV], qv[GV];
#pragma unroll
        for (int g = 0; g < GV; ++g)
          if (i + g * T < r1) {
            xv[g] = x[i + g * T];
            pv[g] = p[i + g * T];
            rv[g] = r[i + g * T];
            qv[g] = q[i + g * T];
          }
#pragma unroll
        for (int g = 0; g < GV; ++g)
          if (i + g * T < r1) {
            x[i + g * T] = xv[g] + alpha * pv[g];
            const double ri = rv[g] - alpha * qv[g];
            r[i + g * T] = ri;
            tr += ri * ri;
          }
      }
      wg_sum2<T>(tr, none2, ws);
      rr = tr;
      ++it;
      if (rr == 0.0 || rr < a.tol2 * bb) {
        residual();  // the true residual decides
        conv = rr == 0.0 || rr < a.tol2 * bb;
        if (!conv && rechecks < MG_MAX_RECHECKS) {
          ++rechecks;
          restart = true;
        }
        break;
      }
      if (it >= a.max_iter) break;
      const double rz_new = precondition();
      const double beta = rz_new / rz;
      rz = rz_new;
      direction(beta, false);
    }
    if (!restart) break;
  }
  if (tid == 0) {
    a.iters[blockIdx.x] = it;
    a.converged[blockIdx.x] = conv ? 1 : 0;
    a.relres[blockIdx.x] = bb > 0.0 ? sqrt(rr / bb) : 0.0;
  }
}

}  // namespace

template <int SM>
hipError_t launch_sm(const BatchMGArgs& a, int nprob, int threads, hipStream_t s) {
  switch (threads) {
    case 256:
      hipLaunchKernelGGL((k_batch_mg<256, SM>), dim3(nprob), dim3(256), 0, s, a);
      break;
    case 512:
      hipLaunchKernelGGL((k_batch_mg<512, SM>), dim3(nprob), dim3(512), 0, s, a);
      break;
    case 1024:
      hipLaunchKernelGGL((k_batch_mg<1024, SM>), dim3(nprob), dim3(1024), 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_mg(const BatchMGArgs& a, int nprob, int threads, int max_slots, hipStream_t s) {
  if (nprob <= 0 || !a.prob || !a.lev || !a.dinv0 || !a.cinv || !a.val0_64 || !a.val0_32 || !a.r || !a.p || !a.q ||
      !a.b0 || !a.x0 || !a.t0 || max_slots < 1 || max_slots > 27)
    return hipErrorInvalidValue;
  return max_slots <= 16 ? launch_sm<16>(a, nprob, threads, s) : launch_sm<27>(a, nprob, threads, s);
}

}  // namespace mpic
