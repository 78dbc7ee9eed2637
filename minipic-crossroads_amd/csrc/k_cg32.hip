// Precision form of ES::solve (ES.cpp:216-231): conjugate gradients in FP32
// inside FP64 iterative refinement, to the reference's stopping rule
// |r|^2 < tol^2 |b|^2 (CG_solve.hpp:71-72,103-107) on the FP64 residual.
//
// Outer pass (FP64): r = b - A x; stop when r.r < tol^2 b.b; otherwise an inner
// FP32 pass on A d = r/|r| from d = 0 to the relative residual
// max(in_tol_min, 0.3 tol |b|/|r|), then x += |r| d.  The FP32 matrix holds the
// FP64 stiffness values rounded to FP32 in the same sliced ELL layout.  Passes
// are reliable updates of one CG run (Sleijpen and van der Vorst; van der
// Vorst and Ye): the true FP64 residual replaces the recursive FP32 one, and
// the search direction p and s = A p carry over (rescaled to the new pass's
// units), so the iteration count stays near the FP64 CG's instead of paying
// for a restart at every pass.
//
// The inner solve is the single-reduction form of CG (Chronopoulos and Gear):
// with w = A r and s = A p carried as vectors, both dot products of an
// iteration (r.r and w.r) are reduced together, so one launch per iteration
// suffices.  A launch reduces the previous launch's per-workgroup partials
// redundantly in every workgroup, updates its own rows (p, s, d, r), and forms
// w = A r_new for its rows by recomputing r_new at each column from the previous
// r, w and s (r_new = r - alpha (w + beta s_prev)), so no other workgroup's new
// values are read.  Vectors and partials alternate between two buffers by
// iteration parity.  Dot products accumulate in FP64.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

// Two workgroup sums with one pair of barriers; each is the same tree as block_sum.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double ws[2][BLOCK / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    ws[0][threadIdx.x >> 6] = a;
    ws[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  a = (ws[0][0] + ws[0][1]) + (ws[0][2] + ws[0][3]);
  b = (ws[1][0] + ws[1][1]) + (ws[1][2] + ws[1][3]);
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double ws[BLOCK / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__device__ __forceinline__ double uniform(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double reduce(const double* __restrict__ part, int P) {
  return block_sum(thread_partials(part, P));
}

// FP64 row of A x (sliced ELL, entries in CSR order).
// Loads of a row's first STAGE entries issued together; later entries in a loop.
constexpr int STAGE = 15;

// FP64 row of A x: acc = 0; acc += val_k * x(col_k) in stored order.
__device__ __forceinline__ double row64(const DevMesh& m, int row, const double* __restrict__ x) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  int c[STAGE];
#pragma unroll
  for (int k = 0; k < STAGE; ++k) c[k] = k < len ? cp[k * 64] : row;
  double acc = 0;
#pragma unroll
  for (int k = 0; k < STAGE; ++k) {
    const double v = k < len ? vp[k * 64] : 0.0;
    const double xv = x[c[k]];
    if (k < len) acc += v * xv;
  }
  for (int k = STAGE; k < len; ++k) acc += vp[k * 64] * x[cp[k * 64]];
  return acc;
}

#pragma clang fp contract(fast)

// Inner state of iteration i, one 16-B word per row: (r_i, w_i = A r_i,
// s_{i-1}, p_{i-1}).  An iteration reads its row's word and gathers its
// columns' words with one load each, and writes the next iteration's word with
// one store (the words alternate between two buffers by iteration parity).
// One ELL row staged in registers: columns, values and the gathered words are
// loaded before anything waits on a reduction.  Rows longer than ROW_REGS read
// their tail directly.  Fifteen entries (the longest row of a Kuhn mesh) keep
// k_ir_cg within 96 VGPRs: five waves per SIMD, 1280 resident workgroups, so a
// 65^3-node mesh (1073 chunks) runs in one round (14% less time per iteration
// than at four waves).
constexpr int ROW_REGS = 15;

struct Row32 {
  int len;
  int c[ROW_REGS];
  float v[ROW_REGS];
  float3 g[ROW_REGS];  // (r, w, s) at column c
  float p0;            // p of entry 0: the row's own word (its diagonal comes first)
};

__device__ __forceinline__ void stage_cols(const DevMesh& m, int row, Row32& R) {
  const int slice = __builtin_amdgcn_readfirstlane(row >> 6), lane = row & 63;  // one slice per wave
  R.len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
#pragma unroll
  for (int k = 0; k < ROW_REGS; ++k) R.c[k] = k < R.len ? cp[k * 64] : row;
}

__device__ __forceinline__ void stage_gathers(const DevMesh& m, const float* __restrict__ val, int row,
                                              const float4* __restrict__ q, Row32& R) {
  const int slice = __builtin_amdgcn_readfirstlane(row >> 6), lane = row & 63;
  const float* __restrict__ vp = val + m.ell_off[slice] + lane;
#pragma unroll
  for (int k = 0; k < ROW_REGS; ++k) {
    R.v[k] = k < R.len ? vp[k * 64] : 0.0f;
    const float4 t = q[R.c[k]];
    R.g[k] = make_float3(t.x, t.y, t.z);
    if (k == 0) R.p0 = t.w;
  }
}

struct IRArgs {
  DevMesh m;
  const float* val32;
  double *x, *b, *r;
  float* d;
  const float4* q_cur;         // (r_i, w_i, s_{i-1}, p_{i-1}) (read)
  float4* q_nxt;               // (r_{i+1}, w_{i+1}, s_i, p_i) (written)
  double *part_bb, *part_rr;   // outer partials
  double *part_cur, *part_nxt; // inner partials [2*nchunks]: r.r then w.r
  double* part_e;              // [2*nchunks]: r.s', p.s' at the start of a continued pass
  double *rho_hist, *alpha_hist;
  IRState* st;
  int it;
  int spec;     // inner launch queued past the predicted end of the pass
  int nchunks;  // row chunks of BLOCK rows; partials are indexed by chunk
};

__global__ void k_ir_setup(IRState* st, double tol2, int max_iter) {
  IRState h{};
  h.tol2 = tol2;
  h.max_iter = max_iter;
  *st = h;
}

// r = b - A x; partials of b.b and r.r.
__global__ __launch_bounds__(BLOCK) void k_ir_resid(IRArgs a) {
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  if (chunk < 0 || a.st->done) return;
  const int i = chunk * BLOCK + threadIdx.x;
  double tb = 0, tr = 0;
  if (i < a.m.nn) {
    const double bi = a.b[i];
    const double ri = bi - row64(a.m, i, a.x);
    a.r[i] = ri;
    tb = bi * bi;
    tr = ri * ri;
  }
  double sb = tb, sr = tr;
  block_sum2(sb, sr);
  if (threadIdx.x == 0) {
    a.part_bb[chunk] = sb;
    a.part_rr[chunk] = sr;
  }
}

// Outer convergence test; otherwise starts an inner pass on r' = r/|r| (FP32)
// with w' = A r' (each column's r' formed from the FP64 residual), d = 0, and
// the partials of r'.r' and w'.r'.  When the previous pass stopped at its
// tolerance, the pass continues its direction (reliable update of the
// residual): p' = c p and s' = c s with c = |r_prev| / |r| (the previous
// pass's units to this one's), the previous last r.r rescaled by c^2 for beta,
// and the partials of r'.s' and p'.s' for the exact p.Ap of the first step.
// Otherwise p' = s' = 0 (a fresh start).  Reads only state fields that no
// workgroup of this launch writes.
__global__ __launch_bounds__(BLOCK) void k_ir_start(IRArgs a, double in_tol_min) {
  IRState* st = a.st;
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  if (chunk < 0 || st->done) return;
  const int outer = st->outer;
  const int cont = outer > 0 && st->cont_ok;
  const double scale_prev = st->scale_prev, rho_last = st->rho_last, tol2 = st->tol2;
  const int in_last = st->in_last;
  const double bb = reduce(a.part_bb, a.nchunks);
  const double rr = reduce(a.part_rr, a.nchunks);
  const bool conv = rr == 0 || rr < tol2 * bb;
  const bool stop = conv || outer >= IR_MAX_OUTER;
  const double scale = sqrt(rr);
  const double cs = cont ? scale_prev / scale : 0.0;
  if (chunk == 0 && threadIdx.x == 0) {
    st->bb = bb;
    st->rr = rr;
    if (stop) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
    } else {
      st->scale = scale;
      st->cont = cont;
      st->rho_c = cs * cs * rho_last;
      const double want = 0.3 * sqrt(tol2 * bb / rr);
      const double eps = want > in_tol_min ? want : in_tol_min;
      st->in_tol2 = eps * eps;
    }
  }
  if (stop) return;
  const int i = chunk * BLOCK + threadIdx.x;
  double tg = 0, td = 0, te = 0, tf = 0;
  if (i < a.m.nn) {
    const double inv = 1.0 / scale;
    const int slice = i >> 6, lane = i & 63;
    const int len = a.m.ell_len[slice];
    const int* __restrict__ cp = a.m.ell_col + a.m.ell_off[slice] + lane;
    const float* __restrict__ vp = a.val32 + a.m.ell_off[slice] + lane;
    int c[STAGE];
#pragma unroll
    for (int k = 0; k < STAGE; ++k) c[k] = k < len ? cp[k * 64] : i;
    float w = 0;
#pragma unroll
    for (int k = 0; k < STAGE; ++k) {
      const float v = k < len ? vp[k * 64] : 0.0f;
      const double rk = a.r[c[k]];
      if (k < len) w = fmaf(v, (float)(rk * inv), w);
    }
    for (int k = STAGE; k < len; ++k) w = fmaf(vp[k * 64], (float)(a.r[cp[k * 64]] * inv), w);
    const float ri = (float)(a.r[i] * inv);
    float sc = 0.0f, pc = 0.0f;
    if (cont) {
      // the previous pass stopped at iteration in_iters: its last (s, p) are in
      // the words of that parity (q_cur = parity 1, q_nxt = parity 0 here)
      const float4 last = (in_last & 1) ? a.q_cur[i] : a.q_nxt[i];
      const float cf = (float)cs;
      sc = cf * last.z;
      pc = cf * last.w;
      te = (double)ri * (double)sc;
      tf = (double)pc * (double)sc;
    }
    a.q_nxt[i] = make_float4(ri, w, sc, pc);  // the state of iteration 0
    a.d[i] = 0.0f;
    tg = (double)ri * (double)ri;
    td = (double)w * (double)ri;
  }
  double sg = tg, sd = td;
  block_sum2(sg, sd);
  if (threadIdx.x == 0) {
    a.part_nxt[chunk] = sg;
    a.part_nxt[a.nchunks + chunk] = sd;
  }
  if (cont) {
    double se = te, sf = tf;
    block_sum2(se, sf);
    if (threadIdx.x == 0) {
      a.part_e[chunk] = se;
      a.part_e[a.nchunks + chunk] = sf;
    }
  }
}

// One inner iteration i (Chronopoulos-Gear CG):
//   gamma = r_i.r_i, delta = w_i.r_i (partials of the previous launch);
//   beta = gamma / gamma_{i-1}, alpha = gamma / (delta - beta gamma / alpha_{i-1});
//   at i = 0 of a fresh pass beta = 0, alpha = gamma / delta; at i = 0 of a
//   continued pass beta = gamma / rho_c and p.Ap = delta + 2 beta r.s' +
//   beta^2 p'.s' exactly (the previous pass's direction is not A-conjugate to
//   the replaced residual, so the short form of p.Ap does not hold);
//   p_i = r_i + beta p_{i-1}; s_i = w_i + beta s_{i-1}; d += alpha p_i;
//   r_{i+1} = r_i - alpha s_i; w_{i+1} = A r_{i+1}; partials of r_{i+1}.r_{i+1}
//   and w_{i+1}.r_{i+1}.
__global__ __launch_bounds__(BLOCK, 5) void k_ir_cg(IRArgs a) {
  // Every load that does not depend on another is issued first (state, scalar
  // history, partials, the row and its gathers); the stop test comes after, so
  // one iteration costs about two memory round trips.
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  if (chunk < 0) return;
  IRState* st = a.st;
  // A launch queued past the predicted end of the pass first checks whether the
  // pass has stopped (then it costs a state read instead of a full row load);
  // the others issue their loads before that test.
  if (a.spec && (st->done | st->in_done)) return;
  const int it = a.it;
  const int done = st->done | st->in_done;
  const double rho0_prev = st->in_rho0, tol2 = st->in_tol2;
  const int max_iter = st->max_iter;
  const bool cont0 = it == 0 && st->cont;
  const bool carry = it > 0 || cont0;  // beta != 0: the update uses p and s
  // unconditional loads (see thread_partials), selected after
  const double rho_h = a.rho_hist[it > 0 ? it - 1 : 0], alpha_h = a.alpha_hist[it > 0 ? it - 1 : 0];
  const double rho_prev = it > 0 ? rho_h : st->rho_c;
  const double alpha_prev = it > 0 ? alpha_h : 1.0;
  const int i = chunk * BLOCK + threadIdx.x;
  const bool row = i < a.m.nn;
  // loads in flight together: the previous launch's partials and the row's
  // columns; the partials are added while the gathers are issued
  double va[PART_BATCH], vb[PART_BATCH];
  partial_loads2(a.part_cur, a.part_cur + a.nchunks, a.nchunks, va, vb);
  Row32 R;
  if (row) stage_cols(a.m, i, R);
  double pg, pd;
  partial_sums2(a.part_cur, a.part_cur + a.nchunks, a.nchunks, va, vb, pg, pd);
  float di = 0;
  if (row) {
    stage_gathers(a.m, a.val32, i, a.q_cur, R);
    di = a.d[i];
  }
  if (done) return;
  // the sums are equal in every lane: scalar registers for them and what
  // follows from them leave the vector registers to the staged row
  const double gamma = uniform(block_sum(pg));
  const double delta = uniform(block_sum(pd));
  const double rho0 = it == 0 ? gamma : rho0_prev;
  double beta = 0, alpha = 0, denom = delta;
  if (it > 0) {
    beta = gamma / rho_prev;
    denom = delta - beta * gamma / alpha_prev;
  } else if (cont0) {
    const double rs = reduce(a.part_e, a.nchunks);
    const double ps = reduce(a.part_e + a.nchunks, a.nchunks);
    beta = gamma / rho_prev;
    denom = fma(beta * beta, ps, fma(2.0 * beta, rs, delta));  // same rounding in every build
  }
  const bool conv = gamma == 0 || (it > 0 && gamma < tol2 * rho0);
  const bool stop = conv || it >= max_iter || !(denom > 0);
  if (!stop) alpha = gamma / denom;
  if (chunk == 0 && threadIdx.x == 0) {
    st->in_rho = gamma;
    if (it == 0) st->in_rho0 = rho0;
    a.rho_hist[it] = gamma;
    a.alpha_hist[it] = alpha;
    if (stop) {
      st->in_done = 1;
      st->in_ok = conv && it > 0 ? 1 : 0;
      st->in_iters = it;
      st->total_inner += it;
    }
  }
  if (stop) return;
  const float bf = (float)beta, af = (float)alpha;
  double tg = 0, td = 0;
  if (row) {
    // the row's own word is entry 0 (the precision form's matrix keeps each
    // row's diagonal first)
    const float4 own = make_float4(R.g[0].x, R.g[0].y, R.g[0].z, R.p0);
    const float ri = own.x, wi = own.y;
    const float p = carry ? fmaf(bf, own.w, ri) : ri;
    const float s = carry ? fmaf(bf, own.z, wi) : wi;
    const float rn = fmaf(-af, s, ri);
    a.d[i] = fmaf(af, p, di);
    float w = 0;
#pragma unroll
    for (int k = 0; k < ROW_REGS; ++k) {
      const float sk = carry ? fmaf(bf, R.g[k].z, R.g[k].y) : R.g[k].y;
      w = fmaf(R.v[k], fmaf(-af, sk, R.g[k].x), w);
    }
    if (R.len > ROW_REGS) {
      const int slice = __builtin_amdgcn_readfirstlane(i >> 6), lane = i & 63;  // one slice per wave
      const int* __restrict__ cp = a.m.ell_col + a.m.ell_off[slice] + lane;
      const float* __restrict__ vp = a.val32 + a.m.ell_off[slice] + lane;
      for (int k = ROW_REGS; k < R.len; ++k) {
        const float4 g = a.q_cur[cp[k * 64]];
        const float sk = carry ? fmaf(bf, g.z, g.y) : g.y;
        w = fmaf(vp[k * 64], fmaf(-af, sk, g.x), w);
      }
    }
    a.q_nxt[i] = make_float4(rn, w, s, p);
    tg = (double)rn * (double)rn;
    td = (double)w * (double)rn;
  }
  double sg = tg, sd = td;
  block_sum2(sg, sd);
  if (threadIdx.x == 0) {
    a.part_nxt[chunk] = sg;
    a.part_nxt[a.nchunks + chunk] = sd;
  }
}

// x += |r| d in FP64; workgroup 0 then records what the next pass's start reads
// (the pass count, this pass's scale, its last r.r, whether it may be
// continued) and clears the inner state.  A pass that has not stopped by itself
// ends here after pass_len iterations (the launches queued for it), as if its
// stop test had passed at iteration pass_len.
__global__ __launch_bounds__(BLOCK) void k_ir_update(IRArgs a, int pass_len) {
  const int chunk = xcd_chunk(blockIdx.x, a.nchunks);
  IRState* st = a.st;
  if (chunk < 0 || st->done) return;
  const int i = chunk * BLOCK + threadIdx.x;
  if (i < a.m.nn) a.x[i] = fma(st->scale, (double)a.d[i], a.x[i]);
  if (chunk == 0 && threadIdx.x == 0) {
    const int ended = st->in_done;
    const int last = ended ? st->in_iters : pass_len;
    if (!ended) st->total_inner += pass_len;
    st->outer = st->outer + 1;
    st->cont_ok = (ended ? st->in_ok : 1) && last > 0 ? 1 : 0;
    st->scale_prev = st->scale;
    st->rho_last = last > 0 ? a.rho_hist[last - 1] : 0.0;
    st->in_last = last;
    st->in_done = 0;
    st->in_ok = 0;
    st->in_iters = 0;
  }
}

// Arguments of inner iteration it (k_ir_start fills the parity-0 buffers: it = -1).
IRArgs make_args(const DevMesh& m, const Mesh32& m32, IRBuffers& c, int it) {
  IRArgs a;
  a.m = diag_first(m, m32);
  a.val32 = m32.ell_val;
  a.x = c.x;
  a.b = c.b;
  a.r = c.r;
  a.d = c.d;
  const int cur = it < 0 ? 1 : (it & 1), nxt = cur ^ 1;
  a.q_cur = c.q[cur];
  a.q_nxt = c.q[nxt];
  a.part_bb = c.part_a;
  a.part_rr = c.part_a + c.nchunks;
  a.part_cur = c.part_b + (size_t)2 * c.nchunks * cur;
  a.part_nxt = c.part_b + (size_t)2 * c.nchunks * nxt;
  a.part_e = c.part_e;
  a.rho_hist = c.rho_hist;
  a.alpha_hist = c.alpha_hist;
  a.st = c.st;
  a.it = it;
  a.spec = 0;
  a.nchunks = c.nchunks;
  return a;
}

}  // namespace

hipError_t launch_ir_outer(const DevMesh& m, const Mesh32& m32, IRBuffers& c, bool first, double tol, int max_iter,
                           double in_tol_min, hipStream_t s) {
  if (!c.x || !c.b || !c.r || !c.d || !c.q[0] || !c.q[1] || !c.part_a || !c.part_b || !c.part_e || !c.rho_hist ||
      !c.alpha_hist || !c.st || !m32.ell_val || c.nchunks <= 0)
    return hipErrorInvalidValue;
  if (first) hipLaunchKernelGGL(k_ir_setup, dim3(1), dim3(1), 0, s, c.st, tol * tol, max_iter);
  const IRArgs a = make_args(m, m32, c, -1);
  const dim3 grid(xcd_grid(c.nchunks));
  hipLaunchKernelGGL(k_ir_resid, grid, dim3(BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_ir_start, grid, dim3(BLOCK), 0, s, a, in_tol_min);
  return hipGetLastError();
}

hipError_t launch_ir_inner(const DevMesh& m, const Mesh32& m32, IRBuffers& c, int it, bool spec, hipStream_t s) {
  IRArgs a = make_args(m, m32, c, it);
  a.spec = spec ? 1 : 0;
  hipLaunchKernelGGL(k_ir_cg, dim3(xcd_grid(c.nchunks)), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ir_setup(IRBuffers& c, double tol, int max_iter, hipStream_t s) {
  if (!c.st) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ir_setup, dim3(1), dim3(1), 0, s, c.st, tol * tol, max_iter);
  return hipGetLastError();
}

hipError_t launch_ir_update(const DevMesh& m, IRBuffers& c, int pass_len, hipStream_t s) {
  Mesh32 none{};
  hipLaunchKernelGGL(k_ir_update, dim3(xcd_grid(c.nchunks)), dim3(BLOCK), 0, s, make_args(m, none, c, 0), pass_len);
  return hipGetLastError();
}

}  // namespace mpic
