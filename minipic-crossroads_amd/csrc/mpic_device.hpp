// Device-only helpers shared by the kernel files.
#pragma once

#include "mpic_internal.hpp"

namespace mpic {

// The MI355X dispatches workgroups to its 8 XCDs round-robin (workgroup b runs
// on XCD b mod 8), so a fixed workgroup -> chunk map keeps every chunk on one
// XCD from launch to launch.  This map gives each XCD one contiguous eighth of
// the n row chunks: the matrix rows an XCD reads and nearly all the vector
// entries its rows gather then stay in that XCD's L2 across the iterations of a
// solve (with the plain map each XCD gathers from the whole vector).  Returns -1
// for the padding workgroups of a grid of xcd_grid(n).
__device__ __forceinline__ int xcd_chunk(int b, int n) {
  const int per = (n + NUM_XCD - 1) / NUM_XCD;
  const int c = (b % NUM_XCD) * per + b / NUM_XCD;
  return c < n ? c : -1;
}

// Two sums over a workgroup of T threads: xor butterflies in each wave, then the
// wave totals added in wave order by every thread (the same order everywhere,
// so every thread holds the same sums).  Starts and ends with barriers around
// its LDS words ws[2 T / 64].
template <int T>
__device__ __forceinline__ void wg_sum2(double& a, double& b, double* ws) {
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

// Rows take their first ELL_BATCH entries in one batch (column indices, then values
// and the gathered vector entries, all loads in flight before the sum) and any
// further entries in a loop.  The sum runs over the row's entries in stored
// order, one fma per entry, as a plain loop over the row would.
constexpr int ELL_BATCH = 15;

__device__ __forceinline__ double ell_row64(const DevMesh& m, int row, const double* __restrict__ x) {
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int* __restrict__ cp = m.ell_col + m.ell_off[slice] + lane;
  const double* __restrict__ vp = m.ell_val + m.ell_off[slice] + lane;
  int c[ELL_BATCH];
#pragma unroll
  for (int k = 0; k < ELL_BATCH; ++k) c[k] = k < len ? cp[k * 64] : row;
  double acc = 0;
#pragma unroll
  for (int k = 0; k < ELL_BATCH; ++k) {
    const double v = k < len ? vp[k * 64] : 0.0;
    const double xv = x[c[k]];
    if (k < len) acc += v * xv;
  }
  for (int k = ELL_BATCH; k < len; ++k) acc += vp[k * 64] * x[cp[k * 64]];
  return acc;
}

// Row sums of NR rows at once, x indexed by global node number (it may point
// into LDS, offset by the problem's first node, which the compiler resolves
// after inlining).  Rows with ok[r] false take no entries (their w is 0);
// their row[r] must still be a row of the mesh.
template <int NR>
__device__ __forceinline__ void ell_rows32(const DevMesh& m, const float* __restrict__ val, const int (&row)[NR],
                                           const bool (&ok)[NR], const float* __restrict__ x, float (&w)[NR]) {
  int len[NR];
  int64_t off[NR];
  int c[NR][ELL_BATCH];
  float v[NR][ELL_BATCH];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int slice = row[r] >> 6;
    len[r] = ok[r] ? m.ell_len[slice] : 0;
    off[r] = m.ell_off[slice] + (row[r] & 63);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int k = 0; k < ELL_BATCH; ++k) c[r][k] = k < len[r] ? m.ell_col[off[r] + k * 64] : row[r];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int k = 0; k < ELL_BATCH; ++k) v[r][k] = k < len[r] ? val[off[r] + k * 64] : 0.0f;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    float acc = 0;
#pragma unroll
    for (int k = 0; k < ELL_BATCH; ++k) {
      const float xv = x[c[r][k]];
      if (k < len[r]) acc = fmaf(v[r][k], xv, acc);
    }
    for (int k = ELL_BATCH; k < len[r]; ++k) acc = fmaf(val[off[r] + k * 64], x[m.ell_col[off[r] + k * 64]], acc);
    w[r] = acc;
  }
}

// FP64 row sums of NR rows at once, the operand of column c given by f(c);
// the same batching and entry order as ell_rows32.
template <int NR, class F>
__device__ __forceinline__ void ell_rows64(const DevMesh& m, const int (&row)[NR], const bool (&ok)[NR], F f,
                                           double (&w)[NR]) {
  int len[NR];
  int64_t off[NR];
  int c[NR][ELL_BATCH];
  double v[NR][ELL_BATCH];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int slice = row[r] >> 6;
    len[r] = ok[r] ? m.ell_len[slice] : 0;
    off[r] = m.ell_off[slice] + (row[r] & 63);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int k = 0; k < ELL_BATCH; ++k) c[r][k] = k < len[r] ? m.ell_col[off[r] + k * 64] : row[r];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int k = 0; k < ELL_BATCH; ++k) v[r][k] = k < len[r] ? m.ell_val[off[r] + k * 64] : 0.0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double acc = 0;
#pragma unroll
    for (int k = 0; k < ELL_BATCH; ++k) {
      const double xv = f(c[r][k]);
      if (k < len[r]) acc += v[r][k] * xv;
    }
    for (int k = ELL_BATCH; k < len[r]; ++k) acc += m.ell_val[off[r] + k * 64] * f(m.ell_col[off[r] + k * 64]);
    w[r] = acc;
  }
}

// A thread's share of P chunk partials, part[t] + part[t + BLOCK] + ... in that
// order (t = threadIdx.x), with the first PART_BATCH loads issued together: a
// plain loop waits for each load before the next (one memory round trip per
// BLOCK partials, five at a 65^3-node mesh).  The loads are unconditional (at
// a clamped index, P > 0): a conditional load's register would have to be
// waited for where the paths join.
constexpr int PART_BATCH = 5;
// Element slots of a node read together by the node-charge kernels.
constexpr int NODE_BATCH = 8;
__device__ __forceinline__ double thread_partials(const double* __restrict__ part, int P) {
  double v[PART_BATCH];
#pragma unroll
  for (int u = 0; u < PART_BATCH; ++u) v[u] = part[min((int)threadIdx.x + u * BLOCK, P - 1)];
  double acc = 0;
#pragma unroll
  for (int u = 0; u < PART_BATCH; ++u)
    if (threadIdx.x + u * BLOCK < P) acc += v[u];
  for (int i = threadIdx.x + PART_BATCH * BLOCK; i < P; i += BLOCK) acc += part[i];
  return acc;
}
// The same for two partial arrays read together, in two steps so that other
// loads can be issued in between: partial_loads2 issues the batch,
// partial_sums2 adds it (and any partials past it).
__device__ __forceinline__ void partial_loads2(const double* __restrict__ pa, const double* __restrict__ pb, int P,
                                               double (&va)[PART_BATCH], double (&vb)[PART_BATCH]) {
#pragma unroll
  for (int u = 0; u < PART_BATCH; ++u) {
    const int i = min((int)threadIdx.x + u * BLOCK, P - 1);
    va[u] = pa[i];
    vb[u] = pb[i];
  }
}
__device__ __forceinline__ void partial_sums2(const double* __restrict__ pa, const double* __restrict__ pb, int P,
                                              const double (&va)[PART_BATCH], const double (&vb)[PART_BATCH],
                                              double& a, double& b) {
  a = 0;
  b = 0;
#pragma unroll
  for (int u = 0; u < PART_BATCH; ++u)
    if (threadIdx.x + u * BLOCK < P) {
      a += va[u];
      b += vb[u];
    }
  for (int i = threadIdx.x + PART_BATCH * BLOCK; i < P; i += BLOCK) {
    a += pa[i];
    b += pb[i];
  }
}

// 12 / 16 consecutive doubles of element e (records are 96 / 128 B, 16-B aligned).
__device__ __forceinline__ void load12(const double* __restrict__ base, int e, double o[12]) {
  const double2* q = reinterpret_cast<const double2*>(base + (size_t)e * 12);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double2 t = q[i];
    o[2 * i] = t.x;
    o[2 * i + 1] = t.y;
  }
}
__device__ __forceinline__ void load16(const double* __restrict__ base, int e, double o[16]) {
  const double2* q = reinterpret_cast<const double2*>(base + (size_t)e * 16);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double2 t = q[i];
    o[2 * i] = t.x;
    o[2 * i + 1] = t.y;
  }
}

// ---- wavefront helpers for records exchanged through LDS ----
// Orders this wavefront's LDS accesses across lanes (a wavefront executes its
// LDS instructions in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// This lane's index, opaque to the optimiser, so that lane-dependent LDS
// offsets are computed where they are used instead of occupying registers
// across the particle loop.
__device__ __forceinline__ int fresh_lane() {
  int l = __lane_id();
  asm volatile("" : "+v"(l));
  return l;
}

// Particle records as 16-B words.
__device__ __forceinline__ Rec64 load_rec(const Rec64* __restrict__ base, int64_t i) {
  const double2* q = reinterpret_cast<const double2*>(base + i);
  const double2 a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3], a4 = q[4];
  const int4 a5 = reinterpret_cast<const int4*>(base + i)[5];
  Rec64 r;
  r.x[0] = a0.x;
  r.x[1] = a0.y;
  r.x[2] = a1.x;
  r.xr[0] = a1.y;
  r.xr[1] = a2.x;
  r.xr[2] = a2.y;
  r.v[0] = a3.x;
  r.v[1] = a3.y;
  r.v[2] = a4.x;
  r.w = a4.y;
  r.elem = a5.x;
  r.id = (uint32_t)a5.y;
  r.type = (uint32_t)a5.z;
  r.pad = 0;
  return r;
}
__device__ __forceinline__ void store_rec(Rec64* __restrict__ base, int64_t i, const Rec64& r) {
  double2* q = reinterpret_cast<double2*>(base + i);
  q[0] = double2{r.x[0], r.x[1]};
  q[1] = double2{r.x[2], r.xr[0]};
  q[2] = double2{r.xr[1], r.xr[2]};
  q[3] = double2{r.v[0], r.v[1]};
  q[4] = double2{r.v[2], r.w};
  reinterpret_cast<int4*>(base + i)[5] = int4{r.elem, (int)r.id, (int)r.type, 0};
}
__device__ __forceinline__ Rec32 load_rec(const Rec32* __restrict__ base, int64_t i) {
  const float4* q = reinterpret_cast<const float4*>(base + i);
  const float4 a0 = q[0], a1 = q[1];
  const int4 a2 = reinterpret_cast<const int4*>(base + i)[2];
  Rec32 r;
  r.x[0] = a0.x;
  r.x[1] = a0.y;
  r.x[2] = a0.z;
  r.elem = __float_as_int(a0.w);
  r.v[0] = a1.x;
  r.v[1] = a1.y;
  r.v[2] = a1.z;
  r.type = (uint32_t)__float_as_int(a1.w);
  r.w = __hiloint2double(a2.y, a2.x);
  r.id = (uint32_t)a2.z;
  r.pad = 0;
  return r;
}
__device__ __forceinline__ void store_rec(Rec32* __restrict__ base, int64_t i, const Rec32& r) {
  float4* q = reinterpret_cast<float4*>(base + i);
  q[0] = float4{r.x[0], r.x[1], r.x[2], __int_as_float(r.elem)};
  q[1] = float4{r.v[0], r.v[1], r.v[2], __int_as_float((int)r.type)};
  reinterpret_cast<int4*>(base + i)[2] = int4{__double2loint(r.w), __double2hiint(r.w), (int)r.id, 0};
}

// physToReferenceVector, mesh.hpp:245-250: out_r = ((0 + J_r0 v0) + J_r1 v1) + J_r2 v2.
__device__ __forceinline__ void p2r_vec(const double* Ji, const double v[3], double o[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double acc = 0.0;
    acc += Ji[r * 3 + 0] * v[0];
    acc += Ji[r * 3 + 1] * v[1];
    acc += Ji[r * 3 + 2] * v[2];
    o[r] = acc;
  }
}

// 128-bit atomic add: 64-bit add on the low word, carry folded into the high word.
// Each low-word add sees an exact old value, so the carries sum exactly and the
// final value is exact in any order.
__device__ __forceinline__ void atomic_add_fx(fx128* dst, fx128 v) {
  if (v.lo == 0ull && v.hi == 0ll) return;
  const unsigned long long old =
      __hip_atomic_fetch_add(&dst->lo, v.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long carry = (old + v.lo) < old ? 1ull : 0ull;
  const unsigned long long h = (unsigned long long)v.hi + carry;
  if (h != 0ull)
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(&dst->hi), h, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Groups the lanes of a wavefront by key.  Returns the lane index of the
// group's leader (its lowest lane) for lanes with key >= 0, -1 otherwise; the
// loop runs once per distinct key in the wavefront.
__device__ __forceinline__ int wave_group_leader(int key, unsigned long long* group_mask) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(key >= 0);
  int leader = -1;
  unsigned long long mine = 0;
  while (pending) {
    const int l = __ffsll((long long)pending) - 1;
    const int k = __shfl(key, l);
    const unsigned long long grp = __ballot(key == k) & pending;
    if ((grp >> lane) & 1ull) {
      leader = l;
      mine = grp;
    }
    pending &= ~grp;
  }
  *group_mask = mine;
  return leader;
}

// 128-bit add into LDS (64-bit LDS atomics, carry folded into the high word).
__device__ __forceinline__ void lds_add_fx(fx128* dst, fx128 v) {
  const unsigned long long old = atomicAdd(&dst->lo, v.lo);
  const unsigned long long carry = (old + v.lo) < old ? 1ull : 0ull;
  const unsigned long long h = (unsigned long long)v.hi + carry;
  if (h != 0ull) atomicAdd(reinterpret_cast<unsigned long long*>(&dst->hi), h);
}

// Kinetic energy term of one particle, PIC.cpp:191-200: 0.5 (w m) |v|^2.
__device__ __forceinline__ double ke_term(double w, double mass, const double V[3]) {
  const double tm = w * mass;
  double vs = 0;
  vs += V[0] * V[0];
  vs += V[1] * V[1];
  vs += V[2] * V[2];
  return 0.5 * tm * vs;
}

// Walk counters of one thread.
struct WalkCount {
  int crossings, capped, nonpos, reflections, relocated;
};

// Workgroup reduction of the push counters and the KE sum into partial[blockIdx.x]
// (NT threads per workgroup).
template <int NT = BLOCK>
__device__ __forceinline__ void block_partial(PushPartial* out, fx128 ke, unsigned long long kmax, const WalkCount& wc,
                                              long long killed) {
  __shared__ PushPartial sh[NT / 64];
  unsigned long long klo = ke.lo;
  long long khi = ke.hi;
  long long c0 = wc.crossings, c1 = killed, c2 = wc.capped, c3 = wc.nonpos, c4 = wc.reflections, c5 = wc.relocated;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    unsigned long long olo = __shfl_xor(klo, off);
    long long ohi = __shfl_xor(khi, off);
    unsigned long long nlo = klo + olo;
    khi = khi + ohi + (nlo < klo ? 1 : 0);
    klo = nlo;
    unsigned long long om = __shfl_xor(kmax, off);
    kmax = om > kmax ? om : kmax;
    c0 += __shfl_xor(c0, off);
    c1 += __shfl_xor(c1, off);
    c2 += __shfl_xor(c2, off);
    c3 += __shfl_xor(c3, off);
    c4 += __shfl_xor(c4, off);
    c5 += __shfl_xor(c5, off);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wave] = PushPartial{klo, khi, kmax, c0, c1, c2, c3, c4, c5};
  __syncthreads();
  if (threadIdx.x == 0) {
    PushPartial r = sh[0];
    for (int w = 1; w < NT / 64; ++w) {
      fx128 t = fx_add(fx128{r.ke_lo, r.ke_hi}, fx128{sh[w].ke_lo, sh[w].ke_hi});
      r.ke_lo = t.lo;
      r.ke_hi = t.hi;
      r.ke_max_bits = sh[w].ke_max_bits > r.ke_max_bits ? sh[w].ke_max_bits : r.ke_max_bits;
      r.crossings += sh[w].crossings;
      r.killed += sh[w].killed;
      r.capped += sh[w].capped;
      r.nonpos += sh[w].nonpos;
      r.reflections += sh[w].reflections;
      r.relocated += sh[w].relocated;
    }
    out[blockIdx.x] = r;
  }
}

// Each workgroup handles one contiguous range of particles in chunks of BLOCK
// (a wavefront covers 64 consecutive particles); element charges of the new
// positions are accumulated in an LDS table and flushed every FLUSH_CHUNKS.
constexpr int FLUSH_CHUNKS = 4;

template <int NT = BLOCK>
__device__ __forceinline__ void block_range(int64_t n, int64_t* lo, int64_t* hi) {
  const int64_t chunks = (n + NT - 1) / NT;
  const int64_t per = (chunks + gridDim.x - 1) / gridDim.x;
  *lo = (int64_t)blockIdx.x * per * NT;
  const int64_t h = *lo + per * NT;
  *hi = h < n ? h : n;
}

// ---------------------------------------------------------------------------
// Block-level charge accumulation.
//
// Element charges are accumulated exactly (128-bit) in an LDS hash table keyed
// by element, over a contiguous run of particles, and flushed to the global
// accumulator with non-returning atomics.  The global accumulator holds each
// (element, node) value as four 64-bit words of 32-bit digits: bits 0-31,
// 32-63, 64-95 of the 128-bit two's-complement value (unsigned) and bits 96-127
// (signed).  A word receives at most one digit per flush and there are fewer
// flushes per slot than particles (< 2^31), so no word can overflow and no carry
// is needed; the exact value is d0 + d1*2^32 + d2*2^64 + d3*2^96 (mod 2^128).
// The 16 words of an element (4 nodes x 4 digits) are one 128-byte line, and a
// flush writes them with 16 consecutive lanes.
// ---------------------------------------------------------------------------
constexpr int DEP_TABLE = 512;  // LDS hash entries per workgroup (power of two)
constexpr int DEP_EMPTY = -1;

struct DepTable {
  int key[DEP_TABLE];
  int cnt[DEP_TABLE];       // particles added to the entry (element counts for the sort)
  fx128 val[DEP_TABLE][4];
};

template <int NT = BLOCK>
__device__ __forceinline__ void dep_table_clear(DepTable& t) {
  for (int k = threadIdx.x; k < DEP_TABLE; k += NT) {
    t.key[k] = DEP_EMPTY;
    t.cnt[k] = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) t.val[k][i] = fx128{0, 0};
  }
}

// Digit d (0..3) of a 128-bit value as a 64-bit word (see above).
__device__ __forceinline__ unsigned long long fx_digit(fx128 v, int d) {
  switch (d) {
    case 0: return v.lo & 0xffffffffull;
    case 1: return v.lo >> 32;
    case 2: return (unsigned long long)v.hi & 0xffffffffull;
    default: return (unsigned long long)(v.hi >> 32);  // arithmetic shift: signed top digit
  }
}

// Direct global add of one particle's 4 node charges and count (table overflow
// path).  Counts are kept per ckey[element] (the sort bucket).
__device__ __forceinline__ void dep_global_add(unsigned long long* __restrict__ acc, int* __restrict__ count,
                                               const int32_t* __restrict__ ckey, int key, const fx128 c[4]) {
  if (count) atomicAdd(count + ckey[key], 1);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const unsigned long long w = fx_digit(c[i], d);
      if (w) __hip_atomic_fetch_add(acc + (size_t)key * 16 + i * 4 + d, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Adds one particle's node charges to the table entry of element key (key < 0:
// nothing) and counts it.  Uses LDS atomics; falls back to global atomics if the
// table is full.
__device__ __forceinline__ void dep_table_add(DepTable& t, unsigned long long* __restrict__ acc,
                                              int* __restrict__ count, const int32_t* __restrict__ ckey, int key,
                                              const fx128 c[4]) {
  if (key < 0) return;
  unsigned h = ((unsigned)key * 2654435761u) >> (32 - 9);  // 9 = log2(DEP_TABLE)
  for (int probe = 0; probe < DEP_TABLE; ++probe) {
    int k = __hip_atomic_load(&t.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == DEP_EMPTY) {
      int expected = DEP_EMPTY;
      __hip_atomic_compare_exchange_strong(&t.key[h], &expected, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
      k = expected == DEP_EMPTY ? key : expected;
    }
    if (k == key) {
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_add_fx(&t.val[h][i], c[i]);
      atomicAdd(&t.cnt[h], 1);
      return;
    }
    h = (h + 1) & (DEP_TABLE - 1);
  }
  dep_global_add(acc, count, ckey, key, c);
}

// Flushes every used entry with coalesced non-returning atomics (charges; counts
// when count != null) and clears the table.  Must be called by the whole workgroup
// of NT threads.  The counts go first, each thread issuing all its bucket loads
// before any atomic: a wavefront's vector memory operations complete in order, so
// a load issued after the charge atomics would wait for their round trips (one
// wait per 16 entries in the charge loop).
template <int NT = BLOCK>
__device__ __forceinline__ void dep_table_flush(DepTable& t, unsigned long long* __restrict__ acc,
                                                int* __restrict__ count, const int32_t* __restrict__ ckey) {
  __syncthreads();
  if (count) {
    constexpr int PER = (DEP_TABLE + NT - 1) / NT;
    int key[PER], bkt[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int k = threadIdx.x + u * NT;
      key[u] = k < DEP_TABLE ? t.key[k] : DEP_EMPTY;
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) bkt[u] = key[u] != DEP_EMPTY ? ckey[key[u]] : 0;
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if (key[u] != DEP_EMPTY) atomicAdd(count + bkt[u], t.cnt[threadIdx.x + u * NT]);
  }
  const int sub = threadIdx.x & 15;  // node = sub/4, digit = sub%4
  for (int e = threadIdx.x >> 4; e < DEP_TABLE; e += NT / 16) {
    const int key = t.key[e];
    if (key != DEP_EMPTY) {
      const unsigned long long w = fx_digit(t.val[e][sub >> 2], sub & 3);
      if (w) __hip_atomic_fetch_add(acc + (size_t)key * 16 + sub, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  dep_table_clear<NT>(t);
  __syncthreads();
}

// Exact 128-bit value of one accumulator slot (4 digit words).
__device__ __forceinline__ fx128 dep_slot_value(const unsigned long long* w) {
  // d0 + d1*2^32 + d2*2^64 + d3*2^96 in 128-bit two's complement arithmetic
  const unsigned long long lo0 = w[0], mid = w[1], hi2 = w[2], top = w[3];
  fx128 r{lo0, 0};
  r = fx_add(r, fx128{mid << 32, (long long)(mid >> 32)});
  r = fx_add(r, fx128{0ull, (long long)hi2});
  r = fx_add(r, fx128{0ull, (long long)(top << 32)});
  return r;
}

}  // namespace mpic
