// Precision form of the particle phases of PIC::time_step (PIC.cpp:245-267):
//   field  = the P1 field of the FP64 potential, constant on each tet
//            (E_e = sum_i phi_i G[e][i]; the reference's gather, PIC.cpp:168-183,
//            weights the same products by barycentric weights that sum to 1),
//            computed in FP64 once per step and stored as float4;
//   push   = Boris kick (B = 0) and barycentric tet walk in FP32
//            (particle_move.hpp:136-225,364-392), then an FP64 point-location test
//            for every particle that ends within LOC_EPS of a face of its tet,
//            then the charge deposit of the new position and the KE;
//   deposit = one 64-bit fixed-point integer per element: a particle adds
//            q w (at scale 2^S) to its element, and each node of the element
//            later receives a quarter of the element sum (every node of a tet gets
//            -q w/4 in the reference, PIC.cpp:121-131 with
//            sum_c wbv(e,i,c) = detJ/24).  Integer sums are exact, so node
//            charges do not depend on the order of the atomics.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

// Below this minimum barycentric coordinate the FP64 test decides the owner.
// The FP32 position carries ~3e-8 absolute error, i.e. a few 1e-6 in reference
// coordinates of a 1/64 tet, so the band is set well above that.
constexpr float LOC_EPS = 2e-4f;
// Threads per k_push32 workgroup: eight wavefronts take adjacent batches of one range, so
// the element records their walks load are shared in the CU's L1 (per launch at 100M
// particles: 256 threads 4.09 ms, 384 4.49, 512 4.00, 768 4.15, 1024 5.13).
constexpr int PUSH_BLOCK = 512;
constexpr int LOC_CAP = 16;        // element moves of one point location
constexpr double OWNER_TOL = 1e-9; // ownership check: tolerated negative coordinate

// The walk's FP32 element record (64 B): inverse Jacobian, node 0 and the face
// neighbours, so a crossing needs one load.
struct ElemRec32 {
  float J[9];
  float n0[3];
  int nb[4];
};

__device__ __forceinline__ ElemRec32 load_erec32(const float* __restrict__ erec, int e) {
  const float4* q = reinterpret_cast<const float4*>(erec + (size_t)e * 16);
  const float4 a0 = q[0], a1 = q[1], a2 = q[2];
  const int4 a3 = reinterpret_cast<const int4*>(erec + (size_t)e * 16)[3];
  ElemRec32 r;
  r.J[0] = a0.x;
  r.J[1] = a0.y;
  r.J[2] = a0.z;
  r.J[3] = a0.w;
  r.J[4] = a1.x;
  r.J[5] = a1.y;
  r.J[6] = a1.z;
  r.J[7] = a1.w;
  r.J[8] = a2.x;
  r.n0[0] = a2.y;
  r.n0[1] = a2.z;
  r.n0[2] = a2.w;
  r.nb[0] = a3.x;
  r.nb[1] = a3.y;
  r.nb[2] = a3.z;
  r.nb[3] = a3.w;
  return r;
}

// FP64 point location of the FP32 position x (widened exactly), starting in e:
// moves across the face of the most negative barycentric coordinate until none
// is negative, a wall is reached, the walk would step back, or LOC_CAP moves.
// Returns the final element, its reference position and the number of moves.
// Face of each coordinate: lambda_0 = 1 - sum R -> face 1, R0 -> face 2,
// R1 -> face 0, R2 -> face 3 (faces f0 = (0,1,3), f1 = (1,2,3), f2 = (0,3,2),
// f3 = (0,2,1)).  Arguments and result travel in registers: the cold FP64 path
// stays out of line without putting the hot loop's state in scratch memory.
struct Located {
  float r0, r1, r2;
  int e, moves;
};

__device__ __noinline__ Located locate64(const double* __restrict__ geo, const int4* __restrict__ nbrs, int e,
                                         float X0, float X1, float X2) {
  const double x[3] = {(double)X0, (double)X1, (double)X2};
  double r[3];
  int prev = -1, moves = 0;
  for (int k = 0;; ++k) {
    double g[12];
    load12(geo, e, g);
    const double dl[3] = {x[0] - g[9], x[1] - g[10], x[2] - g[11]};
    p2r_vec(g, dl, r);
    const double lam[4] = {1.0 - r[0] - r[1] - r[2], r[0], r[1], r[2]};
    int j = 0;
#pragma unroll
    for (int i = 1; i < 4; ++i)
      if (lam[i] < lam[j]) j = i;
    if (!(lam[j] < 0.0) || k == LOC_CAP) break;
    const int4 nb = nbrs[e];
    const int e2 = j == 0 ? nb.y : j == 1 ? nb.z : j == 2 ? nb.x : nb.w;
    if (e2 < 0 || e2 == prev) break;
    prev = e;
    e = e2;
    moves++;
  }
  return Located{(float)r[0], (float)r[1], (float)r[2], e, moves};
}

#pragma clang fp contract(fast)

__device__ __forceinline__ void p2r_vecf(const float* Ji, const float v[3], float o[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) o[r] = Ji[r * 3 + 0] * v[0] + Ji[r * 3 + 1] * v[1] + Ji[r * 3 + 2] * v[2];
}

__device__ __forceinline__ bool near_face(const float R[3]) {
  const float l0 = 1.0f - R[0] - R[1] - R[2];
  return fminf(fminf(l0, R[0]), fminf(R[1], R[2])) < LOC_EPS;
}

// Rare events of the push, counted in LDS where they happen so that the walk
// keeps a single counter (crossings) in registers.
struct RareCount {
  int capped, nonpos, reflections, relocated, killed;
};

// The tet walk of MoveParticle (particle_move.hpp:103-124,136-225) in FP32,
// over the remaining time tau = dt * time_remaining instead of the fraction.
// The exit time through face f is a_f / b_f with b_f > 0:
//   zeta (face 3): R2 / -VR2;  eta (face 0): R1 / -VR1;  xi (face 2): R0 / -VR0;
//   sum (face 1): (1 - R0 - R1 - R2) / (VR0 + VR1 + VR2).
// Starting from the candidate tau / 1 (no face), a face replaces the current
// candidate A / B when b_f > 0 and a_f B < A b_f, tested in the reference's
// order (zeta, eta, xi, sum).  Against tau this is the reference's "the end
// point lies beyond the face" test, between faces its strictly-smaller
// replacement; the products avoid a division per face, and the chosen time is
// A * rcp(B) (1 ulp).  The fraction s of the reference is t / tau.  The
// neighbour across the chosen face is selected with the face (a select per
// face keeps the element record in registers).  At WALK_CAP face hits the walk
// stops (counted as capped).
// g holds the record of element e on entry and of the final element on return.
__device__ __forceinline__ bool walk32(const DevMesh& m, const float* __restrict__ erec, int& e, ElemRec32& g,
                                       float X[3], float R[3], float V[3], float VR[3], float dt, int kill,
                                       int& crossings, RareCount& rare) {
  float tau = dt;
  for (int hits = 0;; ++hits) {
    if (hits >= WALK_CAP) {
      atomicAdd(&rare.capped, 1);
      break;
    }
    float A = tau, B = 1.0f;
    int face = -1, nbs = -1;  // the face taken and the neighbour across it
    {
      const float a = R[2], b = -VR[2];
      const bool take = b > 0.0f && a * B < A * b;
      A = take ? a : A;
      B = take ? b : B;
      face = take ? 3 : face;
      nbs = take ? g.nb[3] : nbs;
    }
    {
      const float a = R[1], b = -VR[1];
      const bool take = b > 0.0f && a * B < A * b;
      A = take ? a : A;
      B = take ? b : B;
      face = take ? 0 : face;
      nbs = take ? g.nb[0] : nbs;
    }
    {
      const float a = R[0], b = -VR[0];
      const bool take = b > 0.0f && a * B < A * b;
      A = take ? a : A;
      B = take ? b : B;
      face = take ? 2 : face;
      nbs = take ? g.nb[2] : nbs;
    }
    {
      const float a = 1.0f - R[0] - R[1] - R[2], b = VR[0] + VR[1] + VR[2];
      const bool take = b > 0.0f && a * B < A * b;
      A = take ? a : A;
      B = take ? b : B;
      face = take ? 1 : face;
      nbs = take ? g.nb[1] : nbs;
    }
    if (face < 0) {
      // no face before the end of the step
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        X[d] += tau * V[d];
        R[d] += tau * VR[d];
      }
      break;
    }
    const float t = A * __builtin_amdgcn_rcpf(B);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      X[d] += t * V[d];
      R[d] += t * VR[d];
    }
    if (nbs >= 0) {
      e = nbs;
      crossings++;
      g = load_erec32(erec, e);
      const float dl[3] = {X[0] - g.n0[0], X[1] - g.n0[1], X[2] - g.n0[2]};
      p2r_vecf(g.J, dl, R);
    } else {
      if (kill) return false;
      atomicAdd(&rare.reflections, 1);
      const double* nrm = m.normals + (size_t)e * 12 + face * 3;
      const float n0 = (float)nrm[0], n1 = (float)nrm[1], n2 = (float)nrm[2];
      const float vn = V[0] * n0 + V[1] * n1 + V[2] * n2;
      V[0] -= 2 * vn * n0;
      V[1] -= 2 * vn * n1;
      V[2] -= 2 * vn * n2;
    }
    p2r_vecf(g.J, V, VR);
    if (!(t > 0.0f)) atomicAdd(&rare.nonpos, 1);
    tau -= t;
    if (!(tau > 0.0f)) break;
  }
  return true;
}

// ---- element charge table (LDS), 64-bit fixed point --------------------------
// N entries (a power of two) keyed by element, linear probing; a key that finds
// neither itself nor a free entry within 64 probes goes straight to the global
// accumulator (the table is flushed long before it fills).
template <int N>
struct DepTable64 {
  static_assert((N & (N - 1)) == 0, "power of two");
  int key[N];
  int cnt[N];
  unsigned long long val[N];
};

template <int N>
__device__ __forceinline__ void dep64_clear(DepTable64<N>& t) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    t.key[k] = DEP_EMPTY;
    t.cnt[k] = 0;
    t.val[k] = 0;
  }
}

template <int N>
__device__ __forceinline__ void dep64_add(DepTable64<N>& t, unsigned long long* __restrict__ acc,
                                          int* __restrict__ count, const int32_t* __restrict__ ckey, int key,
                                          long long c) {
  if (key < 0) return;
  unsigned h = ((unsigned)key * 2654435761u) >> (32 - __builtin_ctz(N));
  for (int probe = 0; probe < (N < 64 ? N : 64); ++probe) {
    int k = __hip_atomic_load(&t.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == DEP_EMPTY) {
      int expected = DEP_EMPTY;
      __hip_atomic_compare_exchange_strong(&t.key[h], &expected, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
      k = expected == DEP_EMPTY ? key : expected;
    }
    if (k == key) {
      atomicAdd(&t.val[h], (unsigned long long)c);
      atomicAdd(&t.cnt[h], 1);
      return;
    }
    h = (h + 1) & (N - 1);
  }
  __hip_atomic_fetch_add(acc + key, (unsigned long long)c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (count) atomicAdd(count + ckey[key], 1);
}

template <int N>
__device__ __forceinline__ void dep64_flush(DepTable64<N>& t, unsigned long long* __restrict__ acc,
                                            int* __restrict__ count, const int32_t* __restrict__ ckey) {
  __syncthreads();
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const int key = t.key[k];
    if (key != DEP_EMPTY) {
      // the bucket load goes before the charge atomic: a wavefront's vector
      // memory operations complete in order, so a load after it would wait for it
      const int bkt = count ? ckey[key] : 0;
      const unsigned long long v = t.val[k];
      if (v) __hip_atomic_fetch_add(acc + key, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (count) atomicAdd(count + bkt, t.cnt[k]);
    }
  }
  __syncthreads();
  dep64_clear(t);
  __syncthreads();
}
constexpr int DEP64 = 1024;       // k_deposit32: a flush covers FLUSH_CHUNKS * BLOCK = 1024 particles
constexpr int DEP64_PUSH = 512;   // k_push32 (its LDS also holds the wavefront exchange areas)

// ---- particle records through LDS ---------------------------------------------
// The vector L1 address path costs about 16 cycles per wavefront load or store
// of 16 B per lane, plus more for every further 64-byte line that a group of
// four lanes touches.  A wavefront's 64 consecutive 48-byte records are
// 3 KB: read or written one record per lane they touch three lines per lane
// group in each of three instructions; as contiguous 16-byte pieces (lane l
// takes pieces l, 64 + l, 128 + l) they touch one.  The pieces pass to and
// from the lanes that own the records through a per-wavefront LDS area.
constexpr int REC_PIECES = 3 * 64;

// Records s0 .. s0 + nv - 1 to lanes 0 .. nv - 1 (others: elem = -1).
__device__ __forceinline__ Rec32 wave_load_recs(float4* __restrict__ sl, const Rec32* __restrict__ P, int64_t s0,
                                                int nv) {
  const int lane = fresh_lane(), n3 = 3 * nv;
  const float4* in = reinterpret_cast<const float4*>(P + s0);
  float4 v0, v1, v2;
  if (lane < n3) v0 = in[lane];
  if (64 + lane < n3) v1 = in[64 + lane];
  if (128 + lane < n3) v2 = in[128 + lane];
  if (lane < n3) sl[lane] = v0;
  if (64 + lane < n3) sl[64 + lane] = v1;
  if (128 + lane < n3) sl[128 + lane] = v2;
  wave_sync();
  Rec32 r{};
  r.elem = -1;
  if (lane < nv) {
    const float4 a0 = sl[3 * lane], a1 = sl[3 * lane + 1];
    const int4 a2 = reinterpret_cast<const int4*>(sl)[3 * lane + 2];
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
  }
  wave_sync();
  return r;
}

// Lanes 0 .. nv - 1 receive records src (their own slot each, on steps that
// read through the ranking); the wavefront reads them as pieces j = 64 k + lane
// (part j % 3 of the record of lane j / 3), so that a lane group touches two
// records instead of four.
__device__ __forceinline__ Rec32 wave_gather_recs(float4* __restrict__ sl, int* __restrict__ key,
                                                  const Rec32* __restrict__ P, int src, int nv) {
  const int lane = fresh_lane(), n3 = 3 * nv;
  if (lane < nv) key[lane] = src;
  wave_sync();
  auto piece = [&](int j) -> float4 {
    const int r = j / 3;
    return reinterpret_cast<const float4*>(P + key[r])[j - 3 * r];
  };
  float4 v0, v1, v2;
  if (lane < n3) v0 = piece(lane);
  if (64 + lane < n3) v1 = piece(64 + lane);
  if (128 + lane < n3) v2 = piece(128 + lane);
  if (lane < n3) sl[lane] = v0;
  if (64 + lane < n3) sl[64 + lane] = v1;
  if (128 + lane < n3) sl[128 + lane] = v2;
  wave_sync();
  Rec32 r{};
  r.elem = -1;
  if (lane < nv) {
    const float4 a0 = sl[3 * lane], a1 = sl[3 * lane + 1];
    const int4 a2 = reinterpret_cast<const int4*>(sl)[3 * lane + 2];
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
  }
  wave_sync();
  return r;
}

// Lane l's record (l < nv) to slot s0 + l: whole records, or only the position
// and velocity words (the weight/id word unchanged).
__device__ __forceinline__ void wave_store_recs(float4* __restrict__ sl, Rec32* __restrict__ Q, int64_t s0, int nv,
                                                bool whole, const Rec32& r) {
  const int lane = fresh_lane(), n3 = 3 * nv;
  if (lane < nv) {
    sl[3 * lane] = float4{r.x[0], r.x[1], r.x[2], __int_as_float(r.elem)};
    sl[3 * lane + 1] = float4{r.v[0], r.v[1], r.v[2], __int_as_float((int)r.type)};
    reinterpret_cast<int4*>(sl)[3 * lane + 2] = int4{__double2loint(r.w), __double2hiint(r.w), (int)r.id, 0};
  }
  wave_sync();
  float4* out = reinterpret_cast<float4*>(Q + s0);
  // piece 64 k + lane is part (lane + k) mod 3 of its record; part 2 is the weight/id word
  const int part = lane % 3;
  const float4 w0 = sl[lane], w1 = sl[64 + lane], w2 = sl[128 + lane];
  if (lane < n3 && (whole || part != 2)) out[lane] = w0;
  if (64 + lane < n3 && (whole || part != 1)) out[64 + lane] = w1;
  if (128 + lane < n3 && (whole || part != 0)) out[128 + lane] = w2;
  wave_sync();
}

// ---- kernels -----------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_field32(DevMesh m, const double* __restrict__ phi, float4* __restrict__ E) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= m.ne) return;
  const int4 nd = m.conn[e];
  const double f[4] = {phi[nd.x], phi[nd.y], phi[nd.z], phi[nd.w]};
  double G[12];
  load12(m.grad, e, G);
  double E0 = 0, E1 = 0, E2 = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    E0 += f[i] * G[i * 3 + 0];
    E1 += f[i] * G[i * 3 + 1];
    E2 += f[i] * G[i * 3 + 2];
  }
  E[e] = make_float4((float)E0, (float)E1, (float)E2, 0.0f);
}

__global__ __launch_bounds__(PUSH_BLOCK, 6) void k_push32(PushArgs32 a) {
  __shared__ DepTable64<DEP64_PUSH> table;
  __shared__ float4 recx[PUSH_BLOCK / 64][REC_PIECES];
  __shared__ int recsrc[PUSH_BLOCK / 64][64];
  __shared__ RareCount rare;
  __shared__ double sp[48];
  float4* sl = recx[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
  int* skey = recsrc[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
  dep64_clear(table);
  if (threadIdx.x < 48) sp[threadIdx.x] = a.species[threadIdx.x];
  if (threadIdx.x == 0) rare = RareCount{0, 0, 0, 0, 0};
  const bool gather = a.perm != nullptr;
  int crossings = 0;
  fx128 ke{0, 0};
  unsigned kmax_hi = 0;  // high word of the largest KE term (terms are >= 0)
  int64_t lo, hi;
  block_range<PUSH_BLOCK>(a.n, &lo, &hi);
  __syncthreads();
  int chunk = 0;
  for (int64_t base = lo; base < hi; base += PUSH_BLOCK, ++chunk) {
    const int64_t p = base + threadIdx.x;  // output position
    const int64_t s0 = base + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 64;  // the wavefront's first
    const int64_t rem = hi - s0;
    const int nv = rem <= 0 ? 0 : rem >= 64 ? 64 : (int)rem;
    int enew = -1;
    double w = 0;
    float V[3] = {0, 0, 0};
    int t = 0;
    Rec32 rc{};
    rc.elem = -1;
    if (nv > 0) {
      if (gather)
        rc = wave_gather_recs(sl, skey, a.P.r, p < hi ? a.perm[p] : 0, nv);
      else
        rc = wave_load_recs(sl, a.P.r, s0, nv);
    }
    if (p < hi) {
      int e = rc.elem;
      if (e >= 0) {
        float X[3], R[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          X[d] = rc.x[d];
          V[d] = rc.v[d];
        }
        t = (int)rc.type;
        w = rc.w;
        const float4 E = a.m32.E[e];
        ElemRec32 g = load_erec32(a.m32.erec, e);
        {
          // reference position of X in e (the record keeps no copy)
          const float dl[3] = {X[0] - g.n0[0], X[1] - g.n0[1], X[2] - g.n0[2]};
          p2r_vecf(g.J, dl, R);
        }
        // Boris push with B = 0 (two half kicks), particle_move.hpp:364-392
        const float hq = a.dt * a.scale * (float)sp[32 + t] * 0.5f;
        const float Ev[3] = {E.x, E.y, E.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          V[d] += hq * Ev[d];
          V[d] += hq * Ev[d];
        }
        float VR[3];
        p2r_vecf(g.J, V, VR);
        const bool alive = walk32(a.m, a.m32.erec, e, g, X, R, V, VR, a.dt, a.kill, crossings, rare);
        if (alive && near_face(R)) {
          const Located L = locate64(a.m.geo, a.m.nbrs, e, X[0], X[1], X[2]);
          e = L.e;
          R[0] = L.r0;
          R[1] = L.r1;
          R[2] = L.r2;
          if (L.moves) atomicAdd(&rare.relocated, L.moves);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          rc.x[d] = X[d];
          rc.v[d] = V[d];
        }
        enew = alive ? e : -1;
        if (!alive) atomicAdd(&rare.killed, 1);
        rc.elem = enew;
      }
    }
    // records back (a dead particle's as read: the ranking drops it), then the element mirror
    if (nv > 0) wave_store_recs(sl, a.Q.r, s0, nv, gather, rc);
    if (p < hi) a.Q.elem[p] = rc.elem < 0 ? -1 : rc.elem;
    if (enew >= 0) {
      dep64_add(table, a.acc, a.count, a.m.bucket, enew, to_fixed64(w * sp[t], a.S_charge));
      const double Vd[3] = {(double)V[0], (double)V[1], (double)V[2]};
      const double term = ke_term(w, sp[16 + t], Vd);
      ke = fx_add(ke, to_fixed(term, a.S_ke));
      const unsigned tb = (unsigned)__double2hiint(term);
      kmax_hi = tb > kmax_hi ? tb : kmax_hi;
    }
    if ((chunk + 1) % FLUSH_CHUNKS == 0) dep64_flush(table, a.acc, a.count, a.m.bucket);
  }
  dep64_flush(table, a.acc, a.count, a.m.bucket);
  // dep64_flush ends with a barrier: every count is in
  WalkCount wc{crossings, 0, 0, 0, 0};
  long long killed = 0;
  if (threadIdx.x == 0) {
    wc.capped = rare.capped;
    wc.nonpos = rare.nonpos;
    wc.reflections = rare.reflections;
    wc.relocated = rare.relocated;
    killed = rare.killed;
  }
  // an upper bound of the largest term: its high word with the low word all ones
  const unsigned long long kmax = kmax_hi ? ((unsigned long long)kmax_hi << 32) | 0xffffffffull : 0ull;
  block_partial<PUSH_BLOCK>(a.partial, ke, kmax, wc, killed);
}

// Deposit of the current positions (weight_charge, PIC.cpp:139-160).
__global__ __launch_bounds__(BLOCK) void k_deposit32(DevMesh m, DevParticles32 P, int64_t n, const double* sp, int S,
                                                     unsigned long long* acc) {
  __shared__ DepTable64<DEP64> table;
  dep64_clear(table);
  int64_t lo, hi;
  block_range(n, &lo, &hi);
  __syncthreads();
  int chunk = 0;
  for (int64_t base = lo; base < hi; base += BLOCK, ++chunk) {
    const int64_t p = base + threadIdx.x;
    if (p < hi) {
      const Rec32 rc = load_rec(P.r, p);
      if (rc.elem >= 0) dep64_add(table, acc, nullptr, nullptr, rc.elem, to_fixed64(rc.w * sp[rc.type], S));
    }
    if ((chunk + 1) % FLUSH_CHUNKS == 0) dep64_flush(table, acc, nullptr, nullptr);
  }
  dep64_flush(table, acc, nullptr, nullptr);
}

// Node charges from the element sums (the finalize half of PIC::weight_charge,
// PIC.cpp:139-160): sum exactly in 64-bit, one rounding to FP64, a quarter per
// node, background added, boundary rows zeroed (PIC.cpp:155-158).
__global__ __launch_bounds__(BLOCK) void k_node_charge64(DevMesh m, const unsigned long long* __restrict__ acc, int S,
                                                         const double* __restrict__ bg, double* __restrict__ rhs) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= m.nn) return;
  // the node's element slots, NODE_BATCH at a time with all their loads in
  // flight (indices clamped to the last slot, whose extra reads are not added)
  unsigned long long s = 0;
  const int b = m.adj_ptr[i], e = m.adj_ptr[i + 1];
  for (int k = b; k < e; k += NODE_BATCH) {
    int a[NODE_BATCH];
#pragma unroll
    for (int u = 0; u < NODE_BATCH; ++u) a[u] = m.adj[min(k + u, e - 1)];
    unsigned long long v[NODE_BATCH];
#pragma unroll
    for (int u = 0; u < NODE_BATCH; ++u) v[u] = acc[a[u] >> 2];
#pragma unroll
    for (int u = 0; u < NODE_BATCH; ++u)
      if (k + u < e) s += v[u];
  }
  const double q = ldexp((double)(long long)s, -S);
  rhs[i] = m.bnd[i] ? 0.0 : (bg ? bg[i] : 0.0) + -0.25 * q;
}

__global__ __launch_bounds__(BLOCK) void k_charge_bound32(DevParticles32 P, int64_t n, const double* sp,
                                                          PushPartial* partial) {
  unsigned long long qmax = 0;
  WalkCount live{};
  for (int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (int64_t)gridDim.x * BLOCK) {
    const Rec32 rc = load_rec(P.r, p);
    if (rc.elem < 0) continue;
    const unsigned long long qb = __double_as_longlong(fabs(rc.w * sp[rc.type]));
    qmax = qb > qmax ? qb : qmax;
    live.crossings++;
  }
  block_partial(partial, fx128{0, 0}, qmax, live, 0);
}

__global__ __launch_bounds__(BLOCK) void k_ke32(DevParticles32 P, int64_t n, const double* sp, int S, int want_sum,
                                                PushPartial* partial) {
  fx128 ke{0, 0};
  unsigned long long kmax = 0;
  WalkCount live{};
  for (int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (int64_t)gridDim.x * BLOCK) {
    const Rec32 rc = load_rec(P.r, p);
    if (rc.elem < 0) continue;
    const double V[3] = {(double)rc.v[0], (double)rc.v[1], (double)rc.v[2]};
    const double term = ke_term(rc.w, sp[16 + rc.type], V);
    if (want_sum) ke = fx_add(ke, to_fixed(term, S));
    const unsigned long long tb = __double_as_longlong(term);
    kmax = tb > kmax ? tb : kmax;
    live.crossings++;
  }
  block_partial(partial, ke, kmax, live, 0);
}

__global__ __launch_bounds__(BLOCK) void k_permute32(DevParticles32 src, DevParticles32 dst,
                                                     const int32_t* __restrict__ perm, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (q >= n) return;
  const Rec32 r = load_rec(src.r, perm[q]);
  store_rec(dst.r, q, r);
  dst.elem[q] = r.elem;
}

__global__ __launch_bounds__(BLOCK) void k_convert(DevMesh m, DevParticles P64, DevParticles32 P32, int64_t off, int64_t n,
                                                   int to32) {
  const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (j >= n) return;
  const int64_t i = off + j;
  if (to32) {
    const Rec64 a = load_rec(P64.r, i);
    Rec32 b;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b.x[d] = (float)a.x[d];
      b.v[d] = (float)a.v[d];
    }
    b.w = a.w;
    b.elem = a.elem;
    b.id = a.id;
    b.type = a.type;
    b.pad = 0;
    store_rec(P32.r, i, b);
    P32.elem[i] = a.elem;
  } else {
    const Rec32 b = load_rec(P32.r, i);
    Rec64 a;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a.x[d] = (double)b.x[d];
      a.xr[d] = 0.0;
      a.v[d] = (double)b.v[d];
    }
    if (b.elem >= 0) {
      // FP64 reference position of the (widened) FP32 position in its element
      double g[12];
      load12(m.geo, b.elem, g);
      const double dl[3] = {a.x[0] - g[9], a.x[1] - g[10], a.x[2] - g[11]};
      p2r_vec(g, dl, a.xr);
    }
    a.w = b.w;
    a.elem = b.elem;
    a.id = b.id;
    a.type = b.type;
    a.pad = 0;
    store_rec(P64.r, i, a);
    P64.elem[i] = b.elem;
  }
}

// FP32 walk records from the FP64 ones: J^-1 and node 0 rounded, neighbours copied.
__global__ __launch_bounds__(BLOCK) void k_erec32(DevMesh m, float* __restrict__ out) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= m.ne) return;
  const double* src = m.erec + (size_t)e * 16;
  float4* q = reinterpret_cast<float4*>(out + (size_t)e * 16);
  q[0] = make_float4((float)src[0], (float)src[1], (float)src[2], (float)src[3]);
  q[1] = make_float4((float)src[4], (float)src[5], (float)src[6], (float)src[7]);
  q[2] = make_float4((float)src[8], (float)src[9], (float)src[10], (float)src[11]);
  reinterpret_cast<int4*>(q)[3] = reinterpret_cast<const int4*>(src)[6];
}

// One row per thread: the row's diagonal entry (its first entry with column ==
// row; padding entries, which repeat the row with value 0, come after it) moves
// to position 0, the others keep their order.
__global__ __launch_bounds__(BLOCK) void k_ell_diag_first(DevMesh m, int32_t* __restrict__ col,
                                                          double* __restrict__ val64, float* __restrict__ val32,
                                                          int* __restrict__ missing) {
  const int row = blockIdx.x * BLOCK + threadIdx.x;
  if (row >= m.nn) return;
  const int slice = row >> 6, lane = row & 63;
  const int len = m.ell_len[slice];
  const int64_t base = m.ell_off[slice] + lane;
  int kd = -1;
  for (int k = 0; k < len && kd < 0; ++k)
    if (m.ell_col[base + (int64_t)k * 64] == row) kd = k;
  if (kd < 0) {
    atomicOr(missing, 1);
    kd = 0;
  }
  int o = 0;
  for (int k = -1; k < len; ++k) {
    const int src = k < 0 ? kd : k;
    if (k == kd) continue;
    const int64_t from = base + (int64_t)src * 64, to = base + (int64_t)o * 64;
    col[to] = m.ell_col[from];
    val64[to] = m.ell_val[from];
    val32[to] = (float)m.ell_val[from];
    ++o;
  }
}

__global__ void k_to_float(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// Ownership check: a live particle's element must contain its position in FP64
// (a coordinate below -OWNER_TOL across an interior face is a mismatch; across
// a wall face the position is outside the domain by rounding and is not counted).
template <class PT>
__global__ __launch_bounds__(BLOCK) void k_check_owner(DevMesh m, PT P, int64_t n, unsigned long long* bad) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  int mismatch = 0;
  if (p < n) {
    const auto rc = load_rec(P.r, p);
    const int e = rc.elem;
    if (e >= 0) {
      double g[12];
      load12(m.geo, e, g);
      const double dl[3] = {(double)rc.x[0] - g[9], (double)rc.x[1] - g[10], (double)rc.x[2] - g[11]};
      double r[3];
      p2r_vec(g, dl, r);
      const double lam[4] = {1.0 - r[0] - r[1] - r[2], r[0], r[1], r[2]};
      const int4 nb = m.nbrs[e];
      const int across[4] = {nb.y, nb.z, nb.x, nb.w};
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (lam[j] < -OWNER_TOL && across[j] >= 0) mismatch = 1;
    }
  }
  const int cnt = __syncthreads_count(mismatch);
  if (threadIdx.x == 0 && cnt) atomicAdd(bad, (unsigned long long)cnt);
}

}  // namespace

hipError_t launch_field32(const DevMesh& m, const double* phi, float4* E, hipStream_t s) {
  hipLaunchKernelGGL(k_field32, dim3(grid_for(m.ne)), dim3(BLOCK), 0, s, m, phi, E);
  return hipGetLastError();
}
// Default grid of k_push32: 24 rounds of the workgroups resident at once (three
// 512-thread workgroups per CU at 80 VGPRs: 18,432 on 256 CUs), capped by
// PUSH32_MAX_GRID and the number of 512-particle batches.  Measured per launch at
// 100M particles: 6 rounds 4.16 ms, 10 3.98, 16 3.93, 24 3.90, 32 3.91, 48 3.94
// (smaller ranges balance the walk's uneven cost better).
constexpr int PUSH32_ROUNDS = 24;
int push32_grid(int64_t n) {
  static int cached[16];
  int dev = 0;
  int g = DEFAULT_GRID;
  if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 16) {
    if (cached[dev] == 0) {
      int per_cu = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_push32, PUSH_BLOCK, 0) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 &&
          cus > 0)
        cached[dev] = std::min(PUSH32_MAX_GRID, PUSH32_ROUNDS * per_cu * cus);
      else
        cached[dev] = DEFAULT_GRID;
    }
    g = cached[dev];
  }
  const int64_t chunks = (n + PUSH_BLOCK - 1) / PUSH_BLOCK;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, chunks));
}

hipError_t launch_push32(const PushArgs32& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  if (!a.m32.erec || !a.m32.E || !a.P.r || !a.Q.r || !a.P.elem || !a.Q.elem || !a.acc || !a.partial || !a.species ||
      !a.m.bucket)
    return hipErrorInvalidValue;
  const unsigned g =
      a.grid > 0 ? (unsigned)(a.grid < PUSH32_MAX_GRID ? a.grid : PUSH32_MAX_GRID) : (unsigned)push32_grid(a.n);
  hipLaunchKernelGGL(k_push32, dim3(g), dim3(PUSH_BLOCK), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_deposit32(const DevMesh& m, const DevParticles32& P, int64_t n, const double* species, int S,
                            unsigned long long* acc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_deposit32, dim3(grid_stride(n)), dim3(BLOCK), 0, s, m, P, n, species, S, acc);
  return hipGetLastError();
}
hipError_t launch_node_charge64(const DevMesh& m, const unsigned long long* acc, int S, const double* background,
                                double* rhs, hipStream_t s) {
  hipLaunchKernelGGL(k_node_charge64, dim3(grid_for(m.nn)), dim3(BLOCK), 0, s, m, acc, S, background, rhs);
  return hipGetLastError();
}
hipError_t launch_charge_bound32(const DevParticles32& P, int64_t n, const double* species, PushPartial* partial,
                                 hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_charge_bound32, dim3(grid_stride(n)), dim3(BLOCK), 0, s, P, n, species, partial);
  return hipGetLastError();
}
hipError_t launch_ke32(const DevParticles32& P, int64_t n, const double* species, int S, int want_sum,
                       PushPartial* partial, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ke32, dim3(grid_stride(n)), dim3(BLOCK), 0, s, P, n, species, S, want_sum, partial);
  return hipGetLastError();
}
hipError_t launch_permute32(const DevParticles32& src, const DevParticles32& dst, const int32_t* perm, int64_t n,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_permute32, dim3(grid_for(n)), dim3(BLOCK), 0, s, src, dst, perm, n);
  return hipGetLastError();
}
hipError_t launch_convert(const DevMesh& m, const DevParticles& P64, const DevParticles32& P32, int64_t off, int64_t n,
                          bool to32,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_convert, dim3(grid_for(n)), dim3(BLOCK), 0, s, m, P64, P32, off, n, to32 ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_erec32(const DevMesh& m, float* out, hipStream_t s) {
  if (m.ne <= 0 || !m.erec || !out) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_erec32, dim3(grid_for(m.ne)), dim3(BLOCK), 0, s, m, out);
  return hipGetLastError();
}
hipError_t launch_ell_diag_first(const DevMesh& m, int32_t* col, double* val64, float* val32, int* missing,
                                 hipStream_t s) {
  if (m.nn <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ell_diag_first, dim3(grid_for(m.nn)), dim3(BLOCK), 0, s, m, col, val64, val32, missing);
  return hipGetLastError();
}

hipError_t launch_to_float(const double* src, float* dst, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t g = (n + BLOCK - 1) / BLOCK;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_to_float, dim3((unsigned)g), dim3(BLOCK), 0, s, src, dst, n);
  return hipGetLastError();
}
hipError_t launch_check_owner(const DevMesh& m, const DevParticles& P, int64_t n, unsigned long long* bad,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_owner<DevParticles>, dim3(grid_for(n)), dim3(BLOCK), 0, s, m, P, n, bad);
  return hipGetLastError();
}
hipError_t launch_check_owner32(const DevMesh& m, const DevParticles32& P, int64_t n, unsigned long long* bad,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_owner<DevParticles32>, dim3(grid_for(n)), dim3(BLOCK), 0, s, m, P, n, bad);
  return hipGetLastError();
}

}  // namespace mpic
