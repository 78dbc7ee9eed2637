// Particle kernels of the MiniPIC timestep on gfx950:
//   fused push  = E gather + Boris kick + barycentric tet walk + charge deposit of
//                 the new position + kinetic energy + sort-bucket counts of the
//                 new positions, one thread per particle; when a permutation from
//                 the previous step's ranking is given, each thread gathers its
//                 particle through it and writes the result in that order;
//   split phases = the same arithmetic as separate launches (weight_Efield,
//                 accel, move) for the reference's phase-by-phase API.
// All FP64 expressions keep the reference's association order and are compiled
// without FMA contraction, so results are bit-identical to the CPU restatement.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"
#include "p1_cubature.hpp"

namespace mpic {
namespace {

// The walk's element record (one 128-B line): inverse Jacobian, node 0, face
// neighbours and 6/detJ.  A crossing needs one load: the neighbours of the new
// element arrive with its Jacobian.
struct ElemRec {
  double J[9];
  double n0[3];
  int nb[4];
  double sixodet;
  double meas;  // detJ / 24 (DevMesh::wbv_factored)
};

__device__ __forceinline__ ElemRec load_erec(const double* __restrict__ erec, int e) {
  const double2* q = reinterpret_cast<const double2*>(erec + (size_t)e * 16);
  const double2 a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3], a4 = q[4], a5 = q[5], a7 = q[7];
  const int4 a6 = reinterpret_cast<const int4*>(erec + (size_t)e * 16)[6];
  ElemRec r;
  r.J[0] = a0.x;
  r.J[1] = a0.y;
  r.J[2] = a1.x;
  r.J[3] = a1.y;
  r.J[4] = a2.x;
  r.J[5] = a2.y;
  r.J[6] = a3.x;
  r.J[7] = a3.y;
  r.J[8] = a4.x;
  r.n0[0] = a4.y;
  r.n0[1] = a5.x;
  r.n0[2] = a5.y;
  r.nb[0] = a6.x;
  r.nb[1] = a6.y;
  r.nb[2] = a6.z;
  r.nb[3] = a6.w;
  r.sixodet = a7.x;
  r.meas = a7.y;
  return r;
}

// Gather, PIC.cpp:162-184: weights from getWeights (mesh.hpp:170-176),
// E_d = sum_i sum_j (phi_i*w_j)*G[e][i]_d.  The element's nodal potentials come
// from m.ephi (copied from phi after the solve), so the gradient and the
// potentials are loaded together.
__device__ __forceinline__ void gather_E(const DevMesh& m, int e, const double R[3], double E[3]) {
  const double wt[4] = {1 - R[0] - R[1] - R[2], R[0], R[1], R[2]};
  const double2* fp = reinterpret_cast<const double2*>(m.ephi + (size_t)e * 4);
  const double2 f01 = fp[0], f23 = fp[1];
  const double dof[4] = {f01.x, f01.y, f23.x, f23.y};
  double G[12];
  load12(m.grad, e, G);
  E[0] = E[1] = E[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int d = 0; d < 3; ++d) E[d] += dof[i] * wt[j] * G[i * 3 + d];
}

// The same gather with the gradients formed from the element's inverse Jacobian
// (DevMesh::grad_derived: bit for bit the mesh's gradients).
__device__ __forceinline__ void gather_E_from_jinv(const DevMesh& m, int e, const double* Ji, const double R[3],
                                                   double E[3]) {
  const double wt[4] = {1 - R[0] - R[1] - R[2], R[0], R[1], R[2]};
  const double2* fp = reinterpret_cast<const double2*>(m.ephi + (size_t)e * 4);
  const double2 f01 = fp[0], f23 = fp[1];
  const double dof[4] = {f01.x, f01.y, f23.x, f23.y};
  double G[12];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double g = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) g += Ji[j * 3 + d] * kGradRef[i][j];
      G[i * 3 + d] = g;
    }
  E[0] = E[1] = E[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int d = 0; d < 3; ++d) E[d] += dof[i] * wt[j] * G[i * 3 + d];
}

// Boris push with B = 0, particle_move.hpp:364-392; v_ref = J^-1 v.
__device__ __forceinline__ void kick(const double* Ji, double V[3], const double E[3], double qm, double dt,
                                     double scale, double VR[3]) {
  const double B[3] = {0.0, 0.0, 0.0};
  const double sdt = dt * scale;
  double t[3];
  t[0] = qm * 0.5 * B[0];
  t[1] = qm * 0.5 * B[1];
  t[2] = qm * 0.5 * B[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) V[d] += sdt * E[d] * qm * 0.5;
  double tmp[3] = {V[0], V[1], V[2]};
  tmp[0] += V[1] * t[2] - V[2] * t[1];
  tmp[1] += V[2] * t[0] - V[0] * t[2];
  tmp[2] += V[0] * t[1] - V[1] * t[0];
  const double tsq = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) t[d] *= 2. / (1 + tsq);
  V[0] += tmp[1] * t[2] - tmp[2] * t[1];
  V[1] += tmp[2] * t[0] - tmp[0] * t[2];
  V[2] += tmp[0] * t[1] - tmp[1] * t[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) V[d] += sdt * E[d] * qm * 0.5;
  p2r_vec(Ji, V, VR);
}


// MoveParticle::operator() + find_intersection, particle_move.hpp:103-124,136-225.
// g holds the record of element e on entry and of the final element on return.
// Returns false when the particle is killed on a wall.
__device__ __forceinline__ bool walk(const DevMesh& m, int& e, ElemRec& g, double X[3], double R[3], double V[3],
                                     double VR[3], double dt, int kill, WalkCount& wc) {
  double tr = 1.0;
  int iters = 0;
  while (tr > 0) {
    if (iters++ >= WALK_CAP) {
      wc.capped++;
      break;
    }
    const double ldt = dt * tr;
    double s = 1.0;
    int face = -1;
    double dxr[3], xf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      dxr[d] = ldt * VR[d];
      xf[d] = R[d] + dxr[d];
    }
    if (VR[2] < 0 && xf[2] < 0) {
      face = 3;
      s = -R[2] / (dxr[2]);
    }
    if (VR[1] < 0 && xf[1] < 0) {
      const double sfy = -R[1] / (dxr[1]);
      if (sfy < s) {
        face = 0;
        s = sfy;
      }
    }
    if (VR[0] < 0 && xf[0] < 0) {
      const double sfx = -R[0] / (dxr[0]);
      if (sfx < s) {
        face = 2;
        s = sfx;
      }
    }
    const double xfs = xf[0] + xf[1] + xf[2];
    const double vrs = VR[0] + VR[1] + VR[2];
    if (vrs > 0 && xfs > 1) {
      const double sfs = (1 - R[0] - R[1] - R[2]) / (dxr[0] + dxr[1] + dxr[2]);
      if (sfs < s) {
        face = 1;
        s = sfs;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      X[d] = X[d] + s * ldt * V[d];
      R[d] = R[d] + s * ldt * VR[d];
    }
    if (s < 1) {
      const int nbv = face == 0 ? g.nb[0] : face == 1 ? g.nb[1] : face == 2 ? g.nb[2] : g.nb[3];
      if (nbv != -1) {  // a neighbour (-2 - e: across the period, particle_move.hpp:190-191)
        const int e2 = nbv >= 0 ? nbv : -2 - nbv;
        if (nbv < -1) X[0] = m.period_x - X[0];
        e = e2;
        wc.crossings++;
        g = load_erec(m.erec, e);  // inverse Jacobian, node 0 and neighbours of the new element
        // physToReference, mesh.hpp:197-211
        const double dl[3] = {X[0] - g.n0[0], X[1] - g.n0[1], X[2] - g.n0[2]};
        double dr[3];
        R[0] = R[1] = R[2] = 0.0;
        p2r_vec(g.J, dl, dr);
#pragma unroll
        for (int d = 0; d < 3; ++d) R[d] += dr[d];
        p2r_vec(g.J, V, VR);
      } else {
        if (kill) return false;
        wc.reflections++;
        const double* nrm = m.normals + (size_t)e * 12 + face * 3;
        const double n0 = nrm[0], n1 = nrm[1], n2 = nrm[2];
        double vn = 0;
        vn += V[0] * n0;
        vn += V[1] * n1;
        vn += V[2] * n2;
        V[0] -= 2 * vn * n0;
        V[1] -= 2 * vn * n1;
        V[2] -= 2 * vn * n2;
        p2r_vec(g.J, V, VR);
      }
    }
    if (!(s > 0)) wc.nonpos++;
    tr -= s * tr;
  }
  return true;
}

// Per-node charge of one particle, PIC.cpp:121-131, in fixed point
// (sixodet = 6/detJ of element e).
__device__ __forceinline__ void particle_charge(const DevMesh& m, int e, double sixodet, double w, double q, int S,
                                                fx128 c[4]) {
  const double tq = w * q * sixodet;
  double W[16];
  load16(m.wbv, e, W);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double acc = 0;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) acc += -tq * W[i * 4 + cc];
    c[i] = to_fixed(acc, S);
  }
}

// The same charge with the weighted basis values formed from the element's
// detJ/24 (meas) and the cubature rule's basis values: wbv[i][c] = meas * N_i(x_c)
// bit for bit where DevMesh::wbv_factored says so, so the 16 products and their
// sums are the reference's.
__device__ __forceinline__ void particle_charge_factored(double sixodet, double meas, double w, double q, int S,
                                                         fx128 c[4]) {
  const double tq = w * q * sixodet;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double acc = 0;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) acc += -tq * (meas * cub_basis(i, cc));
    c[i] = to_fixed(acc, S);
  }
}

// ephi[e][k] = phi[conn[e][k]] (after every solve).
__global__ __launch_bounds__(BLOCK) void k_elem_phi(DevMesh m, const double* __restrict__ phi) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= m.ne) return;
  const int4 nd = m.conn[e];
  double2* out = reinterpret_cast<double2*>(m.ephi + (size_t)e * 4);
  out[0] = double2{phi[nd.x], phi[nd.y]};
  out[1] = double2{phi[nd.z], phi[nd.w]};
}


// ---- FP64 particle records through LDS -------------------------------------------
// As in the precision form: a wavefront's 64 consecutive 96-byte records (6 KB)
// are read and written as contiguous 16-byte pieces (lane l takes pieces
// 64 k + l), one 64-byte line per lane quad instead of six, and pass to and from
// the lanes that own them through a 3 KB LDS area, half the wavefront's records
// at a time (pieces 192 h .. 192 h + 191 are records 32 h .. 32 h + 31).
constexpr int REC64_HALF = 192;  // 16-byte pieces of 32 records

__device__ __forceinline__ Rec64 rec64_from(const float4* sl) {
  const double2 a0 = reinterpret_cast<const double2*>(sl)[0], a1 = reinterpret_cast<const double2*>(sl)[1],
                a2 = reinterpret_cast<const double2*>(sl)[2], a3 = reinterpret_cast<const double2*>(sl)[3],
                a4 = reinterpret_cast<const double2*>(sl)[4];
  const int4 a5 = reinterpret_cast<const int4*>(sl)[5];
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

// Lanes 0 .. nv - 1 receive records: slot s0 + lane (in place), or slot src of
// each lane (through the ranking; key holds them).  Others get elem = -1.
__device__ __forceinline__ Rec64 wave_load_recs64(float4* __restrict__ sl, int* __restrict__ key,
                                                  const Rec64* __restrict__ P, bool gather, int64_t s0, int src,
                                                  int nv) {
  const int lane = fresh_lane(), n6 = 6 * nv;
  if (gather) {
    if (lane < nv) key[lane] = src;
    wave_sync();
  }
  const float4* in = reinterpret_cast<const float4*>(P + s0);
  auto piece = [&](int j) -> float4 {
    if (!gather) return in[j];
    const int r = j / 6;
    return reinterpret_cast<const float4*>(P + key[r])[j - 6 * r];
  };
  Rec64 r{};
  r.elem = -1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int b = REC64_HALF * h;
    if (b < n6) {
      float4 v0, v1, v2;
      if (b + lane < n6) v0 = piece(b + lane);
      if (b + 64 + lane < n6) v1 = piece(b + 64 + lane);
      if (b + 128 + lane < n6) v2 = piece(b + 128 + lane);
      if (b + lane < n6) sl[lane] = v0;
      if (b + 64 + lane < n6) sl[64 + lane] = v1;
      if (b + 128 + lane < n6) sl[128 + lane] = v2;
      wave_sync();
      const int rr = lane - 32 * h;
      if (rr >= 0 && rr < 32 && lane < nv) r = rec64_from(sl + 6 * rr);
      wave_sync();
    }
  }
  return r;
}

// Lane l's record (l < nv) to slot s0 + l.
__device__ __forceinline__ void wave_store_recs64(float4* __restrict__ sl, Rec64* __restrict__ Q, int64_t s0, int nv,
                                                  const Rec64& r) {
  const int lane = fresh_lane(), n6 = 6 * nv;
  float4* out = reinterpret_cast<float4*>(Q + s0);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int b = REC64_HALF * h;
    if (b < n6) {
      const int rr = lane - 32 * h;
      if (rr >= 0 && rr < 32 && lane < nv) {
        double2* q = reinterpret_cast<double2*>(sl + 6 * rr);
        q[0] = double2{r.x[0], r.x[1]};
        q[1] = double2{r.x[2], r.xr[0]};
        q[2] = double2{r.xr[1], r.xr[2]};
        q[3] = double2{r.v[0], r.v[1]};
        q[4] = double2{r.v[2], r.w};
        reinterpret_cast<int4*>(sl + 6 * rr)[5] = int4{r.elem, (int)r.id, (int)r.type, 0};
      }
      wave_sync();
      const float4 w0 = sl[lane], w1 = sl[64 + lane], w2 = sl[128 + lane];
      if (b + lane < n6) out[b + lane] = w0;
      if (b + 64 + lane < n6) out[b + 64 + lane] = w1;
      if (b + 128 + lane < n6) out[b + 128 + lane] = w2;
      wave_sync();
    }
  }
}

struct KArgs {
  DevMesh m;
  DevParticles P;   // input particles
  DevParticles Q;   // output particles of the fused push (== P when pushing in place)
  const int32_t* perm;  // fused push: output position -> input position, or null (in place)
  int32_t* count;       // fused push: per-element counts of the new positions, or null
  int64_t n;
  const double* phi;
  double dt, scale;
  int kill, S_charge, S_ke;
  unsigned long long* acc;
  PushPartial* partial;
  double* E;
  double* vr;
  const double* sp;  // [48]: q[16], m[16], q/m[16]
  minipic::ParticleListView view;  // k_move: wall kills go through add_particles_to_migrate_list
};

__global__ __launch_bounds__(BLOCK, 3) void k_push_fused(KArgs a) {
  __shared__ DepTable table;
  __shared__ float4 recx[BLOCK / 64][REC64_HALF];
  __shared__ int recsrc[BLOCK / 64][64];
  __shared__ double sp[48];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float4* sl = recx[wave];
  int* skey = recsrc[wave];
  dep_table_clear(table);
  if (threadIdx.x < 48) sp[threadIdx.x] = a.sp[threadIdx.x];
  const bool gather = a.perm != nullptr;
  WalkCount wc{};
  long long killed = 0;
  fx128 ke{0, 0};
  unsigned long long kmax = 0;
  int64_t lo, hi;
  block_range(a.n, &lo, &hi);
  __syncthreads();
  int chunk = 0;
  for (int64_t base = lo; base < hi; base += BLOCK, ++chunk) {
    const int64_t p = base + threadIdx.x;  // output position
    const int64_t s0 = base + wave * 64;   // the wavefront's first
    const int64_t rem = hi - s0;
    const int nv = rem <= 0 ? 0 : rem >= 64 ? 64 : (int)rem;
    int enew = -1;
    double w = 0, V[3] = {0, 0, 0};
    double sixodet = 0, meas = 0;  // of the final element (its walk record)
    int t = 0;
    // records in: sorted order gathers through the permutation
    Rec64 rc{};
    rc.elem = -1;
    if (nv > 0) rc = wave_load_recs64(sl, skey, a.P.r, gather, s0, gather && p < hi ? a.perm[p] : 0, nv);
    if (p < hi) {
      ElemRec g;
      int e = rc.elem;
      if (e >= 0) {
        double X[3], R[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          X[d] = rc.x[d];
          R[d] = rc.xr[d];
          V[d] = rc.v[d];
        }
        t = (int)rc.type;
        w = rc.w;
        g = load_erec(a.m.erec, e);
        double E[3];
        if (a.m.grad_derived)
          gather_E_from_jinv(a.m, e, g.J, R, E);
        else
          gather_E(a.m, e, R, E);
        double VR[3];
        kick(g.J, V, E, sp[32 + t], a.dt, a.scale, VR);
        const bool alive = walk(a.m, e, g, X, R, V, VR, a.dt, a.kill, wc);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          rc.x[d] = X[d];
          rc.xr[d] = R[d];
          rc.v[d] = V[d];
        }
        enew = alive ? e : -1;
        killed += alive ? 0 : 1;
        rc.elem = enew;
        sixodet = g.sixodet;
        meas = g.meas;
      }
    }
    // node charges from the walk record before the stores: a wavefront's vector
    // memory operations complete in order, so a use of the record's last loads
    // after the stores would wait for the stores as well
    fx128 c[4];
    if (enew >= 0 && a.m.wbv_factored) particle_charge_factored(sixodet, meas, w, sp[t], a.S_charge, c);
    // records out (a dead particle's as read: the ranking drops it), then the element mirror
    if (nv > 0) wave_store_recs64(sl, a.Q.r, s0, nv, rc);
    if (p < hi) a.Q.elem[p] = rc.elem < 0 ? -1 : rc.elem;
    if (enew >= 0) {
      if (!a.m.wbv_factored) particle_charge(a.m, enew, a.m.sixodet[enew], w, sp[t], a.S_charge, c);
      dep_table_add(table, a.acc, a.count, a.m.bucket, enew, c);
      const double term = ke_term(w, sp[16 + t], V);
      ke = fx_add(ke, to_fixed(term, a.S_ke));
      const unsigned long long tb = __double_as_longlong(term);
      kmax = tb > kmax ? tb : kmax;
    }
    if ((chunk + 1) % FLUSH_CHUNKS == 0) dep_table_flush(table, a.acc, a.count, a.m.bucket);
  }
  dep_table_flush(table, a.acc, a.count, a.m.bucket);
  block_partial(a.partial, ke, kmax, wc, killed);
}

// weight_Efield: E of every live particle (PIC.cpp:186-189).
__global__ __launch_bounds__(BLOCK) void k_gather(KArgs a) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (p >= a.n) return;
  const Rec64 rc = load_rec(a.P.r, p);
  const int e = rc.elem;
  if (e < 0) return;
  const int64_t cap = a.P.cap;
  double E[3];
  gather_E(a.m, e, rc.xr, E);
  for (int d = 0; d < 3; ++d) a.E[d * cap + p] = E[d];
}

// AccelParticlesTag: kick v from the stored E and store v_ref.
__global__ __launch_bounds__(BLOCK) void k_accel(KArgs a) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (p >= a.n) return;
  Rec64 rc = load_rec(a.P.r, p);
  const int e = rc.elem;
  if (e < 0) return;
  const int64_t cap = a.P.cap;
  double V[3], E[3], VR[3];
  for (int d = 0; d < 3; ++d) {
    V[d] = rc.v[d];
    E[d] = a.E[d * cap + p];
  }
  const ElemRec g = load_erec(a.m.erec, e);
  kick(g.J, V, E, a.sp[32 + rc.type], a.dt, a.scale, VR);
  for (int d = 0; d < 3; ++d) {
    rc.v[d] = V[d];
    a.vr[d * cap + p] = VR[d];
  }
  store_rec(a.P.r, p, rc);
}

// MoveParticle: the tet walk from the stored v and v_ref (no deposit).
// MoveParticle::operator() (particle_move.hpp:103-124): walk each particle, then
// the team hands its wall kills (KILL_ONLY) to the team-collective
// ParticleList::add_particles_to_migrate_list, which marks them dead; the host
// then compacts the list (cleanup_list).
__global__ __launch_bounds__(BLOCK) void k_move(KArgs a) {
  __shared__ long team_id[BLOCK];
  __shared__ int team_proc[BLOCK];
  __shared__ int scratch[MPIC_TEAM_SCRATCH];
  const minipic::member_type dev(scratch);
  const int64_t cap = a.P.cap;
  WalkCount wc{};
  long long killed = 0;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < a.n; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t p = base + threadIdx.x;
    int proc = minipic::ParticleListView::DO_NOTHING;
    if (p < a.n) {
      Rec64 rc = load_rec(a.P.r, p);
      int e = rc.elem;
      if (e >= 0) {
        double X[3], R[3], V[3], VR[3];
        for (int d = 0; d < 3; ++d) {
          X[d] = rc.x[d];
          R[d] = rc.xr[d];
          V[d] = rc.v[d];
          VR[d] = a.vr[d * cap + p];
        }
        ElemRec g = load_erec(a.m.erec, e);
        const bool alive = walk(a.m, e, g, X, R, V, VR, a.dt, a.kill, wc);
        for (int d = 0; d < 3; ++d) {
          rc.x[d] = X[d];
          rc.xr[d] = R[d];
          rc.v[d] = V[d];
        }
        rc.elem = e;
        store_rec(a.P.r, p, rc);
        a.P.elem[p] = e;
        if (!alive) {
          proc = minipic::ParticleListView::KILL_ONLY;
          killed++;
        }
      }
    }
    team_id[threadIdx.x] = (long)p;
    team_proc[threadIdx.x] = proc;
    a.view.add_particles_to_migrate_list(dev, team_id, team_proc, BLOCK);
  }
  block_partial(a.partial, fx128{0, 0}, 0ull, wc, killed);
}

// Deposit of the current positions (weight_charge, PIC.cpp:139-160).
__global__ __launch_bounds__(BLOCK) void k_deposit(DevMesh m, DevParticles P, int64_t n, const double* sp, int S,
                                                   unsigned long long* acc) {
  __shared__ DepTable table;
  dep_table_clear(table);
  int64_t lo, hi;
  block_range(n, &lo, &hi);
  __syncthreads();
  int chunk = 0;
  for (int64_t base = lo; base < hi; base += BLOCK, ++chunk) {
    const int64_t p = base + threadIdx.x;
    if (p < hi) {
      const Rec64 rc = load_rec(P.r, p);
      const int e = rc.elem;
      if (e >= 0) {
        fx128 c[4];
        particle_charge(m, e, m.sixodet[e], rc.w, sp[rc.type], S, c);
        dep_table_add(table, acc, nullptr, nullptr, e, c);
      }
    }
    if ((chunk + 1) % FLUSH_CHUNKS == 0) dep_table_flush(table, acc, nullptr, nullptr);
  }
  dep_table_flush(table, acc, nullptr, nullptr);
}

// KE terms (PIC.cpp:191-200): the per-workgroup max term, the live count (in
// the crossings slot) and, when want_sum, the fixed-point sum at scale S.
__global__ __launch_bounds__(BLOCK) void k_ke(DevParticles P, int64_t n, const double* sp, int S, int want_sum,
                                              PushPartial* partial) {
  fx128 ke{0, 0};
  unsigned long long kmax = 0;
  WalkCount live{};
  for (int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (int64_t)gridDim.x * BLOCK) {
    const Rec64 rc = load_rec(P.r, p);
    if (rc.elem < 0) continue;
    const double term = ke_term(rc.w, sp[16 + rc.type], rc.v);
    if (want_sum) ke = fx_add(ke, to_fixed(term, S));
    const unsigned long long tb = __double_as_longlong(term);
    kmax = tb > kmax ? tb : kmax;
    live.crossings++;
  }
  block_partial(partial, ke, kmax, live, 0);
}

// Charge bound for the fixed-point scale: max |w q| (in the max slot) and the
// live count (in the crossings slot).
__global__ __launch_bounds__(BLOCK) void k_charge_bound(DevParticles P, int64_t n, const double* sp,
                                                        PushPartial* partial) {
  unsigned long long qmax = 0;
  WalkCount live{};
  for (int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (int64_t)gridDim.x * BLOCK) {
    const Rec64 rc = load_rec(P.r, p);
    if (rc.elem < 0) continue;
    const unsigned long long qb = __double_as_longlong(fabs(rc.w * sp[rc.type]));
    qmax = qb > qmax ? qb : qmax;
    live.crossings++;
  }
  block_partial(partial, fx128{0, 0}, qmax, live, 0);
}

// Final reduction of per-workgroup partials (one workgroup, integer sums and maxima).
// The sum of the workgroup partials (exact: fixed-point KE, integer counters,
// and a maximum, so the order does not matter).  One workgroup of FINISH_T
// threads; each thread's FINISH_U partials are loaded together, then waves
// reduce by shuffles and one thread adds the waves' results.
constexpr int FINISH_T = 1024;
constexpr int FINISH_U = 4;

struct FinishAcc {
  fx128 ke;
  unsigned long long mx;
  long long c[6];
  __device__ void add(const PushPartial& q) {
    ke = fx_add(ke, fx128{q.ke_lo, q.ke_hi});
    mx = q.ke_max_bits > mx ? q.ke_max_bits : mx;
    c[0] += q.crossings;
    c[1] += q.killed;
    c[2] += q.capped;
    c[3] += q.nonpos;
    c[4] += q.reflections;
    c[5] += q.relocated;
  }
  __device__ PushPartial out() const { return PushPartial{ke.lo, ke.hi, mx, c[0], c[1], c[2], c[3], c[4], c[5]}; }
};

__global__ __launch_bounds__(FINISH_T) void k_push_finish(const PushPartial* __restrict__ part, int nparts,
                                                          PushPartial* out) {
  FinishAcc acc{{0, 0}, 0, {0, 0, 0, 0, 0, 0}};
  for (int i0 = threadIdx.x; i0 < nparts; i0 += FINISH_T * FINISH_U) {
    PushPartial q[FINISH_U];
#pragma unroll
    for (int u = 0; u < FINISH_U; ++u) q[u] = part[min(i0 + u * FINISH_T, nparts - 1)];
#pragma unroll
    for (int u = 0; u < FINISH_U; ++u)
      if (i0 + u * FINISH_T < nparts) acc.add(q[u]);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    PushPartial o;
    o.ke_lo = __shfl_xor(acc.ke.lo, off);
    o.ke_hi = __shfl_xor(acc.ke.hi, off);
    o.ke_max_bits = __shfl_xor(acc.mx, off);
    o.crossings = __shfl_xor(acc.c[0], off);
    o.killed = __shfl_xor(acc.c[1], off);
    o.capped = __shfl_xor(acc.c[2], off);
    o.nonpos = __shfl_xor(acc.c[3], off);
    o.reflections = __shfl_xor(acc.c[4], off);
    o.relocated = __shfl_xor(acc.c[5], off);
    acc.add(o);
  }
  __shared__ PushPartial sh[FINISH_T / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc.out();
  __syncthreads();
  if (threadIdx.x == 0) {
    FinishAcc r{{0, 0}, 0, {0, 0, 0, 0, 0, 0}};
    for (int w = 0; w < FINISH_T / 64; ++w) r.add(sh[w]);
    *out = r.out();
  }
}

KArgs kargs(const PushArgs& a) {
  return KArgs{a.m, a.P, a.perm ? a.Q : a.P, a.perm, a.count, a.n, a.phi, a.dt, a.scale, a.kill, a.S_charge,
               a.S_ke, a.acc, a.partial, a.E, a.vr, a.species, a.view};
}

}  // namespace

int partials_for(int64_t n) { return (int)grid_stride(n); }

hipError_t launch_charge_bound(const DevParticles& P, int64_t n, const double* species, PushPartial* partial,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_charge_bound, dim3(grid_stride(n)), dim3(BLOCK), 0, s, P, n, species, partial);
  return hipGetLastError();
}

hipError_t launch_push_fused(const PushArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_push_fused, dim3(grid_stride(a.n)), dim3(BLOCK), 0, s, kargs(a));
  return hipGetLastError();
}
hipError_t launch_elem_phi(const DevMesh& m, const double* phi, hipStream_t s) {
  if (m.ne <= 0 || !m.ephi) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_elem_phi, dim3(grid_for(m.ne)), dim3(BLOCK), 0, s, m, phi);
  return hipGetLastError();
}
hipError_t launch_gather(const PushArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(grid_for(a.n)), dim3(BLOCK), 0, s, kargs(a));
  return hipGetLastError();
}
hipError_t launch_accel(const PushArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_accel, dim3(grid_for(a.n)), dim3(BLOCK), 0, s, kargs(a));
  return hipGetLastError();
}
hipError_t launch_move(const PushArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_move, dim3(grid_stride(a.n)), dim3(BLOCK), 0, s, kargs(a));
  return hipGetLastError();
}
hipError_t launch_push_finish(const PushPartial* part, int nparts, PushPartial* out, hipStream_t s) {
  hipLaunchKernelGGL(k_push_finish, dim3(1), dim3(FINISH_T), 0, s, part, nparts, out);
  return hipGetLastError();
}
hipError_t launch_deposit(const DevMesh& m, const DevParticles& P, int64_t n, const double* species, int S,
                          unsigned long long* acc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_deposit, dim3(grid_stride(n)), dim3(BLOCK), 0, s, m, P, n, species, S, acc);
  return hipGetLastError();
}
hipError_t launch_ke(const DevParticles& P, int64_t n, const double* species, int S, int want_sum,
                     PushPartial* partial, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ke, dim3(grid_stride(n)), dim3(BLOCK), 0, s, P, n, species, S, want_sum, partial);
  return hipGetLastError();
}

}  // namespace mpic
