// Internal declarations of the MI355X MiniPIC library: device data layout,
// exact fixed-point arithmetic and launch wrappers shared by the .hip files.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "minipic_c.h"
#include "minipic/particle_list_view.hpp"
#include "mg.hpp"

namespace mpic {

constexpr int BLOCK = 256;         // threads per workgroup for particle and row kernels
constexpr int WALK_CAP = 1000;     // crossings per particle per step (the reference has no cap)
constexpr int ELL_SLICE = 64;      // rows per ELL slice = one wavefront
constexpr int MAX_PARTIALS = 8192; // workgroups of the grid-stride particle kernels (upper bound)
constexpr int PUSH32_MAX_GRID = 32768;  // workgroups of k_push32's grid (upper bound)
constexpr int PARTIAL_SLOTS = PUSH32_MAX_GRID > MAX_PARTIALS ? PUSH32_MAX_GRID : MAX_PARTIALS;
constexpr int DEFAULT_GRID = 4096; // their default number of workgroups
constexpr int IR_MAX_OUTER = 40;   // refinement passes of the precision-form solve
constexpr int NUM_XCD = 8;         // accelerator dies of the MI355X, each with its own L2

// Row-chunk kernels launch xcd_grid(n) workgroups for n chunks of BLOCK rows and
// take their chunk from xcd_chunk (mpic_device.hpp).
inline int xcd_grid(int n) { return (n + NUM_XCD - 1) / NUM_XCD * NUM_XCD; }

// 128-bit two's-complement fixed-point value (lo word first).
struct fx128 {
  unsigned long long lo;
  long long hi;
};

// ---------------------------------------------------------------------------
// Device view of the mesh and discretisation (all FP64 unless noted).
// ---------------------------------------------------------------------------
struct DevMesh {
  int32_t nn, ne;
  const double* geo;       // [ne][12]  inverse Jacobian (row-major 3x3) then node-0 coordinates
  // [ne][16]  the walk's element record, one 128-B line: inverse Jacobian (9),
  // node 0 (3), face neighbours (4 x int32 in 2 slots), 6/detJ, detJ/24 (see wbv_factored)
  const double* erec;
  double* ephi;            // [ne][4]   potential at the element's nodes (refreshed after each solve)
  const double* grad;      // [ne][12]  physical P1 gradients G[e][i][d]
  const int4* conn;        // [ne]      element -> 4 node ids
  const int4* nbrs;        // [ne]      face f -> neighbour element, -1 wall
  const double* normals;   // [ne][12]  unit face normals
  const double* sixodet;   // [ne]      6/detJ
  const double* detj;      // [ne]      det J
  const double* wbv;       // [ne][16]  weighted basis values wbv[i][c]
  const double* nodes;     // [nn][3]
  const uint8_t* bnd;      // [nn]      Dirichlet flag
  const int32_t* adj_ptr;  // [nn+1]    node -> slots e*4+k of elements touching it
  const int32_t* adj;      // [4*ne]
  const fx128* bgfix;      // [nn]      background load vector in fixed point at the current scale
  const int32_t* bucket;   // [ne]      spatial bucket of each element (particle ordering key)
  int32_t nbuckets;
  // Stiffness matrix in sliced ELL (64-row slices, entries column-major inside
  // a slice so a wavefront reads one entry of 64 rows per load).  Each row keeps
  // the CSR entry order; padding entries are (col = row, val = +0).
  int32_t nslices;
  const int64_t* ell_off;  // [nslices+1] entry offset of each slice
  const int32_t* ell_len;  // [nslices]   entries per row in the slice
  const int32_t* ell_col;
  const double* ell_val;
  double lo[3], hi[3];
  double period_x;  // lo[0] + hi[0]: x -> period_x - x across a periodic face (neighbour stored as -2 - e)
  // 1 when every weighted basis value is wbv[e][i][c] = (detJ_e * w) * N_i(x_c) bit for bit (the rule of
  // p1_cubature.hpp); erec[e][15] then holds detJ_e * w, and the FP64 push forms the 16 values itself
  int wbv_factored;
  // 1 when every gradient is grad[e][i][d] = sum_j J^-1[j][d] kGradRef[i][j] bit for bit (p1_cubature.hpp):
  // the FP64 push then forms them from the inverse Jacobian of its walk record
  int grad_derived;
};

// Particles on the device are records (one 16-B aligned record per particle,
// read and written as 16-B words) plus a mirrored array of owning elements for
// the sort.  A push that reads its particles through a permutation then touches
// one or two cache lines per particle instead of one line per field.

// FP64 particle record, 96 B.
struct alignas(16) Rec64 {
  double x[3];    // ParticleList::x      particle_list.hpp:71
  double xr[3];   // ParticleList::x_ref  particle_list.hpp:73
  double v[3];    // ParticleList::v      particle_list.hpp:75
  double w;       // ParticleList::weight particle_list.hpp:85
  int32_t elem;   // ParticleList::ielement, -1 dead
  uint32_t id;    // identity key (fill order)
  uint32_t type;  // ParticleList::type
  uint32_t pad;
};
static_assert(sizeof(Rec64) == 96, "FP64 particle record is six 16-B words");
// The device ParticleList view (include/minipic/particle_list_view.hpp) reads
// the same records.
static_assert(offsetof(Rec64, xr) == offsetof(minipic::ParticleRecord, x_ref) &&
                  offsetof(Rec64, v) == offsetof(minipic::ParticleRecord, v) &&
                  offsetof(Rec64, w) == offsetof(minipic::ParticleRecord, weight) &&
                  offsetof(Rec64, elem) == offsetof(minipic::ParticleRecord, ielement) &&
                  offsetof(Rec64, id) == offsetof(minipic::ParticleRecord, id) &&
                  offsetof(Rec64, type) == offsetof(minipic::ParticleRecord, type),
              "Rec64 and minipic::ParticleRecord differ");
// Entries of the view's migrate list (migrations only arise between ranks).
constexpr int64_t MIGRATE_CAP = 1 << 16;

struct DevParticles {
  Rec64* r;       // [cap]
  int32_t* elem;  // [cap] copy of r[].elem, read by the sort
  int64_t cap;
};

// Precision-form particle record, 48 B: FP32 position and velocity; the weight
// stays FP64 because charges are summed exactly from it.  The reference
// position is not stored: the push derives it from the position and the
// element's inverse Jacobian, and the FP64 export computes it in FP64.
struct alignas(16) Rec32 {
  float x[3];
  int32_t elem;
  float v[3];
  uint32_t type;
  double w;
  uint32_t id;
  uint32_t pad;
};
static_assert(sizeof(Rec32) == 48, "precision-form particle record is three 16-B words");

struct DevParticles32 {
  Rec32* r;       // [cap]
  int32_t* elem;  // [cap] copy of r[].elem, read by the sort
  int64_t cap;
};

// Precision-form mesh tables.
struct Mesh32 {
  const float* erec;      // [ne][16] inverse Jacobian (9), node 0 (3), face neighbours (4 x int32), FP32
  float4* E;              // [ne]     field of the current potential (constant per tet)
  // The precision form's copy of the sliced ELL matrix (same slices and
  // lengths as DevMesh) with each row's diagonal entry first and the others in
  // their stored order: columns, FP64 and FP32 values.
  const int32_t* ell_col;
  const double* ell_val64;
  const float* ell_val;
};
// DevMesh whose ELL columns and FP64 values are the precision form's
// diagonal-first copy.
inline DevMesh diag_first(const DevMesh& m, const Mesh32& m32) {
  DevMesh d = m;
  d.ell_col = m32.ell_col;
  d.ell_val = m32.ell_val64;
  return d;
}

struct Species {
  double q[MPIC_MAX_SPECIES];
  double m[MPIC_MAX_SPECIES];
  double qm[MPIC_MAX_SPECIES];
};

// Per-workgroup partial results of the fused push (integer sums: order-free).
struct PushPartial {
  unsigned long long ke_lo;
  long long ke_hi;
  unsigned long long ke_max_bits;  // max KE term (non-negative double bits)
  long long crossings, killed, capped, nonpos, reflections, relocated;
};

struct CGState {
  int done, iters, converged, max_iter;
  double rho0, rho, tol2;
};

// ---------------------------------------------------------------------------
// Exact fixed point: value v -> round-to-nearest-even(v * 2^S) as a 128-bit
// integer.  Sums of such integers are exact and independent of order.
// Back-conversion: |a| = hi*2^64 + lo, (double)hi * 2^64 + (double)lo, * 2^-S.
// ---------------------------------------------------------------------------
__host__ __device__ inline fx128 fx_neg(fx128 a) {
  unsigned long long lo = ~a.lo + 1ull;
  unsigned long long hi = ~(unsigned long long)a.hi + (lo == 0ull ? 1ull : 0ull);
  return {lo, (long long)hi};
}
__host__ __device__ inline fx128 fx_add(fx128 a, fx128 b) {
  unsigned long long lo = a.lo + b.lo;
  unsigned long long c = lo < a.lo ? 1ull : 0ull;
  return {lo, (long long)((unsigned long long)a.hi + (unsigned long long)b.hi + c)};
}
__host__ __device__ inline fx128 to_fixed(double v, int S) {
  double r = rint(ldexp(v, S));
  bool neg = r < 0;
  double mg = fabs(r);
  double hd = floor(mg * 5.421010862427522170e-20);  // * 2^-64
  double ld = mg - hd * 18446744073709551616.0;
  fx128 a{(unsigned long long)ld, (long long)(unsigned long long)hd};
  return neg ? fx_neg(a) : a;
}
__host__ __device__ inline double from_fixed(fx128 a, int S) {
  bool neg = a.hi < 0;
  if (neg) a = fx_neg(a);
  double d = (double)(unsigned long long)a.hi * 18446744073709551616.0 + (double)a.lo;
  d = ldexp(d, -S);
  return neg ? -d : d;
}
// Scale so that any sum bounded by `bound` stays below 2^124 in fixed point.
inline int scale_for_bound(double bound) {
  if (!(bound > 0) || !std::isfinite(bound)) return 100;
  int e = std::ilogb(bound) + 1;
  int S = 124 - e;
  if (S > 1000) S = 1000;
  if (S < -1000) S = -1000;
  return S;
}

// 64-bit fixed point of the precision-form deposit: scale so that any sum
// bounded by `bound` stays below 2^62.
inline int scale_for_bound64(double bound) {
  if (!(bound > 0) || !std::isfinite(bound)) return 40;
  int S = 62 - (std::ilogb(bound) + 1);
  if (S > 1000) S = 1000;
  if (S < -1000) S = -1000;
  return S;
}
__host__ __device__ inline long long to_fixed64(double v, int S) { return (long long)rint(ldexp(v, S)); }

// Launch shapes: one thread per item, or a grid-stride grid capped at MAX_PARTIALS.
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
inline unsigned grid_stride(int64_t n) {
  const int64_t g = (n + BLOCK - 1) / BLOCK;
  return (unsigned)(g < DEFAULT_GRID ? g : DEFAULT_GRID);
}


// ---------------------------------------------------------------------------
// Launch wrappers (one .hip file each).  All are stream-ordered and return
// hipError_t of the launch.
// ---------------------------------------------------------------------------
struct PushArgs {
  DevMesh m;
  DevParticles P;          // input particles
  DevParticles Q;          // output (fused push with perm)
  const int32_t* perm;     // output position -> input position, or null (push in place)
  int32_t* count;          // bucket counts (m.bucket) of the new positions (fused push), or null
  int64_t n;
  const double* phi;
  double dt, scale;
  int kill;
  int S_charge, S_ke;
  unsigned long long* acc; // [ne*16] element charge accumulators (4 nodes x 4 digit words)
  PushPartial* partial;    // [gridDim.x]
  double* E;               // transient [3][cap] for the split entry points
  double* vr;              // transient [3][cap]
  const double* species;   // device [48]: q[16], m[16], q/m[16]
  minipic::ParticleListView view;  // the list as the move's kill path sees it (split entry points)
};

hipError_t launch_push_fused(const PushArgs& a, hipStream_t s);   // gather+kick+walk+deposit+KE
hipError_t launch_elem_phi(const DevMesh& m, const double* phi, hipStream_t s);  // m.ephi from phi
hipError_t launch_gather(const PushArgs& a, hipStream_t s);       // weight_Efield -> E
hipError_t launch_accel(const PushArgs& a, hipStream_t s);        // kick from E -> v, vr
hipError_t launch_move(const PushArgs& a, hipStream_t s);         // walk from v, vr
hipError_t launch_push_finish(const PushPartial* part, int nparts, PushPartial* out, hipStream_t s);
// KE terms over live particles into per-workgroup partials (want_sum = 0: max only).
hipError_t launch_ke(const DevParticles& P, int64_t n, const double* species, int S, int want_sum,
                     PushPartial* partial, hipStream_t s);

hipError_t launch_deposit(const DevMesh& m, const DevParticles& P, int64_t n, const double* species, int S,
                          unsigned long long* acc, hipStream_t s);
hipError_t launch_node_charge(const DevMesh& m, unsigned long long* acc, int S, double* rhs, hipStream_t s);
// max |w q| (ke_max_bits slot) and live count (crossings slot) into partials.
hipError_t launch_charge_bound(const DevParticles& P, int64_t n, const double* species, PushPartial* partial,
                               hipStream_t s);
// Number of workgroup partials the particle kernels write for n particles.
int partials_for(int64_t n);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);

struct CGBuffers {
  double *x, *b, *r, *p0, *p1, *q;  // r, p0, p1: the persistent solve's (and BiCGStab's scratch)
  double2* pr[2];                   // (p_{i-1}, r_i) words by iteration parity (per-iteration kernels)
  double *part_bb, *part_rr, *part_pq;
  double* rho_hist;  // [max_iter+1]
  CGState* st;
  int nchunks;
};
hipError_t launch_cg_init(const DevMesh& m, CGBuffers& c, double tol, int max_iter, hipStream_t s);
// Resets the solver state (tolerance, max_iter, done = 0) without the first residual.
hipError_t launch_cg_setup(CGBuffers& c, double tol, int max_iter, hipStream_t s);
// The whole CG as one cooperative launch (k_cg_persist.hip); epoch [nchunks] and
// err must be zero.  Results equal the per-iteration kernels bit for bit.
hipError_t launch_cg_persist(const DevMesh& m, CGBuffers& c, double tol, int max_iter, unsigned* epoch, int* err,
                             hipStream_t s);
// Workgroups of a persistent solve kernel resident at once (0: no cooperative
// launch); which = 0 FP64 CG, 1 precision-form refinement.
int persist_capacity(int which);
// Iteration it: convergence test + p/q update, then (it < max_iter) the x/r update.
hipError_t launch_cg_iter(const DevMesh& m, CGBuffers& c, int it, int max_iter, hipStream_t s);

// BiCGStab of the reference (k_bicgstab.hip).
struct BiState {
  int done, iters, status, max_iter;  // status 0 converged, 1 max_iter, 2 early exit
  double rho0, rho, resid, tol;
};
struct BiBuffers {
  double *x, *b, *r, *rn, *t, *p0, *p1, *v0, *v1;
  double* part_bb;  // [nchunks]   b.b
  double* part_r;   // [2*nchunks] r.r and r~.r per chunk
  double* part_bv;  // [nchunks]   b.v
  double* part_rt;  // [3*nchunks] r'.r', t.t and t.r' per chunk
  double* hist;     // [3*(max_iter+1)] rho, alpha, omega of each iteration
  BiState* st;
  int nchunks;
};
hipError_t launch_bi_init(const DevMesh& m, BiBuffers& c, double tol, int max_iter, hipStream_t s);
hipError_t launch_bi_iter(const DevMesh& m, BiBuffers& c, int it, hipStream_t s);

// ---- precision form (k_mixed.hip, k_cg32.hip) ----
struct PushArgs32 {
  DevMesh m;
  Mesh32 m32;
  DevParticles32 P;        // input particles
  DevParticles32 Q;        // output (with perm)
  const int32_t* perm;     // output position -> input position, or null (in place)
  int32_t* count;          // bucket counts of the new positions, or null
  int64_t n;
  float dt, scale;
  int kill, S_charge, S_ke;
  unsigned long long* acc;  // [ne] element charge sums, 64-bit fixed point at 2^-S_charge
  PushPartial* partial;
  const double* species;
  int grid;                 // workgroups (0: the default grid-stride size), at most PUSH32_MAX_GRID
};
hipError_t launch_field32(const DevMesh& m, const double* phi, float4* E, hipStream_t s);
hipError_t launch_push32(const PushArgs32& a, hipStream_t s);
// Workgroups of k_push32 when PushArgs32::grid is 0 (= its partial count).
int push32_grid(int64_t n);
hipError_t launch_deposit32(const DevMesh& m, const DevParticles32& P, int64_t n, const double* species, int S,
                            unsigned long long* acc, hipStream_t s);
// rhs[i] = boundary ? 0 : background[i] - (sum of acc over elements of node i) * 2^-S / 4.
hipError_t launch_node_charge64(const DevMesh& m, const unsigned long long* acc, int S, const double* background,
                                double* rhs, hipStream_t s);
hipError_t launch_charge_bound32(const DevParticles32& P, int64_t n, const double* species, PushPartial* partial,
                                 hipStream_t s);
hipError_t launch_ke32(const DevParticles32& P, int64_t n, const double* species, int S, int want_sum,
                       PushPartial* partial, hipStream_t s);
hipError_t launch_permute32(const DevParticles32& src, const DevParticles32& dst, const int32_t* perm, int64_t n,
                            hipStream_t s);
// FP64 <-> precision-form copies of particles [off, off+n) (to32: P64 -> P32);
// the FP64 reference position is computed from the position in the element.
hipError_t launch_convert(const DevMesh& m, const DevParticles& P64, const DevParticles32& P32, int64_t off, int64_t n,
                          bool to32, hipStream_t s);
hipError_t launch_to_float(const double* src, float* dst, size_t n, hipStream_t s);
// The diagonal-first copy of m's ELL matrix; *missing (zeroed by the caller)
// becomes nonzero when a row has no diagonal entry.
hipError_t launch_ell_diag_first(const DevMesh& m, int32_t* col, double* val64, float* val32, int* missing,
                                 hipStream_t s);
hipError_t launch_erec32(const DevMesh& m, float* out, hipStream_t s);  // FP32 walk records
// Ownership check: adds to *bad the live particles whose element does not
// contain their position under the FP64 barycentric test.
hipError_t launch_check_owner(const DevMesh& m, const DevParticles& P, int64_t n, unsigned long long* bad,
                              hipStream_t s);
hipError_t launch_check_owner32(const DevMesh& m, const DevParticles32& P, int64_t n, unsigned long long* bad,
                                hipStream_t s);

// Mixed-precision CG: FP32 CG inner solves inside FP64 iterative refinement.
struct IRState {
  int done, converged, outer, in_done, in_iters, total_inner, max_iter;
  int in_ok;    // the last inner pass stopped at its tolerance
  int cont_ok;  // ... and its direction may be continued (set between passes)
  int cont;     // the current pass continues the previous direction
  int in_last;  // iterations of the previous pass (parity of its last state words)
  int pad;
  double bb, rr, tol2, in_tol2, in_rho0, in_rho;
  double scale;       // |r| of the current pass (x += scale * d)
  double scale_prev;  // |r| of the previous pass (set between passes)
  double rho_last;    // r.r of the previous pass's last completed iteration
  double rho_c;       // rho_last in the current pass's units
};
struct IRBuffers {
  double *x, *b, *r;              // FP64 solution, right-hand side, residual
  float *d, *p;                   // FP32 inner correction; direction p (persistent solve)
  float *rv[2], *wv[2], *sv[2];   // FP32 r, w = A r, s = A p by iteration parity (persistent solve)
  float4* q[2];                   // the same state as (r, w, s, p) words by parity (per-iteration kernels)
  double* part_a;                 // [2*nchunks]: b.b, r.r of the outer residual
  double* part_b;                 // [4*nchunks]: r.r, w.r of the inner residual, by parity
  double* part_e;                 // [2*nchunks]: r.s, p.s at the start of a continued pass
  double *rho_hist, *alpha_hist;  // [max_iter+1]
  IRState* st;
  int nchunks;
};
// Outer step: r = b - A x (FP64) and the convergence test; starts an inner solve
// on r/|r| unless converged.
hipError_t launch_ir_outer(const DevMesh& m, const Mesh32& m32, IRBuffers& c, bool first, double tol, int max_iter,
                           double in_tol_min, hipStream_t s);
// spec: a launch past the predicted end of the pass (checks the stop first).
hipError_t launch_ir_inner(const DevMesh& m, const Mesh32& m32, IRBuffers& c, int it, bool spec, hipStream_t s);
// x += |r| d (FP64); ends a pass that has not stopped by itself after pass_len
// inner iterations.
hipError_t launch_ir_update(const DevMesh& m, IRBuffers& c, int pass_len, hipStream_t s);
hipError_t launch_ir_setup(IRBuffers& c, double tol, int max_iter, hipStream_t s);
// The whole precision-form refinement solve as one cooperative launch; same
// results as launch_ir_outer / launch_ir_inner / launch_ir_update.
hipError_t launch_ir_persist(const DevMesh& m, const Mesh32& m32, IRBuffers& c, double tol, int max_iter,
                             double in_tol_min, int pass_len, unsigned* epoch, int* err, hipStream_t s);

// ---- multigrid-preconditioned CG of the precision form (k_mg.hip, mg.hpp) ----
// One coarse level on the device: lattice cells, unknowns, the 27-entry ELL
// ([27][n] column-major), the diagonal inverse and three work vectors: b (the
// restricted residual, overwritten by the level's V-cycle result), x = omega
// D^-1 b (the first smoothing step) and t (the residual after it, then the
// prolonged correction).
struct MGLevelDev {
  int cx, cy, cz, n;
  const float* val;  // 27-point stencil, [27][n], entry (di + 1) 9 + (dj + 1) 3 + (dk + 1)
  const float* dinv;
  float *b, *x, *t;
};
struct MGState {
  int done, converged, iters, max_iter;
  double tol2, bb, rr;
  int armed;  // the last recheck launch found the solve converged and tested the true residual
  int pad;
};
struct MGBuffers {
  int nlev = 0;             // coarse levels 1..nlev; the last is solved by its dense inverse
  int nx = 0, ny = 0, nz = 0;
  int nchunks = 0;          // level 0 row chunks of BLOCK rows (partials are indexed by chunk)
  const MGLevelDev* lev = nullptr;            // device [nlev + 1] (entry 0 unused)
  MGLevelDev host_lev[MG_MAX_LEVELS + 1] = {};  // the same on the host (grid sizes)
  const float* dinv0 = nullptr;  // [nn]
  const float* cinv = nullptr;   // [n_nlev][n_nlev]
  float *b0 = nullptr, *x0 = nullptr, *t0 = nullptr;  // level 0 vectors (b0 also receives z)
  double *r = nullptr, *q = nullptr;
  double* p[2] = {nullptr, nullptr};  // search directions by iteration parity
  double *part_bb = nullptr, *part_rr = nullptr, *part_rz = nullptr, *part_pq = nullptr;
  double* rz_hist = nullptr;  // [max_iter + 1]
  MGState* st = nullptr;
};
// Resets the state, r = b - A x (FP64), and the V-cycle of iteration 0 (its
// first kernel tests |r| < tol |b| and the iteration limit).
hipError_t launch_mg_start(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, double tol,
                           int max_iter, hipStream_t s);
// Iteration it: q = A p (p = z + beta p_old; beta = 0 when restart), the x and r
// update, then the V-cycle of iteration it + 1 (which first tests convergence).
hipError_t launch_mg_iter(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, int it,
                          bool restart, hipStream_t s);
// Queued after every batch of iterations: when the recursive residual has met the
// tolerance (done and converged), reopens the solve, recomputes r = b - A x in
// FP64 and tests it (done again if it meets the tolerance or the limit) and sets
// armed; otherwise clears armed and leaves the solve as it is.
hipError_t launch_mg_recheck(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b,
                             hipStream_t s);
// The V-cycle of iteration it on the current residual (after a failed recheck:
// the iteration then restarts from it).
hipError_t launch_mg_vcycle(const DevMesh& m, const Mesh32& m32, MGBuffers& g, double* x, const double* b, int it,
                            hipStream_t s);

struct SortBuffers {
  int32_t* count;    // [ne+1]
  int32_t* start;    // [ne+1]
  int32_t* cursor;   // [ne]
  int32_t* blocksum; // scan scratch
  int scan_blocks;
};
// Counting sort by element, as a permutation: count (unless counts_ready), scan
// into sb.start, then perm[start[e] + rank] = position of each live particle.
// by_bucket: the key is m.bucket[elem] instead of elem (counts and starts are
// then per bucket).
hipError_t launch_rank(const DevMesh& m, const DevParticles& P, int64_t n, bool by_bucket, bool counts_ready,
                       SortBuffers& sb, int32_t* perm, hipStream_t s);
// dst[i] = src[perm[i]] for every field, i < n.
hipError_t launch_permute(const DevParticles& src, const DevParticles& dst, const int32_t* perm, int64_t n,
                          hipStream_t s);

struct FillArgs {
  DevMesh m;
  DevParticles P;
  int64_t offset;
  int32_t e0, ecount;  // filled elements [e0, e0 + ecount)
  int nppe, type;
  double density, vel[3], temp, pert_amp, pert_k, lo0, mass;
  uint64_t seed_offset;
  minipic::ParticleListView view;  // the list the fill appends to (push_back; its count is set to offset)
};
// Appends ecount * nppe particles through a.view.push_back; the view's
// particle counter is first set to a.offset.
hipError_t launch_fill(const FillArgs& a, hipStream_t s);
// View counters := {n, 0, 0, 0}.
hipError_t launch_view_reset(unsigned long long* counters, int64_t n, hipStream_t s);
// P.elem[i] = P.r[i].elem for i < n.
hipError_t launch_elem_copy(const DevParticles& P, int64_t n, hipStream_t s);

// ---- batching form (k_batch.hip) ----
struct BatchSolveArgs {
  DevMesh m;
  const float* val32;       // FP32 stiffness values (ELL layout of m)
  const int32_t* node_off;  // [nprob+1] node range of each problem
  double *x, *b, *r64;      // potential, right-hand side, FP64 residual
  float *rf, *wf, *pf, *sf, *df;  // inner CG vectors
  int* iters;               // [nprob] inner iterations of the last solve
  int* converged;           // [nprob]
  double* relres;           // [nprob] |r| / |b| at the end
  double tol2, in_tol_min;
  int max_iter;
  int max_nodes;            // nodes of the largest problem
};
// The batching form's multigrid solve (k_batch_mg.hip): conjugate gradients in
// FP64 preconditioned by an FP32 V-cycle, one workgroup per problem, each
// problem with its own hierarchy (mg_build on its brick, coarsest level at most
// BATCH_MG_DIRECT_MAX unknowns) in pooled device arrays.
constexpr int BATCH_MG_DIRECT_MAX = 128;
struct BatchMGProb {
  int nx, ny, nz;    // cells of the problem's brick (level 0)
  int r0, nn;        // its node range in the union
  int lev0, nlev;    // its coarse levels 1..nlev are lev[lev0 .. lev0 + nlev - 1]
  int64_t cinv_off;  // inverse of its coarsest operator at cinv + cinv_off (column-major)
  int64_t v0_off;    // its level-0 stencil values at val0_* + v0_off ([s0][nn])
  int s0;            // stencil slots of level 0
  int off0[27];      // node offset of each slot's column
};
struct BatchMGArgs {
  DevMesh m;                 // diagonal-first ELL, FP64 values
  const float* val32;        // FP32 values of the same ELL
  const BatchMGProb* prob;   // [nprob]
  const MGLevelDev* lev;     // the coarse levels of every problem
  const float* dinv0;        // [nn] level-0 diagonal inverse of every problem
  const float* cinv;         // the coarsest inverses
  const double* val0_64;     // level-0 stencil values of every problem, FP64
  const float* val0_32;      // the same in FP32
  double* x;                 // potential (warm start)
  const double* b;           // right-hand side
  double *r, *p, *q;         // FP64 CG vectors [nn]
  float *b0, *x0, *t0;       // level-0 V-cycle vectors [nn]; b0 holds z after a cycle
  int* iters;                // [nprob] iterations of the last solve
  int* converged;            // [nprob]
  double* relres;            // [nprob] |r| / |b| at the end
  double tol2;
  int max_iter;
};
// Every problem's multigrid solve, one workgroup of `threads` (256, 512 or 1024) per problem;
// max_slots: the largest level-0 stencil (<= 27).
hipError_t launch_batch_mg(const BatchMGArgs& a, int nprob, int threads, int max_slots, hipStream_t s);
// LDS capacities (nodes) of the batch solve's inner residual: problems up to
// BATCH_LDS_SMALL nodes (64 KB) or BATCH_LDS_LARGE (144 KB); larger ones keep
// it in device memory.
constexpr int BATCH_LDS_SMALL = 16384;
constexpr int BATCH_LDS_LARGE = 36864;
// Every problem's refinement solve, one workgroup of `threads` (256, 512 or 1024) per problem.
hipError_t launch_batch_ir(const BatchSolveArgs& a, int nprob, int threads, hipStream_t s);
// Per-problem kinetic energy (exact fixed-point sum at 2^S) and live count over
// the particle ranges poff[p]..poff[p+1], cut into chunks of STAT_CHUNK
// particles: chunk_off[p] is problem p's first chunk, chunk_off[nprob] ==
// nchunks; part holds one partial per chunk.
constexpr int STAT_CHUNK = 8192;
struct StatPartial {
  unsigned long long lo;
  long long hi;
  long long n;
  long long pad;
};
hipError_t launch_batch_stats(const Rec32* r, const int64_t* poff, const int32_t* chunk_off, int nchunks,
                              const double* species, int S, StatPartial* part, double* ke, int64_t* count, int nprob,
                              hipStream_t s);

}  // namespace mpic
