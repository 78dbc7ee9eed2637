// Host-side context of the library: device memory, stream, timers and the
// state of both forms.  Shared by ctx.hip (C ABI, FP64 form) and ctx_mixed.hip
// (precision form).
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "mpic_internal.hpp"

namespace mpic {
enum Timer { T_DEPOSIT, T_SOLVE, T_PUSH, T_SORT, T_KE, T_PUSH_KERNEL, T_FIELD, T_COUNT };
std::string& global_error();
}  // namespace mpic

struct mpic_ctx {
  using DevMesh = mpic::DevMesh;
  using DevParticles = mpic::DevParticles;
  using DevParticles32 = mpic::DevParticles32;
  using Mesh32 = mpic::Mesh32;
  using fx128 = mpic::fx128;
  using PushPartial = mpic::PushPartial;
  using CGBuffers = mpic::CGBuffers;
  using CGState = mpic::CGState;
  using SortBuffers = mpic::SortBuffers;
  using IRBuffers = mpic::IRBuffers;
  using IRState = mpic::IRState;

  int device = 0;
  hipStream_t stream = nullptr;
  mpic_config cfg{};
  std::string err;
  std::vector<void*> allocs;  // mesh-lifetime device buffers

  DevMesh m{};
  bool periodic = false;  // the mesh joins faces across a period (BiCGStab, FP64 form only)
  std::vector<double> background;  // host copy
  double bg_abs_sum = 0;
  fx128* bgfix = nullptr;
  double* species = nullptr;  // [48]

  // particles (two SoA sets per form: the fused push writes the sorted output
  // into the other one).  The live state is P[cur] (FP64) or P32[cur32]
  // (precision form, state32); n counts its positions.
  int64_t cap = 0, n = 0;
  int32_t* perm = nullptr;  // [cap] sort order of the live set from the last ranking (slot -> position)
  bool perm_pending = false;
  bool perm_by_bucket = false;  // perm orders by sort bucket, not by element
  int64_t perm_n = 0;       // live particles listed in perm
  DevParticles P[2]{};
  int cur = 0;
  void* pmem[2] = {nullptr, nullptr};
  double* E = nullptr;   // transient [3][cap]
  double* vr = nullptr;  // transient [3][cap]
  double* tr = nullptr;  // transient [cap]: time_remaining of the device view
  // device ParticleList view: counters [4] (particles, deleted, migrate
  // entries, dropped appends) and the migrate list
  void* noalloc_scratch = nullptr;  // this device's static scratch of mpic_solve_noalloc (null: unavailable)
  unsigned long long* view_counters = nullptr;
  int64_t* migrate_index = nullptr;
  int32_t* migrate_proc = nullptr;
  PushPartial* partial = nullptr;
  PushPartial* finish = nullptr;   // device [1]
  PushPartial* finish_h = nullptr; // pinned host [1]

  // fields
  double* phi = nullptr;
  double* rhs = nullptr;
  unsigned long long* acc = nullptr;  // [ne*16] element charge digits
  CGBuffers cg{};
  CGState* cg_h = nullptr;  // pinned
  // persistent solves: grid epochs and error flag (device), error flag (pinned),
  // resident-grid capacity (-1 = not yet queried), disabled after a failed launch
  unsigned* sync_epoch = nullptr;
  int* sync_err = nullptr;
  int* sync_err_h = nullptr;
  int persist_cap = -1, persist_cap32 = 0;
  bool persist_off = false;
  int last_cg_iters = 0;
  int rho_hist_cap = 0;
  // BiCGStab (allocated on first use)
  mpic::BiBuffers bi{};
  mpic::BiState* bi_h = nullptr;  // pinned
  int bi_hist_cap = 0;

  SortBuffers sb{};
  int32_t* start_h = nullptr;  // pinned [1]

  // precision form's multigrid solve: the host CSR kept from mpic_create for the
  // hierarchy's build (done at the first preparation of the precision form)
  int lattice[3] = {0, 0, 0};
  std::vector<int32_t> mg_rowptr, mg_col;
  std::vector<double> mg_val;
  std::vector<uint8_t> mg_bnd;
  bool mg_tried = false;  // the build ran (mg_ok: it succeeded)
  bool mg_ok = false;
  std::string mg_why;     // why multigrid does not apply
  mpic::MGBuffers mg{};
  mpic::MGState* mg_h = nullptr;  // pinned
  int mg_hist_cap = 0;
  int mg_last_iters = 0;

  // precision form
  bool mixed_ready = false;  // tables allocated
  bool state32 = false;      // the live particle state is P32[cur32]
  DevParticles32 P32[2]{};
  int cur32 = 0;
  void* pmem32[2] = {nullptr, nullptr};
  Mesh32 m32{};
  int64_t ell_entries = 0;
  unsigned long long* acc64 = nullptr;  // [ne] element charge sums
  bool acc64_deposited = false;
  double* bg64 = nullptr;               // [nn] background, or null
  int S64 = 40;
  IRBuffers ir{};
  IRState* ir_h = nullptr;              // pinned
  int ir_hist_cap = 0;
  std::vector<int> last_inner;          // inner iterations per pass of the last solve
  bool solved_once = false;             // last_inner holds a solve's passes
  unsigned long long* bad = nullptr;    // device counter
  unsigned long long* bad_h = nullptr;  // pinned

  // state
  bool acc_deposited = false;
  bool sort_timed = false;  // T_SORT events hold a ranking not yet read
  int S = 100;
  double ke_max_prev = 0;
  int64_t steps = 0;
  hipEvent_t ev[2 * mpic::T_COUNT] = {};
  double ms[mpic::T_COUNT] = {};
};

namespace mpic {

inline int fail(mpic_ctx* c, int code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    global_error() = msg;
  return code;
}

#define HIPCHK(c, call)                                                                           \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(c, MPIC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(mpic_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  if (count == 0) count = 1;
  HIPCHK(c, hipMalloc(&q, count * sizeof(T)));
  c->allocs.push_back(q);
  *p = (T*)q;
  return MPIC_OK;
}

template <class T>
int dupload(mpic_ctx* c, T** p, const T* host, size_t count) {
  int rc = dalloc(c, p, count);
  if (rc) return rc;
  if (count) HIPCHK(c, hipMemcpy(*p, host, count * sizeof(T), hipMemcpyHostToDevice));
  return MPIC_OK;
}

inline void tstart(mpic_ctx* c, int t, hipStream_t s = nullptr) { (void)hipEventRecord(c->ev[2 * t], s ? s : c->stream); }
inline void tstop(mpic_ctx* c, int t, hipStream_t s = nullptr) { (void)hipEventRecord(c->ev[2 * t + 1], s ? s : c->stream); }
inline double telapsed(mpic_ctx* c, int t) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, c->ev[2 * t], c->ev[2 * t + 1]) != hipSuccess) return 0;
  return ms;
}

// ---- shared steps (ctx.hip) ----
int finish_partials(mpic_ctx* c, int parts, PushPartial* out);
int materialize(mpic_ctx* c);
int rank_now(mpic_ctx* c, bool by_bucket, bool counts_ready, int64_t n_live);
int rank_enqueue(mpic_ctx* c, bool by_bucket, bool counts_ready);
// The per-iteration FP64 CG on buffers b (x = phi, b = rhs of c); allocates nothing.
int cg_run(mpic_ctx* c, CGBuffers& b, double tol, int max_iter, int* iters, double* relres);
// Device address of the allocation-free solve's static scratch on `device`
// (queried once per device, with it current); null if the query failed.
void* noalloc_prepare(int device);
// The device ParticleList view of the FP64 set P[cur] (counters not reset).
minipic::ParticleListView particle_view(mpic_ctx* c);
void rank_finish(mpic_ctx* c, bool by_bucket);
hipError_t wait_stream(hipStream_t s);
// Buffers and capacities of the persistent solves (once per context).
int persist_prepare(mpic_ctx* c);

// ---- precision form (ctx_mixed.hip) ----
// Allocates the precision-form tables and particle sets (idempotent).
int mixed_prepare(mpic_ctx* c);
// Moves the live particle state to the FP32 set (to32) or back to the FP64 set.
int mixed_state(mpic_ctx* c, bool to32);
// Converts particles [off, off+cnt) of the FP64 set into the FP32 set.
int mixed_import(mpic_ctx* c, int64_t off, int64_t cnt);
// The FP32 state widened into the FP64 set (which then mirrors it), for reads.
int mixed_export(mpic_ctx* c);
int mixed_charge_scale(mpic_ctx* c);
int mixed_weight_charge(mpic_ctx* c);
int mixed_solve(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres);
int mixed_time_step(mpic_ctx* c, double dt, mpic_step_stats* out);
int mixed_compute_ke(mpic_ctx* c, double* ke);
int mixed_check_owner(mpic_ctx* c, int64_t* bad);
void mixed_free_particles(mpic_ctx* c);

}  // namespace mpic
