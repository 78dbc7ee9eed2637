// Precision form (MPIC_FORM_MIXED) of the library's host side: FP32 particle
// sets, the 64-bit fixed-point deposit, the FP32-in-FP64 refinement solve and the
// fused time step over them (PIC::time_step, PIC.cpp:245-267).  The FP64 particle
// set stays allocated as the staging area of fills, uploads and downloads.
#include <algorithm>
#include <cmath>

#include "mpic_ctx.hpp"

namespace mpic {

int update_charge_scale(mpic_ctx* c);
int invalidate_deposit(mpic_ctx* c);

namespace {

// Inner iterations of one FP32 pass before the FP64 residual replaces the
// recursive one (reliable update; the direction is kept across passes).  A
// pass ends earlier only at the final tolerance (or IR_IN_TOL_MIN, far below
// what FP32 reaches in IR_PASS_LEN iterations).  A CPU model at Kuhn 64^3
// (warm start, tol 1e-10) takes 269 inner iterations for fixed passes of 30 to
// 60 (296 at 80) against 259 for FP64 CG; a fixed length lets the launches of
// several passes be queued without reading the device state in between.
constexpr int IR_PASS_LEN = 50;
constexpr double IR_IN_TOL_MIN = 1e-6;
// Launches of the expected final pass past its expected stop + IR_SPEC_SLACK
// first test whether the pass has stopped: about 3 us when it has (9 us
// without the early test) and 1-2 us more than a plain launch when it has not.
constexpr int IR_SPEC_SLACK = 2;

int alloc_particles32(mpic_ctx* c) {
  if (c->cap == 0 || c->pmem32[0]) return MPIC_OK;
  const int64_t cap = c->cap;
  for (int k = 0; k < 2; ++k) {
    char* base = nullptr;
    HIPCHK(c, hipMalloc((void**)&base, (sizeof(Rec32) + sizeof(int32_t)) * (size_t)cap + 1024));
    c->pmem32[k] = base;
    DevParticles32& P = c->P32[k];
    P.cap = cap;
    P.r = (Rec32*)base;
    P.elem = (int32_t*)(P.r + cap);
  }
  c->cur32 = 0;
  return MPIC_OK;
}

void record_times(mpic_ctx* c) {
  c->ms[T_DEPOSIT] = telapsed(c, T_DEPOSIT);
  c->ms[T_SOLVE] = telapsed(c, T_SOLVE);
  c->ms[T_PUSH] = telapsed(c, T_PUSH);
  c->ms[T_PUSH_KERNEL] = telapsed(c, T_PUSH_KERNEL);
  c->ms[T_KE] = telapsed(c, T_KE);
  c->ms[T_FIELD] = telapsed(c, T_FIELD);
  if (c->sort_timed) {
    c->ms[T_SORT] = telapsed(c, T_SORT);
    c->sort_timed = false;
  }
}

// The multigrid hierarchy of the solve on the device, once per context; the
// host copy of the matrix is released afterwards.  A mesh without a lattice
// (or one that does not coarsen) keeps the refinement CG.
int mg_prepare(mpic_ctx* c) {
  c->mg_tried = true;
  if (c->mg_rowptr.empty()) {
    c->mg_why = c->cfg.multigrid ? "the mesh descriptor names no lattice" : "multigrid switched off";
    return MPIC_OK;
  }
  MGHost h;
  std::string why;
  const bool ok = mg_build(c->lattice[0], c->lattice[1], c->lattice[2], c->m.nn, c->mg_bnd.data(),
                           c->mg_rowptr.data(), c->mg_col.data(), c->mg_val.data(), h, why);
  std::vector<int32_t>().swap(c->mg_rowptr);
  std::vector<int32_t>().swap(c->mg_col);
  std::vector<double>().swap(c->mg_val);
  std::vector<uint8_t>().swap(c->mg_bnd);
  if (!ok || h.levels.size() > (size_t)MG_MAX_LEVELS) {
    c->mg_why = ok ? "too many levels" : why;
    return MPIC_OK;
  }
  MGBuffers& g = c->mg;
  g.nx = h.nx;
  g.ny = h.ny;
  g.nz = h.nz;
  g.nlev = (int)h.levels.size();
  g.nchunks = c->cg.nchunks;
  const int32_t nn = c->m.nn;
  int rc;
  for (int l = 1; l <= g.nlev; ++l) {
    const MGHostLevel& H = h.levels[l - 1];
    MGLevelDev& L = g.host_lev[l];
    float *val, *dinv;
    L.cx = H.cx;
    L.cy = H.cy;
    L.cz = H.cz;
    L.n = H.n;
    if ((rc = dupload(c, &val, H.val.data(), H.val.size())) ||
        (rc = dupload(c, &dinv, H.dinv.data(), H.dinv.size())) || (rc = dalloc(c, &L.b, (size_t)H.n)) ||
        (rc = dalloc(c, &L.x, (size_t)H.n)) || (rc = dalloc(c, &L.t, (size_t)H.n)))
      return rc;
    L.val = val;
    L.dinv = dinv;
  }
  MGLevelDev* lev;
  float *dinv0, *cinv;
  if ((rc = dupload(c, &lev, g.host_lev, (size_t)g.nlev + 1)) ||
      (rc = dupload(c, &dinv0, h.dinv0.data(), h.dinv0.size())) ||
      (rc = dupload(c, &cinv, h.coarse_inv.data(), h.coarse_inv.size())) || (rc = dalloc(c, &g.b0, (size_t)nn)) ||
      (rc = dalloc(c, &g.x0, (size_t)nn)) || (rc = dalloc(c, &g.t0, (size_t)nn)) || (rc = dalloc(c, &g.r, (size_t)nn)) ||
      (rc = dalloc(c, &g.q, (size_t)nn)) || (rc = dalloc(c, &g.p[0], (size_t)nn)) ||
      (rc = dalloc(c, &g.p[1], (size_t)nn)) || (rc = dalloc(c, &g.part_bb, (size_t)g.nchunks)) ||
      (rc = dalloc(c, &g.part_rr, (size_t)g.nchunks)) || (rc = dalloc(c, &g.part_rz, (size_t)g.nchunks)) ||
      (rc = dalloc(c, &g.part_pq, (size_t)g.nchunks)) || (rc = dalloc(c, &g.st, 1)))
    return rc;
  g.lev = lev;
  g.dinv0 = dinv0;
  g.cinv = cinv;
  if (!c->mg_h && hipHostMalloc((void**)&c->mg_h, sizeof(MGState)) != hipSuccess)
    return fail(c, MPIC_ERR_HIP, "pinned host allocation failed");
  c->mg_ok = true;
  return MPIC_OK;
}

// Multigrid-preconditioned CG (k_mg.hip): the iterations the previous solve
// took are queued at once, then the state is read; after that two at a time.
// When the recursive residual meets the tolerance, the true residual b - A x is
// formed in FP64 and tested; if it misses, the iteration restarts from it.
int mg_solve(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres) {
  MGBuffers& g = c->mg;
  if (max_iter + 2 > c->mg_hist_cap) {
    if (g.rz_hist) (void)hipFree(g.rz_hist);
    g.rz_hist = nullptr;
    HIPCHK(c, hipMalloc((void**)&g.rz_hist, sizeof(double) * (size_t)(max_iter + 2)));
    c->mg_hist_cap = max_iter + 2;
  }
  HIPCHK(c, launch_mg_start(c->m, c->m32, g, c->phi, c->rhs, tol, max_iter, c->stream));
  int it = 0;  // the next iteration to queue
  bool restart = false;
  int rechecks = 0;
  int batch = c->mg_last_iters > 0 ? c->mg_last_iters : 12;
  for (;;) {
    for (int k = 0; k < batch && it < max_iter; ++k, ++it) {
      HIPCHK(c, launch_mg_iter(c->m, c->m32, g, c->phi, c->rhs, it, restart, c->stream));
      restart = false;
    }
    batch = 2;
    // the check of the true residual is queued behind the batch and runs only if
    // the batch met the tolerance: one wait per batch
    HIPCHK(c, launch_mg_recheck(c->m, c->m32, g, c->phi, c->rhs, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->mg_h, g.st, sizeof(MGState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, wait_stream(c->stream));
    const MGState s = *c->mg_h;
    if (!s.armed) {
      if (s.done) break;  // iteration limit or breakdown (done, not converged)
      // the V-cycle after iteration max_iter - 1 stops the solve
      if (it >= max_iter) return fail(c, MPIC_ERR_STATE, "multigrid CG: no stop at the iteration limit");
      continue;
    }
    if (s.done || ++rechecks > MG_MAX_RECHECKS) break;
    // the true residual missed the tolerance: continue from it
    HIPCHK(c, launch_mg_vcycle(c->m, c->m32, g, c->phi, c->rhs, s.iters, c->stream));
    it = s.iters;
    restart = true;
  }
  const MGState& s = *c->mg_h;
  c->mg_last_iters = s.iters;
  c->last_cg_iters = s.iters;
  if (iters) *iters = s.iters;
  if (relres) *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  if (!(s.done && s.converged)) return fail(c, MPIC_ERR_NOT_CONVERGED, "multigrid CG did not converge");
  return MPIC_OK;
}

}  // namespace

int mixed_prepare(mpic_ctx* c) {
  if (!c->mixed_ready) {
    const int32_t nn = c->m.nn, ne = c->m.ne;
    float* erec32;
    float* ell32;
    int32_t* col32;
    double* val64p;
    float4* E32;
    int rc;
    if ((rc = dalloc(c, &erec32, (size_t)ne * 16)) || (rc = dalloc(c, &ell32, (size_t)c->ell_entries)) ||
        (rc = dalloc(c, &col32, (size_t)c->ell_entries)) || (rc = dalloc(c, &val64p, (size_t)c->ell_entries)) ||
        (rc = dalloc(c, &E32, (size_t)ne)) || (rc = dalloc(c, &c->acc64, (size_t)ne)) ||
        (rc = dalloc(c, &c->ir.r, nn)) || (rc = dalloc(c, &c->ir.d, nn)) || (rc = dalloc(c, &c->ir.p, nn)) ||
        (rc = dalloc(c, &c->ir.rv[0], nn)) || (rc = dalloc(c, &c->ir.rv[1], nn)) ||
        (rc = dalloc(c, &c->ir.wv[0], nn)) || (rc = dalloc(c, &c->ir.wv[1], nn)) ||
        (rc = dalloc(c, &c->ir.sv[0], nn)) || (rc = dalloc(c, &c->ir.sv[1], nn)) ||
        (rc = dalloc(c, &c->ir.q[0], nn)) || (rc = dalloc(c, &c->ir.q[1], nn)) ||
        (rc = dalloc(c, &c->ir.part_a, (size_t)2 * c->cg.nchunks)) ||
        (rc = dalloc(c, &c->ir.part_b, (size_t)4 * c->cg.nchunks)) ||
        (rc = dalloc(c, &c->ir.part_e, (size_t)2 * c->cg.nchunks)) || (rc = dalloc(c, &c->ir.st, 1)) ||
        (rc = dalloc(c, &c->bad, 1)))
      return rc;
    if (!c->background.empty() && (rc = dupload(c, &c->bg64, c->background.data(), c->background.size()))) return rc;
    HIPCHK(c, launch_erec32(c->m, erec32, c->stream));
    HIPCHK(c, launch_zero(c->bad, 8, c->stream));
    HIPCHK(c, launch_ell_diag_first(c->m, col32, val64p, ell32, reinterpret_cast<int*>(c->bad), c->stream));
    HIPCHK(c, launch_zero(c->acc64, (size_t)ne * 8, c->stream));
    HIPCHK(c, launch_zero(c->ir.st, sizeof(IRState), c->stream));
    if (!c->ir_h && hipHostMalloc((void**)&c->ir_h, sizeof(IRState)) != hipSuccess)
      return fail(c, MPIC_ERR_HIP, "pinned host allocation failed");
    if (!c->bad_h && hipHostMalloc((void**)&c->bad_h, sizeof(unsigned long long)) != hipSuccess)
      return fail(c, MPIC_ERR_HIP, "pinned host allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->bad_h, c->bad, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, wait_stream(c->stream));
    if (*c->bad_h != 0)
      return fail(c, MPIC_ERR_STATE, "the precision form needs a diagonal entry in every row of the stiffness matrix");
    c->ir.nchunks = c->cg.nchunks;
    c->m32.erec = erec32;
    c->m32.E = E32;
    c->m32.ell_val = ell32;
    c->m32.ell_col = col32;
    c->m32.ell_val64 = val64p;
    c->mixed_ready = true;
  }
  int rc;
  if (!c->mg_tried && (rc = mg_prepare(c))) return rc;
  return alloc_particles32(c);
}

void mixed_free_particles(mpic_ctx* c) {
  for (int k = 0; k < 2; ++k)
    if (c->pmem32[k]) {
      (void)hipFree(c->pmem32[k]);
      c->pmem32[k] = nullptr;
    }
  c->P32[0] = c->P32[1] = DevParticles32{};
  c->state32 = false;
}

int mixed_state(mpic_ctx* c, bool to32) {
  if (to32 == c->state32) return MPIC_OK;
  int rc;
  if ((rc = materialize(c))) return rc;
  if ((rc = invalidate_deposit(c))) return rc;
  if (c->cap > 0) {
    if (to32 && (rc = alloc_particles32(c))) return rc;
    HIPCHK(c, launch_convert(c->m, c->P[c->cur], c->P32[c->cur32], 0, c->n, to32, c->stream));
  }
  c->state32 = to32;
  c->ke_max_prev = 0;
  return update_charge_scale(c);
}

int mixed_import(mpic_ctx* c, int64_t off, int64_t cnt) {
  HIPCHK(c, launch_convert(c->m, c->P[c->cur], c->P32[c->cur32], off, cnt, true, c->stream));
  return MPIC_OK;
}

int mixed_export(mpic_ctx* c) {
  int rc;
  if ((rc = materialize(c))) return rc;
  HIPCHK(c, launch_convert(c->m, c->P[c->cur], c->P32[c->cur32], 0, c->n, false, c->stream));
  return MPIC_OK;
}

// bound = n_live * max |w q|: element sums then stay below 2^62.
int mixed_charge_scale(mpic_ctx* c) {
  PushPartial r{};
  if (c->n > 0) {
    HIPCHK(c, launch_charge_bound32(c->P32[c->cur32], c->n, c->species, c->partial, c->stream));
    int rc = finish_partials(c, partials_for(c->n), &r);
    if (rc) return rc;
  }
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  c->S64 = scale_for_bound64((double)r.crossings * mx);
  return MPIC_OK;
}

int mixed_weight_charge(mpic_ctx* c) {
  if (!c->acc64_deposited)
    HIPCHK(c, launch_deposit32(c->m, c->P32[c->cur32], c->n, c->species, c->S64, c->acc64, c->stream));
  HIPCHK(c, launch_node_charge64(c->m, c->acc64, c->S64, c->bg64, c->rhs, c->stream));
  HIPCHK(c, launch_zero(c->acc64, (size_t)c->m.ne * 8, c->stream));
  c->acc64_deposited = false;
  return MPIC_OK;
}

// Iterative refinement: passes of FP32 inner iterations between FP64
// residuals, queued without reading the device state in between.
int mixed_solve(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres) {
  int rc;
  if ((rc = mixed_prepare(c))) return rc;
  if (c->mg_ok) return mg_solve(c, tol, max_iter, iters, relres);
  c->ir.x = c->phi;
  c->ir.b = c->rhs;
  if (max_iter + 1 > c->ir_hist_cap) {
    if (c->ir.rho_hist) (void)hipFree(c->ir.rho_hist);
    if (c->ir.alpha_hist) (void)hipFree(c->ir.alpha_hist);
    c->ir.rho_hist = c->ir.alpha_hist = nullptr;
    HIPCHK(c, hipMalloc((void**)&c->ir.rho_hist, sizeof(double) * (size_t)(max_iter + 1)));
    HIPCHK(c, hipMalloc((void**)&c->ir.alpha_hist, sizeof(double) * (size_t)(max_iter + 1)));
    c->ir_hist_cap = max_iter + 1;
  }
  // One cooperative launch for the whole refinement solve when the grid is
  // resident at once; the same arithmetic as the per-iteration kernels below.
  if (c->cfg.persistent_solve && !c->persist_off) {
    if ((rc = persist_prepare(c))) return rc;
    if (c->persist_cap32 >= c->ir.nchunks) {
      HIPCHK(c, launch_ir_setup(c->ir, tol, max_iter, c->stream));
      HIPCHK(c, launch_zero(c->sync_epoch, ((size_t)c->ir.nchunks + 2) / 2 * 8, c->stream));
      HIPCHK(c, launch_zero(c->sync_err, 8, c->stream));
      const hipError_t e = launch_ir_persist(c->m, c->m32, c->ir, tol, max_iter, IR_IN_TOL_MIN,
                                             std::min(IR_PASS_LEN, max_iter + 1), c->sync_epoch, c->sync_err,
                                             c->stream);
      if (e == hipSuccess) {
        HIPCHK(c, hipMemcpyAsync(c->ir_h, c->ir.st, sizeof(IRState), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->sync_err_h, c->sync_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (*c->sync_err_h == 0 && c->ir_h->done) {
          const IRState& s = *c->ir_h;
          c->last_cg_iters = s.total_inner;
          if (iters) *iters = s.total_inner;
          if (relres) *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
          if (!s.converged) return fail(c, MPIC_ERR_NOT_CONVERGED, "refinement did not converge");
          return MPIC_OK;
        }
      } else {
        (void)hipGetLastError();
      }
      c->persist_off = true;  // the per-iteration kernels from here on
    }
  }
  // The passes the previous solve took are queued at once (each pass: its
  // inner launches, the update that ends it, the next outer test), then the
  // state is read; after that one pass per read.  Launches after convergence
  // return at once.
  const int L = std::min(IR_PASS_LEN, max_iter + 1);
  const int expect = c->solved_once ? (int)c->last_inner.size() : 1;
  const int final_len = c->last_inner.empty() ? L : c->last_inner.back();
  HIPCHK(c, launch_ir_outer(c->m, c->m32, c->ir, true, tol, max_iter, IR_IN_TOL_MIN, c->stream));
  int pass = 0;
  for (int round = 0;; ++round) {
    const int np = round == 0 ? expect : 1;
    for (int k = 0; k < np; ++k, ++pass) {
      const bool final_pred = pass == expect - 1;
      for (int it = 0; it < L; ++it)
        HIPCHK(c, launch_ir_inner(c->m, c->m32, c->ir, it, final_pred && it > final_len + IR_SPEC_SLACK, c->stream));
      HIPCHK(c, launch_ir_update(c->m, c->ir, L, c->stream));
      HIPCHK(c, launch_ir_outer(c->m, c->m32, c->ir, false, tol, max_iter, IR_IN_TOL_MIN, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->ir_h, c->ir.st, sizeof(IRState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, wait_stream(c->stream));
    if (c->ir_h->done) break;
  }
  const IRState& s = *c->ir_h;
  c->solved_once = true;
  c->last_inner.assign(s.outer, L);
  if (s.outer > 0) c->last_inner.back() = s.in_last;
  c->last_cg_iters = s.total_inner;
  if (iters) *iters = s.total_inner;
  if (relres) *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  if (!s.converged) return fail(c, MPIC_ERR_NOT_CONVERGED, "refinement did not converge");
  return MPIC_OK;
}

int mixed_compute_ke(mpic_ctx* c, double* ke) {
  int rc;
  if ((rc = materialize(c))) return rc;
  PushPartial r{};
  HIPCHK(c, launch_ke32(c->P32[c->cur32], c->n, c->species, 0, 0, c->partial, c->stream));
  if ((rc = finish_partials(c, partials_for(c->n), &r))) return rc;
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  const int S = scale_for_bound((double)r.crossings * mx);
  HIPCHK(c, launch_ke32(c->P32[c->cur32], c->n, c->species, S, 1, c->partial, c->stream));
  if ((rc = finish_partials(c, partials_for(c->n), &r))) return rc;
  *ke = from_fixed(fx128{r.ke_lo, r.ke_hi}, S);
  c->ke_max_prev = mx;
  return MPIC_OK;
}

int mixed_check_owner(mpic_ctx* c, int64_t* bad) {
  int rc;
  if ((rc = materialize(c))) return rc;
  HIPCHK(c, launch_zero(c->bad, 8, c->stream));
  if (c->state32)
    HIPCHK(c, launch_check_owner32(c->m, c->P32[c->cur32], c->n, c->bad, c->stream));
  else
    HIPCHK(c, launch_check_owner(c->m, c->P[c->cur], c->n, c->bad, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->bad_h, c->bad, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *bad = (int64_t)*c->bad_h;
  return MPIC_OK;
}

int mixed_time_step(mpic_ctx* c, double dt, mpic_step_stats* out) {
  int rc;
  mpic_step_stats st{};
  c->ms[T_SORT] = 0;  // set again below if this step reorders
  // 1. charge of the current positions
  tstart(c, T_DEPOSIT);
  if ((rc = mixed_weight_charge(c))) return rc;
  tstop(c, T_DEPOSIT);
  // 2. refinement solve, warm start from the previous potential
  tstart(c, T_SOLVE);
  rc = mixed_solve(c, c->cfg.cg_tol, c->cfg.cg_max_iter, &st.cg_iters, &st.cg_relres);
  tstop(c, T_SOLVE);
  if (rc) return rc;
  // 3. per-tet field, then the fused FP32 kick + walk + FP64 location + deposit + KE
  if (!(c->ke_max_prev > 0) && c->n > 0) {
    PushPartial r{};
    HIPCHK(c, launch_ke32(c->P32[c->cur32], c->n, c->species, 0, 0, c->partial, c->stream));
    if ((rc = finish_partials(c, partials_for(c->n), &r))) return rc;
    std::memcpy(&c->ke_max_prev, &r.ke_max_bits, 8);
  }
  const bool rank_after = c->cfg.sort_period > 0 && (c->steps + 1) % c->cfg.sort_period == 0;
  PushArgs32 a{};
  a.m = c->m;
  a.m32 = c->m32;
  a.P = c->P32[c->cur32];
  a.Q = a.P;
  a.n = c->n;
  a.dt = (float)dt;
  a.scale = 1.0f;
  a.kill = c->cfg.kill_on_wall;
  a.S_charge = c->S64;
  a.S_ke = scale_for_bound((double)(c->perm_pending ? c->perm_n : c->n) * c->ke_max_prev * 256.0);
  a.acc = c->acc64;
  a.partial = c->partial;
  a.species = c->species;
  if (c->perm_pending) {
    a.perm = c->perm;
    a.Q = c->P32[1 - c->cur32];
    a.n = c->perm_n;
  }
  if (rank_after) {
    HIPCHK(c, launch_zero(c->sb.count, (size_t)(((int64_t)c->m.nbuckets + 2) / 2) * 8, c->stream));
    a.count = c->sb.count;
  }
  tstart(c, T_PUSH);
  tstart(c, T_FIELD);
  HIPCHK(c, launch_field32(c->m, c->phi, c->m32.E, c->stream));
  tstop(c, T_FIELD);
  tstart(c, T_PUSH_KERNEL);
  HIPCHK(c, launch_push32(a, c->stream));
  tstop(c, T_PUSH_KERNEL);
  if (c->perm_pending) {
    c->cur32 = 1 - c->cur32;
    c->n = c->perm_n;
    c->perm_pending = false;
  }
  // the push's partials (over its grid) and, on a reordering step, the ranking
  // by sort bucket from the push's counts, then one wait for both
  PushPartial r{};
  if (a.n > 0) {
    HIPCHK(c, launch_push_finish(c->partial, push32_grid(a.n), c->finish, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->finish_h, c->finish, sizeof(PushPartial), hipMemcpyDeviceToHost, c->stream));
  }
  tstop(c, T_PUSH);
  if (c->sort_timed) {
    c->ms[T_SORT] = telapsed(c, T_SORT);
    c->sort_timed = false;
  }
  if (rank_after && (rc = rank_enqueue(c, true, true))) return rc;
  HIPCHK(c, wait_stream(c->stream));
  if (a.n > 0) r = *c->finish_h;
  c->acc64_deposited = true;
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  tstart(c, T_KE);
  const int S_safe = scale_for_bound((double)c->n * mx);
  if (a.S_ke > S_safe) {
    PushPartial k{};
    HIPCHK(c, launch_ke32(c->P32[c->cur32], c->n, c->species, S_safe, 1, c->partial, c->stream));
    if ((rc = finish_partials(c, partials_for(c->n), &k))) return rc;
    st.kinetic_energy = from_fixed(fx128{k.ke_lo, k.ke_hi}, S_safe);
  } else {
    st.kinetic_energy = from_fixed(fx128{r.ke_lo, r.ke_hi}, a.S_ke);
  }
  tstop(c, T_KE);
  c->ke_max_prev = mx;
  st.move = mpic_move_stats{r.crossings, r.killed, r.capped, r.nonpos, r.reflections, r.relocated};
  c->steps++;
  // 4. reordering by sort bucket (ranked above), also after kills
  const int64_t n_live = c->n - r.killed;
  if (rank_after) {
    rank_finish(c, true);
    if (c->perm_n != n_live) return fail(c, MPIC_ERR_STATE, "ranking: live count differs from the push's");
  } else if (r.killed > 0) {
    if ((rc = rank_now(c, true, false, n_live))) return rc;
  }
  HIPCHK(c, wait_stream(c->stream));
  record_times(c);
  st.num_particles = c->perm_pending ? c->perm_n : c->n;
  st.charge_scale_bits = c->S64;
  st.ms_deposit = c->ms[T_DEPOSIT];
  st.ms_solve = c->ms[T_SOLVE];
  st.ms_push = c->ms[T_PUSH];
  st.ms_sort = c->ms[T_SORT];
  st.ms_ke = c->ms[T_KE];
  st.ms_field = c->ms[T_FIELD];
  if (out) *out = st;
  // The same contract as the FP64 form: the step completes (a particle stopped
  // by the crossing cap stays where the cap found it, inside its element after
  // the FP64 location test), the statistics count it, and the call reports it.
  if (r.capped > 0) return fail(c, MPIC_ERR_WALK_CAP, "a particle exceeded the crossing cap");
  return MPIC_OK;
}

}  // namespace mpic
