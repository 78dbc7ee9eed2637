// C ABI of the MI355X MiniPIC library (include/minipic_c.h): context, device
// memory, and the host-side ordering of the timestep phases (PIC::time_step,
// PIC.cpp:245-267).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpic_ctx.hpp"
#include "p1_cubature.hpp"

namespace mpic {
std::string& global_error() {
  static thread_local std::string e;
  return e;
}
}  // namespace mpic

using namespace mpic;

namespace mpic {

// Spatial buckets for the timestep's particle ordering: the cells of a uniform
// grid over the node bounding box, about `per` elements each, holding the
// elements whose centroid lies in them, numbered in Morton order of the cell
// (nearby buckets are nearby in memory).  per <= 1: one bucket per element.
std::vector<int32_t> make_buckets(const mpic_mesh_desc* d, const std::vector<int32_t>& conn, int per,
                                  const double lo[3], const double hi[3], int32_t* nb) {
  const int32_t ne = d->num_elems;
  std::vector<int32_t> b(ne);
  if (per <= 1) {
    for (int32_t e = 0; e < ne; ++e) b[e] = e;
    *nb = ne;
    return b;
  }
  double L[3], vol = 1.0;
  for (int a = 0; a < 3; ++a) {
    L[a] = std::max(hi[a] - lo[a], 1e-300);
    vol *= L[a];
  }
  const double side = std::cbrt(vol * per / ne);
  int cells[3];
  for (int a = 0; a < 3; ++a) cells[a] = (int)std::min(1024.0, std::max(1.0, std::round(L[a] / side)));
  std::vector<uint32_t> code(ne);
  for (int32_t e = 0; e < ne; ++e) {
    uint32_t m = 0;
    for (int a = 0; a < 3; ++a) {
      double cen = 0;
      for (int k = 0; k < 4; ++k) cen += d->nodes[(size_t)conn[(size_t)e * 4 + k] * 3 + a];
      int q = (int)std::floor((cen * 0.25 - lo[a]) / L[a] * cells[a]);
      q = std::min(std::max(q, 0), cells[a] - 1);
      for (int bit = 0; bit < 10; ++bit) m |= (uint32_t)((q >> bit) & 1) << (3 * bit + a);
    }
    code[e] = m;
  }
  std::vector<uint32_t> u(code);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  for (int32_t e = 0; e < ne; ++e) b[e] = (int32_t)(std::lower_bound(u.begin(), u.end(), code[e]) - u.begin());
  *nb = (int32_t)u.size();
  return b;
}

PushArgs push_args(mpic_ctx* c) {
  PushArgs a{};
  a.m = c->m;
  a.P = c->P[c->cur];
  a.n = c->n;
  a.phi = c->phi;
  a.dt = 0;
  a.scale = 1.0;
  a.kill = c->cfg.kill_on_wall;
  a.S_charge = c->S;
  a.S_ke = 100;
  a.acc = c->acc;
  a.partial = c->partial;
  a.E = c->E;
  a.vr = c->vr;
  a.species = c->species;
  a.view = particle_view(c);
  return a;
}

minipic::ParticleListView particle_view(mpic_ctx* c) {
  minipic::ParticleListView v{};
  v.rec = reinterpret_cast<minipic::ParticleRecord*>(c->P[c->cur].r);
  v.elem_copy = c->P[c->cur].elem;
  v.E_ = c->E;
  v.v_ref_ = c->vr;
  v.time_remaining_ = c->tr;
  v.counters = c->view_counters;
  v.migrate_index = c->migrate_index;
  v.migrate_proc = c->migrate_proc;
  v.capacity = c->cap;
  v.migrate_capacity = MIGRATE_CAP;
  return v;
}


// Reduce the per-workgroup partials of the last particle launch; synchronises.
// Waits for the stream by polling it: a blocking wait adds tens of
// microseconds of wake-up after the device has finished, on each of the few
// waits of a time step.
hipError_t wait_stream(hipStream_t s) {
  hipError_t e;
  while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
  }
  return e;
}

int finish_partials(mpic_ctx* c, int parts, PushPartial* out) {
  HIPCHK(c, launch_push_finish(c->partial, parts, c->finish, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->finish_h, c->finish, sizeof(PushPartial), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, wait_stream(c->stream));
  *out = *c->finish_h;
  return MPIC_OK;
}

// Charge fixed-point scale from the particles: bound = sum|bg| + n_live * max|w q|.
int update_charge_scale(mpic_ctx* c) {
  if (c->state32) return mixed_charge_scale(c);
  PushPartial r{};
  if (c->n > 0) {
    HIPCHK(c, launch_charge_bound(c->P[c->cur], c->n, c->species, c->partial, c->stream));
    int rc = finish_partials(c, partials_for(c->n), &r);
    if (rc) return rc;
  }
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  const double live = (double)r.crossings;
  const int S = scale_for_bound(c->bg_abs_sum + live * mx);
  if (S != c->S || !c->bgfix) {
    c->S = S;
    std::vector<fx128> h(c->m.nn, fx128{0, 0});
    if (!c->background.empty())
      for (int32_t i = 0; i < c->m.nn; ++i) h[i] = to_fixed(c->background[i], S);
    if (!c->bgfix) {
      int rc = dalloc(c, &c->bgfix, c->m.nn);
      if (rc) return rc;
      c->m.bgfix = c->bgfix;
    }
    HIPCHK(c, hipMemcpy(c->bgfix, h.data(), h.size() * sizeof(fx128), hipMemcpyHostToDevice));
  }
  return MPIC_OK;
}

// Particle positions changed outside the fused path: drop a stale deposit.
int invalidate_deposit(mpic_ctx* c) {
  if (c->acc_deposited) {
    HIPCHK(c, launch_zero(c->acc, (size_t)c->m.ne * 16 * sizeof(unsigned long long), c->stream));
    c->acc_deposited = false;
  }
  if (c->acc64_deposited) {
    HIPCHK(c, launch_zero(c->acc64, (size_t)c->m.ne * sizeof(unsigned long long), c->stream));
    c->acc64_deposited = false;
  }
  return MPIC_OK;
}

int ensure_transients(mpic_ctx* c) {
  if (!c->E) {
    HIPCHK(c, hipMalloc(&c->E, (size_t)3 * c->cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->vr, (size_t)3 * c->cap * sizeof(double)));
  }
  return MPIC_OK;
}

int free_particles(mpic_ctx* c) {
  for (int k = 0; k < 2; ++k)
    if (c->pmem[k]) {
      (void)hipFree(c->pmem[k]);
      c->pmem[k] = nullptr;
    }
  if (c->E) (void)hipFree(c->E);
  if (c->vr) (void)hipFree(c->vr);
  if (c->tr) (void)hipFree(c->tr);
  if (c->perm) (void)hipFree(c->perm);
  c->E = c->vr = c->tr = nullptr;
  c->perm = nullptr;
  c->perm_pending = false;
  mixed_free_particles(c);
  c->cap = c->n = 0;
  return MPIC_OK;
}

// Applies a pending ranking: the particle list becomes element-ordered and
// compacted (dead particles dropped).
int materialize(mpic_ctx* c) {
  if (!c->perm_pending) return MPIC_OK;
  if (c->state32) {
    HIPCHK(c, launch_permute32(c->P32[c->cur32], c->P32[1 - c->cur32], c->perm, c->perm_n, c->stream));
    c->cur32 = 1 - c->cur32;
    c->n = c->perm_n;
    c->perm_pending = false;
    return MPIC_OK;
  }
  HIPCHK(c, launch_permute(c->P[c->cur], c->P[1 - c->cur], c->perm, c->perm_n, c->stream));
  c->cur = 1 - c->cur;
  c->n = c->perm_n;
  c->perm_pending = false;
  return MPIC_OK;
}

// Counting sort of the current list as a permutation (applied by the next
// fused push or by materialize), by element or by sort bucket.  n_live < 0: not
// known, read back from the scan.
// The ranking's launches and the copy of its live count to start_h; the
// ranking is pending once rank_finish has read that count.
int rank_enqueue(mpic_ctx* c, bool by_bucket, bool counts_ready) {
  DevParticles live = c->P[c->cur];  // only the element array is read
  if (c->state32) live.elem = c->P32[c->cur32].elem;
  tstart(c, T_SORT);
  HIPCHK(c, launch_rank(c->m, live, c->n, by_bucket, counts_ready, c->sb, c->perm, c->stream));
  tstop(c, T_SORT);
  const int32_t nkeys = by_bucket ? c->m.nbuckets : c->m.ne;
  HIPCHK(c, hipMemcpyAsync(c->start_h, c->sb.start + nkeys, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  return MPIC_OK;
}

// After the stream has been waited for.
void rank_finish(mpic_ctx* c, bool by_bucket) {
  c->perm_n = *c->start_h;
  c->perm_pending = true;
  c->perm_by_bucket = by_bucket;
  c->sort_timed = true;
}

int rank_now(mpic_ctx* c, bool by_bucket, bool counts_ready, int64_t n_live) {
  int rc;
  if ((rc = rank_enqueue(c, by_bucket, counts_ready))) return rc;
  HIPCHK(c, wait_stream(c->stream));
  rank_finish(c, by_bucket);
  if (n_live >= 0 && n_live != c->perm_n)
    return fail(c, MPIC_ERR_STATE, "ranking: live count differs from the push's");
  return MPIC_OK;
}

// Element order (ParticleList::sort + compute_start_particles_by_element).
int do_sort(mpic_ctx* c, int32_t* start_opt) {
  int rc;
  if (c->perm_pending && c->perm_by_bucket && (rc = materialize(c))) return rc;
  if (!c->perm_pending && (rc = rank_now(c, false, false, -1))) return rc;
  if ((rc = materialize(c))) return rc;
  if (start_opt)
    HIPCHK(c, hipMemcpyAsync(start_opt, c->sb.start, ((size_t)c->m.ne + 1) * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ms[T_SORT] = telapsed(c, T_SORT);
  c->sort_timed = false;
  return MPIC_OK;
}

int do_weight_charge(mpic_ctx* c) {
  if (!c->acc_deposited) {
    HIPCHK(c, launch_deposit(c->m, c->P[c->cur], c->n, c->species, c->S, c->acc, c->stream));
  }
  HIPCHK(c, launch_node_charge(c->m, c->acc, c->S, c->rhs, c->stream));
  c->acc_deposited = false;
  return MPIC_OK;
}

int finish_solve(mpic_ctx* c, int* iters, double* relres);

// Buffers of the persistent solves and the resident-grid capacity (once).
int persist_prepare(mpic_ctx* c) {
  if (c->persist_cap >= 0) return MPIC_OK;
  int rc;
  if ((rc = dalloc(c, &c->sync_epoch, (size_t)c->cg.nchunks + 2)) || (rc = dalloc(c, &c->sync_err, 2))) return rc;
  if (hipHostMalloc((void**)&c->sync_err_h, sizeof(int)) != hipSuccess)
    return fail(c, MPIC_ERR_HIP, "pinned host allocation failed");
  c->persist_cap = persist_capacity(0);
  c->persist_cap32 = persist_capacity(1);
  return MPIC_OK;
}

int do_solve(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres) {
  c->cg.x = c->phi;
  c->cg.b = c->rhs;
  if (max_iter + 1 > c->rho_hist_cap) {
    if (c->cg.rho_hist) (void)hipFree(c->cg.rho_hist);
    c->cg.rho_hist = nullptr;
    HIPCHK(c, hipMalloc((void**)&c->cg.rho_hist, sizeof(double) * (size_t)(max_iter + 1)));
    c->rho_hist_cap = max_iter + 1;
  }
  // One cooperative launch for the whole solve when the grid is resident at once.
  if (c->cfg.persistent_solve && !c->persist_off) {
    int rc = persist_prepare(c);
    if (rc) return rc;
    if (c->persist_cap >= c->cg.nchunks) {
      HIPCHK(c, launch_cg_setup(c->cg, tol, max_iter, c->stream));
      HIPCHK(c, launch_zero(c->sync_epoch, ((size_t)c->cg.nchunks + 2) / 2 * 8, c->stream));
      HIPCHK(c, launch_zero(c->sync_err, 8, c->stream));
      const hipError_t e = launch_cg_persist(c->m, c->cg, tol, max_iter, c->sync_epoch, c->sync_err, c->stream);
      if (e == hipSuccess) {
        HIPCHK(c, hipMemcpyAsync(c->cg_h, c->cg.st, sizeof(CGState), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->sync_err_h, c->sync_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (*c->sync_err_h == 0 && c->cg_h->done) return finish_solve(c, iters, relres);
      } else {
        (void)hipGetLastError();
      }
      c->persist_off = true;  // the per-iteration kernels from here on (x is unchanged)
    }
  }
  return cg_run(c, c->cg, tol, max_iter, iters, relres);
}

// The per-iteration CG (CG_solve.hpp:54-114) on buffers b: iterations are
// launched in batches sized from the previous solve (kernels after
// convergence return at once), the state is read after each batch.  Allocates
// nothing.
int cg_run(mpic_ctx* c, CGBuffers& b, double tol, int max_iter, int* iters, double* relres) {
  b.x = c->phi;
  b.b = c->rhs;
  HIPCHK(c, launch_cg_init(c->m, b, tol, max_iter, c->stream));
  int it = 0;
  int batch = std::max(8, std::min(max_iter + 1, c->last_cg_iters + 1));
  while (true) {
    const int upto = std::min(max_iter + 1, it + batch);
    for (; it < upto; ++it) HIPCHK(c, launch_cg_iter(c->m, b, it, max_iter, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cg_h, b.st, sizeof(CGState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, wait_stream(c->stream));
    if (c->cg_h->done || it > max_iter) break;
    batch = 8;
  }
  return finish_solve(c, iters, relres);
}

// BiCGStab_solve (CG_solve.hpp:115-206) from x = 0.  status 0 converged, 1
// stopped at max_iter, 2 early exit (rho, omega or t.t zero), as the reference.
int do_bicgstab(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres, int* status) {
  BiBuffers& b = c->bi;
  const int32_t nn = c->m.nn;
  const int nch = c->cg.nchunks;
  if (!b.st) {
    int rc;
    if ((rc = dalloc(c, &b.rn, nn)) || (rc = dalloc(c, &b.t, nn)) || (rc = dalloc(c, &b.v0, nn)) ||
        (rc = dalloc(c, &b.v1, nn)) || (rc = dalloc(c, &b.part_bb, nch)) || (rc = dalloc(c, &b.part_r, 2 * nch)) ||
        (rc = dalloc(c, &b.part_bv, nch)) || (rc = dalloc(c, &b.part_rt, 3 * nch)) || (rc = dalloc(c, &b.st, 1)))
      return rc;
    if (hipHostMalloc((void**)&c->bi_h, sizeof(BiState)) != hipSuccess)
      return fail(c, MPIC_ERR_HIP, "pinned host allocation failed");
    b.nchunks = nch;
  }
  if (max_iter + 1 > c->bi_hist_cap) {
    if (b.hist) (void)hipFree(b.hist);
    b.hist = nullptr;
    HIPCHK(c, hipMalloc((void**)&b.hist, sizeof(double) * 3 * (size_t)(max_iter + 1)));
    c->bi_hist_cap = max_iter + 1;
  }
  b.x = c->phi;
  b.b = c->rhs;
  b.r = c->cg.r;
  b.p0 = c->cg.p0;
  b.p1 = c->cg.p1;
  HIPCHK(c, launch_bi_init(c->m, b, tol, max_iter, c->stream));
  int it = 0;
  while (true) {
    const int upto = std::min(max_iter + 1, it + 16);
    for (; it < upto; ++it) HIPCHK(c, launch_bi_iter(c->m, b, it, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->bi_h, b.st, sizeof(BiState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, wait_stream(c->stream));
    if (c->bi_h->done || it > max_iter) break;
  }
  const BiState& s = *c->bi_h;
  c->last_cg_iters = s.iters;
  if (iters) *iters = s.iters;
  if (relres) *relres = s.rho0 > 0 ? std::sqrt(s.resid / s.rho0) : 0.0;
  if (status) *status = s.done ? s.status : 1;
  return MPIC_OK;
}

int finish_solve(mpic_ctx* c, int* iters, double* relres) {
  const CGState& s = *c->cg_h;
  c->last_cg_iters = s.iters;
  if (iters) *iters = s.iters;
  if (relres) *relres = s.rho0 > 0 ? std::sqrt(s.rho / s.rho0) : 0.0;
  if (!s.converged) return fail(c, MPIC_ERR_NOT_CONVERGED, "CG did not converge within max_iter");
  return MPIC_OK;
}

}  // namespace mpic

#define FP64_ONLY \
  "the precision form runs the fused time step; select MPIC_FORM_FP64 for the phase-by-phase API"

extern "C" {

void mpic_config_default(mpic_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->form = MPIC_FORM_FP64;
  cfg->cg_tol = 1e-10;
  cfg->cg_max_iter = 1000;
  cfg->kill_on_wall = 0;
  cfg->sort_period = 1;
  cfg->sort_bucket = 6;  // about one hex cell of the brick splits (5 or 6 tets)
  cfg->persistent_solve = 0;
  cfg->multigrid = 1;
  cfg->num_species = 3;
  for (int t = 0; t < MPIC_MAX_SPECIES; ++t) cfg->species_m[t] = 1.0;
  cfg->species_q[0] = -1.0;  // e-  (particle_type.hpp:69-71)
  cfg->species_q[1] = 1.0;   // e+
  cfg->species_q[2] = 1.0;   // H+
  cfg->species_m[2] = 1837.0;
}

const char* mpic_last_error(const mpic_ctx* ctx) { return ctx ? ctx->err.c_str() : global_error().c_str(); }

int mpic_create(const mpic_mesh_desc* d, const mpic_config* cfg, mpic_ctx** out) {
  if (!d || !out || d->num_nodes <= 0 || d->num_elems <= 0 || !d->nodes || !d->elem_nodes || !d->elem_nbrs ||
      !d->normals || !d->inv_jacobian || !d->det_jacobian || !d->grad_basis || !d->wbv || !d->boundary ||
      !d->row_ptr || !d->col_idx || !d->values)
    return fail(nullptr, MPIC_ERR_ARG, "mpic_create: incomplete mesh description");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, MPIC_ERR_NO_DEVICE, "mpic_create: no HIP device");
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, MPIC_ERR_NO_DEVICE, std::string("mpic_create: device is not gfx950: ") + prop.gcnArchName);

  mpic_ctx* c = new mpic_ctx();
  c->device = dev;
  // the address of the allocation-free solve's scratch on this device (once per device)
  c->noalloc_scratch = noalloc_prepare(dev);
  if (cfg)
    c->cfg = *cfg;
  else
    mpic_config_default(&c->cfg);
  auto bail = [&](int rc) {
    global_error() = c->err;
    mpic_destroy(c);
    return rc;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(MPIC_ERR_HIP);
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return bail(MPIC_ERR_HIP);

  const int32_t nn = d->num_nodes, ne = d->num_elems;
  c->m.nn = nn;
  c->m.ne = ne;
  // Per-element records.
  std::vector<double> geo((size_t)ne * 12), sixodet(ne), erec((size_t)ne * 16, 0.0);
  std::vector<int32_t> conn((size_t)ne * 4), nbrs((size_t)ne * 4);
  for (int64_t e = 0; e < ne; ++e) {
    std::memcpy(&geo[e * 12], d->inv_jacobian + e * 9, 9 * sizeof(double));
    const int32_t n0 = d->elem_nodes[e * 4];
    if (n0 < 0 || n0 >= nn) return bail(fail(c, MPIC_ERR_ARG, "mpic_create: element node out of range"));
    std::memcpy(&geo[e * 12 + 9], d->nodes + (size_t)n0 * 3, 3 * sizeof(double));
    sixodet[e] = 6 / d->det_jacobian[e];  // the (6/ detJ) of PIC.cpp:125
    for (int k = 0; k < 4; ++k) {
      const int32_t v = d->elem_nodes[e * 4 + k], b = d->elem_nbrs[e * 4 + k];
      // b: neighbour, -1 wall, -2 - neighbour across a periodic face
      if (v < 0 || v >= nn || b >= ne || (b < -1 && -2 - (int64_t)b >= ne))
        return bail(fail(c, MPIC_ERR_ARG, "mpic_create: connectivity out of range"));
      if (b < -1) c->periodic = true;
      conn[e * 4 + k] = v;
      nbrs[e * 4 + k] = b;
    }
    // walk record: J^-1 (9), node 0 (3), neighbours (4 x int32 in slots 12-13), 6/detJ
    std::memcpy(&erec[e * 16], &geo[e * 12], 12 * sizeof(double));
    std::memcpy(&erec[e * 16 + 12], &nbrs[e * 4], 4 * sizeof(int32_t));
    erec[e * 16 + 14] = sixodet[e];
  }
  // The weighted basis values of the configurations' meshes are (detJ w) N_i(x_c), each one IEEE product; when
  // every value of this mesh is, bit for bit, the FP64 push forms them from detJ w in the walk record instead of
  // reading 16 doubles per particle (the charge is the same: PIC.cpp:125-130 multiplies by each value).
  bool factored = ne > 0;
  for (int64_t e = 0; e < ne && factored; ++e) {
    const double meas = d->det_jacobian[e] * kCubW;
    for (int i = 0; i < 4 && factored; ++i)
      for (int q = 0; q < 4; ++q) {
        const double v = meas * cub_basis(i, q);
        if (std::memcmp(&v, d->wbv + e * 16 + i * 4 + q, sizeof(double)) != 0) {
          factored = false;
          break;
        }
      }
    erec[e * 16 + 15] = meas;
  }
  c->m.wbv_factored = factored ? 1 : 0;
  // Likewise the physical basis gradients: sum_j J^-1[j][d] kGradRef[i][j] from 0 in order j = 0..2 (PIC::PIC),
  // formed by the FP64 push from the inverse Jacobian of its walk record when they are exactly that.
  bool derived = ne > 0;
  for (int64_t e = 0; e < ne && derived; ++e)
    for (int i = 0; i < 4 && derived; ++i)
      for (int dd = 0; dd < 3; ++dd) {
        double g = 0;
        for (int j = 0; j < 3; ++j) g += d->inv_jacobian[e * 9 + j * 3 + dd] * kGradRef[i][j];
        if (std::memcmp(&g, d->grad_basis + e * 12 + i * 3 + dd, sizeof(double)) != 0) {
          derived = false;
          break;
        }
      }
  c->m.grad_derived = derived ? 1 : 0;
  // Node -> element slot adjacency (slot = e*4 + k), element order.  The charge
  // of a periodic image node goes to its partner's row (PIC.cpp:132-134): the
  // partner gathers the image's slots, the image none.
  std::vector<int32_t> row_of(conn);
  if (d->node_pair) {
    for (int32_t i = 0; i < nn; ++i)
      if (d->node_pair[i] >= nn || (d->node_pair[i] >= 0 && d->node_pair[d->node_pair[i]] >= 0))
        return bail(fail(c, MPIC_ERR_ARG, "mpic_create: node pair out of range"));
    for (auto& v : row_of)
      if (d->node_pair[v] >= 0) {
        v = d->node_pair[v];
        c->periodic = true;
      }
  }
  if (c->periodic && c->cfg.solver != MPIC_SOLVER_BICGSTAB)
    return bail(fail(c, MPIC_ERR_ARG,
                     "mpic_create: a periodic mesh has a non-symmetric stiffness matrix; use MPIC_SOLVER_BICGSTAB "
                     "(the reference's solver) in the FP64 form"));
  // A lattice mesh keeps its CSR on the host for the precision form's multigrid
  // hierarchy (built when that form is first prepared).
  if (c->cfg.multigrid && !c->periodic && d->lattice[0] > 0 && d->lattice[1] > 0 && d->lattice[2] > 0 &&
      (int64_t)(d->lattice[0] + 1) * (d->lattice[1] + 1) * (d->lattice[2] + 1) == nn) {
    for (int a = 0; a < 3; ++a) c->lattice[a] = d->lattice[a];
    c->mg_rowptr.assign(d->row_ptr, d->row_ptr + nn + 1);
    c->mg_col.assign(d->col_idx, d->col_idx + d->nnz);
    c->mg_val.assign(d->values, d->values + d->nnz);
    c->mg_bnd.assign(d->boundary, d->boundary + nn);
  }
  std::vector<int32_t> adj_ptr(nn + 1, 0), adj((size_t)ne * 4);
  for (int64_t s = 0; s < (int64_t)ne * 4; ++s) adj_ptr[row_of[s] + 1]++;
  for (int32_t i = 0; i < nn; ++i) adj_ptr[i + 1] += adj_ptr[i];
  {
    std::vector<int32_t> pos(adj_ptr.begin(), adj_ptr.end() - 1);
    for (int64_t s = 0; s < (int64_t)ne * 4; ++s) adj[pos[row_of[s]]++] = (int32_t)s;
  }
  // Sliced ELL from CSR, each row's entries in their CSR order without the
  // stored exact zeros (the Kuhn split of a uniform brick stores 15 entries a
  // row of which 8 cancel to 0.0 exactly).  A zero entry adds a signed zero to
  // a row sum, which changes nothing but the sign of an exactly zero sum, so
  // every row sum is the CSR's.  The diagonal is always kept.
  const int32_t nsl = (nn + ELL_SLICE - 1) / ELL_SLICE;
  std::vector<int32_t> keep_ptr(nn + 1, 0), keep;
  keep.reserve(d->nnz);
  for (int32_t r = 0; r < nn; ++r) {
    for (int32_t e = d->row_ptr[r]; e < d->row_ptr[r + 1]; ++e) {
      const int32_t cc = d->col_idx[e];
      if (cc < 0 || cc >= nn) return bail(fail(c, MPIC_ERR_ARG, "mpic_create: CSR column out of range"));
      if (d->values[e] != 0.0 || cc == r) keep.push_back(e);
    }
    keep_ptr[r + 1] = (int32_t)keep.size();
  }
  std::vector<int64_t> ell_off(nsl + 1, 0);
  std::vector<int32_t> ell_len(nsl, 0);
  for (int32_t s = 0; s < nsl; ++s) {
    int32_t L = 0;
    for (int l = 0; l < ELL_SLICE; ++l) {
      const int64_t r = (int64_t)s * ELL_SLICE + l;
      if (r < nn) L = std::max(L, keep_ptr[r + 1] - keep_ptr[r]);
    }
    ell_len[s] = L;
    ell_off[s + 1] = ell_off[s] + (int64_t)L * ELL_SLICE;
  }
  c->ell_entries = ell_off[nsl];
  std::vector<int32_t> ell_col(ell_off[nsl]);
  std::vector<double> ell_val(ell_off[nsl]);
  for (int32_t s = 0; s < nsl; ++s)
    for (int l = 0; l < ELL_SLICE; ++l) {
      const int64_t r = (int64_t)s * ELL_SLICE + l;
      for (int k = 0; k < ell_len[s]; ++k) {
        const int64_t at = ell_off[s] + (int64_t)k * ELL_SLICE + l;
        if (r < nn && k < keep_ptr[r + 1] - keep_ptr[r]) {
          const int32_t e = keep[keep_ptr[r] + k];
          ell_col[at] = d->col_idx[e];
          ell_val[at] = d->values[e];
        } else {
          ell_col[at] = r < nn ? (int32_t)r : 0;
          ell_val[at] = 0.0;
        }
      }
    }
  double* dgeo;
  double* derec;
  double* dgrad;
  double* dnormals;
  double* dsix;
  double* ddet;
  double* dwbv;
  double* dnodes;
  int32_t *dconn, *dnbrs, *dadjp, *dadj, *dlen, *dcol;
  int64_t* doff;
  double* dval;
  uint8_t* dbnd;
  int rc;
#define UP(ptr, vec) \
  if ((rc = dupload(c, &ptr, vec, count_))) return bail(rc);
  size_t count_;
  count_ = geo.size(); UP(dgeo, geo.data());
  count_ = erec.size(); UP(derec, erec.data());
  count_ = (size_t)ne * 12; UP(dgrad, d->grad_basis);
  count_ = (size_t)ne * 12; UP(dnormals, d->normals);
  count_ = (size_t)ne; UP(dsix, sixodet.data());
  count_ = (size_t)ne; UP(ddet, d->det_jacobian);
  count_ = (size_t)ne * 16; UP(dwbv, d->wbv);
  count_ = (size_t)nn * 3; UP(dnodes, d->nodes);
  count_ = conn.size(); UP(dconn, conn.data());
  count_ = nbrs.size(); UP(dnbrs, nbrs.data());
  count_ = adj_ptr.size(); UP(dadjp, adj_ptr.data());
  count_ = adj.size(); UP(dadj, adj.data());
  count_ = ell_len.size(); UP(dlen, ell_len.data());
  count_ = ell_off.size(); UP(doff, ell_off.data());
  count_ = ell_col.size(); UP(dcol, ell_col.data());
  count_ = ell_val.size(); UP(dval, ell_val.data());
  count_ = (size_t)nn; UP(dbnd, d->boundary);
#undef UP
  c->m.geo = dgeo;
  c->m.erec = derec;
  c->m.grad = dgrad;
  c->m.normals = dnormals;
  c->m.sixodet = dsix;
  c->m.detj = ddet;
  c->m.wbv = dwbv;
  c->m.nodes = dnodes;
  c->m.conn = reinterpret_cast<const int4*>(dconn);
  c->m.nbrs = reinterpret_cast<const int4*>(dnbrs);
  c->m.adj_ptr = dadjp;
  c->m.adj = dadj;
  c->m.nslices = nsl;
  c->m.ell_len = dlen;
  c->m.ell_off = doff;
  c->m.ell_col = dcol;
  c->m.ell_val = dval;
  c->m.bnd = dbnd;
  c->m.bgfix = nullptr;
  for (int a = 0; a < 3; ++a) {
    c->m.lo[a] = 1e300;
    c->m.hi[a] = -1e300;
  }
  for (int32_t i = 0; i < nn; ++i)
    for (int a = 0; a < 3; ++a) {
      c->m.lo[a] = std::min(c->m.lo[a], d->nodes[(size_t)i * 3 + a]);
      c->m.hi[a] = std::max(c->m.hi[a], d->nodes[(size_t)i * 3 + a]);
    }
  // A particle crossing a periodic face moves to the other side: x -> lo + hi - x
  // (x -> -x on the reference's [-1,1] box, particle_move.hpp:190-191).
  c->m.period_x = c->m.lo[0] + c->m.hi[0];
  {
    int32_t nb = 0;
    std::vector<int32_t> bk = make_buckets(d, conn, c->cfg.sort_bucket, c->m.lo, c->m.hi, &nb);
    int32_t* dbk;
    if ((rc = dupload(c, &dbk, bk.data(), bk.size()))) return bail(rc);
    c->m.bucket = dbk;
    c->m.nbuckets = nb;
  }
  if (d->background) {
    c->background.assign(d->background, d->background + nn);
    for (int32_t i = 0; i < nn; ++i) c->bg_abs_sum += std::fabs(c->background[i]);
  }
  // Species table [q | m | q/m].
  double sp[48];
  for (int t = 0; t < MPIC_MAX_SPECIES; ++t) {
    sp[t] = c->cfg.species_q[t];
    sp[16 + t] = c->cfg.species_m[t];
    sp[32 + t] = c->cfg.species_m[t] != 0 ? c->cfg.species_q[t] / c->cfg.species_m[t] : 0.0;  // particle_type.hpp:51
  }
  if ((rc = dupload(c, &c->species, sp, 48))) return bail(rc);
  // Fields, CG and sort buffers.
  const int nchunks = (nn + BLOCK - 1) / BLOCK;
  if ((rc = dalloc(c, &c->m.ephi, (size_t)ne * 4))) return bail(rc);
  if ((rc = dalloc(c, &c->phi, nn)) || (rc = dalloc(c, &c->rhs, nn)) || (rc = dalloc(c, &c->acc, (size_t)ne * 16)) ||
      (rc = dalloc(c, &c->cg.r, nn)) || (rc = dalloc(c, &c->cg.p0, nn)) || (rc = dalloc(c, &c->cg.p1, nn)) ||
      (rc = dalloc(c, &c->cg.q, nn)) || (rc = dalloc(c, &c->cg.pr[0], nn)) || (rc = dalloc(c, &c->cg.pr[1], nn)) ||
      (rc = dalloc(c, &c->cg.part_bb, nchunks)) ||
      (rc = dalloc(c, &c->cg.part_rr, nchunks)) || (rc = dalloc(c, &c->cg.part_pq, nchunks)) ||
      (rc = dalloc(c, &c->cg.st, 1)) || (rc = dalloc(c, &c->sb.count, (size_t)ne + 2)) ||
      (rc = dalloc(c, &c->sb.start, (size_t)ne + 2)) || (rc = dalloc(c, &c->sb.cursor, (size_t)ne + 2)) ||
      (rc = dalloc(c, &c->sb.blocksum, (size_t)ne / 2048 + 64)) || (rc = dalloc(c, &c->finish, 1)) ||
      (rc = dalloc(c, &c->partial, PARTIAL_SLOTS)) || (rc = dalloc(c, &c->view_counters, 4)) ||
      (rc = dalloc(c, &c->migrate_index, (size_t)MIGRATE_CAP)) || (rc = dalloc(c, &c->migrate_proc, (size_t)MIGRATE_CAP)))
    return bail(rc);
  c->cg.nchunks = nchunks;
  if (hipHostMalloc((void**)&c->cg_h, sizeof(CGState)) != hipSuccess ||
      hipHostMalloc((void**)&c->finish_h, sizeof(PushPartial)) != hipSuccess ||
      hipHostMalloc((void**)&c->start_h, sizeof(int32_t)) != hipSuccess)
    return bail(fail(c, MPIC_ERR_HIP, "pinned host allocation failed"));
  if (hipMemset(c->phi, 0, (size_t)nn * 8) != hipSuccess || hipMemset(c->rhs, 0, (size_t)nn * 8) != hipSuccess ||
      hipMemset(c->acc, 0, (size_t)ne * 16 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->cg.st, 0, sizeof(CGState)) != hipSuccess)
    return bail(fail(c, MPIC_ERR_HIP, "hipMemset failed"));
  if ((rc = update_charge_scale(c))) return bail(rc);
  if (c->cfg.form == MPIC_FORM_MIXED && (rc = mixed_prepare(c))) return bail(rc);
  if (c->cfg.form != MPIC_FORM_FP64 && c->cfg.form != MPIC_FORM_MIXED)
    return bail(fail(c, MPIC_ERR_ARG, "mpic_create: unknown form"));
  if (c->cfg.solver != MPIC_SOLVER_CG && c->cfg.solver != MPIC_SOLVER_BICGSTAB)
    return bail(fail(c, MPIC_ERR_ARG, "mpic_create: unknown solver"));
  if (c->cfg.solver == MPIC_SOLVER_BICGSTAB && c->cfg.form == MPIC_FORM_MIXED)
    return bail(fail(c, MPIC_ERR_ARG, "mpic_create: BiCGStab runs in the FP64 form only"));
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(c, MPIC_ERR_HIP, "setup synchronisation failed"));
  *out = c;
  return MPIC_OK;
}

void mpic_destroy(mpic_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_particles(c);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->cg.rho_hist) (void)hipFree(c->cg.rho_hist);
  if (c->cg_h) (void)hipHostFree(c->cg_h);
  if (c->finish_h) (void)hipHostFree(c->finish_h);
  if (c->ir_h) (void)hipHostFree(c->ir_h);
  if (c->bad_h) (void)hipHostFree(c->bad_h);
  if (c->bi_h) (void)hipHostFree(c->bi_h);
  if (c->bi.hist) (void)hipFree(c->bi.hist);
  if (c->sync_err_h) (void)hipHostFree(c->sync_err_h);
  if (c->ir.rho_hist) (void)hipFree(c->ir.rho_hist);
  if (c->ir.alpha_hist) (void)hipFree(c->ir.alpha_hist);
  if (c->start_h) (void)hipHostFree(c->start_h);
  if (c->mg_h) (void)hipHostFree(c->mg_h);
  if (c->mg.rz_hist) (void)hipFree(c->mg.rz_hist);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mpic_set_form(mpic_ctx* c, int form) {
  if (!c) return MPIC_ERR_ARG;
  if (form != MPIC_FORM_FP64 && form != MPIC_FORM_MIXED) return fail(c, MPIC_ERR_ARG, "mpic_set_form: unknown form");
  if (form == MPIC_FORM_MIXED && c->cfg.solver == MPIC_SOLVER_BICGSTAB)
    return fail(c, MPIC_ERR_ARG, "mpic_set_form: BiCGStab runs in the FP64 form only");
  int rc;
  if (form == MPIC_FORM_MIXED && (rc = mixed_prepare(c))) return rc;
  if ((rc = mixed_state(c, form == MPIC_FORM_MIXED))) return rc;
  c->cfg.form = form;
  return MPIC_OK;
}

int mpic_set_kill_on_wall(mpic_ctx* c, int kill) {
  if (!c) return MPIC_ERR_ARG;
  c->cfg.kill_on_wall = kill ? 1 : 0;
  return MPIC_OK;
}

int mpic_set_species(mpic_ctx* c, int n, const double* q, const double* m) {
  if (!c || n < 1 || n > MPIC_MAX_SPECIES || !q || !m) return fail(c, MPIC_ERR_ARG, "mpic_set_species: bad arguments");
  for (int t = 0; t < n; ++t)
    if (!std::isfinite(q[t]) || !std::isfinite(m[t])) return fail(c, MPIC_ERR_ARG, "mpic_set_species: non-finite value");
  double sp[48];
  for (int t = 0; t < MPIC_MAX_SPECIES; ++t) {
    c->cfg.species_q[t] = t < n ? q[t] : 0.0;
    c->cfg.species_m[t] = t < n ? m[t] : 1.0;
    sp[t] = c->cfg.species_q[t];
    sp[16 + t] = c->cfg.species_m[t];
    sp[32 + t] = c->cfg.species_m[t] != 0 ? c->cfg.species_q[t] / c->cfg.species_m[t] : 0.0;  // particle_type.hpp:51
  }
  c->cfg.num_species = n;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->species, sp, sizeof(sp), hipMemcpyHostToDevice));
  int rc;
  if ((rc = invalidate_deposit(c))) return rc;
  c->ke_max_prev = 0;
  return update_charge_scale(c);
}

int mpic_get_config(mpic_ctx* c, mpic_config* cfg) {
  if (!c || !cfg) return MPIC_ERR_ARG;
  *cfg = c->cfg;
  return MPIC_OK;
}

int mpic_particles_reserve(mpic_ctx* c, int64_t cap) {
  if (!c || cap <= 0 || cap > ((int64_t)1 << 31) - 1) return fail(c, MPIC_ERR_ARG, "mpic_particles_reserve: bad capacity");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_particles(c);
  for (int k = 0; k < 2; ++k) {
    char* base = nullptr;
    HIPCHK(c, hipMalloc((void**)&base, (sizeof(Rec64) + sizeof(int32_t)) * (size_t)cap + 1024));
    c->pmem[k] = base;
    DevParticles& P = c->P[k];
    P.cap = cap;
    P.r = (Rec64*)base;
    P.elem = (int32_t*)(P.r + cap);
  }
  HIPCHK(c, hipMalloc((void**)&c->perm, sizeof(int32_t) * (size_t)cap));
  c->cap = cap;
  c->n = 0;
  c->cur = 0;
  c->perm_pending = false;
  c->acc_deposited = false;
  c->ke_max_prev = 0;
  HIPCHK(c, hipMemset(c->acc, 0, (size_t)c->m.ne * 16 * sizeof(unsigned long long)));
  c->state32 = false;
  if (c->cfg.form == MPIC_FORM_MIXED) {
    int rc = mixed_prepare(c);
    if (rc) return rc;
    c->state32 = true;
  }
  return MPIC_OK;
}

int mpic_particles_upload(mpic_ctx* c, const mpic_soa_host* h, int64_t n) {
  if (!c || !h || n < 0 || !h->x || !h->x_ref || !h->v || !h->elem || !h->weight || !h->type || h->stride < n)
    return fail(c, MPIC_ERR_ARG, "mpic_particles_upload: bad arguments");
  if (n > c->cap) return fail(c, MPIC_ERR_CAPACITY, "mpic_particles_upload: capacity exceeded");
  // Records are packed on the host and copied in one transfer.
  std::vector<Rec64> rec(n);
  for (int64_t p = 0; p < n; ++p) {
    if (h->type[p] < 0 || h->type[p] >= c->cfg.num_species)
      return fail(c, MPIC_ERR_ARG, "mpic_particles_upload: species out of range");
    if (h->elem[p] < -1 || h->elem[p] >= c->m.ne) return fail(c, MPIC_ERR_ARG, "mpic_particles_upload: element out of range");
    Rec64& r = rec[p];
    for (int d = 0; d < 3; ++d) {
      r.x[d] = h->x[d * h->stride + p];
      r.xr[d] = h->x_ref[d * h->stride + p];
      r.v[d] = h->v[d * h->stride + p];
    }
    r.w = h->weight[p];
    r.elem = h->elem[p];
    r.id = h->id ? h->id[p] : (uint32_t)p;
    r.type = (uint32_t)h->type[p];
    r.pad = 0;
  }
  // every check passed: the list is replaced (a pending ranking of the old list is dropped)
  c->perm_pending = false;
  DevParticles& P = c->P[c->cur];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n) {
    HIPCHK(c, hipMemcpy(P.r, rec.data(), n * sizeof(Rec64), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(P.elem, h->elem, n * 4, hipMemcpyHostToDevice));
  }
  c->n = n;
  c->ke_max_prev = 0;
  int rc;
  if (c->state32 && (rc = mixed_import(c, 0, n))) return rc;
  if ((rc = invalidate_deposit(c))) return rc;
  return update_charge_scale(c);
}

int mpic_particles_download(mpic_ctx* c, mpic_soa_host* h, int64_t* n) {
  if (!c || !h) return fail(c, MPIC_ERR_ARG, "mpic_particles_download: bad arguments");
  if (materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (h->stride < c->n) return fail(c, MPIC_ERR_ARG, "mpic_particles_download: stride below the particle count");
  if (c->state32 && mixed_export(c) != MPIC_OK) return MPIC_ERR_HIP;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevParticles& P = c->P[c->cur];
  const int64_t cnt = c->n;
  std::vector<Rec64> rec(cnt);
  if (cnt) HIPCHK(c, hipMemcpy(rec.data(), P.r, cnt * sizeof(Rec64), hipMemcpyDeviceToHost));
  for (int64_t p = 0; p < cnt; ++p) {
    const Rec64& r = rec[p];
    for (int d = 0; d < 3; ++d) {
      if (h->x) h->x[d * h->stride + p] = r.x[d];
      if (h->x_ref) h->x_ref[d * h->stride + p] = r.xr[d];
      if (h->v) h->v[d * h->stride + p] = r.v[d];
    }
    if (h->weight) h->weight[p] = r.w;
    if (h->elem) h->elem[p] = r.elem;
    if (h->id) h->id[p] = r.id;
    if (h->type) h->type[p] = (int32_t)r.type;
  }
  if (n) *n = cnt;
  return MPIC_OK;
}

int mpic_num_particles(mpic_ctx* c, int64_t* n) {
  if (!c || !n) return MPIC_ERR_ARG;
  *n = c->perm_pending ? c->perm_n : c->n;
  return MPIC_OK;
}

int mpic_fill(mpic_ctx* c, int nppe, int type, double density, const double* vel, double temp, uint64_t seed_offset,
              double pert_amp) {
  if (c && materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (!c || nppe <= 0 || type < 0 || type >= c->cfg.num_species) return fail(c, MPIC_ERR_ARG, "mpic_fill: bad arguments");
  const int64_t add = (int64_t)c->m.ne * nppe;
  if (c->n + add > c->cap) return fail(c, MPIC_ERR_CAPACITY, "Cannot create that number of particles, no fill performed");
  FillArgs a{};
  a.m = c->m;
  a.P = c->P[c->cur];
  a.offset = c->n;
  a.e0 = 0;
  a.ecount = c->m.ne;
  a.lo0 = c->m.lo[0];
  a.nppe = nppe;
  a.type = type;
  a.density = density;
  for (int d = 0; d < 3; ++d) a.vel[d] = vel ? vel[d] : 0.0;
  a.temp = temp;
  a.pert_amp = pert_amp;
  a.pert_k = (2.0 * M_PI) / (c->m.hi[0] - c->m.lo[0]);
  a.mass = c->cfg.species_m[type];
  a.seed_offset = seed_offset;
  a.view = particle_view(c);
  HIPCHK(c, launch_fill(a, c->stream));
  int rc;
  if (c->state32 && (rc = mixed_import(c, c->n, add))) return rc;
  c->n += add;
  c->ke_max_prev = 0;
  if ((rc = invalidate_deposit(c))) return rc;
  return update_charge_scale(c);
}

int mpic_weight_charge(mpic_ctx* c) {
  if (c && materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (!c) return MPIC_ERR_ARG;
  tstart(c, T_DEPOSIT);
  int rc = c->state32 ? mixed_weight_charge(c) : do_weight_charge(c);
  tstop(c, T_DEPOSIT);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ms[T_DEPOSIT] = telapsed(c, T_DEPOSIT);
  return MPIC_OK;
}

int mpic_solve(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres) {
  if (!c || !(tol >= 0) || max_iter < 0) return fail(c, MPIC_ERR_ARG, "mpic_solve: bad arguments");
  tstart(c, T_SOLVE);
  int rc;
  if (c->cfg.form == MPIC_FORM_MIXED) {
    rc = mixed_solve(c, tol, max_iter, iters, relres);
  } else if (c->cfg.solver == MPIC_SOLVER_BICGSTAB) {
    int status = 0;
    rc = do_bicgstab(c, tol, max_iter, iters, relres, &status);
    if (rc == MPIC_OK && status != 0)
      rc = fail(c, MPIC_ERR_NOT_CONVERGED, status == 1 ? "BiCGStab reached max_iter" : "BiCGStab stopped early");
  } else {
    rc = do_solve(c, tol, max_iter, iters, relres);
  }
  tstop(c, T_SOLVE);
  (void)hipStreamSynchronize(c->stream);
  c->ms[T_SOLVE] = telapsed(c, T_SOLVE);
  return rc;
}

int mpic_weight_efield(mpic_ctx* c) {
  if (c && materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (!c || c->cap == 0) return fail(c, MPIC_ERR_STATE, "mpic_weight_efield: no particle list");
  if (c->state32) return fail(c, MPIC_ERR_STATE, "mpic_weight_efield: " FP64_ONLY);
  int rc = ensure_transients(c);
  if (rc) return rc;
  HIPCHK(c, launch_elem_phi(c->m, c->phi, c->stream));
  HIPCHK(c, launch_gather(push_args(c), c->stream));
  return MPIC_OK;
}

int mpic_accel(mpic_ctx* c, double dt, double scale) {
  if (c && materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (!c || c->cap == 0) return fail(c, MPIC_ERR_STATE, "mpic_accel: no particle list");
  if (c->state32) return fail(c, MPIC_ERR_STATE, "mpic_accel: " FP64_ONLY);
  int rc = ensure_transients(c);
  if (rc) return rc;
  PushArgs a = push_args(c);
  a.dt = dt;
  a.scale = scale;
  HIPCHK(c, launch_accel(a, c->stream));
  return MPIC_OK;
}

int mpic_move(mpic_ctx* c, double dt, mpic_move_stats* st) {
  if (c && materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (!c || c->cap == 0) return fail(c, MPIC_ERR_STATE, "mpic_move: no particle list");
  if (c->state32) return fail(c, MPIC_ERR_STATE, "mpic_move: " FP64_ONLY);
  int rc = ensure_transients(c);
  if (rc) return rc;
  if ((rc = invalidate_deposit(c))) return rc;
  PushArgs a = push_args(c);
  a.dt = dt;
  tstart(c, T_PUSH);
  HIPCHK(c, launch_move(a, c->stream));
  tstop(c, T_PUSH);
  PushPartial r{};
  if (c->n > 0 && (rc = finish_partials(c, partials_for(c->n), &r))) return rc;
  c->ms[T_PUSH] = c->n > 0 ? telapsed(c, T_PUSH) : 0;
  if (st) *st = mpic_move_stats{r.crossings, r.killed, r.capped, r.nonpos, r.reflections, r.relocated};
  if (r.killed > 0 && (rc = do_sort(c, nullptr))) return rc;  // cleanup_list, particle_list.cpp:70-73
  if (r.capped > 0) return fail(c, MPIC_ERR_WALK_CAP, "a particle exceeded the crossing cap");
  return MPIC_OK;
}

int mpic_compute_ke(mpic_ctx* c, double* ke) {
  if (c && materialize(c) != MPIC_OK) return MPIC_ERR_HIP;
  if (!c || !ke) return MPIC_ERR_ARG;
  if (c->n == 0) {
    *ke = 0;
    return MPIC_OK;
  }
  if (c->state32) return mixed_compute_ke(c, ke);
  const DevParticles& P = c->P[c->cur];
  PushPartial r{};
  HIPCHK(c, launch_ke(P, c->n, c->species, 0, 0, c->partial, c->stream));
  int rc = finish_partials(c, partials_for(c->n), &r);
  if (rc) return rc;
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  const int S = scale_for_bound((double)r.crossings * mx);
  HIPCHK(c, launch_ke(P, c->n, c->species, S, 1, c->partial, c->stream));
  if ((rc = finish_partials(c, partials_for(c->n), &r))) return rc;
  *ke = from_fixed(fx128{r.ke_lo, r.ke_hi}, S);
  c->ke_max_prev = mx;
  return MPIC_OK;
}

int mpic_sort_by_element(mpic_ctx* c, int32_t* start_opt) {
  if (!c || c->cap == 0) return fail(c, MPIC_ERR_STATE, "mpic_sort_by_element: no particle list");
  return do_sort(c, start_opt);
}

int mpic_time_step(mpic_ctx* c, double dt, mpic_step_stats* out) {
  if (!c || c->cap == 0) return fail(c, MPIC_ERR_STATE, "mpic_time_step: no particle list");
  if (c->state32) return mixed_time_step(c, dt, out);
  int rc;
  mpic_step_stats st{};
  c->ms[T_SORT] = 0;  // set again below if this step reorders
  // 1. charge of the current positions (PIC.cpp:250); a pending ranking does not
  //    matter here (the sums are order-free)
  tstart(c, T_DEPOSIT);
  if ((rc = do_weight_charge(c))) return rc;
  tstop(c, T_DEPOSIT);
  // 2. field solve, warm start from the previous potential (PIC.cpp:251)
  tstart(c, T_SOLVE);
  if (c->cfg.solver == MPIC_SOLVER_BICGSTAB) {
    // PIC::time_step ignores the solver's outcome (PIC.cpp:251): only a failed
    // launch stops the step; the residual is reported in the statistics
    int status = 0;
    rc = do_bicgstab(c, c->cfg.cg_tol, c->cfg.cg_max_iter, &st.cg_iters, &st.cg_relres, &status);
  } else {
    rc = do_solve(c, c->cfg.cg_tol, c->cfg.cg_max_iter, &st.cg_iters, &st.cg_relres);
  }
  tstop(c, T_SOLVE);
  if (rc) return rc;
  tstart(c, T_FIELD);
  HIPCHK(c, launch_elem_phi(c->m, c->phi, c->stream));
  tstop(c, T_FIELD);
  // 3. fused gather + kick + walk + deposit of the new positions + KE (PIC.cpp:252-258);
  //    with a pending ranking the push reads through it and writes element order
  if (!(c->ke_max_prev > 0) && c->n > 0) {
    PushPartial r{};
    HIPCHK(c, launch_ke(c->P[c->cur], c->n, c->species, 0, 0, c->partial, c->stream));
    if ((rc = finish_partials(c, partials_for(c->n), &r))) return rc;
    std::memcpy(&c->ke_max_prev, &r.ke_max_bits, 8);
  }
  const bool rank_after = c->cfg.sort_period > 0 && (c->steps + 1) % c->cfg.sort_period == 0;
  PushArgs a = push_args(c);
  a.dt = dt;
  a.S_ke = scale_for_bound((double)(c->perm_pending ? c->perm_n : c->n) * c->ke_max_prev * 256.0);
  if (c->perm_pending) {
    a.perm = c->perm;
    a.Q = c->P[1 - c->cur];
    a.n = c->perm_n;
  }
  if (rank_after) {
    HIPCHK(c, launch_zero(c->sb.count, (size_t)(((int64_t)c->m.nbuckets + 2) / 2) * 8, c->stream));
    a.count = c->sb.count;
  }
  tstart(c, T_PUSH);
  tstart(c, T_PUSH_KERNEL);
  HIPCHK(c, launch_push_fused(a, c->stream));
  tstop(c, T_PUSH_KERNEL);
  if (c->perm_pending) {
    c->cur = 1 - c->cur;
    c->n = c->perm_n;
    c->perm_pending = false;
  }
  // the push's partials (over its grid) and, on a reordering step, the ranking
  // by sort bucket from the push's counts, then one wait for both
  PushPartial r{};
  if (a.n > 0) {
    HIPCHK(c, launch_push_finish(c->partial, partials_for(a.n), c->finish, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->finish_h, c->finish, sizeof(PushPartial), hipMemcpyDeviceToHost, c->stream));
  }
  tstop(c, T_PUSH);
  if (c->sort_timed) {  // the previous step's ranking has completed
    c->ms[T_SORT] = telapsed(c, T_SORT);
    c->sort_timed = false;
  }
  if (rank_after && (rc = rank_enqueue(c, true, true))) return rc;
  HIPCHK(c, wait_stream(c->stream));
  if (a.n > 0) r = *c->finish_h;
  c->acc_deposited = true;
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  tstart(c, T_KE);
  const int S_safe = scale_for_bound((double)c->n * mx);
  if (a.S_ke > S_safe) {
    // The provisional scale could have overflowed: sum again at a safe scale.
    PushPartial k{};
    HIPCHK(c, launch_ke(c->P[c->cur], c->n, c->species, S_safe, 1, c->partial, c->stream));
    if ((rc = finish_partials(c, partials_for(c->n), &k))) return rc;
    st.kinetic_energy = from_fixed(fx128{k.ke_lo, k.ke_hi}, S_safe);
  } else {
    st.kinetic_energy = from_fixed(fx128{r.ke_lo, r.ke_hi}, a.S_ke);
  }
  tstop(c, T_KE);
  c->ke_max_prev = mx;
  st.move = mpic_move_stats{r.crossings, r.killed, r.capped, r.nonpos, r.reflections, r.relocated};
  c->steps++;
  // 4. reordering of the new positions by sort bucket (counting sort; the counts
  //    came from the push), also whenever particles were killed (cleanup_list)
  const int64_t n_live = c->n - r.killed;
  if (rank_after) {
    rank_finish(c, true);
    if (c->perm_n != n_live) return fail(c, MPIC_ERR_STATE, "ranking: live count differs from the push's");
  } else if (r.killed > 0) {
    if ((rc = rank_now(c, true, false, n_live))) return rc;
  }
  HIPCHK(c, wait_stream(c->stream));
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
  st.num_particles = c->perm_pending ? c->perm_n : c->n;
  st.charge_scale_bits = c->S;
  st.ms_field = c->ms[T_FIELD];
  st.ms_deposit = c->ms[T_DEPOSIT];
  st.ms_solve = c->ms[T_SOLVE];
  st.ms_push = c->ms[T_PUSH];
  st.ms_sort = c->ms[T_SORT];
  st.ms_ke = c->ms[T_KE];
  if (out) *out = st;
  if (r.capped > 0) return fail(c, MPIC_ERR_WALK_CAP, "a particle exceeded the crossing cap");
  return MPIC_OK;
}

int mpic_get_rhs(mpic_ctx* c, double* rhs) {
  if (!c || !rhs) return MPIC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(rhs, c->rhs, (size_t)c->m.nn * 8, hipMemcpyDeviceToHost));
  return MPIC_OK;
}

int mpic_get_phi(mpic_ctx* c, double* phi) {
  if (!c || !phi) return MPIC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(phi, c->phi, (size_t)c->m.nn * 8, hipMemcpyDeviceToHost));
  return MPIC_OK;
}

int mpic_set_phi(mpic_ctx* c, const double* phi) {
  if (!c || !phi) return MPIC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->phi, phi, (size_t)c->m.nn * 8, hipMemcpyHostToDevice));
  return MPIC_OK;
}

int mpic_set_rhs(mpic_ctx* c, const double* rhs) {
  if (!c || !rhs) return MPIC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->rhs, rhs, (size_t)c->m.nn * 8, hipMemcpyHostToDevice));
  return MPIC_OK;
}

int mpic_get_efield(mpic_ctx* c, double* E, int64_t stride) {
  if (!c || !E || stride < c->n || !c->E || c->state32)
    return fail(c, MPIC_ERR_STATE, "mpic_get_efield: no field gathered");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int d = 0; d < 3; ++d)
    HIPCHK(c, hipMemcpy(E + d * stride, c->E + d * c->cap, c->n * 8, hipMemcpyDeviceToHost));
  return MPIC_OK;
}

int mpic_charge_scale(mpic_ctx* c, int* bits) {
  if (!c || !bits) return MPIC_ERR_ARG;
  *bits = c->state32 ? c->S64 : c->S;
  return MPIC_OK;
}

int mpic_check_ownership(mpic_ctx* c, int64_t* mismatches) {
  if (!c || !mismatches) return MPIC_ERR_ARG;
  *mismatches = 0;
  if (c->cap == 0 || c->n == 0) return MPIC_OK;
  if (!c->bad) {
    int rc;
    if ((rc = dalloc(c, &c->bad, 1))) return rc;
    if (!c->bad_h && hipHostMalloc((void**)&c->bad_h, sizeof(unsigned long long)) != hipSuccess)
      return fail(c, MPIC_ERR_HIP, "pinned host allocation failed");
  }
  return mixed_check_owner(c, mismatches);
}

int mpic_particles_view(mpic_ctx* c, mpic_particle_view* out) {
  if (!c || !out) return fail(c, MPIC_ERR_ARG, "mpic_particles_view: bad arguments");
  if (c->cap == 0) return fail(c, MPIC_ERR_STATE, "mpic_particles_view: no particle list");
  if (c->state32) return fail(c, MPIC_ERR_STATE, "mpic_particles_view: " FP64_ONLY);
  int rc;
  if ((rc = materialize(c)) || (rc = ensure_transients(c))) return rc;
  if (!c->tr) HIPCHK(c, hipMalloc((void**)&c->tr, (size_t)c->cap * sizeof(double)));
  HIPCHK(c, launch_view_reset(c->view_counters, c->n, c->stream));
  const minipic::ParticleListView v = particle_view(c);
  out->records = v.rec;
  out->elem_copy = v.elem_copy;
  out->E = v.E_;
  out->v_ref = v.v_ref_;
  out->time_remaining = v.time_remaining_;
  out->counters = v.counters;
  out->migrate_index = v.migrate_index;
  out->migrate_proc = v.migrate_proc;
  out->capacity = v.capacity;
  out->migrate_capacity = v.migrate_capacity;
  return MPIC_OK;
}

int mpic_particles_view_commit(mpic_ctx* c, int64_t* num_particles, int64_t* num_migrate) {
  if (!c) return MPIC_ERR_ARG;
  if (c->state32) return fail(c, MPIC_ERR_STATE, "mpic_particles_view_commit: " FP64_ONLY);
  unsigned long long cnt[4] = {0, 0, 0, 0};
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(cnt, c->view_counters, sizeof(cnt), hipMemcpyDeviceToHost));
  c->perm_pending = false;
  c->n = (int64_t)std::min<unsigned long long>(cnt[0], (unsigned long long)c->cap);
  // the element copy from the records (kernels may have written ielement)
  HIPCHK(c, launch_elem_copy(c->P[c->cur], c->n, c->stream));
  int rc;
  if ((rc = invalidate_deposit(c))) return rc;
  c->ke_max_prev = 0;
  if (cnt[1] > 0 && (rc = do_sort(c, nullptr))) return rc;  // cleanup_list: compact the dead
  if ((rc = update_charge_scale(c))) return rc;
  if (num_particles) *num_particles = c->n;
  if (num_migrate) *num_migrate = (int64_t)std::min<unsigned long long>(cnt[2], (unsigned long long)MIGRATE_CAP);
  if (cnt[3] > 0)
    return fail(c, MPIC_ERR_CAPACITY, "ParticleList::push_back: " + std::to_string(cnt[3]) +
                                          " particle(s) beyond the capacity were not added");
  return MPIC_OK;
}

int mpic_multigrid_levels(mpic_ctx* c, int* levels) {
  if (!c || !levels) return MPIC_ERR_ARG;
  *levels = 0;
  if (c->cfg.form != MPIC_FORM_MIXED && !c->state32) {
    c->err = "the FP64 form solves with the reference's solver";
    return MPIC_OK;
  }
  int rc;
  if ((rc = mixed_prepare(c))) return rc;
  if (c->mg_ok)
    *levels = c->mg.nlev + 1;
  else
    c->err = c->mg_why;
  return MPIC_OK;
}

int mpic_stream(mpic_ctx* c, void** stream) {
  if (!c || !stream) return MPIC_ERR_ARG;
  *stream = (void*)c->stream;
  return MPIC_OK;
}

int mpic_kernel_time(mpic_ctx* c, int which, double* ms) {
  if (!c || !ms) return MPIC_ERR_ARG;
  switch (which) {
    case 0: *ms = c->ms[T_PUSH_KERNEL]; break;
    case 1: *ms = c->ms[T_DEPOSIT]; break;
    case 3: *ms = c->last_cg_iters > 0 ? c->ms[T_SOLVE] / c->last_cg_iters : 0; break;
    case 4: *ms = c->ms[T_SORT]; break;
    default: return MPIC_ERR_ARG;
  }
  return MPIC_OK;
}

}  // extern "C"
