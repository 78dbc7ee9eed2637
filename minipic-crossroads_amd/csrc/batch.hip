// Batching form (mpic_batch_* of include/minipic_c.h): an ensemble of
// independent problems advanced together in the precision form.
//
// The problems are stored as one disjoint mesh -- problem p owns nodes
// [node_off[p], node_off[p+1]) and elements [elem_off[p], elem_off[p+1]); no
// face links one problem to another, so particles never leave their problem and
// the element-sorted particle list keeps every problem in one contiguous range.
// A step is a fixed sequence of launches, one per phase for all problems and no
// host synchronisation in between, so it can be captured in a HIP graph:
//   node charge of the deposit -> clear the deposit -> per-problem solve (one
//   workgroup per problem: conjugate gradients preconditioned by a multigrid
//   V-cycle on each problem's own hierarchy when every problem's brick
//   coarsens, else FP32 CG inside FP64 refinement) -> per-tet field -> push
//   through the last ranking
//   (kick, walk, FP64 location, deposit, counts) -> ranking by element ->
//   per-problem kinetic energy.
// Particles are reflected at walls (no kills), so every problem keeps its
// particle count and the graph replays unchanged.  The tile shapes (workgroups
// of the push, threads per problem of the solve) are chosen by mpic_batch_tune.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mg.hpp"
#include "mpic_ctx.hpp"

using namespace mpic;

struct mpic_batch {
  mpic_ctx* c = nullptr;
  int nprob = 0;
  std::vector<int32_t> node_off, elem_off;  // [nprob+1]
  std::vector<double> lo0, lx;              // x origin and length of each problem's box
  std::vector<int64_t> part_off;            // [nprob+1] particle range of each problem
  int32_t* d_node_off = nullptr;
  int64_t* d_part_off = nullptr;
  double* r64 = nullptr;
  float *rf = nullptr, *wf = nullptr, *pf = nullptr, *sf = nullptr, *df = nullptr;
  int *d_iters = nullptr, *d_conv = nullptr;
  double *d_relres = nullptr, *d_ke = nullptr;
  int64_t* d_count = nullptr;
  int32_t* d_chunk_off = nullptr;           // [nprob+1] first statistics chunk of each problem
  StatPartial* d_stat_part = nullptr;       // one partial per statistics chunk
  int nchunks = 0;
  int max_nodes = 0;                        // nodes of the largest problem
  double* phi_save = nullptr;               // tuner: potential before the trials
  unsigned long long* tune_acc = nullptr;   // tuner: scratch deposit
  int32_t* tune_count = nullptr;
  // the multigrid solve: per-problem hierarchies in pooled device arrays
  bool mg_on = false;
  std::string mg_why;                       // why the refinement solve runs instead
  BatchMGProb* d_mg_prob = nullptr;
  MGLevelDev* d_mg_lev = nullptr;
  float *d_mg_dinv0 = nullptr, *d_mg_cinv = nullptr;
  double* d_mg_val0_64 = nullptr;
  float* d_mg_val0_32 = nullptr;
  int mg_max_slots = 0;
  double *mg_p = nullptr, *mg_q = nullptr;
  int push_grid = 0, solve_threads = 256;
  bool tuned = false, filled = false;
  bool graph_on = false;
  bool captured[2] = {false, false};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  double graph_dt = 0;
  int S_ke = 100;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  double ms_step = 0;
  std::vector<void*> allocs;
};

namespace {

int bfail(mpic_batch* b, int code, const std::string& msg) {
  if (b)
    b->err = msg;
  else
    global_error() = msg;
  return code;
}

#define BCHK(b, call)                                                                                \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return bfail(b, MPIC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A failure inside the context: carry its message.
int cfail(mpic_batch* b, int rc) {
  if (rc != MPIC_OK) b->err = b->c->err;
  return rc;
}

template <class T>
int balloc(mpic_batch* b, T** p, size_t count) {
  void* q = nullptr;
  BCHK(b, hipMalloc(&q, (count ? count : 1) * sizeof(T)));
  b->allocs.push_back(q);
  *p = (T*)q;
  return MPIC_OK;
}

void drop_graphs(mpic_batch* b) {
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    b->exec[k] = nullptr;
    b->captured[k] = false;
  }
}

// The ranking of the live list by element (the sort key of the batch: one
// bucket per element) must be pending at the start of a step.
int ensure_ranked(mpic_batch* b) {
  mpic_ctx* c = b->c;
  if (c->perm_pending) return MPIC_OK;
  return cfail(b, rank_now(c, true, false, c->n));
}

BatchSolveArgs solve_args(mpic_batch* b) {
  mpic_ctx* c = b->c;
  BatchSolveArgs a;
  a.m = diag_first(c->m, c->m32);
  a.val32 = c->m32.ell_val;
  a.node_off = b->d_node_off;
  a.x = c->phi;
  a.b = c->rhs;
  a.r64 = b->r64;
  a.rf = b->rf;
  a.wf = b->wf;
  a.pf = b->pf;
  a.sf = b->sf;
  a.df = b->df;
  a.iters = b->d_iters;
  a.converged = b->d_conv;
  a.relres = b->d_relres;
  a.tol2 = c->cfg.cg_tol * c->cfg.cg_tol;
  a.in_tol_min = 1e-5;
  a.max_iter = c->cfg.cg_max_iter;
  a.max_nodes = b->max_nodes;
  return a;
}

PushArgs32 push_args(mpic_batch* b, double dt, unsigned long long* acc, int32_t* count, int grid) {
  mpic_ctx* c = b->c;
  PushArgs32 a{};
  a.m = c->m;
  a.m32 = c->m32;
  a.P = c->P32[c->cur32];
  a.Q = c->P32[1 - c->cur32];
  a.perm = c->perm;
  a.count = count;
  a.n = c->n;
  a.dt = (float)dt;
  a.scale = 1.0f;
  a.kill = 0;
  a.S_charge = c->S64;
  a.S_ke = b->S_ke;
  a.acc = acc;
  a.partial = c->partial;
  a.species = c->species;
  a.grid = grid;
  return a;
}

BatchMGArgs mg_args(mpic_batch* b) {
  mpic_ctx* c = b->c;
  BatchMGArgs a;
  a.m = diag_first(c->m, c->m32);
  a.val32 = c->m32.ell_val;
  a.prob = b->d_mg_prob;
  a.lev = b->d_mg_lev;
  a.dinv0 = b->d_mg_dinv0;
  a.cinv = b->d_mg_cinv;
  a.val0_64 = b->d_mg_val0_64;
  a.val0_32 = b->d_mg_val0_32;
  a.x = c->phi;
  a.b = c->rhs;
  a.r = b->r64;
  a.p = b->mg_p;
  a.q = b->mg_q;
  a.b0 = b->rf;
  a.x0 = b->wf;
  a.t0 = b->pf;
  a.iters = b->d_iters;
  a.converged = b->d_conv;
  a.relres = b->d_relres;
  a.tol2 = c->cfg.cg_tol * c->cfg.cg_tol;
  a.max_iter = c->cfg.cg_max_iter;
  return a;
}

// Every problem's solve with `threads` per workgroup.
hipError_t launch_solve(mpic_batch* b, int threads, hipStream_t s) {
  return b->mg_on ? launch_batch_mg(mg_args(b), b->nprob, threads, b->mg_max_slots, s)
                  : launch_batch_ir(solve_args(b), b->nprob, threads, s);
}

// The multigrid hierarchy of every problem (mg_build on its brick, the
// coarsest level at most BATCH_MG_DIRECT_MAX unknowns) in pooled device
// arrays.  b->mg_on stays false, with the reason in b->mg_why, when multigrid
// is switched off or some problem's mesh does not coarsen: then every problem
// runs the refinement solve.
int build_mg(mpic_batch* b, const std::vector<mpic_mesh_desc>& d, const mpic_config& cc) {
  b->mg_on = false;
  if (!cc.multigrid) {
    b->mg_why = "multigrid switched off (cfg.multigrid = 0)";
    return MPIC_OK;
  }
  const int P = b->nprob;
  std::vector<MGHost> hs(P);
  for (int p = 0; p < P; ++p) {
    const mpic_mesh_desc& m = d[p];
    std::string why;
    if (m.lattice[0] <= 0 || m.lattice[1] <= 0 || m.lattice[2] <= 0) {
      b->mg_why = "problem " + std::to_string(p) + ": its mesh descriptor names no lattice";
      return MPIC_OK;
    }
    if (!mg_build(m.lattice[0], m.lattice[1], m.lattice[2], m.num_nodes, m.boundary, m.row_ptr, m.col_idx, m.values,
                  hs[p], why, BATCH_MG_DIRECT_MAX)) {
      b->mg_why = "problem " + std::to_string(p) + ": " + why;
      return MPIC_OK;
    }
  }
  int64_t npts = 0, ninv = 0, nv0 = 0;
  int nlev = 0;
  b->mg_max_slots = 0;
  for (const MGHost& h : hs) {
    for (const MGHostLevel& L : h.levels) npts += L.n;
    nlev += (int)h.levels.size();
    ninv += (int64_t)h.coarse_inv.size();
    nv0 += (int64_t)h.val0.size();
    b->mg_max_slots = std::max(b->mg_max_slots, (int)h.slot0.size());
  }
  const int32_t nn = b->c->m.nn;
  float *val, *dinv, *vb, *vx, *vt;
  int rc;
  if ((rc = balloc(b, &val, (size_t)27 * npts)) || (rc = balloc(b, &dinv, (size_t)npts)) ||
      (rc = balloc(b, &vb, (size_t)npts)) || (rc = balloc(b, &vx, (size_t)npts)) || (rc = balloc(b, &vt, (size_t)npts)) ||
      (rc = balloc(b, &b->d_mg_cinv, (size_t)ninv)) || (rc = balloc(b, &b->d_mg_dinv0, (size_t)nn)) ||
      (rc = balloc(b, &b->mg_p, (size_t)nn)) || (rc = balloc(b, &b->mg_q, (size_t)nn)) ||
      (rc = balloc(b, &b->d_mg_lev, (size_t)nlev)) || (rc = balloc(b, &b->d_mg_prob, (size_t)P)) ||
      (rc = balloc(b, &b->d_mg_val0_64, (size_t)nv0)) || (rc = balloc(b, &b->d_mg_val0_32, (size_t)nv0)))
    return rc;
  std::vector<float> hval((size_t)27 * npts), hdinv(npts), hcinv(ninv), hdinv0(nn, 1.0f), hv32(nv0);
  std::vector<double> hv64(nv0);
  int64_t vo = 0;
  std::vector<MGLevelDev> hlev(nlev);
  std::vector<BatchMGProb> hprob(P);
  int64_t po = 0, co = 0;
  int lo = 0;
  for (int p = 0; p < P; ++p) {
    const MGHost& h = hs[p];
    BatchMGProb& Q = hprob[p];
    Q.nx = h.nx;
    Q.ny = h.ny;
    Q.nz = h.nz;
    Q.r0 = b->node_off[p];
    Q.nn = b->node_off[p + 1] - b->node_off[p];
    Q.lev0 = lo;
    Q.nlev = (int)h.levels.size();
    Q.cinv_off = co;
    Q.v0_off = vo;
    Q.s0 = (int)h.slot0.size();
    for (int k = 0; k < 27; ++k) {
      const int sl = k < Q.s0 ? h.slot0[k] : 13;
      Q.off0[k] = ((sl / 9 - 1) * (h.ny + 1) + (sl / 3) % 3 - 1) * (h.nz + 1) + sl % 3 - 1;
    }
    std::copy(h.val0.begin(), h.val0.end(), hv64.begin() + vo);
    for (size_t k = 0; k < h.val0.size(); ++k) hv32[vo + k] = (float)h.val0[k];
    vo += (int64_t)h.val0.size();
    std::copy(h.dinv0.begin(), h.dinv0.end(), hdinv0.begin() + Q.r0);
    for (const MGHostLevel& L : h.levels) {
      MGLevelDev& D = hlev[lo++];
      D.cx = L.cx;
      D.cy = L.cy;
      D.cz = L.cz;
      D.n = L.n;
      D.val = val + 27 * po;
      D.dinv = dinv + po;
      D.b = vb + po;
      D.x = vx + po;
      D.t = vt + po;
      std::copy(L.val.begin(), L.val.end(), hval.begin() + 27 * po);
      std::copy(L.dinv.begin(), L.dinv.end(), hdinv.begin() + po);
      po += L.n;
    }
    std::copy(h.coarse_inv.begin(), h.coarse_inv.end(), hcinv.begin() + co);
    co += (int64_t)h.coarse_inv.size();
  }
  BCHK(b, hipMemcpy(val, hval.data(), hval.size() * 4, hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(dinv, hdinv.data(), hdinv.size() * 4, hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_mg_cinv, hcinv.data(), hcinv.size() * 4, hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_mg_dinv0, hdinv0.data(), hdinv0.size() * 4, hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_mg_lev, hlev.data(), hlev.size() * sizeof(MGLevelDev), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_mg_prob, hprob.data(), hprob.size() * sizeof(BatchMGProb), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_mg_val0_64, hv64.data(), hv64.size() * 8, hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_mg_val0_32, hv32.data(), hv32.size() * 4, hipMemcpyHostToDevice));
  b->mg_on = true;
  b->mg_why.clear();
  return MPIC_OK;
}

int push_grid_used(mpic_batch* b) {
  return b->push_grid > 0 ? std::min(b->push_grid, PUSH32_MAX_GRID) : push32_grid(b->c->n);
}

// One step of every problem as a fixed sequence of launches (no host
// synchronisation), valid for capture.  Requires a pending element ranking and
// the deposit of the current positions in acc64.
int enqueue_step(mpic_batch* b, double dt) {
  mpic_ctx* c = b->c;
  hipStream_t s = c->stream;
  const int32_t ne = c->m.ne;
  BCHK(b, launch_node_charge64(c->m, c->acc64, c->S64, c->bg64, c->rhs, s));
  BCHK(b, launch_zero(c->acc64, (size_t)ne * 8, s));
  BCHK(b, launch_solve(b, b->solve_threads, s));
  BCHK(b, launch_field32(c->m, c->phi, c->m32.E, s));
  BCHK(b, launch_zero(c->sb.count, (size_t)(((int64_t)c->m.nbuckets + 2) / 2) * 8, s));
  BCHK(b, launch_push32(push_args(b, dt, c->acc64, c->sb.count, b->push_grid), s));
  BCHK(b, launch_push_finish(c->partial, push_grid_used(b), c->finish, s));
  DevParticles live = c->P[c->cur];
  live.elem = c->P32[1 - c->cur32].elem;
  BCHK(b, launch_rank(c->m, live, c->n, true, true, c->sb, c->perm, s));
  BCHK(b, launch_batch_stats(c->P32[1 - c->cur32].r, b->d_part_off, b->d_chunk_off, b->nchunks, c->species,
                             b->S_ke, b->d_stat_part, b->d_ke, b->d_count, b->nprob, s));
  return MPIC_OK;
}

// Host bookkeeping after a step: the output set is live and its ranking pending.
void after_step(mpic_batch* b) {
  mpic_ctx* c = b->c;
  c->cur32 = 1 - c->cur32;
  c->perm_pending = true;
  c->perm_by_bucket = true;
  c->perm_n = c->n;
  c->acc64_deposited = true;
  c->steps++;
}

}  // namespace

extern "C" {

const char* mpic_batch_last_error(const mpic_batch* b) { return b ? b->err.c_str() : global_error().c_str(); }

int mpic_batch_create(int nprob, mpic_mesh* const* meshes, const mpic_config* cfg, mpic_batch** out) {
  if (nprob <= 0 || !meshes || !out) return bfail(nullptr, MPIC_ERR_ARG, "mpic_batch_create: bad arguments");
  std::vector<mpic_mesh_desc> d(nprob);
  int64_t nn = 0, ne = 0, nnz = 0;
  for (int p = 0; p < nprob; ++p) {
    if (!meshes[p] || mpic_mesh_describe(meshes[p], &d[p]) != MPIC_OK)
      return bfail(nullptr, MPIC_ERR_ARG, "mpic_batch_create: bad mesh");
    if (d[p].node_pair)  // the batch solve is the precision form's CG
      return bfail(nullptr, MPIC_ERR_ARG, "mpic_batch_create: periodic meshes need BiCGStab (single contexts)");
    nn += d[p].num_nodes;
    ne += d[p].num_elems;
    nnz += d[p].nnz;
  }
  if (nn >= (1ll << 31) - 64 || ne >= (1ll << 31) || nnz >= (1ll << 31))
    return bfail(nullptr, MPIC_ERR_ARG, "mpic_batch_create: ensemble too large for 32-bit indices");
  mpic_batch* b = new mpic_batch();
  b->nprob = nprob;
  b->node_off.assign(nprob + 1, 0);
  b->elem_off.assign(nprob + 1, 0);
  b->lo0.assign(nprob, 0.0);
  b->lx.assign(nprob, 1.0);
  // the disjoint union of the problems' meshes and matrices
  std::vector<double> nodes(nn * 3), normals(ne * 12), jinv(ne * 9), detj(ne), grad(ne * 12), wbv(ne * 16), val(nnz);
  std::vector<int32_t> en(ne * 4), nb(ne * 4), rowptr(nn + 1, 0), col(nnz);
  std::vector<uint8_t> bnd(nn);
  int64_t no = 0, eo = 0, zo = 0;
  for (int p = 0; p < nprob; ++p) {
    const mpic_mesh_desc& m = d[p];
    b->node_off[p] = (int32_t)no;
    b->elem_off[p] = (int32_t)eo;
    double lo = 1e300, hi = -1e300;
    for (int32_t i = 0; i < m.num_nodes; ++i) {
      lo = std::min(lo, m.nodes[(size_t)i * 3]);
      hi = std::max(hi, m.nodes[(size_t)i * 3]);
    }
    b->lo0[p] = lo;
    b->lx[p] = hi > lo ? hi - lo : 1.0;
    std::memcpy(&nodes[no * 3], m.nodes, sizeof(double) * 3 * m.num_nodes);
    std::memcpy(&bnd[no], m.boundary, m.num_nodes);
    std::memcpy(&normals[eo * 12], m.normals, sizeof(double) * 12 * m.num_elems);
    std::memcpy(&jinv[eo * 9], m.inv_jacobian, sizeof(double) * 9 * m.num_elems);
    std::memcpy(&detj[eo], m.det_jacobian, sizeof(double) * m.num_elems);
    std::memcpy(&grad[eo * 12], m.grad_basis, sizeof(double) * 12 * m.num_elems);
    std::memcpy(&wbv[eo * 16], m.wbv, sizeof(double) * 16 * m.num_elems);
    for (int64_t k = 0; k < (int64_t)m.num_elems * 4; ++k) {
      en[eo * 4 + k] = m.elem_nodes[k] + (int32_t)no;
      nb[eo * 4 + k] = m.elem_nbrs[k] >= 0 ? m.elem_nbrs[k] + (int32_t)eo : -1;
    }
    for (int32_t i = 0; i < m.num_nodes; ++i) rowptr[no + i + 1] = m.row_ptr[i + 1] + (int32_t)zo;
    for (int64_t k = 0; k < m.nnz; ++k) col[zo + k] = m.col_idx[k] + (int32_t)no;
    std::memcpy(&val[zo], m.values, sizeof(double) * m.nnz);
    no += m.num_nodes;
    eo += m.num_elems;
    zo += m.nnz;
  }
  b->node_off[nprob] = (int32_t)no;
  b->elem_off[nprob] = (int32_t)eo;
  for (int p = 0; p < nprob; ++p) b->max_nodes = std::max(b->max_nodes, b->node_off[p + 1] - b->node_off[p]);
  mpic_mesh_desc u{};
  u.num_nodes = (int32_t)nn;
  u.num_elems = (int32_t)ne;
  u.nnz = nnz;
  u.nodes = nodes.data();
  u.elem_nodes = en.data();
  u.elem_nbrs = nb.data();
  u.normals = normals.data();
  u.inv_jacobian = jinv.data();
  u.det_jacobian = detj.data();
  u.grad_basis = grad.data();
  u.wbv = wbv.data();
  u.boundary = bnd.data();
  u.row_ptr = rowptr.data();
  u.col_idx = col.data();
  u.values = val.data();
  u.background = nullptr;
  mpic_config cc;
  if (cfg)
    cc = *cfg;
  else
    mpic_config_default(&cc);
  cc.form = MPIC_FORM_MIXED;  // the batching form runs the precision form
  cc.kill_on_wall = 0;        // particle counts stay fixed
  cc.sort_period = 1;
  cc.sort_bucket = 1;         // element order keeps every problem contiguous
  int rc = mpic_create(&u, &cc, &b->c);
  if (rc != MPIC_OK) {
    b->err = global_error();
    global_error() = b->err;
    delete b;
    return rc;
  }
  if ((rc = balloc(b, &b->d_node_off, nprob + 1)) || (rc = balloc(b, &b->d_part_off, nprob + 1)) ||
      (rc = balloc(b, &b->r64, nn)) || (rc = balloc(b, &b->rf, nn)) || (rc = balloc(b, &b->wf, nn)) ||
      (rc = balloc(b, &b->pf, nn)) || (rc = balloc(b, &b->sf, nn)) || (rc = balloc(b, &b->df, nn)) ||
      (rc = balloc(b, &b->d_iters, nprob)) || (rc = balloc(b, &b->d_conv, nprob)) ||
      (rc = balloc(b, &b->d_relres, nprob)) || (rc = balloc(b, &b->d_ke, nprob)) ||
      (rc = balloc(b, &b->d_count, nprob)) || (rc = balloc(b, &b->phi_save, nn)) ||
      (rc = balloc(b, &b->tune_acc, ne)) || (rc = balloc(b, &b->tune_count, ne + 2))) {
    global_error() = b->err;
    mpic_batch_destroy(b);
    return rc;
  }
  if (hipMemcpy(b->d_node_off, b->node_off.data(), sizeof(int32_t) * (nprob + 1), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(b->pf, 0, nn * 4) != hipSuccess || hipMemset(b->sf, 0, nn * 4) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    global_error() = "mpic_batch_create: device setup failed";
    mpic_batch_destroy(b);
    return MPIC_ERR_HIP;
  }
  if ((rc = build_mg(b, d, cc))) {
    global_error() = b->err;
    mpic_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MPIC_OK;
}

void mpic_batch_destroy(mpic_batch* b) {
  if (!b) return;
  if (b->c && b->c->stream) (void)hipStreamSynchronize(b->c->stream);
  drop_graphs(b);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->c) mpic_destroy(b->c);
  delete b;
}

int mpic_batch_fill(mpic_batch* b, const mpic_batch_fill_desc* f) {
  if (!b || !f) return bfail(b, MPIC_ERR_ARG, "mpic_batch_fill: bad arguments");
  if (b->filled) return bfail(b, MPIC_ERR_STATE, "mpic_batch_fill: the ensemble is already filled");
  mpic_ctx* c = b->c;
  const int P = b->nprob;
  b->part_off.assign(P + 1, 0);
  for (int p = 0; p < P; ++p) {
    if (f[p].nppe <= 0 || !(f[p].density > 0) || f[p].temp < 0)
      return bfail(b, MPIC_ERR_ARG, "mpic_batch_fill: bad fill of problem " + std::to_string(p));
    b->part_off[p + 1] = b->part_off[p] + (int64_t)(b->elem_off[p + 1] - b->elem_off[p]) * f[p].nppe;
  }
  const int64_t total = b->part_off[P];
  if (total >= (1ll << 31) - 1) return bfail(b, MPIC_ERR_CAPACITY, "mpic_batch_fill: more than 2^31 particles");
  int rc;
  if ((rc = mpic_particles_reserve(c, total))) return cfail(b, rc);
  // the fills write FP64 records in fill order: problem by problem, element-major
  for (int p = 0; p < P; ++p) {
    const int32_t e0 = b->elem_off[p], ecount = b->elem_off[p + 1] - b->elem_off[p];
    const bool two = f[p].drift != 0.0;
    const int na = two ? f[p].nppe / 2 : f[p].nppe;
    const int nbn = f[p].nppe - na;
    for (int half = 0; half < (two ? 2 : 1); ++half) {
      const int nppe = half == 0 ? na : nbn;
      if (nppe <= 0) continue;
      FillArgs a{};
      a.m = c->m;
      a.P = c->P[c->cur];
      a.offset = b->part_off[p] + (half == 0 ? 0 : (int64_t)ecount * na);
      a.e0 = e0;
      a.ecount = ecount;
      a.nppe = nppe;
      a.type = 0;
      a.density = two ? f[p].density / 2 : f[p].density;
      a.vel[0] = two ? (half == 0 ? f[p].drift : -f[p].drift) : 0.0;
      a.vel[1] = a.vel[2] = 0.0;
      a.temp = f[p].temp;
      a.pert_amp = f[p].pert_amp;
      a.pert_k = 2.0 * M_PI / b->lx[p];
      a.lo0 = b->lo0[p];
      a.mass = c->cfg.species_m[0];
      a.seed_offset = f[p].seed_offset + (uint64_t)half;
      a.view = particle_view(c);
      BCHK(b, launch_fill(a, c->stream));
    }
  }
  c->n = total;
  if ((rc = mixed_import(c, 0, total))) return cfail(b, rc);
  // neutralising background of each problem: the load vector of its fill density
  std::vector<double> bg(c->m.nn, 0.0);
  {
    std::vector<double> wbv((size_t)c->m.ne * 16);
    std::vector<int32_t> conn((size_t)c->m.ne * 4);
    std::vector<uint8_t> bnd(c->m.nn);
    BCHK(b, hipMemcpy(wbv.data(), c->m.wbv, wbv.size() * 8, hipMemcpyDeviceToHost));
    BCHK(b, hipMemcpy(conn.data(), c->m.conn, conn.size() * 4, hipMemcpyDeviceToHost));
    BCHK(b, hipMemcpy(bnd.data(), c->m.bnd, bnd.size(), hipMemcpyDeviceToHost));
    for (int p = 0; p < P; ++p)
      for (int64_t e = b->elem_off[p]; e < b->elem_off[p + 1]; ++e)
        for (int i = 0; i < 4; ++i) {
          double v = 0;
          for (int q = 0; q < 4; ++q) v += -f[p].density * wbv[e * 16 + i * 4 + q];  // ES.cpp:248-253
          const int32_t row = conn[e * 4 + i];
          if (bnd[row]) v = 0;
          bg[row] += v;
        }
  }
  c->background = bg;
  c->bg_abs_sum = 0;
  for (double v : bg) c->bg_abs_sum += std::fabs(v);
  if (!c->bg64 && (rc = dalloc(c, &c->bg64, bg.size()))) return cfail(b, rc);
  BCHK(b, hipMemcpy(c->bg64, bg.data(), bg.size() * 8, hipMemcpyHostToDevice));
  if ((rc = mixed_charge_scale(c))) return cfail(b, rc);
  // initial deposit and ranking
  BCHK(b, launch_zero(c->acc64, (size_t)c->m.ne * 8, c->stream));
  BCHK(b, launch_deposit32(c->m, c->P32[c->cur32], c->n, c->species, c->S64, c->acc64, c->stream));
  c->acc64_deposited = true;
  if ((rc = rank_now(c, true, false, total))) return cfail(b, rc);
  BCHK(b, hipMemcpy(b->d_part_off, b->part_off.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice));
  {
    std::vector<int32_t> co(P + 1, 0);
    for (int p = 0; p < P; ++p)
      co[p + 1] = co[p] + (int32_t)((b->part_off[p + 1] - b->part_off[p] + STAT_CHUNK - 1) / STAT_CHUNK);
    b->nchunks = co[P];
    if ((rc = balloc(b, &b->d_chunk_off, P + 1)) || (rc = balloc(b, &b->d_stat_part, b->nchunks))) return rc;
    BCHK(b, hipMemcpy(b->d_chunk_off, co.data(), sizeof(int32_t) * (P + 1), hipMemcpyHostToDevice));
  }
  // fixed-point scale of the kinetic-energy sums (the push's and the per-problem
  // ones): every term may grow to 256 times the fill's largest before a sum
  // could leave the 128-bit range
  PushPartial r{};
  BCHK(b, launch_ke32(c->P32[c->cur32], c->n, c->species, 0, 0, c->partial, c->stream));
  if ((rc = finish_partials(c, partials_for(c->n), &r))) return cfail(b, rc);
  double mx;
  std::memcpy(&mx, &r.ke_max_bits, 8);
  b->S_ke = scale_for_bound((double)c->n * mx * 256.0);
  b->filled = true;
  drop_graphs(b);
  return MPIC_OK;
}

int mpic_batch_tune(mpic_batch* b, double dt) {
  if (!b || !b->filled) return bfail(b, MPIC_ERR_STATE, "mpic_batch_tune: fill the ensemble first");
  mpic_ctx* c = b->c;
  hipStream_t s = c->stream;
  int rc;
  if ((rc = ensure_ranked(b))) return rc;
  const int32_t nn = c->m.nn;
  BCHK(b, hipMemcpyAsync(b->phi_save, c->phi, (size_t)nn * 8, hipMemcpyDeviceToDevice, s));
  BCHK(b, launch_node_charge64(c->m, c->acc64, c->S64, c->bg64, c->rhs, s));
  // solve: threads per problem
  const int cand_t[3] = {256, 512, 1024};
  float best = 1e30f;
  for (int k = 0; k < 3; ++k) {
    BCHK(b, hipMemcpyAsync(c->phi, b->phi_save, (size_t)nn * 8, hipMemcpyDeviceToDevice, s));
    BCHK(b, hipEventRecord(b->ev0, s));
    BCHK(b, launch_solve(b, cand_t[k], s));
    BCHK(b, hipEventRecord(b->ev1, s));
    BCHK(b, hipEventSynchronize(b->ev1));
    float ms = 0;
    BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (ms < best) {
      best = ms;
      b->solve_threads = cand_t[k];
    }
  }
  // push: workgroups (particles per workgroup), on the field of the current
  // potential, into scratch deposit and counts; the input set is not modified
  BCHK(b, launch_field32(c->m, c->phi, c->m32.E, s));
  const int cand_g[4] = {4096, 8192, 18432, 32768};  // 512-thread workgroups
  best = 1e30f;
  for (int k = 0; k < 4; ++k) {
    BCHK(b, launch_zero(b->tune_acc, (size_t)c->m.ne * 8, s));
    BCHK(b, launch_zero(b->tune_count, (size_t)(((int64_t)c->m.nbuckets + 2) / 2) * 8, s));
    BCHK(b, hipEventRecord(b->ev0, s));
    BCHK(b, launch_push32(push_args(b, dt, b->tune_acc, b->tune_count, cand_g[k]), s));
    BCHK(b, hipEventRecord(b->ev1, s));
    BCHK(b, hipEventSynchronize(b->ev1));
    float ms = 0;
    BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (ms < best) {
      best = ms;
      b->push_grid = cand_g[k];
    }
  }
  BCHK(b, hipMemcpyAsync(c->phi, b->phi_save, (size_t)nn * 8, hipMemcpyDeviceToDevice, s));
  BCHK(b, hipStreamSynchronize(s));
  b->tuned = true;
  drop_graphs(b);
  return MPIC_OK;
}

int mpic_batch_graph(mpic_batch* b, int enable) {
  if (!b) return MPIC_ERR_ARG;
  b->graph_on = enable != 0;
  if (!b->graph_on) drop_graphs(b);
  return MPIC_OK;
}

int mpic_batch_step(mpic_batch* b, double dt, mpic_batch_stats* st) {
  if (!b || !b->filled) return bfail(b, MPIC_ERR_STATE, "mpic_batch_step: fill the ensemble first");
  mpic_ctx* c = b->c;
  hipStream_t s = c->stream;
  int rc;
  if ((rc = ensure_ranked(b))) return rc;
  if (!c->acc64_deposited) {
    BCHK(b, launch_zero(c->acc64, (size_t)c->m.ne * 8, s));
    BCHK(b, launch_deposit32(c->m, c->P32[c->cur32], c->n, c->species, c->S64, c->acc64, s));
    c->acc64_deposited = true;
  }
  BCHK(b, hipEventRecord(b->ev0, s));
  if (b->graph_on) {
    const int k = c->cur32;
    if (b->captured[k] && b->graph_dt != dt) drop_graphs(b);
    if (!b->captured[k]) {
      hipGraph_t g = nullptr;
      BCHK(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = enqueue_step(b, dt);
      hipGraph_t tmp = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &tmp);
      if (rc) return rc;
      BCHK(b, e);
      g = tmp;
      BCHK(b, hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      b->captured[k] = true;
      b->graph_dt = dt;
    }
    BCHK(b, hipGraphLaunch(b->exec[k], s));
  } else {
    if ((rc = enqueue_step(b, dt))) return rc;
  }
  BCHK(b, hipEventRecord(b->ev1, s));
  after_step(b);
  BCHK(b, hipMemcpyAsync(c->finish_h, c->finish, sizeof(PushPartial), hipMemcpyDeviceToHost, s));
  BCHK(b, hipStreamSynchronize(s));
  float ms = 0;
  BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->ms_step = ms;
  // The push's totals (crossing counters, largest KE term).  The per-problem KE
  // sums of this step used the scale fixed at fill; if the largest term has
  // outgrown it they are summed again at a safe scale, which also becomes the
  // scale of later steps (the captured graphs hold the old one: recaptured).
  const PushPartial fin = *c->finish_h;
  double kmax;
  std::memcpy(&kmax, &fin.ke_max_bits, 8);
  const bool rescale = scale_for_bound((double)c->n * kmax) < b->S_ke;
  if (rescale) {
    b->S_ke = scale_for_bound((double)c->n * kmax * 256.0);
    drop_graphs(b);
    BCHK(b, launch_batch_stats(c->P32[c->cur32].r, b->d_part_off, b->d_chunk_off, b->nchunks, c->species, b->S_ke,
                               b->d_stat_part, b->d_ke, b->d_count, b->nprob, s));
    BCHK(b, hipStreamSynchronize(s));
  }
  if (st) {
    std::vector<int> it(b->nprob), cv(b->nprob);
    BCHK(b, hipMemcpy(it.data(), b->d_iters, sizeof(int) * b->nprob, hipMemcpyDeviceToHost));
    BCHK(b, hipMemcpy(cv.data(), b->d_conv, sizeof(int) * b->nprob, hipMemcpyDeviceToHost));
    st->push_grid = push_grid_used(b);
    st->solve_threads = b->solve_threads;
    st->tuned = b->tuned ? 1 : 0;
    st->graph = b->graph_on ? 1 : 0;
    st->max_inner_iters = *std::max_element(it.begin(), it.end());
    st->all_converged = std::all_of(cv.begin(), cv.end(), [](int v) { return v != 0; }) ? 1 : 0;
    st->num_particles = c->n;
    st->ms_step = ms;
    st->crossings = fin.crossings;
    st->capped = fin.capped;
    st->nonpositive = fin.nonpos;
    st->ke_rescaled = rescale ? 1 : 0;
    st->multigrid = b->mg_on ? 1 : 0;
  }
  if (fin.capped > 0) return bfail(b, MPIC_ERR_WALK_CAP, "a particle exceeded the crossing cap");
  return MPIC_OK;
}

int mpic_batch_problem_stats(mpic_batch* b, double* ke, int* iters, double* relres, int* converged, int64_t* nparts) {
  if (!b) return MPIC_ERR_ARG;
  BCHK(b, hipStreamSynchronize(b->c->stream));
  const size_t P = b->nprob;
  if (ke) BCHK(b, hipMemcpy(ke, b->d_ke, P * 8, hipMemcpyDeviceToHost));
  if (iters) BCHK(b, hipMemcpy(iters, b->d_iters, P * 4, hipMemcpyDeviceToHost));
  if (relres) BCHK(b, hipMemcpy(relres, b->d_relres, P * 8, hipMemcpyDeviceToHost));
  if (converged) BCHK(b, hipMemcpy(converged, b->d_conv, P * 4, hipMemcpyDeviceToHost));
  if (nparts) BCHK(b, hipMemcpy(nparts, b->d_count, P * 8, hipMemcpyDeviceToHost));
  return MPIC_OK;
}

int mpic_batch_get_phi(mpic_batch* b, int prob, double* phi) {
  if (!b || !phi || prob < 0 || prob >= b->nprob) return bfail(b, MPIC_ERR_ARG, "mpic_batch_get_phi: bad arguments");
  BCHK(b, hipStreamSynchronize(b->c->stream));
  const int32_t n0 = b->node_off[prob], n1 = b->node_off[prob + 1];
  BCHK(b, hipMemcpy(phi, b->c->phi + n0, (size_t)(n1 - n0) * 8, hipMemcpyDeviceToHost));
  return MPIC_OK;
}

int mpic_batch_get_rhs(mpic_batch* b, int prob, double* rhs) {
  if (!b || !rhs || prob < 0 || prob >= b->nprob) return bfail(b, MPIC_ERR_ARG, "mpic_batch_get_rhs: bad arguments");
  BCHK(b, hipStreamSynchronize(b->c->stream));
  const int32_t n0 = b->node_off[prob], n1 = b->node_off[prob + 1];
  BCHK(b, hipMemcpy(rhs, b->c->rhs + n0, (size_t)(n1 - n0) * 8, hipMemcpyDeviceToHost));
  return MPIC_OK;
}

int mpic_batch_download(mpic_batch* b, int prob, mpic_soa_host* h, int64_t* n) {
  if (!b || !h || prob < 0 || prob >= b->nprob || !b->filled)
    return bfail(b, MPIC_ERR_ARG, "mpic_batch_download: bad arguments");
  mpic_ctx* c = b->c;
  const int64_t p0 = b->part_off[prob], cnt = b->part_off[prob + 1] - p0;
  if (h->stride < cnt) return bfail(b, MPIC_ERR_ARG, "mpic_batch_download: stride below the problem's particle count");
  int rc;
  if ((rc = mixed_export(c))) return cfail(b, rc);  // applies the ranking: element order
  BCHK(b, hipStreamSynchronize(c->stream));
  std::vector<Rec64> rec(cnt);
  if (cnt) BCHK(b, hipMemcpy(rec.data(), c->P[c->cur].r + p0, cnt * sizeof(Rec64), hipMemcpyDeviceToHost));
  const int32_t e0 = b->elem_off[prob];
  const uint32_t id0 = (uint32_t)p0;
  for (int64_t i = 0; i < cnt; ++i) {
    const Rec64& r = rec[i];
    for (int d = 0; d < 3; ++d) {
      if (h->x) h->x[d * h->stride + i] = r.x[d];
      if (h->x_ref) h->x_ref[d * h->stride + i] = r.xr[d];
      if (h->v) h->v[d * h->stride + i] = r.v[d];
    }
    if (h->weight) h->weight[i] = r.w;
    if (h->elem) h->elem[i] = r.elem >= 0 ? r.elem - e0 : -1;
    if (h->id) h->id[i] = r.id - id0;
    if (h->type) h->type[i] = (int32_t)r.type;
  }
  if (n) *n = cnt;
  return MPIC_OK;
}

int mpic_batch_check_ownership(mpic_batch* b, int64_t* mismatches) {
  if (!b || !mismatches) return MPIC_ERR_ARG;
  return cfail(b, mpic_check_ownership(b->c, mismatches));
}

}  // extern "C"
