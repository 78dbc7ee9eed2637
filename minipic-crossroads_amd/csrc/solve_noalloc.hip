// mpic_solve_noalloc: the FP64 CG of mpic_solve (CG_solve.hpp:54-114) for
// callers that cannot allocate device memory during the call.  The solve's
// working memory beyond the context's vectors -- the solver state, the history
// of r.r of every iteration and the per-chunk partials of the three
// reductions -- lives in one file-scope device array (`scratch` below) instead
// of buffers that mpic_solve allocates when max_iter grows.
//
// Limits: max_iter <= NOALLOC_MAX_ITER, nodes <= NOALLOC_MAX_CHUNKS * BLOCK.
// Each GPU holds its own copy of the array (about 400 KB) until the process
// ends.  Every call on one GPU uses that GPU's copy: calls that overlap on one
// GPU (two host threads, or two streams) race on it.  The solver state is reset
// by a kernel at the start of each call; the partials and the history are
// written before they are read, so nothing else needs clearing.
#include <mutex>

#include "mpic_ctx.hpp"

namespace mpic {
namespace {

constexpr int NOALLOC_MAX_ITER = 2047;
constexpr int NOALLOC_MAX_CHUNKS = 16384;
constexpr int NOALLOC_MAX_DEVICES = 64;

struct NoAllocLayout {
  CGState st;
  double pad[8];
  double rho_hist[NOALLOC_MAX_ITER + 1];
  double part_bb[NOALLOC_MAX_CHUNKS];
  double part_rr[NOALLOC_MAX_CHUNKS];
  double part_pq[NOALLOC_MAX_CHUNKS];
};

// The scratch array.  In host code its name is only a stand-in: kernels get
// the device address from noalloc_prepare().
static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocLayout)];

std::mutex g_mu;
void* g_ptr[NOALLOC_MAX_DEVICES];
bool g_done[NOALLOC_MAX_DEVICES];

}  // namespace

void* noalloc_prepare(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    int cur = -1;
    void* p = nullptr;
    if (hipGetDevice(&cur) == hipSuccess && (cur == device || hipSetDevice(device) == hipSuccess)) {
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
      }
      if (cur != device) (void)hipSetDevice(cur);
    }
    g_ptr[device] = p;
    g_done[device] = true;
  }
  return g_ptr[device];
}

}  // namespace mpic

using namespace mpic;

extern "C" int mpic_solve_noalloc(mpic_ctx* c, double tol, int max_iter, int* iters, double* relres) {
  if (!c || !(tol >= 0) || max_iter < 0) return fail(c, MPIC_ERR_ARG, "mpic_solve_noalloc: bad arguments");
  if (max_iter > NOALLOC_MAX_ITER)
    return fail(c, MPIC_ERR_ARG, "mpic_solve_noalloc: max_iter " + std::to_string(max_iter) +
                                     " above the scratch's limit of " + std::to_string(NOALLOC_MAX_ITER));
  if (c->cg.nchunks > NOALLOC_MAX_CHUNKS)
    return fail(c, MPIC_ERR_ARG, "mpic_solve_noalloc: " + std::to_string(c->m.nn) +
                                     " nodes, above the scratch's limit of " +
                                     std::to_string(NOALLOC_MAX_CHUNKS * BLOCK));
  if (c->cfg.form == MPIC_FORM_MIXED || c->state32)
    return fail(c, MPIC_ERR_STATE, "mpic_solve_noalloc: the FP64 form's CG only");
  if (c->cfg.solver != MPIC_SOLVER_CG)
    return fail(c, MPIC_ERR_STATE, "mpic_solve_noalloc: the CG solver only (the context selects BiCGStab)");
  if (!c->noalloc_scratch) return fail(c, MPIC_ERR_HIP, "mpic_solve_noalloc: scratch address unavailable");
  NoAllocLayout* S = static_cast<NoAllocLayout*>(c->noalloc_scratch);
  CGBuffers b = c->cg;
  b.st = &S->st;
  b.rho_hist = S->rho_hist;
  b.part_bb = S->part_bb;
  b.part_rr = S->part_rr;
  b.part_pq = S->part_pq;
  tstart(c, T_SOLVE);
  const int rc = cg_run(c, b, tol, max_iter, iters, relres);
  tstop(c, T_SOLVE);
  (void)hipStreamSynchronize(c->stream);
  c->ms[T_SOLVE] = telapsed(c, T_SOLVE);
  return rc;
}
