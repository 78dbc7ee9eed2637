// HIP code written against the reference's ParticleList functor API
// (particle_list.hpp:54-213) through include/minipic/particle_list_view.hpp,
// launched on a context of libminipic.so.  Usage: view_check
//
//   1. device fill (ParticleFill, whose teams append with push_back), FP64 form;
//   2. a team functor reads every particle through the accessors and
//      - appends, with push_back, a copy of each particle whose id % 5 == 0
//        (new id = id + ID_SHIFT),
//      - deletes, with delete_particle, each particle whose id % 7 == 3,
//      - hands each particle whose id % 11 == 5 to add_particles_to_migrate_list
//        for rank 1, and each whose id % 13 == 6 with KILL_ONLY;
//   3. a reduction functor sums 0.5 w m |v|^2 through the accessors;
//   4. mpic_particles_view_commit (cleanup_list) compacts the list.
// Checks: the particle count, the surviving ids, every copy equal to its
// original field by field, the migrate count, the reduction against
// mpic_compute_ke, and a dropped append reported as MPIC_ERR_CAPACITY.
// Prints "view_check ok ..." and exits 0, or prints the failure and exits 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <vector>

#include "minipic_c.h"
#include "minipic/particle_list_view.hpp"

using minipic::member_type;
using minipic::Particle;
using minipic::ParticleListView;

constexpr int TEAM = 256;
constexpr uint32_t ID_SHIFT = 1u << 30;

__global__ __launch_bounds__(TEAM) void k_edit(ParticleListView parts, long n) {
  __shared__ int scratch[MPIC_TEAM_SCRATCH];
  __shared__ Particle p[TEAM];
  __shared__ long add[TEAM];
  __shared__ size_t del[TEAM];
  __shared__ long ids[TEAM];
  __shared__ int procs[TEAM];
  __shared__ int ndel;
  const member_type dev(scratch);
  const int k = dev.team_rank();
  if (k == 0) ndel = 0;
  dev.team_barrier();
  const long i = (long)dev.league_rank() * dev.team_size() + k;
  add[k] = 0;
  ids[k] = i;
  procs[k] = ParticleListView::DO_NOTHING;
  if (i < n) {
    const uint32_t id = parts.id(i);
    if (id % 5 == 0) {
      for (int d = 0; d < 3; ++d) {
        p[k].x[d] = parts.x(i)[d];
        p[k].x_ref[d] = parts.x_ref(i)[d];
        p[k].v[d] = parts.v(i)[d];
      }
      p[k].ielement = parts.ielement(i);
      p[k].weight = parts.weight(i);
      p[k].type = parts.type(i);
      p[k].is_dead = parts.is_dead(i);
      p[k].id = id + ID_SHIFT;
      add[k] = 1;
    }
    if (id % 7 == 3) del[atomicAdd(&ndel, 1)] = (size_t)i;
    if (id % 11 == 5) procs[k] = 1;
    else if (id % 13 == 6) procs[k] = ParticleListView::KILL_ONLY;
  }
  parts.push_back(dev, p, add, dev.team_size());
  parts.delete_particle(dev, del, (size_t)ndel);
  parts.add_particles_to_migrate_list(dev, ids, procs, dev.team_size());
}

// PICComputeKETag (PIC.cpp:191-200) through the accessors; mass of species 0..2.
__global__ __launch_bounds__(TEAM) void k_ke(ParticleListView parts, long n, double m0, double m1, double m2,
                                             double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || parts.is_dead(i)) return;
  const double m = parts.type(i) == 0 ? m0 : parts.type(i) == 1 ? m1 : m2;
  double v2 = 0;
  for (int d = 0; d < 3; ++d) v2 += parts.v(i)[d] * parts.v(i)[d];
  atomicAdd(out, 0.5 * (parts.weight(i) * m) * v2);
}

// Appends a copy of particle 0 from every member: beyond the capacity when full.
__global__ __launch_bounds__(TEAM) void k_flood(ParticleListView parts) {
  __shared__ int scratch[MPIC_TEAM_SCRATCH];
  __shared__ Particle p[TEAM];
  __shared__ long add[TEAM];
  const member_type dev(scratch);
  const int k = dev.team_rank();
  for (int d = 0; d < 3; ++d) {
    p[k].x[d] = parts.x(0)[d];
    p[k].x_ref[d] = parts.x_ref(0)[d];
    p[k].v[d] = parts.v(0)[d];
  }
  p[k].ielement = parts.ielement(0);
  p[k].weight = parts.weight(0);
  p[k].type = parts.type(0);
  p[k].is_dead = false;
  p[k].id = 0x7fffffffu;
  add[k] = 1;
  parts.push_back(dev, p, add, dev.team_size());
}

#define FAIL(...)                   \
  do {                              \
    std::printf("view_check: ");    \
    std::printf(__VA_ARGS__);       \
    std::printf("\n");              \
    return 1;                       \
  } while (0)
#define CHK(call)                                                                               \
  do {                                                                                          \
    const int rc_ = (call);                                                                     \
    if (rc_ != MPIC_OK) FAIL("%s -> %d (%s)", #call, rc_, mpic_last_error(ctx));                \
  } while (0)

struct Host {
  std::vector<double> x, xr, v, w;
  std::vector<int32_t> elem, type;
  std::vector<uint32_t> id;
  int64_t n = 0;
};

int download(mpic_ctx* ctx, int64_t n, Host& h) {
  h.x.assign(3 * n, 0);
  h.xr.assign(3 * n, 0);
  h.v.assign(3 * n, 0);
  h.w.assign(n, 0);
  h.elem.assign(n, 0);
  h.type.assign(n, 0);
  h.id.assign(n, 0);
  mpic_soa_host s{h.x.data(), h.xr.data(), h.v.data(), h.elem.data(), h.w.data(), h.type.data(), h.id.data(), n};
  return mpic_particles_download(ctx, &s, &h.n);
}

int main() {
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  mpic_mesh* mesh = nullptr;
  mpic_ctx* ctx = nullptr;
  if (mpic_mesh_brick(4, 4, 4, 6, lo, hi, 0.0, 0, 0, &mesh) != MPIC_OK) FAIL("mesh: %s", mpic_last_error(nullptr));
  mpic_mesh_desc desc;
  mpic_mesh_describe(mesh, &desc);
  mpic_config cfg;
  mpic_config_default(&cfg);
  const int rc0 = mpic_create(&desc, &cfg, &ctx);
  if (rc0 == MPIC_ERR_NO_DEVICE) {
    std::printf("view_check: no HIP device\n");
    return 1;
  }
  if (rc0 != MPIC_OK) FAIL("create: %s", mpic_last_error(nullptr));
  const int nppe = 8;
  const int64_t cap = (int64_t)desc.num_elems * nppe * 2;
  CHK(mpic_particles_reserve(ctx, cap));
  CHK(mpic_fill(ctx, nppe, 0, 2.0 / 3.0 * M_PI * M_PI, nullptr, 0.02, 1, 0.0));
  int64_t n0 = 0;
  CHK(mpic_num_particles(ctx, &n0));
  Host before;
  CHK(download(ctx, n0, before));
  void* sp = nullptr;
  CHK(mpic_stream(ctx, &sp));
  hipStream_t s = (hipStream_t)sp;
  // --- edit functor ---
  mpic_particle_view d;
  CHK(mpic_particles_view(ctx, &d));
  const ParticleListView parts(d);
  hipLaunchKernelGGL(k_edit, dim3((unsigned)((n0 + TEAM - 1) / TEAM)), dim3(TEAM), 0, s, parts, (long)n0);
  if (hipGetLastError() != hipSuccess) FAIL("k_edit launch");
  int64_t n1 = 0, nmig = 0;
  CHK(mpic_particles_view_commit(ctx, &n1, &nmig));
  // expected survivors and copies, by id
  std::map<uint32_t, int64_t> orig;
  for (int64_t i = 0; i < n0; ++i) orig[before.id[i]] = i;
  int64_t keep = 0, copies = 0, migrate = 0;
  for (const auto& kv : orig) {
    const uint32_t id = kv.first;
    if (id % 5 == 0) ++copies;
    const bool gone = id % 7 == 3 || id % 11 == 5 || id % 13 == 6;
    if (!gone) ++keep;
    if (id % 11 == 5) ++migrate;
  }
  if (n1 != keep + copies) FAIL("count %lld, expected %lld", (long long)n1, (long long)(keep + copies));
  if (nmig != migrate) FAIL("migrate entries %lld, expected %lld", (long long)nmig, (long long)migrate);
  Host after;
  CHK(download(ctx, n1, after));
  int64_t seen_copies = 0;
  for (int64_t j = 0; j < after.n; ++j) {
    const uint32_t id = after.id[j];
    const bool copy = id >= ID_SHIFT;
    const uint32_t oid = copy ? id - ID_SHIFT : id;
    auto it = orig.find(oid);
    if (it == orig.end()) FAIL("unknown id %u", id);
    const int64_t i = it->second;
    if (!copy && (oid % 7 == 3 || oid % 11 == 5 || oid % 13 == 6)) FAIL("deleted particle %u survived", oid);
    if (copy) {
      if (oid % 5 != 0) FAIL("unexpected copy of %u", oid);
      ++seen_copies;
    }
    for (int dd = 0; dd < 3; ++dd)
      if (after.x[dd * after.n + j] != before.x[dd * n0 + i] || after.xr[dd * after.n + j] != before.xr[dd * n0 + i] ||
          after.v[dd * after.n + j] != before.v[dd * n0 + i])
        FAIL("particle %u: fields differ from the original", id);
    if (after.w[j] != before.w[i] || after.elem[j] != before.elem[i] || after.type[j] != before.type[i])
      FAIL("particle %u: weight/element/type differ", id);
  }
  if (seen_copies != copies) FAIL("copies %lld, expected %lld", (long long)seen_copies, (long long)copies);
  // --- reduction functor against the library's KE ---
  double* dke = nullptr;
  if (hipMalloc(&dke, sizeof(double)) != hipSuccess) FAIL("hipMalloc");
  CHK(mpic_particles_view(ctx, &d));
  const ParticleListView parts2(d);
  if (hipMemsetAsync(dke, 0, sizeof(double), s) != hipSuccess) FAIL("memset");
  hipLaunchKernelGGL(k_ke, dim3((unsigned)((n1 + TEAM - 1) / TEAM)), dim3(TEAM), 0, s, parts2, (long)n1,
                     cfg.species_m[0], cfg.species_m[1], cfg.species_m[2], dke);
  CHK(mpic_particles_view_commit(ctx, nullptr, nullptr));
  double ke_view = 0, ke_lib = 0;
  if (hipMemcpy(&ke_view, dke, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) FAIL("memcpy");
  (void)hipFree(dke);
  CHK(mpic_compute_ke(ctx, &ke_lib));
  if (!(std::fabs(ke_view - ke_lib) <= 1e-12 * ke_lib)) FAIL("KE %.17g vs %.17g", ke_view, ke_lib);
  // --- appends beyond the capacity: dropped and reported ---
  CHK(mpic_particles_view(ctx, &d));
  const ParticleListView parts3(d);
  const int64_t room = cap - n1;
  const unsigned teams = (unsigned)(room / TEAM + 2);
  hipLaunchKernelGGL(k_flood, dim3(teams), dim3(TEAM), 0, s, parts3);
  int64_t n2 = 0;
  const int rc = mpic_particles_view_commit(ctx, &n2, nullptr);
  if (rc != MPIC_ERR_CAPACITY) FAIL("overflow commit returned %d", rc);
  if (n2 != cap) FAIL("after overflow count %lld, expected the capacity %lld", (long long)n2, (long long)cap);
  // the list still steps
  mpic_step_stats st;
  CHK(mpic_time_step(ctx, 0.05, &st));
  std::printf("view_check ok: %lld -> %lld particles (%lld copies, %lld migrate entries), KE %.12e\n",
              (long long)n0, (long long)n1, (long long)copies, (long long)nmig, ke_lib);
  mpic_destroy(ctx);
  mpic_mesh_free(mesh);
  return 0;
}
