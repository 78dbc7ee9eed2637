// particle_list_view.hpp -- the device side of ParticleList (particle_list.hpp:54-213)
// for HIP kernels on gfx950: accessors that return references into the
// particle storage, and the team-collective mutators push_back,
// delete_particle and add_particles_to_migrate_list.
//
// A team is a workgroup.  member_type wraps the workgroup's rank and size and
// an LDS scratch area of at least MPIC_TEAM_SCRATCH ints, as Kokkos' team member
// wraps team_rank/team_size/team_shmem:
//
//   __global__ void k(minipic::ParticleListView parts) {
//     __shared__ int scratch[MPIC_TEAM_SCRATCH];
//     const minipic::member_type dev(scratch);
//     ...
//     parts.x(i)[0] += dt * parts.v(i)[0];          // particle_list.hpp:71-75
//     parts.push_back(dev, p, add, dev.team_size()); // particle_list.hpp:103-139
//   }
//
// The view is obtained from the library with mpic_particles_view() (FP64 form)
// and returned with mpic_particles_view_commit(); kernels that use it run on
// the context's stream (mpic_stream()).  Storage is the library's: one 96-B
// record per particle (ParticleRecord) plus a copy of the element ids that the
// counting sort reads; the transient per-particle fields E, v_ref and
// time_remaining are component planes [3][capacity] / [capacity].
//
// Semantics kept from the reference:
//   push_back     every member of the team calls it with the team's arrays p[]
//                 and add[] (add[k] != 0: append p[k]); on return add[k] holds
//                 the new index of p[k], or -1.  One atomic per workgroup after
//                 a prefix count of the flags (ballots + LDS).  Appends beyond
//                 the capacity are dropped and reported by the commit
//                 (MPIC_ERR_CAPACITY) where the reference asserts.
//   delete_particle / add_particles_to_migrate_list with KILL_ONLY mark the
//                 particle dead (ielement = -1 in the record and in the copy);
//                 the commit (the reference's cleanup_list) compacts the list.
//   add_particles_to_migrate_list with a rank >= 0 records (index, rank) in the
//                 migrate list and marks the particle dead; one GPU has no other
//                 rank, so the commit reports the entries (num_migrate) and
//                 nothing is sent.
// is_dead(i) reads and writes the dead mark; B(i) is zero (PIC.cpp:186-189:
// the electrostatic gather sets B = 0) and cannot be written.
#ifndef MINIPIC_PARTICLE_LIST_VIEW_HPP_
#define MINIPIC_PARTICLE_LIST_VIEW_HPP_

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "minipic_c.h"

// ints of LDS scratch a member_type needs: two per wavefront of a 1024-thread
// workgroup, plus a 64-bit base index
#define MPIC_TEAM_SCRATCH 36

namespace minipic {

// One particle in the library's storage (96 B, six 16-B words).
struct ParticleRecord {
  double x[3];      // ParticleList::x      particle_list.hpp:71
  double x_ref[3];  // ParticleList::x_ref  particle_list.hpp:73
  double v[3];      // ParticleList::v      particle_list.hpp:75
  double weight;    // ParticleList::weight particle_list.hpp:85
  int32_t ielement; // ParticleList::ielement particle_list.hpp:69, -1 = dead
  uint32_t id;      // identity key
  int32_t type;     // ParticleList::type   particle_list.hpp:87
  uint32_t pad;
};
static_assert(sizeof(ParticleRecord) == 96, "particle record is 96 B");

// particle.hpp: the fields a team appends with push_back.
struct Particle {
  double x[3], x_ref[3], v[3], v_ref[3], E[3], B[3];
  double time_remaining, weight;
  long ielement;
  int type;
  bool is_dead;
  uint32_t id;
};

// A particle's 3 components in [3][stride] planes, used as VECTOR&.
struct PlaneVec {
  double* p;
  size_t stride;
  __device__ double& operator[](int d) const { return p[(size_t)d * stride]; }
};
// B of the electrostatic problem.
struct ZeroVec {
  __device__ double operator[](int) const { return 0.0; }
};
// is_dead(i) as bool&: reads ielement < 0; assigning true marks the particle dead.
struct DeadRef {
  int32_t* rec_elem;
  int32_t* mirror;
  __device__ operator bool() const { return *rec_elem < 0; }
  __device__ const DeadRef& operator=(bool dead) const {
    if (dead) {
      *rec_elem = -1;
      *mirror = -1;
    }
    return *this;
  }
};

// The workgroup as a Kokkos team member.
class member_type {
 public:
  __device__ explicit member_type(int* scratch) : scratch_(scratch) {}
  __device__ int team_rank() const { return (int)threadIdx.x; }
  __device__ int team_size() const { return (int)blockDim.x; }
  __device__ int league_rank() const { return (int)blockIdx.x; }
  __device__ int league_size() const { return (int)gridDim.x; }
  __device__ void team_barrier() const { __syncthreads(); }
  __device__ int* team_shmem() const { return scratch_; }

 private:
  int* scratch_;
};

struct ParticleListView {
  enum KILL_MIGRATE_FLAG { KILL_ONLY = -2, DO_NOTHING = -1 };  // particle_list.hpp:58

  ParticleRecord* rec;
  int32_t* elem_copy;
  double* E_;
  double* v_ref_;
  double* time_remaining_;
  unsigned long long* counters;  // [0] particles, [1] deleted, [2] migrate entries, [3] dropped appends
  int64_t* migrate_index;
  int32_t* migrate_proc;
  int64_t capacity, migrate_capacity;

  ParticleListView() = default;
  explicit ParticleListView(const mpic_particle_view& d)
      : rec(static_cast<ParticleRecord*>(d.records)),
        elem_copy(d.elem_copy),
        E_(d.E),
        v_ref_(d.v_ref),
        time_remaining_(d.time_remaining),
        counters(d.counters),
        migrate_index(d.migrate_index),
        migrate_proc(d.migrate_proc),
        capacity(d.capacity),
        migrate_capacity(d.migrate_capacity) {}

  // ---- accessors (particle_list.hpp:66-91) ----
  __device__ int32_t& ielement(size_t i) const { return rec[i].ielement; }
  __device__ double (&x(size_t i) const)[3] { return rec[i].x; }
  __device__ double (&x_ref(size_t i) const)[3] { return rec[i].x_ref; }
  __device__ double (&v(size_t i) const)[3] { return rec[i].v; }
  __device__ PlaneVec v_ref(size_t i) const { return PlaneVec{v_ref_ + i, (size_t)capacity}; }
  __device__ PlaneVec E(size_t i) const { return PlaneVec{E_ + i, (size_t)capacity}; }
  __device__ ZeroVec B(size_t) const { return ZeroVec{}; }
  __device__ double& time_remaining(size_t i) const { return time_remaining_[i]; }
  __device__ double& weight(size_t i) const { return rec[i].weight; }
  __device__ int32_t& type(size_t i) const { return rec[i].type; }
  __device__ uint32_t& id(size_t i) const { return rec[i].id; }
  __device__ DeadRef is_dead(size_t i) const { return DeadRef{&rec[i].ielement, &elem_copy[i]}; }
  __device__ size_t num_particles() const { return (size_t)counters[0]; }
  __device__ size_t num_migrate() const { return (size_t)counters[2]; }

  // ---- team-collective mutators ----

  // particle_list.hpp:103-139.  All members call it; list_size <= team_size
  // (member k handles entry k, where the reference's rank 0 loops over them).
  __device__ void push_back(const member_type& dev, const Particle* p, long* add, int list_size) const {
    const int k = dev.team_rank();
    const int lane = k & 63, wave = k >> 6, nwaves = (dev.team_size() + 63) >> 6;
    int* sc = dev.team_shmem();
    dev.team_barrier();  // the team's p[] and add[] are written
    const bool f = k < list_size && add[k] != 0;
    const unsigned long long b = __ballot(f);
    if (lane == 0) sc[wave] = __popcll(b);
    dev.team_barrier();
    if (k == 0) {
      int total = 0;
      for (int w = 0; w < nwaves; ++w) {
        const int c = sc[w];
        sc[w] = total;
        total += c;
      }
      const unsigned long long base = total ? atomicAdd(counters, (unsigned long long)total) : 0ull;
      sc[nwaves] = (int)(unsigned)base;
      sc[nwaves + 1] = (int)(unsigned)(base >> 32);
    }
    dev.team_barrier();
    if (k < list_size) {
      long idx = -1;
      if (f) {
        const unsigned long long base =
            (unsigned long long)(unsigned)sc[nwaves] | ((unsigned long long)(unsigned)sc[nwaves + 1] << 32);
        const long want = (long)base + sc[wave] + __popcll(b & ((1ull << lane) - 1ull));
        if (want < capacity) {
          idx = want;
          const Particle& q = p[k];
          ParticleRecord& r = rec[idx];
          for (int d = 0; d < 3; ++d) {
            r.x[d] = q.x[d];
            r.x_ref[d] = q.x_ref[d];
            r.v[d] = q.v[d];
          }
          r.weight = q.weight;
          r.ielement = q.is_dead ? -1 : (int32_t)q.ielement;
          r.id = q.id;
          r.type = q.type;
          r.pad = 0;
          elem_copy[idx] = r.ielement;
        } else {
          atomicAdd(&counters[3], 1ull);
        }
      }
      add[k] = idx;
    }
    dev.team_barrier();
  }

  // particle_list.hpp:144-156: marks particles i[0..num_deleted) dead (the
  // commit compacts the list).  All members call it.
  __device__ void delete_particle(const member_type& dev, size_t* i, size_t num_deleted) const {
    dev.team_barrier();
    if (dev.team_rank() == 0 && num_deleted) atomicAdd(&counters[1], (unsigned long long)num_deleted);
    for (size_t k = (size_t)dev.team_rank(); k < num_deleted; k += (size_t)dev.team_size()) is_dead(i[k]) = true;
    dev.team_barrier();
  }

  // particle_list.hpp:172-205.  proc_list[k] >= 0: migrate particle_id_list[k]
  // to that rank; KILL_ONLY: delete it; DO_NOTHING: keep it.  All members call
  // it; cnt <= team_size (member k handles entry k).
  __device__ void add_particles_to_migrate_list(const member_type& dev, long* particle_id_list, int* proc_list,
                                                int cnt) const {
    const int k = dev.team_rank();
    const int lane = k & 63, wave = k >> 6, nwaves = (dev.team_size() + 63) >> 6;
    int* sc = dev.team_shmem();
    dev.team_barrier();
    // per member: one entry (cnt <= team_size), counted by ballots
    const int pr = k < cnt ? proc_list[k] : DO_NOTHING;
    const bool mig = pr > DO_NOTHING, kill = pr == KILL_ONLY;
    const unsigned long long bm = __ballot(mig), bk = __ballot(kill);
    if (lane == 0) {
      sc[wave] = __popcll(bm);
      sc[nwaves + wave] = __popcll(bk);
    }
    dev.team_barrier();
    if (k == 0) {
      int nm = 0, nk = 0;
      for (int w = 0; w < nwaves; ++w) {
        const int c = sc[w];
        sc[w] = nm;
        nm += c;
        nk += sc[nwaves + w];
      }
      if (nk + nm) atomicAdd(&counters[1], (unsigned long long)(nk + nm));
      const unsigned long long base = nm ? atomicAdd(&counters[2], (unsigned long long)nm) : 0ull;
      sc[2 * nwaves] = (int)base;
    }
    dev.team_barrier();
    if (mig) {
      const long slot = (long)sc[2 * nwaves] + sc[wave] + __popcll(bm & ((1ull << lane) - 1ull));
      if (slot < migrate_capacity) {
        migrate_index[slot] = particle_id_list[k];
        migrate_proc[slot] = pr;
      } else {
        atomicAdd(&counters[3], 1ull);
      }
    }
    if (mig || kill) is_dead(particle_id_list[k]) = true;
    dev.team_barrier();
  }
};

}  // namespace minipic

#endif  // MINIPIC_PARTICLE_LIST_VIEW_HPP_
