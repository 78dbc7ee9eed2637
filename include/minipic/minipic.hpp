// minipic.hpp -- the reference's C++ class interface for the electrostatic PIC
// timestep, over the C ABI of minipic_c.h (libminipic.so, gfx950 HIP kernels).
//
// The classes keep the names, arguments and handle semantics of
// npe9/miniPIC-crossroads:
//   DataWarehouse  data_warehouse.hpp:51-90   owns the problem; accessors return shallow handles
//   Mesh           mesh.hpp:60-250            host geometry of the brick mesh
//   ES             ES.hpp:48-116              potential / rhs access, solve(tol)
//   ParticleTypeList / ParticleType  particle_type.hpp:44-86
//   ParticleList   particle_list.hpp:54-320   host view of the device particle list
//   ParticleFill   particle_fill.hpp:59-75    seeded fill on the device
//   ParticleMove   particle_move.hpp:351-408  kick + tet walk
//   PIC            PIC.hpp:54-106             weight_charge / weight_Efield / time_step / timers
// Copies of a handle share one problem (as Kokkos views share memory).  Errors
// throw minipic::Error carrying the library status and message, where the
// reference asserts.  One problem runs on one GPU: where the reference takes a
// Teuchos communicator, any communicator object is accepted and the single
// rank is used (minipic::default_comm() gives one with getRank()/getSize()).
#ifndef MINIPIC_MINIPIC_HPP_
#define MINIPIC_MINIPIC_HPP_

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "minipic_c.h"

// ---- scalar and vector types (types.hpp:47-60, const_len_vector.h) -----------
typedef int LO;
typedef long GO;
typedef double FLOAT;

template <class T, int N>
class ConstLenVector {
 public:
  ConstLenVector() {
    for (int i = 0; i < N; ++i) v_[i] = T();
  }
  explicit ConstLenVector(int) : ConstLenVector() {}
  ConstLenVector(int, const T& val) {
    for (int i = 0; i < N; ++i) v_[i] = val;
  }
  T& operator[](int i) { return v_[i]; }
  const T& operator[](int i) const { return v_[i]; }
  int size() const { return N; }
  T* data() { return v_; }
  const T* data() const { return v_; }

 private:
  T v_[N];
};
typedef ConstLenVector<FLOAT, 3> VECTOR;

namespace minipic {

// The communicator of a one-GPU run: rank 0 of 1 (the calls of
// Teuchos::Comm<int> that the reference's driver makes on it).
struct SingleRankComm {
  int getRank() const { return 0; }
  int getSize() const { return 1; }
  void barrier() const {}
};
// Teuchos::DefaultComm<int>::getComm() of a one-GPU run (an RCP-like handle).
inline std::shared_ptr<const SingleRankComm> default_comm() { return std::make_shared<const SingleRankComm>(); }

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {

inline void check(int rc, const mpic_ctx* ctx, const char* what) {
  if (rc != MPIC_OK) throw Error(rc, std::string(what) + ": " + mpic_last_error(ctx));
}

// State shared by every handle of one problem.
struct Core {
  mpic_mesh* mesh = nullptr;
  mpic_mesh_desc desc{};
  mpic_ctx* ctx = nullptr;
  mpic_config cfg{};
  std::vector<std::string> species_names{"e-", "e+", "H+"};
  double background_density = 0.0;
  bool es_created = false;
  int64_t capacity = 0;
  // version of the device state; host mirrors record the version they copied
  uint64_t version = 1;
  uint64_t parts_v = 0, start_v = 0, phi_v = 0, rhs_v = 0, efield_v = 0;
  int64_t nparts = 0;
  std::vector<double> x, xr, v, w, E;
  std::vector<int32_t> elem, type, start;
  std::vector<uint32_t> id;
  std::vector<double> phi, rhs;

  Core() { mpic_config_default(&cfg); }
  ~Core() {
    if (ctx) mpic_destroy(ctx);
    if (mesh) mpic_mesh_free(mesh);
  }
  Core(const Core&) = delete;
  Core& operator=(const Core&) = delete;

  void changed() { ++version; }
  mpic_ctx* context() {
    if (!ctx) throw Error(MPIC_ERR_STATE, "no problem on the device: call DataWarehouse::create_ES first");
    return ctx;
  }
  void sync_particles() {
    if (parts_v == version) return;
    mpic_ctx* c = context();
    int64_t n = 0;
    check(mpic_num_particles(c, &n), c, "ParticleList");
    x.assign(3 * n, 0.0);
    xr.assign(3 * n, 0.0);
    v.assign(3 * n, 0.0);
    w.assign(n, 0.0);
    elem.assign(n, -1);
    type.assign(n, 0);
    id.assign(n, 0);
    mpic_soa_host h{x.data(), xr.data(), v.data(), elem.data(), w.data(), type.data(), id.data(), n};
    check(mpic_particles_download(c, &h, &n), c, "ParticleList download");
    nparts = n;
    parts_v = version;
  }
  void sync_field(std::vector<double>& f, uint64_t& fv, bool want_phi) {
    if (fv == version) return;
    mpic_ctx* c = context();
    f.assign(desc.num_nodes, 0.0);
    check(want_phi ? mpic_get_phi(c, f.data()) : mpic_get_rhs(c, f.data()), c, want_phi ? "ES::phi" : "ES::rhs");
    fv = version;
  }
};

}  // namespace detail
}  // namespace minipic

// ---- Mesh (mesh.hpp:60-250) ------------------------------------------------
class Mesh {
 public:
  Mesh() = default;
  explicit Mesh(std::shared_ptr<minipic::detail::Core> c) : c_(std::move(c)) {
    num_nodes = c_->desc.num_nodes;
    num_elems = c_->desc.num_elems;
    num_global_elems = num_elems;
  }
  LO num_nodes = 0, num_elems = 0;
  GO num_global_elems = 0;

  VECTOR nodes(LO i) const {
    VECTOR r;
    for (int d = 0; d < 3; ++d) r[d] = c_->desc.nodes[(size_t)i * 3 + d];
    return r;
  }
  LO elem_to_node_lids(LO e, int k) const { return c_->desc.elem_nodes[(size_t)e * 4 + k]; }
  GO elem_face_to_elem_gids(LO e, int f) const { return c_->desc.elem_nbrs[(size_t)e * 4 + f]; }
  FLOAT determinate_jacobian(LO e) const { return c_->desc.det_jacobian[e]; }
  FLOAT inverse_jacobian(LO e, int r, int col) const { return c_->desc.inv_jacobian[(size_t)e * 9 + r * 3 + col]; }
  VECTOR normals(LO e, int f) const {
    VECTOR r;
    for (int d = 0; d < 3; ++d) r[d] = c_->desc.normals[(size_t)e * 12 + f * 3 + d];
    return r;
  }
  bool is_boundary_node(LO i) const { return c_->desc.boundary[i] != 0; }
  // Barycentric weights of a reference position (mesh.hpp:170-176).
  static void getWeights(FLOAT w[4], const VECTOR& xr) {
    w[0] = 1 - xr[0] - xr[1] - xr[2];
    w[1] = xr[0];
    w[2] = xr[1];
    w[3] = xr[2];
  }
  // refToPhysical, mesh.hpp:183-190.
  void refToPhysical(VECTOR& phys, const VECTOR& ref, LO e) const {
    FLOAT w[4];
    getWeights(w, ref);
    for (int d = 0; d < 3; ++d) phys[d] = 0;
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) phys[d] += w[i] * c_->desc.nodes[(size_t)elem_to_node_lids(e, i) * 3 + d];
  }
  // physToReference, mesh.hpp:197-211: one affine step from node 0.
  void physToReference(VECTOR& ref, const VECTOR& phys, LO e) const {
    VECTOR dl;
    const LO n0 = elem_to_node_lids(e, 0);
    for (int d = 0; d < 3; ++d) dl[d] = phys[d] - c_->desc.nodes[(size_t)n0 * 3 + d];
    physToReferenceVector(ref, dl, e);
  }
  // physToReferenceVector, mesh.hpp:245-250.
  void physToReferenceVector(VECTOR& out, const VECTOR& v, LO e) const {
    for (int r = 0; r < 3; ++r) {
      FLOAT acc = 0;
      for (int k = 0; k < 3; ++k) acc += inverse_jacobian(e, r, k) * v[k];
      out[r] = acc;
    }
  }

 private:
  std::shared_ptr<minipic::detail::Core> c_;
};

// ---- ParticleType / ParticleTypeList (particle_type.hpp:44-86) ---------------
class ParticleType {
 public:
  FLOAT q, m, q_over_m;
  ParticleType(FLOAT q_ = -1, FLOAT m_ = 1) : q(q_), m(m_), q_over_m(q_ / m_) {}
};

class ParticleTypeList {
 public:
  ParticleTypeList() = default;
  explicit ParticleTypeList(std::shared_ptr<minipic::detail::Core> c) : c_(std::move(c)) {}
  void add_particle_type(const ParticleType t, const std::string& name) {
    mpic_config& cfg = c_->cfg;
    if (cfg.num_species >= MPIC_MAX_SPECIES) throw minipic::Error(MPIC_ERR_ARG, "too many particle types");
    cfg.species_q[cfg.num_species] = t.q;
    cfg.species_m[cfg.num_species] = t.m;
    cfg.num_species++;
    c_->species_names.push_back(name);
    if (c_->ctx)
      minipic::detail::check(mpic_set_species(c_->ctx, cfg.num_species, cfg.species_q, cfg.species_m), c_->ctx,
                             "add_particle_type");
    c_->changed();
  }
  ParticleType particle_info(int i) const { return ParticleType(c_->cfg.species_q[i], c_->cfg.species_m[i]); }
  int count() const { return c_->cfg.num_species; }
  const std::string& name(int i) const { return c_->species_names.at(i); }

 private:
  std::shared_ptr<minipic::detail::Core> c_;
};

// ---- Particle (particle.hpp) and ParticleList (particle_list.hpp:54-320) -------
struct Particle {
  VECTOR x, x_ref, v, E, B;
  GO ielement = -1;
  FLOAT weight = 0;
  int type = 0;
  bool is_dead = false;
  uint32_t id = 0;
};

// Host-side ParticleList.  Accessors read a host mirror of the device list that
// is refreshed after any call that changed the particles; the device list itself
// is ordered by the library (element order after a sort).
class ParticleList {
 public:
  enum KILL_MIGRATE_FLAG { KILL_ONLY = -2, DO_NOTHING };
  ParticleList() = default;
  explicit ParticleList(std::shared_ptr<minipic::detail::Core> c) : c_(std::move(c)) {}

  std::size_t num_particles() const {
    if (!c_->ctx) return 0;
    int64_t n = 0;
    minipic::detail::check(mpic_num_particles(c_->ctx, &n), c_->ctx, "num_particles");
    return (std::size_t)n;
  }
  std::size_t num_migrate() const { return 0; }  // one rank: nothing migrates
  GO max_capacity() const { return (GO)c_->capacity; }
  GO current_capacity() const { return (GO)c_->capacity; }
  void set_time_remaining(FLOAT) const {}  // the walk keeps time_remaining per particle in registers
  // Dead particles are removed by the library's reordering after every move
  // that killed any (particle_list.cpp:148-191); nothing is left to clean.
  void cleanup_list() const {}

  // ParticleList::sort + compute_start_particles_by_element (particle_list.cpp:90-146).
  void sort() { compute_start_particles_by_element(); }
  void compute_start_particles_by_element() {
    mpic_ctx* c = c_->context();
    c_->start.assign((size_t)c_->desc.num_elems + 1, 0);
    minipic::detail::check(mpic_sort_by_element(c, c_->start.data()), c, "sort");
    c_->changed();
    c_->start_v = c_->version;
  }
  LO start(LO elem) const { return start_table()[elem]; }
  LO count_for_element(LO elem) const { return start_table()[elem + 1] - start_table()[elem]; }

  // Host copies of one particle's fields.
  GO ielement(std::size_t i) const { return host().elem.at(i); }
  VECTOR x(std::size_t i) const { return vec(host().x, i); }
  VECTOR x_ref(std::size_t i) const { return vec(host().xr, i); }
  VECTOR v(std::size_t i) const { return vec(host().v, i); }
  VECTOR E(std::size_t i) const {
    minipic::detail::Core& h = host();
    if (c_->efield_v != c_->version) {
      h.E.assign(3 * h.nparts, 0.0);
      minipic::detail::check(mpic_get_efield(c_->ctx, h.E.data(), h.nparts), c_->ctx, "ParticleList::E");
      c_->efield_v = c_->version;
    }
    return vec(h.E, i);
  }
  VECTOR B(std::size_t) const { return VECTOR(3, 0.0); }  // electrostatic: B = 0 (PIC.cpp:186-189)
  FLOAT weight(std::size_t i) const { return host().w.at(i); }
  int type(std::size_t i) const { return host().type.at(i); }
  bool is_dead(std::size_t i) const { return host().elem.at(i) < 0; }
  uint32_t id(std::size_t i) const { return host().id.at(i); }
  Particle particle(std::size_t i) const {
    Particle p;
    p.x = x(i);
    p.x_ref = x_ref(i);
    p.v = v(i);
    p.B = B(i);
    p.ielement = ielement(i);
    p.weight = weight(i);
    p.type = type(i);
    p.is_dead = is_dead(i);
    p.id = id(i);
    return p;
  }
  // dump_particles, particle_list.hpp:219-225 (same line format).
  void dump_particles(std::ostream& out) const {
    minipic::detail::Core& h = host();
    for (int64_t i = 0; i < h.nparts; ++i)
      out << h.elem[i] << "(" << h.x[i] << ", " << h.x[h.nparts + i] << ", " << h.x[2 * h.nparts + i] << ") "
          << "(" << h.v[i] << ", " << h.v[h.nparts + i] << ", " << h.v[2 * h.nparts + i] << ") " << h.type[i]
          << "\n";
  }

 private:
  minipic::detail::Core& host() const {
    c_->sync_particles();
    return *c_;
  }
  VECTOR vec(const std::vector<double>& a, std::size_t i) const {
    const int64_t n = c_->nparts;
    if ((int64_t)i >= n) throw minipic::Error(MPIC_ERR_ARG, "particle index out of range");
    VECTOR r;
    for (int d = 0; d < 3; ++d) r[d] = a[d * n + i];
    return r;
  }
  const std::vector<int32_t>& start_table() const {
    if (c_->start_v != c_->version) throw minipic::Error(MPIC_ERR_STATE, "call compute_start_particles_by_element first");
    return c_->start;
  }
  std::shared_ptr<minipic::detail::Core> c_;
};

// ---- ES (ES.hpp:48-116) ---------------------------------------------------------
class ES {
 public:
  ES() = default;
  explicit ES(std::shared_ptr<minipic::detail::Core> c) : c_(std::move(c)) {}
  FLOAT phi(LO lid) {
    c_->sync_field(c_->phi, c_->phi_v, true);
    return c_->phi.at(lid);
  }
  FLOAT rhs(LO lid) {
    c_->sync_field(c_->rhs, c_->rhs_v, false);
    return c_->rhs.at(lid);
  }
  // Adds the load vector of f to the rhs (ES.cpp:235-257).
  void set_analytic_RHS(FLOAT (*f)(VECTOR& x)) {
    std::vector<double> load(c_->desc.num_nodes);
    minipic::detail::check(mpic_mesh_load_vector(c_->mesh, &ES::call, (void*)&f, load.data()), nullptr,
                           "set_analytic_RHS");
    c_->sync_field(c_->rhs, c_->rhs_v, false);
    for (size_t i = 0; i < load.size(); ++i) load[i] += c_->rhs[i];
    minipic::detail::check(mpic_set_rhs(c_->context(), load.data()), c_->ctx, "set_analytic_RHS");
    c_->changed();
  }
  // Adds f at every node to the potential (ES.cpp:259-270).
  void set_analytic_phi(FLOAT (*f)(VECTOR& x)) {
    c_->sync_field(c_->phi, c_->phi_v, true);
    std::vector<double> p(c_->phi);
    for (LO i = 0; i < c_->desc.num_nodes; ++i) {
      VECTOR x;
      for (int d = 0; d < 3; ++d) x[d] = c_->desc.nodes[(size_t)i * 3 + d];
      p[i] += f(x);
    }
    minipic::detail::check(mpic_set_phi(c_->context(), p.data()), c_->ctx, "set_analytic_phi");
    c_->changed();
  }
  // The configured solver (CG from the current potential, CG_solve.hpp:54-114,
  // or BiCGStab from zero, :115-206) to |r| < tol |b|; false when it stops
  // without converging.
  bool solve(FLOAT tol = 1.e-5) {
    int iters = 0;
    double relres = 0;
    const int rc = mpic_solve(c_->context(), tol, c_->cfg.cg_max_iter, &iters, &relres);
    c_->changed();
    last_iters_ = iters;
    last_relres_ = relres;
    if (rc == MPIC_ERR_NOT_CONVERGED) return false;
    minipic::detail::check(rc, c_->ctx, "ES::solve");
    return true;
  }
  int last_iterations() const { return last_iters_; }
  double last_relative_residual() const { return last_relres_; }
  // The stiffness matrix, rhs and potential as MatrixMarket files (the writers
  // of ES::solve, ES.cpp:224-228: matrix.mm, b.mm, x.mm), 1-based coordinates,
  // 17 significant digits.
  void write_matrix_market(const std::string& prefix = "") {
    c_->sync_field(c_->rhs, c_->rhs_v, false);
    c_->sync_field(c_->phi, c_->phi_v, true);
    const mpic_mesh_desc& d = c_->desc;
    std::FILE* f = std::fopen((prefix + "matrix.mm").c_str(), "w");
    if (!f) throw minipic::Error(MPIC_ERR_ARG, "write_matrix_market: cannot open " + prefix + "matrix.mm");
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %lld\n", d.num_nodes, d.num_nodes,
                 (long long)d.nnz);
    for (LO r = 0; r < d.num_nodes; ++r)
      for (int32_t k = d.row_ptr[r]; k < d.row_ptr[r + 1]; ++k)
        std::fprintf(f, "%d %d %.17g\n", r + 1, d.col_idx[k] + 1, d.values[k]);
    std::fclose(f);
    const std::vector<double>* vec[2] = {&c_->rhs, &c_->phi};
    const char* name[2] = {"b.mm", "x.mm"};
    for (int v = 0; v < 2; ++v) {
      f = std::fopen((prefix + name[v]).c_str(), "w");
      if (!f) throw minipic::Error(MPIC_ERR_ARG, "write_matrix_market: cannot open " + prefix + name[v]);
      std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d 1\n", d.num_nodes);
      for (double x : *vec[v]) std::fprintf(f, "%.17g\n", x);
      std::fclose(f);
    }
  }

 private:
  static double call(const double* x, void* user) {
    FLOAT (*f)(VECTOR&) = *(FLOAT(**)(VECTOR&))user;
    VECTOR v;
    for (int d = 0; d < 3; ++d) v[d] = x[d];
    return f(v);
  }
  std::shared_ptr<minipic::detail::Core> c_;
  int last_iters_ = 0;
  double last_relres_ = 0;
};

// ---- DataWarehouse (data_warehouse.hpp:51-90, data_warehouse.cpp) -------------
class DataWarehouse {
 public:
  DataWarehouse()
      : c_(std::make_shared<minipic::detail::Core>()),
        mesh_(c_),
        particles_(c_),
        types_(c_),
        es_(c_),
        kill_particles_on_wall_impact_(false) {}
  ~DataWarehouse() = default;
  DataWarehouse(const DataWarehouse&) = delete;             // data_warehouse.hpp:81-82
  DataWarehouse& operator=(const DataWarehouse&) = delete;

  Mesh mesh() const { return mesh_; }
  ParticleList particles() const { return particles_; }
  ParticleTypeList particle_type_list() const { return types_; }
  class ES ES() const { return es_; }
  Mesh& mesh_reference() { return mesh_; }
  ParticleList& particles_reference() { return particles_; }
  ParticleTypeList& particle_type_list_reference() { return types_; }
  class ES& ES_reference() { return es_; }

  // create_mesh (data_warehouse.cpp:50-54, mesh.cpp:78-89): the brick file holds
  // "nx ny nz mapping"; [-1,1]^3 split into 5 tets per hex.  periodic = true joins
  // the x = -1 and x = +1 faces (the two-stream problem's mesh; nx even): its
  // stiffness matrix is not symmetric, so the solver becomes BiCGStab.
  // data_warehouse.hpp:66, create_mesh(std::string&, RCP<Comm>, bool): main.cpp:91
  // calls create_mesh(meshfile, comm, periodic).  The communicator is accepted
  // and unused (one rank).
  template <class Comm>
  void create_mesh(const std::string& in_name, const Comm& comm, bool periodic = false) {
    (void)comm;
    create_mesh(in_name, periodic);
  }
  void create_mesh(const std::string& in_name, bool periodic = false) {
    std::ifstream in(in_name.c_str());
    int nx = 0, ny = 0, nz = 0, mapping = 0;
    if (!(in >> nx >> ny >> nz)) throw minipic::Error(MPIC_ERR_ARG, "cannot read mesh file " + in_name);
    if (!(in >> mapping)) mapping = 0;
    const double lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
    if (periodic)
      create_mesh_periodic_brick(nx, ny, nz, 5, lo, hi, mapping);
    else
      create_mesh_brick(nx, ny, nz, 5, lo, hi, 0.0, 0, mapping);
  }
  // Brick joined in x (mpic_mesh_brick_periodic); selects BiCGStab.
  void create_mesh_periodic_brick(int nx, int ny, int nz, int split, const double lo[3], const double hi[3],
                                  int mapping = 0) {
    if (c_->mesh) throw minipic::Error(MPIC_ERR_STATE, "create_mesh called twice");
    minipic::detail::check(mpic_mesh_brick_periodic(nx, ny, nz, split, lo, hi, mapping, &c_->mesh), nullptr,
                           "create_mesh");
    minipic::detail::check(mpic_mesh_describe(c_->mesh, &c_->desc), nullptr, "create_mesh");
    c_->cfg.solver = MPIC_SOLVER_BICGSTAB;
    mesh_ = Mesh(c_);
  }
  // Brick of nx*ny*nz hexes on [lo,hi] split into 5 (reference) or 6 (Kuhn) tets,
  // interior nodes jittered by +-jitter*h.
  void create_mesh_brick(int nx, int ny, int nz, int split, const double lo[3], const double hi[3],
                         double jitter = 0.0, uint64_t jitter_seed = 0, int mapping = 0) {
    if (c_->mesh) throw minipic::Error(MPIC_ERR_STATE, "create_mesh called twice");
    minipic::detail::check(mpic_mesh_brick(nx, ny, nz, split, lo, hi, jitter, jitter_seed, mapping, &c_->mesh),
                           nullptr, "create_mesh");
    minipic::detail::check(mpic_mesh_describe(c_->mesh, &c_->desc), nullptr, "create_mesh");
    mesh_ = Mesh(c_);
  }
  // Uniform neutralising background density (an rhs load vector); before create_ES.
  void set_background_density(FLOAT f) {
    if (!c_->mesh || c_->ctx) throw minipic::Error(MPIC_ERR_STATE, "set_background_density: after create_mesh, before create_ES");
    minipic::detail::check(mpic_mesh_set_background(c_->mesh, f), nullptr, "set_background_density");
    minipic::detail::check(mpic_mesh_describe(c_->mesh, &c_->desc), nullptr, "set_background_density");
    c_->background_density = f;
  }
  // Library configuration (form, CG tolerance, solver, sort period); changes take
  // effect at create_ES.
  mpic_config& config() { return c_->cfg; }

  void create_particle_list(std::size_t max_capacity = 32768 * 1024) {
    c_->capacity = (int64_t)max_capacity;
    if (c_->ctx) reserve();
  }
  // create_ES (ES.cpp:47-213 + PIC.cpp:203-242): uploads the mesh, stiffness and
  // gradient tables to the GPU.
  void create_ES() {
    if (!c_->mesh) throw minipic::Error(MPIC_ERR_STATE, "create_mesh must precede create_ES");
    if (c_->ctx) return;
    c_->cfg.kill_on_wall = kill_particles_on_wall_impact_ ? 1 : 0;
    minipic::detail::check(mpic_create(&c_->desc, &c_->cfg, &c_->ctx), nullptr, "create_ES");
    c_->es_created = true;
    if (c_->capacity > 0) reserve();
  }
  void set_kill_particles_on_wall_impact(bool flag) {
    kill_particles_on_wall_impact_ = flag;
    c_->cfg.kill_on_wall = flag ? 1 : 0;
    if (c_->ctx) minipic::detail::check(mpic_set_kill_on_wall(c_->ctx, flag), c_->ctx, "set_kill");
  }
  bool kill_particles_on_wall_impact() { return kill_particles_on_wall_impact_; }
  // dump_mesh_data (data_warehouse.hpp, Mesh::dump_data mesh.cpp:794-910): the
  // rank-0 files of the reference, in its line formats (default stream
  // precision, a space after every value):
  //   nodes-000.dat  gid x y z          elems-000.dat  gid n0 n1 n2 n3
  //   faces-000.dat  gid a b c          edges-000.dat  gid a b
  //   owned-faces-000.dat, owned-edges-000.dat, boundary-faces-000.dat,
  //   boundary-edges-000.dat: one gid per line.
  // Faces and edges are numbered as Mesh::Mesh numbers them (mesh.cpp:233-315):
  // the sorted node triples (pairs) of every element's faces (edges), in
  // lexicographic order; on a periodic mesh a face lying on x = hi is keyed by
  // its x = lo image (Mesh::periodic_face, mesh.cpp:62-76).  A face (edge) is
  // on the boundary when its centroid is within 1 - 0.99999999 of a wall in
  // the box's [-1, 1] coordinates (mesh.cpp:604-631).  One rank owns all.
  void dump_mesh_data(const std::string& dir = ".") {
    if (!c_->mesh) return;
    const mpic_mesh_desc& d = c_->desc;
    const LO nn = d.num_nodes, ne = d.num_elems;
    const std::string base = dir.empty() ? std::string(".") : dir;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = hi[a] = d.nodes[a];
      for (LO i = 1; i < nn; ++i) {
        lo[a] = std::min(lo[a], d.nodes[(size_t)i * 3 + a]);
        hi[a] = std::max(hi[a], d.nodes[(size_t)i * 3 + a]);
      }
    }
    auto X = [&](LO n, int a) { return d.nodes[(size_t)n * 3 + a]; };
    // x = hi node -> its x = lo image (same y, z) on a periodic mesh
    std::map<std::pair<double, double>, LO> lo_plane;
    if (d.node_pair)
      for (LO i = 0; i < nn; ++i)
        if (X(i, 0) == lo[0]) lo_plane[{X(i, 1), X(i, 2)}] = i;
    std::map<std::array<LO, 3>, GO> faces;
    std::map<std::array<LO, 2>, GO> edges;
    static const int fv[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
    for (LO e = 0; e < ne; ++e) {
      const int32_t* en = d.elem_nodes + (size_t)e * 4;
      for (int f = 0; f < 4; ++f) {
        std::array<LO, 3> k{en[fv[f][0]], en[fv[f][1]], en[fv[f][2]]};
        if (d.node_pair && X(k[0], 0) == hi[0] && X(k[1], 0) == hi[0] && X(k[2], 0) == hi[0])
          for (LO& n : k) n = lo_plane.at({X(n, 1), X(n, 2)});
        std::sort(k.begin(), k.end());
        faces.emplace(k, 0);
      }
      for (int j = 0; j < 3; ++j)
        for (int k = j + 1; k < 4; ++k) {
          std::array<LO, 2> ed{std::min(en[j], en[k]), std::max(en[j], en[k])};
          edges.emplace(ed, 0);
        }
    }
    GO g = 0;
    for (auto& f : faces) f.second = g++;
    g = 0;
    for (auto& ed : edges) ed.second = g++;
    const double almost_one = 0.99999999;  // mesh.cpp:367
    auto on_wall = [&](const double* c) {
      for (int a = 0; a < 3; ++a) {
        const double s = (2.0 * c[a] - (lo[a] + hi[a])) / (hi[a] - lo[a]);
        if (s <= -almost_one || s >= almost_one) return true;
      }
      return false;
    };
    std::ofstream out;
    out.open(base + "/nodes-000.dat");
    for (LO i = 0; i < nn; ++i) out << i << ' ' << X(i, 0) << ' ' << X(i, 1) << ' ' << X(i, 2) << ' ' << std::endl;
    out.close();
    out.open(base + "/elems-000.dat");
    for (LO e = 0; e < ne; ++e) {
      out << e << ' ';
      for (int k = 0; k < 4; ++k) out << d.elem_nodes[(size_t)e * 4 + k] << ' ';
      out << std::endl;
    }
    out.close();
    std::vector<GO> bface, bedge;
    out.open(base + "/faces-000.dat");
    for (const auto& f : faces) {
      out << f.second << ' ' << f.first[0] << ' ' << f.first[1] << ' ' << f.first[2] << ' ' << std::endl;
      double c[3];
      for (int a = 0; a < 3; ++a) c[a] = (X(f.first[0], a) + X(f.first[1], a) + X(f.first[2], a)) / 3;
      if (on_wall(c)) bface.push_back(f.second);
    }
    out.close();
    out.open(base + "/edges-000.dat");
    for (const auto& ed : edges) {
      out << ed.second << ' ' << ed.first[0] << ' ' << ed.first[1] << ' ' << std::endl;
      double c[3];
      for (int a = 0; a < 3; ++a) c[a] = (X(ed.first[0], a) + X(ed.first[1], a)) / 2;
      if (on_wall(c)) bedge.push_back(ed.second);
    }
    out.close();
    out.open(base + "/owned-faces-000.dat");
    for (GO i = 0; i < (GO)faces.size(); ++i) out << i << std::endl;
    out.close();
    out.open(base + "/owned-edges-000.dat");
    for (GO i = 0; i < (GO)edges.size(); ++i) out << i << std::endl;
    out.close();
    out.open(base + "/boundary-faces-000.dat");
    for (GO i : bface) out << i << std::endl;
    out.close();
    out.open(base + "/boundary-edges-000.dat");
    for (GO i : bedge) out << i << std::endl;
    out.close();
  }

  // The library context behind the classes (for calls the classes do not wrap).
  mpic_ctx* context() { return c_->context(); }
  std::shared_ptr<minipic::detail::Core> core() { return c_; }

 private:
  void reserve() {
    minipic::detail::check(mpic_particles_reserve(c_->ctx, c_->capacity), c_->ctx, "create_particle_list");
    c_->changed();
  }
  std::shared_ptr<minipic::detail::Core> c_;
  Mesh mesh_;
  ParticleList particles_;
  ParticleTypeList types_;
  class ES es_;
  bool kill_particles_on_wall_impact_;
};

// ---- ParticleFill (particle_fill.hpp:59-164) ------------------------------------
// Appends num_parts_per_element particles of `type` to every element, seeded as
// the reference's FillFunctor (seed_offset 0 = the reference's seeds); velocities
// Maxwellian for temp > 0, vel for temp == 0, the two-stream perturbation for
// temp < 0.  pert_amp scales weights by 1 + pert_amp sin(2 pi (x - x_lo)/L_x).
class ParticleFill {
 public:
  ParticleFill(DataWarehouse& data, int num_parts_per_element, int type, FLOAT density, VECTOR vel = VECTOR(3, 0),
               FLOAT temp = 1.0, uint64_t seed_offset = 0, FLOAT pert_amp = 0.0) {
    mpic_ctx* c = data.context();
    const int rc = mpic_fill(c, num_parts_per_element, type, density, vel.data(), temp, seed_offset, pert_amp);
    if (rc == MPIC_ERR_CAPACITY) {
      std::cout << "Cannot create that number of particles, no fill performed\n";  // particle_fill.hpp:62
      return;
    }
    minipic::detail::check(rc, c, "ParticleFill");
    data.core()->changed();
  }
};

// ---- ParticleMove (particle_move.hpp:351-408) -----------------------------------
class ParticleMove {
 public:
  explicit ParticleMove(DataWarehouse& data) : data_(data), scale_(1.0) {}
  // Boris kick from the gathered E (AccelParticlesTag), then the tet walk.
  void move(double dt) {
    mpic_ctx* c = data_.context();
    minipic::detail::check(mpic_accel(c, dt, scale_), c, "ParticleMove::accel");
    const int rc = mpic_move(c, dt, &stats_);
    data_.core()->changed();
    minipic::detail::check(rc, c, "ParticleMove::move");
  }
  void set_force_scale(FLOAT scale = 1) { scale_ = scale; }
  const mpic_move_stats& last_stats() const { return stats_; }

 private:
  DataWarehouse& data_;
  FLOAT scale_;
  mpic_move_stats stats_{};
};

// ---- PIC (PIC.hpp:54-106, PIC.cpp:120-267) --------------------------------------
class PIC {
 public:
  explicit PIC(DataWarehouse& data) : data_(data) { data_.create_ES(); }
  // PIC::weight_charge: node charge of the current particles (+ background), 0 on
  // Dirichlet nodes.
  void weight_charge() {
    mpic_ctx* c = data_.context();
    minipic::detail::check(mpic_weight_charge(c), c, "weight_charge");
    data_.core()->changed();
  }
  // The solve of PIC::time_step at the configured tolerance.
  void solve_for_potential() { data_.ES_reference().solve(data_.config().cg_tol); }
  // PIC::weight_Efield: E of every particle from the potential.
  void weight_Efield() {
    mpic_ctx* c = data_.context();
    minipic::detail::check(mpic_weight_efield(c), c, "weight_Efield");
    data_.core()->changed();
  }
  // PIC::time_step: deposit, solve, gather, kick, walk, KE -- one fused device pass.
  void time_step(double dt) {
    mpic_ctx* c = data_.context();
    const int rc = mpic_time_step(c, dt, &last_);
    data_.core()->changed();
    minipic::detail::check(rc, c, "time_step");
    move_timer_ += last_.ms_push * 1e-3;
    weight_charge_timer_ += last_.ms_deposit * 1e-3;
    sort_timer_ += last_.ms_sort * 1e-3;
    solve_timer_ += last_.ms_solve * 1e-3;
    KE_timer_ += last_.ms_ke * 1e-3;
    weight_field_timer_ += last_.ms_field * 1e-3;
    if (verbose_) std::cout << "Total KE is " << last_.kinetic_energy << std::endl;  // PIC.cpp:262-263
  }
  // PIC::output_timers (PIC.hpp:85-88).  "Weight EField" is the per-element part
  // of the gather (potential or field per element); the per-particle gather is
  // fused into the move.
  void output_timers() {
    std::cout << "Timers are:\nMove " << move_timer_ << " Weight charge " << weight_charge_timer_ << ", Sorting "
              << sort_timer_ << ", Solve " << solve_timer_ << ", Weight EField " << weight_field_timer_
              << ", and KE compute " << KE_timer_ << std::endl;
  }
  const mpic_step_stats& last_step_stats() const { return last_; }
  void set_verbose(bool v) { verbose_ = v; }

 private:
  DataWarehouse& data_;
  mpic_step_stats last_{};
  bool verbose_ = true;
  double move_timer_ = 0, weight_charge_timer_ = 0, sort_timer_ = 0, solve_timer_ = 0, weight_field_timer_ = 0,
         KE_timer_ = 0;
};

#endif  // MINIPIC_MINIPIC_HPP_
