/*
 * minipic_c.h -- C ABI of the MI355X-native MiniPIC electrostatic timestep.
 *
 * One context owns one mesh on one GPU: all device memory and one HIP stream.
 * Calls are stream-ordered; a call that returns a host scalar synchronises.
 * One host thread per context.  Every function returns an int status
 * (MPIC_OK or a negative MPIC_ERR_*); exceptions never cross this ABI and
 * mpic_last_error() returns the message of the last failure on a context.
 *
 * The reference exposes this path as C++ classes, not an FFI.  Each entry point
 * below names the reference interface it replaces (file:line in
 * npe9/miniPIC-crossroads).  The C++ classes of include/minipic/ mirror those
 * reference classes on top of this ABI.
 *
 * Array layouts (host and device):
 *   particle positions / reference positions / velocities: component planes,
 *     element [d*stride + p] for component d of particle p;
 *   per-element arrays: element-major ([elem][...]).
 */
#ifndef MINIPIC_C_H_
#define MINIPIC_C_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPIC_OK 0
#define MPIC_ERR_ARG (-1)           /* bad argument or shape */
#define MPIC_ERR_CAPACITY (-2)      /* particle capacity exceeded (particle_fill.hpp:61-64) */
#define MPIC_ERR_NOT_CONVERGED (-3) /* solver hit max_iter (CG_solve.hpp:73,113) */
#define MPIC_ERR_HIP (-4)           /* a HIP runtime call failed */
#define MPIC_ERR_WALK_CAP (-5)      /* a particle exceeded the per-step crossing cap */
#define MPIC_ERR_STATE (-6)         /* call out of order (e.g. no particles) */
#define MPIC_ERR_NO_DEVICE (-7)     /* no usable gfx950 device */

#define MPIC_FORM_FP64 0  /* plain form: FP64 throughout, the drop-in */
#define MPIC_FORM_MIXED 1 /* precision form: FP32 particles, FP32 inner CG in FP64 refinement */

#define MPIC_SOLVER_CG 0       /* CG_solve, CG_solve.hpp:54-114: warm start from the current potential */
#define MPIC_SOLVER_BICGSTAB 1 /* BiCGStab_solve, CG_solve.hpp:115-206: from x = 0 (PIC::time_step, PIC.cpp:251) */

#define MPIC_MAX_SPECIES 16

typedef struct mpic_ctx mpic_ctx;   /* device state of one problem */
typedef struct mpic_mesh mpic_mesh; /* host-side mesh + discretisation */

/* Host arrays describing a P1 tet mesh and its discretisation.  Borrowed:
 * mpic_create copies them to the device during the call. */
typedef struct {
  int32_t num_nodes;
  int32_t num_elems;
  int64_t nnz;
  const double* nodes;          /* [num_nodes][3]          Mesh::nodes, mesh.hpp:111 */
  const int32_t* elem_nodes;    /* [num_elems][4]          Mesh::elem_to_node_lids, mesh.hpp:136 */
  const int32_t* elem_nbrs;     /* [num_elems][4]          Mesh::elem_face_to_elem_gids, mesh.hpp:144: the neighbour
                                   across face f, -1 at a wall, -2-e for neighbour e across a periodic face
                                   (Mesh::elem_face_periodic, mesh.cpp:376-389) */
  const double* normals;        /* [num_elems][4][3]       Mesh::normals, mesh.hpp:112 */
  const double* inv_jacobian;   /* [num_elems][3][3]       Mesh::inverse_jacobian(e,0), mesh.hpp:153 */
  const double* det_jacobian;   /* [num_elems]             Mesh::determinate_jacobian, mesh.hpp:154 */
  const double* grad_basis;     /* [num_elems][4][3]       PIC::grad_phi_basis_values_(e,i,*), PIC.cpp:213-241 */
  const double* wbv;            /* [num_elems][4][4]       ES::weighted_basis_values_, ES.cpp:193-195 */
  const uint8_t* boundary;      /* [num_nodes] 1 = Dirichlet node, Mesh::boundary_node_map, mesh.cpp:408-412 */
  const int32_t* row_ptr;       /* [num_nodes+1] stiffness CSR (ES::matrix_, ES.cpp:83-205) */
  const int32_t* col_idx;       /* [nnz], ascending within a row */
  const double* values;         /* [nnz] */
  const double* background;     /* [num_nodes] constant rhs term (ES::set_analytic_RHS, ES.cpp:235-257) or NULL */
  const int32_t* node_pair;     /* [num_nodes] or NULL: for a periodic image node, the node whose row carries its
                                   charge and stiffness (Mesh::node_pairs, mesh.cpp:178-189), else -1 */
  int32_t lattice[3];           /* cells (nx, ny, nz) when the nodes are the brick lattice numbered
                                   (i*(ny+1) + j)*(nz+1) + k (mesh.cpp:52-54) with the Dirichlet nodes on its walls
                                   (mpic_mesh_describe sets it for non-periodic bricks); all 0 otherwise.  The
                                   precision form's multigrid solve uses it (checked against num_nodes and the
                                   boundary flags; ignored when they disagree). */
} mpic_mesh_desc;

typedef struct {
  int form;                              /* MPIC_FORM_* */
  double cg_tol;                         /* relative residual tolerance (ES::solve(tol), ES.cpp:216) */
  int cg_max_iter;                       /* CG_solve.hpp:73 uses 1000 */
  int kill_on_wall;                      /* DataWarehouse::set_kill_particles_on_wall_impact, data_warehouse.hpp:76 */
  int sort_period;                       /* reorder the particle list every N time steps (0 = never) */
  int sort_bucket;                       /* elements per spatial bucket of that reordering (<= 1: by element;
                                            default 6, about one hex cell of the brick splits) */
  int num_species;                       /* ParticleTypeList, particle_type.hpp:58-86 */
  double species_q[MPIC_MAX_SPECIES];
  double species_m[MPIC_MAX_SPECIES];
  int solver;                            /* MPIC_SOLVER_*: the solve of mpic_solve and mpic_time_step (FP64 form;
                                            the precision form always runs its own CG, see multigrid) */
  int persistent_solve;                  /* 1: a solve is one cooperative launch when its grid fits on the GPU
                                            (same results, but slower on MI355X: its grid-wide waits cost more
                                            than a kernel boundary); 0 (default): one launch per iteration */
  int multigrid;                         /* precision form: 1 (default) solves with conjugate gradients in FP64
                                            preconditioned by an FP32 multigrid V-cycle when the mesh descriptor names
                                            a lattice that coarsens (mpic_mesh_desc.lattice); 0, or a mesh without
                                            one: the FP32 refinement CG.  Same stopping rule |r| < cg_tol |b|. */
} mpic_config;

/* Host particle SoA (planes as described above).  Borrowed for the call. */
typedef struct {
  double* x;       /* [3][stride]  ParticleList::x      particle_list.hpp:71 */
  double* x_ref;   /* [3][stride]  ParticleList::x_ref  particle_list.hpp:73 */
  double* v;       /* [3][stride]  ParticleList::v      particle_list.hpp:75 */
  int32_t* elem;   /* [stride]     ParticleList::ielement, -1 = dead  particle_list.hpp:69 */
  double* weight;  /* [stride]     ParticleList::weight particle_list.hpp:85 */
  int32_t* type;   /* [stride]     ParticleList::type   particle_list.hpp:87 */
  uint32_t* id;    /* [stride]     identity key (fill order), kept through sorts */
  int64_t stride;
} mpic_soa_host;

typedef struct {
  int64_t crossings;   /* element-face crossings in the tet walk */
  int64_t killed;      /* particles removed at a wall (kill_on_wall) */
  int64_t capped;      /* particles that hit the crossing cap (reference: unbounded loop) */
  int64_t nonpositive; /* sub-steps with s <= 0 (the reference asserts s > 0, particle_move.hpp:223) */
  int64_t reflections; /* specular wall reflections (particle_move.hpp:213-220) */
  int64_t relocated;   /* precision form: element changes made by the FP64 point-location test */
} mpic_move_stats;

typedef struct {
  int cg_iters;
  double cg_relres;
  double kinetic_energy;   /* PIC.cpp:255-264 */
  mpic_move_stats move;
  int64_t num_particles;   /* after the step */
  int charge_scale_bits;   /* fixed-point scale used for the node charge sum */
  double ms_deposit, ms_solve, ms_push, ms_sort, ms_ke;  /* GPU time of the phases; ms_sort is the
                                                            reordering done in this step (0 on steps without one) */
  double ms_field;         /* the per-element part of weight_Efield (PIC.cpp:162-189): potential at the
                              element's nodes (FP64 form) or the element's field (precision form); the
                              per-particle gather runs inside the push (ms_push) */
} mpic_step_stats;

/* ---- configuration -------------------------------------------------------- */
/* Defaults: FP64 form, CG tol 1e-10, max_iter 1000, reflect on walls, reorder
 * by element every step, species e-(-1,1) e+(1,1) H+(1,1837) as
 * particle_type.hpp:65-74. */
void mpic_config_default(mpic_config* cfg);
const char* mpic_last_error(const mpic_ctx* ctx); /* ctx may be NULL: last create/mesh error */

/* ---- host mesh (Mesh::Mesh mesh.cpp:78-714, ES::ES ES.cpp:47-213, PIC::PIC PIC.cpp:203-242) ---- */
/* Brick of nx*ny*nz hexes on [lo,hi], split = 5 (reference, mesh.cpp:83) or 6
 * (Kuhn); interior nodes jittered by +-jitter*h (0 = none); node mapping 0-3 as
 * the reference's mesh file field (mesh.cpp:85-89,763-792). */
int mpic_mesh_brick(int nx, int ny, int nz, int split, const double* lo, const double* hi, double jitter,
                    uint64_t jitter_seed, int mapping, mpic_mesh** out);
/* The same brick, periodic in x (the two-stream problem's mesh, Mesh::Mesh with
 * periodic = true, mesh.cpp:91-94,178-189,365-412): nx must be even; the x = lo
 * and x = hi faces are joined (a particle crossing one moves to the other with
 * x -> lo + hi - x, particle_move.hpp:190-191); nodes on x = hi (away from the
 * y and z walls) are images of their x = lo partners, whose rows take their
 * charge and stiffness (ES.cpp:91-110,168-189; PIC.cpp:132-134); an image row
 * is -phi_image + phi_partner = 0.  Dirichlet nodes: the y and z walls only.
 * The stiffness matrix is then not symmetric: solve with MPIC_SOLVER_BICGSTAB
 * (the reference's solver) in the FP64 form. */
int mpic_mesh_brick_periodic(int nx, int ny, int nz, int split, const double* lo, const double* hi, int mapping,
                             mpic_mesh** out);
/* Uniform background charge density f as an rhs load vector (ES::set_analytic_RHS with a constant). */
int mpic_mesh_set_background(mpic_mesh* mesh, double f);
/* Load vector of a source density f: out[i] = sum over elements and cubature points c of
 * -f(x_c) wbv[e][i][c], 0 at Dirichlet nodes (ES::set_analytic_RHS, ES.cpp:235-257).
 * out has num_nodes entries. */
int mpic_mesh_load_vector(const mpic_mesh* mesh, double (*f)(const double* x, void* user), void* user,
                          double* out);
/* Borrowed views of the mesh arrays (valid until mpic_mesh_free). */
int mpic_mesh_describe(const mpic_mesh* mesh, mpic_mesh_desc* out);
void mpic_mesh_free(mpic_mesh* mesh);

/* ---- context -------------------------------------------------------------- */
/* DataWarehouse::create_mesh/create_ES + PIC::PIC upload (data_warehouse.cpp:50-73). */
int mpic_create(const mpic_mesh_desc* mesh, const mpic_config* cfg, mpic_ctx** out);
void mpic_destroy(mpic_ctx* ctx);
/* MPIC_FORM_FP64: the drop-in (bit-identical to the reference's arithmetic).
 * MPIC_FORM_MIXED: FP32 particle state, push, walk and gather; FP64 point location
 * near faces; 64-bit fixed-point element deposit; FP32 CG inside FP64 iterative
 * refinement to the same |r| < tol |b| rule.  Switching converts the particle
 * state.  In the mixed form the phase-by-phase calls weight_efield, accel and
 * move return MPIC_ERR_STATE; weight_charge, solve, compute_ke, sort and
 * time_step run in that form. */
int mpic_set_form(mpic_ctx* ctx, int form);
/* DataWarehouse::set_kill_particles_on_wall_impact, data_warehouse.hpp:76. */
int mpic_set_kill_on_wall(mpic_ctx* ctx, int kill);
/* ParticleTypeList (particle_type.hpp:58-86): the charge and mass of species 0..n-1
 * (n <= MPIC_MAX_SPECIES); q/m is formed as ParticleType's constructor does (:51). */
int mpic_set_species(mpic_ctx* ctx, int n, const double* q, const double* m);
/* The configuration in effect (form, tolerances, species, ...). */
int mpic_get_config(mpic_ctx* ctx, mpic_config* cfg);

/* ---- particles (ParticleList, particle_list.hpp:54-320) -------------------- */
int mpic_particles_reserve(mpic_ctx* ctx, int64_t capacity);          /* DataWarehouse::create_particle_list */
int mpic_particles_upload(mpic_ctx* ctx, const mpic_soa_host* h, int64_t n);
int mpic_particles_download(mpic_ctx* ctx, mpic_soa_host* h, int64_t* n);
int mpic_num_particles(mpic_ctx* ctx, int64_t* n);                    /* ParticleList::num_particles */
/* ParticleFill(data, nppe, type, density, vel, temp), particle_fill.hpp:59-164, on
 * the device; appends.  seed_offset 0 = the reference's seeds; pert_amp scales
 * weights by 1 + pert_amp*sin(2 pi (x - lo_x)/L_x). */
int mpic_fill(mpic_ctx* ctx, int nppe, int type, double density, const double* vel, double temp,
              uint64_t seed_offset, double pert_amp);

/* ---- device access to the particle list (ParticleList's functor API,
 * particle_list.hpp:54-213) ------------------------------------------------------
 * For HIP kernels written against the reference's ParticleList accessors and
 * team-collective mutators: include/minipic/particle_list_view.hpp builds
 * minipic::ParticleListView from this descriptor.  All pointers are device
 * memory owned by the context. */
typedef struct {
  void* records;                 /* [capacity] 96-B particle records (minipic::ParticleRecord) */
  int32_t* elem_copy;            /* [capacity] element ids as the counting sort reads them */
  double* E;                     /* [3][capacity] ParticleList::E (written by mpic_weight_efield) */
  double* v_ref;                 /* [3][capacity] ParticleList::v_ref (written by mpic_accel) */
  double* time_remaining;        /* [capacity]    ParticleList::time_remaining */
  unsigned long long* counters;  /* [4]: particles, deleted, migrate entries, dropped appends */
  int64_t* migrate_index;        /* [migrate_capacity] migrate list: particle index ... */
  int32_t* migrate_proc;         /* [migrate_capacity] ... and destination rank */
  int64_t capacity, migrate_capacity;
} mpic_particle_view;

/* Opens the device list for user kernels (FP64 form): applies a pending
 * reordering, so records [0, n) are the live particles, and resets the
 * counters (particles = n).  Kernels that use the view run on mpic_stream(). */
int mpic_particles_view(mpic_ctx* ctx, mpic_particle_view* view);
/* Closes it (the reference's cleanup_list, particle_list.cpp:148-191): waits
 * for the stream, takes the particle count from the device counter, refreshes
 * the element copy from the records and compacts deleted particles; reports
 * the migrate entries (one GPU: nothing is sent).  MPIC_ERR_CAPACITY if an
 * append was dropped for lack of capacity (the rest is committed). */
int mpic_particles_view_commit(mpic_ctx* ctx, int64_t* num_particles, int64_t* num_migrate);
/* The context's HIP stream (hipStream_t), on which every call is ordered. */
int mpic_stream(mpic_ctx* ctx, void** stream);

/* ---- phases of PIC::time_step (PIC.cpp:245-267) ----------------------------- */
int mpic_weight_charge(mpic_ctx* ctx);                                 /* PIC::weight_charge, PIC.cpp:139-160 */
/* ES::solve, ES.cpp:216-231, with cfg.solver: CG_solve.hpp:54-114 (warm start) or
 * BiCGStab_solve, CG_solve.hpp:115-206 (from 0; MPIC_ERR_NOT_CONVERGED also for
 * its early exits).  In mpic_time_step a BiCGStab that does not converge does
 * not stop the step (PIC.cpp:251 ignores it); cg_relres reports the residual. */
int mpic_solve(mpic_ctx* ctx, double tol, int max_iter, int* iters, double* relres);
/* Which solve ES::solve runs in the precision form (no counterpart in the
 * reference, whose solver is fixed): *levels = the multigrid levels of its
 * preconditioned CG (fine mesh included), or 0 when it runs the refinement CG
 * (FP64 form, multigrid switched off, or a mesh whose descriptor names no
 * lattice that coarsens; mpic_last_error then gives the reason).  Prepares the
 * precision form's tables when needed. */
int mpic_multigrid_levels(mpic_ctx* ctx, int* levels);
/* No counterpart in the reference: mpic_solve's FP64 CG for callers that
 * cannot allocate device memory during a call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The solve's working
 * memory (solver state, per-iteration history, reduction partials) is a
 * file-scope device array of the library, one per GPU (about 400 KB, held
 * until the process ends); the call allocates nothing.  Same results as
 * mpic_solve bit for bit.  Limits: max_iter <= 2047 and at most 4,194,304
 * nodes (MPIC_ERR_ARG beyond); FP64 form with the CG solver (MPIC_ERR_STATE
 * otherwise).  Calls that overlap on one GPU race on the array. */
int mpic_solve_noalloc(mpic_ctx* ctx, double tol, int max_iter, int* iters, double* relres);
int mpic_weight_efield(mpic_ctx* ctx);                                 /* PIC::weight_Efield, PIC.cpp:162-189 */
int mpic_accel(mpic_ctx* ctx, double dt, double scale);                /* ParticleMove AccelParticlesTag, particle_move.hpp:364-392 */
int mpic_move(mpic_ctx* ctx, double dt, mpic_move_stats* stats);       /* MoveParticle::execute, particle_move.hpp:72-124 */
int mpic_compute_ke(mpic_ctx* ctx, double* ke);                        /* PICComputeKETag, PIC.cpp:191-200 */
int mpic_sort_by_element(mpic_ctx* ctx, int32_t* start_opt);           /* ParticleList::sort + compute_start_particles_by_element, particle_list.cpp:90-146 */
int mpic_time_step(mpic_ctx* ctx, double dt, mpic_step_stats* stats);  /* PIC::time_step, fused */

/* ---- field access ----------------------------------------------------------- */
int mpic_get_rhs(mpic_ctx* ctx, double* rhs);        /* ES::rhs(lid), ES.hpp:63 */
int mpic_get_phi(mpic_ctx* ctx, double* phi);        /* ES::phi(lid), ES.hpp:58 */
int mpic_set_phi(mpic_ctx* ctx, const double* phi);  /* ES::set_analytic_phi, ES.cpp:259-270 */
/* Replaces the right-hand side of the next mpic_solve (ES::set_analytic_RHS, ES.cpp:235-257);
 * the next weight_charge or time_step recomputes it from the particles. */
int mpic_set_rhs(mpic_ctx* ctx, const double* rhs);
int mpic_get_efield(mpic_ctx* ctx, double* E, int64_t stride); /* ParticleList::E after weight_efield, [3][stride] */
int mpic_charge_scale(mpic_ctx* ctx, int* bits);
/* Live particles whose element does not contain their position under the FP64
 * barycentric test (coordinate below -1e-9 across an interior face): the
 * ownership-mismatch count of the precision form (0 by construction). */
int mpic_check_ownership(mpic_ctx* ctx, int64_t* mismatches);

/* ---- batching form: an ensemble of independent problems --------------------------
 * Many problems of uneven size advanced together in the precision form: one
 * launch per phase for all problems (node charge, solve, field, push, sort,
 * per-problem statistics), the solve one workgroup per problem, the whole step
 * capturable in a HIP graph, and the tile shapes (push workgroups, threads per
 * problem in the solve) chosen by mpic_batch_tune.  Each problem follows the
 * reference's PIC::time_step (PIC.cpp:245-267) on its own mesh; particles never
 * leave their problem.  No counterpart in the reference (one problem per run). */
typedef struct mpic_batch mpic_batch;

typedef struct {
  int nppe;             /* particles of species 0 per element (ParticleFill, particle_fill.hpp:59) */
  double temp;          /* fill temperature (2 vth^2) */
  double density;       /* fill density; also the uniform neutralising background */
  double pert_amp;      /* weights scaled by 1 + pert_amp sin(2 pi (x - x_lo)/L_x) */
  double drift;         /* two counter-drifting halves at +-drift along x (0: one Maxwellian) */
  uint64_t seed_offset; /* 0 = the reference's seeds */
} mpic_batch_fill_desc;

typedef struct {
  int push_grid;        /* workgroups of the push (particles per workgroup = particles / push_grid) */
  int solve_threads;    /* threads per problem in the solve (CSR rows per wavefront = rows / (threads / 64)) */
  int tuned;            /* 1 after mpic_batch_tune */
  int graph;            /* 1 while steps replay captured HIP graphs */
  int max_inner_iters;  /* largest inner iteration count over the problems in the last step */
  int all_converged;
  int64_t num_particles;
  double ms_step;       /* device time of the last step */
  int64_t crossings;    /* element-face crossings of the last step, all problems */
  int64_t capped;       /* particles stopped by the crossing cap (the step then returns MPIC_ERR_WALK_CAP) */
  int64_t nonpositive;  /* sub-steps with a non-positive exit time */
  int ke_rescaled;      /* 1: the largest kinetic-energy term outgrew the fixed-point scale set at fill;
                           the per-problem sums were redone at a safe scale and the graphs recaptured */
  int multigrid;        /* 1: every problem solves by FP64 CG preconditioned by an FP32 multigrid V-cycle on
                           its own hierarchy (cfg.multigrid and every mesh a brick lattice of >= 3 cells per
                           direction); 0: FP32 CG inside FP64 refinement.  max_inner_iters then counts the
                           preconditioned iterations. */
} mpic_batch_stats;

/* Builds the ensemble from host meshes (mpic_mesh_brick), borrowed for the call. */
int mpic_batch_create(int nprob, mpic_mesh* const* meshes, const mpic_config* cfg, mpic_batch** out);
void mpic_batch_destroy(mpic_batch* b);
/* Fills every problem (fills[nprob]) and deposits the initial charge. */
int mpic_batch_fill(mpic_batch* b, const mpic_batch_fill_desc* fills);
/* Times the candidate tile shapes on the current state (nothing is advanced) and keeps the fastest. */
int mpic_batch_tune(mpic_batch* b, double dt);
/* One PIC::time_step of every problem.  With graphs enabled the step is a replay.
 * Returns MPIC_ERR_WALK_CAP, after completing the step, when a particle hit the
 * crossing cap (as mpic_time_step does). */
int mpic_batch_step(mpic_batch* b, double dt, mpic_batch_stats* st);
/* enable = 1: capture the step (dt of the next call) and replay it from then on. */
int mpic_batch_graph(mpic_batch* b, int enable);
/* Per-problem results of the last step: kinetic energy, inner iterations, outer
 * relative residual, converged flag, particle count (arrays of nprob; any may be NULL). */
int mpic_batch_problem_stats(mpic_batch* b, double* ke, int* iters, double* relres, int* converged, int64_t* nparts);
int mpic_batch_get_phi(mpic_batch* b, int prob, double* phi);   /* the problem's node count of entries */
int mpic_batch_get_rhs(mpic_batch* b, int prob, double* rhs);
/* The problem's particles (element ids local to the problem). */
int mpic_batch_download(mpic_batch* b, int prob, mpic_soa_host* h, int64_t* n);
int mpic_batch_check_ownership(mpic_batch* b, int64_t* mismatches);
const char* mpic_batch_last_error(const mpic_batch* b);

/* ---- instrumentation (PIC::output_timers, PIC.hpp:85-88) --------------------- */
/* Average device duration (ms) of the last launch of the named kernel family:
 * 0 push, 1 deposit, 2 node gather, 3 CG iteration, 4 sort scatter. */
int mpic_kernel_time(mpic_ctx* ctx, int which, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* MINIPIC_C_H_ */
