// minipic_main -- the reference's driver (main.cpp:58-148) on the MI355X library,
// written against the reference's class interface (include/minipic/*.hpp).
//
// Options of the reference (main.cpp:66-83), as --name=value or --name value:
//   --nparts N        particles per element per fill           (10000)
//   --mesh FILE       brick file "nx ny nz mapping" on [-1,1]^3, 5 tets per hex (brick.txt)
//   --dt DT           time step                                (0.1)
//   --tfinal T        final time; steps while t < tfinal, t += dt (2.0)
//   --output K        write output.<t> every K steps (0 = never)
//   --kill-particles 1  remove particles that hit a wall
//   --problem NAME    box | none (e- and e+ fills, temp 1) | two-stream (mesh periodic in x, two
//                     counter-streaming electron beams and ions, BiCGStab at 1e-6) | plasma (see below)
//   --density D       fill density                             (2 pi^2 / 3)
//   --dump_mesh 1     write the mesh as text
// Options of the configurations this library is measured on:
//   --nx/--ny/--nz N --split 5|6 --unit-cube 1   generated brick instead of a file
//   --jitter J --jitter-seed S                   interior node jitter (+-J h)
//   --steps N         exactly N steps (overrides --tfinal)
//   --form fp64|mixed --cg-tol TOL --solver cg|bicgstab --sort-period N
//   --problem plasma  electrons only, temp 2 vth^2, uniform neutralising background,
//                     --vth V --pert A (weights 1 + A sin(2 pi x/L)) --seed S
//   --drift U         plasma: two counter-drifting halves of the electrons at +-U along x
//   --capacity-factor F  particle capacity = F * elements * nparts + 1000 (reference: 3)
//   --quiet 1         no per-step KE line
// Prints the reference's "Fill time", per-step "Total KE is", "Move time ...
// updates/second" and timer lines.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "minipic.hpp"

namespace {

struct Options {
  std::map<std::string, std::string> kv;
  std::string get(const std::string& k, const std::string& d) const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
  double num(const std::string& k, double d) const {
    auto it = kv.find(k);
    return it == kv.end() ? d : std::atof(it->second.c_str());
  }
  long integer(const std::string& k, long d) const {
    auto it = kv.find(k);
    return it == kv.end() ? d : std::atol(it->second.c_str());
  }
  bool has(const std::string& k) const { return kv.count(k) != 0; }
};

const char* kUsage =
    "usage: minipic_main [--nparts N] [--mesh FILE | --nx N --ny N --nz N [--split 5|6] [--unit-cube 1]]\n"
    "                    [--dt DT] [--tfinal T | --steps N] [--output K] [--kill-particles 1]\n"
    "                    [--problem box|none|two-stream|plasma] [--density D] [--dump_mesh 1]\n"
    "                    [--form fp64|mixed] [--cg-tol TOL] [--solver cg|bicgstab] [--sort-period N]\n"
    "                    [--vth V] [--pert A] [--drift U] [--seed S] [--jitter J] [--jitter-seed S]\n"
    "                    [--capacity-factor F] [--quiet 1]\n";

bool parse(int argc, char** argv, Options& o) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--help" || a == "-h") return false;
    if (a.rfind("--", 0) != 0) throw std::runtime_error("unexpected argument " + a);
    a = a.substr(2);
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      o.kv[a.substr(0, eq)] = a.substr(eq + 1);
    } else {
      if (i + 1 >= argc) throw std::runtime_error("missing value for --" + a);
      o.kv[a] = argv[++i];
    }
  }
  return true;
}

int run(const Options& o) {
  const int nppe = (int)o.integer("nparts", 10000);
  const double dt = o.num("dt", 0.1);
  const double t_final = o.num("tfinal", 2.0);
  const std::string problem = o.get("problem", "none");
  const double density = o.num("density", 2. / 3. * M_PI * M_PI);
  const int output = (int)o.integer("output", 0);
  const bool kill = o.integer("kill-particles", 0) == 1;
  const bool two_stream = problem == "two-stream";

  DataWarehouse data;
  if (o.has("nx")) {
    const int nx = (int)o.integer("nx", 8), ny = (int)o.integer("ny", nx), nz = (int)o.integer("nz", nx);
    const bool unit = o.integer("unit-cube", 0) == 1;
    const double lo[3] = {unit ? 0.0 : -1.0, unit ? 0.0 : -1.0, unit ? 0.0 : -1.0};
    const double hi[3] = {1.0, 1.0, 1.0};
    if (two_stream)
      data.create_mesh_periodic_brick(nx, ny, nz, (int)o.integer("split", 5), lo, hi, 0);
    else
      data.create_mesh_brick(nx, ny, nz, (int)o.integer("split", 5), lo, hi, o.num("jitter", 0.0),
                             (uint64_t)o.integer("jitter-seed", 0), 0);
  } else {
    auto comm = minipic::default_comm();  // main.cpp:70 (one rank)
    std::string meshfile = o.get("mesh", "brick.txt");
    const bool periodic = two_stream;
    data.create_mesh(meshfile, comm, periodic);  // main.cpp:89-92
  }
  mpic_config& cfg = data.config();
  cfg.form = o.get("form", "fp64") == "mixed" ? MPIC_FORM_MIXED : MPIC_FORM_FP64;
  // the two-stream problem runs the reference's solver: BiCGStab at 1e-6 (PIC.cpp:251)
  cfg.cg_tol = o.num("cg-tol", two_stream ? 1e-6 : cfg.cg_tol);
  cfg.solver = two_stream || o.get("solver", "cg") == "bicgstab" ? MPIC_SOLVER_BICGSTAB : MPIC_SOLVER_CG;
  cfg.sort_period = (int)o.integer("sort-period", cfg.sort_period);
  const bool plasma = problem == "plasma";
  if (plasma) data.set_background_density(density);
  const double cap_factor = o.num("capacity-factor", 3.0);
  data.create_particle_list((std::size_t)(cap_factor * data.mesh().num_elems * (double)nppe) + 1000);
  data.create_ES();
  data.set_kill_particles_on_wall_impact(kill);

  const auto t0 = std::chrono::steady_clock::now();
  if (two_stream) {  // main.cpp:98-106
    VECTOR v(3, 0.0);
    v[0] = 0.5;
    const FLOAT temp = -.05;
    ParticleFill fill1a(data, nppe / 2, 0, density / 2, v, temp);
    v[0] *= -1;
    ParticleFill fill1b(data, nppe / 2, 0, density / 2, v, temp);
    ParticleFill fill2(data, nppe, 2, density, VECTOR(3, 0.0), 0);
  } else if (plasma) {
    const double vth = o.num("vth", 0.1), drift = o.num("drift", 0.0), pert = o.num("pert", 0.0);
    const uint64_t seed = (uint64_t)o.integer("seed", 1);
    if (drift != 0.0) {
      VECTOR v(3, 0.0);
      v[0] = drift;
      ParticleFill a(data, nppe / 2, 0, density / 2, v, 2 * vth * vth, seed, pert);
      v[0] = -drift;
      ParticleFill b(data, nppe - nppe / 2, 0, density / 2, v, 2 * vth * vth, seed + 1, pert);
    } else {
      ParticleFill a(data, nppe, 0, density, VECTOR(3, 0.0), 2 * vth * vth, seed, pert);
    }
  } else {
    ParticleFill fill1(data, nppe, 0, density);  // main.cpp:108-109
    ParticleFill fill2(data, nppe, 1, density);
  }
  if (o.integer("dump_mesh", 0)) data.dump_mesh_data();
  PIC pic(data);
  pic.set_verbose(o.integer("quiet", 0) == 0);
  const double fill_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const long total_parts = (long)data.particles().num_particles();
  std::printf("Fill time %lf for %ld particles\n", fill_time, total_parts);

  const long steps_opt = o.integer("steps", -1);
  const auto t1 = std::chrono::steady_clock::now();
  int output_cnt = 0;
  long steps = 0;
  std::cout.precision(17);
  for (double t = 0; steps_opt >= 0 ? steps < steps_opt : t < t_final; t += dt, ++steps) {
    if (output && (output_cnt++) % output == 0) {  // main.cpp:125-136
      // The reference writes list positions 0, nppe, 2 nppe, ...; its list
      // keeps the fill order (the OpenMP path never reorders it), so these are
      // the particles of fill index 0, nppe, 2 nppe, ...: this build's list is
      // reordered by element, and the same particles are found by their
      // identity key (fill index), written in that order.
      char name[256];
      std::snprintf(name, sizeof(name), "output.%f", t);
      std::ofstream out(name);
      ParticleList& pl = data.particles_reference();
      const long nparts = (long)pl.num_particles();
      std::vector<std::pair<uint32_t, long>> pick;
      for (long i = 0; i < nparts; ++i)
        if (pl.id(i) % (uint32_t)nppe == 0) pick.emplace_back(pl.id(i), i);
      std::sort(pick.begin(), pick.end());
      for (const auto& q : pick)
        if (pl.type(q.second) == 0) {
          const VECTOR x = pl.x(q.second), v = pl.v(q.second);
          out << x[0] << " " << x[1] << " " << x[2] << " " << v[0] << " " << v[1] << " " << v[2] << std::endl;
        }
    }
    pic.time_step(dt);
  }
  const double loop_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  const long parts_end = (long)data.particles().num_particles();
  // main.cpp:140-146: the reference's rate counts one species over the nominal steps.
  std::printf("Move time %lf for %ld parts, or %5.1lfE6 updates/second \n", loop_time, parts_end,
              1e-6 * data.mesh().num_global_elems * nppe * t_final / (dt * loop_time));
  std::printf("Particle-steps %ld in %lf s: %.6e particle-steps/second (all particles, %ld steps)\n",
              total_parts * steps, loop_time, (double)total_parts * steps / loop_time, steps);
  pic.output_timers();
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    Options o;
    if (!parse(argc, argv, o)) {
      std::cout << kUsage;
      return 0;
    }
    return run(o);
  } catch (const minipic::Error& e) {
    std::cerr << "minipic_main: " << e.what() << std::endl;
    return 2;
  } catch (const std::exception& e) {
    std::cerr << "minipic_main: " << e.what() << "\n" << kUsage;
    return 2;
  }
}
