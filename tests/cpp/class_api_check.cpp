// Exercises the reference's class interface (include/minipic) on the GPU and
// writes what it computed as text for tests/test_cpp_api.py to compare with the
// oracle.  Usage: class_api_check OUTDIR
//   OUTDIR/split.txt      rhs, phi after weight_charge + ES::solve (one row per node)
//   OUTDIR/moved.txt      particles after weight_Efield + ParticleMove::move
//   OUTDIR/counts.txt     count_for_element(e) after compute_start_particles_by_element
//   OUTDIR/stepped.txt    particles after two PIC::time_step calls
//   OUTDIR/analytic.txt   phi of ES::solve on the load vector of sin(pi x) sin(pi y) sin(pi z)
//   OUTDIR/efield.txt     ParticleList::E after set_analytic_phi(1 + 2x + 3y - z) + weight_Efield
//   OUTDIR/matrix.mm, b.mm, x.mm  ES::write_matrix_market after that solve
//   OUTDIR/info.txt       key value lines
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "data_warehouse.hpp"
#include "particle_fill.hpp"
#include "particle_move.hpp"
#include "PIC.hpp"

namespace {

const double kDensity = 2.0 / 3.0 * M_PI * M_PI;

FLOAT source(VECTOR& x) { return std::sin(M_PI * x[0]) * std::sin(M_PI * x[1]) * std::sin(M_PI * x[2]); }
FLOAT linear(VECTOR& x) { return 1.0 + 2.0 * x[0] + 3.0 * x[1] - x[2]; }

void write_particles(const std::string& path, ParticleList& pl) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  const std::size_t n = pl.num_particles();
  for (std::size_t i = 0; i < n; ++i) {
    const Particle p = pl.particle(i);
    std::fprintf(f, "%u %ld %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p.id, p.ielement,
                 p.type, p.x[0], p.x[1], p.x[2], p.x_ref[0], p.x_ref[1], p.x_ref[2], p.v[0], p.v[1], p.v[2],
                 p.weight);
  }
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: class_api_check OUTDIR\n");
    return 2;
  }
  const std::string out = argv[1];
  try {
    std::ofstream info(out + "/info.txt");
    info.precision(17);
    {
      DataWarehouse data;
      const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      data.create_mesh_brick(4, 4, 4, 6, lo, hi);
      data.set_background_density(kDensity);
      data.config().cg_tol = 1e-10;
      Mesh mesh = data.mesh();
      data.create_particle_list((std::size_t)mesh.num_elems * 16 * 2);
      data.create_ES();
      data.set_kill_particles_on_wall_impact(false);
      ParticleFill fill(data, 16, 0, kDensity, VECTOR(3, 0.0), 0.02, 1, 0.0);
      PIC pic(data);
      pic.set_verbose(false);
      int S = 0;
      mpic_charge_scale(data.context(), &S);
      info << "charge_scale " << S << "\n";
      info << "num_particles " << data.particles().num_particles() << "\n";
      info << "num_elems " << mesh.num_elems << "\nnum_nodes " << mesh.num_nodes << "\n";
      // phase by phase: PIC.cpp:250-253
      pic.weight_charge();
      class ES& es = data.ES_reference();
      const bool ok = es.solve(1e-10);
      info << "split_solve_converged " << ok << "\nsplit_solve_iters " << es.last_iterations() << "\n";
      {
        std::FILE* f = std::fopen((out + "/split.txt").c_str(), "w");
        for (LO i = 0; i < mesh.num_nodes; ++i) std::fprintf(f, "%.17g %.17g\n", es.rhs(i), es.phi(i));
        std::fclose(f);
      }
      pic.weight_Efield();
      ParticleMove move(data);
      move.move(0.1);
      info << "move_crossings " << move.last_stats().crossings << "\n";
      ParticleList& pl = data.particles_reference();
      write_particles(out + "/moved.txt", pl);
      pl.compute_start_particles_by_element();
      {
        std::FILE* f = std::fopen((out + "/counts.txt").c_str(), "w");
        for (LO e = 0; e < mesh.num_elems; ++e) std::fprintf(f, "%d %d\n", pl.start(e), pl.count_for_element(e));
        std::fclose(f);
      }
      // fused steps: PIC::time_step
      for (int s = 0; s < 2; ++s) {
        pic.time_step(0.1);
        info << "step" << s << "_ke " << pic.last_step_stats().kinetic_energy << "\n";
        info << "step" << s << "_cg_iters " << pic.last_step_stats().cg_iters << "\n";
      }
      write_particles(out + "/stepped.txt", pl);
      std::ostringstream dump;
      pl.dump_particles(dump);
      const std::string text = dump.str();
      info << "dump_lines " << std::count(text.begin(), text.end(), '\n') << "\n";
    }
    {
      // ES on its own: analytic source and potential (ES.cpp:235-270).
      DataWarehouse data;
      const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      data.create_mesh_brick(8, 8, 8, 6, lo, hi);
      data.create_particle_list(data.mesh().num_elems * 4);
      data.create_ES();
      class ES& es = data.ES_reference();
      es.set_analytic_RHS(source);
      const bool ok = es.solve(1e-12);
      info << "analytic_converged " << ok << "\n";
      es.write_matrix_market(out + "/");  // matrix.mm, b.mm, x.mm
      Mesh mesh = data.mesh();
      {
        std::FILE* f = std::fopen((out + "/analytic.txt").c_str(), "w");
        for (LO i = 0; i < mesh.num_nodes; ++i) {
          const VECTOR x = mesh.nodes(i);
          std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g\n", x[0], x[1], x[2], es.rhs(i), es.phi(i));
        }
        std::fclose(f);
      }
      // replace phi by a linear function: the gathered field is its gradient
      ParticleFill fill(data, 4, 1, kDensity, VECTOR(3, 0.0), 0.0);
      std::vector<double> zero(mesh.num_nodes, 0.0);
      mpic_set_phi(data.context(), zero.data());
      data.core()->changed();
      es.set_analytic_phi(linear);
      PIC pic(data);
      pic.weight_Efield();
      ParticleList& pl = data.particles_reference();
      std::FILE* f = std::fopen((out + "/efield.txt").c_str(), "w");
      for (std::size_t i = 0; i < pl.num_particles(); ++i) {
        const VECTOR E = pl.E(i);
        std::fprintf(f, "%.17g %.17g %.17g\n", E[0], E[1], E[2]);
      }
      std::fclose(f);
    }
    info << "done 1\n";
  } catch (const minipic::Error& e) {
    std::fprintf(stderr, "class_api_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
