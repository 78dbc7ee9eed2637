// The reference driver's mesh setup, main.cpp:88-92, written against
// include/minipic: a communicator, DataWarehouse::create_mesh(meshfile, comm,
// periodic), then Mesh::dump_data's files (DataWarehouse::dump_mesh_data).
// Host code only: no GPU is touched.  Usage: mesh_dump_check MESHFILE PERIODIC OUTDIR
#include <cstdio>
#include <cstdlib>
#include <string>

#include "data_warehouse.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: mesh_dump_check MESHFILE PERIODIC OUTDIR\n");
    return 2;
  }
  try {
    auto comm = minipic::default_comm();  // Teuchos::DefaultComm<int>::getComm()
    std::string meshfile = argv[1];
    const bool periodic = std::atoi(argv[2]) != 0;
    DataWarehouse data;
    data.create_mesh(meshfile, comm, periodic);
    data.dump_mesh_data(argv[3]);
    if (comm->getRank() == 0)
      std::printf("rank %d of %d: %d nodes, %d elems\n", comm->getRank(), comm->getSize(), data.mesh().num_nodes,
                  data.mesh().num_elems);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
