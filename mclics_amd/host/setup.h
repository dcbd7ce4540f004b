// Lattice setup of the mclics host program: point types, surface normals, anchorings and initial directors.
#ifndef MCLICS_SETUP_H_
#define MCLICS_SETUP_H_

#include <string>
#include <vector>

#include "mclgpu.h"
#include "params.h"

struct Lattice {
  int Nx = 0, Ny = 0, Nz = 0;
  std::vector<int> pt;        // 0 empty, 1 bulk, n >= 2 surface n - 2
  std::vector<float> ns;      // outward normals (3 per site, zero off the surfaces)
  std::vector<float> ni;      // directors (3 per site)
  std::vector<mcl_surface> surfaces;
  std::vector<std::string> surface_names;
  int periodic[3] = {0, 0, 0};
  long nvalid = 0;
};

// Setup in the order of simulator.cpp:157-251: evolution mode, geometry, boundaries, checkerboard check,
// anchorings, initial condition and potential. Every input error exits here with the reference's status, before
// any output file is touched.
Lattice setup_simulation(Params &p);

#endif
