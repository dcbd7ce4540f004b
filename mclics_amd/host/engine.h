// What the mode loops of the mclics program drive: one lattice's sweeps, its temperature and field, the FP64
// moments of its measurements and the snapshots. Two implementations: libmclgpu on the GPU (the evol modes with
// the GPU suffix) and the program's own OpenMP sweep on the host CPU (the modes without it).
#ifndef MCLICS_ENGINE_H_
#define MCLICS_ENGINE_H_

#include <memory>

#include "mclgpu.h"
#include "params.h"
#include "setup.h"

class Engine {
 public:
  virtual ~Engine() = default;
  // Label of the mode header and the run summary: "GPU" or "CPU".
  virtual const char *device() const = 0;
  virtual void set_temperature(float T) = 0;
  virtual void set_field(float elecE) = 0;
  // n sweeps; with measure_every > 0, a measurement is added to the moments after every measure_every-th sweep.
  // Returns when the sweeps are done.
  virtual void sweeps(int n, int measure_every) = 0;
  virtual void reset_moments() = 0;
  virtual mcl_moments moments() = 0;
  // Current directors (AoS, 3 per site) and the local order parameter of every site (parameter_order.cpp:75-127;
  // any value on empty sites).
  virtual void snapshot(float *ni, float *local_S) = 0;
};

// The libmclgpu engine; a device error exits with status 20 before any output file is touched.
std::unique_ptr<Engine> make_gpu_engine(const Params &p, const Lattice &L);
// The host engine of the CPU modes.
std::unique_ptr<Engine> make_cpu_engine(const Params &p, const Lattice &L);

#endif
