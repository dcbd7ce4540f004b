// The GPU evolution of one simulation: a libmclgpu context holding the lattice, the mode loops and the writers.
#ifndef MCLICS_RUN_H_
#define MCLICS_RUN_H_

#include "mclgpu.h"
#include "params.h"
#include "setup.h"

class GpuRun {
 public:
  // Creates the device context; a device error exits with status 20 before any output file is touched.
  GpuRun(Params &p, Lattice &L);
  ~GpuRun();
  GpuRun(const GpuRun &) = delete;
  GpuRun &operator=(const GpuRun &) = delete;
  // Snapshot of the current directors (also used for ic.csv).
  int print_n(const char *fname);
  // The mode loop selected by `evol`: thermal, step, quench or electric (GPU suffix).
  int run();

 private:
  void sweeps(int n, int measure_every);
  Params &p_;
  Lattice &L_;
  void *ctx_ = nullptr;
  long long sweeps_ = 0;
  double sweep_seconds_ = 0.0;
};

#endif
