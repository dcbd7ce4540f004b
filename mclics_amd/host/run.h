// The evolution of one simulation: the mode loops and the writers, over a GPU or CPU engine.
#ifndef MCLICS_RUN_H_
#define MCLICS_RUN_H_

#include <memory>

#include "engine.h"
#include "params.h"
#include "setup.h"

class Run {
 public:
  Run(Params &p, Lattice &L, std::unique_ptr<Engine> engine);
  Run(const Run &) = delete;
  Run &operator=(const Run &) = delete;
  // Snapshot of the current directors (also used for ic.csv).
  int print_n(const char *fname);
  // The mode loop selected by `evol`: thermal, step, quench or electric (with or without the GPU suffix).
  int run();

 private:
  void sweeps(int n, int measure_every);
  Params &p_;
  Lattice &L_;
  std::unique_ptr<Engine> eng_;
  long long sweeps_ = 0;
  double sweep_seconds_ = 0.0;
};

#endif
