// mclics: Monte Carlo simulation of nematic and polar liquid crystals on a cubic lattice, with the checkerboard
// Metropolis sweep on an MI355X through libmclgpu (evol modes with the GPU suffix) or on the host CPU (the modes
// without it). Usage: mclics [input_file]. Outputs in the working directory: po.dat, ic.csv and
// director_field_*.csv, as the reference program (src/main.cpp:32-34).
#include <dirent.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include "params.h"
#include "run.h"
#include "setup.h"

namespace {

bool has_suffix(const char *name, const char *suffix) {
  const size_t n = strlen(name), s = strlen(suffix);
  return s <= n && strcmp(name + (n - s), suffix) == 0;
}

// Removes ic.csv, po.dat and director_field_*.csv of a previous run; called only after the whole setup has
// succeeded, so an input error never deletes results (main.cpp:46-81).
void clean_previous_output() {
  printf("  Removing output files from previous run... ");
  fflush(stdout);
  int removed = 0;
  DIR *dir = opendir(".");
  if (dir) {
    struct dirent *e;
    while ((e = readdir(dir)) != nullptr) {
      const char *name = e->d_name;
      if (strcmp(name, "ic.csv") == 0 || strcmp(name, "po.dat") == 0 ||
          (strncmp(name, "director_field_", 15) == 0 && has_suffix(name, ".csv"))) {
        remove(name);
        removed++;
      }
    }
    closedir(dir);
  }
  printf("%d file(s) removed.\n\n", removed);
}

}  // namespace

int main(int argc, char **argv) {
  print_header("McLics - Monte Carlo Liquid Crystal Simulator (MI355X engine)", 60, "=");
  printf("  GPU modes: sweeps and observables on the MI355X through libmclgpu (HIP, gfx950).\n");
  printf("  CPU modes: the host's OpenMP sweep.\n\n");
  Params p = read_input_file(argc > 1 ? argv[1] : nullptr);
  derive_parameters(p);
  validate_parameters(p);
  print_parameters(p);
  Lattice L = setup_simulation(p);
  // the engine is set up before the previous output is removed, so a device error leaves it in place
  const bool gpu = p.evol.size() > 3 && strcasecmp(p.evol.c_str() + p.evol.size() - 3, "GPU") == 0;
  Run run(p, L, gpu ? make_gpu_engine(p, L) : make_cpu_engine(p, L));
  clean_previous_output();
  run.print_n((char *)"ic.csv");
  run.run();
  print_header("SIMULATION COMPLETE", 60, "=");
  printf("\n");
  return 0;
}
