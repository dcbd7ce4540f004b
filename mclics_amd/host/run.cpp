// Mode loops of the mclics host program (thermal, step, quench, electric, each with or without the GPU suffix) and
// the output writers (po.dat and director_field_*.csv). The loops follow the reference's src/evolve_{thermal,step,
// quench,electric}.cu and .cpp: the same schedule, moments, column layout, number formats and file names. The
// sweeps, the FP64 measurements and the local order parameter of the snapshots come from the engine: libmclgpu for
// the GPU modes, the program's OpenMP sweep for the CPU modes.
#include "run.h"

#include "csvfmt.h"

#include <omp.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <string>
#include <vector>

namespace {

// Sweep schedule (include/sweep.h:13-34): points counted, not accumulated; snapshot names int(S x).
int sweep_points(double a, double b, double d) { return (int)floor((b - a) / d + 1e-3) + 1; }
double sweep_value(double a, double b, double d, int n, int npoints) {
  if (n == npoints - 1 && fabs((b - a) / d - (npoints - 1)) < 1e-3) return b;
  return a + n * d;
}
int snapshot_scale(double step) {
  int S = 100;
  while (S * fabs(step) < 1 - 1e-3 && S < 100000000) S *= 10;
  return S;
}
long snapshot_index(double x, int S) { return (long)(S * (x + 1e-3 / S)); }

struct Extras {
  bool pear, sph, cyl;
  void header(FILE *f) const {
    if (pear) fprintf(f, " P dtheta divn");
    if (sph) fprintf(f, pear ? " Sr_sph Pr_sph" : " Sr_sph");
    if (cyl) fprintf(f, pear ? " Sr_cyl Pr_cyl" : " Sr_cyl");
    fprintf(f, "\n");
  }
  void values(FILE *f, const mcl_moments &m) const {
    const double n = (double)m.n;
    if (pear) fprintf(f, " %g %g %g", m.P / n, m.dtheta / n, m.divn / n);
    if (sph) {
      fprintf(f, " %g", m.Sr_sph / n);
      if (pear) fprintf(f, " %g", m.Pr_sph / n);
    }
    if (cyl) {
      fprintf(f, " %g", m.Sr_cyl / n);
      if (pear) fprintf(f, " %g", m.Pr_cyl / n);
    }
    fprintf(f, "\n");
  }
};

}  // namespace

Run::Run(Params &p, Lattice &L, std::unique_ptr<Engine> engine) : p_(p), L_(L), eng_(std::move(engine)) {}

void Run::sweeps(int n, int measure_every) {
  if (n <= 0) return;
  const auto t0 = std::chrono::steady_clock::now();
  eng_->sweeps(n, measure_every);
  sweep_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  sweeps_ += n;
}

// Snapshot and ic.csv format (io.cpp:56-89): header x,y,z,nx,ny,nz,S,pt, x fastest, %g, S = local order
// parameter, NaN on empty sites. The local order parameter comes from the engine; the text is
// formatted by up to 16 threads, each on its own chunk of lines, and the chunks are written in order, so the file is
// byte-identical to printing the lines one by one.
int Run::print_n(const char *fname) {
  const size_t Nt = L_.pt.size();
  std::vector<float> S(Nt);
  eng_->snapshot(L_.ni.data(), S.data());
  FILE *out = fopen(fname, "w");
  if (!out) {
    perror(fname);
    return 1;
  }
  fputs("x,y,z,nx,ny,nz,S,pt\n", out);
  const size_t kChunk = 65536;  // lines per chunk
  const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<char>> bufs(nthreads, std::vector<char>(kChunk * mclcsv::kMaxLine));
  std::vector<size_t> used(nthreads);
  const int Nx = L_.Nx, Ny = L_.Ny;
  auto format = [&](int t, size_t s0, size_t s1) {
    char *p = bufs[t].data();
    int i = (int)(s0 % Nx), j = (int)((s0 / Nx) % Ny), k = (int)(s0 / ((size_t)Nx * Ny));
    for (size_t s = s0; s < s1; s++) {
      const int pt = L_.pt[s];
      p = mclcsv::put_line(p, i, j, k, L_.ni[3 * s], L_.ni[3 * s + 1], L_.ni[3 * s + 2], pt ? S[s] : NAN, pt);
      if (++i == Nx) {
        i = 0;
        if (++j == Ny) {
          j = 0;
          k++;
        }
      }
    }
    used[t] = (size_t)(p - bufs[t].data());
  };
  for (size_t b0 = 0; b0 < Nt; b0 += kChunk * nthreads) {
    std::vector<std::thread> pool;
    int n = 0;
    for (; n < nthreads && b0 + n * kChunk < Nt; n++) {
      const size_t s0 = b0 + n * kChunk, s1 = std::min(Nt, s0 + kChunk);
      if (n == 0) continue;  // chunk 0 is formatted by this thread below
      pool.emplace_back(format, n, s0, s1);
    }
    format(0, b0, std::min(Nt, b0 + kChunk));
    for (auto &th : pool) th.join();
    for (int t = 0; t < n; t++)
      if (fwrite(bufs[t].data(), 1, used[t], out) != used[t]) {
        perror(fname);
        fclose(out);
        return 1;
      }
  }
  printf("  Snapshot -> %s\n", fname);
  fflush(stdout);
  fclose(out);
  return 0;
}

int Run::run() {
  const std::string mode = mode_base(p_.evol);
  Extras ex{iequals(p_.potential, "pear"), iequals(p_.radial, "sphere") || iequals(p_.radial, "both"),
            iequals(p_.radial, "cylinder") || iequals(p_.radial, "both")};
  FILE *po = fopen("po.dat", "a");
  if (!po) {
    perror("po.dat");
    return 1;
  }
  char fname[1000];
  mcl_moments m;
  auto measure_block = [&](void) {
    eng_->reset_moments();
    sweeps(p_.MCS, 1);
    m = eng_->moments();
  };
  // the CPU modes name their thread count after the po.dat header (evolve_thermal.cpp:56-58)
  const bool cpu = strcmp(eng_->device(), "CPU") == 0;
  auto threads_line = [&](void) {
    if (!cpu) return;
    print_field("Threads:", omp_get_max_threads());
    printf("\n");
    fflush(stdout);
  };
  char label[32];
  auto title = [&](const char *name) {
    snprintf(label, sizeof(label), cpu ? "  %s evolution\n" : "  %s evolution (GPU)\n", name);
    printf("%s", label);
  };
  auto write_line = [&](const char *first) {
    const double S1 = m.S / m.n, S2 = m.S2 / m.n, E = m.E / m.n, E2 = m.E2 / m.n;
    fprintf(po, "%s %g %g %g %g", first, S1, S2 - S1 * S1, E, E2 - E * E);
    ex.values(po, m);
    fflush(po);
  };
  char first[64];
  if (iequals(mode, "thermal")) {
    title("Thermal");
    print_field("Ti:", p_.Ti);
    print_field("Tf:", p_.Tf);
    print_field("dT:", p_.dT);
    printf("\n");
    fprintf(po, "T S varS E varE");
    ex.header(po);
    fflush(po);
    threads_line();
    const int snapS = snapshot_scale(p_.dT);
    if (snapS != 100) printf("  Snapshot names: director_field_<int(%d T)>.csv (|dT| < 0.01)\n", snapS);
    const int nT = sweep_points(p_.Ti, p_.Tf, p_.dT);
    for (int it = 0; it < nT; it++) {
      p_.T = (float)sweep_value(p_.Ti, p_.Tf, p_.dT, it, nT);
      eng_->set_temperature(p_.T);
      sweeps(p_.MCT, 0);
      measure_block();
      snprintf(fname, sizeof(fname), "director_field_%ld.csv", snapshot_index(p_.T, snapS));
      print_n(fname);
      snprintf(first, sizeof(first), "%g", p_.T);
      write_line(first);
    }
  } else if (iequals(mode, "step")) {
    title("Step");
    print_field("First file:", p_.first_file);
    print_field("Last file:", p_.first_file + p_.fn);
    printf("\n");
    fprintf(po, "ii S varS E varE");
    ex.header(po);
    fflush(po);
    threads_line();
    p_.T = p_.Ti;
    eng_->set_temperature(p_.T);
    for (int ii = p_.first_file; ii < p_.fn + p_.first_file; ii++) {
      sweeps(p_.MCT, 0);
      measure_block();
      snprintf(fname, sizeof(fname), "director_field_%d.csv", ii);
      print_n(fname);
      snprintf(first, sizeof(first), "%d", ii);
      write_line(first);
    }
  } else if (iequals(mode, "quench")) {
    title("Quench");
    printf("  %-13s%g (MCT = %d steps)\n", "Ti:", p_.Ti, p_.MCT);
    printf("  %-13s%g (MCQ = %d steps, then MCS = %d measurement steps)\n", "Tf:", p_.Tf, p_.MCQ, p_.MCS);
    printf("  %-13s%d\n", "Repetitions:", p_.fn);
    printf("  %-13s%d\n\n", "First file:", p_.first_file);
    fprintf(po, "ii S varS E varE");
    ex.header(po);
    fflush(po);
    threads_line();
    for (int ii = p_.first_file; ii < p_.fn + p_.first_file; ii++) {
      p_.T = p_.Ti;
      eng_->set_temperature(p_.T);
      sweeps(p_.MCT, 0);
      p_.T = p_.Tf;
      eng_->set_temperature(p_.T);
      sweeps(p_.MCQ, 0);
      measure_block();
      snprintf(fname, sizeof(fname), "director_field_%d.csv", ii);
      print_n(fname);
      snprintf(first, sizeof(first), "%d", ii);
      write_line(first);
    }
  } else {  // electric
    title("Electric");
    print_field("Ei:", p_.elecEi);
    print_field("Ef:", p_.elecEf);
    print_field("dE:", p_.elecdE);
    printf("\n");
    p_.T = p_.Ti;
    eng_->set_temperature(p_.T);
    fprintf(po, "field S varS E varE");
    ex.header(po);
    fflush(po);
    threads_line();
    const int snapS = snapshot_scale(p_.elecdE);
    printf("  Snapshot names: director_field_<int(%d E)>.csv\n", snapS);
    const int nE = sweep_points(p_.elecEi, p_.elecEf, p_.elecdE);
    for (int ie = 0; ie < nE; ie++) {
      p_.elecE = (float)sweep_value(p_.elecEi, p_.elecEf, p_.elecdE, ie, nE);
      eng_->set_field(p_.elecE);
      sweeps(p_.MCT, 0);
      measure_block();
      snprintf(fname, sizeof(fname), "director_field_%ld.csv", snapshot_index(p_.elecE, snapS));
      print_n(fname);
      snprintf(first, sizeof(first), "%g", p_.elecE);
      write_line(first);
    }
  }
  fclose(po);
  const double updates = (double)sweeps_ * (double)L_.nvalid;
  printf("\n  %s sweeps: %lld, site updates: %.6g, sweep time: %.3f s (%.3f site-updates/ns)\n\n", eng_->device(),
         (long long)sweeps_, updates, sweep_seconds_,
         sweep_seconds_ > 0 ? updates / (sweep_seconds_ * 1e9) : 0.0);
  return 0;
}
