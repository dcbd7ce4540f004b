// Input-file reader, derived parameters and validation (formats and messages of the reference, src/io.cpp).
#include "params.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <fstream>
#include <iostream>
#include <string>

bool iequals(const std::string &a, const char *b) { return strcasecmp(a.c_str(), b) == 0; }

std::string mode_base(const std::string &evol) {
  std::string m = evol;
  if (m.size() > 3 && strcasecmp(m.c_str() + m.size() - 3, "GPU") == 0) m.resize(m.size() - 3);
  return m;
}

void print_separator(const char *seg, int count) {
  for (int i = 0; i < count; i++) fputs(seg, stdout);
  putchar('\n');
}

void print_header(const char *title, int count, const char *seg) {
  print_separator(seg, count);
  int pad = (count - (int)strlen(title) - 2) / 2;
  if (pad < 0) pad = 0;
  printf("%*s %s %*s\n", pad, "", title, pad, "");
  print_separator(seg, count);
}

static const int kFieldWidth = 13;
void print_field(const char *label, const char *value) { printf("  %-*s%s\n", kFieldWidth, label, value); }
void print_field(const char *label, double value) { printf("  %-*s%g\n", kFieldWidth, label, value); }
void print_field(const char *label, int value) { printf("  %-*s%d\n", kFieldWidth, label, value); }

namespace {

// The reference keeps these values in fixed char arrays and rejects longer words (io.cpp:112-121).
void read_word(std::ifstream &in, std::string &dest, size_t size, const char *name) {
  std::string value;
  if (!(in >> value)) return;  // reported by check_stream
  if (value.size() >= size) {
    fprintf(stderr, "  [ERROR] Value of '%s' is too long (%zu characters, maximum %zu): %.60s...\n", name,
            value.size(), size - 1, value.c_str());
    exit(1);
  }
  dest = value;
}

// Discards the rest of the line; reports text other than a comment (io.cpp:123-130).
void finish_line(std::ifstream &in, const char *name) {
  if (in.fail()) return;
  std::string rest;
  std::getline(in, rest);
  size_t start = rest.find_first_not_of(" \t\r");
  if (start != std::string::npos && rest[start] != '#')
    fprintf(stderr, "  [WARNING] Ignoring extra text after parameter '%s': %s\n", name, rest.c_str() + start);
}

void check_stream(std::ifstream &in, const char *name) {
  if (in.eof()) return;
  if (in.fail() || in.bad()) {
    fprintf(stderr, "  [ERROR] Invalid value for parameter: %s\n", name);
    exit(1);
  }
}

template <typename T>
void read_value(std::ifstream &in, T &dest, const char *name) {
  in >> dest;
  finish_line(in, name);
}

template <typename T>
void read_surface_value(std::ifstream &in, std::map<int, T> &dest, const char *name) {
  int n = 0;
  in >> n;
  T v{};
  in >> v;
  if (!in.fail()) dest[n] = v;
  finish_line(in, name);
}

bool any_of(const char *name, std::initializer_list<const char *> options) {
  for (const char *o : options)
    if (strcasecmp(name, o) == 0) return true;
  return false;
}

int n_input_errors = 0;
void input_error(const char *msg) {
  fprintf(stderr, "  [ERROR] %s\n", msg);
  n_input_errors++;
}

}  // namespace

Params read_input_file(const char *fname) {
  Params p;
  if (fname == nullptr) {
    printf("  [!] No parameter file specified - using default parameter values.\n\n");
    return p;
  }
  std::ifstream in(fname);
  if (!in) {
    perror(fname);
    exit(5);
  }
  printf("  Parameter file : %s\n\n", fname);
  std::string token;
  while (in >> token) {
    if (token.size() >= 200) {
      fprintf(stderr, "  [ERROR] Parameter name too long: %.60s...\n", token.c_str());
      exit(1);
    }
    const char *t = token.c_str();
    if (any_of(t, {"nx"})) read_value(in, p.Nx, t);
    else if (any_of(t, {"ny"})) read_value(in, p.Ny, t);
    else if (any_of(t, {"nz"})) read_value(in, p.Nz, t);
    else if (any_of(t, {"A"})) read_value(in, p.A, t);
    else if (any_of(t, {"B1"})) read_value(in, p.B1, t);
    else if (any_of(t, {"C"})) read_value(in, p.C, t);
    else if (any_of(t, {"Ti"})) read_value(in, p.Ti, t);
    else if (any_of(t, {"Tf"})) read_value(in, p.Tf, t);
    else if (any_of(t, {"dT"})) read_value(in, p.dT, t);
    else if (any_of(t, {"MCS"})) read_value(in, p.MCS, t);
    else if (any_of(t, {"MCT"})) read_value(in, p.MCT, t);
    else if (any_of(t, {"MCQ", "quench_steps"})) {
      read_value(in, p.MCQ, t);
      if (!in.fail() && p.MCQ < 0) {
        fprintf(stderr, "  [ERROR] MCQ (quench_steps) must be >= 0 (got %d)\n", p.MCQ);
        exit(1);
      }
    } else if (any_of(t, {"seed"})) read_value(in, p.seed, t);
    else if (any_of(t, {"ic"})) { read_word(in, p.ic, 50, t); finish_line(in, t); }
    else if (any_of(t, {"ic_file"})) { read_word(in, p.ic_file, 1024, t); finish_line(in, t); }
    else if (any_of(t, {"xbound"})) { read_word(in, p.xbound, 20, t); finish_line(in, t); }
    else if (any_of(t, {"ybound"})) { read_word(in, p.ybound, 20, t); finish_line(in, t); }
    else if (any_of(t, {"zbound"})) { read_word(in, p.zbound, 20, t); finish_line(in, t); }
    else if (any_of(t, {"potential"})) { read_word(in, p.potential, 20, t); finish_line(in, t); }
    else if (any_of(t, {"evol", "mode"})) { read_word(in, p.evol, 20, t); finish_line(in, t); }
    else if (any_of(t, {"k11"})) read_value(in, p.k11, t);
    else if (any_of(t, {"k22"})) read_value(in, p.k22, t);
    else if (any_of(t, {"k33"})) read_value(in, p.k33, t);
    else if (any_of(t, {"p0"})) read_value(in, p.p0, t);
    else if (any_of(t, {"phi_0"})) read_value(in, p.phi_0, t);
    else if (any_of(t, {"theta_0"})) read_value(in, p.theta_0, t);
    else if (any_of(t, {"p0_i"})) read_value(in, p.p0_i, t);
    else if (any_of(t, {"fn"})) read_value(in, p.fn, t);
    else if (any_of(t, {"first_file_number", "initial_output"})) read_value(in, p.first_file, t);
    else if (t[0] == '#') {
      std::string comment;
      std::getline(in, comment);
    } else if (any_of(t, {"geometry"})) { read_word(in, p.geometry, 20, t); finish_line(in, t); }
    else if (any_of(t, {"boundary_file"})) { read_word(in, p.bound_file, 1024, t); finish_line(in, t); }
    else if (any_of(t, {"anchoring_type"})) read_surface_value(in, p.anchoring_type, t);
    else if (any_of(t, {"W"})) read_surface_value(in, p.W, t);
    else if (any_of(t, {"phi_s"})) read_surface_value(in, p.phi_s, t);
    else if (any_of(t, {"theta_s"})) read_surface_value(in, p.theta_s, t);
    else if (any_of(t, {"elecX", "Electric_field_x", "EF_x"})) read_value(in, p.elecX, t);
    else if (any_of(t, {"elecY", "Electric_field_y", "EF_y"})) read_value(in, p.elecY, t);
    else if (any_of(t, {"elecZ", "Electric_field_z", "EF_z"})) read_value(in, p.elecZ, t);
    else if (any_of(t, {"Ei", "initial_E"})) read_value(in, p.elecEi, t);
    else if (any_of(t, {"Ef", "final_E"})) read_value(in, p.elecEf, t);
    else if (any_of(t, {"dE", "E_variation"})) read_value(in, p.elecdE, t);
    else if (any_of(t, {"elecA", "dielectric_anisotropy", "Aniso_E"})) read_value(in, p.elecA, t);
    else if (any_of(t, {"elecP", "dipolar_coupling", "Polar_E"})) read_value(in, p.elecP, t);
    else if (any_of(t, {"radial"})) { read_word(in, p.radial, 20, t); finish_line(in, t); }
    else {
      fprintf(stderr, "  [ERROR] Unknown parameter: '%s'\n", t);
      exit(3);
    }
    check_stream(in, t);
  }
  return p;
}

void derive_parameters(Params &p) {
  // polar models need n -> -n moves (main.cpp:88)
  p.flipMove = (iequals(p.potential, "pear") || p.elecP != 0) ? 1 : 0;
  if (p.MCQ < 0) p.MCQ = p.MCT;
  if (p.radial.empty()) p.radial = iequals(p.geometry, "sphere") ? "sphere" : "none";
  if (!iequals(p.radial, "sphere") && !iequals(p.radial, "cylinder") && !iequals(p.radial, "both") &&
      !iequals(p.radial, "none")) {
    fprintf(stderr, "  [ERROR] radial must be sphere, cylinder, both or none (got '%s')\n", p.radial.c_str());
    exit(1);
  }
  if (p.elecA != 0 || p.elecP != 0) {
    float norm = sqrtf(p.elecX * p.elecX + p.elecY * p.elecY + p.elecZ * p.elecZ);
    if (norm == 0) {
      fprintf(stderr, "  [ERROR] elecA or elecP is set but the field direction (elecX, elecY, elecZ) is zero\n");
      exit(1);
    }
    p.elecX /= norm;
    p.elecY /= norm;
    p.elecZ /= norm;
  }
}

void validate_parameters(const Params &p) {
  char msg[400];
  const std::string mode = mode_base(p.evol);
  const char *m = mode.c_str();
  if (p.Nx < 1 || p.Ny < 1 || p.Nz < 1) {
    snprintf(msg, sizeof(msg), "Grid sizes must be >= 1 (got %d x %d x %d).", p.Nx, p.Ny, p.Nz);
    input_error(msg);
  }
  if (p.MCS < 1) {
    snprintf(msg, sizeof(msg), "MCS must be >= 1 (got %d): averages are taken over the MCS steps.", p.MCS);
    input_error(msg);
  }
  if (p.MCT < 0) {
    snprintf(msg, sizeof(msg), "MCT must be >= 0 (got %d).", p.MCT);
    input_error(msg);
  }
  if ((strcasecmp(m, "step") == 0 || strcasecmp(m, "quench") == 0) && p.fn < 1) {
    snprintf(msg, sizeof(msg), "fn must be >= 1 in %s mode (got %d).", m, p.fn);
    input_error(msg);
  }
  if (p.Ti < 0) {
    snprintf(msg, sizeof(msg), "Ti must be >= 0 (got %g).", p.Ti);
    input_error(msg);
  }
  if ((strcasecmp(m, "thermal") == 0 || strcasecmp(m, "quench") == 0) && p.Tf < 0) {
    snprintf(msg, sizeof(msg), "Tf must be >= 0 (got %g).", p.Tf);
    input_error(msg);
  }
  if (strcasecmp(m, "thermal") == 0) {
    if (p.dT == 0)
      input_error("dT must be nonzero in thermal mode.");
    else if ((p.Tf - p.Ti) * p.dT < 0) {
      snprintf(msg, sizeof(msg), "dT = %g does not lead from Ti = %g to Tf = %g (use dT < 0 to cool, dT > 0 to heat).",
               p.dT, p.Ti, p.Tf);
      input_error(msg);
    }
  }
  if (strcasecmp(m, "electric") == 0) {
    if (p.elecdE == 0)
      input_error("dE must be nonzero in electric mode.");
    else if ((p.elecEf - p.elecEi) * p.elecdE < 0) {
      snprintf(msg, sizeof(msg), "dE = %g does not lead from Ei = %g to Ef = %g.", p.elecdE, p.elecEi, p.elecEf);
      input_error(msg);
    }
  }
  if (iequals(p.potential, "ghrl") || iequals(p.potential, "gruhn-hess")) {
    if (p.k11 <= 0 || p.k22 <= 0 || p.k33 <= 0) {
      snprintf(msg, sizeof(msg), "k11, k22 and k33 must be > 0 (got %g, %g, %g).", p.k11, p.k22, p.k33);
      input_error(msg);
    } else if (p.k11 - 3 * p.k22 - p.k33 >= -1e-6 * p.k22) {
      snprintf(msg, sizeof(msg),
               "The GHRL potential needs k11 < 3 k22 + k33 (got k11 = %g, k22 = %g, k33 = %g): otherwise its energy "
               "scale nu is not negative and the pair potential is not nematic.",
               p.k11, p.k22, p.k33);
      input_error(msg);
    }
    if (p.p0 != 0 && fabs(p.p0) < 2 * M_PI)
      fprintf(stderr,
              "  [WARNING] |p0| = %g < 2 pi: the twist per lattice bond exceeds 90 degrees and the lattice has no\n"
              "            equilibrium helix (sin(phi) = 2 pi / p0 > 1).\n",
              fabs(p.p0));
  }
  if (iequals(p.ic, "cholesteric") && p.p0_i == 0 && p.p0 == 0)
    input_error("ic cholesteric needs a nonzero pitch: set p0_i (or p0).");
  if (n_input_errors > 0) {
    fprintf(stderr, "  %d error(s) in the input file. Nothing was run and no file was changed.\n", n_input_errors);
    exit(1);
  }
}

void check_checkerboard(const Params &p) {
  const char *axis[3] = {"x", "y", "z"};
  const std::string *type[3] = {&p.xbound, &p.ybound, &p.zbound};
  const int n[3] = {p.Nx, p.Ny, p.Nz};
  for (int a = 0; a < 3; a++) {
    if (iequals(*type[a], "periodic") && n[a] % 2 != 0) {
      fprintf(stderr,
              "  [ERROR] Periodic boundary along %s requires an even number of sites (N%s = %d).\n"
              "          The checkerboard update needs it; use an even size or a free boundary.\n",
              axis[a], axis[a], n[a]);
      exit(2);
    }
  }
}

void set_ghrl(Params &p) {
  // lambda = (L/9)(2 K1 - 3 K2 + K3), mu = L (K2 - K1), nu = (L/9)(K1 - 3 K2 - K3), rho = (L/9)(K1 - K3),
  // sigma = -L K2 2 pi / p0, with L chosen so that |nu| = 1 (the LL energy unit).
  const float k11 = p.k11, k22 = p.k22, k33 = p.k33, p0 = p.p0;
  const float Scale = (float)fabs(9.0 / (k11 - 3 * k22 - k33));
  p.ghrl_sigma = p0 ? (float)(-Scale * k22 * (2 * M_PI / p0)) : 0.0f;
  p.ghrl_lambda = (Scale / 9) * (2 * k11 - 3 * k22 + k33);
  p.ghrl_mu = Scale * (k22 - k11);
  p.ghrl_rho = (Scale / 9) * (k11 - k33);
  p.ghrl_nu = (Scale / 9) * (k11 - 3 * k22 - k33);
  print_header("ELASTIC PARAMETERS");
  printf("  %-12s %g\n", "k11:", k11);
  printf("  %-12s %g\n", "k22:", k22);
  printf("  %-12s %g\n", "k33:", k33);
  printf("  %-12s %g\n", "lambda:", p.ghrl_lambda);
  printf("  %-12s %g\n", "mu:", p.ghrl_mu);
  printf("  %-12s %g\n", "rho:", p.ghrl_rho);
  printf("  %-12s %g\n", "nu:", p.ghrl_nu);
  printf("  %-12s %g\n", "sigma:", p.ghrl_sigma);
  printf("  %-12s %g\n", "scale:", Scale);
  print_separator();
  printf("\n");
}

void print_parameters(const Params &p) {
  print_header("SIMULATION PARAMETERS");
  printf("  %-20s %dx%dx%d\n", "Grid:", p.Nx, p.Ny, p.Nz);
  printf("  %-20s %s\n", "Geometry:", p.geometry.c_str());
  printf("  %-20s %s\n", "Potential:", p.potential.c_str());
  printf("  %-20s %s\n", "Mode:", p.evol.c_str());
  printf("  %-20s %d steps\n", "MCS:", p.MCS);
  printf("  %-20s %d steps\n", "MCT:", p.MCT);
  if (strncasecmp(p.evol.c_str(), "quench", 6) == 0) printf("  %-20s %d steps\n", "MCQ:", p.MCQ);
  printf("  %-20s %d\n", "Seed:", p.seed);
  if (iequals(p.ic, "ic_file")) printf("  %-20s %s\n", "IC file:", p.ic_file.c_str());
  if (p.flipMove) printf("  %-20s %s\n", "Flip move:", "on (n -> -n, polar model)");
  if (!iequals(p.radial, "none")) printf("  %-20s %s\n", "Radial obs.:", p.radial.c_str());
  if (p.elecA || p.elecP) {
    printf("\n  Electric Field\n");
    printf("  %-20s (%g, %g, %g)\n", "Direction:", p.elecX, p.elecY, p.elecZ);
    printf("  %-20s %g\n", "Anisotropy:", p.elecA);
    printf("  %-20s %g\n", "Dipolar coupling:", p.elecP);
  }
  print_separator();
  printf("\n");
}
