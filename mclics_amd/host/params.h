// Run parameters of the mclics host program and the input-file reader.
//
// The input format, parameter names, aliases, defaults and error behaviour are those of McLics (reference
// include/parameters.h:18-98, src/io.cpp:107-310 and 373-475, src/main.cpp:86-108), so existing input files run
// unchanged.
#ifndef MCLICS_PARAMS_H_
#define MCLICS_PARAMS_H_

#include <map>
#include <string>

struct Params {
  // lattice
  std::string geometry = "bulk";
  int Nx = 16, Ny = 16, Nz = 16;
  std::string xbound = "free", ybound = "free", zbound = "free";
  std::string bound_file;
  // surfaces, by surface number
  std::map<int, float> theta_s, phi_s, W;
  std::map<int, std::string> anchoring_type;
  // potential
  std::string potential = "ll";
  float A = 1.0f, B1 = 0.04f, C = 0.3f;
  float Ti = 1.2f, Tf = 0.0f, dT = -0.02f, T = 0.0f;
  float ghrl_rho = 0, ghrl_lambda = 0, ghrl_mu = 0, ghrl_nu = 1, ghrl_sigma = 0;
  // Monte Carlo
  std::string evol = "thermal";
  int MCS = 100, MCT = 100, MCQ = -1, fn = 1, first_file = 0, seed = 1;
  // initial condition
  std::string ic = "random", ic_file;
  float phi_0 = 0, theta_0 = 0, p0_i = 0;
  float k11 = 1, k22 = 1, k33 = 1, p0 = 0;
  // field
  float elecX = 0, elecY = 0, elecZ = 0, elecA = 0, elecP = 0;
  int flipMove = 0;
  std::string radial;
  float elecEi = 0, elecEf = 0, elecdE = 0, elecE = 1;
};

// Reads `name value` lines (case-insensitive names, aliases, per-surface values, '#' comments). A missing file
// exits with status 5, an unknown name with 3, an unreadable value with 1.
Params read_input_file(const char *fname);
// Derived parameters of main.cpp:88-106 (flip move, MCQ default, radial default, field normalization).
void derive_parameters(Params &p);
// Collects every impossible value, reports them all and exits with status 1 (io.cpp:381-454).
void validate_parameters(const Params &p);
// Periodic axes need an even size for the checkerboard update; exits with status 2 (io.cpp:462-475).
void check_checkerboard(const Params &p);
// GHRL coefficients from k11, k22, k33 and p0 (io.cpp:346-371).
void set_ghrl(Params &p);
void print_parameters(const Params &p);

bool iequals(const std::string &a, const char *b);
// Evolution mode without a trailing "GPU" suffix (io.cpp:384-385).
std::string mode_base(const std::string &evol);

// Terminal helpers (the "  Label: value" rows of io.cpp:23-52).
void print_separator(const char *seg = "-", int count = 60);
void print_header(const char *title, int count = 60, const char *seg = "-");
void print_field(const char *label, const char *value);
void print_field(const char *label, double value);
void print_field(const char *label, int value);

#endif
