// Geometry, anchoring and initial-condition setup (behaviour of the reference's src/geometry_*.cpp,
// src/anchoring*.cpp, src/ic.cpp and src/simulator.cpp; file formats and error statuses kept).
#include "setup.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <fstream>
#include <string>
#include <vector>

#include "philox.h"

namespace {

inline size_t lin(const Lattice &L, int i, int j, int k) {
  return (size_t)i + (size_t)L.Nx * ((size_t)j + (size_t)L.Ny * k);
}

// Splits one line of a comma-separated file into trimmed fields (io.cpp:92-105).
std::vector<std::string> split_csv(const std::string &line) {
  std::vector<std::string> out;
  std::string field;
  for (size_t k = 0; k <= line.size(); k++) {
    if (k == line.size() || line[k] == ',') {
      size_t a = field.find_first_not_of(" \t\r"), b = field.find_last_not_of(" \t\r");
      out.push_back(a == std::string::npos ? "" : field.substr(a, b - a + 1));
      field.clear();
    } else {
      field += line[k];
    }
  }
  return out;
}

int boundary_kind(const std::string &type, const char *axis) {
  if (iequals(type, "free")) return 0;
  if (iequals(type, "periodic")) return 1;
  fprintf(stderr, "  [ERROR] %s boundary '%s' not implemented. Options: free | periodic\n", axis, type.c_str());
  exit(2);
}

// ---- geometries (geometry_bulk.cpp:61-65, geometry_slab.cpp:51-71, geometry_sphere.cpp:32-58) ----

void geometry_slab(Lattice &L) {
  for (int i = 0; i < L.Nx; i++)
    for (int j = 0; j < L.Ny; j++) {
      size_t s = lin(L, i, j, 0);
      L.pt[s] = 2;
      L.ns[3 * s + 0] = 0;
      L.ns[3 * s + 1] = 0;
      L.ns[3 * s + 2] = -1;
      for (int k = 1; k < L.Nz - 1; k++) L.pt[lin(L, i, j, k)] = 1;
      s = lin(L, i, j, L.Nz - 1);
      L.pt[s] = 3;
      L.ns[3 * s + 0] = 0;
      L.ns[3 * s + 1] = 0;
      L.ns[3 * s + 2] = 1;
    }
}

void geometry_sphere(Lattice &L) {
  const float Hx = (float)(L.Nx / 2), Hy = (float)(L.Ny / 2), Hz = (float)(L.Nz / 2);
  for (int i = 0; i < L.Nx; i++)
    for (int j = 0; j < L.Ny; j++)
      for (int k = 0; k < L.Nz; k++) {
        const float Rx = i - Hx, Ry = j - Hy, Rz = k - Hz;
        const float radius = (float)sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
        const size_t s = lin(L, i, j, k);
        if (radius < Hz - 1) {
          L.pt[s] = 1;
        } else if (radius < Hz + 1) {
          L.pt[s] = 2;
          L.ns[3 * s + 0] = Rx / radius;
          L.ns[3 * s + 1] = Ry / radius;
          L.ns[3 * s + 2] = Rz / radius;
        } else {
          L.pt[s] = 0;
        }
      }
}

void geometry_error(const char *file, int line, const char *msg) {
  fprintf(stderr, "  [ERROR] %s, line %d: %s\n", file, line, msg);
  exit(10);
}

// Custom geometry file: header naming the columns x, y, z, nx, ny, nz, pt (any order, other columns ignored),
// then one line per listed point; unlisted points are bulk (geometry_custom.cpp:62-163).
int geometry_custom(Lattice &L, const std::string &fname) {
  std::ifstream in(fname);
  if (!in) {
    perror(fname.c_str());
    exit(2);
  }
  printf("  Boundary file: %s\n", fname.c_str());
  const char *f = fname.c_str();
  std::string line;
  if (!std::getline(in, line)) geometry_error(f, 1, "empty file (a header line is expected).");
  std::vector<std::string> header = split_csv(line);
  const char *names[7] = {"x", "y", "z", "nx", "ny", "nz", "pt"};
  int col[7];
  std::string missing;
  for (int c = 0; c < 7; c++) {
    col[c] = -1;
    for (size_t h = 0; h < header.size(); h++)
      if (strcasecmp(header[h].c_str(), names[c]) == 0) col[c] = (int)h;
    if (col[c] < 0) missing += std::string(missing.empty() ? "" : ", ") + names[c];
  }
  if (!missing.empty()) {
    std::string msg = "missing column(s) in the header: " + missing + " (required: x,y,z,nx,ny,nz,pt).";
    geometry_error(f, 1, msg.c_str());
  }
  printf("  Column positions: x=%d y=%d z=%d nx=%d ny=%d nz=%d pt=%d\n", col[0], col[1], col[2], col[3], col[4],
         col[5], col[6]);
  const size_t Nt = L.pt.size();
  std::vector<char> seen(Nt, 0);
  std::fill(L.pt.begin(), L.pt.end(), 1);
  int max_kind = 0, line_number = 1, renormalized = 0;
  size_t listed = 0;
  char msg[300];
  while (std::getline(in, line)) {
    line_number++;
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::vector<std::string> fl = split_csv(line);
    if (fl.size() != header.size()) {
      snprintf(msg, sizeof(msg), "%zu fields, but the header has %zu.", fl.size(), header.size());
      geometry_error(f, line_number, msg);
    }
    double v[7];
    for (int c = 0; c < 7; c++) {
      char *end;
      v[c] = strtod(fl[col[c]].c_str(), &end);
      if (fl[col[c]].empty() || *end != '\0') {
        snprintf(msg, sizeof(msg), "column %s is not a number: '%s'.", names[c], fl[col[c]].c_str());
        geometry_error(f, line_number, msg);
      }
    }
    const int ii = (int)v[0], jj = (int)v[1], kk = (int)v[2], type = (int)v[6];
    if (v[0] != ii || v[1] != jj || v[2] != kk || v[6] != type)
      geometry_error(f, line_number, "x, y, z and pt must be integers.");
    if (ii < 0 || ii >= L.Nx || jj < 0 || jj >= L.Ny || kk < 0 || kk >= L.Nz) {
      snprintf(msg, sizeof(msg), "point (%d, %d, %d) outside the %d x %d x %d grid.", ii, jj, kk, L.Nx, L.Ny, L.Nz);
      geometry_error(f, line_number, msg);
    }
    if (type < 0) geometry_error(f, line_number, "negative point type (use 0 empty, 1 bulk, 2+ surface).");
    const size_t idx = lin(L, ii, jj, kk);
    if (seen[idx]) {
      snprintf(msg, sizeof(msg), "point (%d, %d, %d) is listed twice.", ii, jj, kk);
      geometry_error(f, line_number, msg);
    }
    seen[idx] = 1;
    listed++;
    double nx = v[3], ny = v[4], nz = v[5], norm = sqrt(nx * nx + ny * ny + nz * nz);
    if (type >= 2) {
      if (norm < 1e-6) geometry_error(f, line_number, "surface point with a zero normal vector.");
      if (fabs(norm - 1) > 1e-3) {
        renormalized++;
        nx /= norm;
        ny /= norm;
        nz /= norm;
      }
    }
    L.ns[3 * idx + 0] = (float)nx;
    L.ns[3 * idx + 1] = (float)ny;
    L.ns[3 * idx + 2] = (float)nz;
    L.pt[idx] = type;
    max_kind = std::max(max_kind, type);
  }
  if (listed < Nt)
    fprintf(stderr, "  [WARNING] %zu of the %zu grid points are not listed in %s; they were set to bulk (pt = 1).\n",
            Nt - listed, Nt, f);
  if (renormalized > 0)
    fprintf(stderr, "  [WARNING] %d surface normal(s) in %s did not have unit length and were normalized.\n",
            renormalized, f);
  const int nsurf = max_kind > 1 ? max_kind - 1 : 0;
  printf("  %d boundary surface(s) found.\n\n", nsurf);
  return nsurf;
}

// ---- anchorings (Geometry::Boundary_Init, geometry.cpp:78-107, and the Anchoring constructors) ----

void missing_parameter(int id, const char *name, const char *param) {
  printf("Parameter %s not defined for the boundary condition #%d.\n", param, id);
  printf("The boundary condition %s needs the aforementioned parameter defined.\n", name);
  printf("Please, set it in your input file, or check if it is misspelled.\n");
  printf("Aborting the program.\n");
  exit(1);
}

void lab_axis(float theta_deg, float phi_deg, float out[3]) {
  const float toRad = (float)(M_PI / 180);
  out[0] = sinf(toRad * theta_deg) * cosf(toRad * phi_deg);
  out[1] = sinf(toRad * theta_deg) * sinf(toRad * phi_deg);
  out[2] = cosf(toRad * theta_deg);
}

void init_surfaces(Lattice &L, const Params &p, int nsurf) {
  struct Kind {
    const char *input, *name;
    int type;
    bool axis;
  };
  static const Kind kinds[] = {
      {"rp", "Rapini-Papoular Anchoring", MCL_ANCH_RP, true},
      {"rp_tangent", "Rapini-Papoular Tangent Anchoring", MCL_ANCH_RP_TANGENT, true},
      {"fg", "Fournier-Galatola Anchoring", MCL_ANCH_FG, false},
      {"homeotropic", "Homeotropic Anchoring", MCL_ANCH_HOMEOTROPIC, false},
      {"rp_ghrl", "Rapini-Papoular Anchoring GHRL", MCL_ANCH_RP_GHRL, true},
      {"fg_ghrl", "Fournier-Galatola Anchoring GHRL", MCL_ANCH_FG_GHRL, false},
      {"homeotropic_ghrl", "Homeotropic Anchoring GHRL", MCL_ANCH_HOMEOTROPIC_GHRL, false},
  };
  for (int s = 0; s < nsurf; s++) {
    auto at = p.anchoring_type.find(s);
    if (at == p.anchoring_type.end()) {
      fprintf(stderr,
              "  [ERROR] The geometry has %d surface(s), but anchoring_type %d is not defined in the input file.\n",
              nsurf, s);
      exit(2);
    }
    const Kind *k = nullptr;
    for (const Kind &c : kinds)
      if (iequals(at->second, c.input)) k = &c;
    if (!k) {
      fprintf(stderr,
              "  [ERROR] Anchoring type '%s' not implemented. Options: homeotropic | rp | rp_tangent | fg |"
              " homeotropic_ghrl | rp_ghrl | fg_ghrl\n",
              at->second.c_str());
      exit(2);
    }
    char label[24];
    snprintf(label, sizeof(label), "Surface %d:", s);
    print_field(label, k->name);
    mcl_surface sf{};
    sf.type = k->type;
    auto w = p.W.find(s);
    if (w == p.W.end()) missing_parameter(s, k->name, "W");
    sf.W = w->second;
    print_field("W:", sf.W);
    if (k->axis) {
      auto ph = p.phi_s.find(s);
      if (ph == p.phi_s.end()) missing_parameter(s, k->name, "phi_s");
      print_field("phi_s:", ph->second);
      auto th = p.theta_s.find(s);
      if (th == p.theta_s.end()) missing_parameter(s, k->name, "theta_s");
      print_field("theta_s:", th->second);
      lab_axis(th->second, ph->second, sf.axis);
    }
    printf("\n");
    L.surfaces.push_back(sf);
    L.surface_names.push_back(k->name);
    if (k->type == MCL_ANCH_RP_TANGENT) {
      // sites whose normal is parallel to the axis have no tangent easy axis (anchoring_rp_tangent.cpp:49-63)
      int singular = 0;
      for (size_t q = 0; q < L.pt.size(); q++) {
        if (L.pt[q] != s + 2) continue;
        const float *n = &L.ns[3 * q];
        const float ds = sf.axis[0] * n[0] + sf.axis[1] * n[1] + sf.axis[2] * n[2];
        const float e[3] = {sf.axis[0] - ds * n[0], sf.axis[1] - ds * n[1], sf.axis[2] - ds * n[2]};
        if (sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) < 1e-3f) singular++;
      }
      if (singular > 0)
        printf("rp_tangent surface %d: %d site(s) with normal parallel to the reference axis (no anchoring there).\n",
               s, singular);
    }
  }
}

// ---- initial conditions (ic.cpp) ----

void set_empty(Lattice &L, size_t s) {
  L.ni[3 * s + 0] = 0;
  L.ni[3 * s + 1] = 0;
  L.ni[3 * s + 2] = 1;
}

// Isotropic unit vectors: z uniform in (-1, 1), azimuth uniform (Archimedes), from Philox4x32-10 keyed by the seed
// with the site index as counter (the reference draws them from GSL taus + gsl_ran_dir_3d, ic.cpp:38-65).
void ic_random(Lattice &L, int seed) {
  printf("  Initial condition: random (seed %d)\n\n", seed);
  for (size_t s = 0; s < L.pt.size(); s++) {
    if (!L.pt[s]) {
      set_empty(L, s);
      continue;
    }
    const mcl::Philox4 R = mcl::philox4x32_10((uint32_t)s, (uint32_t)(s >> 32), 0u, 0x49434943u, (uint32_t)seed,
                                              0x1C1C1C1Cu);
    const double z = 2.0 * mcl::u01_open(R.v[0]) - 1.0;
    const double phi = 2.0 * M_PI * mcl::u01_open(R.v[1]);
    const double r = sqrt(std::max(0.0, 1.0 - z * z));
    L.ni[3 * s + 0] = (float)(r * cos(phi));
    L.ni[3 * s + 1] = (float)(r * sin(phi));
    L.ni[3 * s + 2] = (float)z;
  }
}

void ic_homogeneous(Lattice &L, float theta, float phi) {
  printf("  Initial condition: homogeneous\n");
  printf("  %-10s %g\n", "phi0:", phi);
  printf("  %-10s %g\n\n", "theta0:", theta);
  for (size_t s = 0; s < L.pt.size(); s++) {
    if (!L.pt[s]) {
      set_empty(L, s);
      continue;
    }
    L.ni[3 * s + 0] = (float)(sin(theta * M_PI / 180) * cos(phi * M_PI / 180));
    L.ni[3 * s + 1] = (float)(sin(theta * M_PI / 180) * sin(phi * M_PI / 180));
    L.ni[3 * s + 2] = (float)cos(theta * M_PI / 180);
  }
}

void ic_cholesteric(Lattice &L, const Params &p) {
  const float theta = p.theta_0;
  const float phi_0 = (float)(p.phi_0 * M_PI / 180);
  const float p0 = p.p0_i ? p.p0_i : p.p0;
  printf("  Initial condition: cholesteric\n");
  printf("  %-10s %g\n", "p0:", p0);
  printf("  %-10s %g\n", "phi0:", p.phi_0);
  printf("  %-10s %g\n\n", "theta0:", theta);
  for (int i = 0; i < L.Nx; i++)
    for (int j = 0; j < L.Ny; j++)
      for (int k = 0; k < L.Nz; k++) {
        const size_t s = lin(L, i, j, k);
        if (!L.pt[s]) {
          set_empty(L, s);
          continue;
        }
        const float phi = (float)(phi_0 - 2 * M_PI * k / p0);
        L.ni[3 * s + 0] = (float)(sin(theta * M_PI / 180) * cos(phi));
        L.ni[3 * s + 1] = (float)(sin(theta * M_PI / 180) * sin(phi));
        L.ni[3 * s + 2] = (float)cos(theta * M_PI / 180);
      }
}

void ic_error(const char *file, int line, const char *msg) {
  fprintf(stderr, "  [ERROR] %s, line %d: %s\n", file, line, msg);
  exit(3);
}

// Initial-condition file: header with x, y, z, nx, ny, nz (any order, other columns ignored), one line per point;
// every non-empty site must be listed; non-unit directors (|n| off by more than 1e-3) are normalized
// (ic.cpp:103-200).
void ic_file(Lattice &L, const std::string &fname) {
  const char *f = fname.c_str();
  std::ifstream in(fname);
  if (!in) {
    perror(f);
    exit(2);
  }
  printf("  Initial condition: file (%s)\n", f);
  fflush(stdout);
  std::string line;
  if (!std::getline(in, line)) ic_error(f, 1, "empty file (a header line is expected).");
  std::vector<std::string> header = split_csv(line);
  const char *names[6] = {"x", "y", "z", "nx", "ny", "nz"};
  int col[6];
  std::string missing;
  for (int c = 0; c < 6; c++) {
    col[c] = -1;
    for (size_t h = 0; h < header.size(); h++)
      if (strcasecmp(header[h].c_str(), names[c]) == 0) col[c] = (int)h;
    if (col[c] < 0) missing += std::string(missing.empty() ? "" : ", ") + names[c];
  }
  if (!missing.empty()) {
    std::string msg = "missing column(s) in the header: " + missing + " (required: x,y,z,nx,ny,nz).";
    ic_error(f, 1, msg.c_str());
  }
  const size_t Nt = L.pt.size();
  std::vector<char> seen(Nt, 0);
  int line_number = 1, normalized = 0;
  char msg[300];
  while (std::getline(in, line)) {
    line_number++;
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::vector<std::string> fl = split_csv(line);
    if (fl.size() != header.size()) {
      snprintf(msg, sizeof(msg), "%zu fields, but the header has %zu.", fl.size(), header.size());
      ic_error(f, line_number, msg);
    }
    double v[6];
    for (int c = 0; c < 6; c++) {
      char *end;
      v[c] = strtod(fl[col[c]].c_str(), &end);
      if (fl[col[c]].empty() || *end != '\0') {
        snprintf(msg, sizeof(msg), "column %s is not a number: '%s'.", names[c], fl[col[c]].c_str());
        ic_error(f, line_number, msg);
      }
    }
    const int i = (int)v[0], j = (int)v[1], k = (int)v[2];
    if (v[0] != i || v[1] != j || v[2] != k) ic_error(f, line_number, "x, y and z must be integers.");
    if (i < 0 || i >= L.Nx || j < 0 || j >= L.Ny || k < 0 || k >= L.Nz) {
      snprintf(msg, sizeof(msg), "point (%d, %d, %d) outside the %d x %d x %d grid.", i, j, k, L.Nx, L.Ny, L.Nz);
      ic_error(f, line_number, msg);
    }
    const size_t idx = lin(L, i, j, k);
    if (seen[idx]) {
      snprintf(msg, sizeof(msg), "point (%d, %d, %d) is listed twice.", i, j, k);
      ic_error(f, line_number, msg);
    }
    seen[idx] = 1;
    double nx = v[3], ny = v[4], nz = v[5], norm = sqrt(nx * nx + ny * ny + nz * nz);
    if (L.pt[idx]) {
      if (norm < 1e-6) ic_error(f, line_number, "zero director on a non-empty site.");
      if (fabs(norm - 1) > 1e-3) {
        normalized++;
        nx /= norm;
        ny /= norm;
        nz /= norm;
      }
    } else {
      nx = 0;
      ny = 0;
      nz = 1;
    }
    L.ni[3 * idx + 0] = (float)nx;
    L.ni[3 * idx + 1] = (float)ny;
    L.ni[3 * idx + 2] = (float)nz;
  }
  int unlisted = 0;
  for (size_t idx = 0; idx < Nt; idx++) {
    if (seen[idx]) continue;
    if (L.pt[idx]) unlisted++;
    set_empty(L, idx);
  }
  if (unlisted > 0) {
    fprintf(stderr, "  [ERROR] %s: %d non-empty grid point(s) are not listed (the grid is %d x %d x %d).\n", f,
            unlisted, L.Nx, L.Ny, L.Nz);
    exit(3);
  }
  if (normalized > 0)
    fprintf(stderr, "  [WARNING] %d director(s) in %s did not have unit length and were normalized.\n", normalized, f);
}

}  // namespace

Lattice setup_simulation(Params &p) {
  // evolution mode (simulator.cpp:162-188): this program runs the GPU modes
  const std::string mode = mode_base(p.evol);
  const bool gpu = p.evol.size() > 3 && strcasecmp(p.evol.c_str() + p.evol.size() - 3, "GPU") == 0;
  const bool known = iequals(mode, "thermal") || iequals(mode, "step") || iequals(mode, "quench") ||
                     iequals(mode, "electric");
  if (!known) {
    fprintf(stderr,
            "  [ERROR] Evol mode '%s' not implemented. Options: thermal | step | quench | electric, or the same"
            " with the GPU suffix (thermalGPU, ...)\n",
            p.evol.c_str());
    exit(2);
  }
  if (!gpu) {
    fprintf(stderr,
            "  [ERROR] Evol mode '%s' runs on the CPU, which this MI355X build does not provide; use %sGPU.\n",
            p.evol.c_str(), mode.c_str());
    exit(2);
  }
  Lattice L;
  L.Nx = p.Nx;
  L.Ny = p.Ny;
  L.Nz = p.Nz;
  const size_t Nt = (size_t)p.Nx * p.Ny * p.Nz;
  L.pt.assign(Nt, 1);
  L.ns.assign(3 * Nt, 0.0f);
  L.ni.assign(3 * Nt, 0.0f);
  int nsurf = 0;
  if (iequals(p.geometry, "bulk")) {
    print_field("Geometry:", "Bulk");
  } else if (iequals(p.geometry, "slab")) {
    print_field("Geometry:", "Slab");
    nsurf = 2;
    geometry_slab(L);
    p.zbound = "free";
  } else if (iequals(p.geometry, "sphere")) {
    print_field("Geometry:", "Sphere");
    nsurf = 1;
    p.xbound = p.ybound = p.zbound = "free";
    geometry_sphere(L);
  } else if (iequals(p.geometry, "custom")) {
    print_field("Geometry:", "Custom");
    nsurf = geometry_custom(L, p.bound_file);
  } else {
    fprintf(stderr, "  [ERROR] Geometry '%s' not implemented. Options: bulk | slab | sphere | custom\n",
            p.geometry.c_str());
    exit(2);
  }
  L.periodic[0] = boundary_kind(p.xbound, "X");
  L.periodic[1] = boundary_kind(p.ybound, "Y");
  L.periodic[2] = boundary_kind(p.zbound, "Z");
  if (!iequals(p.geometry, "sphere")) {
    print_field("X boundary:", p.xbound.c_str());
    print_field("Y boundary:", p.ybound.c_str());
    if (!iequals(p.geometry, "slab")) print_field("Z boundary:", p.zbound.c_str());
  }
  printf("\n");
  check_checkerboard(p);
  init_surfaces(L, p, nsurf);
  if (iequals(p.ic, "random"))
    ic_random(L, p.seed);
  else if (iequals(p.ic, "homogeneous"))
    ic_homogeneous(L, p.theta_0, p.phi_0);
  else if (iequals(p.ic, "ic_file"))
    ic_file(L, p.ic_file);
  else if (iequals(p.ic, "cholesteric"))
    ic_cholesteric(L, p);
  else {
    fprintf(stderr,
            "  [ERROR] Initial condition '%s' not implemented. Options: random | homogeneous | ic_file | cholesteric\n",
            p.ic.c_str());
    exit(2);
  }
  if (iequals(p.potential, "ll") || iequals(p.potential, "lebwohl-lasher")) {
    print_field("Potential:", "Lebwohl-Lasher");
  } else if (iequals(p.potential, "ghrl") || iequals(p.potential, "gruhn-hess")) {
    set_ghrl(p);
    print_field("Potential:", "Gruhn-Hess (GHRL)");
  } else if (iequals(p.potential, "pear")) {
    print_field("Potential:", "Pear (Dhakal-Selinger)");
  } else {
    fprintf(stderr, "  [ERROR] Potential '%s' not implemented. Options: ll | ghrl | pear\n", p.potential.c_str());
    exit(2);
  }
  printf("\n");
  for (int v : L.pt) L.nvalid += v ? 1 : 0;
  return L;
}
