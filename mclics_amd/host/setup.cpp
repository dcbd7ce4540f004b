// Geometry, anchoring and initial-condition setup (behaviour of the reference's src/geometry_*.cpp,
// src/anchoring*.cpp, src/ic.cpp and src/simulator.cpp; file formats and error statuses kept).
#include "setup.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "coltable.h"
#include "philox.h"

namespace {

inline size_t lin(const Lattice &L, int i, int j, int k) {
  return (size_t)i + (size_t)L.Nx * ((size_t)j + (size_t)L.Ny * k);
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

// A record's grid coordinate (or point type) is usable when it is a whole number that fits an int.
bool whole(double v) { return v == std::trunc(v) && std::fabs(v) < 2147483648.0; }

// Sites named by the records of a table: each must lie in the grid and be named at most once.
class SiteRoll {
 public:
  explicit SiteRoll(const Lattice &L) : L_(L), named_(L.pt.size(), 0) {}

  // Site of the whole coordinates r[0..2], or a table error when it is outside the grid.
  size_t locate(const ColumnTable &t, const double *r) const {
    const int i = (int)r[0], j = (int)r[1], k = (int)r[2];
    if (i < 0 || i >= L_.Nx || j < 0 || j >= L_.Ny || k < 0 || k >= L_.Nz)
      t.fail("point (%d, %d, %d) outside the %d x %d x %d grid.", i, j, k, L_.Nx, L_.Ny, L_.Nz);
    return lin(L_, i, j, k);
  }

  // Marks site s as named, or a table error when a record named it before.
  void take(const ColumnTable &t, size_t s) {
    if (named_[s]) {
      const long nx = L_.Nx, ny = L_.Ny;
      t.fail("point (%ld, %ld, %ld) is listed twice.", (long)s % nx, (long)(s / nx) % ny, (long)(s / nx / ny));
    }
    named_[s] = 1;
    count_++;
  }

  bool named(size_t s) const { return named_[s] != 0; }
  size_t count() const { return count_; }

 private:
  const Lattice &L_;
  std::vector<char> named_;
  size_t count_ = 0;
};

// Rescales (x, y, z) to unit length when it is clearly off (|length - 1| > 1e-3), so that files written with rounded
// digits are taken as they are. Returns whether it rescaled.
bool rescale_if_off(double &x, double &y, double &z) {
  const double len = std::sqrt(x * x + y * y + z * z);
  if (std::fabs(len - 1.0) <= 1e-3) return false;
  x /= len;
  y /= len;
  z /= len;
  return true;
}

// Custom geometry (behaviour of geometry_custom.cpp:62-163): a table with the columns x, y, z, nx, ny, nz, pt. Each
// record sets one site's point type (0 empty, 1 bulk, n >= 2 surface n - 2) and, on surface sites, its outward
// normal; sites without a record are bulk. Format errors end the program with status 10. Returns the number of
// surfaces, the largest point type less one.
int geometry_custom(Lattice &L, const std::string &path) {
  ColumnTable table(path, 10);
  printf("  Boundary file: %s\n", path.c_str());
  table.bind({"x", "y", "z", "nx", "ny", "nz", "pt"}, "x,y,z,nx,ny,nz,pt");
  printf("  Column positions: x=%d y=%d z=%d nx=%d ny=%d nz=%d pt=%d\n", table.position(0), table.position(1),
         table.position(2), table.position(3), table.position(4), table.position(5), table.position(6));
  std::fill(L.pt.begin(), L.pt.end(), 1);
  SiteRoll roll(L);
  int top_type = 0, rescaled = 0;
  double r[7];
  while (table.next(r)) {
    if (!whole(r[0]) || !whole(r[1]) || !whole(r[2]) || !whole(r[6]))
      table.fail("x, y, z and pt must be integers.");
    const size_t s = roll.locate(table, r);
    const int type = (int)r[6];
    if (type < 0) table.fail("negative point type (use 0 empty, 1 bulk, 2+ surface).");
    roll.take(table, s);
    double nx = r[3], ny = r[4], nz = r[5];
    if (type >= 2) {
      if (std::sqrt(nx * nx + ny * ny + nz * nz) < 1e-6) table.fail("surface point with a zero normal vector.");
      if (rescale_if_off(nx, ny, nz)) rescaled++;
    }
    L.pt[s] = type;
    L.ns[3 * s + 0] = (float)nx;
    L.ns[3 * s + 1] = (float)ny;
    L.ns[3 * s + 2] = (float)nz;
    top_type = std::max(top_type, type);
  }
  const size_t Nt = L.pt.size();
  if (roll.count() < Nt)
    fprintf(stderr, "  [WARNING] %zu of the %zu grid points are not listed in %s; they were set to bulk (pt = 1).\n",
            Nt - roll.count(), Nt, path.c_str());
  if (rescaled > 0)
    fprintf(stderr, "  [WARNING] %d surface normal(s) in %s did not have unit length and were normalized.\n",
            rescaled, path.c_str());
  const int nsurf = top_type > 1 ? top_type - 1 : 0;
  printf("  %d boundary surface(s) found.\n\n", nsurf);
  fflush(stdout);
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

// Cholesteric helix along z (behaviour of ic.cpp:202-232): polar angle theta_0, azimuth phi_0 - 2 pi k / p0 on plane
// k (p0_i replaces p0 when it is set). Every site of a plane gets the plane's director; empty sites get (0, 0, 1).
// The angle steps keep the reference's precisions: the azimuth is rounded to float, the components are computed in
// double from it and stored as float.
void ic_cholesteric(Lattice &L, const Params &p) {
  const float pitch = p.p0_i ? p.p0_i : p.p0;
  printf("  Initial condition: cholesteric\n");
  printf("  %-10s %g\n", "p0:", pitch);
  printf("  %-10s %g\n", "phi0:", p.phi_0);
  printf("  %-10s %g\n\n", "theta0:", p.theta_0);
  const double polar = p.theta_0 * M_PI / 180;
  const float azimuth_0 = (float)(p.phi_0 * M_PI / 180);
  const size_t plane = (size_t)L.Nx * L.Ny;
  for (int k = 0; k < L.Nz; k++) {
    const float azimuth = (float)(azimuth_0 - 2 * M_PI * k / pitch);
    const float d[3] = {(float)(sin(polar) * cos(azimuth)), (float)(sin(polar) * sin(azimuth)), (float)cos(polar)};
    for (size_t s = plane * k; s < plane * (k + 1); s++) {
      if (L.pt[s])
        std::copy(d, d + 3, &L.ni[3 * s]);
      else
        set_empty(L, s);
    }
  }
}

// Initial condition from a file (behaviour of ic.cpp:103-200): a table with the columns x, y, z, nx, ny, nz (a
// director_field snapshot qualifies). Every non-empty site must have a record; directors off unit length by more
// than 1e-3 are rescaled; empty sites get (0, 0, 1) whatever their record says. Format errors end the program with
// status 3.
void ic_file(Lattice &L, const std::string &path) {
  ColumnTable table(path, 3);
  printf("  Initial condition: file (%s)\n", path.c_str());
  fflush(stdout);
  table.bind({"x", "y", "z", "nx", "ny", "nz"}, "x,y,z,nx,ny,nz");
  const size_t Nt = L.pt.size();
  for (size_t s = 0; s < Nt; s++) set_empty(L, s);
  SiteRoll roll(L);
  int rescaled = 0;
  double r[6];
  while (table.next(r)) {
    if (!whole(r[0]) || !whole(r[1]) || !whole(r[2])) table.fail("x, y and z must be integers.");
    const size_t s = roll.locate(table, r);
    roll.take(table, s);
    if (!L.pt[s]) continue;
    double nx = r[3], ny = r[4], nz = r[5];
    if (std::sqrt(nx * nx + ny * ny + nz * nz) < 1e-6) table.fail("zero director on a non-empty site.");
    if (rescale_if_off(nx, ny, nz)) rescaled++;
    L.ni[3 * s + 0] = (float)nx;
    L.ni[3 * s + 1] = (float)ny;
    L.ni[3 * s + 2] = (float)nz;
  }
  long unlisted = 0;
  for (size_t s = 0; s < Nt; s++) unlisted += (L.pt[s] && !roll.named(s)) ? 1 : 0;
  if (unlisted > 0) {
    fprintf(stderr, "  [ERROR] %s: %ld non-empty grid point(s) are not listed (the grid is %d x %d x %d).\n",
            path.c_str(), unlisted, L.Nx, L.Ny, L.Nz);
    exit(3);
  }
  if (rescaled > 0)
    fprintf(stderr, "  [WARNING] %d director(s) in %s did not have unit length and were normalized.\n", rescaled,
            path.c_str());
}

}  // namespace

Lattice setup_simulation(Params &p) {
  // evolution mode (simulator.cpp:36-53): thermal, step, quench or electric, on the CPU or with the GPU suffix
  const std::string mode = mode_base(p.evol);
  const bool known = iequals(mode, "thermal") || iequals(mode, "step") || iequals(mode, "quench") ||
                     iequals(mode, "electric");
  if (!known) {
    fprintf(stderr,
            "  [ERROR] Evol mode '%s' not implemented. Options: thermal | step | quench | electric, or the same"
            " with the GPU suffix (thermalGPU, ...)\n",
            p.evol.c_str());
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
