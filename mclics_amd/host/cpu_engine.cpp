// Host CPU engine of the evol modes without the GPU suffix (thermal, step, quench, electric): the reference runs them
// with its OpenMP sweep (src/evolve.cpp:85-221) and host measurements (src/evolve.cpp:33-75,
// src/parameter_order.cpp). This engine is the program's own:
//
//  - Sweep: two colours of the checkerboard in turn ((i + j + k) even, then odd); sites of one colour share no bond,
//    so the planes of a colour are split over the OpenMP threads. Every trial draws its four uniforms from one
//    Philox4x32-10 block keyed by the seed, with the site and the sweep number as counter, so a trajectory does not
//    depend on the number of threads (the reference's per-thread generators make its CPU trajectories change with
//    OMP_NUM_THREADS).
//  - Trial (the reference's Markov chain): rotation by pi va, va uniform in (-ang_var, ang_var), about x, y or z
//    (thresholds 0.333 and 0.666), renormalization, Metropolis acceptance from the FP32 site energies (T = 0 accepts
//    when the energy does not increase), then the n -> -n trial when flip moves are on; the ang_var rule after each
//    sweep.
//  - Site energy: the six bonds summed in FP64 and rounded to FP32, plus the anchoring and field terms
//    (geometry.cpp:23-76, potential.cpp:19-63, anchoring_*.cpp).
//  - Measurements in FP64, reduced per plane and summed in plane order (reproducible for any thread count): mean
//    site energy, Q tensor with its largest eigenvalue and eigenvector (cyclic Jacobi rotations), radial order,
//    polarization and the two splay measures of the pear model, local order for the snapshots.
#include <math.h>
#include <omp.h>
#include <stdio.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"
#include "philox.h"

namespace {

struct V3 {
  float x, y, z;
};
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

enum { kLL, kGHRL, kPear };

struct Couplings {
  float A, B1, C;
  float lam, mu, nu, rho, sigma;
  V3 e;  // unit field direction
  float elecA, elecP;
  bool field;
};

// GHRL bond of director a at a site with director b at its neighbour along the unit vector r (potential.cpp:40-56);
// also the phantom bond of the *_ghrl anchorings.
inline float ghrl_bond(const Couplings &c, V3 a, V3 b, V3 r) {
  const float ai = dot(a, r), aj = dot(b, r), ab = dot(a, b);
  const float p2 = 1.5f * ab * ab - 0.5f;
  const float twist = r.x * (b.y * a.z - b.z * a.y) + r.y * (b.z * a.x - b.x * a.z) + r.z * (b.x * a.y - b.y * a.x);
  const float e1 = 1.5f * ai * ai + 1.5f * aj * aj - 1.0f;
  return e1 * (c.rho * p2 + c.lam) + c.mu * (ai * aj * ab - 1.0f / 9.0f) + c.nu * p2 +
         c.sigma * (ab > 0 ? 1.0f : -1.0f) * twist;
}

// Bulk bond energies (potential.cpp:19-56).
template <int POT>
inline float bond(const Couplings &c, V3 a, V3 b, V3 r) {
  if constexpr (POT == kLL) {
    const float ab = dot(a, b);
    return -1.5f * c.A * ab * ab;
  } else if constexpr (POT == kGHRL) {
    return ghrl_bond(c, a, b, r);
  } else {
    const float ab = dot(a, b);
    const float along = r.x * (b.x - a.x) + r.y * (b.y - a.y) + r.z * (b.z - a.z);
    return -c.A * ab * ab - c.B1 * ab - 0.25f * c.C * (1.0f + ab) * (1.0f + ab) * along;
  }
}

constexpr V3 kAxis[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// Largest eigenvalue of the symmetric 3x3 matrix (xx, xy, xz, yy, yz, zz) and its unit eigenvector, by cyclic Jacobi
// rotations (each zeroes one off-diagonal element; a few sweeps reach FP64 rounding).
double largest_eigen(const double q[6], double vec[3]) {
  double a[3][3] = {{q[0], q[1], q[2]}, {q[1], q[3], q[4]}, {q[2], q[4], q[5]}};
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 32; sweep++) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (const auto &pq : pairs) {
      const int p = pq[0], r = pq[1];
      if (a[p][r] == 0.0) continue;
      const double theta = (a[r][r] - a[p][p]) / (2.0 * a[p][r]);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
      const double apr = a[p][r];
      a[p][p] -= t * apr;
      a[r][r] += t * apr;
      a[p][r] = a[r][p] = 0.0;
      for (int o = 0; o < 3; o++) {
        if (o != p && o != r) {
          const double aop = a[o][p], aor = a[o][r];
          a[o][p] = a[p][o] = cs * aop - sn * aor;
          a[o][r] = a[r][o] = sn * aop + cs * aor;
        }
        const double vop = v[o][p], vor = v[o][r];
        v[o][p] = cs * vop - sn * vor;
        v[o][r] = sn * vop + cs * vor;
      }
    }
  }
  int m = 0;
  for (int d = 1; d < 3; d++)
    if (a[d][d] > a[m][m]) m = d;
  for (int d = 0; d < 3; d++) vec[d] = v[d][m];
  return a[m][m];
}

class CpuEngine final : public Engine {
 public:
  CpuEngine(const Params &p, const Lattice &L);
  const char *device() const override { return "CPU"; }
  void set_temperature(float T) override { T_ = T; }
  void set_field(float elecE) override { elecE_ = elecE; }
  void sweeps(int n, int measure_every) override;
  void reset_moments() override { mom_ = mcl_moments{}; }
  mcl_moments moments() override { return mom_; }
  void snapshot(float *ni, float *local_S) override;

 private:
  // Neighbour coordinate of i along one axis, or -1 past a free boundary (periodic: wrap).
  static std::vector<int> neighbours(int n, bool periodic, int step) {
    std::vector<int> out(n);
    for (int i = 0; i < n; i++) {
      const int j = i + step;
      out[i] = (j >= 0 && j < n) ? j : (periodic ? (j + n) % n : -1);
    }
    return out;
  }
  size_t site(int i, int j, int k) const { return (size_t)i + (size_t)nx_ * ((size_t)j + (size_t)ny_ * k); }
  // Site indices of the six neighbours of (i, j, k) (+x, -x, +y, -y, +z, -z), -1 where there is none or it is empty.
  void neighbours_of(int i, int j, int k, long long nb[6]) const {
    const int c[3] = {i, j, k};
    for (int d = 0; d < 6; d++) {
      int o[3] = {c[0], c[1], c[2]};
      const int axis = d >> 1;
      o[axis] = step_[d][c[axis]];
      nb[d] = -1;
      if (o[axis] < 0) continue;
      const size_t q = site(o[0], o[1], o[2]);
      if (pt_[q]) nb[d] = (long long)q;
    }
  }
  V3 dir(size_t s) const { return V3{n_[3 * s], n_[3 * s + 1], n_[3 * s + 2]}; }
  float anchoring(size_t s, V3 m) const;
  template <int POT>
  float site_energy(size_t s, const long long nb[6], V3 m) const;
  template <int POT>
  long long colour_pass(int colour, float av);
  template <int POT>
  void measure();
  template <int POT>
  double energy_sum() const;

  int nx_, ny_, nz_;
  bool periodic_[3];
  int pot_;
  bool flip_;
  std::vector<int> step_[6];
  std::vector<int> pt_;
  std::vector<float> n_, normal_, tangent_;
  std::vector<mcl_surface> surf_;
  Couplings c_{};
  float T_, elecE_, ang_var_ = 0.5f;  // evolve_thermal.cpp:40
  uint64_t seed_;
  unsigned long long sweep_ = 0;
  long long nvalid_ = 0;
  bool pear_, sph_, cyl_;
  mcl_moments mom_{};
};

CpuEngine::CpuEngine(const Params &p, const Lattice &L)
    : nx_(L.Nx), ny_(L.Ny), nz_(L.Nz), pt_(L.pt), n_(L.ni), normal_(L.ns), surf_(L.surfaces) {
  for (int a = 0; a < 3; a++) periodic_[a] = L.periodic[a] != 0;
  const int len[3] = {nx_, ny_, nz_};
  for (int d = 0; d < 6; d++) step_[d] = neighbours(len[d >> 1], periodic_[d >> 1], (d & 1) ? -1 : 1);
  pot_ = iequals(p.potential, "pear") ? kPear : (iequals(p.potential, "ghrl") || iequals(p.potential, "gruhn-hess"))
                                                    ? kGHRL
                                                    : kLL;
  flip_ = p.flipMove != 0;
  c_.A = p.A;
  c_.B1 = p.B1;
  c_.C = p.C;
  c_.lam = p.ghrl_lambda;
  c_.mu = p.ghrl_mu;
  c_.nu = p.ghrl_nu;
  c_.rho = p.ghrl_rho;
  c_.sigma = p.ghrl_sigma;
  c_.e = V3{p.elecX, p.elecY, p.elecZ};
  c_.elecA = p.elecA;
  c_.elecP = p.elecP;
  c_.field = p.elecA != 0 || p.elecP != 0;
  T_ = p.Ti;
  elecE_ = p.elecE;
  seed_ = (uint64_t)(uint32_t)p.seed;
  pear_ = iequals(p.potential, "pear");
  sph_ = iequals(p.radial, "sphere") || iequals(p.radial, "both");
  cyl_ = iequals(p.radial, "cylinder") || iequals(p.radial, "both");
  for (int v : pt_) nvalid_ += v ? 1 : 0;
  // rp_tangent: the easy axis of each site is the lab axis projected on the site's tangent plane; none (zero vector,
  // no anchoring) where the axis is within 1e-3 of the normal (anchoring_rp_tangent.cpp:49-63)
  tangent_.assign(normal_.size(), 0.0f);
  for (size_t s = 0; s < pt_.size(); s++) {
    if (pt_[s] < 2 || pt_[s] - 2 >= (int)surf_.size() || surf_[pt_[s] - 2].type != MCL_ANCH_RP_TANGENT) continue;
    const float *d = surf_[pt_[s] - 2].axis;
    const V3 sn{normal_[3 * s], normal_[3 * s + 1], normal_[3 * s + 2]};
    const float ds = d[0] * sn.x + d[1] * sn.y + d[2] * sn.z;
    const V3 t{d[0] - ds * sn.x, d[1] - ds * sn.y, d[2] - ds * sn.z};
    const float len_t = std::sqrt(dot(t, t));
    if (len_t < 1e-3f) continue;
    tangent_[3 * s] = t.x / len_t;
    tangent_[3 * s + 1] = t.y / len_t;
    tangent_[3 * s + 2] = t.z / len_t;
  }
}

// Anchoring energy of director m at surface site s (W included; anchoring_*.cpp).
float CpuEngine::anchoring(size_t s, V3 m) const {
  const mcl_surface &sf = surf_[pt_[s] - 2];
  const V3 nrm{normal_[3 * s], normal_[3 * s + 1], normal_[3 * s + 2]};
  const V3 lab{sf.axis[0], sf.axis[1], sf.axis[2]};
  switch (sf.type) {
    case MCL_ANCH_HOMEOTROPIC: {
      const float c = dot(m, nrm);
      return -sf.W * c * c;
    }
    case MCL_ANCH_RP: {
      const float c = dot(m, lab);
      return -sf.W * c * c;
    }
    case MCL_ANCH_RP_TANGENT: {
      const float c = dot(m, V3{tangent_[3 * s], tangent_[3 * s + 1], tangent_[3 * s + 2]});
      return -sf.W * c * c;
    }
    case MCL_ANCH_FG: {
      const float c = dot(m, nrm);
      return sf.W * c * c;
    }
    case MCL_ANCH_HOMEOTROPIC_GHRL:
      return sf.W * ghrl_bond(c_, m, nrm, nrm);
    case MCL_ANCH_RP_GHRL:
      return sf.W * ghrl_bond(c_, m, lab, nrm);
    case MCL_ANCH_FG_GHRL: {
      const float along = dot(m, nrm);
      V3 t{m.x - along * nrm.x, m.y - along * nrm.y, m.z - along * nrm.z};
      const float len_t = std::sqrt(dot(t, t));
      if (len_t < 1e-6f) return 0.0f;
      t = V3{t.x / len_t, t.y / len_t, t.z / len_t};
      return sf.W * ghrl_bond(c_, m, t, nrm);
    }
  }
  return 0.0f;
}

template <int POT>
float CpuEngine::site_energy(size_t s, const long long nb[6], V3 m) const {
  double bonds = 0.0;
  for (int d = 0; d < 6; d++)
    if (nb[d] >= 0) bonds += bond<POT>(c_, m, dir((size_t)nb[d]), kAxis[d]);
  float E = (float)bonds;
  if (pt_[s] > 1) E += anchoring(s, m);
  if (c_.field) {
    const float f = elecE_ * dot(m, c_.e);
    E += -c_.elecP * f - c_.elecA * f * f;
  }
  return E;
}

// One colour of a sweep: a Metropolis trial on every non-empty site of that colour. Returns the accepted rotations
// (the flip trial is not counted, as it must not steer ang_var).
template <int POT>
long long CpuEngine::colour_pass(int colour, float av) {
  long long accepted = 0;
  const float T = T_;
  const uint32_t sw_lo = (uint32_t)sweep_, sw_hi = (uint32_t)(sweep_ >> 32) ^ 0x43505553u;
  const uint32_t key0 = (uint32_t)seed_, key1 = (uint32_t)(seed_ >> 32) ^ 0x3C6EF372u;
#pragma omp parallel for schedule(static) reduction(+ : accepted)
  for (int k = 0; k < nz_; k++) {
    for (int j = 0; j < ny_; j++) {
      for (int i = (colour + j + k) & 1; i < nx_; i += 2) {
        const size_t s = site(i, j, k);
        if (!pt_[s]) continue;
        const mcl::Philox4 R = mcl::philox4x32_10((uint32_t)s, (uint32_t)((uint64_t)s >> 32), sw_lo, sw_hi, key0, key1);
        const float va = (2.0f * mcl::u01_open(R.v[0]) - 1.0f) * av;
        const float axis = mcl::u01_open(R.v[1]);
        float sn, cs;
        sincosf((float)M_PI * va, &sn, &cs);
        const V3 n = dir(s);
        V3 t;
        if (axis < 0.333f)
          t = V3{n.x, n.y * cs + n.z * sn, n.z * cs - n.y * sn};
        else if (axis < 0.666f)
          t = V3{n.x * cs - n.z * sn, n.y, n.z * cs + n.x * sn};
        else
          t = V3{n.x * cs + n.y * sn, n.y * cs - n.x * sn, n.z};
        const float len = std::sqrt(dot(t, t));
        t = V3{t.x / len, t.y / len, t.z / len};
        long long nb[6];
        neighbours_of(i, j, k, nb);
        const float E_old = site_energy<POT>(s, nb, n), E_new = site_energy<POT>(s, nb, t);
        const float u = mcl::u01_open(R.v[2]);
        const bool take = T > 0 ? u < expf(-(E_new - E_old) / T) : E_new <= E_old;
        V3 cur = take ? t : n;
        float E_cur = take ? E_new : E_old;
        accepted += take ? 1 : 0;
        if (flip_) {
          const V3 f{-cur.x, -cur.y, -cur.z};
          const float E_flip = site_energy<POT>(s, nb, f);
          const float uf = mcl::u01_open(R.v[3]);
          if (T > 0 ? uf < expf(-(E_flip - E_cur) / T) : E_flip <= E_cur) cur = f;
        }
        n_[3 * s] = cur.x;
        n_[3 * s + 1] = cur.y;
        n_[3 * s + 2] = cur.z;
      }
    }
  }
  return accepted;
}

template <int POT>
double CpuEngine::energy_sum() const {
  std::vector<double> plane(nz_, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz_; k++) {
    double e = 0.0;
    for (int j = 0; j < ny_; j++)
      for (int i = 0; i < nx_; i++) {
        const size_t s = site(i, j, k);
        if (!pt_[s]) continue;
        long long nb[6];
        neighbours_of(i, j, k, nb);
        e += site_energy<POT>(s, nb, dir(s));
      }
    plane[k] = e;
  }
  double sum = 0.0;
  for (double e : plane) sum += e;
  return sum;
}

// One measurement added to the moments (evolve_thermal.cpp:80-94 and ExtraObservables, parameter_order.cpp:249-295).
template <int POT>
void CpuEngine::measure() {
  enum { QXX, QXY, QXZ, QYY, QYZ, QZZ, SRS, PRS, NRS, SRC, PRC, NRC, DIV, NQ };
  std::vector<double> part((size_t)nz_ * NQ, 0.0);
  const double cx = nx_ / 2, cy = ny_ / 2, cz = nz_ / 2;  // the sphere's centre (integer halves)
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz_; k++) {
    double *q = &part[(size_t)k * NQ];
    for (int j = 0; j < ny_; j++)
      for (int i = 0; i < nx_; i++) {
        const size_t s = site(i, j, k);
        if (!pt_[s]) continue;
        const V3 n = dir(s);
        const double x = n.x, y = n.y, z = n.z;
        q[QXX] += 3 * x * x;
        q[QXY] += 3 * x * y;
        q[QXZ] += 3 * x * z;
        q[QYY] += 3 * y * y;
        q[QYZ] += 3 * y * z;
        q[QZZ] += 3 * z * z;
        for (int cyl = 0; cyl < 2; cyl++) {
          if (!(cyl ? cyl_ : sph_)) continue;
          const double rx = i - cx, ry = j - cy, rz = cyl ? 0.0 : k - cz;
          const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
          if (r < 0.5) continue;
          const double c = (x * rx + y * ry + z * rz) / r;
          q[cyl ? SRC : SRS] += 1.5 * c * c - 0.5;
          q[cyl ? PRC : PRS] += c;
          q[cyl ? NRC : NRS] += 1;
        }
        if (pear_) {  // <div n>: per axis the mean of the one-sided differences to the existing neighbours
          const float own[3] = {n.x, n.y, n.z};
          long long nb[6];
          neighbours_of(i, j, k, nb);
          double div = 0;
          for (int axis = 0; axis < 3; axis++) {
            double d = 0;
            int cnt = 0;
            for (int side = 0; side < 2; side++) {
              const long long o = nb[2 * axis + side];
              if (o < 0) continue;
              const float *m = &n_[3 * (size_t)o];
              d += (side ? -1.0 : 1.0) * ((double)m[axis] - own[axis]);
              cnt++;
            }
            if (cnt) div += d / cnt;
          }
          q[DIV] += div;
        }
      }
  }
  double tot[NQ] = {};
  for (int k = 0; k < nz_; k++)
    for (int c = 0; c < NQ; c++) tot[c] += part[(size_t)k * NQ + c];
  const double nv = (double)nvalid_;
  const double Q[6] = {(tot[QXX] - nv) / (2 * nv), tot[QXY] / (2 * nv), tot[QXZ] / (2 * nv),
                       (tot[QYY] - nv) / (2 * nv), tot[QYZ] / (2 * nv), (tot[QZZ] - nv) / (2 * nv)};
  double director[3];
  const double S = largest_eigen(Q, director);
  const double E = energy_sum<POT>() / nv;
  mom_.n++;
  mom_.S += S;
  mom_.S2 += S * S;
  mom_.E += E;
  mom_.E2 += E * E;
  if (sph_) {
    mom_.Sr_sph += tot[NRS] > 0 ? tot[SRS] / tot[NRS] : 0.0;
    mom_.Pr_sph += tot[NRS] > 0 ? tot[PRS] / tot[NRS] : 0.0;
  }
  if (cyl_) {
    mom_.Sr_cyl += tot[NRC] > 0 ? tot[SRC] / tot[NRC] : 0.0;
    mom_.Pr_cyl += tot[NRC] > 0 ? tot[PRC] / tot[NRC] : 0.0;
  }
  if (pear_) {
    // polarization |<d.n>| with d the director (parameter_order.cpp:131-141), and the splay magnitude: the bond
    // splay vectors (b (b.r) - a (a.r) + (a.b)(a (b.r) - b (a.r))) / 2 summed over both directions of every bond
    // (each direction gives the same vector), 1.5 |sum| / number of directed bonds (parameter_order.cpp:147-176)
    std::vector<double> sp((size_t)nz_ * 5, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz_; k++) {
      double *w = &sp[(size_t)k * 5];
      for (int j = 0; j < ny_; j++)
        for (int i = 0; i < nx_; i++) {
          const size_t s = site(i, j, k);
          if (!pt_[s]) continue;
          const V3 a = dir(s);
          w[0] += director[0] * a.x + director[1] * a.y + director[2] * a.z;
          long long nb[6];
          neighbours_of(i, j, k, nb);
          for (int d = 0; d < 6; d += 2) {  // +x, +y, +z: each bond once, counted for both directions
            const long long o = nb[d];
            if (o < 0) continue;
            const V3 b = dir((size_t)o), r = kAxis[d];
            const float ra = dot(a, r), rb = dot(b, r), ab = dot(a, b);
            w[1] += (double)(b.x * rb - a.x * ra + a.x * ab * rb - b.x * ab * ra);
            w[2] += (double)(b.y * rb - a.y * ra + a.y * ab * rb - b.y * ab * ra);
            w[3] += (double)(b.z * rb - a.z * ra + a.z * ab * rb - b.z * ab * ra);
            w[4] += 2;
          }
        }
    }
    double t5[5] = {};
    for (int k = 0; k < nz_; k++)
      for (int c = 0; c < 5; c++) t5[c] += sp[(size_t)k * 5 + c];
    mom_.P += std::fabs(t5[0]) / nv;
    mom_.dtheta += t5[4] > 0 ? 1.5 * std::sqrt(t5[1] * t5[1] + t5[2] * t5[2] + t5[3] * t5[3]) / t5[4] : 0.0;
    mom_.divn += tot[DIV] / nv;
  }
}

void CpuEngine::sweeps(int n, int measure_every) {
  for (int s = 0; s < n; s++) {
    long long accepted = 0;
    for (int colour = 0; colour < 2; colour++) {
      switch (pot_) {
        case kGHRL: accepted += colour_pass<kGHRL>(colour, ang_var_); break;
        case kPear: accepted += colour_pass<kPear>(colour, ang_var_); break;
        default: accepted += colour_pass<kLL>(colour, ang_var_); break;
      }
    }
    sweep_++;
    // ang_var rule (evolve.cpp:211-221)
    const double rate = (double)accepted / (double)nvalid_;
    if (rate < 0.5) {
      ang_var_ *= 0.99f;
      if (ang_var_ < 0.01f) ang_var_ = 0.01f;
    } else if (rate > 0.5) {
      ang_var_ /= 0.99f;
      if (ang_var_ > 1.0f) ang_var_ -= 0.5f;
    }
    if (measure_every > 0 && (s + 1) % measure_every == 0) {
      switch (pot_) {
        case kGHRL: measure<kGHRL>(); break;
        case kPear: measure<kPear>(); break;
        default: measure<kLL>(); break;
      }
    }
  }
}

// Local order parameter of every non-empty site: the largest eigenvalue of Q over the site and its existing
// non-empty neighbours (parameter_order.cpp:75-127).
void CpuEngine::snapshot(float *ni, float *local_S) {
  std::copy(n_.begin(), n_.end(), ni);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz_; k++)
    for (int j = 0; j < ny_; j++)
      for (int i = 0; i < nx_; i++) {
        const size_t s = site(i, j, k);
        if (!pt_[s]) {
          local_S[s] = NAN;
          continue;
        }
        double q[6] = {}, cnt = 0;
        long long nb[6];
        neighbours_of(i, j, k, nb);
        for (int d = -1; d < 6; d++) {
          const long long o = d < 0 ? (long long)s : nb[d];
          if (o < 0) continue;
          const V3 m = dir((size_t)o);
          const double x = m.x, y = m.y, z = m.z;
          q[0] += 3 * x * x;
          q[1] += 3 * x * y;
          q[2] += 3 * x * z;
          q[3] += 3 * y * y;
          q[4] += 3 * y * z;
          q[5] += 3 * z * z;
          cnt++;
        }
        const double Q[6] = {(q[0] - cnt) / (2 * cnt), q[1] / (2 * cnt), q[2] / (2 * cnt),
                             (q[3] - cnt) / (2 * cnt), q[4] / (2 * cnt), (q[5] - cnt) / (2 * cnt)};
        double v[3];
        local_S[s] = (float)largest_eigen(Q, v);
      }
}

}  // namespace

std::unique_ptr<Engine> make_cpu_engine(const Params &p, const Lattice &L) {
  return std::unique_ptr<Engine>(new CpuEngine(p, L));
}
