/*
 * CPU oracle for the checkerboard Metropolis sweep of McLics -- TEST INFRASTRUCTURE ONLY (see mcl_oracle.h).
 *
 * Each function cites the reference file:line it restates. Loaded by tests/, the smoke check and bench.py's
 * cpu_baseline leg only; never by the product (libmclgpu, mclics).
 */
#include "mcl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static double orc_wtime(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Pair potentials: potential.cpp:19-56. Float inputs; the double literals of the reference (1.5, 0.25, 1.0/9)
 * promote the arithmetic to double exactly as there, and the result is returned as float. */

float orc_pot_ll(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]) {
  (void)rij;
  float nij = ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2]; /* potential.cpp:29 */
  return (float)(-(double)p->A * 1.5 * nij * nij);            /* potential.cpp:30 */
}

float orc_pot_ghrl(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]) {
  const float el = p->ghrl_lambda, em = p->ghrl_mu, en = p->ghrl_nu, er = p->ghrl_rho, es = p->ghrl_sigma;
  float ai = ni[0] * rij[0] + ni[1] * rij[1] + ni[2] * rij[2]; /* potential.cpp:46 */
  float aj = nj[0] * rij[0] + nj[1] * rij[1] + nj[2] * rij[2]; /* potential.cpp:47 */
  float nij = ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2];   /* potential.cpp:48 */
  float pij = 1.5f * nij * nij - 0.5f;                          /* potential.cpp:49 */
  float cross = (ni[2] * nj[1] - ni[1] * nj[2]) * rij[0]        /* potential.cpp:50-52: r.(nj x ni) */
                + (ni[0] * nj[2] - ni[2] * nj[0]) * rij[1] + (ni[1] * nj[0] - ni[0] * nj[1]) * rij[2];
  float E1 = (float)((1.5 * ai * ai) + (1.5 * aj * aj) - 1); /* potential.cpp:54: P2(ai) + P2(aj) */
  return (float)((E1 * (er * pij + el) + em * (ai * aj * nij - 1.0 / 9)) + en * pij +
                 es * (nij > 0 ? 1 : -1) * cross); /* potential.cpp:55 */
}

float orc_pot_pear(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]) {
  float nij = ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2]; /* potential.cpp:21 */
  float polar_splay = (1 + 2 * nij + nij * nij) *             /* potential.cpp:22 */
                      ((nj[0] - ni[0]) * rij[0] + (nj[1] - ni[1]) * rij[1] + (nj[2] - ni[2]) * rij[2]);
  return (float)(-p->A * nij * nij - p->B1 * nij - (p->C * 0.25) * polar_splay); /* potential.cpp:23 */
}

float orc_pair(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]) {
  switch (p->potential) {
    case ORC_POT_GHRL: return orc_pot_ghrl(p, ni, nj, rij);
    case ORC_POT_PEAR: return orc_pot_pear(p, ni, nj, rij);
    default: return orc_pot_ll(p, ni, nj, rij);
  }
}

/* Field coupling: potential.cpp:59-63. */
float orc_field(const orc_params *p, const float ni[3]) {
  float nDotE = p->elecE * (ni[0] * p->elecX + ni[1] * p->elecY + ni[2] * p->elecZ);
  return -p->elecP * nDotE - p->elecA * nDotE * nDotE;
}

/* Anchoring::lab_axis, anchoring.cpp:19-24. */
void orc_lab_axis(float theta_deg, float phi_deg, float out[3]) {
  const float toRad = (float)(M_PI / 180);
  out[0] = sinf(toRad * theta_deg) * cosf(toRad * phi_deg);
  out[1] = sinf(toRad * theta_deg) * sinf(toRad * phi_deg);
  out[2] = cosf(toRad * theta_deg);
}

/* RP_Tangent_Anchoring::easy_axis, anchoring_rp_tangent.cpp:49-63 (singular_tol 1e-3, anchoring.h:56). */
int orc_rp_tangent_axis(const float d[3], const float s[3], float e[3]) {
  float ds = d[0] * s[0] + d[1] * s[1] + d[2] * s[2];
  e[0] = d[0] - ds * s[0];
  e[1] = d[1] - ds * s[1];
  e[2] = d[2] - ds * s[2];
  float mod = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  if (mod < 1e-3f) {
    e[0] = e[1] = e[2] = 0;
    return 0;
  }
  e[0] /= mod;
  e[1] /= mod;
  e[2] /= mod;
  return 1;
}

/* Anchoring energies of the CPU classes (anchoring_*.cpp), W included. */
float orc_surface(const orc_params *p, int surf, const float ni[3], const float s[3]) {
  const float W = p->surf_W[surf];
  const float *ax = p->surf_axis[surf];
  float nij, e[3];
  switch (p->surf_kind[surf]) {
    case ORC_ANCH_HOMEOTROPIC: /* anchoring_homeotropic.cpp:33-36 */
      nij = ni[0] * s[0] + ni[1] * s[1] + ni[2] * s[2];
      return -W * nij * nij;
    case ORC_ANCH_RP: /* anchoring_rp.cpp:29-32 */
      nij = ni[0] * ax[0] + ni[1] * ax[1] + ni[2] * ax[2];
      return -W * nij * nij;
    case ORC_ANCH_RP_TANGENT: /* anchoring_rp_tangent.cpp:65-70 */
      if (!orc_rp_tangent_axis(ax, s, e)) return 0;
      nij = ni[0] * e[0] + ni[1] * e[1] + ni[2] * e[2];
      return -W * nij * nij;
    case ORC_ANCH_FG: /* anchoring_fg.cpp:17-20 */
      nij = ni[0] * s[0] + ni[1] * s[1] + ni[2] * s[2];
      return +W * nij * nij;
    case ORC_ANCH_HOMEOTROPIC_GHRL: /* anchoring_homeotropic.cpp:53-56 */
      return W * orc_pot_ghrl(p, ni, s, s);
    case ORC_ANCH_RP_GHRL: /* anchoring_rp.cpp:62-65 */
      return W * orc_pot_ghrl(p, ni, ax, s);
    case ORC_ANCH_FG_GHRL: { /* anchoring_fg.cpp:37-47 */
      float ns_ = ni[0] * s[0] + ni[1] * s[1] + ni[2] * s[2];
      float nj[3] = {ni[0] - ns_ * s[0], ni[1] - ns_ * s[1], ni[2] - ns_ * s[2]};
      float mod = sqrtf(nj[0] * nj[0] + nj[1] * nj[1] + nj[2] * nj[2]);
      if (mod < 1e-6f) return 0;
      nj[0] /= mod;
      nj[1] /= mod;
      nj[2] /= mod;
      return W * orc_pot_ghrl(p, ni, nj, s);
    }
  }
  return 0;
}

/* setGHRL, io.cpp:346-371. */
void orc_set_ghrl(orc_params *p, float k11, float k22, float k33, float p0, float *scale_out) {
  float Scale = (float)fabs((9.0) / (k11 - 3 * k22 - k33));
  p->ghrl_sigma = p0 ? (float)(-Scale * k22 * (2 * M_PI / p0)) : 0;
  p->ghrl_lambda = (Scale / 9) * (2 * k11 - 3 * k22 + k33);
  p->ghrl_mu = Scale * (k22 - k11);
  p->ghrl_rho = (Scale / 9) * (k11 - k33);
  p->ghrl_nu = (Scale / 9) * (k11 - 3 * k22 - k33);
  if (scale_out) *scale_out = Scale;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Neighbour lookup: Periodic_Boundary / Free_Boundary (potential.cpp:64-74) and setNni (evolve.cpp:23-32). */

static int orc_bound(int periodic, int *ii, int NN) {
  if (periodic) {
    *ii = (*ii + NN) % NN;
    return 1;
  }
  return (*ii == -1 || *ii == NN) ? 0 : 1;
}

/* Fills the 6 neighbour indices of site (i, j, k) in the order +x, -x, +y, -y, +z, -z (evolve.cpp:133-146);
 * -1 where the neighbour does not exist across a free boundary. */
static void orc_neighbours(const orc_params *p, int i, int j, int k, int nb[6]) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1, km = k - 1;
  int e0 = orc_bound(p->periodic[0], &ip, Nx), e1 = orc_bound(p->periodic[0], &im, Nx);
  int e2 = orc_bound(p->periodic[1], &jp, Ny), e3 = orc_bound(p->periodic[1], &jm, Ny);
  int e4 = orc_bound(p->periodic[2], &kp, Nz), e5 = orc_bound(p->periodic[2], &km, Nz);
  nb[0] = e0 ? ip + Nx * (j + Ny * k) : -1;
  nb[1] = e1 ? im + Nx * (j + Ny * k) : -1;
  nb[2] = e2 ? i + Nx * (jp + Ny * k) : -1;
  nb[3] = e3 ? i + Nx * (jm + Ny * k) : -1;
  nb[4] = e4 ? i + Nx * (j + Ny * kp) : -1;
  nb[5] = e5 ? i + Nx * (j + Ny * km) : -1;
}

static const float ORC_RIJ[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

/* Site energy: Geometry::bond_energy (geometry.cpp:23-76, double accumulator, float result) plus the surface
 * term of Slab/Sphere/Custom site_energy (geometry_slab.cpp:54-67 etc.) and the field (geometry_bulk.cpp:67-75). */
float orc_site_energy(const orc_params *p, const float *ni, const int *pt, const float *ns, int idx, const float n0[3]) {
  const int Nx = p->Nx, Ny = p->Ny;
  int i = idx % Nx, j = (idx / Nx) % Ny, k = idx / (Nx * Ny);
  int nb[6];
  orc_neighbours(p, i, j, k, nb);
  double Eb = 0;
  for (int d = 0; d < 6; d++) {
    if (nb[d] < 0 || pt[nb[d]] == 0) continue; /* absent neighbour or empty site: no bond */
    const float *nj = &ni[3 * nb[d]];
    Eb += orc_pair(p, n0, nj, ORC_RIJ[d]);
  }
  float E = (float)Eb;
  if (pt[idx] > 1 && ns) E += orc_surface(p, pt[idx] - 2, n0, &ns[3 * idx]);
  if (p->elecA != 0 || p->elecP != 0) E += orc_field(p, n0);
  return E;
}

/* EvolveN::energy_calculator, evolve.cpp:33-75. */
double orc_energy_mean(const orc_params *p, const float *ni, const int *pt, const float *ns) {
  const int Nt = p->Nx * p->Ny * p->Nz;
  double Etot = 0;
  int nValid = 0;
#pragma omp parallel for reduction(+ : Etot, nValid) schedule(static)
  for (int idx = 0; idx < Nt; idx++) {
    if (pt[idx] == 0) continue;
    Etot += orc_site_energy(p, ni, pt, ns, idx, &ni[3 * idx]);
    nValid++;
  }
  return nValid ? Etot / nValid : 0;
}

/* Total energy, each bond once: sum over sites of the bonds to +x, +y, +z only, plus anchoring and field. This is
 * an independent loop (not the site energy), so dE = E_total(after) - E_total(before) checks the site energy. */
double orc_energy_total(const orc_params *p, const float *ni, const int *pt, const float *ns) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  double E = 0;
  for (int k = 0; k < Nz; k++)
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        int idx = i + Nx * (j + Ny * k);
        if (!pt[idx]) continue;
        const float *n0 = &ni[3 * idx];
        int nb[6];
        orc_neighbours(p, i, j, k, nb);
        for (int d = 0; d < 6; d += 2) { /* +x, +y, +z */
          if (nb[d] < 0 || !pt[nb[d]]) continue;
          E += orc_pair(p, n0, &ni[3 * nb[d]], ORC_RIJ[d]);
        }
        if (pt[idx] > 1 && ns) E += orc_surface(p, pt[idx] - 2, n0, &ns[3 * idx]);
        if (p->elecA != 0 || p->elecP != 0) E += orc_field(p, n0);
      }
  return E;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Geometries: point types and outward normals. */

int orc_geometry(int geom, int Nx, int Ny, int Nz, int *pt, float *ns) {
  const int Nt = Nx * Ny * Nz;
  if (ns) memset(ns, 0, sizeof(float) * 3 * (size_t)Nt);
  if (geom == ORC_GEOM_BULK) { /* geometry_bulk.cpp:61-65 */
    for (int ii = 0; ii < Nt; ii++) pt[ii] = 1;
    return 0;
  }
  if (geom == ORC_GEOM_SLAB) { /* geometry_slab.cpp:51-71 */
    for (int ii = 0; ii < Nx; ii++)
      for (int jj = 0; jj < Ny; jj++) {
        int kk = 0;
        pt[ii + Nx * (jj + Ny * kk)] = 2;
        if (ns) {
          ns[(ii + Nx * (jj + Ny * kk)) * 3 + 0] = 0;
          ns[(ii + Nx * (jj + Ny * kk)) * 3 + 1] = 0;
          ns[(ii + Nx * (jj + Ny * kk)) * 3 + 2] = -1;
        }
        for (kk = 1; kk < Nz - 1; kk++) pt[ii + Nx * (jj + Ny * kk)] = 1;
        kk = Nz - 1;
        pt[ii + Nx * (jj + Ny * kk)] = 3;
        if (ns) {
          ns[(ii + Nx * (jj + Ny * kk)) * 3 + 0] = 0;
          ns[(ii + Nx * (jj + Ny * kk)) * 3 + 1] = 0;
          ns[(ii + Nx * (jj + Ny * kk)) * 3 + 2] = 1;
        }
      }
    return 0;
  }
  if (geom == ORC_GEOM_SPHERE) { /* geometry_sphere.cpp:32-58 */
    float Hx = (float)(Nx / 2), Hy = (float)(Ny / 2), Hz = (float)(Nz / 2);
    for (int ii = 0; ii < Nx; ii++)
      for (int jj = 0; jj < Ny; jj++)
        for (int kk = 0; kk < Nz; kk++) {
          float Rx = ii - Hx, Ry = jj - Hy, Rz = kk - Hz;
          float radius = (float)sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
          int idx = ii + Nx * (jj + Ny * kk);
          if (radius < Hz - 1)
            pt[idx] = 1;
          else if (radius < Hz + 1) {
            pt[idx] = 2;
            if (ns) {
              ns[idx * 3 + 0] = Rx / radius;
              ns[idx * 3 + 1] = Ry / radius;
              ns[idx * 3 + 2] = Rz / radius;
            }
          } else
            pt[idx] = 0;
        }
    return 0;
  }
  return -1; /* custom geometries are passed in as arrays */
}

/* homogeneous_ic, ic.cpp:67-92. */
void orc_ic_homogeneous(int Nt, const int *pt, float *ni, float theta_deg, float phi_deg) {
  for (int idx = 0; idx < Nt; idx++) {
    if (pt[idx]) {
      ni[3 * idx + 0] = (float)(sin(theta_deg * M_PI / 180) * cos(phi_deg * M_PI / 180));
      ni[3 * idx + 1] = (float)(sin(theta_deg * M_PI / 180) * sin(phi_deg * M_PI / 180));
      ni[3 * idx + 2] = (float)cos(theta_deg * M_PI / 180);
    } else {
      ni[3 * idx + 0] = 0;
      ni[3 * idx + 1] = 0;
      ni[3 * idx + 2] = 1;
    }
  }
}

/* random_ic, ic.cpp:38-65: isotropic unit vectors (z uniform in [-1, 1), azimuth uniform; GSL taus and
 * gsl_ran_dir_3d are not available, so the stream differs). Empty sites get (0, 0, 1). */
void orc_ic_random(int Nt, const int *pt, float *ni, uint64_t seed) {
  orc_rng r;
  orc_rng_seed(&r, seed, 0x1C);
  for (int idx = 0; idx < Nt; idx++) {
    if (pt[idx]) {
      double z = 2.0 * orc_rng_uniform(&r) - 1.0;
      double phi = 2.0 * M_PI * orc_rng_uniform(&r);
      double s = sqrt(fmax(0.0, 1.0 - z * z));
      ni[3 * idx + 0] = (float)(s * cos(phi));
      ni[3 * idx + 1] = (float)(s * sin(phi));
      ni[3 * idx + 2] = (float)z;
    } else {
      ni[3 * idx + 0] = 0;
      ni[3 * idx + 1] = 0;
      ni[3 * idx + 2] = 1;
    }
  }
}

int orc_count_valid(int Nt, const int *pt) {
  int n = 0;
  for (int i = 0; i < Nt; i++) n += pt[i] ? 1 : 0;
  return n;
}

/* evolve.cu:348-391: the per-site vector of the GPU anchoring kernels. */
int orc_gpu_surface_vectors(const orc_params *p, int Nt, const int *pt, const float *ns, float *sv) {
  int nSingular = 0;
  memcpy(sv, ns, sizeof(float) * 3 * (size_t)Nt);
  for (int idx = 0; idx < Nt; idx++) {
    if (pt[idx] < 2) continue;
    int s = pt[idx] - 2;
    if (p->surf_kind[s] == ORC_ANCH_RP) {
      sv[3 * idx + 0] = p->surf_axis[s][0];
      sv[3 * idx + 1] = p->surf_axis[s][1];
      sv[3 * idx + 2] = p->surf_axis[s][2];
    } else if (p->surf_kind[s] == ORC_ANCH_RP_TANGENT) {
      if (!orc_rp_tangent_axis(p->surf_axis[s], &ns[3 * idx], &sv[3 * idx])) nSingular++;
    }
  }
  return nSingular;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* RNG: xoshiro128** (Blackman & Vigna, "Scrambled linear pseudorandom number generators", 2018), seeded by
 * splitmix64. Stands in for GSL ranlxs0; uniform in [0, 1) with 24 random bits like a float draw. */

static uint64_t orc_splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void orc_rng_seed(orc_rng *r, uint64_t seed, uint64_t stream) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ULL ^ (stream + 0x632BE59BD9B4E019ULL);
  uint64_t a = orc_splitmix64(&x), b = orc_splitmix64(&x);
  r->s[0] = (uint32_t)a;
  r->s[1] = (uint32_t)(a >> 32);
  r->s[2] = (uint32_t)b;
  r->s[3] = (uint32_t)(b >> 32);
  if (!(r->s[0] | r->s[1] | r->s[2] | r->s[3])) r->s[0] = 1;
}

static inline uint32_t orc_rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

static inline uint32_t orc_rng_next(orc_rng *r) {
  uint32_t *s = r->s;
  const uint32_t result = orc_rotl(s[1] * 5, 7) * 9;
  const uint32_t t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = orc_rotl(s[3], 11);
  return result;
}

float orc_rng_uniform(orc_rng *r) { return (float)(orc_rng_next(r) >> 8) * (1.0f / 16777216.0f); }

/* ------------------------------------------------------------------------------------------------------------ */
/* Trial rotation: evolve.cpp:153-176 (rotation by pi*va about x, y or z; thresholds 0.333 and 0.666). */
void orc_rotate(const float n[3], float va, float rot, float out[3]) {
  float c = cosf((float)(M_PI * va)), s = sinf((float)(M_PI * va));
  if (rot < 0.333) {
    out[0] = n[0];
    out[1] = n[1] * c + n[2] * s;
    out[2] = n[2] * c - n[1] * s;
  } else if (rot < 0.666) {
    out[0] = n[0] * c - n[2] * s;
    out[1] = n[1];
    out[2] = n[2] * c + n[0] * s;
  } else {
    out[0] = n[0] * c + n[1] * s;
    out[1] = n[1] * c - n[0] * s;
    out[2] = n[2];
  }
  float norm = sqrtf(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]); /* evolve.cpp:171-176 */
  out[0] /= norm;
  out[1] /= norm;
  out[2] /= norm;
}

/* evolve.cpp:211-220. */
float orc_ang_var_update(float ang_var, long acceptance, int nvalid) {
  if (1.0 * acceptance / nvalid < 0.5) {
    ang_var *= 0.99;
    if (ang_var < 0.01) ang_var = 0.01f;
  } else if (1.0 * acceptance / nvalid > 0.5) {
    ang_var /= 0.99;
    if (ang_var > 1.0) ang_var -= 0.5;
  }
  return ang_var;
}

/* EvolveN::Monte_Carlo_Step, evolve.cpp:85-221: 8 sublattice ticks (di = tick%2, dj = (tick/2)%2, dk = tick/4),
 * one omp-for over the 2x2x2 boxes per tick, one RNG per thread. */
long orc_sweep(const orc_params *p, float *ni, const int *pt, const float *ns, float *ang_var_io, orc_rng *rngs,
               int nthreads) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  const float T = p->T;
  const float ang_var = *ang_var_io;
  const int nBoxX = Nx / 2 + Nx % 2, nBoxY = Ny / 2 + Ny % 2, nBoxZ = Nz / 2 + Nz % 2;
  const int Ntt = nBoxX * nBoxY * nBoxZ;
  long acceptance = 0;
  int nValid = orc_count_valid(Nx * Ny * Nz, pt);
#pragma omp parallel num_threads(nthreads) reduction(+ : acceptance)
  {
#ifdef _OPENMP
    orc_rng *r = &rngs[omp_get_thread_num()];
#else
    orc_rng *r = &rngs[0];
#endif
    for (int tick = 0; tick < 8; tick++) {
      const int di = tick % 2, dj = (tick / 2) % 2, dk = tick / 4;
#pragma omp for schedule(static)
      for (int nt = 0; nt < Ntt; nt++) {
        int iBox = nt % nBoxX, jBox = (nt / nBoxX) % nBoxY, kBox = nt / nBoxX / nBoxY;
        int i = di + 2 * iBox, j = dj + 2 * jBox, k = dk + 2 * kBox;
        if (i >= Nx || j >= Ny || k >= Nz) continue;
        int idx = i + Nx * (j + Ny * k);
        if (pt[idx] == 0) continue;
        float *n = &ni[3 * idx];
        float va = (2 * orc_rng_uniform(r) - 1) * ang_var; /* evolve.cpp:153 */
        float rot = orc_rng_uniform(r);                       /* evolve.cpp:154 */
        float nNew[3];
        orc_rotate(n, va, rot, nNew);
        float n0[3] = {n[0], n[1], n[2]};
        float E_old = orc_site_energy(p, ni, pt, ns, idx, n0); /* evolve.cpp:178 */
        float E_new = orc_site_energy(p, ni, pt, ns, idx, nNew); /* evolve.cpp:182 */
        float E_cur = E_old;
        float u = orc_rng_uniform(r); /* evolve.cpp:187 */
        if (T > 0 ? u < exp(-(E_new - E_old) / T) : E_new <= E_old) {
          n[0] = nNew[0];
          n[1] = nNew[1];
          n[2] = nNew[2];
          E_cur = E_new;
          acceptance++;
        }
        if (p->flipMove) { /* evolve.cpp:196-207 */
          float nf[3] = {-n[0], -n[1], -n[2]};
          float E_flip = orc_site_energy(p, ni, pt, ns, idx, nf);
          float uf = orc_rng_uniform(r);
          if (T > 0 ? uf < exp(-(E_flip - E_cur) / T) : E_flip <= E_cur) {
            n[0] = nf[0];
            n[1] = nf[1];
            n[2] = nf[2];
          }
        }
      }
    }
  }
  *ang_var_io = orc_ang_var_update(ang_var, acceptance, nValid);
  return acceptance;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Observables. */

/* Q_tensor, parameter_order.cpp:40-71 (double sums). */
void orc_q_tensor(int Nt, const float *ni, const int *pt, double Q[9]) {
  double q0 = 0, q1 = 0, q2 = 0, q4 = 0, q5 = 0;
  long points = 0;
  for (int idx = 0; idx < Nt; idx++) {
    if (!pt[idx]) continue;
    double x = ni[3 * idx], y = ni[3 * idx + 1], z = ni[3 * idx + 2];
    q0 += 3.0 * x * x - 1;
    q1 += 3.0 * x * y;
    q2 += 3.0 * x * z;
    q4 += 3.0 * y * y - 1;
    q5 += 3.0 * y * z;
    points++;
  }
  double d = points ? 2.0 * points : 1.0;
  Q[0] = q0 / d;
  Q[1] = q1 / d;
  Q[2] = q2 / d;
  Q[4] = q4 / d;
  Q[5] = q5 / d;
  Q[3] = Q[1];
  Q[6] = Q[2];
  Q[7] = Q[5];
  Q[8] = -Q[0] - Q[4];
}

/* Symmetric 3x3 eigen-solve by cyclic Jacobi rotations (Golub & Van Loan, Alg. 8.4.3), replacing GSL symmv +
 * symmv_sort(VAL_DESC) of parameter_order.cpp:19-37. Returns the largest eigenvalue; vec = its eigenvector. */
double orc_eig_sym3(const double Ain[9], double vec[3], double evals_desc[3]) {
  double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (Ain[3 * i + j] + Ain[3 * j + i]);
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off < 1e-300) break;
    for (int pq = 0; pq < 3; pq++) {
      int pp = pq == 2 ? 1 : 0, qq = pq == 0 ? 1 : 2;
      if (fabs(a[pp][qq]) < 1e-300) continue;
      double theta = (a[qq][qq] - a[pp][pp]) / (2 * a[pp][qq]);
      double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
      double c = 1 / sqrt(t * t + 1), s = t * c;
      for (int kk = 0; kk < 3; kk++) { /* A <- J^T A J */
        double akp = a[kk][pp], akq = a[kk][qq];
        a[kk][pp] = c * akp - s * akq;
        a[kk][qq] = s * akp + c * akq;
      }
      for (int kk = 0; kk < 3; kk++) {
        double apk = a[pp][kk], aqk = a[qq][kk];
        a[pp][kk] = c * apk - s * aqk;
        a[qq][kk] = s * apk + c * aqk;
      }
      for (int kk = 0; kk < 3; kk++) {
        double vkp = v[kk][pp], vkq = v[kk][qq];
        v[kk][pp] = c * vkp - s * vkq;
        v[kk][qq] = s * vkp + c * vkq;
      }
    }
  }
  int order[3] = {0, 1, 2};
  for (int x = 0; x < 3; x++)
    for (int y = x + 1; y < 3; y++)
      if (a[order[y]][order[y]] > a[order[x]][order[x]]) {
        int t = order[x];
        order[x] = order[y];
        order[y] = t;
      }
  if (evals_desc)
    for (int x = 0; x < 3; x++) evals_desc[x] = a[order[x]][order[x]];
  if (vec)
    for (int x = 0; x < 3; x++) vec[x] = v[x][order[0]];
  return a[order[0]][order[0]];
}

/* Radial_order, parameter_order.cpp:223-245. */
void orc_radial_order(const orc_params *p, const float *ni, const int *pt, int cylinder, double *Sr, double *Pr) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  const float cx = (float)(Nx / 2), cy = (float)(Ny / 2), cz = (float)(Nz / 2);
  double s = 0, pp = 0;
  long valid = 0;
  for (int k = 0; k < Nz; k++)
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        const int idx = i + Nx * (j + Ny * k);
        if (!pt[idx]) continue;
        const float rx = i - cx, ry = j - cy, rz = cylinder ? 0 : k - cz;
        const float r = sqrtf(rx * rx + ry * ry + rz * rz);
        if (r < 0.5f) continue;
        const double c = (ni[3 * idx] * rx + ni[3 * idx + 1] * ry + ni[3 * idx + 2] * rz) / r;
        s += 1.5 * c * c - 0.5;
        pp += c;
        valid++;
      }
  *Sr = valid ? s / valid : 0;
  *Pr = valid ? pp / valid : 0;
}

/* lattice_order_parameter, parameter_order.cpp:75-127. */
double orc_local_order(const orc_params *p, const float *ni, const int *pt, int i, int j, int k) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  double Q[9] = {0};
  int points = 0;
  for (int di = -1; di < 2; di++)
    for (int dj = -1; dj < 2; dj++)
      for (int dk = -1; dk < 2; dk++) {
        if (abs(di) + abs(dj) + abs(dk) > 1) continue;
        int ii = i + di, jj = j + dj, kk = k + dk;
        if (orc_bound(p->periodic[0], &ii, Nx) && orc_bound(p->periodic[1], &jj, Ny) &&
            orc_bound(p->periodic[2], &kk, Nz) && pt[ii + Nx * (jj + Ny * kk)] != 0) {
          const float *n = &ni[(ii + Nx * (jj + Ny * kk)) * 3];
          for (int l = 0; l < 3; l++) {
            for (int m = 0; m <= l; m++) Q[3 * m + l] += 3.0 * (n[l] * n[m]);
            Q[4 * l] -= 1;
          }
          points++;
        }
      }
  const int upper[6] = {0, 1, 2, 4, 5, 8};
  for (int u = 0; u < 6; u++) Q[upper[u]] /= 2.0 * points;
  Q[3] = Q[1];
  Q[6] = Q[2];
  Q[7] = Q[5];
  return orc_eig_sym3(Q, NULL, NULL);
}

/* Polarization, parameter_order.cpp:131-141. */
double orc_polarization(int Nt, const float *ni, const int *pt, const double dir[3]) {
  double sum = 0;
  int valid = 0;
  for (int idx = 0; idx < Nt; idx++) {
    if (!pt[idx]) continue;
    sum += dir[0] * ni[3 * idx] + dir[1] * ni[3 * idx + 1] + dir[2] * ni[3 * idx + 2];
    valid++;
  }
  return valid ? fabs(sum) / valid : 0;
}

/* Splay_magnitude, parameter_order.cpp:147-176. */
double orc_splay_magnitude(const orc_params *p, const float *ni, const int *pt) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  const int dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  double s[3] = {0, 0, 0};
  long nBonds = 0;
  for (int k = 0; k < Nz; k++)
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        const int idx = i + Nx * (j + Ny * k);
        if (!pt[idx]) continue;
        const float *a = &ni[3 * idx];
        for (int d = 0; d < 6; d++) {
          int ii = i + dirs[d][0], jj = j + dirs[d][1], kk = k + dirs[d][2];
          if (!(orc_bound(p->periodic[0], &ii, Nx) && orc_bound(p->periodic[1], &jj, Ny) &&
                orc_bound(p->periodic[2], &kk, Nz)))
            continue;
          const int jdx = ii + Nx * (jj + Ny * kk);
          if (!pt[jdx]) continue;
          const float *b = &ni[3 * jdx];
          const float ra = a[0] * dirs[d][0] + a[1] * dirs[d][1] + a[2] * dirs[d][2];
          const float rb = b[0] * dirs[d][0] + b[1] * dirs[d][1] + b[2] * dirs[d][2];
          const float ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
          for (int c = 0; c < 3; c++) s[c] += 0.5 * (b[c] * rb - a[c] * ra + a[c] * ab * rb - b[c] * ab * ra);
          nBonds++;
        }
      }
  if (nBonds == 0) return 0;
  return 1.5 * sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) / nBonds;
}

/* Splay_divergence, parameter_order.cpp:183-215. */
double orc_splay_divergence(const orc_params *p, const float *ni, const int *pt) {
  const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  double sum = 0;
  long valid = 0;
  for (int k = 0; k < Nz; k++)
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        const int idx = i + Nx * (j + Ny * k);
        if (!pt[idx]) continue;
        double div = 0;
        for (int ax = 0; ax < 3; ax++) {
          double d = 0;
          int nb = 0;
          for (int s = -1; s <= 1; s += 2) {
            int ii = i, jj = j, kk = k;
            if (ax == 0)
              ii += s;
            else if (ax == 1)
              jj += s;
            else
              kk += s;
            if (!(orc_bound(p->periodic[0], &ii, Nx) && orc_bound(p->periodic[1], &jj, Ny) &&
                  orc_bound(p->periodic[2], &kk, Nz)))
              continue;
            const int jdx = ii + Nx * (jj + Ny * kk);
            if (!pt[jdx]) continue;
            d += s * (ni[3 * jdx + ax] - ni[3 * idx + ax]);
            nb++;
          }
          if (nb) div += d / nb;
        }
        sum += div;
        valid++;
      }
  return valid ? sum / valid : 0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Sweep schedule helpers: include/sweep.h:13-34. */
int orc_sweep_points(double a, double b, double d) { return (int)floor((b - a) / d + 1e-3) + 1; }
double orc_sweep_value(double a, double b, double d, int n, int npoints) {
  if (n == npoints - 1 && fabs((b - a) / d - (npoints - 1)) < 1e-3) return b;
  return a + n * d;
}
int orc_snapshot_scale(double step) {
  int S = 100;
  while (S * fabs(step) < 1 - 1e-3 && S < 100000000) S *= 10;
  return S;
}
long orc_snapshot_index(double x, int S) { return (long)(S * (x + 1e-3 / S)); }

/* One point of a mode loop: MCT sweeps, then MCS sweeps with one measurement each (evolve_thermal.cpp:160-200,
 * evolve_electric.cpp:68-90). row = (x, S, varS, E, varE, Sr_sph or NaN). */
static void orc_point(orc_params *p, float *ni, const int *pt, const float *ns, int MCT, int MCS, float *ang_var,
                      orc_rng *rngs, int nthreads, int radial, double x, double *row) {
  const int Nt = p->Nx * p->Ny * p->Nz;
  for (int step = 0; step < MCT; step++) orc_sweep(p, ni, pt, ns, ang_var, rngs, nthreads);
  double S1 = 0, S2 = 0, E = 0, E2 = 0, Sr = 0;
  for (int step = 0; step < MCS; step++) {
    orc_sweep(p, ni, pt, ns, ang_var, rngs, nthreads);
    double e = orc_energy_mean(p, ni, pt, ns);
    E += e;
    E2 += e * e;
    double Q[9], v[3];
    orc_q_tensor(Nt, ni, pt, Q);
    double s = orc_eig_sym3(Q, v, NULL);
    S1 += s;
    S2 += s * s;
    if (radial) {
      double sr, pr;
      orc_radial_order(p, ni, pt, 0, &sr, &pr);
      Sr += sr;
    }
  }
  E /= MCS;
  E2 /= MCS;
  S1 /= MCS;
  S2 /= MCS;
  row[0] = x;
  row[1] = S1;
  row[2] = S2 - S1 * S1;
  row[3] = E;
  row[4] = E2 - E * E;
  row[5] = radial ? Sr / MCS : NAN;
}

static orc_rng *orc_thread_rngs(uint64_t seed, int nthreads) {
  orc_rng *rngs = (orc_rng *)calloc((size_t)nthreads, sizeof(orc_rng));
  for (int t = 0; t < nthreads; t++) orc_rng_seed(&rngs[t], seed * 104729ULL + (uint64_t)t, 0x5EED);
  return rngs;
}

/* thermalEvolveN::run, evolve_thermal.cpp:144-212 (one RNG per thread seeded from seed*104729 + thread). */
int orc_thermal_run(orc_params *p, float *ni, const int *pt, const float *ns, double Ti, double Tf, double dT,
                    int MCT, int MCS, uint64_t seed, int nthreads, double *rows, int max_rows) {
  orc_rng *rngs = orc_thread_rngs(seed, nthreads);
  float ang_var = 0.5f;
  const int nT = orc_sweep_points(Ti, Tf, dT);
  for (int it = 0; it < nT && it < max_rows; it++) {
    p->T = (float)orc_sweep_value(Ti, Tf, dT, it, nT);
    orc_point(p, ni, pt, ns, MCT, MCS, &ang_var, rngs, nthreads, 1, p->T, rows + 6 * it);
  }
  free(rngs);
  return nT < max_rows ? nT : max_rows;
}

/* electricEvolveN::run, evolve_electric.cpp:31-103: field magnitude swept Ei, Ei + dE, ..., Ef at the fixed
 * temperature p->T, ang_var carried from point to point (starts at 0.5). rows: nE x 6 (field, S, varS, E, varE,
 * NaN); returns nE. */
int orc_electric_run(orc_params *p, float *ni, const int *pt, const float *ns, double Ei, double Ef, double dE,
                     int MCT, int MCS, uint64_t seed, int nthreads, double *rows, int max_rows) {
  orc_rng *rngs = orc_thread_rngs(seed, nthreads);
  float ang_var = 0.5f;
  const int nE = orc_sweep_points(Ei, Ef, dE);
  for (int ie = 0; ie < nE && ie < max_rows; ie++) {
    p->elecE = (float)orc_sweep_value(Ei, Ef, dE, ie, nE);
    orc_point(p, ni, pt, ns, MCT, MCS, &ang_var, rngs, nthreads, 0, p->elecE, rows + 6 * ie);
  }
  free(rngs);
  return nE < max_rows ? nE : max_rows;
}

double orc_time_sweeps(const orc_params *p, float *ni, const int *pt, const float *ns, int nsweeps, uint64_t seed,
                       int nthreads, float *ang_var) {
  orc_rng *rngs = (orc_rng *)calloc((size_t)nthreads, sizeof(orc_rng));
  for (int t = 0; t < nthreads; t++) orc_rng_seed(&rngs[t], seed * 104729ULL + (uint64_t)t, 0x5EED);
  double t0 = orc_wtime();
  for (int s = 0; s < nsweeps; s++) orc_sweep(p, ni, pt, ns, ang_var, rngs, nthreads);
  double t1 = orc_wtime();
  free(rngs);
  return t1 - t0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11. Round: (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2; c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0);
 * the key is bumped by the Weyl constants between rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
