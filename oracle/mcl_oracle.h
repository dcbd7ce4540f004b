/*
 * CPU oracle for the checkerboard Metropolis sweep of McLics -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's CPU path (src/evolve.cpp, src/potential.cpp, src/geometry*.cpp,
 * src/anchoring*.cpp, src/parameter_order.cpp, include/sweep.h, src/io.cpp:setGHRL). It is the checker that the
 * GPU library (libmclgpu) is compared against, and the CPU baseline that bench.py times. Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it; the product path (libmclgpu, the mclics binary) never does.
 *
 * Deliberate differences from the reference, none of which changes the sampled distribution:
 *   - RNG: GSL ranlxs0 (one generator per OpenMP thread, evolve_thermal.cpp:43-51) is not available here; each
 *     thread gets a xoshiro128** stream (Blackman & Vigna, 2018) seeded from (seed, thread) by splitmix64. RNG value
 *     parity with the reference is therefore unpinned; only the statistics are compared.
 *   - Q-tensor sums are accumulated in double (the reference sums floats, parameter_order.cpp:45-60).
 *   - The 3x3 eigen-solve (GSL gsl_eigen_symmv in the reference, parameter_order.cpp:19-37) is a cyclic Jacobi
 *     iteration in double.
 *
 * Pinning of each function is listed in the test files under tests/ (test_oracle_*.py); none is "parity unpinned"
 * except the RNG streams (by design).
 */
#ifndef MCL_ORACLE_H_
#define MCL_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_POT_LL = 0, ORC_POT_GHRL = 1, ORC_POT_PEAR = 2 };
enum { ORC_GEOM_BULK = 0, ORC_GEOM_SLAB = 1, ORC_GEOM_SPHERE = 2, ORC_GEOM_CUSTOM = 3 };
/* Anchoring classes of include/anchoring.h:29-104 */
enum {
  ORC_ANCH_HOMEOTROPIC = 0,      /* -W (n.s)^2                     anchoring_homeotropic.cpp:33-36 */
  ORC_ANCH_RP = 1,               /* -W (n.e)^2, e = lab axis       anchoring_rp.cpp:29-32 */
  ORC_ANCH_RP_TANGENT = 2,       /* -W (n.e)^2, e = projected axis anchoring_rp_tangent.cpp:49-70 */
  ORC_ANCH_FG = 3,               /* +W (n.s)^2                     anchoring_fg.cpp:17-20 */
  ORC_ANCH_HOMEOTROPIC_GHRL = 4, /* W U_GHRL(n, s, s)              anchoring_homeotropic.cpp:53-56 */
  ORC_ANCH_RP_GHRL = 5,          /* W U_GHRL(n, e, s)              anchoring_rp.cpp:62-65 */
  ORC_ANCH_FG_GHRL = 6           /* W U_GHRL(n, t(n), s)           anchoring_fg.cpp:37-47 */
};

#define ORC_MAX_SURF 64

/* Run parameters: the subset of Parameters (include/parameters.h:18-98) that the sweep reads. */
typedef struct orc_params {
  int Nx, Ny, Nz;
  int periodic[3]; /* 1 periodic, 0 free (Periodic_Boundary / Free_Boundary, potential.cpp:64-74) */
  int potential;
  float A, B1, C;
  float ghrl_lambda, ghrl_mu, ghrl_nu, ghrl_rho, ghrl_sigma;
  float elecX, elecY, elecZ, elecA, elecP, elecE;
  int flipMove;
  float T;
  int nSurf;
  int surf_kind[ORC_MAX_SURF];
  float surf_W[ORC_MAX_SURF];
  float surf_axis[ORC_MAX_SURF][3]; /* lab easy axis (rp, rp_tangent, rp_ghrl) */
} orc_params;

/* ---- potentials (potential.cpp) ---- */
float orc_pot_ll(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]);
float orc_pot_ghrl(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]);
float orc_pot_pear(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]);
float orc_pair(const orc_params *p, const float ni[3], const float nj[3], const float rij[3]);
float orc_field(const orc_params *p, const float ni[3]);
float orc_surface(const orc_params *p, int surf, const float ni[3], const float s[3]);
void orc_set_ghrl(orc_params *p, float k11, float k22, float k33, float p0, float *scale_out);
void orc_lab_axis(float theta_deg, float phi_deg, float out[3]);
int orc_rp_tangent_axis(const float d[3], const float s[3], float e[3]);

/* ---- lattice energies ---- */
/* Site energy of site idx with director n0 (bonds + anchoring + field), geometry.cpp:23-76 + geometry_*.cpp. */
float orc_site_energy(const orc_params *p, const float *ni, const int *pt, const float *ns, int idx, const float n0[3]);
/* Mean site energy over non-empty sites (each bond counted twice), evolve.cpp:33-75. */
double orc_energy_mean(const orc_params *p, const float *ni, const int *pt, const float *ns);
/* Total energy with every bond counted once (independent unique-bond loop; used to check dE). */
double orc_energy_total(const orc_params *p, const float *ni, const int *pt, const float *ns);

/* ---- geometry and initial conditions ---- */
int orc_geometry(int geom, int Nx, int Ny, int Nz, int *pt, float *ns);
void orc_ic_homogeneous(int Nt, const int *pt, float *ni, float theta_deg, float phi_deg);
void orc_ic_random(int Nt, const int *pt, float *ni, uint64_t seed);
int orc_count_valid(int Nt, const int *pt);
/* Per-site anchoring vector passed to the GPU path: the normal, replaced by the easy axis for rp / rp_tangent
 * (evolve.cu:348-391). Returns the number of singular rp_tangent sites. */
int orc_gpu_surface_vectors(const orc_params *p, int Nt, const int *pt, const float *ns, float *sv);

/* ---- RNG and sweep ---- */
typedef struct orc_rng { uint32_t s[4]; } orc_rng;
void orc_rng_seed(orc_rng *r, uint64_t seed, uint64_t stream);
float orc_rng_uniform(orc_rng *r); /* [0, 1) like gsl_rng_uniform */
/* One sweep: 8 sublattice ticks, trial rotation, Metropolis, flip, ang_var rule (evolve.cpp:85-221).
 * rngs: nthreads generators. Returns the number of accepted rotations. */
long orc_sweep(const orc_params *p, float *ni, const int *pt, const float *ns, float *ang_var, orc_rng *rngs,
               int nthreads);
/* ang_var adaptation after a sweep (evolve.cpp:211-220, evolve.cu:430-436). */
float orc_ang_var_update(float ang_var, long acceptance, int nvalid);
/* Trial rotation of n by pi*va about axis chosen by rot (evolve.cpp:153-176), renormalized. */
void orc_rotate(const float n[3], float va, float rot, float out[3]);

/* ---- observables ---- */
/* Q = <(3nn - I)/2> over non-empty sites, row-major 3x3 (parameter_order.cpp:40-71). */
void orc_q_tensor(int Nt, const float *ni, const int *pt, double Q[9]);
/* Largest eigenvalue and its eigenvector of a symmetric 3x3 matrix (Jacobi, double). */
double orc_eig_sym3(const double A[9], double vec[3], double evals_desc[3]);
/* Radial order about (Nx/2, Ny/2, Nz/2) (parameter_order.cpp:223-245). */
void orc_radial_order(const orc_params *p, const float *ni, const int *pt, int cylinder, double *Sr, double *Pr);
/* Local order parameter of snapshots (parameter_order.cpp:75-127). */
double orc_local_order(const orc_params *p, const float *ni, const int *pt, int i, int j, int k);
/* Pear observables (parameter_order.cpp:131-215). */
double orc_polarization(int Nt, const float *ni, const int *pt, const double dir[3]);
double orc_splay_magnitude(const orc_params *p, const float *ni, const int *pt);
double orc_splay_divergence(const orc_params *p, const float *ni, const int *pt);

/* ---- mode loop ---- */
/* Thermal mode (evolve_thermal.cpp:144-212): for each T of the sweep, MCT sweeps then MCS sweeps with one
 * measurement each. rows: nT x 6 doubles (T, S, varS, E, varE, Sr_sph); returns nT. ang_var starts at 0.5. */
int orc_thermal_run(orc_params *p, float *ni, const int *pt, const float *ns, double Ti, double Tf, double dT,
                    int MCT, int MCS, uint64_t seed, int nthreads, double *rows, int max_rows);
/* Electric mode (evolve_electric.cpp:31-103): field magnitude swept Ei -> Ef at the fixed T = p->T. rows: nE x 6
 * doubles (field, S, varS, E, varE, NaN); returns nE. ang_var starts at 0.5. */
int orc_electric_run(orc_params *p, float *ni, const int *pt, const float *ns, double Ei, double Ef, double dE,
                     int MCT, int MCS, uint64_t seed, int nthreads, double *rows, int max_rows);
/* Timed CPU baseline: nsweeps sweeps; returns wall seconds. */
double orc_time_sweeps(const orc_params *p, float *ni, const int *pt, const float *ns, int nsweeps, uint64_t seed,
                       int nthreads, float *ang_var);

/* ---- sweep schedule (include/sweep.h) ---- */
int orc_sweep_points(double a, double b, double d);
double orc_sweep_value(double a, double b, double d, int n, int npoints);
int orc_snapshot_scale(double step);
long orc_snapshot_index(double x, int S);

/* ---- Philox4x32-10 (Salmon et al., SC'11) for the RNG known-answer test of the GPU generator ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
