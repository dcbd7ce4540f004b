/*
 * libmclgpu -- MI355X (gfx950) checkerboard Metropolis engine for McLics-style nematic lattices. C ABI.
 *
 * The library replaces the GPU sweep engine of McLics behind its mode loops: the EvolveNGPU class
 * (include/evolve.cuh:25-28 of the reference) with Monte_Carlo_Step_GPU (src/evolve.cu:317-440),
 * energy_calculator_GPU (src/evolve.cu:442-478), the host Q_tensor / Eigen_value_evaluation calls of the mode
 * loops (src/evolve_thermal.cu:85-96, src/parameter_order.cpp:19-71) and the device physics of src/potential.cu,
 * src/geometry.cu. The mode loops (thermalGPU, stepGPU, quenchGPU, electricGPU) stay in the host program and call
 * the entry points below.
 *
 * Conventions (the reference's, include/define.h:9-19):
 *   - site (i, j, k) has the linear index idx = i + nx * (j + ny * k) (x fastest);
 *   - directors are passed as AoS float[3 * nx * ny * nz]; the library stores them in its own layout on the device;
 *   - point types: 0 empty, 1 bulk, n >= 2 site of surface n - 2 (int32, as the reference's pt array);
 *   - normals: AoS float[3 * Nt], the outward unit normal on surface sites (Geometry::ns, include/geometry.h:274).
 * Ownership: the caller owns every host buffer; the context owns all device memory (released by mcl_destroy).
 * Errors: every call returns MCL_OK or an error code; mcl_last_error() gives the message of the calling thread's
 * last failure. The library never calls exit(); the host maps MCL_ERR_DEVICE to the reference's exit(20)
 * (src/evolve.cu:325-475).
 * Threading: a context is used by one host thread at a time; several contexts may coexist in a process.
 *
 * A context holds one or more independent lattices ("replicas") that are swept together in the same kernel
 * launches; a single simulation is a context with one lattice.
 */
#ifndef MCLGPU_H_
#define MCLGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCL_ABI_VERSION 1

enum mcl_status {
  MCL_OK = 0,
  MCL_ERR_ARG = 1,      /* invalid argument (sizes, odd periodic axis, unknown type, ...) */
  MCL_ERR_NOMEM = 2,    /* device or host allocation failed */
  MCL_ERR_DEVICE = 20   /* HIP runtime or kernel error (the reference exits with status 20) */
};

/* Pair potentials: Bulk_Energy_Lebwohl_Lasher / _GHRL / _Selinger_Pear (src/potential.cpp:19-56). */
enum mcl_potential { MCL_POT_LL = 0, MCL_POT_GHRL = 1, MCL_POT_PEAR = 2 };

/* Anchoring types, in the order of include/anchoring.h:29-104 (input names in parentheses). */
enum mcl_anchoring {
  MCL_ANCH_HOMEOTROPIC = 0,      /* homeotropic      -W (n.s)^2 */
  MCL_ANCH_RP = 1,               /* rp               -W (n.e)^2, e = axis */
  MCL_ANCH_RP_TANGENT = 2,       /* rp_tangent       -W (n.e)^2, e = axis projected on the tangent plane */
  MCL_ANCH_FG = 3,               /* fg               +W (n.s)^2 */
  MCL_ANCH_HOMEOTROPIC_GHRL = 4, /* homeotropic_ghrl  W U_GHRL(n, s, r = s) */
  MCL_ANCH_RP_GHRL = 5,          /* rp_ghrl           W U_GHRL(n, axis, r = s) */
  MCL_ANCH_FG_GHRL = 6           /* fg_ghrl           W U_GHRL(n, tangent projection of n, r = s) */
};

/* Radial observables (parameter_order.cpp:223-245), selected by the `radial` input (main.cpp:90-96). */
enum mcl_radial { MCL_RADIAL_NONE = 0, MCL_RADIAL_SPHERE = 1, MCL_RADIAL_CYLINDER = 2, MCL_RADIAL_BOTH = 3 };

/* One surface: its anchoring type, strength W and lab easy axis (Anchoring::lab_axis of theta_s, phi_s,
 * anchoring.cpp:19-24; unused by homeotropic, fg, homeotropic_ghrl and fg_ghrl). */
typedef struct mcl_surface {
  int32_t type;
  float W;
  float axis[3];
} mcl_surface;

/* Physics of one lattice: the fields of Parameters (include/parameters.h:39-96) that the energy reads.
 * potential and flip_move must be the same for every lattice of a context. */
typedef struct mcl_model {
  int32_t potential;
  float A, B1, C;                                           /* LL and pear couplings */
  float ghrl_lambda, ghrl_mu, ghrl_nu, ghrl_rho, ghrl_sigma; /* from setGHRL (src/io.cpp:346-371) */
  float elec_dir[3];                                        /* unit field direction (normalized, main.cpp:97-106) */
  float elecA, elecP;                                       /* dielectric and dipolar couplings */
  float elecE;                                              /* field magnitude (1 outside the electric mode) */
  int32_t flip_move;                                        /* n -> -n trial (main.cpp:88) */
} mcl_model;

/* One lattice. */
typedef struct mcl_lattice_desc {
  int32_t nx, ny, nz;
  int32_t periodic[3];       /* 1 periodic, 0 free; periodic axes must have even sizes (io.cpp:462-475) */
  const float *ni;           /* initial directors, AoS */
  const int32_t *pt;         /* point types, or NULL for an all-bulk lattice */
  const float *ns;           /* outward normals, AoS; may be NULL when nsurf == 0 */
  int32_t nsurf;
  const mcl_surface *surf;   /* nsurf entries (surface n = pt - 2) */
  mcl_model model;
  float T;                   /* temperature */
  float ang_var;             /* initial maximum rotation (0.5 in the reference, evolve_thermal.cu:43) */
  uint64_t seed;             /* RNG key: one independent stream per seed (the same seed gives the same stream at any
                                position of any context, so replicas need distinct seeds) */
  int32_t radial;            /* enum mcl_radial */
} mcl_lattice_desc;

/* Running sums over the accumulated measurements of one lattice (evolve_thermal.cu:84-101 keeps S1, S2, E, E2
 * and the ExtraObservables sums on the host; here they are FP64 sums on the device). */
typedef struct mcl_moments {
  int64_t n;                       /* number of measurements accumulated */
  double S, S2, E, E2;             /* sums of S, S^2, E, E^2 (E = mean site energy) */
  double Sr_sph, Pr_sph, Sr_cyl, Pr_cyl; /* sums of the radial observables */
  double P, dtheta, divn;          /* sums of the pear observables (parameter_order.cpp:131-215) */
} mcl_moments;

/* One measurement of the current state. */
typedef struct mcl_observables {
  int64_t nvalid;                  /* non-empty sites */
  double E;                        /* mean site energy (energy_calculator_GPU) */
  double Q[6];                     /* Q tensor xx, xy, xz, yy, yz, zz (Q_tensor) */
  double S;                        /* largest eigenvalue of Q (Eigen_value_evaluation) */
  double director[3];              /* its eigenvector */
  double Sr_sph, Pr_sph, Sr_cyl, Pr_cyl;
  double P, dtheta, divn;
} mcl_observables;

/* Creates a context with nlat lattices on the current HIP device. Replaces the device setup of the mode loops
 * (evolve_thermal.cu:48-66) and the first-call setup of Monte_Carlo_Step_GPU (evolve.cu:329-406). */
int mcl_create(const mcl_lattice_desc *lat, int32_t nlat, void **ctx_out);
void mcl_destroy(void *ctx);

/* Runs nsweeps sweeps of every lattice (Monte_Carlo_Step_GPU, evolve.cu:317-440: one trial rotation per non-empty
 * site, Metropolis acceptance, optional flip trial, ang_var adaptation after each sweep). When measure_every > 0,
 * a measurement (energy_calculator_GPU + Q_tensor + Eigen_value_evaluation + ExtraObservables) is accumulated into
 * the moments after every measure_every-th sweep. Asynchronous: returns once the work is queued. */
int mcl_sweeps(void *ctx, int32_t nsweeps, int32_t measure_every);

/* Sets the temperature (d_T, evolve_thermal.cu:79-80) or the field magnitude (evolve_electric.cu:306-307) of one
 * lattice, or of all lattices when lat < 0. */
int mcl_set_temperature(void *ctx, int32_t lat, float T);
int mcl_set_field(void *ctx, int32_t lat, float elecE);
int mcl_set_ang_var(void *ctx, int32_t lat, float ang_var);
int mcl_get_ang_var(void *ctx, int32_t lat, float *ang_var);
/* Accepted rotations of the last completed sweep and the number of completed sweeps (synchronizes). */
int mcl_get_sweep_stats(void *ctx, int32_t lat, int64_t *last_accepted, int64_t *sweeps_done);

/* Moments: reset (S1 = S2 = E = E2 = 0, evolve_thermal.cu:84) and read (synchronizes). */
int mcl_reset_moments(void *ctx);
int mcl_get_moments(void *ctx, int32_t lat, mcl_moments *out);
/* Measures the current state now without accumulating it (synchronizes). */
int mcl_measure(void *ctx, int32_t lat, mcl_observables *out);
/* *active = 1 when the measurements of mcl_sweeps are folded into the colour-1 sweep pass (no reference
 * counterpart: the reference measures after the sweep, evolve_thermal.cu:87-92). Every context but those measuring
 * the pear observables, GHRL and pear contexts with empty sites, lattices one site wide in x, and lattices measuring
 * both radial orders; the others measure with separate kernels after the sweep. The moments agree either way (E to
 * FP32 rounding, S and the radial order to FP64 rounding). */
int mcl_measure_in_sweep(void *ctx, int32_t *active);

/* Director transfer in the AoS host layout (the cudaMemcpy of d_ni, evolve_thermal.cu:90 and :99). */
int mcl_get_directors(void *ctx, int32_t lat, float *ni);
int mcl_set_directors(void *ctx, int32_t lat, const float *ni);
/* Local order parameter of every site for the snapshots (lattice_order_parameter, parameter_order.cpp:75-127):
 * out[idx] = largest eigenvalue of Q over the site and its non-empty neighbours; NaN on empty sites. */
int mcl_local_order(void *ctx, int32_t lat, float *out);

int mcl_synchronize(void *ctx);
const char *mcl_last_error(void);
int mcl_abi_version(void);

/* ---- verification entry points (no reference counterpart: they expose the device functions to parity tests) ---- */

/* Site energies computed by the sweep kernel's own device energy function: for each k, e_cur[k] is the energy of
 * site sites[k] with its current director and e_trial[k] with director trial[3k..3k+2] (the quantities compared
 * in MC_GPU, evolve.cu:205-209). */
int mcl_site_energies(void *ctx, int32_t lat, const int32_t *sites, const float *trial, int32_t n, float *e_cur,
                      float *e_trial);
/* No counterpart in the reference. mcl_site_energies without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, tight loops of proposals, harnesses that forbid hipMalloc. The staging
 * of at most MCL_SCRATCH_SITES sites per call (24 B each, 1.5 MiB) lives in one static device array per GPU, held
 * until the process ends; larger n is rejected with MCL_ERR_ARG. Calls that overlap on one GPU share that array and
 * race. The results equal mcl_site_energies' bit for bit. */
#define MCL_SCRATCH_SITES 65536
int mcl_site_energies_noalloc(void *ctx, int32_t lat, const int32_t *sites, const float *trial, int32_t n,
                              float *e_cur, float *e_trial);
/* Philox4x32-10 of the device generator: out[4k..4k+3] = philox(ctr[4k..4k+3], key[2k..2k+1]). */
int mcl_philox(const uint32_t *ctr, const uint32_t *key, int32_t n, uint32_t *out);

/* ---- timing (used by the benchmark) ---- */
/* When stride > 0, mcl_sweeps records a HIP event pair around each run of consecutive sweep-kernel launches (at
 * most stride sweeps per run; a measurement ends a run), using events created up front for max_timed_sweeps
 * sweeps. With graph replay on (mcl_set_graph), the pair goes around each replay instead, and the replay's duration
 * counts for its 2 * len sweep-kernel launches (including the measurement work folded into them and the replay's
 * final finalize). mcl_kernel_timing returns the number of timed sweep-kernel launches and their summed duration
 * (synchronizes); with lattice groups (mcl_set_lattice_groups) a launch over one group counts as its share of the
 * context's colour-block positions, so the count is in whole-context colour passes. stride = 0 switches timing off. */
int mcl_set_kernel_timing(void *ctx, int32_t stride, int32_t max_timed_sweeps);
int mcl_kernel_timing(void *ctx, int64_t *launches, double *total_ms);
/* STREAM-style copy of nbytes (rounded down to 16 B) between two fresh device buffers, reps times on the null
 * stream; returns the bandwidth counting bytes read plus written, and the time per copy. It confirms the HBM peak
 * and calibrates the FETCH_SIZE / WRITE_SIZE counters with a known byte count. */
int mcl_copy_bandwidth(int64_t nbytes, int32_t reps, double *gbs, double *ms_per_copy);
/* ---- launch shape, graphs and the autotuner (no reference counterpart: the reference launches 8 kernels of a
 * fixed shape per sweep, evolve.cu:334-339 and 416-422) ---- */

/* Result of the autotuner: threads per workgroup, tile order (1: consecutive tiles on the same XCD), measured
 * time per sweep of the context's lattices, and whether it came from the cache. */
typedef struct mcl_tune_result {
  int32_t threads;
  int32_t xcd_order;
  double us_per_sweep;
  int32_t from_cache;
} mcl_tune_result;

/* Times every candidate launch shape (64..1024 threads per workgroup x tile order) over trial_sweeps sweeps of
 * the context's own lattices and applies the fastest. The simulation state (directors, ang_var, counters, sweep
 * number) is saved before and restored after the trials. Results are cached per shape signature in the process
 * and, when the environment variable MCL_TUNE_CACHE names a file, in that file. */
int mcl_autotune(void *ctx, int32_t trial_sweeps, mcl_tune_result *out);
/* As mcl_autotune, with the trials measured every measure_every sweeps (0: never), the cadence of the run that
 * follows: the measuring passes prefer other shapes than the plain ones (they hold more registers). The moments are
 * saved and restored as well; results are cached per shape signature and cadence. */
int mcl_autotune_ex(void *ctx, int32_t trial_sweeps, int32_t measure_every, mcl_tune_result *out);
/* Sets the launch shape by hand (threads a multiple of 64, at most 1024). */
int mcl_set_launch(void *ctx, int32_t threads, int32_t xcd_order);
/* block_sweeps > 0: mcl_sweeps replays captured HIP graphs instead of launching each kernel: graphs of about
 * block_sweeps sweeps (a multiple of measure_every) while the call has that many left, then one graph of the largest
 * multiple of measure_every that is left (the whole rest when nothing is measured), so calls of any length that is a
 * multiple of measure_every run from graphs; fewer than measure_every remaining sweeps run eagerly. Each context
 * keeps the 8 graphs it replayed last. 0 switches graphs off. Results are identical to the eager launches. */
int mcl_set_graph(void *ctx, int32_t block_sweeps);
/* Captures (without running anything) the graphs that mcl_sweeps(ctx, nsweeps, measure_every) would replay under
 * the current launch shape and graph setting, so that the capture does not fall inside a timed or latency-bound
 * call. No reference counterpart. Does nothing when graphs are off or the call would run resident sweeps. */
int mcl_prepare_sweeps(void *ctx, int32_t nsweeps, int32_t measure_every);
/* Resident sweeps (no reference counterpart; for the small lattices of campaigns such as the reference's paper runs,
 * whose sweeps are otherwise bound by the latency of two launches each). When every lattice of the context has at
 * most 4096 colour-block positions, 2 ceil(nx/2) ny nz (2048 with the flip trial: 16^3, 10 x 10 x 16 and 6 x 6 x 48
 * fit), mcl_sweeps with measure_every = 0 runs whole sweeps with each lattice held in the LDS of one workgroup, many
 * sweeps per launch (one position per thread up to 1024 per colour). The trajectories are bit-identical to the
 * two-pass path. On by default; enable = 0 switches it off. */
int mcl_set_resident(void *ctx, int32_t enable);
/* *active = 1 when plain sweeps of the context run resident (the context qualifies and it is switched on). */
int mcl_resident(void *ctx, int32_t *active);
/* Lattice groups (no reference counterpart; for batches larger than the GPU's 256 MB Infinity Cache). When the
 * directors and point types of the context's lattices, both colours, exceed max_bytes, mcl_sweeps splits the
 * lattices into groups of consecutive lattices of about equal size, each within max_bytes, and runs every sweep of
 * the call for one group before the next, so that each group's colour passes are served from that cache instead of
 * HBM. Results are identical to the ungrouped order (a lattice's trajectory depends only on its own sites and sweep
 * numbers); measurements, graphs and kernel timing cover every group. Default 192 MiB; max_bytes <= 0 keeps one
 * group. Drops the context's captured graphs. */
int mcl_set_lattice_groups(void *ctx, int64_t max_bytes);
/* *ngroups = the number of lattice groups mcl_sweeps runs in turn (1 when the context fits). */
int mcl_get_lattice_groups(void *ctx, int32_t *ngroups);

/* Device stream of the context (a hipStream_t), for callers that record their own events. */
int mcl_get_stream(void *ctx, void **stream);

/* ---- the learned surrogate (no reference counterpart) ----
 * An MLP 5 -> 64 SiLU -> 64 SiLU -> 3 in FP32, trained and evaluated by HIP kernels. Inputs are (T, A, W, field,
 * log N) and outputs (S, E, chi) of the plain Metropolis method, both normalized by the caller. Parameter layout:
 * W1[64][5], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]. It only chooses where the plain method samples and
 * warm-starts directors; it produces no reported number. */
int mcl_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, void **mlp);
void mcl_mlp_destroy(void *mlp);
int mcl_mlp_num_params(void);
int mcl_mlp_get_params(void *mlp, float *params, int32_t n);
/* Sets the parameters and resets the Adam moments. */
int mcl_mlp_set_params(void *mlp, const float *params, int32_t n);
/* y[n][3] = f(x[n][5]) (host arrays). */
int mcl_mlp_forward(void *mlp, const float *x, int32_t n, float *y);
/* Gradient of L = (1/n) sum_s sum_o w_o (f(x_s)_o - t_so)^2 with respect to the parameters, and L. w_out may be
 * NULL (all ones). */
int mcl_mlp_gradient(void *mlp, const float *x, const float *t, const float *w_out, int32_t n, float *grad,
                     float *loss);
/* Uploads the training set and runs `epochs` epochs of mini-batch Adam (beta 0.9 / 0.999, eps 1e-8) with
 * learning rate lr, shuffling with a generator keyed by seed; returns the mean loss of the last epoch. */
int mcl_mlp_set_data(void *mlp, const float *x, const float *t, int32_t n);
int mcl_mlp_train(void *mlp, int32_t epochs, int32_t batch, float lr, const float *w_out, uint64_t seed,
                  float *last_loss);
const char *mcl_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
