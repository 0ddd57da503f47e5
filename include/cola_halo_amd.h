/*
 * cola_halo_amd -- MI355X-native 2LPT / particle-mesh / COLA time stepping.
 *
 * C ABI of the HIP library libcola_halo_amd.so.  Plain pointers and sizes
 * only.  The library owns all device state; host buffers cross the boundary
 * only where the reference hands host arrays around (snapshot output,
 * particle state for checkpoints and tests).
 *
 * Each entry point names the reference interface it replaces
 * (junkoda/cola_halo, file:line).  The reference reports fatal errors with
 * msg_abort(code) -> MPI_Abort (msg.c:36-45); here every entry point returns
 * that same numeric code instead (0 = success) and cola_last_error() gives
 * the message.  One calling thread per context; a context is not re-entrant.
 */
#ifndef COLA_HALO_AMD_H
#define COLA_HALO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* particle.h:6-12 -- 56-byte particle record (AoS, host side only) */
typedef struct {
  float x[3];
  float dx1[3]; /* ZA displacement extrapolated to a=1 */
  float dx2[3]; /* 2LPT displacement extrapolated to a=1 */
  float v[3];   /* COLA-frame velocity */
  long long id;
} cola_particle;

/* particle.h:24-28 -- 32-byte snapshot record */
typedef struct {
  float x[3];
  float v[3]; /* km/s */
  long long id;
} cola_particle_min;

typedef struct cola_context cola_context;

/* Status codes reused from the reference's msg_abort codes. */
enum {
  COLA_OK = 0,
  COLA_ERR_ALLOC = 4001,        /* pm.c:129-131 */
  COLA_ERR_POWER_FILE = 3000,   /* power.c:68-71 */
  COLA_ERR_SIGMA8 = 3010,       /* power.c:111-113 */
  COLA_ERR_MASS = 6040,         /* pm.c:551-553 */
  COLA_ERR_ARG = 9001,          /* invalid argument (no reference counterpart) */
  COLA_ERR_UNSUPPORTED = 9002,  /* grid size with a prime factor > 7 */
  COLA_ERR_HIP = 9003,          /* HIP runtime error */
  COLA_ERR_STATE = 9004,        /* call out of order (e.g. force before ICs) */
  COLA_ERR_NET_FILE = 9005      /* unreadable or malformed net-weights file */
};

/* Last error message of the calling thread ("" if none). */
const char* cola_last_error(void);

/* Context: replaces comm_init (comm.h:8, called main.c:65), pm_init
   (pm.h:6-9, main.c:100), lpt_init (lpt.h:6, main.c:86),
   allocate_particles / allocate_shared_memory (mem.h:12-15, main.c:85-92).
   nc particles per dimension, ngrid PM cells per dimension (ngrid/nc need not
   be an integer: mass per particle is (ngrid/nc)^3, pm.c:216 generalised). */
int cola_context_create(int nc, int ngrid, double boxsize, cola_context** out);
int cola_context_destroy(cola_context* ctx);
int cola_synchronize(cola_context* ctx);

/* power.c:36-60 power_init: reads a (k, P(k)) table, checks sigma8 within 5%
   (returns 3010 beyond) and keeps Norm = 1. */
int cola_power_init(cola_context* ctx, const char* filename, double a_init, double sigma8,
                    double omega_m, double omega_lambda);
/* Same, from arrays already in memory. */
int cola_power_init_table(cola_context* ctx, const double* k, const double* pk, int n,
                          double a_init, double sigma8, double omega_m, double omega_lambda);
/* power.c:120-151 PowerSpec */
double cola_power_spec(const cola_context* ctx, double k);
/* power.c:199-202 GrowthFactor */
double cola_growth_factor(const cola_context* ctx, double astart, double aend);
/* sigma8 of the loaded table by power.c:103-118's integral */
double cola_power_sigma8(const cola_context* ctx);

/* Host-only power-table helpers (no device needed): sigma8 of a (k, P)
   table by power.c:103-118's integral; PowerSpec (power.c:120-151) of that
   table at nq wave numbers; the same read from a file (3000 if unreadable). */
int cola_power_table_sigma8(const double* k, const double* pk, int n, double* sigma8);
int cola_power_table_eval(const double* k, const double* pk, int n, const double* kq, int nq, double* out);
int cola_power_file_sigma8(const char* filename, double* sigma8);

/* Linear P(k) table for the inputs of the configurations: Eisenstein & Hu
   (1998, ApJ 496, 605) no-wiggle transfer function (eqs. 26, 28-31, T_cmb =
   2.728 K) times k^ns, tabulated at n log-spaced k in [kmin, kmax] h/Mpc and
   normalised so that sigma8 evaluated by power.c's own integral
   (power.c:103-118, 204-224) equals sigma8.  No reference counterpart (the
   reference reads a CAMB table, power.c:62-101). */
int cola_eh98_nowiggle_table(double omega_m, double omega_b, double h, double ns, double sigma8,
                             double kmin, double kmax, int n, double* k_out, double* p_out);

/* lpt.c:172-207: the seed table of the white-noise columns (nmesh^2 entries). */
int cola_lpt_seedtable(int nmesh, int seed, unsigned int* table);
/* The first n uniforms of the RANLUX stream seeded with seed (luxury 202,
   GSL ranlxd1 semantics as lpt.c:172-233 uses it). */
int cola_ranlxd1_uniforms(unsigned long seed, int n, double* out);
/* cola.c scalars at a, Om = (float)omega_m: out[0] growthD, out[1] growthD2,
   out[2] growthD2v, out[3] Qfactor, out[4] DprimeQ(a, 1) (cola.c:156-234). */
int cola_growth_scalars(float omega_m, float a, double* out);
/* Per-step factors of cola_kick / cola_drift (cola.c:77-91, 123-134):
   out[0] dda, out[1] q1, out[2] q2, out[3] dyyy, out[4] da1, out[5] da2;
   the drift starts from a_v = avel1. */
int cola_step_factors(float omega_m, float a_x, float a_v, float avel1, float apos1, double* out);

/* lpt.c:145-578 lpt_set_displacement (lpt.h:7, main.c:127): 2LPT initial
   conditions for seed on the device; sets a_x = a_init, a_v = 0. */
int cola_lpt_set_displacement(cola_context* ctx, double a_init, double omega_m, int seed, double boxsize);
/* lpt.c:172-351 (the cdisp fill of lpt_set_displacement, lpt.h:7): the
   white-noise ZA displacement spectra of seed as the device builds them,
   copied into out as 3 complex nc x nc x (nc/2+1) grids ([axis][i][j][kz],
   re/im interleaved; n_out >= 6 nc^2 (nc/2+1) floats).  Needs
   cola_power_init; leaves the particles untouched. */
int cola_lpt_modes(cola_context* ctx, int seed, double boxsize, float* out, size_t n_out);
/* lpt.c:611-614 lpt_get_local_nx (one rank: nc) */
int cola_lpt_get_local_nx(const cola_context* ctx);

/* move.c:15-98 move_particles2 (move.h:7, main.c:167): on one rank only the
   periodic wrap into [0, L) remains. */
int cola_move_particles(cola_context* ctx);

/* pm.c:565-596 pm_calculate_forces (pm.h:10, main.c:171): CIC deposit with
   the reference's ghost-particle x periodicity, mass check (6040), r2c FFT,
   Poisson gradient per axis, c2r FFT, CIC readout into the force array. */
int cola_pm_calculate_forces(cola_context* ctx);

/* cola.c:73-117 cola_kick (cola.h:7, main.c:195) and cola.c:119-153
   cola_drift (cola.h:9, main.c:205); a_x / a_v advance as in the reference. */
int cola_kick(cola_context* ctx, float omega_m, float avel1);
int cola_drift(cola_context* ctx, float omega_m, float apos1);

/* One full time step with no snapshot in it: wrap, force, kick, drift
   (main.c:162-206 with the output checks false).  The readout, kick and drift
   run as one fused kernel; results equal the separate calls bit for bit.
   No single reference counterpart. */
int cola_step(cola_context* ctx, float omega_m, float avel1, float apos1);

/* Graph mode for cola_step (no reference counterpart): the step's kernels are
   captured into a HIP graph on the first call and replayed afterwards; the
   per-step values (kick / drift factors, mass-check slot) reach the kernels
   through one small host-to-device copy before each replay.  Results equal
   the eager path bit for bit.  Changing tiles or FFT shapes drops the graph. */
int cola_set_step_graph(cola_context* ctx, int on);

/* Batched realizations (no reference counterpart): one cola_step on each of n
   contexts (sizes may differ), issued back to back on their own streams so
   the GPU overlaps the realizations' kernels; per-context avel1 / apos1.
   Equals n separate cola_step calls bit for bit. */
int cola_step_batch(cola_context* const* ctxs, int n, float omega_m, const float* avel1, const float* apos1);

/* Ensembles (the reference runs its realizations one after another,
   main.c:117-121 with seed = random_seed + i; no counterpart for running
   them together).  An ensemble steps n contexts of any sizes with one
   kernel launch per phase over the workgroups of every realization: mass
   slot, wrap + tile counts, scan (3), scatter, deposit, grid sums, z r2c,
   y forward, fused x, y inverse (3 axes), z c2r (3 axes), readout + kick +
   drift.  The transform passes launch once per FFT kernel variant the
   members' own launches use (register tile and radix set: one per pass when
   all members share one); cola_ensemble_set_merge(e, 1) makes every pass
   one launch with the widest variant (14 launches per step whatever the
   sizes; slower where the variants differ).  Each realization keeps its own
   tiles, radix schedule and launch shapes, and the particles equal those of
   cola_step on each context bit for bit.  The contexts need their
   particles set, no residual net attached, and an x pass that fuses (every
   grid of 2^a 3^b 5^c with a last radix of 3, 4, 5 or 8 does).  An ensemble
   step orders after the work queued on the member contexts and before
   their later work; changing a member's tiles or FFT shapes rebuilds the
   ensemble's tables at its next step.  A context may be a member once
   (9001 otherwise).  Destroying a member context drops the ensemble's
   tables; its later steps return 9004, and it can still be destroyed. */
typedef struct cola_ensemble cola_ensemble;
int cola_ensemble_create(cola_context* const* ctxs, int n, cola_ensemble** out);
/* One step of every member (cola_step's meaning, per-member avel1 / apos1). */
int cola_ensemble_step(cola_ensemble* e, float omega_m, const float* avel1, const float* apos1);
/* Graph mode: the 14 launches are captured once and replayed; per-step values
   reach the kernels through one host-to-device copy before each replay. */
int cola_ensemble_set_graph(cola_ensemble* e, int on);
/* on = 1: one launch per transform pass for all members; 0 (default): one
   per pass and kernel variant.  Takes effect at the next step. */
int cola_ensemble_set_merge(cola_ensemble* e, int on);
/* Kernel launches per ensemble step (0 before the first step). */
int cola_ensemble_launches(const cola_ensemble* e);
int cola_ensemble_destroy(cola_ensemble* e);

/* cola.c:341-416 cola_set_snapshot (cola.h:18, main.c:346): kick/drift to
   aout, add the 2LPT velocity, km/s; writes np records in lattice-ID order
   into the caller's host array. */
int cola_set_snapshot(cola_context* ctx, double aout, float omega_m, cola_particle_min* out, size_t np_out);
/* cola.c:297-338 set_noncola_initial (cola.h:16, main.c:135) */
int cola_set_noncola_initial(cola_context* ctx, float omega_m, cola_particle_min* out, size_t np_out);

/* State access (Particles in particle.h:14-22 is caller-visible in the
   reference; here it lives on the device). */
long long cola_np(const cola_context* ctx);
int cola_get_times(const cola_context* ctx, float* a_x, float* a_v);
/* Particles.a_x / a_v (particle.h:17), set directly by the driver
   (main.c:145 sets a_v = da after the initial conditions). */
int cola_set_times(cola_context* ctx, float a_x, float a_v);
int cola_get_particles(cola_context* ctx, cola_particle* out, size_t np_out);
int cola_set_particles(cola_context* ctx, const cola_particle* in, size_t np_in, float a_x, float a_v);
int cola_get_forces(cola_context* ctx, float* out, size_t np_out);          /* np x 3 */
/* pm.c:209-279 PtoMesh + pm.c:535-562 check_total_density only: leaves the
   real-space density -1 + rho on the device until the next force call. */
int cola_pm_assign(cola_context* ctx);
int cola_get_density(cola_context* ctx, float* out, size_t n_out);          /* padded real grid */
int cola_get_force_grid(cola_context* ctx, int axis, float* out, size_t n_out);
double cola_last_mass_sum(const cola_context* ctx);                          /* pm.c:535-562 */

/* Device-side copy of the particle state and clocks, and its restore: a
   resume point between steps that never leaves HBM.  No reference
   counterpart (the reference has no checkpoint; its GADGET output lacks dx1,
   dx2 and the COLA-frame velocity). */
int cola_state_save(cola_context* ctx);
int cola_state_restore(cola_context* ctx);

/* Device-time profiling with HIP events on the context's stream, replacing
   the reference's wall-clock phase timers (timer.c:11-84).  Phases are
   finer than the reference's categories: one per kernel family, named by
   cola_phase_name (e.g. "deposit", "fft_x_inv", "readout_kick_drift").
   cola_get_phase_stats fills up to n entries of accumulated milliseconds and
   launch counts since the last reset; it synchronises the stream. */
int cola_set_profiling(cola_context* ctx, int on);
int cola_phase_count(void);
const char* cola_phase_name(int phase);
int cola_get_phase_stats(cola_context* ctx, double* ms, long long* launches, int n);
int cola_reset_phase_stats(cola_context* ctx);

/* ---- Output grids ---------------------------------------------------------
   Positions are read where they already are on the device: the particle
   state (COLA_SOURCE_PARTICLES) or the records of the last cola_set_snapshot
   / cola_set_noncola_initial (COLA_SOURCE_SNAPSHOT, what the reference's
   snapshot_time outputs use, main.c:346-382). */
#define COLA_SOURCE_PARTICLES 0
#define COLA_SOURCE_SNAPSHOT 1

/* Replaces coarse_grid2 (coarse_grid.c:17-93 with cic_mass_assignment,
   coarse_grid.c:95-121) up to the file write: CIC density with periodic wrap
   and no WPAR on an n^3 grid, converted to delta = rho * n^3 / np - 1 (float),
   returned in C order [ix][iy][iz] in delta_out (n^3 floats).  sums_out, if
   not NULL, receives the sum of rho and the sum of delta (coarse_grid.c:64-66).
   Returns COLA_ERR_STATE where the reference asserts
   abs(sum - np) < 0.1 sum (coarse_grid.c:84).  The device summation order
   differs from the reference's serial loop: cells agree to float rounding. */
int cola_coarse_grid(cola_context* ctx, int n, int source, float* delta_out, double* sums_out);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): cola_coarse_grid's kernels with the CIC accumulators in
   a static device array of the library (16 MiB per GPU, held until the
   process ends), so the call allocates nothing.  n <= 128 (COLA_ERR_ARG
   beyond).  delta_dev is DEVICE memory for n^3 floats (C order, as
   cola_coarse_grid's delta_out); sums_dev is NULL or DEVICE memory for the
   two sums.  Work is only enqueued on the context's stream: the call does
   not synchronize, and the mass check of cola_coarse_grid (coarse_grid.c:84)
   is the caller's, from sums_dev after cola_synchronize.  Every call on one
   GPU shares the one scratch array: calls that overlap on a GPU (contexts
   on different streams, host threads) race on it. */
int cola_coarse_grid_noalloc(cola_context* ctx, int n, int source, float* delta_dev, double* sums_dev);

/* Shell-averaged power spectrum of the particle distribution (no reference
   counterpart; the CIC assignment is cola_coarse_grid's): delta on an n^3
   grid (n even, 2^a 3^b 5^c), r2c FFT, P(k) = L^3 abs(delta_k / n^3)^2 over
   all modes but k = 0, averaged in nbins linear bins of abs(k) on
   [kmin, kmax) (h/Mpc).  Per bin: mean abs(k), mean P ((Mpc/h)^3) and the
   number of modes (0 for an empty bin, whose k and P are 0).  flags:
   COLA_PK_DECONVOLVE_CIC divides each mode by the CIC window squared,
   COLA_PK_SUBTRACT_SHOT subtracts L^3 / np from each bin. */
#define COLA_PK_DECONVOLVE_CIC 1
#define COLA_PK_SUBTRACT_SHOT 2
int cola_power_spectrum(cola_context* ctx, int n, int source, double kmin, double kmax, int nbins, int flags,
                        double* k_out, double* p_out, double* nmodes_out);

/* ---- Friends-of-friends halos -------------------------------------------
   One halo of the FoF catalogue: the fields of one line of fof.c:1166-1168
   ("%d %e %e %e %e %e %e": nfof, centre, mean velocity in km/s). */
typedef struct {
  int nfof;
  float x[3];
  float v[3];
} cola_halo;

/* Replaces fof_find_halos (fof.c:1035-1084) and the gathering half of
   fof_write_halos (fof.c:1086-1164) on one rank, on the device copy of the
   last snapshot: the records' positions are first wrapped into [0, L)
   (move.c:33-41, the move_particles2_min of fof.c:1044), then every pair
   with squared distance below ll^2 under the periodic minimum image is
   linked (fof.c:586-587), and groups of at least nfof_min particles
   (fof.c:224 uses 32) become halos: centre = x0 + mean of the folded offsets
   from x0 wrapped into [0, L), velocity = mean record velocity
   (fof.c:1152-1163).  Halos are ordered by the smallest record index of
   their group and x0 is that record's position (the reference lists halos
   in tree-walk order and takes x0 from the walk's first particle).  The
   driver passes ll = (float)(linking_factor * boxsize / nc) (main.c:353). */
int cola_fof_find_halos(cola_context* ctx, float ll, int nfof_min, long long* nhalo_out);
int cola_fof_get_halos(cola_context* ctx, cola_halo* out, long long n);
/* Per snapshot record: the smallest record index of its group when the group
   is a halo, else -1 (no reference counterpart; for tests). */
int cola_fof_get_labels(cola_context* ctx, int* out, size_t n);
/* The device snapshot records as they are now (after cola_fof_find_halos,
   with positions wrapped, as the reference writes them after FoF). */
int cola_get_snapshot(cola_context* ctx, cola_particle_min* out, size_t np_out);

/* fof.c:1147-1175: the halo text file, one "%d %e %e %e %e %e %e" line per
   halo.  Returns 8110 when the file cannot be opened. */
int cola_write_halos(const char* filename, const cola_halo* halos, long long n);

/* ---- Learned force residual (no reference counterpart) --------------------
   A 4-layer periodic 3-D conv net on the coarse PM grid (8 -> 16 -> 16 -> 16
   -> 3 channels, 3^3 kernels, relu between layers), input asinh(delta) *
   in_scale and the scale factor a, output the force residual between the
   coarse grid and a grid twice as fine, in units of out_scale.  bfloat16
   weights and activations on MFMA with fp32 accumulation; fp32 master
   weights and Adam.  Parameters (cola_resnet_nparam floats) are W1 b1 W2 b2
   W3 b3 W4 b4 with W[27][cin][cout], tap t = 9 (dx+1) + 3 (dy+1) + (dz+1). */
typedef struct cola_resnet cola_resnet;
int cola_resnet_nparam(void);
int cola_resnet_create(int n, unsigned int seed, cola_resnet** out); /* n: multiple of 32 */
int cola_resnet_destroy(cola_resnet* net);
int cola_resnet_get_params(cola_resnet* net, float* out);
int cola_resnet_set_params(cola_resnet* net, const float* in);
int cola_resnet_get_scales(cola_resnet* net, float* in_scale, float* out_scale);
int cola_resnet_set_scales(cola_resnet* net, float in_scale, float out_scale);
/* out_scale = RMS of the residual targets of stored samples [0, count), so
   that predicting no correction has loss 1 (training fits it if unset). */
int cola_resnet_fit_scale(cola_resnet* net, int count);
/* Host arrays (tests): x [n^3][8] channel-last input (rounded to bfloat16),
   y [n^3][3] output; the gradient entry point also returns the mean squared
   error against target [n^3][3] (in output units) and its gradient with
   respect to every parameter. */
int cola_resnet_forward_host(cola_resnet* net, const float* x, float* y);
int cola_resnet_gradient_host(cola_resnet* net, const float* x, const float* target, double* loss, float* grad);
/* One Adam step (Kingma & Ba 2015: beta1 0.9, beta2 0.999, eps 1e-8, bias
   corrected by 1 - beta^t) on the gradient of the last
   cola_resnet_gradient_host call, the update cola_resnet_train makes after
   each sample (no reference counterpart; for tests). */
int cola_resnet_adam_step(cola_resnet* net, float lr);
/* The inference path on the same host input (no reference counterpart):
   channel 0 of x is the feature and channel 1 of the first cell the scale
   factor a (x must hold a in every cell of channel 1 and zeros in channels
   2-7); the first layer runs on the feature channel with a folded into its
   bias.  Equals cola_resnet_forward_host up to fp32 summation order. */
int cola_resnet_forward_features_host(cola_resnet* net, const float* x, float* y);
/* Mean device time (HIP events, `reps` launches) of one inference forward
   pass (four layers from the feature grid of the last host input), in ms (no
   reference counterpart). */
int cola_resnet_time_forward(cola_resnet* net, int reps, double* ms);
/* Device copy of the particle state of src into dst (same particle count). */
int cola_copy_particles(cola_context* dst, cola_context* src);
/* Training pair from two contexts holding the same particles: coarse (the
   net's grid) and fine (twice the cells per axis).  Stores asinh(delta) of
   the coarse density and the residual = fine force restricted to the coarse
   cells minus the coarse force (see cola_resnet_set_restriction), with the scale
   factor a_x of the coarse context. */
int cola_resnet_add_pair(cola_resnet* net, cola_context* coarse, cola_context* fine);
/* How cola_resnet_add_pair restricts the fine force to a coarse cell: the
   average over the 3^3 fine neighbours of the coincident fine cell with
   weights (w, 1 - 2w, w) per axis, 0 <= w <= 1/4 (0: the point value, 1/4:
   full weighting; default 1/8). */
int cola_resnet_set_restriction(cola_resnet* net, float side_weight);
int cola_resnet_num_samples(const cola_resnet* net);
/* Stored sample i as fp32 host arrays (tests): feature [n^3] (bf16 asinh(delta)
   times the input scale), target [3][n^3] (bf16 residual, C order per axis)
   and its scale factor a. */
int cola_resnet_get_sample(cola_resnet* net, int i, float* feature, float* target, float* a);
/* Adam, one step per stored sample, epochs over samples [0, ntrain) in a
   shuffled order; out_scale is set to the targets' RMS on the first call.
   epoch_loss (may be NULL) receives the mean loss per epoch (1 = predicting
   zero). */
int cola_resnet_train(cola_resnet* net, int ntrain, int epochs, float lr, unsigned int seed, double* epoch_loss);
int cola_resnet_eval(cola_resnet* net, int first, int count, double* loss);
/* Attach (net) or detach (NULL) the net: every force of the context then adds
   out_scale * net(delta, a) to the three force grids before the readout.  A
   net serves one context at a time (its activation buffers are its own):
   attaching a net that another context holds returns COLA_ERR_STATE;
   destroying an attached net detaches it first. */
int cola_set_residual_net(cola_context* ctx, cola_resnet* net);
/* Launch shape of the inference path in a context's force (no reference
   counterpart): on = 1 (default) makes the features and the first two layers
   in one pass over the density grid, on = 0 runs them as three launches
   through the feature and h1 grids.  Both give the same forces bit for bit. */
int cola_resnet_set_fusion(cola_resnet* net, int on);
/* Loop order of the layers with 16 input channels in the forward passes (no
   reference counterpart): by_input_planes = 1 (default) streams each input
   plane once and adds its taps to the three output planes it feeds
   (layers 2-3 and the fused first pass: k_conv16r, k_conv12r; the last
   layer with the x offsets as output rows: k_conv_head_r); 0 computes each
   output plane from its three input planes (k_conv16, k_conv12, also the
   training path's data gradient).  The two orders sum the same products in
   another order: results agree to float rounding, not bit for bit. */
int cola_resnet_set_plane_order(cola_resnet* net, int by_input_planes);
/* Launch shape of the last two layers in a context's force (no reference
   counterpart): on = 1 (default; with the fused first pass and input-plane
   order) runs layer 3 and the last layer as one launch after the force
   transforms, h3 kept in LDS; on = 0 writes h3 to memory between two
   launches.  Both give the same forces bit for bit. */
int cola_resnet_set_head_fusion(cola_resnet* net, int on);
/* Net-weights file: the trained parameters and scales, independent of the
   grid size (the net is a convolution over cells).  Little-endian layout:
   char magic[8] "COLANET1", int32 nparam, int32 0, float in_scale,
   float out_scale, float restriction side weight, float 0,
   float params[nparam] (the cola_resnet_get_params order), int32 nparam.
   save returns 9000 when the file cannot be written (as write.c:23-24);
   load returns COLA_ERR_NET_FILE for a missing, short or foreign file. */
int cola_resnet_save(cola_resnet* net, const char* filename);
int cola_resnet_load(cola_resnet* net, const char* filename);

/* ---- Driver-side files (host only) -------------------------------------- */

/* Parameters of the driver, parameters.h:4-49 with fixed-size storage.  A
   string key that is absent has strlen_* = 0; otherwise strlen_* counts the
   terminating NUL, as in the reference (so "" gives 1). */
#define COLA_PARAM_MAX_ZOUT 256
#define COLA_PARAM_MAX_STRING 1024
typedef struct {
  int nc;
  int pm_nc_factor;
  double np_alloc_factor;
  int ntimestep;
  int random_seed;
  int nrealization;
  int loglevel;
  double a_final;
  double boxsize;
  double omega_m, sigma8, h;
  int n_zout;
  double zout[COLA_PARAM_MAX_ZOUT];
  char power_spectrum_filename[COLA_PARAM_MAX_STRING];
  int strlen_power_spectrum_filename;
  char fof_filename[COLA_PARAM_MAX_STRING];
  int strlen_fof_filename;
  char snapshot_filename[COLA_PARAM_MAX_STRING];
  int strlen_snapshot_filename;
  char subsample_filename[COLA_PARAM_MAX_STRING];
  int strlen_subsample_filename;
  char cgrid_filename[COLA_PARAM_MAX_STRING];
  int strlen_cgrid_filename;
  char init_filename[COLA_PARAM_MAX_STRING];
  int strlen_init_filename;
  double fof_linking_factor;
  double subsample_factor;
  int cgrid_nc;
  int write_longid;
  /* Keys of this build with no reference counterpart (files without them
     read as before): pm_ngrid = PM cells per side, overriding
     nc * pm_nc_factor (e.g. 768 for nc = 512, a factor of 1.5; 0 = unset);
     residual_net = a net-weights file (cola_resnet_save) whose learned force
     residual every force evaluation adds. */
  int pm_ngrid;
  char residual_net_filename[COLA_PARAM_MAX_STRING];
  int strlen_residual_net_filename;
} cola_params;

/* read_param.c:190-327 read_parameter_file (one rank; no broadcast): same
   syntax ("name = value", "--" comments, {a, b} arrays, "strings",
   t/T/1/f/F/0 booleans, numbers that must start with a digit), defaults and
   abort codes: 1001 unreadable file, 1002 bad boolean, 1003 bad string or
   array, 1004 bad array element, 1007 unknown name.  Also reads the two
   extension keys pm_ngrid (integer) and residual_net (string). */
int cola_read_parameters(const char* filename, cola_params* out);
/* confirm_param.c:8-131 confirm_parameters: returns its abort code (11001-11008,
   11018-11020, 10021-10023) for a missing or empty required value, else 0;
   for the extension keys 10024 (empty residual_net) and 10025 (pm_ngrid
   negative or odd). */
int cola_confirm_parameters(const cola_params* p);

/* write.c:14-107 write_snapshot for one rank: a GADGET-1 file with 4-byte
   record markers, the 256-byte header (gadget_file.h:4-23; particle mass
   omega_m * 27.7455 * L^3 / np, time a, redshift 1/a - 1, one file),
   positions, velocities times 1/sqrt(a) and 4- or 8-byte IDs.  `filename`
   is the full name (the reference appends ".<rank>").  a, boxsize, omega_m
   and h are the Snapshot struct's floats (particle.h:30-42). Returns 9000
   when the file cannot be opened (write.c:23-24). */
int cola_write_gadget(const char* filename, const cola_particle_min* p, long long np, float a, float boxsize,
                      float omega_m, float h, int use_long_id);

/* read.c:30-168 read_snapshot on one rank: a GADGET-1 snapshot written as
   one file `filename` or as filename.0, filename.1, ... (header num_files;
   find_files, read.c:171-201).  Type-1 particles of every file in file
   order: positions wrapped into [0, L) (read.c:99-102), velocities times
   sqrt(a) back to km/s (read.c:76, 130-132), id 0 (read.c:137).  info gets
   the file count (0 when neither name exists: not an error, as read.c
   returns 0), the particles read and the last header's np_total, a,
   boxsize, omega_m and h.  With out == NULL only the headers are read and
   np_read is the particle count.  Errors: 15000 a file cannot be opened,
   15001 more particles than np_out, 15005 a record marker does not match
   (read.c:16-27). */
typedef struct {
  int nfiles;
  long long np_read;
  long long np_total;
  float a, boxsize, omega_m, h;
} cola_gadget_info;
int cola_read_gadget(const char* filename, cola_particle_min* out, long long np_out, cola_gadget_info* info);

/* The snapshot records the output functions read (COLA_SOURCE_SNAPSHOT,
   cola_fof_find_halos, cola_get_snapshot) set from host records, e.g. a
   snapshot read back from a file (halo_main.c's read_snapshot; no separate
   reference function).  np_in must equal the context's particle count. */
int cola_set_snapshot_records(cola_context* ctx, const cola_particle_min* in, size_t np_in);

/* Random particle subsample (subsample.c).  The generator is a RANLUX
   (ranlxd1) stream seeded with 2 * seed (subsample_init, subsample.c:19-29,
   rank 0) that persists across outputs and realizations. */
typedef struct cola_subsampler cola_subsampler;
int cola_subsampler_create(double factor, unsigned int seed, cola_subsampler** out);
int cola_subsampler_destroy(cola_subsampler* s);
/* subsample.c:115-159 write_random_sabsample + write.c:254-311
   write_particles_binary_mpi on one rank: record i is kept when the next
   uniform draw is < factor (one draw per record, in record order).  File:
   6 floats (boxsize, particle mass omega_m 27.7455 L^3 / np, omega_m, h, a,
   redshift), int n, n x {float x[3], v[3]}, int n.  *nsub_out (may be NULL)
   receives n.  Returns 9000 when the file cannot be written. */
int cola_write_subsample(cola_subsampler* s, const char* filename, const cola_particle_min* p, long long np,
                         float a, float boxsize, float omega_m, float h, long long* nsub_out);

/* coarse_grid.c:46-91 file layout: float boxsize, int n, n^3 float delta in
   C order, int n.  Returns 10021 when the file cannot be opened. */
int cola_write_coarse_grid(const char* filename, float boxsize, int n, const float* delta);

/* Deposit tile shape (cells per tile along x, y, z: powers of two dividing
   ngrid, x side >= 2, product <= 4096) and FFT tile width (log2 of the transforms per
   workgroup, -1 = automatic).  Results do not depend on either; the
   autotuner sweeps them for speed. */
int cola_set_tiling(cola_context* ctx, int tx, int ty, int tz, int fft_log2_tile);
/* Adjacent FFT column tiles dispatched to one XCD (a power of two, 1 to 64). */
int cola_set_fft_xcd_group(cola_context* ctx, int group);
/* Consecutive readout workgroups dispatched to one XCD (1 to 4096; default
   one y band of each lattice slab per XCD: max(8, nc^2 / 2048)). */
int cola_set_readout_group(cola_context* ctx, int group);
/* Launch shape of the PM transforms: log2 of the transforms per workgroup of
   the y/x column passes and of the z row passes (-1 = default), and the XCD
   group.  Results do not depend on them. */
int cola_set_fft_tiles(cola_context* ctx, int log2_cols, int log2_rows, int xcd_group);
/* As cola_set_fft_tiles with separate widths for the y and the x column
   passes (log2_xcols = -1: as log2_ycols). */
int cola_set_fft_tiles_xy(cola_context* ctx, int log2_ycols, int log2_xcols, int log2_rows, int xcd_group);

/* Radix schedule of the PM transforms (the order and grouping of the
   Stockham radix passes): 0 radix 8 first (default), 1 radices 3 and 5
   first, 2 radix 4, 3 the powers of two smallest first, 4 the last
   power-of-two radix moved to the second pass.  Every schedule computes the
   same DFT; the float rounding differs (agreement ~1e-6 of the maximum),
   as between the plans FFTW_MEASURE may pick (pm.c:139-147). */
int cola_set_fft_schedule(cola_context* ctx, int schedule);
/* Host-only: the radices of schedule `schedule` for an n-point transform into
   radices[0 .. max-1]; returns the number of passes, 0 if n or the schedule
   is not supported. */
int cola_fft_radices(int n, int schedule, int* radices, int max);

/* Autotuner (no reference counterpart): times the PM force fields (deposit +
   transforms, `reps` per candidate) of the current particle state for deposit
   tile shapes, FFT column / row tile widths, XCD groups and the radix
   schedule by coordinate descent, keeps the fastest and reports it.  The
   launch shapes do not change any result (the deposit sums in
   order-independent fixed point and every transform computes each column
   with the same arithmetic whatever the tile width); the radix schedule
   changes the transforms' rounding only (see cola_set_fft_schedule), like
   FFTW_MEASURE's choice of plan in the reference (pm.c:139-147).
   Wraps the positions first (move.c:33-41), as every step does. */
typedef struct {
  int tile[3];
  int fft_log2_cols, fft_log2_rows, xcd_group;
  int fft_log2_xcols; /* x column pass width, -1 = as fft_log2_cols (the y pass) */
  double ms_before, ms_after; /* mean PM force-field time of the start and the chosen shape */
  int fft_schedule;    /* radix schedule of the PM transforms (cola_set_fft_schedule) */
  int fft_radices[8];  /* its radix passes, first to last (0 after the last) */
} cola_tuning;
int cola_autotune(cola_context* ctx, int reps, cola_tuning* out);

/* Device copy of `bytes` bytes by a plain streaming kernel, repeated reps
   times; returns the mean copy bandwidth (read + write bytes per second,
   GB/s) in *gbps.  Used to calibrate hardware byte counters and as the
   achievable-bandwidth reference.  No reference counterpart. */
int cola_bandwidth_probe(size_t bytes, int reps, double* gbps);

/* Standalone 3-D FFTs on host arrays with FFTW's conventions (unnormalised;
   c2r = c2c over x and y, then c2r over z reading only the real parts of the
   kz = 0 and kz = n/2 elements), in place on the padded real layout
   [n][n][2(n/2+1)].  These are the transforms pm.c:285,353 and
   lpt.c:403,427,501-502 execute through FFTW. */
int cola_fft_r2c_3d(int n, float* data);
/* Device time of this library's transforms on an n^3 grid resident in HBM
   (mean over reps after one warm-up, HIP events on the stream): ms_out[0]
   r2c, ms_out[1] c2r, ms_out[2] the PM force sequence (r2c, then for each of
   three axes the Poisson gradient and a c2r into its own grid).  For
   comparison with other FFT libraries; no reference counterpart. */
int cola_fft_time(int n, int reps, double* ms_out);
int cola_fft_c2r_3d(int n, float* data);

#ifdef __cplusplus
}
#endif
#endif
