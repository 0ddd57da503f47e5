/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, CPU-only restatement of junkoda/cola_halo's time-stepping hot
 * path (2LPT initial conditions, particle-mesh force, COLA kick/drift) used as
 * the checker for the HIP library in cola_halo_amd/.  Only tests/, the
 * package's smoke() and bench.py's cpu_baseline leg may load it, and only to
 * check or time against; the product never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Third-party
 * arithmetic the reference calls is restated from its published algorithm:
 *   - GSL gsl_rng_ranlxd1 / gsl_rng_uniform (Luescher's 48-bit RANLUX),
 *   - GSL gsl_integration_qag (QUADPACK QAG with the GK41 / GK61 rules),
 *   - GSL gsl_sf_hyperg_2F1 (Gauss hypergeometric series),
 *   - FFTW3f unnormalised r2c / c2r (the discrete Fourier transform).
 */
#ifndef COLA_ORACLE_H
#define COLA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* particle.h:6-12 */
typedef struct {
  float x[3];
  float dx1[3];
  float dx2[3];
  float v[3];
  long long id;
} OrcParticle;

/* particle.h:24-28 */
typedef struct {
  float x[3];
  float v[3];
  long long id;
} OrcParticleMin;

/* ---- RANLUX double (GSL ranlxd) ---- */
typedef struct {
  double x[12];
  double carry;
  int ir, jr, ir_old, pr;
} OrcRanlxd;

void orc_ranlxd_set(OrcRanlxd* r, unsigned long seed, int luxury);
double orc_ranlxd_uniform(OrcRanlxd* r);
void orc_seedtable(int nmesh, int seed, unsigned int* table);

/* ---- power.c ---- */
int orc_power_set_table(const double* k, const double* p, int n);
int orc_power_read_file(const char* filename);
double orc_powerspec(double k);
double orc_tophat_sigma2(double R);
int orc_power_check_sigma8(double sigma8, double* sigma8_input);
double orc_growth_factor(double astart, double aend, double omega_m, double omega_lambda);

/* ---- quadrature / special functions ---- */
double orc_qag(double (*f)(double, void*), void* params, double a, double b,
               double epsabs, double epsrel, int limit, int key, double* abserr);
double orc_hyp2f1(double a, double b, double c, double z);

/* ---- cola.c scalars (Om is cola.c's file-static float) ---- */
float orc_growthDtemp(float a, float Om);
float orc_growthD(float a, float Om);
float orc_growthD2(float a, float Om);
float orc_growthD2v(float a, float Om);
float orc_Qfactor(float a, float Om);
double orc_DprimeQ(double a, float nGrowth, float Om);
double orc_Sq(double ai, double af, double aRef, float Om);
double orc_Sphi(double ai, double af, double aRef, float Om);

/* ---- FFT (unnormalised, FFTW sign conventions) ---- */
void orc_fft_1d(int n, double* re_im, int sign);
void orc_fft_r2c_3d(int n, float* data);   /* padded real in, complex out, in place */
void orc_fft_c2r_3d(int n, float* data);   /* complex in, padded real out, in place */

/* ---- lpt.c ---- */
int orc_lpt_set_displacement(int nc, double a_init, double omega_m, int seed,
                             double box, OrcParticle* particles,
                             float* cdisp_out, float* disp_out, float* disp2_out);
/* lpt.c:172-351 alone: the ZA displacement spectra of seed (3 complex
   nc x nc x (nc/2+1) grids, interleaved floats). */
int orc_lpt_modes(int nc, double box, int seed, float* cdisp_out);
/* lpt.c:358-503 on given ZA displacement spectra (3 complex nc x nc x (nc/2+1)
   grids): real ZA displacement (unnormalised c2r, as lpt.c uses it) and 2LPT
   displacement times 1/nc^3, padded real layout. */
int orc_lpt_from_modes(int nc, double box, const float* cdisp_in, float* disp_out, float* disp2_out);

/* ---- pm.c / move.c ---- */
void orc_wrap(OrcParticle* p, int np, float boxsize);
void orc_pm_deposit(int ngrid, float boxsize, float wpar, const OrcParticle* p,
                    int np, float* density);
double orc_density_sum(int ngrid, const float* density);
/* pm.c:359-412 + ghosts (pm.c:416-532): CIC readout of one padded real grid
   at np particles into out[np]. */
void orc_pm_readout(int ngrid, float boxsize, const OrcParticle* p, int np, const float* grid, float* out);
int orc_pm_calculate_forces(int ngrid, float boxsize, float wpar, long long np_total,
                            const OrcParticle* p, int np, float* force,
                            float* density_out, float* force_grid_out, double* mass_sum);

/* ---- cola.c kernels ---- */
void orc_kick(OrcParticle* p, const float* force, int np, float Om,
              float* a_x, float* a_v, float avel1);
void orc_drift(OrcParticle* p, int np, float Om, float* a_x, float* a_v, float apos1);
void orc_set_snapshot(double aout, const OrcParticle* p, const float* force, int np,
                      float Om, float a_x, float a_v, OrcParticleMin* out);
void orc_set_noncola_initial(const OrcParticle* p, int np, float Om, float a_x,
                             OrcParticleMin* out);

/* ---- coarse_grid.c and the shell power spectrum ---- */
/* coarse_grid.c:95-121: zero grid[n^3] (C order), then add the CIC weights of
   np positions (component a of particle i at pos[i*stride + a]), periodic. */
void orc_coarse_grid(const float* pos, long stride, int np, int n, float boxsize, float* grid);
/* coarse_grid.c:59-70: grid = nbar_inv * grid - 1 with nbar_inv = (float)(n^3 / np);
   sums[0] = sum of the input cells, sums[1] = sum of the output cells. */
void orc_density_contrast(float* grid, int n, long long np_total, double* sums);
/* Shell-averaged power spectrum of a real grid delta[n^3] (C order) in a box
   of side L: P(k) = L^3 |DFT(delta)(k) / n^3|^2 over all n^3 modes except
   k = 0, in nbins linear bins of |k| on [kmin, kmax).  flags bit 0 divides
   each mode by the CIC window squared, bit 1 subtracts L^3 / np_shot from the
   bin mean.  Outputs per bin: mean |k|, mean P, number of modes. */
void orc_power_spectrum(const float* delta, int n, double boxsize, double kmin, double kmax, int nbins,
                        int flags, double np_shot, double* k_out, double* p_out, double* nmodes_out);

/* ---- friends-of-friends (fof.c on one rank) ---- */
typedef struct {
  int nfof;
  float x[3];
  float v[3];
} OrcHalo;
/* Wraps rec positions into [0, L) in place (move.c:33-41), links every pair
   with squared float distance below ll*ll under the periodic minimum image
   (strict, fof.c:586-587), and for groups of >= nfof_min members writes
   label[i] = smallest member index (else -1) and one halo per group in order
   of that index: centre x0 + mean folded offset from x0 (x0 = the smallest
   member's position), wrapped into [0, L); mean velocity (rec[i*stride+3..5]).
   Returns the number of halos (halos must hold np entries). */
int orc_fof(float* rec, long stride, int np, float boxsize, float ll, int nfof_min, int* label, OrcHalo* halos);

/* ---- learned force residual: 3-D periodic conv net, double precision ----
   Layers (taps t = 9 (dx+1) + 3 (dy+1) + (dz+1), input at (x+dx, y+dy, z+dz)
   with periodic wrap; activations channel-last [cell][channel]):
     h1 = relu(conv(x, W1[27][8][16]) + b1)      x: 8 input channels
     h2 = relu(conv(h1, W2[27][16][16]) + b2)
     h3 = relu(conv(h2, W3[27][16][16]) + b3)
     y  = conv(h3, W4[27][16][3]) + b4
   Parameters are packed in that order (W1, b1, W2, b2, W3, b3, W4, b4):
   ORC_NET_NPARAM doubles.  Loss = mean over cells and 3 components of
   (y - target)^2.  round_bf16 != 0 rounds the input, the weights and every
   stored activation to bfloat16 (round to nearest even) the way the device
   kernels store them; the arithmetic stays double. */
#define ORC_NET_NPARAM (27 * 8 * 16 + 16 + 2 * (27 * 16 * 16 + 16) + 27 * 16 * 3 + 3)
double orc_bf16_round(double x);
void orc_conv3d(int n, int cin, int cout, const double* in, const double* w, const double* b, double* out);
/* forward; y: [n^3][3]; returns the loss if target != NULL (else 0) */
double orc_net_forward(int n, const double* params, const double* x, const double* target, double* y,
                       int round_bf16);
/* loss and its gradient with respect to every parameter (grad: ORC_NET_NPARAM) */
double orc_net_gradient(int n, const double* params, const double* x, const double* target, double* grad,
                        int round_bf16);

/* ---- main.c schedule ---- */
int orc_run(int nc, int ngrid, double boxsize, double omega_m, double a_final,
            int ntimestep, int seed, const double* zout, int nout,
            OrcParticleMin* snapshots, float* snapshot_a, int* nsnap_written);

#ifdef __cplusplus
}
#endif
#endif
