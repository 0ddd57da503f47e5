// C ABI of libcola_halo_amd.so: context lifetime, the reference's module
// entry points (power_init, lpt_set_displacement, move_particles2,
// pm_calculate_forces, cola_kick, cola_drift, cola_set_snapshot,
// set_noncola_initial) and state access.  All device work of a context runs
// in order on its one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cola_halo_amd.h"
#include "analysis.h"
#include "common.h"
#include "fof.h"
#include "resnet.h"
#include "cosmo.h"
#include "fft.h"
#include "lpt.h"
#include "pm.h"

namespace cola {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

// Profiled phases, one per kernel family.
enum Phase : int {
  kPhWrap = 0,
  kPhBinCount,
  kPhBinScan,
  kPhBinScatter,
  kPhDeposit,
  kPhFftZR2C,
  kPhFftYFwd,
  kPhFftXFwd,
  kPhFftXInv,
  kPhFftYInv,
  kPhFftZC2R,
  kPhFftXFused,
  kPhReadout,
  kPhReadoutKickDrift,
  kPhKick,
  kPhDrift,
  kPhSnapshot,
  kPhLptModes,
  kPhLptFft,
  kPhLptOther,
  kPhNetFeatures,
  kPhNetConv1,
  kPhNetHidden,
  kPhNetHead,
  kPhNetConv12,
  kPhNetConv34,
  kNumPhases
};
const char* const kPhaseNames[kNumPhases] = {
    "wrap",      "bin_count",  "bin_scan",   "bin_scatter", "deposit",   "fft_z_r2c",     "fft_y_fwd",
    "fft_x_fwd", "fft_x_inv",  "fft_y_inv",  "fft_z_c2r",   "fft_x_fused", "readout",   "readout_kick_drift", "kick",
    "drift",     "snapshot",   "lpt_modes",  "lpt_fft",     "lpt_other", "net_features", "net_conv1",
    "net_conv_hidden", "net_conv_out_add", "net_conv12", "net_conv34"};

struct EventPair {
  int phase;
  hipEvent_t a, b;
};

}  // namespace cola

using namespace cola;

struct cola_context {
  int nc = 0, ng = 0;
  double boxsize = 0;
  float boxf = 0;
  long long np = 0;
  float wpar = 0;
  hipStream_t stream = nullptr;

  DevParticles p;
  float* pbuf = nullptr;
  float* density = nullptr;
  float* fgrid[3] = {nullptr, nullptr, nullptr};
  long long grid_floats = 0;

  FftPlan pm_fft, lpt_fft;
  int fft_log2_tile = -1;  // column passes (cola_set_tiling sets both)
  int fft_log2_rows = -1;
  int fft_log2_xtile = -1;  // x column pass when set apart from the y pass (-1: as fft_log2_tile)
  int fft_schedule = 0;     // radix schedule of the PM transforms (fft_factor)
  CicGeom cic;
  TileGeom tiles;
  BinBuffers bins;

  static constexpr int kMassSlots = 64;
  double* d_mass = nullptr;
  StepParams* h_prm = nullptr;  // pinned ring, one entry per mass slot
  StepParams* d_prm = nullptr;  // the values of the step in flight
  bool graph_mode = false;
  ResNet* net = nullptr;  // learned force residual added to the force grids (not owned)
  cola_resnet* net_handle = nullptr;  // the handle `net` belongs to
  hipGraphExec_t step_graph = nullptr;  // captured fused step
  // bumped whenever buffers, tiles or plans a captured step or an ensemble's
  // tables refer to change
  unsigned long long config_gen = 0;
  int mass_slot = 0;     // next slot
  int mass_pending = 0;  // slots written since the last check
  double last_mass = 0;

  PowerTable power;
  double pow_omega_m = 0.3, pow_omega_lambda = 0.7, sigma8_input = 0;
  double* d_logk = nullptr;
  double* d_logD = nullptr;

  float a_x = 0, a_v = 0;
  bool have_particles = false;
  bool have_force = false;
  bool have_grids = false;
  bool fi_force = false;  // the last force was corrected: its grids are the net's interleaved d_fi
  bool density_is_real = false;

  DevParticleMin* d_snap = nullptr;
  bool have_snapshot = false;
  float* saved = nullptr;  // x, v, dx1, dx2 copy of cola_state_save
  float saved_ax = 0, saved_av = 0;

  // Output grids (coarse density grid, power spectrum), sized on first use.
  float* aux_grid = nullptr;
  long long aux_floats = 0;
  unsigned long long* aux_fix = nullptr;  // fixed-point CIC sums, n^3
  long long aux_fix_cells = 0;
  FftPlan aux_fft;
  int aux_fft_n = 0;
  double* d_aux = nullptr;  // 2 sums + 3 * kMaxShellBins shell accumulators
  FofResult fof;            // last friends-of-friends catalogue

  bool profiling = false;
  std::vector<EventPair> events;
  std::vector<hipEvent_t> free_events;
  double phase_ms[kNumPhases] = {0};
  long long phase_n[kNumPhases] = {0};
  int open_phase = -1;
  hipEvent_t open_event = nullptr;
  std::vector<cola_ensemble*> ensembles;  // the ensembles this context is a member of
};

// A net's activation buffers serve one context at a time: `attached` is the
// context whose steps run the net on the context's stream.
struct cola_resnet {
  ResNet r;
  hipStream_t stream = nullptr;
  cola_context* attached = nullptr;
};

// A member context of ensemble e is being destroyed: e drops its tables and
// graph (they hold the context's buffers) and refuses to step from then on.
static void ensemble_forget(cola_ensemble* e, cola_context* c);

namespace {

int get_event(cola_context* c, hipEvent_t* e)
{
  if (!c->free_events.empty()) {
    *e = c->free_events.back();
    c->free_events.pop_back();
    return kOk;
  }
  COLA_HIP(hipEventCreate(e));
  return kOk;
}

void phase_begin(cola_context* c, int ph)
{
  if (!c->profiling) return;
  hipEvent_t e;
  if (get_event(c, &e) != kOk) return;
  (void)hipEventRecord(e, c->stream);
  c->open_phase = ph;
  c->open_event = e;
}

void phase_end(cola_context* c, int ph)
{
  if (!c->profiling || c->open_phase != ph) return;
  hipEvent_t e;
  if (get_event(c, &e) != kOk) return;
  (void)hipEventRecord(e, c->stream);
  c->events.push_back({ph, c->open_event, e});
  c->open_phase = -1;
}

int collect_events(cola_context* c)
{
  if (c->events.empty()) return kOk;
  COLA_HIP(hipStreamSynchronize(c->stream));
  for (const EventPair& ep : c->events) {
    float ms = 0;
    COLA_HIP(hipEventElapsedTime(&ms, ep.a, ep.b));
    c->phase_ms[ep.phase] += ms;
    c->phase_n[ep.phase] += 1;
    c->free_events.push_back(ep.a);
    c->free_events.push_back(ep.b);
  }
  c->events.clear();
  return kOk;
}

void fft_hook(void* user, int pass, int begin)
{
  cola_context* c = static_cast<cola_context*>(user);
  static const int map[7] = {kPhFftZR2C, kPhFftYFwd, kPhFftXFwd, kPhFftXInv, kPhFftYInv, kPhFftZC2R, kPhFftXFused};
  if (begin)
    phase_begin(c, map[pass]);
  else
    phase_end(c, map[pass]);
}

void lpt_fft_hook(void* user, int pass, int begin)
{
  cola_context* c = static_cast<cola_context*>(user);
  if (begin)
    phase_begin(c, kPhLptFft);
  else
    phase_end(c, kPhLptFft);
}

// Check the deposit's mass sums written since the last check (pm.c:535-562:
// |sum| <= 1e-7 * WPAR * Np) and the binning overflow flag.
int check_pending(cola_context* c)
{
  if (c->mass_pending == 0) return kOk;
  const int n = std::min(c->mass_pending, cola_context::kMassSlots);
  std::vector<double> sums(cola_context::kMassSlots);
  int overflow = 0;
  COLA_HIP(hipStreamSynchronize(c->stream));
  COLA_HIP(hipMemcpy(sums.data(), c->d_mass, sizeof(double) * cola_context::kMassSlots, hipMemcpyDeviceToHost));
  COLA_HIP(hipMemcpy(&overflow, c->bins.overflow, sizeof(int), hipMemcpyDeviceToHost));
  c->mass_pending = 0;
  const double tol = 1.0e-7 * (double)c->wpar * (double)c->np;
  int status = kOk;
  for (int k = 0; k < n; ++k) {
    const int slot = (c->mass_slot - 1 - k + cola_context::kMassSlots) % cola_context::kMassSlots;
    const double s = sums[slot];
    if (k == 0) c->last_mass = s;
    if (std::fabs(s) > tol) {
      set_error("total CIC density error is large: " + std::to_string(s) + " > " + std::to_string(tol));
      status = kErrMass;
    }
  }
  if (overflow) {
    const int zero = 0;
    COLA_HIP(hipMemcpy(c->bins.overflow, &zero, sizeof(int), hipMemcpyHostToDevice));
    set_error(overflow & 2 ? "deposit bin capacity exceeded" : "particle position outside the PM grid");
    status = kErrMass;
  }
  return status;
}

int alloc_grids(cola_context* c)
{
  if (c->have_grids) return kOk;
  const size_t bytes = sizeof(float) * c->grid_floats;
  COLA_HIP(hipMalloc(&c->density, bytes));
  for (int a = 0; a < 3; ++a) COLA_HIP(hipMalloc(&c->fgrid[a], bytes));
  // the row padding is never written by the deposit; keep it finite
  COLA_HIP(hipMemset(c->density, 0, bytes));
  for (int a = 0; a < 3; ++a) COLA_HIP(hipMemset(c->fgrid[a], 0, bytes));
  c->have_grids = true;
  return kOk;
}

// pm.c:209-279: binning and the tiled deposit into c->density.  wrap_first
// folds move_particles2 into the binning pass.
// Claims the next mass-check slot and sends this step's values to the device
// (through the pinned ring entry of the slot, which the host does not touch
// again before the stream has passed it: the ring is drained every 62 slots).
int step_params(cola_context* c, const KickFactors* k, const DriftFactors* d)
{
  if (c->mass_pending >= cola_context::kMassSlots - 2) COLA_TRY(check_pending(c));
  const int slot = c->mass_slot;
  c->mass_slot = (c->mass_slot + 1) % cola_context::kMassSlots;
  c->mass_pending += 1;
  StepParams& hp = c->h_prm[slot];
  hp.slot = slot;
  hp.a = c->a_x;
  if (k) hp.k = *k;
  if (d) hp.d = *d;
  COLA_HIP(hipMemcpyAsync(c->d_prm, &hp, sizeof(StepParams), hipMemcpyHostToDevice, c->stream));
  return kOk;
}

// pm.c:209-279 + 535-562 kernels (captured into the step graph as they are).
int assign_kernels(cola_context* c, bool wrap_first)
{
  COLA_TRY(launch_zero_mass(c->d_mass, c->d_prm, c->stream));
  phase_begin(c, kPhBinCount);
  COLA_TRY(launch_bin_count(c->p, c->cic, c->tiles, c->bins, wrap_first, c->boxf, c->stream));
  phase_end(c, kPhBinCount);
  phase_begin(c, kPhBinScan);
  COLA_TRY(launch_bin_scan(c->tiles, c->bins, c->stream));
  phase_end(c, kPhBinScan);
  phase_begin(c, kPhBinScatter);
  COLA_TRY(launch_bin_scatter(c->p, c->cic, c->tiles, c->bins, c->stream));
  phase_end(c, kPhBinScatter);
  phase_begin(c, kPhDeposit);
  COLA_TRY(launch_deposit_tiles(c->cic, c->tiles, c->bins, c->density, c->d_mass, c->d_prm, c->stream));
  phase_end(c, kPhDeposit);
  return kOk;
}

int pm_assign(cola_context* c, bool wrap_first)
{
  COLA_TRY(alloc_grids(c));
  COLA_TRY(step_params(c, nullptr, nullptr));
  COLA_TRY(assign_kernels(c, wrap_first));
  c->density_is_real = true;
  return kOk;
}

// pm.c:282-355: r2c of the density, then per axis the Poisson gradient and
// the c2r into the three force grids.
KOpParams pm_kop(const cola_context* c)
{
  const float scale = 2.0 * kPiD / c->boxf;  // pm.c:315
  KOpParams kp;
  kp.f1 = -1.0 / std::pow((double)c->ng, 3.0) / scale;  // pm.c:317
  return kp;
}

int pm_fields(cola_context* c)
{
  const KOpParams kp = pm_kop(c);
  float* const out[3] = {c->fgrid[0], c->fgrid[1], c->fgrid[2]};
  // the net's features come from the real-space density before the r2c
  // overwrites it; its last layer adds its output to the force grids
  // with fuse34 (by input planes) layer 3 runs with the last layer after the transforms
  const bool fused34 = c->net && c->net->fuse12 && c->net->roll && c->net->fuse34;
  if (c->net && c->net->fuse12) {
    phase_begin(c, kPhNetConv12);
    COLA_TRY(resnet_layers12(c->net, c->density, c->cic.row, &c->d_prm->a, c->stream));
    phase_end(c, kPhNetConv12);
    if (!fused34) {
      phase_begin(c, kPhNetHidden);
      COLA_TRY(resnet_layer3(c->net, c->stream));
      phase_end(c, kPhNetHidden);
    }
  } else if (c->net) {
    phase_begin(c, kPhNetFeatures);
    COLA_TRY(resnet_features(c->net, c->density, c->cic.row, c->stream));
    phase_end(c, kPhNetFeatures);
    phase_begin(c, kPhNetConv1);
    COLA_TRY(resnet_layer1(c->net, &c->d_prm->a, c->stream));
    phase_end(c, kPhNetConv1);
    phase_begin(c, kPhNetHidden);
    COLA_TRY(resnet_layers23(c->net, c->stream));
    phase_end(c, kPhNetHidden);
  }
  c->density_is_real = false;
  COLA_TRY(c->pm_fft.forces(c->density, out, kp, c->stream));
  if (fused34) {
    phase_begin(c, kPhNetConv34);
    COLA_TRY(resnet_layer3_head(c->net, out, c->cic.row, c->stream));
    phase_end(c, kPhNetConv34);
  } else if (c->net) {
    phase_begin(c, kPhNetHead);
    COLA_TRY(resnet_head_add(c->net, out, c->cic.row, c->stream));
    phase_end(c, kPhNetHead);
  }
  c->fi_force = c->net != nullptr;
  return kOk;
}

bool valid(const cola_context* c) { return c != nullptr; }

// The learned residual's sums over the grid of the last force (its mean is
// taken off the particles' forces), or null without a net.
const double* net_sums(const cola_context* c) { return c->net ? c->net->d_fsum : nullptr; }
// The corrected force grids, interleaved [cell][3] (what the readout
// gathers), or null without a net (the readout then reads fgrid[0..2]).
const float* net_fi(const cola_context* c) { return c->net ? c->net->d_fi : nullptr; }

int require_particles(cola_context* c)
{
  if (!c->have_particles) {
    set_error("no particles: call cola_lpt_set_displacement or cola_set_particles first");
    return kErrState;
  }
  return kOk;
}

int upload_power(cola_context* c)
{
  if (c->d_logk) (void)hipFree(c->d_logk);
  if (c->d_logD) (void)hipFree(c->d_logD);
  c->d_logk = c->d_logD = nullptr;
  const size_t n = c->power.logk.size();
  COLA_HIP(hipMalloc(&c->d_logk, sizeof(double) * n));
  COLA_HIP(hipMalloc(&c->d_logD, sizeof(double) * n));
  COLA_HIP(hipMemcpy(c->d_logk, c->power.logk.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  COLA_HIP(hipMemcpy(c->d_logD, c->power.logD.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  return kOk;
}

int finish_power_init(cola_context* c, double sigma8, double omega_m, double omega_lambda)
{
  c->pow_omega_m = omega_m;
  c->pow_omega_lambda = omega_lambda;
  if (c->power.logk.size() < 2) {
    set_error("power spectrum table needs at least two rows");
    return kErrPowerFile;
  }
  c->sigma8_input = c->power.sigma8();
  if (std::fabs(c->sigma8_input - sigma8) / sigma8 > 0.05) {
    set_error("Input sigma8 " + std::to_string(c->sigma8_input) + " is far from target sigma8 " +
              std::to_string(sigma8));
    return kErrSigma8;
  }
  return upload_power(c);
}

}  // namespace

extern "C" {

const char* cola_last_error(void) { return last_error(); }

int cola_context_create(int nc, int ngrid, double boxsize, cola_context** out)
{
  if (!out || nc <= 0 || ngrid <= 0 || !(boxsize > 0) || nc % 2 || ngrid % 2) {
    set_error("cola_context_create: nc and ngrid must be positive and even, boxsize > 0");
    return kErrArg;
  }
  *out = nullptr;
  cola_context* c = new cola_context();
  c->nc = nc;
  c->ng = ngrid;
  c->boxsize = boxsize;
  c->boxf = (float)boxsize;
  c->np = (long long)nc * nc * nc;
  c->wpar = (float)std::pow((double)ngrid / (double)nc, 3.0);
  int st = kOk;
  auto fail = [&](int s) {
    cola_context_destroy(c);
    return s;
  };
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(kErrHip);
  }
  st = c->pm_fft.init(ngrid, -1, -1, true);
  if (st) {
    set_error("unsupported PM grid size " + std::to_string(ngrid) + " (needs factors 2, 3, 5)");
    return fail(st);
  }
  st = c->lpt_fft.init(nc);
  if (st) {
    set_error("unsupported particle grid size " + std::to_string(nc) + " (needs factors 2, 3, 5)");
    return fail(st);
  }
  c->pm_fft.hook = fft_hook;
  c->pm_fft.hook_user = c;
  c->lpt_fft.hook = lpt_fft_hook;
  c->lpt_fft.hook_user = c;
  c->cic = make_cic_geom(ngrid, c->boxf, c->wpar, 2LL * c->pm_fft.nzp);
  // one y band of each lattice slab (nc^2 particles) per XCD
  c->cic.readout_group = (int)std::max(8LL, (long long)nc * nc / 256 / 8);
  c->grid_floats = (long long)ngrid * ngrid * c->cic.row;
  // default deposit tile 8 x 8 x 32 cells, halved until it divides ngrid
  auto pick = [&](int want) {
    for (int t = want; t > 1; t /= 2)
      if (ngrid % t == 0) return t;
    return 1;
  };
  if (!make_tile_geom(ngrid, pick(8), pick(8), pick(32), &c->tiles)) return fail(kErrArg);
  const size_t pbytes = sizeof(float) * (size_t)c->np;
  if (hipMalloc(&c->pbuf, pbytes * 15) != hipSuccess) {
    set_error("cannot allocate particle state");
    return fail(kErrAlloc);
  }
  c->p.np = c->np;
  for (int a = 0; a < 3; ++a) {
    c->p.x[a] = c->pbuf + (0 + a) * c->np;
    c->p.v[a] = c->pbuf + (3 + a) * c->np;
    c->p.dx1[a] = c->pbuf + (6 + a) * c->np;
    c->p.dx2[a] = c->pbuf + (9 + a) * c->np;
    c->p.f[a] = c->pbuf + (12 + a) * c->np;
  }
  st = bin_alloc(&c->bins, c->tiles, c->np);
  if (st) return fail(kErrAlloc);
  if (hipMalloc(&c->d_mass, sizeof(double) * cola_context::kMassSlots) != hipSuccess) return fail(kErrAlloc);
  if (hipMemset(c->d_mass, 0, sizeof(double) * cola_context::kMassSlots) != hipSuccess) return fail(kErrHip);
  if (hipHostMalloc(&c->h_prm, sizeof(StepParams) * cola_context::kMassSlots) != hipSuccess) return fail(kErrAlloc);
  std::memset(c->h_prm, 0, sizeof(StepParams) * cola_context::kMassSlots);
  if (hipMalloc(&c->d_prm, sizeof(StepParams)) != hipSuccess) return fail(kErrAlloc);
  if (hipMemset(c->d_prm, 0, sizeof(StepParams)) != hipSuccess) return fail(kErrHip);
  *out = c;
  return kOk;
}

int cola_context_destroy(cola_context* c)
{
  if (!c) return kOk;
  // the context's stream waits for every ensemble step it was part of
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->net_handle) c->net_handle->attached = nullptr;
  for (cola_ensemble* e : c->ensembles) ensemble_forget(e, c);
  for (const EventPair& ep : c->events) {
    (void)hipEventDestroy(ep.a);
    (void)hipEventDestroy(ep.b);
  }
  for (hipEvent_t e : c->free_events) (void)hipEventDestroy(e);
  if (c->open_event) (void)hipEventDestroy(c->open_event);
  c->pm_fft.destroy();
  c->lpt_fft.destroy();
  bin_free(&c->bins);
  if (c->pbuf) (void)hipFree(c->pbuf);
  if (c->density) (void)hipFree(c->density);
  for (int a = 0; a < 3; ++a)
    if (c->fgrid[a]) (void)hipFree(c->fgrid[a]);
  if (c->d_mass) (void)hipFree(c->d_mass);
  if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
  if (c->h_prm) (void)hipHostFree(c->h_prm);
  if (c->d_prm) (void)hipFree(c->d_prm);
  if (c->d_logk) (void)hipFree(c->d_logk);
  if (c->d_logD) (void)hipFree(c->d_logD);
  if (c->d_snap) (void)hipFree(c->d_snap);
  if (c->saved) (void)hipFree(c->saved);
  c->aux_fft.destroy();
  fof_free(&c->fof);
  if (c->aux_grid) (void)hipFree(c->aux_grid);
  if (c->aux_fix) (void)hipFree(c->aux_fix);
  if (c->d_aux) (void)hipFree(c->d_aux);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return kOk;
}

int cola_synchronize(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  COLA_HIP(hipStreamSynchronize(c->stream));
  return check_pending(c);
}

int cola_power_init(cola_context* c, const char* filename, double a_init, double sigma8, double omega_m,
                    double omega_lambda)
{
  (void)a_init;
  if (!valid(c) || !filename) return kErrArg;
  if (c->power.read(filename) != 0) {
    set_error(std::string("Error: unable to read input power spectrum: ") + filename);
    return kErrPowerFile;
  }
  return finish_power_init(c, sigma8, omega_m, omega_lambda);
}

int cola_power_init_table(cola_context* c, const double* k, const double* pk, int n, double a_init, double sigma8,
                          double omega_m, double omega_lambda)
{
  (void)a_init;
  if (!valid(c) || !k || !pk || n < 2) return kErrArg;
  c->power.set(k, pk, n);
  return finish_power_init(c, sigma8, omega_m, omega_lambda);
}

int cola_eh98_nowiggle_table(double omega_m, double omega_b, double h, double ns, double sigma8, double kmin,
                             double kmax, int n, double* k_out, double* p_out)
{
  if (n < 2 || !k_out || !p_out || !(kmin > 0) || !(kmax > kmin) || !(omega_m > 0) || !(h > 0) || !(sigma8 > 0)) {
    set_error("cola_eh98_nowiggle_table: invalid arguments");
    return kErrArg;
  }
  eh98_nowiggle_table(omega_m, omega_b, h, ns, sigma8, kmin, kmax, n, k_out, p_out);
  return kOk;
}

int cola_power_table_sigma8(const double* k, const double* pk, int n, double* sigma8)
{
  if (!k || !pk || n < 2 || !sigma8) return kErrArg;
  PowerTable t;
  t.set(k, pk, n);
  *sigma8 = t.sigma8();
  return kOk;
}

int cola_power_table_eval(const double* k, const double* pk, int n, const double* kq, int nq, double* out)
{
  if (!k || !pk || n < 2 || (nq > 0 && (!kq || !out))) return kErrArg;
  PowerTable t;
  t.set(k, pk, n);
  for (int i = 0; i < nq; ++i) out[i] = t(kq[i]);
  return kOk;
}

int cola_power_file_sigma8(const char* filename, double* sigma8)
{
  if (!filename || !sigma8) return kErrArg;
  PowerTable t;
  if (t.read(filename) != 0) {
    set_error(std::string("Error: unable to read input power spectrum: ") + filename);
    return kErrPowerFile;
  }
  if (t.logk.size() < 2) return kErrPowerFile;
  *sigma8 = t.sigma8();
  return kOk;
}

int cola_lpt_seedtable(int nmesh, int seed, unsigned int* table)
{
  if (nmesh <= 0 || !table) return kErrArg;
  lpt_seedtable(nmesh, seed, table);
  return kOk;
}

int cola_ranlxd1_uniforms(unsigned long seed, int n, double* out)
{
  if (n < 0 || (n > 0 && !out)) return kErrArg;
  ranlux_uniforms(seed, n, out);
  return kOk;
}

int cola_growth_scalars(float omega_m, float a, double* out)
{
  if (!out) return kErrArg;
  const ColaScalars s(omega_m);
  out[0] = s.growthD(a);
  out[1] = s.growthD2(a);
  out[2] = s.growthD2v(a);
  out[3] = s.Qfactor(a);
  out[4] = s.DprimeQ(a, 1.0f);
  return kOk;
}

int cola_step_factors(float omega_m, float a_x, float a_v, float avel1, float apos1, double* out)
{
  if (!out) return kErrArg;
  const KickFactors k = kick_factors(omega_m, a_x, a_v, avel1);
  const DriftFactors d = drift_factors(omega_m, a_x, avel1, apos1);
  out[0] = k.dda;
  out[1] = k.q1;
  out[2] = k.q2;
  out[3] = d.dyyy;
  out[4] = d.da1;
  out[5] = d.da2;
  return kOk;
}

double cola_power_spec(const cola_context* c, double k)
{
  if (!c || c->power.empty()) return 0.0;
  return c->power(k);
}

double cola_growth_factor(const cola_context* c, double astart, double aend)
{
  if (!c) return 0.0;
  return growth_factor(astart, aend, c->pow_omega_m, c->pow_omega_lambda);
}

double cola_power_sigma8(const cola_context* c) { return c ? c->sigma8_input : 0.0; }

}  // extern "C"

namespace {

// lpt.c:172-351: the seed table on the host, then the white-noise ZA
// displacement spectra of every column on the device into cd[0..2]
// (complex [i][j][kz], N x N x (N/2+1) each).
int fill_modes(cola_context* c, int seed, double box, float2* const cd[3])
{
  const int N = c->nc;
  std::vector<unsigned int> table((size_t)N * N);
  lpt_seedtable(N, seed, table.data());
  unsigned int* d_table = nullptr;
  COLA_HIP(hipMalloc(&d_table, sizeof(unsigned int) * table.size()));
  int st = hipMemcpyAsync(d_table, table.data(), sizeof(unsigned int) * table.size(), hipMemcpyHostToDevice,
                          c->stream) == hipSuccess
               ? kOk
               : kErrHip;
  if (!st) {
    DevPower pw;
    pw.logk = c->d_logk;
    pw.logD = c->d_logD;
    pw.n = (int)c->power.logk.size();
    pw.norm = c->power.norm;
    const double fac = std::pow(2 * kPiD / box, 1.5);  // lpt.c:166
    phase_begin(c, kPhLptModes);
    st = launch_mode_fill(N, d_table, pw, box, fac, cd[0], cd[1], cd[2], c->stream);
    phase_end(c, kPhLptModes);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && !st) st = kErrHip;
  (void)hipFree(d_table);
  return st;
}

}  // namespace

extern "C" {

int cola_lpt_set_displacement(cola_context* c, double InitTime, double omega_m, int seed, double box)
{
  if (!valid(c)) return kErrArg;
  if (c->power.empty() || !c->d_logk) {
    set_error("cola_lpt_set_displacement: call cola_power_init first");
    return kErrState;
  }
  const int N = c->nc;
  const long long nz = N / 2 + 1;
  const long long csize = (long long)N * N * nz;  // complex elements per grid
  float2* work = nullptr;
  COLA_HIP(hipMalloc(&work, sizeof(float2) * csize * 9));
  float2* cd[3];
  float2* cg[6];
  for (int a = 0; a < 3; ++a) cd[a] = work + a * csize;
  for (int g = 0; g < 6; ++g) cg[g] = work + (3 + g) * csize;
  auto rg = [](float2* p) { return reinterpret_cast<float*>(p); };

  int st = kOk;
  auto done = [&](int s) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(work);
    return s;
  };
  st = fill_modes(c, seed, box, cd);
  if (st) return done(st);

  // lpt.c:358-403: i k_j dis_i for (ij) = 00 01 02 11 12 22, c2r each
  static const int src[6] = {0, 0, 0, 1, 1, 2};
  static const int dir[6] = {0, 1, 2, 1, 2, 2};
  KOpParams kp;
  kp.box = box;
  for (int g = 0; g < 6; ++g) {
    kp.axis = dir[g];
    st = c->lpt_fft.c2r(cd[src[g]], rg(cg[g]), KOp::LptGrad, kp, c->stream);
    if (st) return done(st);
  }
  phase_begin(c, kPhLptOther);
  st = launch_lpt_source(N, rg(cg[0]), rg(cg[1]), rg(cg[2]), rg(cg[3]), rg(cg[4]), rg(cg[5]), c->stream);
  phase_end(c, kPhLptOther);
  if (st) return done(st);
  st = c->lpt_fft.r2c(rg(cg[3]), c->stream);
  if (st) return done(st);
  // lpt.c:433-503: second-order displacements into cdigrad[0..2], ZA in place
  for (int a = 0; a < 3; ++a) {
    kp.axis = a;
    st = c->lpt_fft.c2r(cg[3], rg(cg[a]), KOp::Lpt2, kp, c->stream);
    if (st) return done(st);
  }
  for (int a = 0; a < 3; ++a) {
    st = c->lpt_fft.c2r(nullptr, rg(cd[a]), KOp::None, kp, c->stream);
    if (st) return done(st);
  }
  // lpt.c:507-561
  LptParticleFactors f;
  const double Omega = omega_m;
  f.dx = box / N;
  f.Dplus = 1.0 / growth_factor(InitTime, 1.0, c->pow_omega_m, c->pow_omega_lambda);
  f.nmesh3_inv = 1.0 / std::pow((double)N, 3.0);
  const double a = InitTime;
  const double omega = Omega / (Omega + (1.0 - Omega) * a * a * a);
  f.D2 = (float)(f.Dplus * f.Dplus) * std::pow(omega / Omega, -1.0 / 143.0);
  f.D20 = std::pow(Omega, -1.0 / 143.0);
  const float* disp[3] = {rg(cd[0]), rg(cd[1]), rg(cd[2])};
  const float* disp2[3] = {rg(cg[0]), rg(cg[1]), rg(cg[2])};
  phase_begin(c, kPhLptOther);
  st = launch_lpt_particles(N, disp, disp2, f, c->p, c->stream);
  phase_end(c, kPhLptOther);
  if (st) return done(st);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return done(kErrHip);
  c->a_x = InitTime;
  c->a_v = 0.0f;
  c->have_particles = true;
  c->have_force = false;
  return done(kOk);
}

int cola_lpt_modes(cola_context* c, int seed, double box, float* out, size_t n_out)
{
  if (!valid(c) || !out || !(box > 0)) return kErrArg;
  if (c->power.empty() || !c->d_logk) {
    set_error("cola_lpt_modes: call cola_power_init first");
    return kErrState;
  }
  const int N = c->nc;
  const long long csize = (long long)N * N * (N / 2 + 1);
  if (n_out < (size_t)(6 * csize)) {
    set_error("mode buffer too small: needs 6 nc^2 (nc/2+1) floats");
    return kErrArg;
  }
  float2* work = nullptr;
  COLA_HIP(hipMalloc(&work, sizeof(float2) * csize * 3));
  float2* const cd[3] = {work, work + csize, work + 2 * csize};
  int st = fill_modes(c, seed, box, cd);
  if (!st && hipMemcpy(out, work, sizeof(float2) * csize * 3, hipMemcpyDeviceToHost) != hipSuccess) st = kErrHip;
  (void)hipFree(work);
  return st;
}

int cola_lpt_get_local_nx(const cola_context* c) { return c ? c->nc : 0; }

int cola_move_particles(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  phase_begin(c, kPhWrap);
  COLA_TRY(launch_wrap(c->p, c->boxf, c->stream));
  phase_end(c, kPhWrap);
  return kOk;
}

int cola_pm_calculate_forces(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  COLA_TRY(pm_assign(c, false));
  // the reference checks the mass every step before its FFT (pm.c:574)
  COLA_TRY(check_pending(c));
  COLA_TRY(pm_fields(c));
  phase_begin(c, kPhReadout);
  COLA_TRY(launch_readout(c->p, c->cic, c->fgrid[0], c->fgrid[1], c->fgrid[2], net_sums(c), net_fi(c), c->stream));
  phase_end(c, kPhReadout);
  c->have_force = true;
  return kOk;
}

int cola_pm_assign(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  COLA_TRY(pm_assign(c, false));
  return check_pending(c);
}

int cola_kick(cola_context* c, float omega_m, float avel1)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  const KickFactors k = kick_factors(omega_m, c->a_x, c->a_v, avel1);
  phase_begin(c, kPhKick);
  COLA_TRY(launch_kick(c->p, k, c->stream));
  phase_end(c, kPhKick);
  c->a_v = avel1;
  return kOk;
}

int cola_drift(cola_context* c, float omega_m, float apos1)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  const DriftFactors d = drift_factors(omega_m, c->a_x, c->a_v, apos1);
  phase_begin(c, kPhDrift);
  COLA_TRY(launch_drift(c->p, d, c->stream));
  phase_end(c, kPhDrift);
  c->a_x = apos1;
  return kOk;
}

}  // extern "C"

namespace {

// The kernels of one fused step after step_params: wrap + deposit, the force
// fields, readout + kick + drift.
int step_kernels(cola_context* c)
{
  COLA_TRY(assign_kernels(c, true));
  COLA_TRY(pm_fields(c));
  phase_begin(c, kPhReadoutKickDrift);
  COLA_TRY(launch_readout_kick_drift(c->p, c->cic, c->fgrid[0], c->fgrid[1], c->fgrid[2], c->d_prm, net_sums(c),
                                     net_fi(c), c->stream));
  phase_end(c, kPhReadoutKickDrift);
  return kOk;
}

void destroy_step_graph(cola_context* c)
{
  if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
  c->step_graph = nullptr;
}

// The context's buffers, tiles or plans change: captured steps and ensemble
// tables built on them are stale.
void drop_step_graph(cola_context* c)
{
  destroy_step_graph(c);
  c->config_gen += 1;
}

// Captures step_kernels once; every later step replays it after sending its
// values (step_params).  Per-phase profiling is off inside the graph.
int capture_step_graph(cola_context* c)
{
  const bool prof = c->profiling;
  c->profiling = false;
  COLA_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int st = step_kernels(c);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->profiling = prof;
  if (st != kOk || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (st == kOk) set_error(std::string("step capture: ") + hipGetErrorString(e));
    return st != kOk ? st : kErrHip;
  }
  const hipError_t ei = hipGraphInstantiate(&c->step_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    c->step_graph = nullptr;
    set_error(std::string("step graph instantiate: ") + hipGetErrorString(ei));
    return kErrHip;
  }
  return kOk;
}

}  // namespace

extern "C" {

int cola_step(cola_context* c, float omega_m, float avel1, float apos1)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  COLA_TRY(alloc_grids(c));
  const KickFactors k = kick_factors(omega_m, c->a_x, c->a_v, avel1);
  // the drift starts from a_v = avel1 (cola.c:124 after the kick)
  const DriftFactors d = drift_factors(omega_m, c->a_x, avel1, apos1);
  if (c->graph_mode && !c->step_graph) COLA_TRY(capture_step_graph(c));
  COLA_TRY(step_params(c, &k, &d));
  if (c->graph_mode)
    COLA_HIP(hipGraphLaunch(c->step_graph, c->stream));
  else
    COLA_TRY(step_kernels(c));
  c->density_is_real = false;
  c->a_v = avel1;
  c->a_x = apos1;
  c->have_force = false;
  if (c->mass_pending >= cola_context::kMassSlots / 2) return check_pending(c);
  return kOk;
}

int cola_step_batch(cola_context* const* ctxs, int n, float omega_m, const float* avel1, const float* apos1)
{
  if (n < 0 || (n > 0 && (!ctxs || !avel1 || !apos1))) return kErrArg;
  for (int i = 0; i < n; ++i) COLA_TRY(cola_step(ctxs[i], omega_m, avel1[i], apos1[i]));
  return kOk;
}

int cola_set_step_graph(cola_context* c, int on)
{
  if (!valid(c)) return kErrArg;
  COLA_HIP(hipStreamSynchronize(c->stream));
  c->graph_mode = on != 0;
  if (!c->graph_mode) destroy_step_graph(c);
  return kOk;
}

}  // extern "C"

// The steps of many contexts with one launch per kernel (batched tables of
// every member's buffers, tiles and plans).
struct cola_ensemble {
  std::vector<cola_context*> ctx;
  std::vector<unsigned long long> gen;  // each member's config_gen the tables were built for
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<hipEvent_t> ready;  // per member: its queued work before a step
  PmBatch pm;
  FftForcesBatch fft;
  bool built = false;
  static constexpr int kSlots = 64;
  StepParams* h_prm = nullptr;  // pinned ring of kSlots entries of n values
  StepParams* d_prm = nullptr;  // the values of the step in flight, one per member
  int ring = 0;
  int since_check = 0;  // steps since the members' mass sums were checked
  bool graph_mode = false;
  bool merge_all = false;  // one transform launch per pass for every member (else per kernel variant)
  bool warmed = false;  // the kernels ran eagerly since the tables were built
  hipGraphExec_t graph = nullptr;
};

namespace {

void drop_ensemble_graph(cola_ensemble* e)
{
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  e->graph = nullptr;
}

}  // namespace

static void ensemble_forget(cola_ensemble* e, cola_context* c)
{
  (void)hipStreamSynchronize(e->stream);
  drop_ensemble_graph(e);
  e->pm.destroy();
  e->fft.destroy();
  e->built = false;
  for (cola_context*& m : e->ctx)
    if (m == c) m = nullptr;
}

namespace {

// (Re)builds the batched tables from the members' current buffers and plans.
int ensemble_build(cola_ensemble* e)
{
  COLA_HIP(hipStreamSynchronize(e->stream));
  drop_ensemble_graph(e);
  e->pm.destroy();
  e->fft.destroy();
  e->built = false;
  const int n = (int)e->ctx.size();
  std::vector<PmMember> ms(n);
  std::vector<const FftPlan*> plans(n);
  std::vector<float*> dens(n), outs(3 * n);
  std::vector<KOpParams> kps(n);
  for (int m = 0; m < n; ++m) {
    cola_context* c = e->ctx[m];
    COLA_TRY(require_particles(c));
    if (c->net) {
      set_error("ensemble member " + std::to_string(m) + " has a residual net attached");
      return kErrState;
    }
    if (!c->pm_fft.can_fuse()) {
      set_error("ensemble member " + std::to_string(m) + ": its PM transform's x pass does not fuse");
      return kErrUnsupported;
    }
    COLA_TRY(alloc_grids(c));
    PmMember& M = ms[m];
    M.p = c->p;
    M.g = c->cic;
    M.t = c->tiles;
    M.b = c->bins;
    M.density = c->density;
    for (int a = 0; a < 3; ++a) M.f[a] = c->fgrid[a];
    M.ring = c->d_mass;
    M.prm = e->d_prm + m;
    M.boxsize = c->boxf;
    plans[m] = &c->pm_fft;
    dens[m] = c->density;
    for (int a = 0; a < 3; ++a) outs[3 * m + a] = c->fgrid[a];
    kps[m] = pm_kop(c);
    e->gen[m] = c->config_gen;
  }
  COLA_TRY(e->pm.init(ms.data(), n));
  COLA_TRY(e->fft.init(plans.data(), dens.data(), outs.data(), kps.data(), n, e->merge_all));
  e->built = true;
  e->warmed = false;
  return kOk;
}

int ensemble_kernels(cola_ensemble* e)
{
  COLA_TRY(launch_assign_batch(e->pm, e->stream));
  COLA_TRY(e->fft.run(e->stream));
  COLA_TRY(launch_readout_kick_drift_batch(e->pm, e->stream));
  return kOk;
}

int capture_ensemble_graph(cola_ensemble* e)
{
  COLA_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  const int st = ensemble_kernels(e);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->stream, &g);
  if (st != kOk || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (st == kOk) set_error(std::string("ensemble capture: ") + hipGetErrorString(ec));
    return st != kOk ? st : kErrHip;
  }
  const hipError_t ei = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    e->graph = nullptr;
    set_error(std::string("ensemble graph instantiate: ") + hipGetErrorString(ei));
    return kErrHip;
  }
  return kOk;
}

}  // namespace

extern "C" {

int cola_ensemble_create(cola_context* const* ctxs, int n, cola_ensemble** out)
{
  if (!out || n <= 0 || !ctxs) return kErrArg;
  *out = nullptr;
  for (int m = 0; m < n; ++m) {
    if (!valid(ctxs[m])) return kErrArg;
    for (int k = 0; k < m; ++k)
      if (ctxs[k] == ctxs[m]) {
        set_error("ensemble members " + std::to_string(k) + " and " + std::to_string(m) + " are the same context");
        return kErrArg;
      }
  }
  cola_ensemble* e = new cola_ensemble;
  auto fail = [&](int st) {
    cola_ensemble_destroy(e);
    return st;
  };
  e->ctx.assign(ctxs, ctxs + n);
  e->gen.assign(n, 0);
  if (hipStreamCreate(&e->stream) != hipSuccess) return fail(kErrHip);
  if (hipEventCreateWithFlags(&e->done, hipEventDisableTiming) != hipSuccess) return fail(kErrHip);
  e->ready.assign(n, nullptr);
  for (int m = 0; m < n; ++m)
    if (hipEventCreateWithFlags(&e->ready[m], hipEventDisableTiming) != hipSuccess) return fail(kErrHip);
  if (hipHostMalloc(&e->h_prm, sizeof(StepParams) * cola_ensemble::kSlots * n) != hipSuccess) return fail(kErrAlloc);
  std::memset(e->h_prm, 0, sizeof(StepParams) * cola_ensemble::kSlots * n);
  if (hipMalloc(&e->d_prm, sizeof(StepParams) * n) != hipSuccess) return fail(kErrAlloc);
  if (hipMemset(e->d_prm, 0, sizeof(StepParams) * n) != hipSuccess) return fail(kErrHip);
  for (cola_context* c : e->ctx) c->ensembles.push_back(e);
  *out = e;
  return kOk;
}

int cola_ensemble_step(cola_ensemble* e, float omega_m, const float* avel1, const float* apos1)
{
  if (!e || !avel1 || !apos1) return kErrArg;
  const int n = (int)e->ctx.size();
  for (int m = 0; m < n; ++m)
    if (!e->ctx[m]) {
      set_error("ensemble member " + std::to_string(m) + " was destroyed");
      return kErrState;
    }
  bool stale = !e->built;
  for (int m = 0; m < n; ++m) stale = stale || e->ctx[m]->config_gen != e->gen[m] || e->ctx[m]->net;
  if (stale) COLA_TRY(ensemble_build(e));
  // after the work queued on the members
  for (int m = 0; m < n; ++m) {
    COLA_HIP(hipEventRecord(e->ready[m], e->ctx[m]->stream));
    COLA_HIP(hipStreamWaitEvent(e->stream, e->ready[m], 0));
  }
  // every member's step values (cola_step), in the pinned ring entry of this
  // step: the ring holds 64 steps and the members are drained every 32
  StepParams* hp = e->h_prm + (size_t)e->ring * n;
  for (int m = 0; m < n; ++m) {
    cola_context* c = e->ctx[m];
    if (c->mass_pending >= cola_context::kMassSlots - 2) COLA_TRY(check_pending(c));
    const KickFactors k = kick_factors(omega_m, c->a_x, c->a_v, avel1[m]);
    const DriftFactors d = drift_factors(omega_m, c->a_x, avel1[m], apos1[m]);
    hp[m].slot = c->mass_slot;
    hp[m].a = c->a_x;
    hp[m].k = k;
    hp[m].d = d;
    c->mass_slot = (c->mass_slot + 1) % cola_context::kMassSlots;
    c->mass_pending += 1;
  }
  e->ring = (e->ring + 1) % cola_ensemble::kSlots;
  COLA_HIP(hipMemcpyAsync(e->d_prm, hp, sizeof(StepParams) * n, hipMemcpyHostToDevice, e->stream));
  // the first step after a (re)build runs eagerly (one-time kernel set-up
  // stays out of the capture); later ones replay the captured graph
  if (e->graph_mode && !e->graph && e->warmed) COLA_TRY(capture_ensemble_graph(e));
  if (e->graph_mode && e->graph) {
    COLA_HIP(hipGraphLaunch(e->graph, e->stream));
  } else {
    COLA_TRY(ensemble_kernels(e));
    e->warmed = true;
  }
  // the members' later work after this step
  COLA_HIP(hipEventRecord(e->done, e->stream));
  for (int m = 0; m < n; ++m) {
    cola_context* c = e->ctx[m];
    COLA_HIP(hipStreamWaitEvent(c->stream, e->done, 0));
    c->density_is_real = false;
    c->a_v = avel1[m];
    c->a_x = apos1[m];
    c->have_force = false;
    c->fi_force = false;
  }
  if (++e->since_check >= cola_ensemble::kSlots / 2) {
    e->since_check = 0;
    int st = kOk;
    for (int m = 0; m < n; ++m) {
      const int r = check_pending(e->ctx[m]);
      if (st == kOk) st = r;
    }
    return st;
  }
  return kOk;
}

int cola_ensemble_set_graph(cola_ensemble* e, int on)
{
  if (!e) return kErrArg;
  COLA_HIP(hipStreamSynchronize(e->stream));
  e->graph_mode = on != 0;
  if (!e->graph_mode) drop_ensemble_graph(e);
  return kOk;
}

int cola_ensemble_set_merge(cola_ensemble* e, int on)
{
  if (!e || (on != 0 && on != 1)) return kErrArg;
  if (e->merge_all == (on != 0)) return kOk;
  COLA_HIP(hipStreamSynchronize(e->stream));
  e->merge_all = on != 0;
  drop_ensemble_graph(e);
  e->built = false;  // rebuilt at the next step
  return kOk;
}

int cola_ensemble_launches(const cola_ensemble* e)
{
  if (!e || !e->built) return 0;
  // mass slot, count, scan (3), scatter, deposit, sums; the transform passes; readout
  return 8 + (int)e->fft.stages.size() + 1;
}

int cola_ensemble_destroy(cola_ensemble* e)
{
  if (!e) return kOk;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (cola_context* c : e->ctx)
    if (c) c->ensembles.erase(std::remove(c->ensembles.begin(), c->ensembles.end(), e), c->ensembles.end());
  drop_ensemble_graph(e);
  e->pm.destroy();
  e->fft.destroy();
  for (hipEvent_t ev : e->ready)
    if (ev) (void)hipEventDestroy(ev);
  if (e->done) (void)hipEventDestroy(e->done);
  if (e->h_prm) (void)hipHostFree(e->h_prm);
  if (e->d_prm) (void)hipFree(e->d_prm);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return kOk;
}

static int snapshot_common(cola_context* c, const SnapshotFactors& f, bool noncola, cola_particle_min* out,
                           size_t np_out)
{
  if (!out || np_out < (size_t)c->np) {
    set_error("snapshot buffer too small");
    return kErrArg;
  }
  if (!c->d_snap) COLA_HIP(hipMalloc(&c->d_snap, sizeof(DevParticleMin) * c->np));
  phase_begin(c, kPhSnapshot);
  if (noncola)
    COLA_TRY(launch_noncola_initial(c->p, f, c->d_snap, c->stream));
  else
    COLA_TRY(launch_snapshot(c->p, f, c->d_snap, c->stream));
  phase_end(c, kPhSnapshot);
  c->have_snapshot = true;
  static_assert(sizeof(DevParticleMin) == sizeof(cola_particle_min), "snapshot record layout");
  COLA_HIP(hipMemcpyAsync(out, c->d_snap, sizeof(DevParticleMin) * c->np, hipMemcpyDeviceToHost, c->stream));
  COLA_HIP(hipStreamSynchronize(c->stream));
  return check_pending(c);
}

int cola_set_snapshot(cola_context* c, double aout, float omega_m, cola_particle_min* out, size_t np_out)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  if (!c->have_force) {
    set_error("cola_set_snapshot needs the force of the current step (cola_pm_calculate_forces)");
    return kErrState;
  }
  return snapshot_common(c, snapshot_factors(omega_m, c->a_x, c->a_v, aout), false, out, np_out);
}

int cola_set_noncola_initial(cola_context* c, float omega_m, cola_particle_min* out, size_t np_out)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  return snapshot_common(c, noncola_initial_factors(omega_m, c->a_x), true, out, np_out);
}

long long cola_np(const cola_context* c) { return c ? c->np : 0; }

int cola_get_times(const cola_context* c, float* a_x, float* a_v)
{
  if (!c) return kErrArg;
  if (a_x) *a_x = c->a_x;
  if (a_v) *a_v = c->a_v;
  return kOk;
}

int cola_set_times(cola_context* c, float a_x, float a_v)
{
  if (!valid(c)) return kErrArg;
  c->a_x = a_x;
  c->a_v = a_v;
  return kOk;
}

int cola_get_particles(cola_context* c, cola_particle* out, size_t np_out)
{
  if (!valid(c) || !out || np_out < (size_t)c->np) return kErrArg;
  COLA_TRY(require_particles(c));
  std::vector<float> h((size_t)c->np * 12);
  COLA_HIP(hipStreamSynchronize(c->stream));
  COLA_HIP(hipMemcpy(h.data(), c->pbuf, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
  const long long n = c->np;
  for (long long i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      out[i].x[a] = h[(0 + a) * n + i];
      out[i].v[a] = h[(3 + a) * n + i];
      out[i].dx1[a] = h[(6 + a) * n + i];
      out[i].dx2[a] = h[(9 + a) * n + i];
    }
    out[i].id = i + 1;
  }
  return check_pending(c);
}

int cola_set_particles(cola_context* c, const cola_particle* in, size_t np_in, float a_x, float a_v)
{
  if (!valid(c) || !in || np_in != (size_t)c->np) {
    set_error("cola_set_particles: need exactly nc^3 particles in lattice-ID order");
    return kErrArg;
  }
  const long long n = c->np;
  std::vector<float> h((size_t)n * 12);
  for (long long i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      h[(0 + a) * n + i] = in[i].x[a];
      h[(3 + a) * n + i] = in[i].v[a];
      h[(6 + a) * n + i] = in[i].dx1[a];
      h[(9 + a) * n + i] = in[i].dx2[a];
    }
  COLA_HIP(hipStreamSynchronize(c->stream));
  COLA_HIP(hipMemcpy(c->pbuf, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  c->a_x = a_x;
  c->a_v = a_v;
  c->have_particles = true;
  c->have_force = false;
  return kOk;
}

int cola_get_forces(cola_context* c, float* out, size_t np_out)
{
  if (!valid(c) || !out || np_out < (size_t)c->np) return kErrArg;
  if (!c->have_force) {
    set_error("no force computed for the current state");
    return kErrState;
  }
  const long long n = c->np;
  std::vector<float> h((size_t)n * 3);
  COLA_HIP(hipStreamSynchronize(c->stream));
  COLA_HIP(hipMemcpy(h.data(), c->p.f[0], sizeof(float) * h.size(), hipMemcpyDeviceToHost));
  for (long long i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) out[3 * i + a] = h[a * n + i];
  return kOk;
}

// Device grid (rows of cic.row floats) -> host FFTW layout (rows of 2(ng/2+1)).
// The two padding floats of each host row are set to zero (on the device they
// hold what the in-place transforms left there).
static int grid_to_host(const cola_context* c, const float* d, float* out)
{
  const long long hrow = 2 * (c->ng / 2 + 1);
  COLA_HIP(hipMemcpy2D(out, sizeof(float) * hrow, d, sizeof(float) * c->cic.row, sizeof(float) * hrow,
                       (size_t)c->ng * c->ng, hipMemcpyDeviceToHost));
  for (long long r = 0; r < (long long)c->ng * c->ng; ++r) out[r * hrow + c->ng] = out[r * hrow + c->ng + 1] = 0.0f;
  return kOk;
}

static size_t host_grid_floats(const cola_context* c) { return (size_t)c->ng * c->ng * 2 * (c->ng / 2 + 1); }

int cola_get_density(cola_context* c, float* out, size_t n_out)
{
  if (!valid(c) || !out || n_out < host_grid_floats(c)) return kErrArg;
  if (!c->have_grids || !c->density_is_real) {
    set_error("the density grid is transformed in place by the force; call cola_pm_assign first");
    return kErrState;
  }
  COLA_HIP(hipStreamSynchronize(c->stream));
  return grid_to_host(c, c->density, out);
}

int cola_get_force_grid(cola_context* c, int axis, float* out, size_t n_out)
{
  if (!valid(c) || !out || axis < 0 || axis > 2 || n_out < host_grid_floats(c)) return kErrArg;
  if (!c->have_grids) return kErrState;
  COLA_HIP(hipStreamSynchronize(c->stream));
  if (c->fi_force && c->net) {
    // the corrected force lives interleaved in the net's grid
    const long long cells = (long long)c->ng * c->ng * c->ng;
    std::vector<float> fi(3 * (size_t)cells);
    COLA_HIP(hipMemcpy(fi.data(), c->net->d_fi, sizeof(float) * fi.size(), hipMemcpyDeviceToHost));
    const long long hrow = 2 * (c->ng / 2 + 1);
    for (long long line = 0; line < (long long)c->ng * c->ng; ++line) {
      for (int z = 0; z < c->ng; ++z) out[line * hrow + z] = fi[(line * c->ng + z) * 3 + axis];
      out[line * hrow + c->ng] = out[line * hrow + c->ng + 1] = 0.0f;
    }
    return kOk;
  }
  return grid_to_host(c, c->fgrid[axis], out);
}

double cola_last_mass_sum(const cola_context* c) { return c ? c->last_mass : 0.0; }

int cola_state_save(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(require_particles(c));
  const size_t bytes = sizeof(float) * 12 * (size_t)c->np;
  if (!c->saved) COLA_HIP(hipMalloc(&c->saved, bytes));
  COLA_HIP(hipMemcpyAsync(c->saved, c->pbuf, bytes, hipMemcpyDeviceToDevice, c->stream));
  c->saved_ax = c->a_x;
  c->saved_av = c->a_v;
  return kOk;
}

int cola_state_restore(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  if (!c->saved) {
    set_error("cola_state_restore: nothing saved");
    return kErrState;
  }
  const size_t bytes = sizeof(float) * 12 * (size_t)c->np;
  COLA_HIP(hipMemcpyAsync(c->pbuf, c->saved, bytes, hipMemcpyDeviceToDevice, c->stream));
  c->a_x = c->saved_ax;
  c->a_v = c->saved_av;
  c->have_particles = true;
  c->have_force = false;
  return kOk;
}

int cola_set_profiling(cola_context* c, int on)
{
  if (!valid(c)) return kErrArg;
  c->profiling = on != 0;
  return kOk;
}

int cola_phase_count(void) { return kNumPhases; }

const char* cola_phase_name(int phase)
{
  if (phase < 0 || phase >= kNumPhases) return "";
  return kPhaseNames[phase];
}

int cola_get_phase_stats(cola_context* c, double* ms, long long* launches, int n)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(collect_events(c));
  for (int i = 0; i < n && i < kNumPhases; ++i) {
    if (ms) ms[i] = c->phase_ms[i];
    if (launches) launches[i] = c->phase_n[i];
  }
  return kOk;
}

int cola_reset_phase_stats(cola_context* c)
{
  if (!valid(c)) return kErrArg;
  COLA_TRY(collect_events(c));
  for (int i = 0; i < kNumPhases; ++i) {
    c->phase_ms[i] = 0;
    c->phase_n[i] = 0;
  }
  return kOk;
}

}  // extern "C"

namespace {

int set_deposit_tiles(cola_context* c, int tx, int ty, int tz)
{
  TileGeom t;
  if (!make_tile_geom(c->ng, tx, ty, tz, &t)) {
    set_error("deposit tile sides must be powers of two dividing ngrid (x side >= 2), <= 4096 cells");
    return kErrArg;
  }
  if (t.tx == c->tiles.tx && t.ty == c->tiles.ty && t.tz == c->tiles.tz && c->bins.counts) return kOk;
  COLA_HIP(hipStreamSynchronize(c->stream));
  drop_step_graph(c);
  COLA_TRY(check_pending(c));
  BinBuffers nb;
  COLA_TRY(bin_alloc(&nb, t, c->np));
  bin_free(&c->bins);
  c->bins = nb;
  c->tiles = t;
  return kOk;
}

}  // namespace

extern "C" {

int cola_set_tiling(cola_context* c, int tx, int ty, int tz, int fft_log2_tile)
{
  if (!valid(c)) return kErrArg;
  // the captured step graph holds the old plans' twiddle tables and tile shapes
  COLA_HIP(hipStreamSynchronize(c->stream));
  drop_step_graph(c);
  COLA_TRY(set_deposit_tiles(c, tx, ty, tz));
  if (fft_log2_tile != c->fft_log2_tile || fft_log2_tile != c->fft_log2_rows) {
    FftPlan a, b;
    COLA_TRY(a.init(c->ng, fft_log2_tile, fft_log2_tile, true, -1, c->fft_schedule));
    if (b.init(c->nc, fft_log2_tile, fft_log2_tile) != kOk) {
      a.destroy();
      return kErrUnsupported;
    }
    a.xgroup = c->pm_fft.xgroup;
    b.xgroup = c->lpt_fft.xgroup;
    c->pm_fft.destroy();
    c->lpt_fft.destroy();
    c->pm_fft = a;
    c->lpt_fft = b;
    c->pm_fft.hook = fft_hook;
    c->pm_fft.hook_user = c;
    c->lpt_fft.hook = lpt_fft_hook;
    c->lpt_fft.hook_user = c;
    c->fft_log2_rows = fft_log2_tile;
    c->fft_log2_tile = fft_log2_tile;
    c->fft_log2_xtile = -1;
  }
  return kOk;
}

int cola_set_fft_schedule(cola_context* c, int schedule)
{
  if (!valid(c) || schedule < 0 || schedule >= kFftSchedules) return kErrArg;
  if (schedule == c->fft_schedule) return kOk;
  FftPlan a;
  const int st = a.init(c->ng, c->fft_log2_tile, c->fft_log2_rows, true, c->fft_log2_xtile, schedule);
  if (st != kOk) {
    set_error("FFT radix schedule not available for this grid");
    return st;
  }
  COLA_HIP(hipStreamSynchronize(c->stream));
  drop_step_graph(c);
  a.xgroup = c->pm_fft.xgroup;
  c->pm_fft.destroy();
  c->pm_fft = a;
  c->pm_fft.hook = fft_hook;
  c->pm_fft.hook_user = c;
  c->fft_schedule = schedule;
  return kOk;
}

int cola_fft_radices(int n, int schedule, int* radices, int max)
{
  if (!radices || max < 1) return -kErrArg;
  FftRadices r;
  if (n < 2 || n % 2 || !fft_factor(n, &r, schedule)) return 0;
  const int np = std::min(r.npass, max);
  for (int i = 0; i < np; ++i) radices[i] = r.pass[i].R;
  return r.npass;
}

int cola_set_fft_tiles(cola_context* c, int log2_cols, int log2_rows, int xcd_group)
{
  return cola_set_fft_tiles_xy(c, log2_cols, -1, log2_rows, xcd_group);
}

int cola_set_fft_tiles_xy(cola_context* c, int log2_ycols, int log2_xcols, int log2_rows, int xcd_group)
{
  if (!valid(c) || log2_ycols < -1 || log2_ycols > 6 || log2_xcols < -1 || log2_xcols > 6 || log2_rows < -1 ||
      log2_rows > 6 || xcd_group < 1 || xcd_group > 64 || (xcd_group & (xcd_group - 1)))
    return kErrArg;
  COLA_HIP(hipStreamSynchronize(c->stream));
  drop_step_graph(c);
  if (log2_ycols != c->fft_log2_tile || log2_rows != c->fft_log2_rows || log2_xcols != c->fft_log2_xtile) {
    FftPlan a;
    const int st = a.init(c->ng, log2_ycols, log2_rows, true, log2_xcols, c->fft_schedule);
    if (st != kOk) {
      set_error("FFT tile widths not supported for this grid");
      return st;
    }
    c->pm_fft.destroy();
    c->pm_fft = a;
    c->pm_fft.hook = fft_hook;
    c->pm_fft.hook_user = c;
    c->fft_log2_tile = log2_ycols;
    c->fft_log2_rows = log2_rows;
    c->fft_log2_xtile = log2_xcols;
  }
  c->pm_fft.xgroup = xcd_group;
  return kOk;
}

}  // extern "C"

namespace {

// Mean device time of `reps` PM force fields (deposit + transforms) of the
// current particle state.
int time_pm_fields(cola_context* c, int reps, double* ms)
{
  hipEvent_t e0 = nullptr, e1 = nullptr;
  COLA_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return kErrHip;
  }
  int st = pm_assign(c, false);
  if (!st) st = pm_fields(c);
  if (!st && hipEventRecord(e0, c->stream) != hipSuccess) st = kErrHip;
  for (int r = 0; r < reps && !st; ++r) {
    st = pm_assign(c, false);
    if (!st) st = pm_fields(c);
  }
  float t = 0;
  if (!st && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&t, e0, e1) != hipSuccess))
    st = kErrHip;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!st) st = check_pending(c);
  *ms = t / reps;
  return st;
}

}  // namespace

extern "C" {

int cola_autotune(cola_context* c, int reps, cola_tuning* out)
{
  if (!valid(c) || reps < 1 || !out) return kErrArg;
  COLA_TRY(require_particles(c));
  const bool prof = c->profiling;
  c->profiling = false;
  // the shapes tuned here are the PM's; the net's kernels have fixed shapes
  // and would only add their time (and its noise) to every trial
  COLA_TRY(launch_wrap(c->p, c->boxf, c->stream));  // move.c:33-41, as every step begins
  ResNet* const net = c->net;
  c->net = nullptr;  // restored below on every path
  cola_tuning best;
  best.tile[0] = c->tiles.tx;
  best.tile[1] = c->tiles.ty;
  best.tile[2] = c->tiles.tz;
  best.fft_log2_cols = c->fft_log2_tile;
  best.fft_log2_rows = c->fft_log2_rows;
  best.fft_log2_xcols = c->fft_log2_xtile;
  best.xcd_group = c->pm_fft.xgroup;
  best.fft_schedule = c->fft_schedule;
  int st = time_pm_fields(c, reps, &best.ms_before);
  best.ms_after = best.ms_before;
  auto apply = [&](const cola_tuning& t) -> int {
    int r = set_deposit_tiles(c, t.tile[0], t.tile[1], t.tile[2]);
    if (!r) r = cola_set_fft_schedule(c, t.fft_schedule);
    if (!r) r = cola_set_fft_tiles_xy(c, t.fft_log2_cols, t.fft_log2_xcols, t.fft_log2_rows, t.xcd_group);
    return r;
  };
  auto trial = [&](cola_tuning t) {
    if (st) return;
    if (apply(t) != kOk) return;  // shape not supported for this grid: skip it
    double ms = 0;
    st = time_pm_fields(c, reps, &ms);
    if (!st && ms < best.ms_after) {
      best = t;
      best.ms_after = ms;
    }
  };
  // coordinate descent: deposit tile, then FFT column / row tiles, the x
  // column tile, then XCD group
  static const int tiles[][3] = {{8, 8, 32}, {16, 8, 32}, {8, 16, 32}, {16, 16, 16}, {8, 8, 64}, {4, 8, 64}, {16, 16, 32}};
  for (const auto& tl : tiles) {
    cola_tuning t = best;
    t.tile[0] = tl[0], t.tile[1] = tl[1], t.tile[2] = tl[2];
    trial(t);
  }
  // the x column pass (fused with the Poisson operator) apart from the y
  // pass, before and after the y column / z row widths
  auto sweep_x = [&]() {
    for (int xcols = 1; xcols <= 4; ++xcols) {
      cola_tuning t = best;
      t.fft_log2_xcols = xcols;
      trial(t);
    }
  };
  sweep_x();
  for (int cols = 2; cols <= 4; ++cols)
    for (int rows = 0; rows <= 2; ++rows) {
      cola_tuning t = best;
      t.fft_log2_cols = cols;
      t.fft_log2_rows = rows;
      trial(t);
    }
  sweep_x();
  for (int g : {4, 8, 16, 64}) {
    cola_tuning t = best;
    t.xcd_group = g;
    trial(t);
  }
  // the radix schedule of the transforms (FFTW_MEASURE picks among its plans
  // the same way, pm.c:139-147), then the x column width again for it
  for (int sch = 0; sch < kFftSchedules; ++sch) {
    cola_tuning t = best;
    t.fft_schedule = sch;
    trial(t);
  }
  sweep_x();
  if (!st) st = apply(best);
  if (!st) {
    for (int i = 0; i < 8; ++i) best.fft_radices[i] = 0;
    (void)cola_fft_radices(c->ng, best.fft_schedule, best.fft_radices, 8);
  }
  c->profiling = prof;
  c->net = net;
  *out = best;
  return st;
}

}  // extern "C"

namespace {
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, long long n)
{
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace

namespace {

int position_source(cola_context* c, int source, PosSource* src)
{
  if (source == COLA_SOURCE_PARTICLES) {
    COLA_TRY(require_particles(c));
    for (int a = 0; a < 3; ++a) src->x[a] = c->p.x[a];
    src->stride = 1;
  } else if (source == COLA_SOURCE_SNAPSHOT) {
    if (!c->have_snapshot) {
      set_error("no snapshot: call cola_set_snapshot or cola_set_noncola_initial first");
      return kErrState;
    }
    const float* base = reinterpret_cast<const float*>(c->d_snap);
    for (int a = 0; a < 3; ++a) src->x[a] = base + a;
    src->stride = sizeof(DevParticleMin) / sizeof(float);
  } else {
    set_error("unknown position source");
    return kErrArg;
  }
  src->np = c->np;
  return kOk;
}

int ensure_aux(cola_context* c, long long floats, long long cells)
{
  if (!c->d_aux) COLA_HIP(hipMalloc(&c->d_aux, sizeof(double) * (2 + 3 * kMaxShellBins)));
  if (c->aux_floats < floats) {
    if (c->aux_grid) (void)hipFree(c->aux_grid);
    c->aux_grid = nullptr;
    c->aux_floats = 0;
    if (hipMalloc(&c->aux_grid, sizeof(float) * floats) != hipSuccess) {
      (void)hipGetLastError();
      set_error("unable to allocate the output grid");
      return kErrAlloc;
    }
    c->aux_floats = floats;
  }
  if (c->aux_fix_cells < cells) {
    if (c->aux_fix) (void)hipFree(c->aux_fix);
    c->aux_fix = nullptr;
    c->aux_fix_cells = 0;
    if (hipMalloc(&c->aux_fix, sizeof(unsigned long long) * cells) != hipSuccess) {
      (void)hipGetLastError();
      set_error("unable to allocate the output grid");
      return kErrAlloc;
    }
    c->aux_fix_cells = cells;
  }
  return kOk;
}

// coarse_grid.c:49-68 on the device: CIC assignment, then delta = rho n^3 / np - 1.
int deposit_contrast(cola_context* c, const PosSource& src, int n, long long row)
{
  COLA_TRY(launch_cic_periodic(src, n, c->boxf, c->aux_fix, c->stream));
  COLA_TRY(launch_fill_zero(c->d_aux, 2, c->stream));
  const float nbar_inv = (double)n * n * n / (double)c->np;  // coarse_grid.c:56
  return launch_density_contrast(n, row, nbar_inv, c->aux_fix, c->aux_grid, c->d_aux, c->stream);
}

}  // namespace

// Scratch of cola_coarse_grid_noalloc: the 64-bit fixed-point CIC
// accumulators of one grid of up to kNoAllocGrid^3 cells, then two doubles
// for the density sums when the caller passes none.  Every GPU the library
// runs on holds its own copy (16 MiB + 256 B) until the process ends.  Host
// code never passes the array's name to a kernel: it passes the device
// address hipGetSymbolAddress returns for the current device
// (noalloc_scratch()).
constexpr int kNoAllocGrid = 128;
constexpr long long kNoAllocCells = (long long)kNoAllocGrid * kNoAllocGrid * kNoAllocGrid;
static __device__ __align__(256) unsigned long long scratch[kNoAllocCells + 32];

namespace {

constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mutex;
unsigned long long* g_scratch_ptr[kNoAllocDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device.
int noalloc_scratch(unsigned long long** out)
{
  int dev = -1;
  COLA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error("cola_coarse_grid_noalloc: device index beyond the library's table of 64");
    return kErrArg;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    COLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("cola_coarse_grid_noalloc: no device address for the scratch array");
      return kErrHip;
    }
    g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace

extern "C" {

int cola_coarse_grid(cola_context* c, int n, int source, float* delta_out, double* sums_out)
{
  if (!valid(c) || n < 1 || n > 4096 || !delta_out) return kErrArg;
  PosSource src;
  COLA_TRY(position_source(c, source, &src));
  const long long ncell = (long long)n * n * n;
  COLA_TRY(ensure_aux(c, ncell, ncell));
  COLA_TRY(deposit_contrast(c, src, n, n));
  double sums[2] = {0, 0};
  COLA_HIP(hipMemcpyAsync(delta_out, c->aux_grid, sizeof(float) * ncell, hipMemcpyDeviceToHost, c->stream));
  COLA_HIP(hipMemcpyAsync(sums, c->d_aux, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
  COLA_HIP(hipStreamSynchronize(c->stream));
  if (sums_out) {
    sums_out[0] = sums[0];
    sums_out[1] = sums[1];
  }
  COLA_TRY(check_pending(c));
  if (!(std::fabs(sums[0] - (double)c->np) < 0.1 * sums[0])) {  // coarse_grid.c:84
    set_error("coarse grid: assigned mass differs from the particle count by 10% or more");
    return kErrState;
  }
  return kOk;
}

int cola_coarse_grid_noalloc(cola_context* c, int n, int source, float* delta_dev, double* sums_dev)
{
  if (!valid(c) || n < 1 || !delta_dev) return kErrArg;
  if (n > kNoAllocGrid) {
    set_error("cola_coarse_grid_noalloc: n must be at most 128 (the size of its static scratch)");
    return kErrArg;
  }
  PosSource src;
  COLA_TRY(position_source(c, source, &src));
  unsigned long long* fix = nullptr;
  COLA_TRY(noalloc_scratch(&fix));
  double* sums = sums_dev ? sums_dev : reinterpret_cast<double*>(fix + kNoAllocCells);
  // the same kernels as cola_coarse_grid; the accumulators and sums are
  // cleared by fill kernels on the context's stream
  COLA_TRY(launch_cic_periodic(src, n, c->boxf, fix, c->stream));
  COLA_TRY(launch_fill_zero(sums, 2, c->stream));
  const float nbar_inv = (double)n * n * n / (double)c->np;  // coarse_grid.c:56
  return launch_density_contrast(n, n, nbar_inv, fix, delta_dev, sums, c->stream);
}

int cola_power_spectrum(cola_context* c, int n, int source, double kmin, double kmax, int nbins, int flags,
                        double* k_out, double* p_out, double* nmodes_out)
{
  if (!valid(c) || n < 2 || !k_out || !p_out || !nmodes_out) return kErrArg;
  if (nbins < 1 || nbins > kMaxShellBins || !(kmax > kmin) || !(kmin >= 0.0) || (flags & ~3)) {
    set_error("power spectrum: need 1 <= nbins <= 2048, 0 <= kmin < kmax and flags in {0..3}");
    return kErrArg;
  }
  PosSource src;
  COLA_TRY(position_source(c, source, &src));
  if (c->aux_fft_n != n) {
    c->aux_fft.destroy();
    c->aux_fft_n = 0;
    const int st = c->aux_fft.init(n, c->fft_log2_tile);
    if (st != kOk) {
      if (st == kErrUnsupported) set_error("power spectrum: the grid size must be even and of the form 2^a 3^b 5^c");
      return st;
    }
    c->aux_fft_n = n;
  }
  const long long row = 2LL * (n / 2 + 1);
  COLA_TRY(ensure_aux(c, (long long)n * n * row, (long long)n * n * n));
  COLA_TRY(deposit_contrast(c, src, n, row));
  COLA_TRY(c->aux_fft.r2c(c->aux_grid, c->stream));
  ShellBins b;
  b.kmin = kmin;
  b.dk = (kmax - kmin) / nbins;
  b.kf = 2.0 * kPiD / c->boxsize;
  const double vol = c->boxsize * c->boxsize * c->boxsize;
  const double nn = (double)n * n * n;
  b.norm = vol / (nn * nn);
  b.nbins = nbins;
  b.deconvolve = flags & COLA_PK_DECONVOLVE_CIC;
  double* acc = c->d_aux + 2;
  COLA_TRY(launch_fill_zero(acc, 3LL * nbins, c->stream));
  COLA_TRY(launch_power_shells(n, reinterpret_cast<const float2*>(c->aux_grid), b, acc, c->stream));
  std::vector<double> h(3 * (size_t)nbins);
  COLA_HIP(hipMemcpyAsync(h.data(), acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  COLA_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < nbins; ++i) {
    const double cnt = h[3 * i + 2];
    nmodes_out[i] = cnt;
    k_out[i] = cnt > 0 ? h[3 * i] / cnt : 0.0;
    p_out[i] = cnt > 0 ? h[3 * i + 1] / cnt : 0.0;
    if (cnt > 0 && (flags & COLA_PK_SUBTRACT_SHOT)) p_out[i] -= vol / (double)c->np;
  }
  return check_pending(c);
}

int cola_fof_find_halos(cola_context* c, float ll, int nfof_min, long long* nhalo_out)
{
  if (!valid(c) || !nhalo_out || nfof_min < 1) return kErrArg;
  if (!c->have_snapshot) {
    set_error("FoF runs on the snapshot: call cola_set_snapshot or cola_set_noncola_initial first");
    return kErrState;
  }
  FofInput in;
  in.base = reinterpret_cast<float*>(c->d_snap);
  in.stride = sizeof(DevParticleMin) / sizeof(float);
  in.np = c->np;
  in.boxsize = c->boxf;
  in.ll = ll;
  in.nfof_min = nfof_min;
  COLA_TRY(fof_run(in, &c->fof, c->stream));
  *nhalo_out = (long long)c->fof.halos.size();
  return check_pending(c);
}

int cola_fof_get_halos(cola_context* c, cola_halo* out, long long n)
{
  if (!valid(c) || (n > 0 && !out)) return kErrArg;
  if (n < (long long)c->fof.halos.size()) {
    set_error("halo buffer too small");
    return kErrArg;
  }
  std::copy(c->fof.halos.begin(), c->fof.halos.end(), out);
  return kOk;
}

int cola_fof_get_labels(cola_context* c, int* out, size_t n)
{
  if (!valid(c) || !out) return kErrArg;
  if (!c->fof.d_label || n < (size_t)c->fof.np) {
    set_error(c->fof.d_label ? "label buffer too small" : "no FoF result: call cola_fof_find_halos first");
    return c->fof.d_label ? kErrArg : kErrState;
  }
  COLA_HIP(hipMemcpyAsync(out, c->fof.d_label, sizeof(int) * c->fof.np, hipMemcpyDeviceToHost, c->stream));
  COLA_HIP(hipStreamSynchronize(c->stream));
  return kOk;
}

int cola_get_snapshot(cola_context* c, cola_particle_min* out, size_t np_out)
{
  if (!valid(c) || !out || np_out < (size_t)c->np) return kErrArg;
  if (!c->have_snapshot) {
    set_error("no snapshot: call cola_set_snapshot or cola_set_noncola_initial first");
    return kErrState;
  }
  COLA_HIP(hipMemcpyAsync(out, c->d_snap, sizeof(DevParticleMin) * c->np, hipMemcpyDeviceToHost, c->stream));
  COLA_HIP(hipStreamSynchronize(c->stream));
  return kOk;
}

int cola_set_snapshot_records(cola_context* c, const cola_particle_min* in, size_t np_in)
{
  if (!valid(c) || !in) return kErrArg;
  if (np_in != (size_t)c->np) {
    set_error("snapshot of " + std::to_string(np_in) + " records for a context of " + std::to_string(c->np) +
              " particles");
    return kErrArg;
  }
  if (!c->d_snap) COLA_HIP(hipMalloc(&c->d_snap, sizeof(DevParticleMin) * c->np));
  static_assert(sizeof(DevParticleMin) == sizeof(cola_particle_min), "snapshot record layout");
  COLA_HIP(hipMemcpyAsync(c->d_snap, in, sizeof(DevParticleMin) * c->np, hipMemcpyHostToDevice, c->stream));
  COLA_HIP(hipStreamSynchronize(c->stream));
  c->have_snapshot = true;
  return kOk;
}

int cola_copy_particles(cola_context* dst, cola_context* src)
{
  if (!valid(dst) || !valid(src) || dst->np != src->np) return kErrArg;
  COLA_TRY(require_particles(src));
  COLA_HIP(hipStreamSynchronize(src->stream));
  COLA_HIP(hipStreamSynchronize(dst->stream));
  COLA_HIP(hipMemcpyAsync(dst->pbuf, src->pbuf, sizeof(float) * 12 * (size_t)src->np, hipMemcpyDeviceToDevice,
                          dst->stream));
  dst->a_x = src->a_x;
  dst->a_v = src->a_v;
  dst->have_particles = true;
  dst->have_force = false;
  return kOk;
}

}  // extern "C"

namespace {

// Work on the net's own stream (host I/O, training, timing) starts only after
// the attached context's queued steps, which use the same activation buffers.
int net_quiesce(cola_resnet* h)
{
  if (h->attached) COLA_HIP(hipStreamSynchronize(h->attached->stream));
  COLA_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

}  // namespace

extern "C" {

int cola_resnet_nparam(void) { return kNetParams; }

int cola_resnet_create(int n, unsigned int seed, cola_resnet** out)
{
  if (!out) return kErrArg;
  *out = nullptr;
  cola_resnet* h = new cola_resnet;
  int st = hipStreamCreate(&h->stream) == hipSuccess ? kOk : kErrHip;
  if (!st) st = resnet_init(&h->r, n, seed);
  if (!st) st = hipStreamSynchronize(h->stream) == hipSuccess ? kOk : kErrHip;
  if (st) {
    resnet_free(&h->r);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return st;
  }
  *out = h;
  return kOk;
}

int cola_resnet_destroy(cola_resnet* h)
{
  if (!h) return kOk;
  if (cola_context* c = h->attached) {  // detach: the context's graph and steps must not see the net again
    (void)hipStreamSynchronize(c->stream);
    drop_step_graph(c);
    c->net = nullptr;
    c->net_handle = nullptr;
  }
  (void)hipStreamSynchronize(h->stream);
  resnet_free(&h->r);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return kOk;
}

int cola_resnet_get_params(cola_resnet* h, float* out)
{
  if (!h || !out) return kErrArg;
  COLA_TRY(net_quiesce(h));
  COLA_HIP(hipMemcpy(out, h->r.d_param, sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
  return kOk;
}

int cola_resnet_set_params(cola_resnet* h, const float* in)
{
  if (!h || !in) return kErrArg;
  COLA_TRY(net_quiesce(h));
  COLA_HIP(hipMemcpy(h->r.d_param, in, sizeof(float) * kNetParams, hipMemcpyHostToDevice));
  COLA_TRY(resnet_pack(&h->r, h->stream));
  COLA_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int cola_resnet_get_scales(cola_resnet* h, float* in_scale, float* out_scale)
{
  if (!h) return kErrArg;
  if (in_scale) *in_scale = h->r.in_scale;
  if (out_scale) *out_scale = h->r.out_scale;
  return kOk;
}

int cola_resnet_set_scales(cola_resnet* h, float in_scale, float out_scale)
{
  if (!h || !(out_scale > 0.0f) || !(in_scale > 0.0f)) return kErrArg;
  COLA_TRY(net_quiesce(h));
  h->r.in_scale = in_scale;
  h->r.out_scale = out_scale;
  h->r.scale_set = true;
  return kOk;
}

int cola_resnet_fit_scale(cola_resnet* h, int count)
{
  if (!h) return kErrArg;
  COLA_TRY(net_quiesce(h));
  return resnet_fit_scale(&h->r, count, h->stream);
}

namespace {

// host x [n^3][8] -> device input (bf16)
int upload_input(cola_resnet* h, const float* x)
{
  const long long nc = (long long)h->r.n * h->r.n * h->r.n;
  std::vector<unsigned short> xb(nc * 8);
  for (long long i = 0; i < nc * 8; ++i) {
    unsigned u;
    std::memcpy(&u, &x[i], 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    xb[i] = (unsigned short)(u >> 16);
  }
  COLA_HIP(hipMemcpy(h->r.d_x, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
  // the feature path's inputs: channel 0 as the feature grid, channel 1 of
  // the first cell as a
  std::vector<unsigned short> f(nc);
  for (long long i = 0; i < nc; ++i) f[i] = xb[i * 8];
  COLA_HIP(hipMemcpy(h->r.d_f1, f.data(), f.size() * 2, hipMemcpyHostToDevice));
  COLA_HIP(hipMemcpy(h->r.d_a, &x[1], sizeof(float), hipMemcpyHostToDevice));
  return kOk;
}

int download_output(cola_resnet* h, float* y)
{
  const long long nc = (long long)h->r.n * h->r.n * h->r.n;
  std::vector<float> yd(3 * nc);
  COLA_HIP(hipMemcpyAsync(yd.data(), h->r.d_y, sizeof(float) * 3 * nc, hipMemcpyDeviceToHost, h->stream));
  COLA_HIP(hipStreamSynchronize(h->stream));
  for (long long cell = 0; cell < nc; ++cell)
    for (int a = 0; a < 3; ++a) y[cell * 3 + a] = yd[a * nc + cell];
  return kOk;
}

}  // namespace

int cola_resnet_forward_host(cola_resnet* h, const float* x, float* y)
{
  if (!h || !x || !y) return kErrArg;
  COLA_TRY(net_quiesce(h));
  COLA_TRY(upload_input(h, x));
  COLA_TRY(resnet_forward(&h->r, h->stream));
  return download_output(h, y);
}

int cola_resnet_forward_features_host(cola_resnet* h, const float* x, float* y)
{
  if (!h || !x || !y) return kErrArg;
  COLA_TRY(net_quiesce(h));
  COLA_TRY(upload_input(h, x));
  COLA_TRY(resnet_forward_features(&h->r, h->r.d_f1, h->r.d_a, 0.0f, h->stream));
  return download_output(h, y);
}

int cola_resnet_set_restriction(cola_resnet* h, float side_weight)
{
  if (!h || !(side_weight >= 0.0f && side_weight <= 0.25f)) return kErrArg;
  h->r.restrict_side = side_weight;
  return kOk;
}

int cola_resnet_time_forward(cola_resnet* h, int reps, double* ms)
{
  if (!h || reps < 1 || !ms) return kErrArg;
  COLA_TRY(net_quiesce(h));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  COLA_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return kErrHip;
  }
  // the inference path: one-channel first layer from the feature grid
  auto fwd = [&]() { return resnet_forward_features(&h->r, h->r.d_f1, h->r.d_a, 0.0f, h->stream); };
  int st = fwd();  // warm
  if (!st && hipEventRecord(e0, h->stream) != hipSuccess) st = kErrHip;
  for (int i = 0; i < reps && !st; ++i) st = fwd();
  float t = 0.0f;
  if (!st && (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&t, e0, e1) != hipSuccess))
    st = kErrHip;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / reps;
  return st;
}

int cola_resnet_gradient_host(cola_resnet* h, const float* x, const float* target, double* loss, float* grad)
{
  if (!h || !x || !target || !loss || !grad) return kErrArg;
  const long long nc = (long long)h->r.n * h->r.n * h->r.n;
  COLA_TRY(net_quiesce(h));
  COLA_TRY(upload_input(h, x));
  std::vector<float> t(3 * nc);
  for (long long cell = 0; cell < nc; ++cell)
    for (int a = 0; a < 3; ++a) t[a * nc + cell] = target[cell * 3 + a] * h->r.out_scale;
  float* dt = nullptr;
  COLA_HIP(hipMalloc(&dt, sizeof(float) * 3 * nc));
  int st = hipMemcpy(dt, t.data(), sizeof(float) * 3 * nc, hipMemcpyHostToDevice) == hipSuccess ? kOk : kErrHip;
  if (!st) st = resnet_forward(&h->r, h->stream);
  if (!st) st = resnet_backward(&h->r, nullptr, dt, h->stream);
  if (!st && hipMemcpyAsync(loss, h->r.d_loss, sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    st = kErrHip;
  if (!st && hipMemcpyAsync(grad, h->r.d_grad, sizeof(float) * kNetParams, hipMemcpyDeviceToHost, h->stream) !=
                 hipSuccess)
    st = kErrHip;
  if (!st && hipStreamSynchronize(h->stream) != hipSuccess) st = kErrHip;
  (void)hipFree(dt);
  return st;
}

int cola_resnet_adam_step(cola_resnet* h, float lr)
{
  if (!h || !(lr > 0.0f)) return kErrArg;
  COLA_TRY(net_quiesce(h));
  COLA_TRY(resnet_adam(&h->r, lr, h->stream));
  COLA_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int cola_resnet_add_pair(cola_resnet* h, cola_context* coarse, cola_context* fine)
{
  if (!h || !valid(coarse) || !valid(fine)) return kErrArg;
  if (coarse->ng != h->r.n || fine->ng != 2 * coarse->ng || fine->np != coarse->np ||
      fine->boxsize != coarse->boxsize) {
    set_error("force pair: need coarse ngrid == net grid, fine ngrid == 2 coarse, same particles and box");
    return kErrArg;
  }
  COLA_TRY(require_particles(coarse));
  COLA_TRY(require_particles(fine));
  COLA_TRY(net_quiesce(h));
  ResNet* const saved = coarse->net;
  coarse->net = nullptr;  // the pair holds the plain coarse force
  int st = pm_assign(coarse, true);
  if (!st) st = hipStreamSynchronize(coarse->stream) == hipSuccess ? kOk : kErrHip;
  float* dcopy = nullptr;
  // keep the coarse real-space density: pm_fields transforms it in place
  if (!st && hipMalloc(&dcopy, sizeof(float) * coarse->grid_floats) != hipSuccess) st = kErrAlloc;
  if (!st && hipMemcpyAsync(dcopy, coarse->density, sizeof(float) * coarse->grid_floats, hipMemcpyDeviceToDevice,
                            coarse->stream) != hipSuccess)
    st = kErrHip;
  if (!st) st = pm_fields(coarse);
  coarse->net = saved;
  if (!st) st = pm_assign(fine, true);
  if (!st) st = pm_fields(fine);
  if (!st && (hipStreamSynchronize(fine->stream) != hipSuccess || hipStreamSynchronize(coarse->stream) != hipSuccess))
    st = kErrHip;
  if (!st) st = check_pending(coarse);
  if (!st) st = check_pending(fine);
  if (!st) {
    const float* const cf[3] = {coarse->fgrid[0], coarse->fgrid[1], coarse->fgrid[2]};
    const float* const ff[3] = {fine->fgrid[0], fine->fgrid[1], fine->fgrid[2]};
    st = resnet_store_pair(&h->r, dcopy, coarse->cic.row, cf, ff, fine->cic.row, coarse->a_x, h->stream);
  }
  if (!st && hipStreamSynchronize(h->stream) != hipSuccess) st = kErrHip;
  if (dcopy) (void)hipFree(dcopy);
  coarse->have_force = false;
  fine->have_force = false;
  return st;
}

int cola_resnet_num_samples(const cola_resnet* h) { return h ? (int)h->r.s_feat.size() : 0; }

int cola_resnet_get_sample(cola_resnet* h, int i, float* feature, float* target, float* a)
{
  if (!h || i < 0 || i >= (int)h->r.s_feat.size() || !feature || !target || !a) return kErrArg;
  const long long nc = (long long)h->r.n * h->r.n * h->r.n;
  std::vector<unsigned short> f(nc), t(3 * nc);
  COLA_TRY(net_quiesce(h));
  COLA_HIP(hipMemcpy(f.data(), h->r.s_feat[i], sizeof(unsigned short) * nc, hipMemcpyDeviceToHost));
  COLA_HIP(hipMemcpy(t.data(), h->r.s_target[i], sizeof(unsigned short) * 3 * nc, hipMemcpyDeviceToHost));
  auto widen = [](unsigned short b) {
    const unsigned u = (unsigned)b << 16;
    float v;
    std::memcpy(&v, &u, sizeof v);
    return v;
  };
  for (long long c = 0; c < nc; ++c) feature[c] = widen(f[c]);
  for (long long c = 0; c < 3 * nc; ++c) target[c] = widen(t[c]);
  *a = h->r.s_a[i];
  return kOk;
}

int cola_resnet_train(cola_resnet* h, int ntrain, int epochs, float lr, unsigned int seed, double* epoch_loss)
{
  if (!h) return kErrArg;
  COLA_TRY(net_quiesce(h));
  return resnet_train(&h->r, ntrain, epochs, lr, seed, epoch_loss, h->stream);
}

int cola_resnet_eval(cola_resnet* h, int first, int count, double* loss)
{
  if (!h || !loss) return kErrArg;
  COLA_TRY(net_quiesce(h));
  return resnet_eval(&h->r, first, count, loss, h->stream);
}

namespace {

constexpr char kNetMagic[8] = {'C', 'O', 'L', 'A', 'N', 'E', 'T', '1'};

}  // namespace

int cola_resnet_save(cola_resnet* h, const char* filename)
{
  if (!h || !filename) return kErrArg;
  std::vector<float> prm(kNetParams);
  COLA_TRY(net_quiesce(h));
  COLA_HIP(hipMemcpy(prm.data(), h->r.d_param, sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
  FILE* f = std::fopen(filename, "wb");
  if (!f) {
    set_error(std::string("cannot write net-weights file ") + filename);
    return 9000;
  }
  const int32_t head[2] = {kNetParams, 0};
  const float scales[4] = {h->r.in_scale, h->r.out_scale, h->r.restrict_side, 0.0f};
  const int32_t tail = kNetParams;
  bool ok = std::fwrite(kNetMagic, 1, 8, f) == 8 && std::fwrite(head, 4, 2, f) == 2 &&
            std::fwrite(scales, 4, 4, f) == 4 && std::fwrite(prm.data(), 4, kNetParams, f) == (size_t)kNetParams &&
            std::fwrite(&tail, 4, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_error(std::string("short write of net-weights file ") + filename);
    return 9000;
  }
  return kOk;
}

int cola_resnet_load(cola_resnet* h, const char* filename)
{
  if (!h || !filename) return kErrArg;
  FILE* f = std::fopen(filename, "rb");
  if (!f) {
    set_error(std::string("cannot open net-weights file ") + filename);
    return kErrNetFile;
  }
  char magic[8];
  int32_t head[2] = {0, 0}, tail = 0;
  float scales[4];
  std::vector<float> prm(kNetParams);
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kNetMagic, 8) == 0 &&
            std::fread(head, 4, 2, f) == 2 && head[0] == kNetParams && std::fread(scales, 4, 4, f) == 4 &&
            std::fread(prm.data(), 4, kNetParams, f) == (size_t)kNetParams && std::fread(&tail, 4, 1, f) == 1 &&
            tail == kNetParams && std::fgetc(f) == EOF;
  std::fclose(f);
  if (ok) ok = scales[0] > 0.0f && scales[1] > 0.0f && scales[2] >= 0.0f && scales[2] <= 0.25f;
  for (int i = 0; ok && i < kNetParams; ++i) ok = std::isfinite(prm[i]);
  if (!ok) {
    set_error(std::string("not a net-weights file of this net (") + std::to_string(kNetParams) +
              " parameters): " + filename);
    return kErrNetFile;
  }
  COLA_TRY(cola_resnet_set_params(h, prm.data()));
  COLA_TRY(cola_resnet_set_scales(h, scales[0], scales[1]));
  h->r.restrict_side = scales[2];
  return kOk;
}

int cola_resnet_set_fusion(cola_resnet* h, int on)
{
  if (!h || (on != 0 && on != 1)) return kErrArg;
  if (h->attached) {
    // a captured step holds the launches of the other path
    COLA_HIP(hipStreamSynchronize(h->attached->stream));
    drop_step_graph(h->attached);
  }
  h->r.fuse12 = on != 0;
  return kOk;
}

int cola_resnet_set_head_fusion(cola_resnet* h, int on)
{
  if (!h || (on != 0 && on != 1)) return kErrArg;
  if (h->attached) {
    COLA_HIP(hipStreamSynchronize(h->attached->stream));
    drop_step_graph(h->attached);
  }
  h->r.fuse34 = on != 0;
  return kOk;
}

int cola_resnet_set_plane_order(cola_resnet* h, int by_input_planes)
{
  if (!h || (by_input_planes != 0 && by_input_planes != 1)) return kErrArg;
  if (h->attached) {
    COLA_HIP(hipStreamSynchronize(h->attached->stream));
    drop_step_graph(h->attached);
  }
  h->r.roll = by_input_planes != 0;
  return kOk;
}

int cola_set_residual_net(cola_context* c, cola_resnet* h)
{
  if (!valid(c)) return kErrArg;
  if (h && h->r.n != c->ng) {
    set_error("residual net grid differs from the context's PM grid");
    return kErrArg;
  }
  if (h && h->attached && h->attached != c) {
    set_error("residual net is attached to another context (each context needs its own net: the activation "
              "buffers belong to the net)");
    return kErrState;
  }
  COLA_HIP(hipStreamSynchronize(c->stream));
  if (h) COLA_HIP(hipStreamSynchronize(h->stream));
  drop_step_graph(c);
  if (c->net_handle && c->net_handle != h) c->net_handle->attached = nullptr;
  c->net = h ? &h->r : nullptr;
  c->net_handle = h;
  if (h) h->attached = c;
  return kOk;
}

int cola_bandwidth_probe(size_t bytes, int reps, double* gbps)
{
  if (bytes < 16 || reps < 1 || !gbps) return kErrArg;
  const long long n = (long long)(bytes / 16);
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  COLA_HIP(hipMalloc(&a, 16 * n));
  if (hipMalloc(&b, 16 * n) != hipSuccess) {
    (void)hipFree(a);
    return kErrAlloc;
  }
  int st = kOk;
  if (hipMemset(a, 0, 16 * n) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    st = kErrHip;
  if (!st) {
    const unsigned blocks = 256 * 8 * 4;
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e1, nullptr);
    float ms = 0;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = kErrHip;
    *gbps = 2.0 * 16.0 * (double)n * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return st;
}

int cola_set_fft_xcd_group(cola_context* c, int group)
{
  if (!valid(c) || group < 1 || group > 64 || (group & (group - 1))) return kErrArg;
  COLA_HIP(hipStreamSynchronize(c->stream));
  drop_step_graph(c);
  c->pm_fft.xgroup = group;
  c->lpt_fft.xgroup = group;
  return kOk;
}

int cola_set_readout_group(cola_context* c, int group)
{
  if (!valid(c) || group < 1 || group > 4096) return kErrArg;
  COLA_HIP(hipStreamSynchronize(c->stream));
  drop_step_graph(c);
  c->cic.readout_group = group;
  return kOk;
}

static int standalone_fft(int n, float* data, bool forward)
{
  if (n <= 0 || n % 2 || !data) return kErrArg;
  FftPlan plan;
  COLA_TRY(plan.init(n));
  const size_t floats = (size_t)n * n * 2 * (n / 2 + 1);
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * floats) != hipSuccess) {
    plan.destroy();
    return kErrAlloc;
  }
  int st = hipMemcpy(d, data, sizeof(float) * floats, hipMemcpyHostToDevice) == hipSuccess ? kOk : kErrHip;
  if (!st) st = forward ? plan.r2c(d, nullptr) : plan.c2r(nullptr, d, KOp::None, KOpParams{}, nullptr);
  if (!st && hipDeviceSynchronize() != hipSuccess) st = kErrHip;
  if (!st && hipMemcpy(data, d, sizeof(float) * floats, hipMemcpyDeviceToHost) != hipSuccess) st = kErrHip;
  (void)hipFree(d);
  plan.destroy();
  return st;
}

__global__ void k_fill_pattern(float* d, long long n, unsigned seed)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)i * 2654435761u ^ seed;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  d[i] = (float)(h & 0xffffff) * (1.0f / 16777216.0f) - 0.5f;
}

int cola_fft_time(int n, int reps, double* ms_out)
{
  if (n <= 0 || n % 2 || reps < 1 || !ms_out) return kErrArg;
  FftPlan plan;
  COLA_TRY(plan.init(n, -1, -1, true));  // the PM layout: rows padded to 128 bytes
  const long long floats = (long long)n * n * 2 * plan.nzp;
  float* buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // work, pristine copy, 3 force grids
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = kOk;
  for (int b = 0; b < 5 && !st; ++b)
    if (hipMalloc(&buf[b], sizeof(float) * floats) != hipSuccess) st = kErrAlloc;
  if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = kErrHip;
  if (!st) {
    hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((floats + 255) / 256)), dim3(256), 0, nullptr, buf[1], floats,
                       12345u);
    KOpParams kp;
    kp.f1 = 1.0f;
    float* const out[3] = {buf[2], buf[3], buf[4]};
    double tot[3] = {0, 0, 0};
    for (int r = 0; r <= reps && !st; ++r) {  // r = 0 warms up
      for (int which = 0; which < 3 && !st; ++which) {
        if (hipMemcpyAsync(buf[0], buf[1], sizeof(float) * floats, hipMemcpyDeviceToDevice, nullptr) != hipSuccess) {
          st = kErrHip;
          break;
        }
        (void)hipEventRecord(e0, nullptr);
        if (which == 0)
          st = plan.r2c(buf[0], nullptr);
        else if (which == 1)
          st = plan.c2r(nullptr, buf[0], KOp::None, KOpParams{}, nullptr);
        else
          st = plan.forces(buf[0], out, kp, nullptr);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0;
        if (!st && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          st = kErrHip;
        if (r > 0) tot[which] += ms;
      }
    }
    for (int w = 0; w < 3; ++w) ms_out[w] = tot[w] / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (float* b : buf)
    if (b) (void)hipFree(b);
  plan.destroy();
  return st;
}

int cola_fft_r2c_3d(int n, float* data) { return standalone_fft(n, data, true); }
int cola_fft_c2r_3d(int n, float* data) { return standalone_fft(n, data, false); }

}  // extern "C"
