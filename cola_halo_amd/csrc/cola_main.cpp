// cola_halo: the reference's driver program (main.c:44-230) on one MI355X,
// over the library's C ABI.
//
//     cola_halo param.lua
//
// Reads the same parameter file (read_param.c syntax), runs nrealization
// realizations of 2LPT + COLA with the linear-in-a schedule, and writes the
// same outputs at each requested redshift: FoF halo catalogue
// (fof%05d%c.txt), GADGET-1 snapshot (snp%05d%c.0), random subsample
// (sub%05d%c.b), coarse density grid (grid%05d%c.b), and the optional 2LPT
// initial condition (init_%05d.0).  Two keys of this build extend the file:
// pm_ngrid (the PM grid when nc * pm_nc_factor cannot express it) and
// residual_net (a net-weights file: the learned force residual is added to
// every force).  A fatal error prints the message and
// exits with the reference's abort code (as MPI_Abort does, the shell sees its
// low 8 bits).  Steps with no output in them run as one fused call
// (cola_step).
#include <sys/time.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/cola_halo_amd.h"

namespace {

int g_loglevel = 0;
enum Level { kDebug = 0, kVerbose = 1, kInfo = 2, kNormal = 3 };

// msg.c:22-33: print when the message level is at least the log level.
void msg(int level, const char* fmt, ...)
{
  if (level < g_loglevel) return;
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stdout, fmt, ap);
  va_end(ap);
  std::fflush(stdout);
}

[[noreturn]] void abort_with(int code)
{
  std::fprintf(stderr, "%s\n", cola_last_error());
  std::fflush(stderr);
  std::exit(code);
}

void check(int code)
{
  if (code != 0) abort_with(code);
}

double now()
{
  timeval tp;
  gettimeofday(&tp, nullptr);
  return tp.tv_sec + tp.tv_usec * 1e-6;
}

// timer.c categories: Init, 2LPT, COLA, Snapshot (device work is drained at
// every category switch so wall time lands in the right one).
struct Timer {
  static constexpr int kN = 4;
  const char* names[kN] = {"Init", "2LPT", "COLA", "Snapshot"};
  double t[kN] = {0, 0, 0, 0};
  int cat = 0;
  double since = now();
  cola_context* ctx = nullptr;
  void set(int c)
  {
    if (ctx) check(cola_synchronize(ctx));
    const double n = now();
    t[cat] += n - since;
    since = n;
    cat = c;
  }
  void print()
  {
    set(cat);
    double total = 0;
    for (double v : t) total += v;
    for (int i = 0; i < kN; ++i)
      msg(kInfo, "%-16s %7.2f   %4.1f%%\n", names[i], t[i], total > 0 ? 100.0 * t[i] / total : 0.0);
    msg(kInfo, "----------------------------------\n");
    msg(kInfo, "%-16s %7.2f sec\n", "Total", total);
  }
};

std::string fmt(const char* f, const char* base, int seed, char suffix)
{
  char buf[4096];
  std::snprintf(buf, sizeof(buf), f, base, seed, suffix);
  return buf;
}

void print_parameters(const cola_params& p, const char* filename)
{
  // confirm_param.c:10-130 listing
  msg(kNormal, "parameter file %s read.\n", filename);
  msg(kNormal, "nc = %d\n", p.nc);
  msg(kNormal, "pm_nc_factor = %d\n", p.pm_nc_factor);
  msg(kNormal, "boxsize = %lg\n", p.boxsize);
  msg(kNormal, "a_final = %lg\n", p.a_final);
  msg(kNormal, "ntimestep = %d\n", p.ntimestep);
  if (p.n_zout > 0) {
    msg(kNormal, "output_redshift = {%lg", p.zout[0]);
    for (int i = 1; i < p.n_zout; i++) msg(kNormal, ", %lg", p.zout[i]);
    msg(kNormal, "}\n");
  }
  msg(kNormal, "random_seed = %d\n", p.random_seed);
  msg(kNormal, "nrealization = %d\n", p.nrealization);
  msg(kNormal, "omega_m = %lg\n", p.omega_m);
  msg(kNormal, "h = %lg\n", p.h);
  msg(kNormal, "sigma8 (check) = %lg\n", p.sigma8);
  msg(kNormal, "np_alloc_factor = %lg\n", p.np_alloc_factor);
  msg(kNormal, "loglevel = %d\n", p.loglevel);
  if (p.strlen_power_spectrum_filename) msg(kNormal, "powerspectrum = \"%s\"\n", p.power_spectrum_filename);
  if (p.strlen_fof_filename)
    msg(kNormal, "fof_filename = \"%s\"\nlinking_factor = %lg\n", p.fof_filename, p.fof_linking_factor);
  else
    msg(kNormal, "No FoF halo finding. fof_filename is not given.\n");
  if (p.strlen_snapshot_filename)
    msg(kNormal, "snapshot = \"%s\"\nwrite_longid = %d\n", p.snapshot_filename, p.write_longid);
  else
    msg(kNormal, "No N-body particle output (snapshot)\n");
  if (p.strlen_subsample_filename)
    msg(kNormal, "subsample = \"%s\"\nsubsample_factor = %lg\n", p.subsample_filename, p.subsample_factor);
  else
    msg(kNormal, "No subsampled particle output (subsample)\n");
  if (p.strlen_cgrid_filename)
    msg(kNormal, "coarse_grid = \"%s\"\ncoarse_grid_nc = %d\n", p.cgrid_filename, p.cgrid_nc);
  else
    msg(kNormal, "No coarse density grid output (coarse_grid)\n");
  if (p.strlen_init_filename) msg(kNormal, "inital= \"%s\"\n", p.init_filename);
  if (p.pm_ngrid > 0) msg(kNormal, "pm_ngrid = %d\n", p.pm_ngrid);
  if (p.strlen_residual_net_filename) msg(kNormal, "residual_net = \"%s\"\n", p.residual_net_filename);
}

struct Run {
  const cola_params& p;
  cola_context* ctx;
  cola_subsampler* sub;
  Timer& timer;
  std::vector<cola_particle_min> snap;
  std::vector<float> grid;
  std::vector<cola_halo> halos;
  float om, boxf, hf;
  int seed = 0;

  // main.c:326-392 snapshot_time
  void snapshot_time(float aout, int iout)
  {
    timer.set(3);
    const char suffix = (char)('a' + iout);
    check(cola_set_snapshot(ctx, aout, om, snap.data(), snap.size()));
    const float a_snap = aout;  // Snapshot.a is a float (cola.c:412)
    if (p.strlen_fof_filename) {
      const float ll = (float)(p.fof_linking_factor * boxf / p.nc);  // main.c:353
      long long nh = 0;
      check(cola_fof_find_halos(ctx, ll, 32, &nh));
      halos.resize(nh);
      check(cola_fof_get_halos(ctx, halos.data(), nh));
      const std::string f = fmt("%s%05d%c.txt", p.fof_filename, seed, suffix);
      check(cola_write_halos(f.c_str(), halos.data(), nh));
      msg(kNormal, "%lld halos written to %s\n", nh, f.c_str());
      // the records are wrapped into [0, L) by the halo finder, as in the
      // reference, before the snapshot and the subsample are written
      check(cola_get_snapshot(ctx, snap.data(), snap.size()));
    }
    if (p.strlen_snapshot_filename) {
      const std::string f = fmt("%s%05d%c", p.snapshot_filename, seed, suffix) + ".0";  // write.c:19-20
      check(cola_write_gadget(f.c_str(), snap.data(), (long long)snap.size(), a_snap, boxf, om, hf,
                              p.write_longid));
      msg(kNormal, "snapshot %s written\n", f.c_str());
    }
    if (p.strlen_subsample_filename) {
      const std::string f = fmt("%s%05d%c.b", p.subsample_filename, seed, suffix);
      long long n = 0;
      check(cola_write_subsample(sub, f.c_str(), snap.data(), (long long)snap.size(), a_snap, boxf, om, hf, &n));
      msg(kNormal, "subsample binary %s written\n", f.c_str());
    }
    if (p.strlen_cgrid_filename) {
      const std::string f = fmt("%s%05d%c.b", p.cgrid_filename, seed, suffix);
      const long long n3 = (long long)p.cgrid_nc * p.cgrid_nc * p.cgrid_nc;
      grid.resize(n3);
      check(cola_coarse_grid(ctx, p.cgrid_nc, COLA_SOURCE_SNAPSHOT, grid.data(), nullptr));
      check(cola_write_coarse_grid(f.c_str(), boxf, p.cgrid_nc, grid.data()));
      msg(kVerbose, "Coarse grid file written %s, nc=%d\n", f.c_str(), p.cgrid_nc);
    }
    msg(kNormal, "snapshot %d (%c) written z=%4.2f a=%5.3f\n", iout + 1, suffix, 1.0 / aout - 1.0, aout);
    timer.set(2);
  }
};

}  // namespace

int main(int argc, char* argv[])
{
  if (argc < 2) {
    std::fprintf(stderr, "Error: Parameter file not specified. cola_code param.lua\n");
    return 1;
  }
  const char* filename = argv[argc - 1];
  cola_params p;
  check(cola_read_parameters(filename, &p));
  g_loglevel = p.loglevel;
  print_parameters(p, filename);
  check(cola_confirm_parameters(&p));

  Timer timer;
  const int nc_factor = p.pm_nc_factor;
  const double OmegaM = p.omega_m, OmegaLambda = 1.0 - OmegaM;
  // pm_ngrid (this build's key) sets the PM grid directly: factors such as
  // 1.5 that pm_nc_factor (an int, parameters.h:6) cannot express
  const int ngrid = p.pm_ngrid > 0 ? p.pm_ngrid : nc_factor * p.nc;
  cola_context* ctx = nullptr;
  check(cola_context_create(p.nc, ngrid, p.boxsize, &ctx));
  timer.ctx = ctx;
  cola_resnet* net = nullptr;
  if (p.strlen_residual_net_filename) {
    // learned force residual on the PM grid (no reference counterpart)
    check(cola_resnet_create(ngrid, 1, &net));
    check(cola_resnet_load(net, p.residual_net_filename));
    check(cola_set_residual_net(ctx, net));
    msg(kNormal, "learned force residual %s on the %d^3 PM grid\n", p.residual_net_filename, ngrid);
  }
  const int nsteps = p.ntimestep;
  const double a_final = p.a_final;
  const double da = a_final / nsteps;  // main.c:70-72
  const double a_init = da;
  check(cola_power_init(ctx, p.power_spectrum_filename, a_init, p.sigma8, OmegaM, OmegaLambda));
  cola_subsampler* sub = nullptr;
  check(cola_subsampler_create(p.subsample_factor, (unsigned)p.random_seed, &sub));

  const int nout = p.n_zout;
  std::vector<float> aout(nout);
  for (int i = 0; i < nout; i++) {
    aout[i] = (float)(1.0 / (1 + p.zout[i]));  // main.c:107-110
    msg(kVerbose, "zout[%d]= %lf, aout= %f\n", i, p.zout[i], aout[i]);
  }

  Run run{p, ctx, sub, timer, std::vector<cola_particle_min>((size_t)cola_np(ctx)), {}, {},
          (float)OmegaM, (float)p.boxsize, (float)p.h};
  const float Om = (float)OmegaM;
  for (int ireal = 0; ireal < p.nrealization; ireal++) {
    msg(kVerbose, "\n%d/%d realization.\n", ireal + 1, p.nrealization);
    const int seed = p.random_seed + ireal;
    run.seed = seed;
    int iout = 0;
    timer.set(1);
    check(cola_lpt_set_displacement(ctx, a_init, OmegaM, seed, p.boxsize));
    if (p.strlen_init_filename) {
      // main.c:131-137: the 2LPT state as a GADGET file at a_init
      char base[4096];
      std::snprintf(base, sizeof(base), "%s_%05d", p.init_filename, seed);
      check(cola_set_noncola_initial(ctx, Om, run.snap.data(), run.snap.size()));
      float a_x = 0, a_v = 0;
      check(cola_get_times(ctx, &a_x, &a_v));
      const std::string f = std::string(base) + ".0";
      check(cola_write_gadget(f.c_str(), run.snap.data(), (long long)run.snap.size(), a_x, run.boxf, Om, run.hf,
                              p.write_longid));
    }
    timer.set(2);
    float a_x = 0, a_v = 0;
    check(cola_get_times(ctx, &a_x, &a_v));
    check(cola_set_times(ctx, a_x, (float)da));  // main.c:145
    if (nout > 0 && nsteps > 1 && a_final > a_init) {
      msg(kNormal, "Time integration a= %g -> %g, %d steps\n", a_init, a_final, nsteps);
      for (int istep = 1; istep <= nsteps; istep++) {
        msg(kNormal, "Timestep %d/%d\n", istep, nsteps);
        const double avel0 = (istep - 0.5) * da, apos0 = istep * da;  // main.c:174-178
        const double avel1 = (istep + 0.5) * da, apos1 = (istep + 1.0) * da;
        const bool out_now = iout < nout && avel0 <= aout[iout] && aout[iout] <= apos0;
        const bool out_mid = iout < nout && apos0 < aout[iout] && aout[iout] <= avel1;
        if (!out_now && !out_mid && iout < nout) {
          check(cola_step(ctx, Om, (float)avel1, (float)apos1));
          continue;
        }
        check(cola_move_particles(ctx));
        check(cola_pm_calculate_forces(ctx));
        while (iout < nout && avel0 <= aout[iout] && aout[iout] <= apos0) {
          run.snapshot_time(aout[iout], iout);
          iout++;
        }
        if (iout >= nout) break;
        check(cola_kick(ctx, Om, (float)avel1));
        while (iout < nout && apos0 < aout[iout] && aout[iout] <= avel1) {
          run.snapshot_time(aout[iout], iout);
          iout++;
        }
        if (iout >= nout) break;
        check(cola_drift(ctx, Om, (float)apos1));
      }
    }
    timer.print();
  }
  check(cola_subsampler_destroy(sub));
  check(cola_context_destroy(ctx));
  if (net) check(cola_resnet_destroy(net));
  return 0;
}
