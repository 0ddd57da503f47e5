// halo: the reference's post-processing program (halo_main.c:51-185) on one
// MI355X, over the library's C ABI.
//
//     halo param.lua
//
// For every directory listed in dirs.txt, reads the GADGET-1 snapshots
// <dir>/<snapshot>_000, _001, ... (one file or .0, .1, ...; read.c) until one
// is missing, and writes for each the random subsample
// (<dir>/<subsample>_NNN.b), the coarse density grid
// (<dir>/<coarse_grid>_NNN.b) when the parameter file names them, and the
// FoF halo catalogue <dir>/fof_NNN.txt (linking length 0.2 L / nc).  A fatal
// error prints the message and exits with the reference's abort code.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cola_halo_amd.h"

namespace {

int g_loglevel = 0;
enum Level { kDebug = 0, kVerbose = 1, kInfo = 2, kNormal = 3 };

void msg(int level, const char* fmt, ...)
{
  if (level < g_loglevel) return;
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stdout, fmt, ap);
  va_end(ap);
  std::fflush(stdout);
}

[[noreturn]] void abort_with(int code, const char* text = nullptr)
{
  std::fprintf(stderr, "%s\n", text ? text : cola_last_error());
  std::fflush(stderr);
  std::exit(code);
}

void check(int code)
{
  if (code != 0) abort_with(code);
}

std::string numbered(const std::string& dir, const char* name, int iout, const char* ext)
{
  char buf[4096];
  std::snprintf(buf, sizeof(buf), "%s/%s_%03d%s", dir.c_str(), name, iout, ext);
  return buf;
}

}  // namespace

int main(int argc, char* argv[])
{
  if (argc < 2) {
    std::fprintf(stderr, "Error: Parameter file not specified. halo param.lua\n");
    return 1;
  }
  cola_params p;
  check(cola_read_parameters(argv[argc - 1], &p));
  g_loglevel = p.loglevel;
  check(cola_confirm_parameters(&p));
  if (!p.strlen_snapshot_filename) abort_with(1, "filename snapshot is not set in the parameter file");

  const int nc = p.nc;
  constexpr int kNcFactor = 3;  // halo_main.c:41
  cola_context* ctx = nullptr;
  check(cola_context_create(nc, kNcFactor * nc, p.boxsize, &ctx));
  cola_subsampler* sub = nullptr;
  check(cola_subsampler_create(p.subsample_factor, (unsigned)p.random_seed, &sub));
  std::vector<cola_particle_min> snap((size_t)cola_np(ctx));
  std::vector<float> grid;
  std::vector<cola_halo> halos;

  FILE* fp_dir = std::fopen("dirs.txt", "r");
  if (!fp_dir) abort_with(0, "Unable to open list of directories dirs.txt");
  char line[256];
  while (std::fgets(line, 255, fp_dir)) {
    const size_t last = std::strlen(line);
    if (last > 0 && line[last - 1] == '\n') line[last - 1] = '\0';
    if (line[0] == '\0') break;
    const std::string dir = line;
    std::printf("Enter directory: %s\n", dir.c_str());
    for (int iout = 0; iout < 100000; ++iout) {
      const std::string base = numbered(dir, p.snapshot_filename, iout, "");
      std::printf("Read %s\n", base.c_str());
      cola_gadget_info info;
      check(cola_read_gadget(base.c_str(), nullptr, 0, &info));  // headers: the particle count
      if (info.nfiles == 0) {
        if (iout == 0) abort_with(2, ("Unable to read snapshot " + base).c_str());
        break;
      }
      if (info.np_read != (long long)snap.size()) {
        std::fprintf(stderr, "Error: snapshot %s has %lld particles, nc^3 = %zu\n", base.c_str(), info.np_read,
                     snap.size());
        std::exit(15002);  // read.c:127-129
      }
      check(cola_read_gadget(base.c_str(), snap.data(), (long long)snap.size(), &info));
      check(cola_set_snapshot_records(ctx, snap.data(), snap.size()));
      const float boxsize = info.boxsize;
      if (p.strlen_subsample_filename) {
        const std::string f = numbered(dir, p.subsample_filename, iout, ".b");
        check(cola_write_subsample(sub, f.c_str(), snap.data(), (long long)snap.size(), info.a, boxsize,
                                   info.omega_m, info.h, nullptr));
      }
      if (p.strlen_cgrid_filename) {
        const std::string f = numbered(dir, p.cgrid_filename, iout, ".b");
        grid.resize((size_t)p.cgrid_nc * p.cgrid_nc * p.cgrid_nc);
        check(cola_coarse_grid(ctx, p.cgrid_nc, COLA_SOURCE_SNAPSHOT, grid.data(), nullptr));
        check(cola_write_coarse_grid(f.c_str(), boxsize, p.cgrid_nc, grid.data()));
      }
      const float ll = (float)(0.2 * boxsize / nc);  // halo_main.c:162
      long long nh = 0;
      check(cola_fof_find_halos(ctx, ll, 32, &nh));
      halos.resize(nh);
      check(cola_fof_get_halos(ctx, halos.data(), nh));
      char fof[4096];
      std::snprintf(fof, sizeof(fof), "%s/fof_%03d.txt", dir.c_str(), iout);
      check(cola_write_halos(fof, halos.data(), nh));
      msg(kNormal, "%lld halos written to %s\n", nh, fof);
    }
  }
  std::fclose(fp_dir);
  msg(kNormal, "Processed all directories. halo done.\n");
  check(cola_subsampler_destroy(sub));
  check(cola_context_destroy(ctx));
  return 0;
}
