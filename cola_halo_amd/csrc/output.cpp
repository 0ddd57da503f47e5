// Output files of the driver: GADGET-1 snapshots (write.c:14-107) and the
// coarse density grid (coarse_grid.c:46-91).  Host only; the records come
// from the device in lattice-ID order.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <new>
#include <vector>

#include "../../include/cola_halo_amd.h"
#include "errors.h"
#include "ranlux.h"

namespace {

using cola::set_error;

// GADGET-1 header, 256 bytes (the file format of gadget_file.h:4-23).
struct GadgetHeader {
  int32_t np[6];
  double mass[6];
  double time;
  double redshift;
  int32_t flag_sfr;
  int32_t flag_feedback;
  uint32_t np_total[6];
  int32_t flag_cooling;
  int32_t num_files;
  double boxsize;
  double omega0;
  double omega_lambda;
  double hubble_param;
  int32_t flag_stellarage;
  int32_t flag_metals;
  uint32_t np_total_highword[6];
  int32_t flag_entropy_instead_u;
  char fill[60];
};
static_assert(sizeof(GadgetHeader) == 256, "GADGET-1 header is 256 bytes");

class File {
 public:
  explicit File(const char* name) : fp_(std::fopen(name, "wb")) {}
  ~File()
  {
    if (fp_) std::fclose(fp_);
  }
  bool ok() const { return fp_ != nullptr && good_; }
  void put(const void* data, size_t bytes)
  {
    if (fp_ && good_ && bytes > 0) good_ = std::fwrite(data, 1, bytes, fp_) == bytes;
  }
  void marker(int32_t len) { put(&len, sizeof(len)); }
  bool close()
  {
    const bool closed = fp_ && std::fclose(fp_) == 0;
    fp_ = nullptr;
    return closed && good_;
  }

 private:
  FILE* fp_;
  bool good_ = true;
};

// Stream one record block in bounded chunks through a staging buffer.
template <typename T, typename F>
void put_block(File& f, long long np, int per_particle, F fill)
{
  constexpr long long kChunk = 1 << 16;
  std::vector<T> buf(kChunk * per_particle);
  for (long long i0 = 0; i0 < np; i0 += kChunk) {
    const long long n = std::min(kChunk, np - i0);
    for (long long i = 0; i < n; ++i) fill(i0 + i, &buf[i * per_particle]);
    f.put(buf.data(), sizeof(T) * n * per_particle);
  }
}

}  // namespace

extern "C" {

int cola_write_gadget(const char* filename, const cola_particle_min* p, long long np, float a, float boxsize_f,
                      float omega_m_f, float h_f, int use_long_id)
{
  if (!filename || (np > 0 && !p) || np < 0) return COLA_ERR_ARG;
  // The record-length markers are 4-byte ints (write.c:61-100).
  const long long id_bytes = use_long_id ? 8 : 4;
  if (np * 12 > INT32_MAX || np * id_bytes > INT32_MAX) {
    set_error("GADGET-1 record longer than 2^31 bytes: too many particles for one file");
    return COLA_ERR_ARG;
  }
  File f(filename);
  if (!f.ok()) {
    set_error(std::string("Error: Unable to write to file: ") + filename);
    return 9000;
  }
  const double h = h_f;
  const double boxsize = boxsize_f;
  const double omega_m = omega_m_f;
  const long long np_total = np;  // one rank

  GadgetHeader header;
  std::memset(&header, 0, sizeof(header));
  const double rho_crit = 27.7455;  // write.c:45
  const double m = omega_m * rho_crit * std::pow(boxsize, 3.0) / np_total;
  header.np[1] = (int32_t)np;
  header.mass[1] = m;
  header.time = a;
  header.redshift = 1.0 / header.time - 1;
  header.np_total[1] = (uint32_t)np_total;
  header.np_total_highword[1] = (uint32_t)(np_total >> 32);
  header.num_files = 1;
  header.boxsize = boxsize;
  header.omega0 = omega_m;
  header.omega_lambda = 1.0 - omega_m;
  header.hubble_param = h;

  f.marker(sizeof(GadgetHeader));
  f.put(&header, sizeof(header));
  f.marker(sizeof(GadgetHeader));

  const int32_t blk_vec = (int32_t)(np * 12);
  f.marker(blk_vec);
  put_block<float>(f, np, 3, [&](long long i, float* o) {
    for (int k = 0; k < 3; ++k) o[k] = p[i].x[k];
  });
  f.marker(blk_vec);

  const float vfac = 1.0 / std::sqrt((double)a);  // write.c:74 (double sqrt), GADGET velocity convention
  f.marker(blk_vec);
  put_block<float>(f, np, 3, [&](long long i, float* o) {
    for (int k = 0; k < 3; ++k) o[k] = vfac * p[i].v[k];
  });
  f.marker(blk_vec);

  const int32_t blk_id = (int32_t)(np * id_bytes);
  f.marker(blk_id);
  if (use_long_id)
    put_block<unsigned long long>(f, np, 1, [&](long long i, unsigned long long* o) { *o = p[i].id; });
  else
    put_block<unsigned int>(f, np, 1, [&](long long i, unsigned int* o) { *o = (unsigned int)p[i].id; });
  f.marker(blk_id);

  if (!f.close()) {
    set_error(std::string("Error: write failed: ") + filename);
    return 9000;
  }
  return COLA_OK;
}

namespace {

// A GADGET-1 file being read: 4-byte record markers around each block.
class Reader {
 public:
  explicit Reader(const std::string& name) : fp_(std::fopen(name.c_str(), "rb")) {}
  ~Reader()
  {
    if (fp_) std::fclose(fp_);
  }
  bool ok() const { return fp_ != nullptr; }
  bool get(void* data, size_t bytes) { return bytes == 0 || std::fread(data, 1, bytes, fp_) == bytes; }
  // read.c:16-27 check_separator: the marker must equal the block length
  int separator(long long expected)
  {
    int32_t v = -1;
    if (!get(&v, sizeof(v)) || (long long)v != expected) {
      set_error("Error: Unable to read snapshot correctly. Separator " + std::to_string(v) +
                " != " + std::to_string(expected));
      return 15005;
    }
    return COLA_OK;
  }

 private:
  FILE* fp_;
};

int read_header(Reader& r, GadgetHeader* h)
{
  int st = r.separator(256);
  if (st) return st;
  if (!r.get(h, sizeof(*h))) {
    set_error("Error: Unable to read snapshot header");
    return 15005;
  }
  return r.separator(256);
}

// read.c:171-201 find_files: the file count of a snapshot written as one
// file `base` or as base.0, base.1, ... (0 when neither exists).
int find_files(const std::string& base, int* nfiles)
{
  *nfiles = 0;
  for (const std::string& name : {base, base + ".0"}) {
    Reader r(name);
    if (!r.ok()) continue;
    GadgetHeader h;
    const int st = read_header(r, &h);
    if (st) return st;
    *nfiles = h.num_files;
    return COLA_OK;
  }
  return COLA_OK;
}

}  // namespace

extern "C" {

int cola_read_gadget(const char* filename, cola_particle_min* out, long long np_out, cola_gadget_info* info)
{
  if (!filename || !info || np_out < 0 || (np_out > 0 && !out)) return COLA_ERR_ARG;
  std::memset(info, 0, sizeof(*info));
  const std::string base = filename;
  int nfiles = 0;
  int st = find_files(base, &nfiles);
  if (st) return st;
  if (nfiles <= 0) return COLA_OK;  // read.c:46-47: not found
  info->nfiles = nfiles;
  GadgetHeader h;
  long long nread = 0;
  for (int num = 0; num < nfiles; ++num) {
    const std::string name = nfiles > 1 ? base + "." + std::to_string(num) : base;
    Reader r(name);
    if (!r.ok()) {
      set_error("Error: can't open snapshot file " + name + " (read.c)");
      return 15000;
    }
    if ((st = read_header(r, &h))) return st;
    const long long np = h.np[1];  // type-1 particles only (read.c:75)
    if (np < 0) {
      set_error("Error: negative particle count in the header of " + name);
      return 15005;
    }
    if (!out) {  // counting pass: headers only
      nread += np;
      continue;
    }
    if (nread + np > np_out) {
      set_error("Error: Not enough space to read snapshot. " + std::to_string(np) + " particles in " + name);
      return 15001;
    }
    const float vfac_back = std::sqrt(h.time);  // read.c:76, GADGET convention back to km/s
    const float boxsize = (float)h.boxsize;
    std::vector<float> buf;
    try {
      buf.resize(3 * (size_t)np);
    } catch (const std::bad_alloc&) {
      set_error("Error: unable to allocate the read buffer for " + name);
      return COLA_ERR_ALLOC;
    }
    if ((st = r.separator(12 * np))) return st;
    if (!r.get(buf.data(), sizeof(float) * buf.size())) {
      set_error("Error: short position block in " + name);
      return 15005;
    }
    if ((st = r.separator(12 * np))) return st;
    for (long long i = 0; i < np; ++i) {
      cola_particle_min& q = out[nread + i];
      for (int k = 0; k < 3; ++k) {
        float x = buf[3 * i + k];
        if (x < 0.0f) x += boxsize;  // read.c:99-102 periodic wrap
        if (x >= boxsize) x -= boxsize;
        q.x[k] = x;
      }
      q.id = 0;  // read.c:137
    }
    if ((st = r.separator(12 * np))) return st;
    if (!r.get(buf.data(), sizeof(float) * buf.size())) {
      set_error("Error: short velocity block in " + name);
      return 15005;
    }
    if ((st = r.separator(12 * np))) return st;
    for (long long i = 0; i < np; ++i)
      for (int k = 0; k < 3; ++k) out[nread + i].v[k] = vfac_back * buf[3 * i + k];
    nread += np;
  }
  // read.c:149-156 (the last file's header)
  info->np_read = nread;
  info->np_total = ((long long)h.np_total_highword[1] << 32) + (long long)h.np_total[1];
  info->a = (float)h.time;
  info->boxsize = (float)h.boxsize;
  info->omega_m = (float)h.omega0;
  info->h = (float)h.hubble_param;
  return COLA_OK;
}

}  // extern "C"

struct cola_subsampler {
  cola::RanluxStream* rng = nullptr;
  double factor = 0.0;
};

int cola_subsampler_create(double factor, unsigned int seed, cola_subsampler** out)
{
  if (!out || !(factor >= 0.0)) return COLA_ERR_ARG;
  cola_subsampler* s = new cola_subsampler;
  s->factor = factor;
  s->rng = cola::ranlux_new(2ULL * seed);  // subsample.c:26, this_node = 0
  *out = s;
  return COLA_OK;
}

int cola_subsampler_destroy(cola_subsampler* s)
{
  if (!s) return COLA_OK;
  cola::ranlux_delete(s->rng);
  delete s;
  return COLA_OK;
}

int cola_write_subsample(cola_subsampler* s, const char* filename, const cola_particle_min* p, long long np, float a,
                         float boxsize, float omega_m, float h, long long* nsub_out)
{
  if (!s || !filename || (np > 0 && !p) || np < 0) return COLA_ERR_ARG;
  std::vector<float> sel;
  for (long long i = 0; i < np; ++i) {
    if (cola::ranlux_next(s->rng) < s->factor) {  // subsample.c:131
      for (int k = 0; k < 3; ++k) sel.push_back(p[i].x[k]);
      for (int k = 0; k < 3; ++k) sel.push_back(p[i].v[k]);
    }
  }
  const long long nsub = (long long)sel.size() / 6;
  if (nsub_out) *nsub_out = nsub;
  if (nsub > INT32_MAX) {
    set_error("subsample larger than 2^31 records");
    return COLA_ERR_ARG;
  }
  const double rho_crit = 27.7455;
  float header[6];
  header[0] = boxsize;
  header[1] = omega_m * rho_crit * std::pow((double)boxsize, 3.0) / np;  // subsample.c:150-151
  header[2] = omega_m;
  header[3] = h;
  header[4] = a;
  header[5] = 1.0 / a - 1.0;
  const int32_t n32 = (int32_t)nsub;
  File f(filename);
  if (!f.ok()) {
    set_error(std::string("Error: Unable to write to file: ") + filename);
    return 9000;
  }
  f.put(header, sizeof(header));
  f.put(&n32, sizeof(n32));
  f.put(sel.data(), sizeof(float) * sel.size());
  f.put(&n32, sizeof(n32));
  if (!f.close()) {
    set_error(std::string("Error: write failed: ") + filename);
    return 9000;
  }
  return COLA_OK;
}

int cola_write_coarse_grid(const char* filename, float boxsize, int n, const float* delta)
{
  if (!filename || n < 1 || !delta) return COLA_ERR_ARG;
  File f(filename);
  if (!f.ok()) {
    set_error(std::string("Unable to write grid to file: ") + filename);
    return 10021;  // coarse_grid.c:41-43
  }
  f.put(&boxsize, sizeof(float));
  f.put(&n, sizeof(int));
  f.put(delta, sizeof(float) * (size_t)n * n * n);
  f.put(&n, sizeof(int));
  if (!f.close()) {
    set_error(std::string("write failed: ") + filename);
    return 10021;
  }
  return COLA_OK;
}

int cola_write_halos(const char* filename, const cola_halo* halos, long long n)
{
  if (!filename || (n > 0 && !halos) || n < 0) return COLA_ERR_ARG;
  FILE* fp = std::fopen(filename, "w");
  if (!fp) {
    set_error(std::string("Error: Unable to write halo to file ") + filename);
    return 8110;  // fof.c:1148-1150
  }
  bool ok = true;
  for (long long i = 0; i < n && ok; ++i) {
    const cola_halo& h = halos[i];
    ok = std::fprintf(fp, "%d %e %e %e %e %e %e\n", h.nfof, h.x[0], h.x[1], h.x[2], h.v[0], h.v[1], h.v[2]) > 0;
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) {
    set_error(std::string("write failed: ") + filename);
    return 8110;
  }
  return COLA_OK;
}

}  // extern "C"
