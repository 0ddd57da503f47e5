// PM deposit / readout and COLA particle kernels.
//
// Deposit design: the reference scatters 8 CIC weights per particle with
// atomics into a grid initialised to -1 (pm.c:209-279).  Here each step
// counting-sorts particle positions into every density tile their stencil
// touches (a particle whose stencil crosses a tile face is listed in both
// tiles), then one workgroup per tile accumulates its particles' weights in
// LDS and writes every cell of the tile exactly once with coalesced stores.
// No global atomics touch the grid, and the -1 initialisation costs no extra
// pass.  The reference's x periodicity through ghost copies
// (pm.c:416-502, 505-532) is reproduced exactly: a particle with
// x > L - L/Ng also deposits (and reads back) through a copy at x - L whose
// weights are recomputed from that shifted position.
#include "common.h"
#include "pm.h"

#include <algorithm>
#include <vector>

namespace cola {

namespace {

constexpr int kThreads = 256;
constexpr int kScanChunk = 2048;  // counts per scan workgroup (8 per thread)


// The 8 CIC cells and weights of one position (pm.c:229-276).  fn(ix, iy, iz, w)
// is called for every cell whose x layer lies in [0, ng); y and z wrap.
// Returns false if y or z fall below -1 cell, where the reference would index
// outside its grid; nothing is emitted then.
template <class Fn>
__device__ inline bool cic_cells(float x, float y, float z, const CicGeom& g, Fn&& fn)
{
  const int ng = g.ng;
  const float X = x * g.scale;
  const float Y = y * g.scale;
  const float Z = z * g.scale;
  const int iI = (int)floorf(X);
  int J = (int)Y;
  int K = (int)Z;
  const float D1 = X - ((float)iI);
  const float D2 = Y - ((float)J);
  const float D3 = Z - ((float)K);
  const float T1 = 1.0f - D1;
  const float T2 = 1.0f - D2;
  const float T3 = 1.0f - D3;
  const float T2W = T2 * g.wpar;
  const float D2W = D2 * g.wpar;
  if (J < 0 || K < 0) return false;
  if (J >= ng) J = 0;
  if (K >= ng) K = 0;
  const int I1 = iI + 1;
  int J1 = J + 1;
  if (J1 >= ng) J1 = 0;
  int K1 = K + 1;
  if (K1 >= ng) K1 = 0;
  if (0 <= iI && iI < ng) {
    fn(iI, J, K, T3 * T1 * T2W);
    fn(iI, J, K1, D3 * T1 * T2W);
    fn(iI, J1, K, T3 * T1 * D2W);
    fn(iI, J1, K1, D3 * T1 * D2W);
  }
  if (0 <= I1 && I1 < ng) {
    fn(I1, J, K, T3 * D1 * T2W);
    fn(I1, J, K1, D3 * D1 * T2W);
    fn(I1, J1, K, T3 * D1 * D2W);
    fn(I1, J1, K1, D3 * D1 * D2W);
  }
  return true;
}

// The particle and, beyond the ghost threshold, its copy at x - L.
template <class Fn>
__device__ inline bool cic_with_ghost(float x, float y, float z, const CicGeom& g, Fn&& fn)
{
  bool ok = cic_cells(x, y, z, g, fn);
  if (x > g.x_right) ok = cic_cells(x + g.shift, y, z, g, fn) && ok;
  return ok;
}

// Tiles touched by a particle's CIC cells, as the lowest and highest tile
// index per axis.  Along y and z the cells are J, J+1 (wrapped), so at most
// two tiles.  Along x the cells are iI, iI+1 within [0, ng) and, for a ghost
// (x > x_right, so X >= ng - 2 up to rounding), the copy's cells, of which
// only 0 lies in the grid: the particle's own valid cells are ng-2 and ng-1,
// in one tile whenever tiles are at least 2 cells wide along x
// (make_tile_geom), and the ghost adds tile 0.  So every axis touches at
// most two tiles, and {lo, hi} lists them exactly.  Returns false (no tiles)
// where cic_cells emits nothing (y or z below -1 cell).
struct TileSpan {
  int x0, x1, y0, y1, z0, z1;  // x0 > x1: no cell in the grid along x
};

__device__ inline bool tile_span(float x, float y, float z, const CicGeom& g, const TileGeom& t, TileSpan& s)
{
  const int ng = g.ng;
  // the same float expressions as cic_cells, so the same cells
  const int iI = (int)floorf(x * g.scale);
  int J = (int)(y * g.scale);
  int K = (int)(z * g.scale);
  if (J < 0 || K < 0) {
    s.x0 = 1;
    s.x1 = 0;
    return false;
  }
  if (J >= ng) J = 0;
  if (K >= ng) K = 0;
  const int J1 = J + 1 >= ng ? 0 : J + 1;
  const int K1 = K + 1 >= ng ? 0 : K + 1;
  s.y0 = min(J, J1) >> t.ly;
  s.y1 = max(J, J1) >> t.ly;
  s.z0 = min(K, K1) >> t.lz;
  s.z1 = max(K, K1) >> t.lz;
  int lo = 0x7fffffff, hi = -1;
  auto add = [&](int i) {
    const bool in = 0 <= i && i < ng;
    lo = in ? min(lo, i >> t.lx) : lo;
    hi = in ? max(hi, i >> t.lx) : hi;
  };
  add(iI);
  add(iI + 1);
  if (x > g.x_right) {
    const int ig = (int)floorf((x + g.shift) * g.scale);
    add(ig);
    add(ig + 1);
  }
  s.x0 = lo;
  s.x1 = hi;
  return true;
}

// Number of (x, y, z) tile combinations of a span (0 to 8).
__device__ inline int span_count(const TileSpan& s)
{
  if (s.x0 > s.x1) return 0;
  return (1 + (s.x1 != s.x0)) * (1 + (s.y1 != s.y0)) * (1 + (s.z1 != s.z0));
}

// Tile index of combination k < span_count(s): z fastest, then y, then x.
// Tile indices fit 32 bits (make_tile_geom rejects more than 2^31 - 1 tiles).
__device__ inline int span_key(const TileSpan& s, const TileGeom& t, int k)
{
  const bool two_z = s.z1 != s.z0, two_y = s.y1 != s.y0;
  const int c = two_z && (k & 1) ? s.z1 : s.z0;
  const int r = two_z ? k >> 1 : k;
  const int b = two_y && (r & 1) ? s.y1 : s.y0;
  const int a = (two_y ? r >> 1 : r) ? s.x1 : s.x0;
  return (a * t.nty + b) * t.ntz + c;
}

__device__ inline void wrap1(float& x, float L)
{
  if (x < 0.0f)
    x += L;
  else if (x >= L)
    x -= L;
}

__global__ void __launch_bounds__(kThreads) k_wrap(DevParticles p, float L)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float x = p.x[a][i];
    wrap1(x, L);
    p.x[a][i] = x;
  }
}

// Workgroup-aggregated counter updates.  The (tile, particle) entries of a
// workgroup's particles (lattice neighbours, so mostly a few tiles) are
// combined in an LDS hash table: one global atomic per distinct tile and
// workgroup, all issued at once, instead of one per entry.  Entries that
// find no free slot within kHashProbes probes (more distinct tiles than a
// workgroup's particles normally touch) take a global atomic of their own.
constexpr int kHashLog2 = 9;
constexpr int kHashSlots = 1 << kHashLog2;
constexpr int kHashProbes = 32;
constexpr int kHashEmpty = -1;

struct BinHash {
  int key[kHashSlots];
  int cnt[kHashSlots];
};

__device__ inline void hash_clear(BinHash& h)
{
  for (int s = threadIdx.x; s < kHashSlots; s += kThreads) {
    h.key[s] = kHashEmpty;
    h.cnt[s] = 0;
  }
}

// Slot holding key (inserted if absent), or -1 after kHashProbes probes.
__device__ inline int hash_slot(BinHash& h, int key)
{
  unsigned s = ((unsigned)key * 2654435761u) >> (32 - kHashLog2);
#pragma unroll 1
  for (int probe = 0; probe < kHashProbes; ++probe) {
    const int prev = atomicCAS(&h.key[s], kHashEmpty, key);
    if (prev == kHashEmpty || prev == key) return (int)s;
    s = (s + 1) & (kHashSlots - 1);
  }
  return -1;
}

// Runs of equal keys in adjacent lanes of a wavefront (lattice neighbours
// mostly share a tile): the run's first lane (head) applies one atomic for
// the whole run, so the LDS table sees one atomic per run instead of one per
// lane (same-address atomics of a wavefront serialise).  For every active
// lane: its run's head lane; for a head: the run's length.  Wave-uniform
// control flow required (ballots).
struct LaneRun {
  int head, len;
  bool is_head;
};

__device__ inline LaneRun lane_run(int key, bool active)
{
  const int lane = __lane_id();
  const int prev_key = __shfl_up(key, 1);
  const unsigned long long act = __ballot(active);
  const bool prev_act = lane > 0 && ((act >> (lane - 1)) & 1ULL);
  LaneRun r;
  r.is_head = active && (!prev_act || prev_key != key);
  const unsigned long long heads = __ballot(r.is_head);
  // the run ends before the next head or inactive lane
  const unsigned long long stop = heads | ~act;
  const unsigned long long after = lane == 63 ? 0ULL : stop >> (lane + 1);
  const int next = after ? lane + __ffsll((long long)after) : 64;
  r.len = next - lane;
  // head: the highest head at or below this lane
  const unsigned long long upto = heads & (lane == 63 ? ~0ULL : (2ULL << lane) - 1ULL);
  r.head = upto ? 63 - __clzll((long long)upto) : lane;
  return r;
}

// Particles per thread in the binning kernels: a workgroup bins kThreads *
// kBinPer lattice-ordered particles (several lattice rows share tiles), so
// its table flush and barriers are paid once per 1024 particles and each
// thread has kBinPer independent loads in flight.  Particle q * kThreads +
// lane of a workgroup's chunk keeps adjacent lanes on adjacent particles.
constexpr int kBinPer = 4;

// Block blk of the counting pass (the kernels below run it for their own
// block index or for a block of one member of a batched launch).
__device__ __forceinline__ void bin_count_block(const DevParticles& p, const CicGeom& g, const TileGeom& t,
                                                int* __restrict__ counts, int* __restrict__ overflow, int wrap_first,
                                                float L, long long blk)
{
  __shared__ BinHash h;
  float px[kBinPer], py[kBinPer], pz[kBinPer];
  bool live[kBinPer];
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
    const long long i = (blk * kBinPer + q) * kThreads + threadIdx.x;
    live[q] = i < p.np;
    px[q] = live[q] ? p.x[0][i] : 0.0f;
    py[q] = live[q] ? p.x[1][i] : 0.0f;
    pz[q] = live[q] ? p.x[2][i] : 0.0f;
  }
  hash_clear(h);
  TileSpan ts[kBinPer];
  int nk[kBinPer];
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
    nk[q] = 0;
    if (live[q]) {
      float x = px[q], y = py[q], z = pz[q];
      if (wrap_first) {
        const long long i = (blk * kBinPer + q) * kThreads + threadIdx.x;
        wrap1(x, L);
        wrap1(y, L);
        wrap1(z, L);
        if (x != px[q]) p.x[0][i] = x;
        if (y != py[q]) p.x[1][i] = y;
        if (z != pz[q]) p.x[2][i] = z;
      }
      if (!tile_span(x, y, z, g, t, ts[q])) atomicOr(overflow, 1);
      nk[q] = span_count(ts[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
    // the slot loop is uniform across the wavefront (up to the wave's max)
    int kmax = nk[q];
    for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
    for (int k = 0; k < kmax; ++k) {
      const bool act = k < nk[q];
      const int key = act ? span_key(ts[q], t, k) : 0;
      const LaneRun run = lane_run(key, act);
      if (run.is_head) {
        const int slot = hash_slot(h, key);
        if (slot >= 0)
          atomicAdd(&h.cnt[slot], run.len);
        else
          atomicAdd(&counts[key], run.len);
      }
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < kHashSlots; s += kThreads)
    if (h.key[s] != kHashEmpty) atomicAdd(&counts[h.key[s]], h.cnt[s]);
}

__global__ void __launch_bounds__(kThreads)
k_bin_count(DevParticles p, CicGeom g, TileGeom t, int* __restrict__ counts, int* __restrict__ overflow, int wrap_first,
            float L)
{
  bin_count_block(p, g, t, counts, overflow, wrap_first, L, blockIdx.x);
}

// Claims the entries' slots by counting each tile's counter back down to zero
// (ready for the next step): a workgroup takes the cnt slots below the
// counter's value c, and its entry of rank r in the tile gets position
// offsets[tile] + c - 1 - r.  Entries that found no table slot claim theirs
// one by one from the counter.
__device__ __forceinline__ void bin_scatter_block(const DevParticles& p, const CicGeom& g, const TileGeom& t,
                                                  int* __restrict__ counts, const int* __restrict__ offsets,
                                                  float* __restrict__ sorted, long long capacity,
                                                  int* __restrict__ overflow, long long blk)
{
  __shared__ BinHash h;
  __shared__ long long base[kHashSlots];
  float px[kBinPer], py[kBinPer], pz[kBinPer];
  bool live[kBinPer];
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
    const long long i = (blk * kBinPer + q) * kThreads + threadIdx.x;
    live[q] = i < p.np;
    px[q] = live[q] ? p.x[0][i] : 0.0f;
    py[q] = live[q] ? p.x[1][i] : 0.0f;
    pz[q] = live[q] ? p.x[2][i] : 0.0f;
  }
  hash_clear(h);
  TileSpan ts[kBinPer];
  int nk[kBinPer];
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
    nk[q] = 0;
    if (live[q]) {
      tile_span(px[q], py[q], pz[q], g, t, ts[q]);
      nk[q] = span_count(ts[q]);
    }
  }
  __syncthreads();
  // per entry: table slot << 16 | rank in the workgroup's share of the tile
  // (ranks < kThreads * kBinPer * 8 = 8192), or -1 without a slot.  Static
  // indices keep the codes in registers.
  const int lane = __lane_id();
  int code[kBinPer][8];
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
    int kmax = nk[q];
    for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      code[q][k] = -1;
      if (k < kmax) {
        const bool act = k < nk[q];
        const int key = act ? span_key(ts[q], t, k) : 0;
        const LaneRun run = lane_run(key, act);
        int sl = -1, b = 0;
        if (run.is_head) {
          sl = hash_slot(h, key);
          if (sl >= 0) b = atomicAdd(&h.cnt[sl], run.len);
        }
        sl = __shfl(sl, run.head);
        b = __shfl(b, run.head);
        if (sl >= 0) code[q][k] = sl << 16 | (b + lane - run.head);
      }
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < kHashSlots; s += kThreads) {
    const int key = h.key[s];
    if (key != kHashEmpty) base[s] = (long long)offsets[key] + atomicSub(&counts[key], h.cnt[s]) - 1;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kBinPer; ++q) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k < nk[q]) {
        long long pos;
        if (code[q][k] >= 0) {
          pos = base[code[q][k] >> 16] - (code[q][k] & 0xffff);
        } else {
          const int key = span_key(ts[q], t, k);
          pos = (long long)offsets[key] + atomicSub(&counts[key], 1) - 1;
        }
        if (pos < capacity) {
          sorted[3 * pos + 0] = px[q];
          sorted[3 * pos + 1] = py[q];
          sorted[3 * pos + 2] = pz[q];
        } else {
          atomicOr(overflow, 2);
        }
      }
    }
  }
}

__global__ void __launch_bounds__(kThreads)
k_bin_scatter(DevParticles p, CicGeom g, TileGeom t, int* __restrict__ counts, const int* __restrict__ offsets,
              float* __restrict__ sorted, long long capacity, int* __restrict__ overflow)
{
  bin_scatter_block(p, g, t, counts, offsets, sorted, capacity, overflow, blockIdx.x);
}

// ---- exclusive scan of the tile counts -------------------------------------
__device__ __forceinline__ void scan_reduce_block(const int* __restrict__ in, long long n, int* __restrict__ sums,
                                                  int blk)
{
  __shared__ int red[kThreads];
  const long long base = (long long)blk * kScanChunk;
  int s = 0;
  for (int k = threadIdx.x; k < kScanChunk; k += kThreads) {
    const long long i = base + k;
    if (i < n) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blk] = red[0];
}

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const int* __restrict__ in, long long n, int* __restrict__ sums)
{
  scan_reduce_block(in, n, sums, blockIdx.x);
}

__device__ inline int block_exclusive_scan(int v, int* sh, int* total)
{
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    const int add = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  const int incl = sh[tid];
  *total = sh[kThreads - 1];
  __syncthreads();
  return incl - v;
}

// single workgroup: exclusive scan of the chunk sums in place
__device__ __forceinline__ void scan_top_block(int* __restrict__ sums, int nsums)
{
  __shared__ int sh[kThreads];
  int carry = 0;
  for (int base = 0; base < nsums; base += kThreads) {
    const int i = base + threadIdx.x;
    const int v = i < nsums ? sums[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, sh, &total);
    if (i < nsums) sums[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(kThreads) k_scan_top(int* __restrict__ sums, int nsums) { scan_top_block(sums, nsums); }

__device__ __forceinline__ void scan_final_block(const int* __restrict__ in, long long n, const int* __restrict__ sums,
                                                 int* __restrict__ out, int blk)
{
  __shared__ int sh[kThreads];
  const long long base = (long long)blk * kScanChunk;
  constexpr int per = kScanChunk / kThreads;
  int v[per];
  int s = 0;
#pragma unroll
  for (int k = 0; k < per; ++k) {
    const long long i = base + (long long)threadIdx.x * per + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int total;
  int run = sums[blk] + block_exclusive_scan(s, sh, &total);
#pragma unroll
  for (int k = 0; k < per; ++k) {
    const long long i = base + (long long)threadIdx.x * per + k;
    if (i < n) out[i] = run;
    run += v[k];
    if (i == n - 1) out[n] = run;
  }
}

__global__ void __launch_bounds__(kThreads)
k_scan_final(const int* __restrict__ in, long long n, const int* __restrict__ sums, int* __restrict__ out)
{
  scan_final_block(in, n, sums, out, blockIdx.x);
}

// ---- deposit: one workgroup per tile ----------------------------------------
__device__ inline double wave_sum(double v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Each CIC weight (pm.c's float products, WPAR included) is added to its
// LDS cell as a 32.32 fixed-point integer: w * 2^32 is an exact integer for
// every w >= 2^-9 and is rounded at 2^-33 below, so a cell holds the exact
// sum of its weights whatever the order of the atomics, and the density
// -1 + sum is rounded once.  The grid is therefore bit-reproducible (the
// reference's serial float sum differs from it by float rounding only).
constexpr float kFix = 4294967296.0f;  // 2^32
constexpr double kUnfix = 1.0 / 4294967296.0;
constexpr int kDepositThreads = 256;

// Tile `tile` of the deposit; cells: the tile's LDS accumulators.
__device__ __forceinline__ void deposit_tile(const CicGeom& g, const TileGeom& t, const int* __restrict__ offsets,
                                             const float* __restrict__ sorted, float* __restrict__ density,
                                             double* __restrict__ tile_sums, unsigned long long* cells, long long tile)
{
  __shared__ double wsum[kDepositThreads / 64];
  const int tz = (int)(tile % t.ntz);
  const int ty = (int)((tile / t.ntz) % t.nty);
  const int tx = (int)(tile / ((long long)t.ntz * t.nty));
  const int x0 = tx << t.lx, y0 = ty << t.ly, z0 = tz << t.lz;
  const int ncell = 1 << (t.lx + t.ly + t.lz);
  for (int e = threadIdx.x; e < ncell; e += kDepositThreads) cells[e] = 0ULL;
  __syncthreads();
  const int beg = offsets[tile], end = offsets[tile + 1];
  for (int k = beg + threadIdx.x; k < end; k += kDepositThreads) {
    const float x = sorted[3LL * k + 0], y = sorted[3LL * k + 1], z = sorted[3LL * k + 2];
    cic_with_ghost(x, y, z, g, [&](int i, int j, int kk, float w) {
      const unsigned li = (unsigned)(i - x0), lj = (unsigned)(j - y0), lk = (unsigned)(kk - z0);
      if (li < (unsigned)t.tx && lj < (unsigned)t.ty && lk < (unsigned)t.tz)
        atomicAdd(&cells[(((li << t.ly) + lj) << t.lz) + lk], (unsigned long long)rintf(w * kFix));
    });
  }
  __syncthreads();
  double s = 0.0;
  const int zmask = t.tz - 1, ymask = t.ty - 1;
  auto cell_value = [&](int e) {
    return (float)(-1.0 + (double)cells[e] * kUnfix);  // pm.c:220 density starts at -1
  };
  if (t.lz >= 2) {
    // four consecutive z cells per thread: one 16-byte store (rows start on
    // 16-byte boundaries and z0 is a multiple of 4)
    for (int e = threadIdx.x << 2; e < ncell; e += kDepositThreads << 2) {
      const int lk = e & zmask;
      const int lj = (e >> t.lz) & ymask;
      const int li = e >> (t.lz + t.ly);
      const float4 v = make_float4(cell_value(e), cell_value(e + 1), cell_value(e + 2), cell_value(e + 3));
      *reinterpret_cast<float4*>(density + ((long long)(x0 + li) * g.ng + (y0 + lj)) * g.row + (z0 + lk)) = v;
      s += (double)v.x;
      s += (double)v.y;
      s += (double)v.z;
      s += (double)v.w;
    }
  } else {
    for (int e = threadIdx.x; e < ncell; e += kDepositThreads) {
      const int lk = e & zmask;
      const int lj = (e >> t.lz) & ymask;
      const int li = e >> (t.lz + t.ly);
      const float v = cell_value(e);
      density[((long long)(x0 + li) * g.ng + (y0 + lj)) * g.row + (z0 + lk)] = v;
      s += (double)v;
    }
  }
  s = wave_sum(s);
  if (__lane_id() == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < kDepositThreads / 64; ++w) tot += wsum[w];
    tile_sums[tile] = tot;
  }
}

__global__ void __launch_bounds__(kDepositThreads)
k_deposit_tiles(CicGeom g, TileGeom t, const int* __restrict__ offsets, const float* __restrict__ sorted,
                float* __restrict__ density, double* __restrict__ tile_sums)
{
  extern __shared__ unsigned long long cells[];
  deposit_tile(g, t, offsets, sorted, density, tile_sums, cells, blockIdx.x);
}

// pm.c:535-562: sum of the per-tile sums into ring[prm->slot] (zeroed first)
__device__ __forceinline__ void sum_doubles_block(const double* __restrict__ v, long long n, double* ring,
                                                  const StepParams* __restrict__ prm, int blk)
{
  double* const out = ring + prm->slot;
  __shared__ double wsum[kThreads / 64];
  const long long base = (long long)blk * kThreads * 16;
  double s = 0.0;
  for (int k = 0; k < 16; ++k) {
    const long long i = base + (long long)k * kThreads + threadIdx.x;
    if (i < n) s += v[i];
  }
  s = wave_sum(s);
  if (__lane_id() == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) tot += wsum[w];
    atomicAdd(out, tot);
  }
}

__global__ void __launch_bounds__(kThreads)
k_sum_doubles(const double* __restrict__ v, long long n, double* ring, const StepParams* __restrict__ prm)
{
  sum_doubles_block(v, n, ring, prm, blockIdx.x);
}

// ---- readout (pm.c:359-412) -------------------------------------------------
// Particles stay in lattice order (a workgroup's 256 particles are a lattice
// row at nc = 256), and lattice neighbours gather from overlapping grid
// rows: consecutive workgroups go in groups (CicGeom::readout_group) to one
// XCD, so the shared grid lines are read once into that XCD's L2 (256^3 on
// 768^3: groups of 8 took the fused readout from 2.0 to 1.68 ms).

// One axis of the CIC gather at one position, term order of pm.c:397-410.
struct Gather3 {
  float f[3];
};

__device__ inline Gather3 cic_gather(float x, float y, float z, const CicGeom& g, const float* __restrict__ g0,
                                     const float* __restrict__ g1, const float* __restrict__ g2)
{
  const int ng = g.ng;
  const float X = x * g.scale;
  const float Y = y * g.scale;
  const float Z = z * g.scale;
  int iI = (int)floorf(X);
  int J = (int)Y;
  int K = (int)Z;
  const float D1 = X - (float)iI;
  const float D2 = Y - (float)J;
  const float D3 = Z - (float)K;
  const float T1 = 1.0f - D1;
  const float T2 = 1.0f - D2;
  const float T3 = 1.0f - D3;
  Gather3 r;
  r.f[0] = r.f[1] = r.f[2] = 0.0f;
  if (J < 0 || K < 0) return r;
  if (J >= ng) J = 0;
  if (K >= ng) K = 0;
  const int I1 = iI + 1;
  int J1 = J + 1;
  if (J1 >= ng) J1 = 0;
  int K1 = K + 1;
  if (K1 >= ng) K1 = 0;
  const float* grids[3] = {g0, g1, g2};
  const long long row = g.row;
  if (0 <= iI && iI < ng) {
    const long long a = ((long long)iI * ng + J) * row, b = ((long long)iI * ng + J1) * row;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const float* m = grids[ax];
      r.f[ax] += m[a + K] * T3 * T1 * T2 + m[a + K1] * D3 * T1 * T2 + m[b + K] * T3 * T1 * D2 + m[b + K1] * D3 * T1 * D2;
    }
  }
  if (0 <= I1 && I1 < ng) {
    const long long a = ((long long)I1 * ng + J) * row, b = ((long long)I1 * ng + J1) * row;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const float* m = grids[ax];
      r.f[ax] += m[a + K] * T3 * D1 * T2 + m[a + K1] * D3 * D1 * T2 + m[b + K] * T3 * D1 * D2 + m[b + K1] * D3 * D1 * D2;
    }
  }
  return r;
}

// The same gather from one interleaved grid fi[cell][3] (cells unpadded,
// C order): the three components of a corner are one 12-byte record, so a
// particle touches four grid lines instead of twelve.  Same index, weight
// and term order as cic_gather.
__device__ inline Gather3 cic_gather_fi(float x, float y, float z, const CicGeom& g, const float* __restrict__ fi)
{
  const int ng = g.ng;
  const float X = x * g.scale;
  const float Y = y * g.scale;
  const float Z = z * g.scale;
  int iI = (int)floorf(X);
  int J = (int)Y;
  int K = (int)Z;
  const float D1 = X - (float)iI;
  const float D2 = Y - (float)J;
  const float D3 = Z - (float)K;
  const float T1 = 1.0f - D1;
  const float T2 = 1.0f - D2;
  const float T3 = 1.0f - D3;
  Gather3 r;
  r.f[0] = r.f[1] = r.f[2] = 0.0f;
  if (J < 0 || K < 0) return r;
  if (J >= ng) J = 0;
  if (K >= ng) K = 0;
  const int I1 = iI + 1;
  int J1 = J + 1;
  if (J1 >= ng) J1 = 0;
  int K1 = K + 1;
  if (K1 >= ng) K1 = 0;
  auto rec = [&](long long line, int k) { return *reinterpret_cast<const float3*>(fi + (line * ng + k) * 3); };
  if (0 <= iI && iI < ng) {
    const long long a = (long long)iI * ng + J, b = (long long)iI * ng + J1;
    const float3 aK = rec(a, K), aK1 = rec(a, K1), bK = rec(b, K), bK1 = rec(b, K1);
    const float va[3][4] = {{aK.x, aK1.x, bK.x, bK1.x}, {aK.y, aK1.y, bK.y, bK1.y}, {aK.z, aK1.z, bK.z, bK1.z}};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      r.f[ax] += va[ax][0] * T3 * T1 * T2 + va[ax][1] * D3 * T1 * T2 + va[ax][2] * T3 * T1 * D2 + va[ax][3] * D3 * T1 * D2;
  }
  if (0 <= I1 && I1 < ng) {
    const long long a = (long long)I1 * ng + J, b = (long long)I1 * ng + J1;
    const float3 aK = rec(a, K), aK1 = rec(a, K1), bK = rec(b, K), bK1 = rec(b, K1);
    const float va[3][4] = {{aK.x, aK1.x, bK.x, bK1.x}, {aK.y, aK1.y, bK.y, bK1.y}, {aK.z, aK1.z, bK.z, bK1.z}};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      r.f[ax] += va[ax][0] * T3 * D1 * T2 + va[ax][1] * D3 * D1 * T2 + va[ax][2] * T3 * D1 * D2 + va[ax][3] * D3 * D1 * D2;
  }
  return r;
}

// Force at a particle: its own gather plus, beyond the ghost threshold, the
// gather at its ghost position added afterwards (pm.c:526-531).
template <bool FI>
__device__ inline Gather3 particle_force(float x, float y, float z, const CicGeom& g, const float* g0, const float* g1,
                                         const float* g2)
{
  Gather3 f = FI ? cic_gather_fi(x, y, z, g, g0) : cic_gather(x, y, z, g, g0, g1, g2);
  if (x > g.x_right) {
    const Gather3 h = FI ? cic_gather_fi(x + g.shift, y, z, g, g0) : cic_gather(x + g.shift, y, z, g, g0, g1, g2);
#pragma unroll
    for (int a = 0; a < 3; ++a) f.f[a] += h.f[a];
  }
  return f;
}

// Mean of a residual added to the force grids (fsum: its three sums over the
// grid, null when there is none): a constant in the grid is a constant in
// every CIC readout, and the force of a periodic field has none
// (pm.c:312-313 zeroes the k = 0 mode), so it is taken back off.
__device__ inline void residual_mean(const double* __restrict__ fsum, float inv_cells, float (&m)[3])
{
#pragma unroll
  for (int a = 0; a < 3; ++a) m[a] = fsum ? (float)(fsum[a] * (double)inv_cells) : 0.0f;
}

// FI: g0 is the interleaved grid fi[cell][3] (g1, g2 unused).
template <bool FI>
__global__ void __launch_bounds__(kThreads)
k_readout(DevParticles p, CicGeom g, const float* __restrict__ g0, const float* __restrict__ g1,
          const float* __restrict__ g2, const double* __restrict__ fsum, float inv_cells)
{
  const long long i = (long long)xcd_group_remap(blockIdx.x, gridDim.x, g.readout_group) * kThreads + threadIdx.x;
  if (i >= p.np) return;
  Gather3 f = particle_force<FI>(p.x[0][i], p.x[1][i], p.x[2][i], g, g0, g1, g2);
  float m[3];
  residual_mean(fsum, inv_cells, m);
#pragma unroll
  for (int a = 0; a < 3; ++a) p.f[a][i] = fsum ? f.f[a] - m[a] : f.f[a];
}

// cola.c:105-111 for one component
__device__ inline float kick1(float v, float f, float dx1, float dx2, const KickFactors& k)
{
  const float ax = (float)(k.m15om * (double)f - (double)(1.0f * (dx1 * k.q1 + dx2 * k.q2)));
  return v + ax * k.dda;
}

// cola.c:143-148 for one component
__device__ inline float drift1(float x, float v, float dx1, float dx2, const DriftFactors& d)
{
  return x + (v * d.dyyy + 1.0f * (dx1 * d.da1 + dx2 * d.da2));
}

// Block blk of nblk (the member's own grid) of the fused readout.
template <bool FI>
__device__ __forceinline__ void readout_kick_drift_block(const DevParticles& p, const CicGeom& g,
                                                         const float* __restrict__ g0, const float* __restrict__ g1,
                                                         const float* __restrict__ g2,
                                                         const StepParams* __restrict__ prm,
                                                         const double* __restrict__ fsum, float inv_cells, int blk,
                                                         int nblk)
{
  const long long i = (long long)xcd_group_remap(blk, nblk, g.readout_group) * kThreads + threadIdx.x;
  if (i >= p.np) return;
  const KickFactors kf = prm->k;
  const DriftFactors df = prm->d;
  float x[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) x[a] = p.x[a][i];
  Gather3 f = particle_force<FI>(x[0], x[1], x[2], g, g0, g1, g2);
  if (fsum) {
    float m[3];
    residual_mean(fsum, inv_cells, m);
#pragma unroll
    for (int a = 0; a < 3; ++a) f.f[a] -= m[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float d1 = p.dx1[a][i], d2 = p.dx2[a][i];
    const float v = kick1(p.v[a][i], f.f[a], d1, d2, kf);
    p.v[a][i] = v;
    p.x[a][i] = drift1(x[a], v, d1, d2, df);
  }
}

template <bool FI>
__global__ void __launch_bounds__(kThreads)
k_readout_kick_drift(DevParticles p, CicGeom g, const float* __restrict__ g0, const float* __restrict__ g1,
                     const float* __restrict__ g2, const StepParams* __restrict__ prm, const double* __restrict__ fsum,
                     float inv_cells)
{
  readout_kick_drift_block<FI>(p, g, g0, g1, g2, prm, fsum, inv_cells, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kThreads) k_kick(DevParticles p, KickFactors kf)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) p.v[a][i] = kick1(p.v[a][i], p.f[a][i], p.dx1[a][i], p.dx2[a][i], kf);
}

__global__ void __launch_bounds__(kThreads) k_drift(DevParticles p, DriftFactors df)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) p.x[a][i] = drift1(p.x[a][i], p.v[a][i], p.dx1[a][i], p.dx2[a][i], df);
}

// cola.c:388-410
__global__ void __launch_bounds__(kThreads) k_snapshot(DevParticles p, SnapshotFactors sf, DevParticleMin* out)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.np) return;
  DevParticleMin o;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float x = p.x[a][i], v = p.v[a][i], d1 = p.dx1[a][i], d2 = p.dx2[a][i], f = p.f[a][i];
    const float ax = (float)(sf.kick.m15om * (double)f - (double)(1.0f * (d1 * sf.kick.q1 + d2 * sf.kick.q2)));
    o.v[a] = sf.vfac * (v + ax * sf.kick.dda + (d1 * sf.Dv + d2 * sf.Dv2) * 1.0f);
    o.x[a] = x + v * sf.drift.dyyy + 1.0f * (d1 * sf.drift.da1 + d2 * sf.drift.da2);
  }
  o.id = i + 1;
  out[i] = o;
}

// cola.c:321-330
__global__ void __launch_bounds__(kThreads) k_noncola_initial(DevParticles p, SnapshotFactors sf, DevParticleMin* out)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.np) return;
  DevParticleMin o;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    o.v[a] = sf.vfac * (p.dx1[a][i] * sf.Dv + p.dx2[a][i] * sf.Dv2);
    o.x[a] = p.x[a][i];
  }
  o.id = i + 1;
  out[i] = o;
}

__global__ void k_zero_doubles(double* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_zero_mass(double* ring, const StepParams* __restrict__ prm)
{
  if (threadIdx.x == 0) ring[prm->slot] = 0.0;
}

// ---- batched kernels: workgroup b runs workgroup b - start[m] of member m --
__global__ void __launch_bounds__(kThreads) k_zero_mass_batch(const PmMember* __restrict__ ms, int n)
{
  const int m = blockIdx.x * kThreads + threadIdx.x;
  if (m < n) ms[m].ring[ms[m].prm->slot] = 0.0;
}

__global__ void __launch_bounds__(kThreads)
k_bin_count_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  bin_count_block(M.p, M.g, M.t, M.b.counts, M.b.overflow, 1, M.boxsize, (long long)blockIdx.x - start[m]);
}

__global__ void __launch_bounds__(kThreads)
k_scan_reduce_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  scan_reduce_block(M.b.counts, M.t.ntiles, M.b.block_sums, (int)blockIdx.x - start[m]);
}

// one workgroup per member
__global__ void __launch_bounds__(kThreads) k_scan_top_batch(const PmMember* __restrict__ ms)
{
  const PmMember& M = ms[blockIdx.x];
  scan_top_block(M.b.block_sums, (int)((M.t.ntiles + kScanChunk - 1) / kScanChunk));
}

__global__ void __launch_bounds__(kThreads)
k_scan_final_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  scan_final_block(M.b.counts, M.t.ntiles, M.b.block_sums, M.b.offsets, (int)blockIdx.x - start[m]);
}

__global__ void __launch_bounds__(kThreads)
k_bin_scatter_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  bin_scatter_block(M.p, M.g, M.t, M.b.counts, M.b.offsets, M.b.sorted, M.b.capacity, M.b.overflow,
                    (long long)blockIdx.x - start[m]);
}

__global__ void __launch_bounds__(kDepositThreads)
k_deposit_tiles_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  extern __shared__ unsigned long long cells[];
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  deposit_tile(M.g, M.t, M.b.offsets, M.b.sorted, M.density, M.b.tile_sums, cells, (long long)blockIdx.x - start[m]);
}

__global__ void __launch_bounds__(kThreads)
k_sum_doubles_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  sum_doubles_block(M.b.tile_sums, M.t.ntiles, M.ring, M.prm, (int)blockIdx.x - start[m]);
}

// Members start on a multiple of 8 x their readout group, so the XCD
// grouping of each member's workgroups is that of its own launch.
__global__ void __launch_bounds__(kThreads)
k_readout_kick_drift_batch(const PmMember* __restrict__ ms, const int* __restrict__ start, int n)
{
  const int m = batch_find(start, n, blockIdx.x);
  const PmMember& M = ms[m];
  const int nb = (int)((M.p.np + kThreads - 1) / kThreads);
  const int lb = (int)blockIdx.x - start[m];
  if (lb >= nb) return;  // alignment padding
  readout_kick_drift_block<false>(M.p, M.g, M.f[0], M.f[1], M.f[2], M.prm, nullptr, 0.0f, lb, nb);
}

unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
unsigned bin_blocks(long long n) { return (unsigned)((n + kThreads * kBinPer - 1) / (kThreads * kBinPer)); }

}  // namespace

CicGeom make_cic_geom(int ng, float boxsize, float wpar, long long row)
{
  CicGeom g;
  g.ng = ng;
  g.scale = ((float)ng) / ((float)boxsize);
  g.wpar = wpar;
  const float eps = boxsize / ng;
  g.x_right = boxsize / ng * (0 + ng) - eps;
  g.shift = -boxsize;
  g.row = row;
  return g;
}

static int log2_exact(int v)
{
  int l = 0;
  while ((1 << l) < v) ++l;
  return (1 << l) == v ? l : -1;
}

bool make_tile_geom(int ng, int tx, int ty, int tz, TileGeom* t)
{
  // tx >= 2: a ghost's x cells then span at most two x tiles (tile_span)
  if (tx < 2 || ty <= 0 || tz <= 0 || ng % tx || ng % ty || ng % tz) return false;
  if ((long long)tx * ty * tz > 4096) return false;
  const int lx = log2_exact(tx), ly = log2_exact(ty), lz = log2_exact(tz);
  if (lx < 0 || ly < 0 || lz < 0) return false;
  t->tx = tx;
  t->ty = ty;
  t->tz = tz;
  t->lx = lx;
  t->ly = ly;
  t->lz = lz;
  t->ntx = ng / tx;
  t->nty = ng / ty;
  t->ntz = ng / tz;
  t->ntiles = (long long)t->ntx * t->nty * t->ntz;
  return t->ntiles <= 0x7fffffffLL;
}

int bin_alloc(BinBuffers* b, const TileGeom& t, long long np)
{
  bin_free(b);
  const long long nsums = (t.ntiles + kScanChunk - 1) / kScanChunk;
  b->capacity = 8 * np + 1024;
  COLA_HIP(hipMalloc(&b->counts, sizeof(int) * t.ntiles));
  COLA_HIP(hipMalloc(&b->offsets, sizeof(int) * (t.ntiles + 1)));
  COLA_HIP(hipMalloc(&b->block_sums, sizeof(int) * (nsums + 1)));
  COLA_HIP(hipMalloc(&b->sorted, sizeof(float) * 3 * b->capacity));
  COLA_HIP(hipMalloc(&b->overflow, sizeof(int)));
  COLA_HIP(hipMalloc(&b->tile_sums, sizeof(double) * t.ntiles));
  COLA_HIP(hipMemset(b->counts, 0, sizeof(int) * t.ntiles));
  COLA_HIP(hipMemset(b->overflow, 0, sizeof(int)));
  return kOk;
}

void bin_free(BinBuffers* b)
{
  if (b->counts) (void)hipFree(b->counts);
  if (b->offsets) (void)hipFree(b->offsets);
  if (b->block_sums) (void)hipFree(b->block_sums);
  if (b->sorted) (void)hipFree(b->sorted);
  if (b->overflow) (void)hipFree(b->overflow);
  if (b->tile_sums) (void)hipFree(b->tile_sums);
  *b = BinBuffers{};
}

int launch_wrap(const DevParticles& p, float boxsize, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_wrap, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, boxsize);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_bin_count(const DevParticles& p, const CicGeom& g, const TileGeom& t, BinBuffers& b, bool wrap_first,
                     float boxsize, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_bin_count, dim3(bin_blocks(p.np)), dim3(kThreads), 0, s, p, g, t, b.counts, b.overflow,
                     wrap_first ? 1 : 0, boxsize);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_bin_scan(const TileGeom& t, BinBuffers& b, hipStream_t s)
{
  const long long n = t.ntiles;
  const int nsums = (int)((n + kScanChunk - 1) / kScanChunk);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nsums), dim3(kThreads), 0, s, b.counts, n, b.block_sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kThreads), 0, s, b.block_sums, nsums);
  hipLaunchKernelGGL(k_scan_final, dim3(nsums), dim3(kThreads), 0, s, b.counts, n, b.block_sums, b.offsets);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_bin_scatter(const DevParticles& p, const CicGeom& g, const TileGeom& t, BinBuffers& b, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_bin_scatter, dim3(bin_blocks(p.np)), dim3(kThreads), 0, s, p, g, t, b.counts, b.offsets,
                     b.sorted, b.capacity, b.overflow);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_deposit_tiles(const CicGeom& g, const TileGeom& t, const BinBuffers& b, float* density, double* ring,
                         const StepParams* prm, hipStream_t s)
{
  const size_t lds = sizeof(unsigned long long) * t.tx * t.ty * t.tz;
  hipLaunchKernelGGL(k_deposit_tiles, dim3((unsigned)t.ntiles), dim3(kDepositThreads), lds, s, g, t, b.offsets, b.sorted,
                     density, b.tile_sums);
  const long long per = (long long)kThreads * 16;
  hipLaunchKernelGGL(k_sum_doubles, dim3((unsigned)((t.ntiles + per - 1) / per)), dim3(kThreads), 0, s, b.tile_sums,
                     t.ntiles, ring, prm);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_readout(const DevParticles& p, const CicGeom& g, const float* fx, const float* fy, const float* fz,
                   const double* fsum, const float* fi, hipStream_t s)
{
  if (p.np == 0) return kOk;
  const float inv_cells = (float)(1.0 / ((double)g.ng * g.ng * g.ng));
  if (fi)
    hipLaunchKernelGGL(k_readout<true>, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, g, fi, fi, fi, fsum,
                       inv_cells);
  else
    hipLaunchKernelGGL(k_readout<false>, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, g, fx, fy, fz, fsum,
                       inv_cells);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_readout_kick_drift(const DevParticles& p, const CicGeom& g, const float* fx, const float* fy,
                              const float* fz, const StepParams* prm, const double* fsum, const float* fi,
                              hipStream_t s)
{
  if (p.np == 0) return kOk;
  const float inv_cells = (float)(1.0 / ((double)g.ng * g.ng * g.ng));
  if (fi)
    hipLaunchKernelGGL(k_readout_kick_drift<true>, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, g, fi, fi, fi, prm,
                       fsum, inv_cells);
  else
    hipLaunchKernelGGL(k_readout_kick_drift<false>, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, g, fx, fy, fz,
                       prm, fsum, inv_cells);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_kick(const DevParticles& p, const KickFactors& k, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_kick, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, k);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_drift(const DevParticles& p, const DriftFactors& d, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_drift, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, d);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_snapshot(const DevParticles& p, const SnapshotFactors& f, DevParticleMin* out, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_snapshot, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, f, out);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_noncola_initial(const DevParticles& p, const SnapshotFactors& f, DevParticleMin* out, hipStream_t s)
{
  if (p.np == 0) return kOk;
  hipLaunchKernelGGL(k_noncola_initial, dim3(blocks_for(p.np)), dim3(kThreads), 0, s, p, f, out);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_zero_mass(double* ring, const StepParams* prm, hipStream_t s)
{
  hipLaunchKernelGGL(k_zero_mass, dim3(1), dim3(64), 0, s, ring, prm);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int PmBatch::init(const PmMember* ms, int nm)
{
  destroy();
  if (nm <= 0 || !ms) return kErrArg;
  n = nm;
  std::vector<int> st((size_t)kPmBatchKernels * (nm + 1));
  size_t lds = 0;
  long long at[kPmBatchKernels] = {};
  for (int m = 0; m < nm; ++m) {
    const PmMember& M = ms[m];
    const long long nsums = (M.t.ntiles + kScanChunk - 1) / kScanChunk;
    const long long per_sum = (long long)kThreads * 16;
    const long long cnt[kPmBatchKernels] = {(long long)bin_blocks(M.p.np), nsums, M.t.ntiles,
                                            (M.t.ntiles + per_sum - 1) / per_sum, (long long)blocks_for(M.p.np)};
    const long long align[kPmBatchKernels] = {1, 1, 1, 1, 8LL * std::max(1, M.g.readout_group)};
    for (int k = 0; k < kPmBatchKernels; ++k) {
      at[k] = (at[k] + align[k] - 1) / align[k] * align[k];
      st[(size_t)k * (nm + 1) + m] = (int)at[k];
      at[k] += cnt[k];
      if (at[k] > 0x7fffffffLL || at[k] * kThreads > 0xffffffffLL) {
        set_error("batched step: too many workgroups for one launch");
        return kErrUnsupported;
      }
    }
    lds = std::max(lds, sizeof(unsigned long long) * M.t.tx * M.t.ty * M.t.tz);
  }
  for (int k = 0; k < kPmBatchKernels; ++k) {
    st[(size_t)k * (nm + 1) + nm] = (int)at[k];
    blocks[k] = (int)at[k];
  }
  deposit_lds = lds;
  const size_t mb = (sizeof(PmMember) * nm + 255) / 256 * 256;
  COLA_HIP(hipMalloc(&d_mem, mb + sizeof(int) * st.size()));
  COLA_HIP(hipMemcpy(d_mem, ms, sizeof(PmMember) * nm, hipMemcpyHostToDevice));
  COLA_HIP(hipMemcpy(static_cast<char*>(d_mem) + mb, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
  members = static_cast<const PmMember*>(d_mem);
  start = reinterpret_cast<const int*>(static_cast<char*>(d_mem) + mb);
  return kOk;
}

void PmBatch::destroy()
{
  if (d_mem) (void)hipFree(d_mem);
  *this = PmBatch{};
}

int launch_assign_batch(const PmBatch& b, hipStream_t s)
{
  const int n = b.n;
  auto tab = [&](int k) { return b.start + (size_t)k * (n + 1); };
  hipLaunchKernelGGL(k_zero_mass_batch, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, b.members, n);
  if (b.blocks[kBatchBin] > 0)
    hipLaunchKernelGGL(k_bin_count_batch, dim3(b.blocks[kBatchBin]), dim3(kThreads), 0, s, b.members, tab(kBatchBin),
                       n);
  hipLaunchKernelGGL(k_scan_reduce_batch, dim3(b.blocks[kBatchScan]), dim3(kThreads), 0, s, b.members,
                     tab(kBatchScan), n);
  hipLaunchKernelGGL(k_scan_top_batch, dim3(n), dim3(kThreads), 0, s, b.members);
  hipLaunchKernelGGL(k_scan_final_batch, dim3(b.blocks[kBatchScan]), dim3(kThreads), 0, s, b.members,
                     tab(kBatchScan), n);
  if (b.blocks[kBatchBin] > 0)
    hipLaunchKernelGGL(k_bin_scatter_batch, dim3(b.blocks[kBatchBin]), dim3(kThreads), 0, s, b.members,
                       tab(kBatchBin), n);
  hipLaunchKernelGGL(k_deposit_tiles_batch, dim3(b.blocks[kBatchTile]), dim3(kDepositThreads), b.deposit_lds, s,
                     b.members, tab(kBatchTile), n);
  hipLaunchKernelGGL(k_sum_doubles_batch, dim3(b.blocks[kBatchSum]), dim3(kThreads), 0, s, b.members, tab(kBatchSum),
                     n);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_readout_kick_drift_batch(const PmBatch& b, hipStream_t s)
{
  if (b.blocks[kBatchRead] == 0) return kOk;
  hipLaunchKernelGGL(k_readout_kick_drift_batch, dim3(b.blocks[kBatchRead]), dim3(kThreads), 0, s, b.members,
                     b.start + (size_t)kBatchRead * (b.n + 1), b.n);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_zero_doubles(double* p, int n, hipStream_t s)
{
  hipLaunchKernelGGL(k_zero_doubles, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  COLA_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cola
