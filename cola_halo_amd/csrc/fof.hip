// Friends-of-friends on the device: cell sort, lock-free union-find over all
// pairs closer than the linking length, then group statistics.
//
// The reference (fof.c:427-627) walks a k-d tree from each unvisited particle
// with a FIFO and, across the slab boundary in x, links exported buffer
// particles and merges halo records (fof.c:713-839, 965-1030).  On one rank
// every one of those paths links exactly the pairs with squared distance
// below ll^2 under the periodic minimum image, so a group here is the same
// set of particles.  Group order and the choice of the first particle x0
// differ: groups are listed by their smallest record index, and x0 is that
// record's position.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <climits>
#include <cmath>

#include "common.h"
#include "fof.h"

namespace cola {
namespace {

constexpr int kThreads = 256;

unsigned blocks(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

__device__ inline int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void st(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Root of x with path halving.  parent[y] <= y holds throughout (links go
// from the larger root to the smaller), so every walk ends.
__device__ int find_root(int* parent, int x)
{
  while (true) {
    const int p = ld(&parent[x]);
    if (p == x) return x;
    const int gp = ld(&parent[p]);
    if (gp == p) return p;
    st(&parent[x], gp);
    x = gp;
  }
}

__device__ void unite(int* parent, int a, int b)
{
  while (true) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

// move.c:33-41 wrap (the move_particles2_min of fof.c:1044), written back to
// the records, and the linking-grid cell key of each record.
__global__ void __launch_bounds__(kThreads)
k_fof_wrap_key(float* base, long long stride, long long np, float box, int nd, float inv_cs, float* __restrict__ px,
               float* __restrict__ py, float* __restrict__ pz, unsigned* __restrict__ keys, int* __restrict__ idx)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= np) return;
  float x[3];
  int c[3];
  for (int a = 0; a < 3; ++a) {
    float v = base[i * stride + a];
    if (v < 0.0f)
      v += box;
    else if (v >= box)
      v -= box;
    base[i * stride + a] = v;
    x[a] = v;
    int q = (int)(v * inv_cs);
    c[a] = q < 0 ? 0 : (q >= nd ? nd - 1 : q);
  }
  px[i] = x[0];
  py[i] = x[1];
  pz[i] = x[2];
  keys[i] = ((unsigned)c[0] * (unsigned)nd + (unsigned)c[1]) * (unsigned)nd + (unsigned)c[2];
  idx[i] = (int)i;
}

__global__ void __launch_bounds__(kThreads)
k_fof_gather(long long np, const int* __restrict__ sidx, const float* __restrict__ px, const float* __restrict__ py,
             const float* __restrict__ pz, float4* __restrict__ spos, int* __restrict__ parent)
{
  const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (k >= np) return;
  const int i = sidx[k];
  spos[k] = make_float4(px[i], py[i], pz[i], 0.0f);
  parent[k] = (int)k;
}

__device__ inline long long lower_bound(const unsigned* a, long long n, unsigned key)
{
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

struct LinkGeom {
  long long np;
  int nd;
  float box, half, ll2;
};

__device__ inline void link_range(const LinkGeom& g, const unsigned* __restrict__ skeys,
                                  const float4* __restrict__ spos, int* parent, long long k, float4 p, unsigned klo,
                                  unsigned khi)
{
  long long lo = lower_bound(skeys, g.np, klo);
  const long long hi = lower_bound(skeys, g.np, khi + 1u);
  if (lo < k + 1) lo = k + 1;  // each pair once, from its lower sorted index
  for (long long j = lo; j < hi; ++j) {
    const float4 q = spos[j];
    float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    if (dx > g.half)
      dx -= g.box;
    else if (dx < -g.half)
      dx += g.box;
    if (dy > g.half)
      dy -= g.box;
    else if (dy < -g.half)
      dy += g.box;
    if (dz > g.half)
      dz -= g.box;
    else if (dz < -g.half)
      dz += g.box;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < g.ll2) unite(parent, (int)k, (int)j);  // fof.c:586-587 (strict)
  }
}

// One thread per sorted particle: the neighbouring cells (x, y columns of
// three cells along z, contiguous in key order) are found by binary search
// in the sorted keys; pairs are linked from their lower index.
__global__ void __launch_bounds__(kThreads)
k_fof_link(LinkGeom g, const unsigned* __restrict__ skeys, const float4* __restrict__ spos, int* parent)
{
  const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (k >= g.np) return;
  const unsigned nd = (unsigned)g.nd;
  const unsigned key = skeys[k];
  const int cz = (int)(key % nd), cy = (int)((key / nd) % nd), cx = (int)(key / (nd * nd));
  const float4 p = spos[k];
  int offs[3], noff;
  if (g.nd >= 3) {
    offs[0] = -1, offs[1] = 0, offs[2] = 1, noff = 3;
  } else {
    offs[0] = 0, offs[1] = 1, noff = g.nd;
  }
  for (int a = 0; a < noff; ++a) {
    const unsigned x = (unsigned)((cx + offs[a] + g.nd) % g.nd);
    for (int b = 0; b < noff; ++b) {
      const unsigned y = (unsigned)((cy + offs[b] + g.nd) % g.nd);
      const unsigned col = (x * nd + y) * nd;
      if (g.nd < 3) {
        link_range(g, skeys, spos, parent, k, p, col, col + nd - 1);
      } else if (cz >= 1 && cz <= g.nd - 2) {
        link_range(g, skeys, spos, parent, k, p, col + cz - 1, col + cz + 1);
      } else if (cz == 0) {
        link_range(g, skeys, spos, parent, k, p, col, col + 1);
        link_range(g, skeys, spos, parent, k, p, col + nd - 1, col + nd - 1);
      } else {
        link_range(g, skeys, spos, parent, k, p, col + nd - 2, col + nd - 1);
        link_range(g, skeys, spos, parent, k, p, col, col);
      }
    }
  }
}

// Group size and smallest record index per root; root of every record.
__global__ void __launch_bounds__(kThreads)
k_fof_roots(long long np, const int* __restrict__ sidx, int* parent, int* __restrict__ count, int* __restrict__ minidx,
            int* __restrict__ rootof)
{
  const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (k >= np) return;
  const int r = find_root(parent, (int)k);
  const int i = sidx[k];
  atomicAdd(&count[r], 1);
  atomicMin(&minidx[r], i);
  rootof[i] = r;
}

__global__ void __launch_bounds__(kThreads)
k_fof_heads(long long np, int nfof_min, const int* __restrict__ rootof, const int* __restrict__ count,
            const int* __restrict__ minidx, unsigned char* __restrict__ head, int* __restrict__ label)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= np) return;
  const int r = rootof[i];
  const bool halo = count[r] >= nfof_min;
  label[i] = halo ? minidx[r] : -1;
  head[i] = halo && minidx[r] == (int)i;
}

__global__ void __launch_bounds__(kThreads)
k_fof_slots(int nhalo, const int* __restrict__ heads, const int* __restrict__ rootof, int* __restrict__ slot)
{
  const int h = blockIdx.x * kThreads + threadIdx.x;
  if (h >= nhalo) return;
  slot[rootof[heads[h]]] = h;
}

// fof.c:284-324 halo_particle: offsets from x0 with the periodic fold of
// HalfBoxSize, and the velocity sum.  Sums are fixed-point integers (offsets
// in units of 2^-32, velocities of 2^-24 km/s), exact for all but the
// smallest values and independent of the order of the atomics, so the
// catalogue is bit-reproducible.
constexpr float kFixX = 4294967296.0f;  // 2^32
constexpr float kFixV = 16777216.0f;    // 2^24

__global__ void __launch_bounds__(kThreads)
k_fof_accumulate(long long np, const float* __restrict__ base, long long stride, float box,
                 const int* __restrict__ rootof, const int* __restrict__ slot, const int* __restrict__ heads,
                 const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                 unsigned long long* __restrict__ sums)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= np) return;
  const int h = slot[rootof[i]];
  if (h < 0) return;
  const int i0 = heads[h];
  const float half = 0.5f * box;
  const float x[3] = {px[i], py[i], pz[i]};
  const float x0[3] = {px[i0], py[i0], pz[i0]};
  for (int a = 0; a < 3; ++a) {
    float dx = x[a] - x0[a];
    if (dx > half)
      dx -= box;
    else if (dx < -half)
      dx += box;
    atomicAdd(&sums[6LL * h + a], (unsigned long long)(long long)rintf(dx * kFixX));
    atomicAdd(&sums[6LL * h + 3 + a], (unsigned long long)(long long)rintf(base[i * stride + 3 + a] * kFixV));
  }
}

// fof.c:1152-1168: centre x0 + mean offset folded into [0, L), mean velocity.
__global__ void __launch_bounds__(kThreads)
k_fof_finalize(int nhalo, const int* __restrict__ heads, const int* __restrict__ rootof,
               const int* __restrict__ count, const float* __restrict__ px, const float* __restrict__ py,
               const float* __restrict__ pz, const unsigned long long* __restrict__ sums, float box,
               cola_halo* __restrict__ out)
{
  const int h = blockIdx.x * kThreads + threadIdx.x;
  if (h >= nhalo) return;
  const int i0 = heads[h];
  const int n = count[rootof[i0]];
  const float x0[3] = {px[i0], py[i0], pz[i0]};
  cola_halo o;
  o.nfof = n;
  for (int a = 0; a < 3; ++a) {
    const double dsum = (double)(long long)sums[6LL * h + a] / (double)kFixX;
    const double vsum = (double)(long long)sums[6LL * h + 3 + a] / (double)kFixV;
    float x = (float)((double)x0[a] + dsum / n);
    if (x < 0.0f) x += box;
    if (x >= box) x -= box;
    o.x[a] = x;
    o.v[a] = (float)(vsum / n);
  }
  out[h] = o;
}

template <typename T>
__global__ void k_fill(T* p, long long n, T v)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// Device buffers of one search, freed on scope exit.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch()
  {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int get(T** p, long long n)
  {
    *p = nullptr;
    if (n <= 0) n = 1;
    if (hipMalloc(p, sizeof(T) * n) != hipSuccess) {
      (void)hipGetLastError();
      set_error("FoF: unable to allocate device memory");
      return kErrAlloc;
    }
    ptrs.push_back(*p);
    return kOk;
  }
};

}  // namespace

int fof_cells_per_axis(float boxsize, float ll)
{
  // Cell side >= 1.01 ll: float rounding of a cell index (x * nd / L) can
  // then never put two particles closer than ll two cells apart.
  int nd = (int)std::floor((double)boxsize / (1.01 * (double)ll));
  if (nd > 1600) nd = 1600;
  return nd < 1 ? 1 : nd;
}

void fof_free(FofResult* r)
{
  if (r->d_label) (void)hipFree(r->d_label);
  r->d_label = nullptr;
  r->halos.clear();
  r->np = 0;
}

int fof_run(const FofInput& in, FofResult* out, hipStream_t s)
{
  if (in.np < 0 || in.np >= INT_MAX || !(in.ll > 0.0f) || !(in.boxsize > 0.0f) || !(in.ll < 0.5f * in.boxsize)) {
    set_error("FoF: need 0 <= np < 2^31 and 0 < ll < L/2");
    return kErrArg;
  }
  fof_free(out);
  const long long np = in.np;
  COLA_HIP(hipMalloc(&out->d_label, sizeof(int) * std::max(np, 1LL)));
  out->np = np;
  if (np == 0) return kOk;

  const int nd = fof_cells_per_axis(in.boxsize, in.ll);
  const float inv_cs = (float)nd / in.boxsize;
  Scratch sc;
  float *px, *py, *pz;
  unsigned *keys, *skeys;
  int *idx, *sidx, *parent, *count, *minidx, *rootof, *slot, *heads, *d_nhalo;
  float4* spos;
  unsigned char* head;
  COLA_TRY(sc.get(&px, np));
  COLA_TRY(sc.get(&py, np));
  COLA_TRY(sc.get(&pz, np));
  COLA_TRY(sc.get(&keys, np));
  COLA_TRY(sc.get(&skeys, np));
  COLA_TRY(sc.get(&idx, np));
  COLA_TRY(sc.get(&sidx, np));
  COLA_TRY(sc.get(&spos, np));
  COLA_TRY(sc.get(&parent, np));
  COLA_TRY(sc.get(&count, np));
  COLA_TRY(sc.get(&minidx, np));
  COLA_TRY(sc.get(&rootof, np));
  COLA_TRY(sc.get(&slot, np));
  COLA_TRY(sc.get(&heads, np));
  COLA_TRY(sc.get(&head, np));
  COLA_TRY(sc.get(&d_nhalo, 1));

  hipLaunchKernelGGL(k_fof_wrap_key, dim3(blocks(np)), dim3(kThreads), 0, s, in.base, in.stride, np, in.boxsize, nd,
                     inv_cs, px, py, pz, keys, idx);
  COLA_HIP(hipGetLastError());

  // sort records by cell key
  const unsigned long long ncell = (unsigned long long)nd * nd * nd;
  int bits = 1;
  while (bits < 32 && (1ULL << bits) < ncell) ++bits;
  size_t tmp_bytes = 0;
  COLA_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, skeys, idx, sidx, (size_t)np, 0, bits, s));
  size_t sel_bytes = 0;
  COLA_HIP(rocprim::select(nullptr, sel_bytes, rocprim::counting_iterator<int>(0), head, heads, d_nhalo, (size_t)np,
                           s));
  void* tmp = nullptr;
  COLA_TRY(sc.get(reinterpret_cast<unsigned char**>(&tmp), (long long)std::max(tmp_bytes, sel_bytes)));
  COLA_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, skeys, idx, sidx, (size_t)np, 0, bits, s));

  hipLaunchKernelGGL(k_fof_gather, dim3(blocks(np)), dim3(kThreads), 0, s, np, sidx, px, py, pz, spos, parent);
  LinkGeom g;
  g.np = np;
  g.nd = nd;
  g.box = in.boxsize;
  g.half = 0.5f * in.boxsize;
  g.ll2 = in.ll * in.ll;  // fof.c:452
  hipLaunchKernelGGL(k_fof_link, dim3(blocks(np)), dim3(kThreads), 0, s, g, skeys, spos, parent);
  hipLaunchKernelGGL(k_fill<int>, dim3(blocks(np)), dim3(kThreads), 0, s, count, np, 0);
  hipLaunchKernelGGL(k_fill<int>, dim3(blocks(np)), dim3(kThreads), 0, s, minidx, np, INT_MAX);
  hipLaunchKernelGGL(k_fill<int>, dim3(blocks(np)), dim3(kThreads), 0, s, slot, np, -1);
  hipLaunchKernelGGL(k_fof_roots, dim3(blocks(np)), dim3(kThreads), 0, s, np, sidx, parent, count, minidx, rootof);
  hipLaunchKernelGGL(k_fof_heads, dim3(blocks(np)), dim3(kThreads), 0, s, np, in.nfof_min, rootof, count, minidx,
                     head, out->d_label);
  COLA_HIP(hipGetLastError());
  COLA_HIP(rocprim::select(tmp, sel_bytes, rocprim::counting_iterator<int>(0), head, heads, d_nhalo, (size_t)np, s));
  int nhalo = 0;
  COLA_HIP(hipMemcpyAsync(&nhalo, d_nhalo, sizeof(int), hipMemcpyDeviceToHost, s));
  COLA_HIP(hipStreamSynchronize(s));
  if (nhalo == 0) return kOk;

  unsigned long long* sums;
  cola_halo* d_halos;
  COLA_TRY(sc.get(&sums, 6LL * nhalo));
  COLA_TRY(sc.get(&d_halos, nhalo));
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(6LL * nhalo)), dim3(kThreads), 0, s, sums, 6LL * nhalo,
                     0ULL);
  hipLaunchKernelGGL(k_fof_slots, dim3(blocks(nhalo)), dim3(kThreads), 0, s, nhalo, heads, rootof, slot);
  hipLaunchKernelGGL(k_fof_accumulate, dim3(blocks(np)), dim3(kThreads), 0, s, np, in.base, in.stride, in.boxsize,
                     rootof, slot, heads, px, py, pz, sums);
  hipLaunchKernelGGL(k_fof_finalize, dim3(blocks(nhalo)), dim3(kThreads), 0, s, nhalo, heads, rootof, count, px, py,
                     pz, sums, in.boxsize, d_halos);
  COLA_HIP(hipGetLastError());
  out->halos.resize(nhalo);
  COLA_HIP(hipMemcpyAsync(out->halos.data(), d_halos, sizeof(cola_halo) * nhalo, hipMemcpyDeviceToHost, s));
  COLA_HIP(hipStreamSynchronize(s));
  return kOk;
}

}  // namespace cola
