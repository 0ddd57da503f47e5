// Coarse density grid (coarse_grid.c:17-121) and the shell power spectrum
// estimator.  Both run on positions already in HBM (the particle state or the
// last snapshot), so an output costs one pass over the particles plus the grid.
#include "analysis.h"
#include "common.h"

namespace cola {
namespace {

constexpr int kThreads = 256;

__device__ inline double wave_sum_d(double v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

constexpr float kFix = 4294967296.0f;  // 2^32
constexpr double kUnfix = 1.0 / 4294967296.0;

__device__ inline void add_fix(unsigned long long* cell, float w)
{
  atomicAdd(cell, (unsigned long long)rintf(w * kFix));
}

// coarse_grid.c:95-121.  One thread per particle; the index arithmetic and
// the weight products follow the reference term by term: X = x * dx_inv in
// float, ix = floor(X), w = 1 - (X - ix), neighbours (ix + nc) % nc and
// (ix + 1 + nc) % nc (computed here with a modulo that also holds below -nc,
// where the reference would index out of bounds).
__global__ void __launch_bounds__(kThreads)
k_cic_periodic(PosSource src, int n, float dx_inv, unsigned long long* __restrict__ fix)
{
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= src.np) return;
  int i0[3], i1[3];
  float w[3];
  for (int j = 0; j < 3; ++j) {
    const float X = src.x[j][i * src.stride] * dx_inv;
    const int ix = (int)floorf(X);
    w[j] = 1.0f - (X - (float)ix);
    int m = ix % n;
    if (m < 0) m += n;
    i0[j] = m;
    i1[j] = m + 1 == n ? 0 : m + 1;
  }
  const long long r00 = ((long long)i0[0] * n + i0[1]) * n, r01 = ((long long)i0[0] * n + i1[1]) * n;
  const long long r10 = ((long long)i1[0] * n + i0[1]) * n, r11 = ((long long)i1[0] * n + i1[1]) * n;
  add_fix(&fix[r00 + i0[2]], w[0] * w[1] * w[2]);
  add_fix(&fix[r01 + i0[2]], w[0] * (1 - w[1]) * w[2]);
  add_fix(&fix[r00 + i1[2]], w[0] * w[1] * (1 - w[2]));
  add_fix(&fix[r01 + i1[2]], w[0] * (1 - w[1]) * (1 - w[2]));
  add_fix(&fix[r10 + i0[2]], (1 - w[0]) * w[1] * w[2]);
  add_fix(&fix[r11 + i0[2]], (1 - w[0]) * (1 - w[1]) * w[2]);
  add_fix(&fix[r10 + i1[2]], (1 - w[0]) * w[1] * (1 - w[2]));
  add_fix(&fix[r11 + i1[2]], (1 - w[0]) * (1 - w[1]) * (1 - w[2]));
}

// coarse_grid.c:63-68: delta = nbar_inv * rho - 1 (float), with the double
// sums of rho and delta.  Cell c of the n^3 grid goes to (c / n) * row + c % n.
__global__ void __launch_bounds__(kThreads)
k_density_contrast(long long ncell, int n, long long row, float nbar_inv, const unsigned long long* __restrict__ fix,
                   float* __restrict__ grid, double* __restrict__ sums)
{
  __shared__ double part[2][kThreads / 64];
  double s_in = 0.0, s_out = 0.0;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < ncell; c += stride) {
    const float rho = (float)((double)fix[c] * kUnfix);
    const float d = nbar_inv * rho - 1.0f;
    grid[(c / n) * row + c % n] = d;
    s_in += (double)rho;
    s_out += (double)d;
  }
  s_in = wave_sum_d(s_in);
  s_out = wave_sum_d(s_out);
  if (__lane_id() == 0) {
    part[0][threadIdx.x >> 6] = s_in;
    part[1][threadIdx.x >> 6] = s_out;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += part[threadIdx.x][w];
    unsafeAtomicAdd(&sums[threadIdx.x], t);
  }
}

__device__ inline double sinc2(double x)
{
  if (x == 0.0) return 1.0;
  const double s = sin(x) / x;
  return s * s;
}

// Shell sums over the half spectrum.  Each workgroup accumulates its share in
// an LDS histogram and flushes the touched bins with one atomic per value.
__global__ void __launch_bounds__(kThreads)
k_power_shells(int n, const float2* __restrict__ dk, ShellBins b, double* __restrict__ acc)
{
  extern __shared__ double hist[];
  for (int e = threadIdx.x; e < 3 * b.nbins; e += kThreads) hist[e] = 0.0;
  __syncthreads();
  const int nzc = n / 2 + 1;
  const long long total = (long long)n * n * nzc;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total; idx += stride) {
    const int kz = (int)(idx % nzc);
    const long long line = idx / nzc;
    const int iy = (int)(line % n), ix = (int)(line / n);
    const int mx = ix <= n / 2 ? ix : ix - n;
    const int my = iy <= n / 2 ? iy : iy - n;
    const long long kk2 = (long long)mx * mx + (long long)my * my + (long long)kz * kz;
    if (kk2 == 0) continue;
    const double k = b.kf * sqrt((double)kk2);
    const double t = (k - b.kmin) / b.dk;
    if (!(t >= 0.0) || t >= (double)b.nbins) continue;
    const int bin = (int)floor(t);
    const float2 v = dk[idx];
    double p = ((double)v.x * v.x + (double)v.y * v.y) * b.norm;
    if (b.deconvolve) {
      const double w = sinc2(kPiD * mx / n) * sinc2(kPiD * my / n) * sinc2(kPiD * kz / n);
      p /= w * w;
    }
    const double wt = (kz == 0 || 2 * kz == n) ? 1.0 : 2.0;
    unsafeAtomicAdd(&hist[3 * bin], wt * k);
    unsafeAtomicAdd(&hist[3 * bin + 1], wt * p);
    unsafeAtomicAdd(&hist[3 * bin + 2], wt);
  }
  __syncthreads();
  for (int bin = threadIdx.x; bin < b.nbins; bin += kThreads) {
    if (hist[3 * bin + 2] == 0.0) continue;
    for (int q = 0; q < 3; ++q) unsafeAtomicAdd(&acc[3 * bin + q], hist[3 * bin + q]);
  }
}

template <typename T>
__global__ void k_fill_zero(T* p, long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T(0);
}

unsigned capped_blocks(long long work, long long per_block, unsigned cap)
{
  const long long b = (work + per_block - 1) / per_block;
  return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

int launch_fill_zero(float* p, long long n, hipStream_t s)
{
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(k_fill_zero<float>, dim3(capped_blocks(n, kThreads * 8LL, 65536)), dim3(kThreads), 0, s, p, n);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_fill_zero(double* p, long long n, hipStream_t s)
{
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(k_fill_zero<double>, dim3(capped_blocks(n, kThreads * 8LL, 65536)), dim3(kThreads), 0, s, p,
                     n);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_cic_periodic(const PosSource& src, int n, float boxsize, unsigned long long* fix, hipStream_t s)
{
  if (n < 1) {
    set_error("coarse grid: invalid grid size");
    return kErrArg;
  }
  const long long ncell = (long long)n * n * n;
  hipLaunchKernelGGL(k_fill_zero<unsigned long long>, dim3(capped_blocks(ncell, kThreads * 8LL, 65536)),
                     dim3(kThreads), 0, s, fix, ncell);
  COLA_HIP(hipGetLastError());
  if (src.np == 0) return kOk;
  const float dx_inv = n / boxsize;  // coarse_grid.c:99
  hipLaunchKernelGGL(k_cic_periodic, dim3((unsigned)((src.np + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, src,
                     n, dx_inv, fix);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_density_contrast(int n, long long row, float nbar_inv, const unsigned long long* fix, float* grid,
                            double* sums, hipStream_t s)
{
  const long long ncell = (long long)n * n * n;
  hipLaunchKernelGGL(k_density_contrast, dim3(capped_blocks(ncell, kThreads * 16LL, 8192)), dim3(kThreads), 0, s,
                     ncell, n, row, nbar_inv, fix, grid, sums);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_power_shells(int n, const float2* delta_k, const ShellBins& b, double* acc, hipStream_t s)
{
  if (b.nbins < 1 || b.nbins > kMaxShellBins) {
    set_error("power spectrum: nbins must be in [1, 2048]");
    return kErrArg;
  }
  const long long total = (long long)n * n * (n / 2 + 1);
  const size_t lds = sizeof(double) * 3 * b.nbins;
  hipLaunchKernelGGL(k_power_shells, dim3(capped_blocks(total, kThreads * 16LL, 4096)), dim3(kThreads), lds, s, n,
                     delta_k, b, acc);
  COLA_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cola
