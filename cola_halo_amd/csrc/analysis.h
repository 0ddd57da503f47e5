// Output-side grid kernels on device particle positions: the coarse density
// grid of coarse_grid.c and a shell-averaged power spectrum estimator built on
// the same assignment and the library's r2c FFT.
#pragma once

#include <hip/hip_runtime.h>

namespace cola {

// np positions, component a of particle i at x[a][i * stride] (stride 1 for
// the structure-of-arrays state, 8 floats for the 32-byte snapshot records).
struct PosSource {
  const float* x[3] = {nullptr, nullptr, nullptr};
  long long stride = 1;
  long long np = 0;
};

// coarse_grid.c:95-121 on an n^3 grid: zero `fix` (n^3 cells, C order),
// then add every particle's eight CIC weights with periodic wrap as 32.32
// fixed-point integers (exact for weights >= 2^-9), so the sums do not depend
// on the order of the atomics and the grid is bit-reproducible.
int launch_cic_periodic(const PosSource& src, int n, float boxsize, unsigned long long* fix, hipStream_t s);

// coarse_grid.c:59-70: rho = (float) fixed-point sum, grid = nbar_inv * rho - 1
// (float, as the reference), written to a grid with rows of `row` floats (n
// for the C-order file layout, 2(n/2+1) for an in-place r2c).  sums[0] += sum
// of rho, sums[1] += sum of delta (sums zeroed by the caller).
int launch_density_contrast(int n, long long row, float nbar_inv, const unsigned long long* fix, float* grid,
                            double* sums, hipStream_t s);

// Shell bins of |k| = kf * |m| on [kmin, kmin + nbins * dk).
struct ShellBins {
  double kmin = 0, dk = 0, kf = 0;
  double norm = 0;  // L^3 / n^6: |DFT|^2 -> P(k)
  int nbins = 0;
  int deconvolve = 0;  // divide by the CIC window squared
};
constexpr int kMaxShellBins = 2048;

// Accumulates, over the half spectrum [x][y][kz] (kz <= n/2) of an r2c
// output, acc[3b] += weight * |k|, acc[3b+1] += weight * P, acc[3b+2] +=
// weight, where weight 2 stands for the conjugate mode of 0 < kz < n/2.
// acc (3 * nbins doubles) is zeroed by the caller; k = 0 is skipped.
int launch_power_shells(int n, const float2* delta_k, const ShellBins& b, double* acc, hipStream_t s);

// Fill n floats / doubles with zero (a kernel, so the call may be captured).
int launch_fill_zero(float* p, long long n, hipStream_t s);
int launch_fill_zero(double* p, long long n, hipStream_t s);

}  // namespace cola
