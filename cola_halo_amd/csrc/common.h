// Shared definitions of the HIP library: error plumbing and small device
// helpers.  All device code is compiled with -ffp-contract=off so every float
// and double operation rounds exactly where the reference's C code rounds
// (the transforms, which the reference leaves to FFTW, call fmaf explicitly).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "errors.h"

namespace cola {

// Status codes reuse the reference's msg_abort codes (see cola_halo_amd.h).
enum Status : int {
  kOk = 0,
  kErrAlloc = 4001,
  kErrPowerFile = 3000,
  kErrSigma8 = 3010,
  kErrMass = 6040,
  kErrArg = 9001,
  kErrUnsupported = 9002,
  kErrHip = 9003,
  kErrState = 9004,
  kErrNetFile = 9005,
};

// Return a status from the enclosing function if a HIP call fails.
#define COLA_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::cola::set_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
      return ::cola::kErrHip;                                                           \
    }                                                                                   \
  } while (0)

#define COLA_TRY(expr)         \
  do {                         \
    int s_ = (expr);           \
    if (s_ != 0) return s_;    \
  } while (0)

__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Workgroups whose data overlap (adjacent FFT column tiles sharing
// 128-byte lines, lattice rows of particles gathering from the same grid
// rows) go in groups of G to the same XCD in consecutive dispatch rounds:
// block 8G q + 8h + x (XCD x, h < G) -> logical block 8G q + G x + h.  A
// bijection of [0, nb); the last nb mod 8G blocks keep their index.
__device__ inline int xcd_group_remap(int b, int nb, int G)
{
  const int span = 8 * G;
  if (G <= 1 || b >= nb - nb % span) return b;
  const int q = b / span, r = b - q * span;
  return span * q + G * (r & 7) + (r >> 3);
}

// Batched launches list their members' workgroups back to back: member m
// owns workgroups [start[m], start[m+1]).  The member of workgroup b (the
// last one whose range starts at or before b; empty ranges are skipped).
__device__ __forceinline__ int batch_find(const int* __restrict__ start, int n, int b)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

constexpr double kPiD = 3.14159265358979323846;

}  // namespace cola
