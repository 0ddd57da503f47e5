// Hand-written Stockham FFT for the PM and LPT grids.
//
// A 3-D transform of the padded real layout [n][n][2(n/2+1)] runs as three
// tile passes.  Each pass loads a tile of T independent 1-D transforms of
// length n into LDS (layout [position][tile column], column fastest), runs
// the radix passes in registers with one LDS exchange per radix, and writes
// the tile back.  The z pass works on rows (two real rows packed into one
// complex transform); the y and x passes work on strided columns, reading T
// adjacent complex columns per row so every global access is a contiguous
// T*8-byte segment.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cola {

constexpr int kFftThreads = 256;
constexpr int kFftEmax = 48;  // complex values one thread holds per pass
constexpr int kFftMaxTile = kFftThreads * kFftEmax;  // n * T limit (96 KiB of LDS)

struct FftRadices {
  int n = 0;
  int npass = 0;
  int radix[8] = {0};
};

// Multiplier applied while loading the first (x) pass of an inverse
// transform: the k-space operators of the PM force and of 2LPT.
enum class KOp : int {
  None = 0,
  PmPoisson = 1,  // pm.c:322-346  F = i f2 delta, f2 = f1 n_axis / |n|^2
  LptGrad = 2,    // lpt.c:379-395 i k_dir * disp
  Lpt2 = 3,       // lpt.c:479-486 source * k_axis / (i k^2)
};

struct KOpParams {
  int axis = 0;       // PmPoisson / Lpt2 axis, LptGrad derivative direction
  float f1 = 0.0f;    // pm.c:317
  double box = 0.0;   // lpt.c kvec = n * 2 * pi / Box
};

// Per-size plan: radix schedule, twiddles and tile widths of the three passes.
struct FftPlan {
  int n = 0;
  FftRadices rad;
  float2* d_tw = nullptr;  // exp(-2 pi i k / n), k < n
  int log2T_rows = 0, log2T_y = 0, log2T_x = 0;

  // Optional per-pass profiling hook: pass 0 z r2c, 1 y fwd, 2 x fwd,
  // 3 x inv, 4 y inv, 5 z c2r; begin = 1 before the launch, 0 after.
  void (*hook)(void* user, int pass, int begin) = nullptr;
  void* hook_user = nullptr;

  int init(int n, int force_log2T = -1);
  void destroy();

  // In place on the padded layout: real -> half-spectrum [x][y][kz].
  int r2c(float* data, hipStream_t s) const;
  // c2c over x (optionally out of place from in_k with a k-space operator),
  // then y, then c2r over z, in place on out.
  int c2r(const float2* in_k, float* out, KOp op, const KOpParams& kp, hipStream_t s) const;
};

// Radix schedule for n = 2^a 3^b 5^c 7^d; returns false otherwise.
bool fft_factor(int n, FftRadices* r);
// Largest tile width (log2) with n*T <= kFftMaxTile that keeps every radix
// pass within the per-thread register budget; capped at cap_log2.
int fft_tile_log2(const FftRadices& r, int cap_log2);

}  // namespace cola
