// Hand-written Stockham FFT for the PM and LPT grids.
//
// A 3-D transform of the padded real layout [n][n][2(n/2+1)] runs as three
// tile passes, one per axis.  A workgroup owns T independent 1-D transforms
// of length n:
//   * the first radix pass loads its butterfly inputs straight from HBM into
//     registers (all loads of a thread in flight at once),
//   * the radix passes exchange data through one LDS tile (one write and one
//     read per radix boundary), with the twiddle table staged in LDS,
//   * the last radix pass stores its outputs straight from registers.
// Column passes (y, x) give each thread a column index fastest, so every
// global access is a contiguous T*8-byte segment of a row; the row pass (z)
// gives each thread a position fastest, so it streams whole rows, two real
// rows packed into one complex transform.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cola {

constexpr int kFftMaxThreads = 1024;
constexpr int kFftMaxTile = 12288;  // n * T complex values per workgroup (96 KiB)

// One radix pass: radix R, L = product of the earlier radices, m = n / R,
// twiddle stride n / (R L) and magic multipliers for division by L and m.
struct FftPass {
  int R = 0, L = 0, m = 0, twstep = 0;
  unsigned magL = 0, magM = 0;
};

struct FftRadices {
  int n = 0;
  int npass = 0;
  FftPass pass[8];
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

// Launch shape of one axis pass: 2^log2T transforms per workgroup, threads
// per workgroup, register slots per thread.
struct FftPassGeom {
  int log2T = 0, threads = 0, emax = 0;
};

// Per-size plan: radix schedule, twiddles and tile shapes of the three passes.
struct FftPlan {
  int n = 0;
  int schedule = 0;
  FftRadices rad;
  float2* d_tw = nullptr;  // exp(-2 pi i k / n), k < n
  FftPassGeom grows, gy, gx;
  int xgroup = 8;  // adjacent column tiles per XCD dispatch group (measured best)
  // Complex row stride of the half spectrum (>= n/2+1): rows padded to a
  // multiple of 16 complex values start on 128-byte lines, so column tiles
  // of 8 values never straddle a line; real rows are 2 nzp floats.
  int nzp = 0;

  // Optional per-pass profiling hook: pass 0 z r2c, 1 y fwd, 2 x fwd,
  // 3 x inv, 4 y inv, 5 z c2r, 6 fused x; begin = 1 before the launch, 0 after.
  void (*hook)(void* user, int pass, int begin) = nullptr;
  void* hook_user = nullptr;

  // log2T_cols >= 0 fixes the tile width (log2 transforms per workgroup) of
  // the y and x column passes, log2T_rows >= 0 that of the z row passes
  // (row pairs); -1 picks the measured defaults.
  // pad_rows: complex rows of round_up(n/2+1, 16) values instead of n/2+1.
  // log2T_xcols >= 0 gives the x column pass (plain or fused) its own width.
  // schedule: the radix schedule (fft_factor).
  int init(int n, int log2T_cols = -1, int log2T_rows = -1, bool pad_rows = false, int log2T_xcols = -1,
           int schedule = 0);
  void destroy();

  // In place on the padded layout (real rows of 2 nzp floats): real ->
  // half-spectrum [x][y][kz] with complex rows of nzp values.
  int r2c(float* data, hipStream_t s) const;
  // c2c over x (optionally out of place from in_k with a k-space operator),
  // then y, then c2r over z, in place on out.
  int c2r(const float2* in_k, float* out, KOp op, const KOpParams& kp, hipStream_t s) const;
  // The PM force fields in one sequence (pm.c:282-355): r2c of the padded
  // real density (in place; destroyed), then for each axis the Poisson
  // gradient and c2r into out[axis].  The forward x pass, the three Poisson
  // operators and the three inverse x passes run as one fused kernel
  // (hook pass 6); delta(k) is never written to memory.
  int forces(float* density, float* const out[3], const KOpParams& kp, hipStream_t s) const;
  bool can_fuse() const;
};

// One launch of a batched force transform: one pass of every grid.
struct FftBatchStage {
  int pass = 0;  // 0 z r2c, 1 y forward, 2 fused x, 3 y inverse (3 axes), 4 z c2r (3 axes)
  int emax = 0, rs = 0, threads = 0;
  size_t lds = 0;
  int blocks = 0, njobs = 0;
  const void* jobs = nullptr;  // device job table
  const int* start = nullptr;  // device: first workgroup of each job, njobs + 1 entries
};

// The force transforms (FftPlan::forces) of many grids, batched per pass
// over the workgroups of every grid: z r2c, y forward, fused x, the y inverse
// of the three axes, the z c2r of the three axes (instead of nine launches
// per grid).  Each grid keeps its own plan's radix schedule and tile
// shapes; a launch runs its jobs with the widest workgroup and register
// tile any of them needs, and the butterflies of a job do not depend on
// either, so the outputs equal each plan's forces() bit for bit.
// merge_all: one launch per pass for all grids (the kernel variant with the
// most registers any grid needs); else one launch per pass and kernel
// variant (register tile, radix set) that the grids' own launches use.
// Every plan must fuse (FftPlan::can_fuse).
struct FftForcesBatch {
  std::vector<FftBatchStage> stages;
  void* d_mem = nullptr;
  // plans[i] transforms density[i] (destroyed) into out[3 i + axis] with
  // the operator parameters kp[i].
  int init(const FftPlan* const* plans, float* const* density, float* const* out, const KOpParams* kp, int n,
           bool merge_all);
  int run(hipStream_t s) const;
  void destroy();
};

// Radix schedules of n = 2^a 3^b 5^c (the order and grouping of the radix
// passes; every schedule computes the same DFT, rounding differs):
//   0  powers of two as radix 8 (a leftover 2 as 4 x 4) first, then 3s, 5s
//   1  3s and 5s first, then the powers of two as in 0
//   2  powers of two as radix 4 (an odd power adds one radix 8), then 3s, 5s
//   3  the powers of two of 0 in reverse order (smallest first), then 3s, 5s
//   4  as 0 with its last power-of-two radix moved to the second pass
constexpr int kFftSchedules = 5;
// Radix schedule `schedule` for n; false if n has another prime factor or
// the schedule needs more than 8 passes.
bool fft_factor(int n, FftRadices* r, int schedule = 0);
// Register slots one thread needs for tile width 2^log2T with the given
// workgroup size (0 if the tile exceeds kFftMaxTile or 24 slots).
int fft_emax(const FftRadices& r, int log2T, int threads);
FftPassGeom pass_geom(const FftRadices& r, int log2T);

}  // namespace cola
