// Learned force residual on MFMA: implicit-GEMM 3-D convolutions in bfloat16
// with fp32 accumulation (forward, data gradient, weight gradient), the loss,
// Adam and the glue to the PM grids.
//
// Layout: activations channel-last, act[cell][16] (input [cell][8]), cell =
// (x n + y) n + z, periodic.  A convolution is the GEMM
//   out[cell][co] = sum over taps t and input channels ci of
//                   in[cell + d_t][ci] W[t][ci][co],   K = 27 cin,
// computed per wave as a 16-cell (along z) x 16-channel tile with
// v_mfma_f32_16x16x32_bf16.  Its lanes hold (g = lane / 16, r = lane % 16):
// A rows r, k = 8g..8g+7; B k = 8g..8g+7, column r; D rows 4g..4g+3,
// column r.  Every cell operand is 8 consecutive channels of one cell (one
// 16-byte load); the K order of each kernel is given with it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "common.h"
#include "resnet.h"

namespace cola {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

__host__ __device__ inline unsigned short f2bf(float f)
{
  unsigned u = __builtin_bit_cast(unsigned, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                         // round to nearest even
  return (unsigned short)(u >> 16);
}

__host__ __device__ inline float bf2f(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

// asinh(x) to a few float ulps, cheaper than the libm-grade asinhf: the odd
// Taylor series to x^7 for |x| < 1/8 (truncation below 1e-9 relative), else
// ln(|x| + sqrt(x^2 + 1)) from the hardware log2 and square root (argument
// >= 1.007, no cancellation), ln(2|x|) beyond 1e9.  The feature is rounded
// to bfloat16 (8 bits) afterwards.
__device__ inline float feature_asinh(float x)
{
  constexpr float kLn2 = 0.693147180559945f;
  const float ax = fabsf(x);
  float r;
  if (ax < 0.125f) {
    const float x2 = ax * ax;
    r = ax * (1.0f + x2 * (-1.0f / 6.0f + x2 * (3.0f / 40.0f + x2 * (-15.0f / 336.0f))));
  } else if (ax < 1e9f) {
    r = __builtin_amdgcn_logf(ax + __builtin_amdgcn_sqrtf(ax * ax + 1.0f)) * kLn2;
  } else {
    r = (__builtin_amdgcn_logf(ax) + 1.0f) * kLn2;
  }
  return copysignf(r, x);
}

// The net's input feature of a density contrast: asinh(delta) * scale in bf16.
__device__ inline unsigned short feature_of(float delta, float scale) { return f2bf(feature_asinh(delta) * scale); }

__device__ inline int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ inline long long cell_at(int x, int y, int z, int t, int sign, int n)
{
  const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
  return ((long long)wrapi(x + sign * dx, n) * n + wrapi(y + sign * dy, n)) * n + wrapi(z + sign * dz, n);
}

enum ConvMode { kFwdRelu = 0, kFwdLinear = 1, kBwdData = 2, kFwdAdd = 3 };

// Outputs of a convolution by mode: bf16 activations (kFwdRelu; kBwdData
// masked by relu'(hprev)), the fp32 net output y [3][n^3] (kFwdLinear), or
// scale * y added in place to three grids with rows of `row` floats
// (kFwdAdd: the corrected PM force, out[a] = out[a] + out_scale * y[a]).
struct ConvOut {
  unsigned short* bf = nullptr;
  float* f = nullptr;
  const unsigned short* hprev = nullptr;
  float* add[3] = {nullptr, nullptr, nullptr};
  long long row = 0;
  float scale = 0.0f;
  double* sum = nullptr;  // kFwdAdd: the three components' sums of what was added
  float* fi = nullptr;    // kFwdAdd: the corrected forces go here, interleaved [cell][3], instead of add[]
  float* trash = nullptr; // k_conv3h: 64 floats the lanes without an output cell store to
};

// 16-channel convolution with operand reuse.  A wave computes kConvTy output
// tiles (16 cells along z x 16 channels) at rows y0 .. y0 + kConvTy - 1 of
// one x plane, streaming the kConvTy + 2 input rows of each of the 3 x
// offsets once.  Each input row is two 16-byte loads per lane: cell z0 + r
// and its z-shifted neighbour (z - 1 for lane groups 0-1, z + 1 for 2-3),
// channels 8 (g % 2) .. + 7 (loading the neighbour directly from L1 measured
// faster than shifting the row across lanes with DPP).  K is ordered so that
// every MFMA operand is a row, its shifted copy or a lane-group select of two
// rows:
//   chunk 3 (dx+1) + (dy+1), dx, dy in {-1,0,1}: taps (dx,dy,-1) | (dx,dy,+1)
//   chunk 9 + (dx+1):                          taps (dx,-1,0) | (dx,0,0)
//   chunk 12:                                  taps (-1,+1,0) | (0,+1,0)
//   chunk 13:                                  taps (+1,+1,0) | none
// (left of | for lane groups 0 and 1, right for 2 and 3).
constexpr int kConvTy = 8;

// A workgroup (4 waves) owns a column of blocks of 2 x 16 cells along z by
// 2 kConvTy rows along y and marches along x through `march` consecutive x
// planes (a unit); wave w computes z tile w % 2 and the kConvTy rows of half
// w / 2 of each plane.  Its input planes (the rows and the z cells with one
// halo cell each side, periodic) live in a ring of 4 LDS slots filled by
// direct global-to-LDS loads (global_load_lds_dwordx4: no registers):
// output plane x reads the slots of planes x - 1, x, x + 1 while plane x + 2
// streams into the fourth, so each input plane is loaded once per unit
// instead of three times.
constexpr int kBlockZ = 32, kBlockY = 2 * kConvTy;
constexpr int kBrickZ = kBlockZ + 2, kBrickY = kBlockY + 2;
constexpr int kPlaneCells = kBrickY * kBrickZ;
constexpr int kRingSlots = 4;

// Ring cell of plane p (p >= -1) at brick row yy and brick z zz.
__device__ inline int ring_cell(int p, int yy, int zz) { return (((p + kRingSlots) & (kRingSlots - 1)) * kBrickY + yy) * kBrickZ + zz; }

// Starts the loads of input plane p (periodic) of the column (y0, z0) into
// its ring slot: QPC 16-byte quarters per cell (2 for 16 channels, 1 for 8).
// Chunk c of the plane goes to its slot's chunk c; a wave's 64 consecutive
// chunks land at its uniform LDS base + 16 lane.  The caller waits.
template <int QPC>
__device__ inline void load_plane(uint4* ring, const unsigned short* __restrict__ in, int n, int p, int y0, int z0)
{
  const int lane = threadIdx.x & 63;
  uint4* slot = ring + ring_cell(p, 0, 0) * QPC;
  const long long xoff = (long long)wrapi(p, n) * n;
  for (int base = threadIdx.x & ~63; base < kPlaneCells * QPC; base += kThreads) {
    const int c = base + lane;
    if (c < kPlaneCells * QPC) {
      const int q = c % QPC, cell = c / QPC;
      const int zz = cell % kBrickZ, yy = cell / kBrickZ;
      const long long src = (xoff + wrapi(y0 + yy - 1, n)) * n + wrapi(z0 + zz - 1, n);
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(in + src * (8 * QPC) + q * 8),
                                       (__attribute__((address_space(3))) void*)(slot + base), 16, 0, 0);
    }
  }
}

// Unit u of a march: column (x-segment, y block, z block).
struct ConvUnit {
  int x0, x1, y0, z0;
};

__device__ inline ConvUnit conv_unit(long long u, int n, int march)
{
  const int nzb = n / kBlockZ, nyb = n / kBlockY;
  ConvUnit c;
  c.z0 = (int)(u % nzb) * kBlockZ;
  const long long t2 = u / nzb;
  c.y0 = (int)(t2 % nyb) * kBlockY;
  c.x0 = (int)(t2 / nyb) * march;
  c.x1 = min(c.x0 + march, n);
  return c;
}

__device__ inline uint4 sel4(bool lo, uint4 a, uint4 b)
{
  return lo ? a : b;
}

// Output tile transposed: the packed weights go in as the A operand (row =
// output channel) and the input cells as B (column = cell), both with the
// register layout [lane g, r][j] = k 8g + j of row/column r.  Lane (g, r)
// then holds output channels 4g..4g+3 of cell r: 8 contiguous bytes.
__device__ inline floatx4 mfma(uint4 x, uint4 w, floatx4 acc)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, w), __builtin_bit_cast(bf16x8, x), acc, 0,
                                                 0, 0);
}

// Four fp32 values -> four bfloat16 (round to nearest even) in 8 bytes.
__device__ inline uint2 pack_bf16x4(float a, float b, float c, float d)
{
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  const bf16x2 lo = {(__bf16)a, (__bf16)b}, hi = {(__bf16)c, (__bf16)d};
  return make_uint2(__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi));
}

// Products of one output plane x in the K order of k_conv16: the kConvTy
// output rows of wave row half wy, lanes at brick z zc (centre) and zs (z
// neighbour), from the 16-channel input planes x - 1, x, x + 1 whose ring
// slots start at cells sb[0], sb[1], sb[2] (slots of kPlaneCells cells, two
// uint4 halves per cell).  Operands of input row j + 1 are read from LDS
// while row j's products run (one row of software pipelining hides the LDS
// latency).
__device__ inline void conv16_plane(const uint4* __restrict__ ring, const int (&sb)[3], int wy, int zc, int zs,
                                    int half, bool lo, const uint4 (&b)[14], floatx4 (&acc)[kConvTy])
{
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
  for (int t = 0; t < kConvTy; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  uint4 prev[3] = {zero4, zero4, zero4};
  uint4 cen[3], sh[3];
  const int row0 = kConvTy * wy;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cen[d] = ring[(sb[d] + row0 * kBrickZ + zc) * 2 + half];
    sh[d] = ring[(sb[d] + row0 * kBrickZ + zs) * 2 + half];
  }
#pragma unroll
  for (int j = 0; j < kConvTy + 2; ++j) {  // input row j of the half feeds output rows j - 2, j - 1, j
    uint4 ncen[3], nsh[3];
    if (j + 1 < kConvTy + 2) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        ncen[d] = ring[(sb[d] + (row0 + j + 1) * kBrickZ + zc) * 2 + half];
        nsh[d] = ring[(sb[d] + (row0 + j + 1) * kBrickZ + zs) * 2 + half];
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of this row's products
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (j >= 2) acc[j - 2] = mfma(sh[d], b[3 * d + 2], acc[j - 2]);                     // dy = +1
      if (j >= 1 && j - 1 < kConvTy) acc[j - 1] = mfma(sh[d], b[3 * d + 1], acc[j - 1]);  // dy = 0
      if (j < kConvTy) acc[j] = mfma(sh[d], b[3 * d], acc[j]);                            // dy = -1
    }
    if (j >= 1 && j - 1 < kConvTy) {
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[j - 1] = mfma(sel4(lo, prev[d], cen[d]), b[9 + d], acc[j - 1]);
    }
    if (j >= 2) {
      acc[j - 2] = mfma(sel4(lo, cen[0], cen[1]), b[12], acc[j - 2]);
      acc[j - 2] = mfma(sel4(lo, cen[2], zero4), b[13], acc[j - 2]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      prev[d] = cen[d];
      if (j + 1 < kConvTy + 2) {
        cen[d] = ncen[d];
        sh[d] = nsh[d];
      }
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(kThreads)
k_conv16(const unsigned short* __restrict__ in, const unsigned short* __restrict__ wpk, const float* __restrict__ bias,
         int cout, int n, int march, ConvOut o)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const bool lo = g < 2;
  uint4 b[14];
#pragma unroll
  for (int c = 0; c < 14; ++c) b[c] = *reinterpret_cast<const uint4*>(wpk + ((long long)c * 64 + lane) * 8);
  float bch[4];  // bias of output channels 4g..4g+3
#pragma unroll
  for (int i = 0; i < 4; ++i) bch[i] = (MODE == kBwdData || 4 * g + i >= cout) ? 0.0f : bias[4 * g + i];
  extern __shared__ uint4 ring[];  // [slot][yy][zz][2 halves of 8 channels]
  const int half = g & 1;
  const int wz = (threadIdx.x >> 6) & 1, wy = threadIdx.x >> 7;
  const int zc = 1 + 16 * wz + r;          // brick z of the lane's cell
  const int zs = lo ? zc - 1 : zc + 1;     // its z neighbour for the dz = -1 | +1 taps
  float added[3] = {0.0f, 0.0f, 0.0f};     // kFwdAdd: sums of the residual this lane added
  const long long units = (long long)((n + march - 1) / march) * (n / kBlockY) * (n / kBlockZ);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const ConvUnit cu = conv_unit(u, n, march);
    __syncthreads();  // the previous unit's operand reads are done
    for (int p = cu.x0 - 1; p <= cu.x0 + 1; ++p) load_plane<2>(ring, in, n, p, cu.y0, cu.z0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int x = cu.x0; x < cu.x1; ++x) {
      if (x + 2 <= cu.x1) load_plane<2>(ring, in, n, x + 2, cu.y0, cu.z0);  // lands during this plane
      // kFwdAdd: the force values this lane updates, loaded while the products run
      float fold[kConvTy][3];
      if (MODE == kFwdAdd && g == 0) {
#pragma unroll
        for (int t = 0; t < kConvTy; ++t) {
          const long long at = ((long long)x * n + cu.y0 + kConvTy * wy + t) * o.row + cu.z0 + 16 * wz + r;
#pragma unroll
          for (int i = 0; i < 3; ++i) fold[t][i] = o.add[i][at];
        }
      }
      floatx4 acc[kConvTy];
      int sb[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) sb[d] = ring_cell(x - 1 + d, 0, 0);
      conv16_plane(ring, sb, wy, zc, zs, half, lo, b, acc);
      // plane x + 2 (and the previous plane's stores) complete; every wave is
      // past its reads of plane x - 1, whose slot the next prefetch takes
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {
        const long long line = (long long)x * n + cu.y0 + kConvTy * wy + t;
        const long long cell = line * n + cu.z0 + 16 * wz + r;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = acc[t][i] + bch[i];
        if (MODE == kFwdRelu) {
          *reinterpret_cast<uint2*>(o.bf + cell * 16 + 4 * g) =
              pack_bf16x4(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f), fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f));
        } else if (MODE == kFwdLinear) {
          if (g == 0) {
            const long long nc = (long long)n * n * n;
#pragma unroll
            for (int i = 0; i < 3; ++i) o.f[i * nc + cell] = v[i];
          }
        } else if (MODE == kFwdAdd) {
          if (g == 0) {
            const long long at = line * o.row + cu.z0 + 16 * wz + r;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const float dv = o.scale * v[i];
              if (o.fi)
                o.fi[(line * n + cu.z0 + 16 * wz + r) * 3 + i] = fold[t][i] + dv;
              else
                o.add[i][at] = fold[t][i] + dv;
              added[i] += dv;
            }
          }
        } else {
          const uint2 hp = *reinterpret_cast<const uint2*>(o.hprev + cell * 16 + 4 * g);
          const float m0 = bf2f((unsigned short)(hp.x & 0xffffu)), m1 = bf2f((unsigned short)(hp.x >> 16));
          const float m2 = bf2f((unsigned short)(hp.y & 0xffffu)), m3 = bf2f((unsigned short)(hp.y >> 16));
          *reinterpret_cast<uint2*>(o.bf + cell * 16 + 4 * g) =
              pack_bf16x4(m0 > 0.0f ? v[0] : 0.0f, m1 > 0.0f ? v[1] : 0.0f, m2 > 0.0f ? v[2] : 0.0f,
                          m3 > 0.0f ? v[3] : 0.0f);
        }
      }
    }
  }
  if (MODE == kFwdAdd && o.sum && g == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float a = added[i];
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) a += __shfl_xor(a, off);
      if (r == 0) atomicAdd(o.sum + i, (double)a);
    }
  }
}

// ---- 16-channel convolution by input planes (inference) -----------------
// k_conv16 reads the three input planes x - 1, x, x + 1 for every output
// plane x, six LDS operands per input row for 14 products.  k_conv16r
// consumes each input plane once: a wave keeps the accumulators of three
// output planes (p + 1, p, p - 1) and input plane p adds its taps of all
// three x offsets to them, two LDS operands per input row (the row and its
// z-shifted copy) for 15 products.  Output plane p - 1 is then complete and
// stored; its accumulators are reused for plane p + 2.  K order, five chunks
// per input-plane offset dx (15 in all, one more than k_conv16's 14):
//   chunk 5 (dx+1) + (dy+1), dy in {-1,0,1}: taps (dx,dy,-1) | (dx,dy,+1)
//   chunk 5 (dx+1) + 3:                     taps (dx,-1,0) | (dx,0,0)
//   chunk 5 (dx+1) + 4:                     taps (dx,+1,0) | none
// With every input plane read once from LDS, the ring holds the plane being
// read and three planes in flight (loads issued three planes ahead), and a
// wave waits only for the next plane's loads (vmcnt), not for its stores.
constexpr int kRollChunks = 15;

__host__ __device__ inline int roll_tap(int c, int h)
{
  const int dx = c / 5, k = c % 5;
  if (k < 3) return 9 * dx + 3 * k + (h ? 2 : 0);  // (dx, dy, -1) | (dx, dy, +1)
  if (k == 3) return 9 * dx + (h ? 4 : 1);         // (dx, -1, 0) | (dx, 0, 0)
  return h ? -1 : 9 * dx + 7;                      // (dx, +1, 0) | none
}

// The rolling kernels' ring: 4 slots of kRollSlotChunks 16-byte chunks
// (the brick's 2 kPlaneCells chunks padded to whole rounds of kThreads), so
// that every lane of every wave issues each of its loads: the compiler then
// knows how many are in flight, and the counts below hold.  The padding
// chunks receive copies of the last cell.  2 x 81,920 bytes fill a CU's
// 160 KiB: two workgroups per CU.
constexpr int kRollLoads = (2 * kPlaneCells + kThreads - 1) / kThreads;  // loads per wave and plane
constexpr int kRollSlotChunks = kRollLoads * kThreads;
constexpr size_t kRollRingBytes = sizeof(uint4) * kRingSlots * kRollSlotChunks;
static_assert(kRollRingBytes * 2 <= 160 * 1024, "two workgroups per CU");

__device__ inline const uint4* roll_slot(const uint4* ring, int p) { return ring + (p & (kRingSlots - 1)) * kRollSlotChunks; }

// Loads of input plane `plane` (periodic) into the ring slot of plane
// `slot_plane` (16 channels: two chunks per cell), with the slot and the
// data plane apart so that the tail of a march can refill a free slot
// (every step then issues the same kRollLoads loads per wave).
__device__ inline void load_plane_to(uint4* ring, const unsigned short* __restrict__ in, int n, int slot_plane,
                                     int plane, int y0, int z0)
{
  const int lane = threadIdx.x & 63;
  uint4* slot = ring + (slot_plane & (kRingSlots - 1)) * kRollSlotChunks;
  const long long xoff = (long long)wrapi(plane, n) * n;
#pragma unroll
  for (int k = 0; k < kRollLoads; ++k) {
    const int base = k * kThreads + (threadIdx.x & ~63);
    const int c = min(base + lane, 2 * kPlaneCells - 1);
    const int q = c & 1, cell = c >> 1;
    const int zz = cell % kBrickZ, yy = cell / kBrickZ;
    const long long src = (xoff + wrapi(y0 + yy - 1, n)) * n + wrapi(z0 + zz - 1, n);
    __builtin_amdgcn_global_load_lds(static_cast<const void*>(in + src * 16 + q * 8),
                                     (__attribute__((address_space(3))) void*)(slot + base), 16, 0, 0);
  }
}

// s_waitcnt immediate: wait until at most n vector-memory operations of the
// wave are outstanding (expcnt and lgkmcnt not waited on).
constexpr int vmcnt_imm(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }
// the counter holds at most 63: a larger allowance waits for more than needed
constexpr int vmcnt_cap(int n) { return n < 63 ? n : 63; }

// Products of one input plane for the wave's kConvTy output rows: taps
// dx = -1 into am (output plane p + 1), dx = 0 into a0 (plane p), dx = +1
// into ap (plane p - 1).  `row` points at the lane's centre cell of brick
// row 0 (its 16-byte half), `shr` at its z neighbour; brick rows are
// 2 kBrickZ uint4 apart.
template <int T>
__device__ inline void roll_plane(const uint4* __restrict__ row, const uint4* __restrict__ shr, bool lo,
                                  const uint4 (&b)[kRollChunks], floatx4 (&am)[T], floatx4 (&a0)[T], floatx4 (&ap)[T])
{
  constexpr int kRow = 2 * kBrickZ;
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  uint4 prev = zero4;
  uint4 cen = row[0], sh = shr[0];
#pragma unroll
  for (int j = 0; j < T + 2; ++j) {  // brick row j feeds output rows j - 2, j - 1, j
    uint4 ncen = zero4, nsh = zero4;
    if (j + 1 < T + 2) {
      ncen = row[(j + 1) * kRow];
      nsh = shr[(j + 1) * kRow];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the next row's reads ahead of this row's products
    if (j < T) {  // dy = -1: output row j
      am[j] = mfma(sh, b[0], am[j]);
      a0[j] = mfma(sh, b[5], a0[j]);
      ap[j] = mfma(sh, b[10], ap[j]);
    }
    if (j >= 1 && j - 1 < T) {  // dy = 0 and (dy,dz) = (-1,0) | (0,0): output row j - 1
      am[j - 1] = mfma(sh, b[1], am[j - 1]);
      a0[j - 1] = mfma(sh, b[6], a0[j - 1]);
      ap[j - 1] = mfma(sh, b[11], ap[j - 1]);
      const uint4 pc = sel4(lo, prev, cen);
      am[j - 1] = mfma(pc, b[3], am[j - 1]);
      a0[j - 1] = mfma(pc, b[8], a0[j - 1]);
      ap[j - 1] = mfma(pc, b[13], ap[j - 1]);
    }
    if (j >= 2) {  // dy = +1 and (dy,dz) = (+1,0) | none: output row j - 2
      am[j - 2] = mfma(sh, b[2], am[j - 2]);
      a0[j - 2] = mfma(sh, b[7], a0[j - 2]);
      ap[j - 2] = mfma(sh, b[12], ap[j - 2]);
      const uint4 qc = sel4(lo, cen, zero4);
      am[j - 2] = mfma(qc, b[4], am[j - 2]);
      a0[j - 2] = mfma(qc, b[9], a0[j - 2]);
      ap[j - 2] = mfma(qc, b[14], ap[j - 2]);
    }
    prev = cen;
    cen = ncen;
    sh = nsh;
  }
}

// Output plane q of a rolling convolution from its accumulators, by mode
// (kFwdRelu: bf16 activations; kFwdLinear: fp32 y; kFwdAdd: added to the
// force grids, old values loaded here), then the accumulators cleared.
template <int MODE>
__device__ inline void roll_store(floatx4 (&acc)[kConvTy], bool store, int q, int n, int ybase, int zcell, int g,
                                  const float (&bch)[4], const ConvOut& o)
{
  if (store) {
#pragma unroll
    for (int t = 0; t < kConvTy; ++t) {
      const long long line = (long long)q * n + ybase + t;
      const long long cell = line * n + zcell;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[t][i] + bch[i];
      if (MODE == kFwdRelu) {
        *reinterpret_cast<uint2*>(o.bf + cell * 16 + 4 * g) =
            pack_bf16x4(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f), fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f));
      } else if (MODE == kFwdLinear) {
        if (g == 0) {
          const long long nc = (long long)n * n * n;
#pragma unroll
          for (int i = 0; i < 3; ++i) o.f[i * nc + cell] = v[i];
        }
      } else {
        if (g == 0) {
          const long long at = line * o.row + zcell;
#pragma unroll
          for (int i = 0; i < 3; ++i) o.add[i][at] = o.add[i][at] + o.scale * v[i];
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kConvTy; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
}

// Vector-memory instructions one step of a rolling march issues per wave
// besides its plane loads (the stores of one output plane; kFwdAdd also
// loads the old force values).
template <int MODE>
constexpr int roll_store_ops()
{
  return MODE == kFwdRelu ? kConvTy : (MODE == kFwdLinear ? 3 * kConvTy : 6 * kConvTy);
}

// Step i of a march (input plane p = x0 - 1 + i) with accumulator sets
// rotated by R = i mod 3: set (R + 1) % 3 holds output plane p + 1, set R
// plane p, set (R + 2) % 3 plane p - 1.
template <int MODE, int R>
__device__ inline void roll_step(uint4* ring, const unsigned short* __restrict__ in, int n, const ConvUnit& cu, int i,
                                 int rowoff, int shoff, bool lo, const uint4 (&b)[kRollChunks],
                                 floatx4 (&acc)[3][kConvTy], int ybase, int zcell, int g, const float (&bch)[4],
                                 const ConvOut& o)
{
  const int p = cu.x0 - 1 + i;
  // plane p + 3 into the slot plane p - 1 left (the march's last plane again
  // once past its end, so every step issues the same loads)
  load_plane_to(ring, in, n, p + 3, min(p + 3, cu.x1), cu.y0, cu.z0);
  const uint4* slot = roll_slot(ring, p);
  roll_plane(slot + rowoff, slot + shoff, lo, b, acc[(R + 1) % 3], acc[R], acc[(R + 2) % 3]);
  roll_store<MODE>(acc[(R + 2) % 3], p - 1 >= cu.x0 && p - 1 < cu.x1, p - 1, n, ybase, zcell, g, bch, o);
  // plane p + 1 landed: the operations issued after its loads may still be
  // in flight (the later planes' loads and up to three steps' stores; the
  // count saturates at the counter's 63, waiting longer)
  constexpr int L = kRollLoads, S = roll_store_ops<MODE>();
  constexpr int w0 = vmcnt_imm(2 * L), w1 = vmcnt_imm(vmcnt_cap(2 * L + S));
  constexpr int w2 = vmcnt_imm(vmcnt_cap(2 * L + 2 * S)), w3 = vmcnt_imm(vmcnt_cap(2 * L + 3 * S));
  const int nst = min(max(i - 1, 0), 3);
  if (nst == 3) {
    __builtin_amdgcn_s_waitcnt(w3);
  } else if (nst == 2) {
    __builtin_amdgcn_s_waitcnt(w2);
  } else if (nst == 1) {
    __builtin_amdgcn_s_waitcnt(w1);
  } else {
    __builtin_amdgcn_s_waitcnt(w0);
  }
  // every wave's plane p + 1 is in LDS and every wave is past its reads of
  // plane p, whose slot the next step's loads take (no fence: the waits
  // above are the ones needed)
  asm volatile("s_barrier" ::: "memory");
}

template <int MODE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_conv16r(const unsigned short* __restrict__ in, const unsigned short* __restrict__ wpk, const float* __restrict__ bias,
          int cout, int n, int march, ConvOut o)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const bool lo = g < 2;
  uint4 b[kRollChunks];
#pragma unroll
  for (int c = 0; c < kRollChunks; ++c) b[c] = *reinterpret_cast<const uint4*>(wpk + ((long long)c * 64 + lane) * 8);
  float bch[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bch[i] = 4 * g + i >= cout ? 0.0f : bias[4 * g + i];
  extern __shared__ uint4 ring[];  // [slot][yy][zz][2 halves of 8 channels]
  const int half = g & 1;
  const int wz = (threadIdx.x >> 6) & 1, wy = threadIdx.x >> 7;
  const int zc = 1 + 16 * wz + r;
  const int zs = lo ? zc - 1 : zc + 1;
  const int rowoff = (kConvTy * wy * kBrickZ + zc) * 2 + half;
  const int shoff = (kConvTy * wy * kBrickZ + zs) * 2 + half;
  floatx4 acc[3][kConvTy];
  const long long units = (long long)((n + march - 1) / march) * (n / kBlockY) * (n / kBlockZ);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const ConvUnit cu = conv_unit(u, n, march);
    const int ybase = cu.y0 + kConvTy * wy, zcell = cu.z0 + 16 * wz + r;
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) acc[s][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // the previous unit's refills have landed and its operand reads are done
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // planes x0 - 1, x0, x0 + 1 (x0 + 1 <= x1: a march has at least one plane)
    for (int k = 0; k < 3; ++k) load_plane_to(ring, in, n, cu.x0 - 1 + k, cu.x0 - 1 + k, cu.y0, cu.z0);
    constexpr int w_pro = vmcnt_imm(2 * kRollLoads);
    __builtin_amdgcn_s_waitcnt(w_pro);
    asm volatile("s_barrier" ::: "memory");
    // input planes x0 - 1 .. x1, rounded up to whole rotations (the extra
    // steps reload the last plane and store nothing)
    const int np = cu.x1 - cu.x0 + 2;
    for (int i = 0; i < np; i += 3) {
      roll_step<MODE, 0>(ring, in, n, cu, i, rowoff, shoff, lo, b, acc, ybase, zcell, g, bch, o);
      roll_step<MODE, 1>(ring, in, n, cu, i + 1, rowoff, shoff, lo, b, acc, ybase, zcell, g, bch, o);
      roll_step<MODE, 2>(ring, in, n, cu, i + 2, rowoff, shoff, lo, b, acc, ybase, zcell, g, bch, o);
    }
  }
  __builtin_amdgcn_s_waitcnt(0);  // the tail's refills land before the workgroup's LDS is released
}

// ---- last layer (16 -> 3) by input planes --------------------------------
// With three output channels the three x offsets become output rows of one
// product: row 4 co + s of the A operand holds W[tap (dx, dy, dz)][ci][co]
// for the x offset dx whose output plane sums in accumulator element s
// (rows with co = 3 or s = 3 zero), so one product per (dy, dz) chunk of an
// input plane (the five chunks of roll_tap for one dx) makes all of that
// plane's contributions.  The weights come in three rotations (step mod 3)
// that keep output plane o in element (o - x0) mod 3 of every lane: output
// plane q = G0(D_{q-1}) + G1(D_q) + G2(D_{q+1}) sums in place over three
// steps, and lane group g holds channel g, one 4-byte access per lane and
// row when the plane is complete (kFwdAdd loads the old force values one
// step ahead).  Five products per 16-cell tile and input plane instead of
// fifteen.
constexpr int kHeadChunks = 5;

// Products of one input plane added into d (one tile per output row: the
// running sums of three output planes, one per accumulator element).  ROW:
// uint4 between input rows (two per cell).
template <int ROW = 2 * kBrickZ, int T = kConvTy>
__device__ inline void head_plane_acc(const uint4* __restrict__ row, const uint4* __restrict__ shr, bool lo,
                                      const uint4 (&b)[kHeadChunks], floatx4 (&d)[T])
{
  constexpr int kRow = ROW;
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  uint4 prev = zero4;
  uint4 cen = row[0], sh = shr[0];
#pragma unroll
  for (int j = 0; j < T + 2; ++j) {
    uint4 ncen = zero4, nsh = zero4;
    if (j + 1 < T + 2) {
      ncen = row[(j + 1) * kRow];
      nsh = shr[(j + 1) * kRow];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (j < T) d[j] = mfma(sh, b[0], d[j]);
    if (j >= 1 && j - 1 < T) {
      d[j - 1] = mfma(sh, b[1], d[j - 1]);
      d[j - 1] = mfma(sel4(lo, prev, cen), b[3], d[j - 1]);
    }
    if (j >= 2) {
      d[j - 2] = mfma(sh, b[2], d[j - 2]);
      d[j - 2] = mfma(sel4(lo, cen, zero4), b[4], d[j - 2]);
    }
    prev = cen;
    cen = ncen;
    sh = nsh;
  }
}

// The last layer's step with rotated weights (R = step mod 3): the products
// of an input plane accumulate in place, the contributions to output plane
// o in accumulator element (o - x0) mod 3 of every lane, lane group g
// holding output channel g, so no partial sums move between lanes and a
// completed plane is already spread one channel per lane group: its
// element is read out and cleared.  Selects the rotation's pack at run
// time (wave-uniform).
template <int ROW = 2 * kBrickZ, int T = kConvTy>
__device__ inline void head_plane_rot(const uint4* __restrict__ row, const uint4* __restrict__ shr, bool lo,
                                      const uint4 (&b)[3][kHeadChunks], int R, floatx4 (&d)[T])
{
  if (R == 0)
    head_plane_acc<ROW, T>(row, shr, lo, b[0], d);
  else if (R == 1)
    head_plane_acc<ROW, T>(row, shr, lo, b[1], d);
  else
    head_plane_acc<ROW, T>(row, shr, lo, b[2], d);
}


// s_waitcnt for "at most n vector-memory operations outstanding", n rounded
// down to the ladder 2L + 8k (L = kRollLoads) that the head's steps produce.
__device__ inline void head_wait(int n)
{
  constexpr int L2 = 2 * kRollLoads;
  if (n >= L2 + 48) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(vmcnt_cap(L2 + 48)));
  } else if (n >= L2 + 40) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(L2 + 40));
  } else if (n >= L2 + 32) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(L2 + 32));
  } else if (n >= L2 + 24) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(L2 + 24));
  } else if (n >= L2 + 16) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(L2 + 16));
  } else if (n >= L2 + 8) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(L2 + 8));
  } else if (n >= L2) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(L2));
  } else {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
  }
}

template <int MODE>  // kFwdAdd or kFwdLinear
__global__ void __launch_bounds__(kThreads)
k_conv_head_r(const unsigned short* __restrict__ in, const unsigned short* __restrict__ wh,
              const float* __restrict__ bias, int n, int march, ConvOut o)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const bool lo = g < 2;
  uint4 b[3][kHeadChunks];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < kHeadChunks; ++c)
      b[q][c] = *reinterpret_cast<const uint4*>(wh + ((long long)(q * kHeadChunks + c) * 64 + lane) * 8);
  const float bo = g < 3 ? bias[g] : 0.0f;
  extern __shared__ uint4 ring[];
  const int half = g & 1;
  const int wz = (threadIdx.x >> 6) & 1, wy = threadIdx.x >> 7;
  const int zc = 1 + 16 * wz + r;
  const int zs = lo ? zc - 1 : zc + 1;
  const int rowoff = (kConvTy * wy * kBrickZ + zc) * 2 + half;
  const int shoff = (kConvTy * wy * kBrickZ + zs) * 2 + half;
  const long long nc = (long long)n * n * n;
  // this lane's output channel: its force grid (kFwdAdd) or y row (kFwdLinear),
  // selected once (indexing the argument struct by g would load it per use)
  float* const gout = MODE == kFwdAdd ? (g == 0 ? o.add[0] : (g == 1 ? o.add[1] : o.add[2]))
                                      : o.f + (g < 3 ? g : 0) * nc;
  constexpr int L = kRollLoads;
  constexpr int F = MODE == kFwdAdd ? kConvTy : 0, S = kConvTy;
  float added = 0.0f;  // kFwdAdd: this lane's sum of the residual it added (component g)
  const long long units = (long long)((n + march - 1) / march) * (n / kBlockY) * (n / kBlockZ);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const ConvUnit cu = conv_unit(u, n, march);
    const int ybase = cu.y0 + kConvTy * wy, zcell = cu.z0 + 16 * wz + r;
    floatx4 acc[kConvTy];  // output plane o in element (o - x0) mod 3, channel g in lane group g
#pragma unroll
    for (int t = 0; t < kConvTy; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int k = 0; k < 3; ++k) load_plane_to(ring, in, n, cu.x0 - 1 + k, cu.x0 - 1 + k, cu.y0, cu.z0);
    constexpr int w_pro = vmcnt_imm(2 * kRollLoads);
    __builtin_amdgcn_s_waitcnt(w_pro);
    asm volatile("s_barrier" ::: "memory");
    const int np = cu.x1 - cu.x0 + 2;  // input planes x0 - 1 .. x1
    for (int i = 0; i < np; ++i) {
      const int p = cu.x0 - 1 + i;
      const bool store = p - 1 >= cu.x0 && p - 1 < cu.x1;  // output plane p - 1 completes at this step
      // kFwdAdd: the old force values of plane p - 1, issued before the plane
      // loads so that their wait leaves those in flight
      float fold[kConvTy];
      if (MODE == kFwdAdd && store) {  // every lane (group 3 reads group 2's grid): no divergent branch
#pragma unroll
        for (int t = 0; t < kConvTy; ++t) fold[t] = gout[((long long)(p - 1) * n + ybase + t) * o.row + zcell];
      }
      load_plane_to(ring, in, n, p + 3, min(p + 3, cu.x1), cu.y0, cu.z0);
      const uint4* slot = roll_slot(ring, p);
      // input plane p adds to output planes p + 1, p, p - 1; p - 1 is complete
      // in accumulator element (i + 1) mod 3
      const int R = i % 3, G = (R + 1) % 3;
      head_plane_rot(slot + rowoff, slot + shoff, lo, b, R, acc);
      if (store) {
#pragma unroll
        for (int t = 0; t < kConvTy; ++t) {
          const float v = (G == 0 ? acc[t][0] : (G == 1 ? acc[t][1] : acc[t][2])) + bo;
          if (g < 3) {
            const long long line = (long long)(p - 1) * n + ybase + t;
            if (MODE == kFwdAdd) {
              const float dv = o.scale * v;
              if (o.fi)
                o.fi[(line * n + zcell) * 3 + g] = fold[t] + dv;  // lanes g = 0-2 of a cell: 12 contiguous bytes
              else
                gout[line * o.row + zcell] = fold[t] + dv;
              added += dv;
            } else {
              gout[line * n + zcell] = v;
            }
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {  // element G of every lane (G is uniform: no dynamic indexing)
        acc[t][0] = G == 0 ? 0.0f : acc[t][0];
        acc[t][1] = G == 1 ? 0.0f : acc[t][1];
        acc[t][2] = G == 2 ? 0.0f : acc[t][2];
      }
      // plane p + 1 landed: step j issues the old-value loads F(j), the plane
      // loads and the stores S(j); count what was issued after the loads of
      // plane p + 1 (the prologue's for i < 2, step i - 2's after)
      auto ocount = [&](int j, int per) { return j >= 2 && j <= cu.x1 - cu.x0 + 1 ? per : 0; };
      int after = 2 * L;
      if (i >= 2)
        after += ocount(i - 2, S) + ocount(i - 1, F) + ocount(i - 1, S) + ocount(i, F) + ocount(i, S);
      head_wait(after);
      asm volatile("s_barrier" ::: "memory");
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  if (MODE == kFwdAdd && o.sum) {
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) added += __shfl_xor(added, off);  // within the 16-lane group
    if (r == 0 && g < 3) atomicAdd(o.sum + g, (double)added);
  }
}

// ---- layer 3 and the last layer in one pass -------------------------------
// k_conv16r<kFwdRelu> (h2 -> h3) and k_conv_head_r<kFwdAdd> (h3 -> corrected
// force) fused, so that h3 never goes to memory (32 bytes per cell written
// and read back).  A unit owns the output cells y0 .. y0 + 15, z0 .. z0 + 29
// (z blocks of 30, the last one ragged: its cells wrap around and are not
// stored) of planes x0 .. x1 - 1.  Layer 3 makes h3 on the region the last
// layer reads, rows y0 - 1 .. y0 + 16 and cells z0 - 1 .. z0 + 30 (18 x 32,
// 1.2 times the cells it owns), from h2 bricks of 20 x 34 cells in a 4-slot
// ring loaded three planes ahead, as in k_conv16r.  Eight waves in two
// roles, one of each per SIMD: waves 0-3 run layer 3 (wave (wy, wz): rows
// 9 wy .. 9 wy + 8 of the 16-cell column wz), waves 4-7 the last layer one
// plane behind (output rows 8 wy .. 8 wy + 7 of output column wz).  At step
// i (h2 plane p = x0 - 2 + i) layer 3 rolls plane p into its three
// accumulator sets and writes the completed h3 plane p - 1 (bias, relu,
// bf16) into LDS slot (p - 1) & 1, while the last layer consumes h3 plane
// p - 2 (written before the previous step's barrier) as k_conv_head_r
// consumes an input plane and stores output plane p - 3; one barrier per
// step.  The last layer's shuffles, loads and stores run beside layer 3's
// products on the same SIMD.  Per output cell the products and sums run in
// the order of the two separate launches: the force is theirs bit for bit.
// The first steps' h3 planes (before x0 - 1, and slot contents of a
// previous unit) only reach output planes before x0, which are not stored.
constexpr int kThreads3 = 512;
constexpr int kT3 = 9;                                  // layer-3 rows per wave
constexpr int kH3Y = 2 * kT3, kH3Z = 32;                // h3 region of a unit
constexpr int kOut3Z = kH3Z - 2;                        // output cells along z
constexpr int kBrick3Y = kH3Y + 2;                      // h2 brick rows
constexpr int kPlane3Cells = kBrick3Y * kBrickZ;        // 680
constexpr int kRoll3Loads = (2 * kPlane3Cells + kThreads3 - 1) / kThreads3;  // 3 loads per wave and plane
constexpr int kRoll3SlotChunks = kRoll3Loads * kThreads3;
constexpr int kH3SlotChunks = 2 * kH3Y * kH3Z;
constexpr size_t kConv3hBytes = sizeof(uint4) * (kRingSlots * kRoll3SlotChunks + 2 * kH3SlotChunks);
static_assert(kConv3hBytes <= 160 * 1024, "one workgroup per CU");

// h2 plane `plane` (periodic) into the ring slot of plane `slot_plane`: brick
// rows y0 - 2 .. y0 + 17, cells z0 - 2 .. z0 + 31 (padding chunks repeat the
// last cell, so every lane issues every load).
__device__ inline void load_plane3_to(uint4* ring, const unsigned short* __restrict__ in, int n, int slot_plane,
                                      int plane, int y0, int z0)
{
  const int lane = threadIdx.x & 63;
  uint4* slot = ring + (slot_plane & (kRingSlots - 1)) * kRoll3SlotChunks;
  const long long xoff = (long long)wrapi(plane, n) * n;
#pragma unroll
  for (int k = 0; k < kRoll3Loads; ++k) {
    const int base = k * kThreads3 + (threadIdx.x & ~63);
    const int c = min(base + lane, 2 * kPlane3Cells - 1);
    const int q = c & 1, cell = c >> 1;
    const int zz = cell % kBrickZ, yy = cell / kBrickZ;
    const long long src = (xoff + wrapi(y0 + yy - 2, n)) * n + wrapi(z0 + zz - 2, n);
    __builtin_amdgcn_global_load_lds(static_cast<const void*>(in + src * 16 + q * 8),
                                     (__attribute__((address_space(3))) void*)(slot + base), 16, 0, 0);
  }
}

// s_waitcnt: at most n vector-memory operations outstanding and every LDS
// access done.
constexpr int vmcnt_lgkm0_imm(int n) { return (n & 15) | (7 << 4) | (0 << 8) | ((n >> 4) << 14); }

__global__ void __launch_bounds__(kThreads3) __attribute__((amdgpu_waves_per_eu(2)))
k_conv3h(const unsigned short* __restrict__ in, const unsigned short* __restrict__ wr3, const float* __restrict__ b3,
         const unsigned short* __restrict__ wh, const float* __restrict__ b4, int n, int march, ConvOut o)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const bool lo = g < 2;
  const int wave = threadIdx.x >> 6;
  const bool head = wave >= 4;  // role (wave-uniform)
  const int wz = wave & 1, wy = (wave >> 1) & 1;
  const int half = g & 1;
  extern __shared__ uint4 smem[];
  uint4* ring = smem;                                   // [slot][20 rows][34 cells][2 halves]
  uint4* h3 = smem + kRingSlots * kRoll3SlotChunks;     // [2 slots][18 rows][32 cells][2 halves]
  const int nzb = (n + kOut3Z - 1) / kOut3Z, nyb = n / kBlockY;
  const long long units = (long long)((n + march - 1) / march) * nyb * nzb;
  constexpr int L = kRoll3Loads;
  if (!head) {
    // ---- layer 3: rows 9 wy .. 9 wy + 8 of the h3 region, cells 16 wz + r
    uint4 b[kRollChunks];
#pragma unroll
    for (int c = 0; c < kRollChunks; ++c) b[c] = *reinterpret_cast<const uint4*>(wr3 + ((long long)c * 64 + lane) * 8);
    float bch[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) bch[i] = b3[4 * g + i];
    const int zc = 1 + 16 * wz + r;
    const int zs = lo ? zc - 1 : zc + 1;
    const int rowoff = (kT3 * wy * kBrickZ + zc) * 2 + half;
    const int shoff = (kT3 * wy * kBrickZ + zs) * 2 + half;
    floatx4 acc[3][kT3];
    for (long long u = blockIdx.x; u < units; u += gridDim.x) {
      const int z0 = (int)(u % nzb) * kOut3Z;
      const long long t2 = u / nzb;
      const int y0 = (int)(t2 % nyb) * kBlockY;
      const int x0 = (int)(t2 / nyb) * march, x1 = min(x0 + march, n);
#pragma unroll
      for (int s3 = 0; s3 < 3; ++s3)
#pragma unroll
        for (int t = 0; t < kT3; ++t) acc[s3][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();  // the previous unit is done
      for (int k = 0; k < 3; ++k) load_plane3_to(ring, in, n, x0 - 2 + k, x0 - 2 + k, y0, z0);
      constexpr int w_pro = vmcnt_imm(2 * L);
      __builtin_amdgcn_s_waitcnt(w_pro);
      asm volatile("s_barrier" ::: "memory");
      auto step = [&](auto rot, int i) {
        constexpr int R = decltype(rot)::value;
        const int p = x0 - 2 + i;
        // h2 plane p + 3 into the slot plane p - 1 left (the last plane again past x1 + 1)
        load_plane3_to(ring, in, n, p + 3, min(p + 3, x1 + 1), y0, z0);
        const uint4* slot = ring + (p & (kRingSlots - 1)) * kRoll3SlotChunks;
        roll_plane<kT3>(slot + rowoff, slot + shoff, lo, b, acc[(R + 1) % 3], acc[R], acc[(R + 2) % 3]);
        // h3 plane p - 1 complete: bias, relu, bf16 into its LDS slot
        uint2* dst = reinterpret_cast<uint2*>(h3 + ((p - 1) & 1) * kH3SlotChunks);
        floatx4(&a)[kT3] = acc[(R + 2) % 3];
#pragma unroll
        for (int t = 0; t < kT3; ++t) {
          const int cell = (kT3 * wy + t) * kH3Z + 16 * wz + r;
          dst[cell * 4 + g] = pack_bf16x4(fmaxf(a[t][0] + bch[0], 0.0f), fmaxf(a[t][1] + bch[1], 0.0f),
                                          fmaxf(a[t][2] + bch[2], 0.0f), fmaxf(a[t][3] + bch[3], 0.0f));
          a[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        // h2 plane p + 1 landed (its loads and this wave's later ones: two
        // steps' L) and the h3 writes are done
        __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0_imm(2 * L));
        asm volatile("s_barrier" ::: "memory");
      };
      const int np = x1 - x0 + 5;  // h2 planes x0 - 2 .. x1 + 2 (the last a repeat), in whole rotations
      for (int i = 0; i < np; i += 3) {
        step(std::integral_constant<int, 0>{}, i);
        step(std::integral_constant<int, 1>{}, i + 1);
        step(std::integral_constant<int, 2>{}, i + 2);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    return;
  }
  // ---- the last layer: output rows 8 wy .. 8 wy + 7, cells z0 + 14 wz + r
  // (the second column's first two cells are the first column's last two)
  uint4 bh[3][kHeadChunks];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < kHeadChunks; ++c)
      bh[q][c] = *reinterpret_cast<const uint4*>(wh + ((long long)(q * kHeadChunks + c) * 64 + lane) * 8);
  const float bo = g < 3 ? b4[g] : 0.0f;
  const int hz = 1 + 14 * wz + r;
  const int hzs = lo ? hz - 1 : hz + 1;
  const int hrowoff = (kConvTy * wy * kH3Z + hz) * 2 + half;
  const int hshoff = (kConvTy * wy * kH3Z + hzs) * 2 + half;
  float* const gout = g == 0 ? o.add[0] : (g == 1 ? o.add[1] : o.add[2]);
  constexpr int F = kConvTy, S = kConvTy;
  static_assert(2 * kRoll3Loads + 3 * S + 2 * F < 64, "the wait's count fits the counter");
  float added = 0.0f;
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const int z0 = (int)(u % nzb) * kOut3Z;
    const long long t2 = u / nzb;
    const int y0 = (int)(t2 % nyb) * kBlockY;
    const int x0 = (int)(t2 / nyb) * march, x1 = min(x0 + march, n);
    const int ybase = y0 + kConvTy * wy;
    const int zcell = z0 + 14 * wz + r;
    const bool zok = zcell < n && (wz == 0 || r >= 2);  // a cell of the grid this lane alone stores
    const int zld = zcell < n ? zcell : zcell - n;       // in range for the old-value loads
    floatx4 acc[kConvTy];  // output plane o in element (o - x0) mod 3, channel g in lane group g
#pragma unroll
    for (int t = 0; t < kConvTy; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // the previous unit is done
    for (int k = 0; k < 3; ++k) load_plane3_to(ring, in, n, x0 - 2 + k, x0 - 2 + k, y0, z0);
    constexpr int w_pro = vmcnt_imm(2 * L);
    __builtin_amdgcn_s_waitcnt(w_pro);
    asm volatile("s_barrier" ::: "memory");
    // output plane p - 3 is stored at steps i = 5 .. x1 - x0 + 4.  Every
    // step issues the same vector-memory operations, so the waits are
    // fixed: F old-value loads for the next step (its plane clamped into
    // the march), L plane loads, S stores (into trash at steps that store
    // nothing); the old values arrive a whole step before their use.
    auto stores = [&](int j) { return j >= 5 && j < x1 - x0 + 5; };
    auto old_plane = [&](int j) { return min(max(x0 - 5 + j, x0), x1 - 1); };  // step j's output plane
    const int np = x1 - x0 + 5;
    float fold[kConvTy], fnext[kConvTy];
#pragma unroll
    for (int t = 0; t < kConvTy; ++t) fold[t] = 0.0f;
    for (int i = 0; i < np; ++i) {  // the same number of steps (and barriers) as layer 3's rotations
      const int p = x0 - 2 + i;
      const bool store = stores(i);
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) fnext[t] = gout[((long long)old_plane(i + 1) * n + ybase + t) * o.row + zld];
      load_plane3_to(ring, in, n, p + 3, min(p + 3, x1 + 1), y0, z0);
      // h3 plane q = p - 2 adds to output planes q + 1, q, q - 1; q - 1 is
      // complete in accumulator element (i + 1) mod 3
      const uint4* hs = h3 + ((p - 2) & 1) * kH3SlotChunks;
      const int R = i % 3, G = (R + 1) % 3;
      head_plane_rot<2 * kH3Z>(hs + hrowoff, hs + hshoff, lo, bh, R, acc);
      const bool put = store && zok;
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {
        const float v = (G == 0 ? acc[t][0] : (G == 1 ? acc[t][1] : acc[t][2])) + bo;
        if (g < 3) {
          const long long line = (long long)(put ? p - 3 : x0) * n + ybase + t;
          const float dv = o.scale * v;
          float* dst2 = put ? o.fi + (line * n + zcell) * 3 + g : o.trash + lane;
          *dst2 = fold[t] + dv;
          if (put) added += dv;
        }
      }
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {  // element G of every lane (G is uniform: no dynamic indexing)
        acc[t][0] = G == 0 ? 0.0f : acc[t][0];
        acc[t][1] = G == 1 ? 0.0f : acc[t][1];
        acc[t][2] = G == 2 ? 0.0f : acc[t][2];
      }
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) fold[t] = fnext[t];
      // h2 plane p + 1 landed: after its loads (step i - 2's L) came that
      // step's S and two more steps' F, L, S; for i < 2 its loads were the
      // prologue's second and third, followed by L (i = 0) and the steps'
      if (i >= 2) {
        __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0_imm(2 * L + 3 * S + 2 * F));
      } else if (i == 1) {
        __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0_imm(2 * L + 2 * S + 2 * F));
      } else {
        __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0_imm(2 * L + S + F));
      }
      asm volatile("s_barrier" ::: "memory");
    }
    // layer 3's steps come in whole rotations: the remaining barriers
    for (int i = np; i < (np + 2) / 3 * 3; ++i) {
      const int p = x0 - 2 + i;
      load_plane3_to(ring, in, n, p + 3, min(p + 3, x1 + 1), y0, z0);
      __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0_imm(0));
      asm volatile("s_barrier" ::: "memory");
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  if (o.sum) {
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) added += __shfl_xor(added, off);  // within the 16-lane group
    if (r == 0 && g < 3) atomicAdd(o.sum + g, (double)added);
  }
}

// 8-channel convolution (the input features) with the operand reuse of
// k_conv16: one 16-byte load holds all 8 channels of a cell, so a K chunk of
// 32 is one input row's three z taps, lane group g loading cell z + g - 1
// (group 3 idle):  chunk 3 (dx+1) + (dy+1): taps (dx,dy,-1) | (dx,dy,0) |
// (dx,dy,+1) | none.
__global__ void __launch_bounds__(kThreads)
k_conv8(const unsigned short* __restrict__ in, const unsigned short* __restrict__ wpk, const float* __restrict__ bias,
        int n, int march, unsigned short* __restrict__ out_bf)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  uint4 b[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) b[c] = *reinterpret_cast<const uint4*>(wpk + ((long long)c * 64 + lane) * 8);
  float bch[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bch[i] = bias[4 * g + i];
  extern __shared__ uint4 ring[];  // [slot][yy][zz], 8 channels per cell
  const int wz = (threadIdx.x >> 6) & 1, wy = threadIdx.x >> 7;
  const int zs = 1 + 16 * wz + r + (g < 3 ? g - 1 : 0);  // brick z of the lane's tap
  const long long units = (long long)((n + march - 1) / march) * (n / kBlockY) * (n / kBlockZ);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const ConvUnit cu = conv_unit(u, n, march);
    __syncthreads();  // the previous unit's operand reads are done
    for (int p = cu.x0 - 1; p <= cu.x0 + 1; ++p) load_plane<1>(ring, in, n, p, cu.y0, cu.z0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int x = cu.x0; x < cu.x1; ++x) {
      if (x + 2 <= cu.x1) load_plane<1>(ring, in, n, x + 2, cu.y0, cu.z0);  // lands during this plane
      floatx4 acc[kConvTy];
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int j = 0; j < kConvTy + 2; ++j) {  // input row j of the half feeds output rows j - 2, j - 1, j
        const int yy = kConvTy * wy + j;
        uint4 op[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const uint4 c = ring[ring_cell(x - 1 + d, yy, zs)];
          op[d] = g < 3 ? c : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (j >= 2) acc[j - 2] = mfma(op[d], b[3 * d + 2], acc[j - 2]);                 // dy = +1
          if (j >= 1 && j - 1 < kConvTy) acc[j - 1] = mfma(op[d], b[3 * d + 1], acc[j - 1]);  // dy = 0
          if (j < kConvTy) acc[j] = mfma(op[d], b[3 * d], acc[j]);                          // dy = -1
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {
        const long long cell = ((long long)x * n + cu.y0 + kConvTy * wy + t) * n + cu.z0 + 16 * wz + r;
        *reinterpret_cast<uint2*>(out_bf + cell * 16 + 4 * g) =
            pack_bf16x4(fmaxf(acc[t][0] + bch[0], 0.0f), fmaxf(acc[t][1] + bch[1], 0.0f),
                        fmaxf(acc[t][2] + bch[2], 0.0f), fmaxf(acc[t][3] + bch[3], 0.0f));
      }
    }
  }
}

// First layer from the feature channel alone.  The first layer's input is 8
// channels: the feature, a (the same in every cell) and six zeros.  A
// constant channel of a periodic convolution adds a * sum_t W[t][1][co] to
// every cell, so it joins the bias, and only the feature's 27 taps remain:
// two K chunks per 16-cell tile instead of nine, 2 bytes per input cell
// instead of 16.  The feature planes sit in LDS as triples (f(z - 1), f(z),
// f(z + 1), -) per cell, so that one 8-byte read gives the three dz taps of
// a (dx, dy) row: lane group g of chunk c holds tap rows 2s and 2s + 1,
// s = 4c + g, of the nine rows R = 3 (dx + 1) + (dy + 1), each as its three
// dz taps and a zero (the fourth element of a triple is masked off); rows
// past 8 have zero weights and re-read row 8 or 4.  a comes from device
// memory (pa) when given.
constexpr int kConv1Tri = 10;  // the first of the two chunks in the layer's forward pack

// tap row of operand slot k (0, 1: chunk 0; 2, 3: chunk 1) in lane group g
__host__ __device__ inline int tri_row(int g, int k) { return k < 2 ? 2 * g + k : (g == 0 ? 8 : 4); }

// feature f of ring cell c (rows of rowz cells, zz = c mod rowz) into its
// own triple and those of its z neighbours in the row
__device__ inline void put_triple(unsigned short* ring16, int c, int zz, int rowz, unsigned short f)
{
  ring16[4 * c + 1] = f;
  if (zz > 0) ring16[4 * (c - 1) + 2] = f;
  if (zz + 1 < rowz) ring16[4 * (c + 1)] = f;
}

// layer-1 products of a tile from the lane's four tap-row triples
__device__ inline floatx4 conv1_tri(const uint2 (&t)[4], const uint4& w0, const uint4& w1)
{
  const uint4 op0 = make_uint4(t[0].x, t[0].y & 0xffffu, t[1].x, t[1].y & 0xffffu);
  const uint4 op1 = make_uint4(t[2].x, t[2].y & 0xffffu, t[3].x, t[3].y & 0xffffu);
  return mfma(op1, w1, mfma(op0, w0, floatx4{0.0f, 0.0f, 0.0f, 0.0f}));
}

// Bias of the first layer's output channels 4g..4g+3 with its constant a
// channel folded in: bias + a * sum_t W[t][1][co], a and W rounded to bf16
// as the 8-channel layer sees them (the products are exact in fp32).
__device__ inline void conv1_bias(const float* __restrict__ w, const float* __restrict__ bias, float a, int g,
                                  float (&bch)[4])
{
  const float abf = bf2f(f2bf(a));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float sa = 0.0f;
    for (int t = 0; t < 27; ++t) sa += abf * bf2f(f2bf(w[(t * 8 + 1) * 16 + 4 * g + i]));
    bch[i] = bias[4 * g + i] + sa;
  }
}

constexpr int kConv1Waves = 3;  // waves per SIMD (= workgroups per CU) requested for k_conv1 (2: 1.5 ms slower at 768^3)

// The march of k_conv16 over a ring of triple planes; the planes stream
// through registers, the loads of plane x + 2 in flight during plane x's
// products.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kConv1Waves)))
k_conv1(const unsigned short* __restrict__ f, const unsigned short* __restrict__ wpk, const float* __restrict__ w,
        const float* __restrict__ bias, const float* __restrict__ pa, float a_host, int n, int march,
        unsigned short* __restrict__ out_bf)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const uint4 w0 = *reinterpret_cast<const uint4*>(wpk + ((long long)kConv1Tri * 64 + lane) * 8);
  const uint4 w1 = *reinterpret_cast<const uint4*>(wpk + ((long long)(kConv1Tri + 1) * 64 + lane) * 8);
  float bch[4];
  conv1_bias(w, bias, pa ? *pa : a_host, g, bch);
  __shared__ uint2 ring[kRingSlots * kPlaneCells];  // [slot][yy][zz] triples
  const int wz = (threadIdx.x >> 6) & 1, wy = threadIdx.x >> 7;
  const int zc = 1 + 16 * wz + r;
  // per operand slot k of this lane: plane offset dx and ring offset of dy
  int tdx[4], toff[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int R = tri_row(g, k);
    tdx[k] = R / 3 - 1;
    toff[k] = (R % 3 - 1) * kBrickZ;
  }
  constexpr int kPer = (kPlaneCells + kThreads - 1) / kThreads;  // cells per thread and plane
  const long long units = (long long)((n + march - 1) / march) * (n / kBlockY) * (n / kBlockZ);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const ConvUnit cu = conv_unit(u, n, march);
    auto fetch = [&](int p, unsigned short (&v)[kPer]) {
      const long long xoff = (long long)wrapi(p, n) * n;
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int c = threadIdx.x + k * kThreads;
        if (c < kPlaneCells) {
          const int zz = c % kBrickZ, yy = c / kBrickZ;
          v[k] = f[(xoff + wrapi(cu.y0 + yy - 1, n)) * n + wrapi(cu.z0 + zz - 1, n)];
        }
      }
    };
    auto put = [&](int p, const unsigned short (&v)[kPer]) {
      unsigned short* slot = reinterpret_cast<unsigned short*>(ring + ring_cell(p, 0, 0));
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int c = threadIdx.x + k * kThreads;
        if (c < kPlaneCells) put_triple(slot, c, c % kBrickZ, kBrickZ, v[k]);
      }
    };
    __syncthreads();  // the previous unit's reads are done
    for (int p = cu.x0 - 1; p <= cu.x0 + 1; ++p) {
      unsigned short v[kPer];
      fetch(p, v);
      put(p, v);
    }
    __syncthreads();
    for (int x = cu.x0; x < cu.x1; ++x) {
      const bool pre = x + 2 <= cu.x1;
      unsigned short nxt[kPer];
      if (pre) fetch(x + 2, nxt);
      int base[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) base[k] = ring_cell(x + tdx[k], kConvTy * wy + 1, zc) + toff[k];
      floatx4 acc[kConvTy];
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {
        uint2 tr[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) tr[k] = ring[base[k] + t * kBrickZ];
        acc[t] = conv1_tri(tr, w0, w1);
      }
      if (pre) put(x + 2, nxt);  // slot of plane x - 2, free since the last barrier
      __syncthreads();
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) {
        const long long cell = ((long long)x * n + cu.y0 + kConvTy * wy + t) * n + cu.z0 + 16 * wz + r;
        *reinterpret_cast<uint2*>(out_bf + cell * 16 + 4 * g) =
            pack_bf16x4(fmaxf(acc[t][0] + bch[0], 0.0f), fmaxf(acc[t][1] + bch[1], 0.0f),
                        fmaxf(acc[t][2] + bch[2], 0.0f), fmaxf(acc[t][3] + bch[3], 0.0f));
      }
    }
  }
}

// The fused first pass's feature planes: the 16 x 32 block with two halo
// cells per side (the first layer's halo on top of the second's).
constexpr int kFeatY = kBlockY + 4, kFeatZ = kBlockZ + 4;
constexpr int kFeatCells = kFeatY * kFeatZ;
constexpr int kFeatSlots = 4;
constexpr int kFeatPer = (kFeatCells + kThreads - 1) / kThreads;

// Features, layer 1 and layer 2 in one pass with layer 2 by input planes
// (k_conv16r's order): h2 from the real-space density grid.  A unit keeps
// feature planes (kFeatY x kFeatZ, 4 ring slots) and two h1 planes (the
// kBrickY x kBrickZ brick, padded to kH1SlotCells) in LDS.  Step i of the
// march (h1 plane p = x0 - 1 + i): the density of feature plane p + 3 is
// loaded into registers; layer 2 adds h1 plane p's taps to the
// accumulators of h2 planes p + 1, p, p - 1 while, row by row, the wave
// makes one 16-cell tile of h1 plane p + 1 (from feature planes p .. p + 2,
// into the other h1 slot; layer 1 as in k_conv1); feature plane p + 3 takes
// the slot of plane p - 1; h2 plane p - 1 is stored.  One barrier per step.
// Every step makes all 40 tiles (the last ones write the pad) and its next
// h1 plane even past the march's end (never read for a stored plane), so
// the loop has no branches inside.  The arithmetic is that of k_features,
// k_conv1 and k_conv16r<kFwdRelu>: h2 equals theirs bit for bit.
constexpr int kH1SlotCells = 4 * (kConvTy + 2) * 16;  // one tile per wave and row of the march step
static_assert(kH1SlotCells >= kPlaneCells, "the h1 slot holds the brick");
constexpr size_t kConv12rBytes = sizeof(uint4) * 2 * 2 * kH1SlotCells + sizeof(uint2) * kFeatSlots * kFeatCells;

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_conv12r(const float* __restrict__ delta, long long row, float in_scale, const unsigned short* __restrict__ wpk1,
          const float* __restrict__ w1, const float* __restrict__ b1, const float* __restrict__ pa, float a_host,
          const unsigned short* __restrict__ wr2, const float* __restrict__ b2, int n, int march,
          unsigned short* __restrict__ out_bf)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const bool lo = g < 2;
  const int wave = threadIdx.x >> 6;
  uint4 b[kRollChunks];
#pragma unroll
  for (int c = 0; c < kRollChunks; ++c) b[c] = *reinterpret_cast<const uint4*>(wr2 + ((long long)c * 64 + lane) * 8);
  const uint4 w10 = *reinterpret_cast<const uint4*>(wpk1 + ((long long)kConv1Tri * 64 + lane) * 8);
  const uint4 w11 = *reinterpret_cast<const uint4*>(wpk1 + ((long long)(kConv1Tri + 1) * 64 + lane) * 8);
  float bch1[4], bch2[4];
  conv1_bias(w1, b1, pa ? *pa : a_host, g, bch1);
#pragma unroll
  for (int i = 0; i < 4; ++i) bch2[i] = b2[4 * g + i];
  extern __shared__ uint4 smem[];
  uint4* h1 = smem;  // [2 slots][kH1SlotCells][2 halves of 8 channels]
  uint2* feat = reinterpret_cast<uint2*>(smem + 2 * 2 * kH1SlotCells);  // [slot][yy][zz] triples
  const int half = g & 1;
  const int wz = wave & 1, wy = wave >> 1;
  const int zc = 1 + 16 * wz + r;
  const int zs = lo ? zc - 1 : zc + 1;
  const int rowoff = (kConvTy * wy * kBrickZ + zc) * 2 + half;
  const int shoff = (kConvTy * wy * kBrickZ + zs) * 2 + half;
  // layer-1 tap rows of this lane's operand slots: plane offset dx and feature-row offset of dy
  int tdx[4], toff[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int R = tri_row(g, k);
    tdx[k] = R / 3 - 1;
    toff[k] = (R % 3 - 1) * kFeatZ;
  }
  floatx4 acc[3][kConvTy];
  const long long units = (long long)((n + march - 1) / march) * (n / kBlockY) * (n / kBlockZ);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const ConvUnit cu = conv_unit(u, n, march);
    const int ybase = cu.y0 + kConvTy * wy, zcell = cu.z0 + 16 * wz + r;
    auto fetch = [&](int q, float (&v)[kFeatPer]) {
      const long long xoff = (long long)wrapi(q, n) * n;
#pragma unroll
      for (int k = 0; k < kFeatPer; ++k) {
        const int c = threadIdx.x + k * kThreads;
        if (c < kFeatCells) {
          const int yy = c / kFeatZ, zz = c - yy * kFeatZ;
          v[k] = delta[(xoff + wrapi(cu.y0 + yy - 2, n)) * row + wrapi(cu.z0 + zz - 2, n)];
        }
      }
    };
    auto put = [&](int q, const float (&v)[kFeatPer]) {
      unsigned short* slot = reinterpret_cast<unsigned short*>(feat + (q & (kFeatSlots - 1)) * kFeatCells);
#pragma unroll
      for (int k = 0; k < kFeatPer; ++k) {
        const int c = threadIdx.x + k * kThreads;
        if (c < kFeatCells) put_triple(slot, c, c % kFeatZ, kFeatZ, feature_of(v[k], in_scale));
      }
    };
    // tile tt (0 .. 39) of h1 plane p from feature planes p - 1 .. p + 1
    // into h1 slot p & 1 (tiles past the brick write the slot's pad)
    auto l1tile = [&](int p, int tt) {
      const int cell = tt * 16 + r;
      const int c = min(cell, kPlaneCells - 1);
      const int yy = c / kBrickZ, zz = c - yy * kBrickZ;
      const int centre = (yy + 1) * kFeatZ + zz + 1;
      uint2 tr[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) tr[k] = feat[((p + tdx[k]) & (kFeatSlots - 1)) * kFeatCells + toff[k] + centre];
      const floatx4 a1 = conv1_tri(tr, w10, w11);
      uint2* dst = reinterpret_cast<uint2*>(h1 + 2 * (p & 1) * kH1SlotCells);
      dst[cell * 4 + g] = pack_bf16x4(fmaxf(a1[0] + bch1[0], 0.0f), fmaxf(a1[1] + bch1[1], 0.0f),
                                      fmaxf(a1[2] + bch1[2], 0.0f), fmaxf(a1[3] + bch1[3], 0.0f));
    };
    // one step of the march with the accumulator sets rotated by R = i mod 3
    auto step = [&](auto rot, int i) {
      constexpr int R = decltype(rot)::value;
      const int p = cu.x0 - 1 + i;
      const bool more = p + 3 <= cu.x1 + 1;  // feature planes x0 - 2 .. x1 + 1
      float v[kFeatPer];
      if (more) fetch(p + 3, v);
      const uint4* slot = h1 + 2 * (p & 1) * kH1SlotCells;
      roll_plane(slot + rowoff, slot + shoff, lo, b, acc[(R + 1) % 3], acc[R], acc[(R + 2) % 3]);
#pragma unroll 1
      for (int j = 0; j < kConvTy + 2; ++j) l1tile(p + 1, wave + 4 * j);
      // the features before the stores: their wait for this step's loads
      // then leaves the stores in flight through the next step
      if (more) put(p + 3, v);
      roll_store<kFwdRelu>(acc[(R + 2) % 3], p - 1 >= cu.x0 && p - 1 < cu.x1, p - 1, n, ybase, zcell, g, bch2,
                           ConvOut{out_bf});
      __syncthreads();
    };
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int t = 0; t < kConvTy; ++t) acc[s][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();  // the previous unit's reads are done
    {
      float v[kFeatPer];
      for (int q = cu.x0 - 2; q <= cu.x0 + 1; ++q) {
        fetch(q, v);
        put(q, v);
      }
    }
    __syncthreads();
    for (int tt = wave; tt < 4 * (kConvTy + 2); tt += kThreads / 64) l1tile(cu.x0 - 1, tt);
    __syncthreads();
    const int np = cu.x1 - cu.x0 + 2;  // h1 planes x0 - 1 .. x1, in whole rotations
    for (int i = 0; i < np; i += 3) {
      step(std::integral_constant<int, 0>{}, i);
      step(std::integral_constant<int, 1>{}, i + 1);
      step(std::integral_constant<int, 2>{}, i + 2);
    }
  }
}

// Weight gradient of one layer: dW[t][ci][co] = sum_cells X[cell + d_t][ci]
// G[cell][co] (M = ci, N = co, K = 32 cells along z per MFMA; all 27 taps
// accumulate in registers) and db[co] = sum_cells G[cell][co].  Each wave
// writes its partial sums to part[wave][27 * 256 + 16].
template <int CIN>
__global__ void __launch_bounds__(kThreads)
k_conv_wgrad(const unsigned short* __restrict__ X, const unsigned short* __restrict__ G, int n,
             float* __restrict__ part)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  floatx4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  const long long nblk = (long long)n * n * (n / 32);
  const long long nwaves = (long long)gridDim.x * (kThreads / 64);
  const long long wave = (long long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  for (long long blk = wave; blk < nblk; blk += nwaves) {
    const int zb = (int)(blk % (n / 32));
    const long long xy = blk / (n / 32);
    const int y = (int)(xy % n), x = (int)(xy / n);
    const int z0 = zb * 32 + 8 * g;
    unsigned short bv[8];
    const long long base = ((long long)x * n + y) * n + z0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      bv[j] = G[(base + j) * 16 + r];
      bsum += bf2f(bv[j]);
    }
    bf16x8 bop;
    __builtin_memcpy(&bop, bv, 16);
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      unsigned short av[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) av[j] = r < CIN ? X[cell_at(x, y, z0 + j, t, 1, n) * CIN + r] : (unsigned short)0;
      bf16x8 aop;
      __builtin_memcpy(&aop, av, 16);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aop, bop, acc[t], 0, 0, 0);
    }
  }
  float* p = part + wave * (27 * 256 + 16);
#pragma unroll
  for (int t = 0; t < 27; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) p[(t * 16 + 4 * g + i) * 16 + r] = acc[t][i];
  // bias: lanes r, r + 16, r + 32, r + 48 hold the same channel
  bsum += __shfl_xor(bsum, 16);
  bsum += __shfl_xor(bsum, 32);
  if (g == 0) p[27 * 256 + r] = bsum;
}

// Sum of the per-wave partials (fixed order) into the fp32 gradient of one
// layer, in the master layout W[t][ci][co] then b[co].
__global__ void k_wgrad_reduce(const float* __restrict__ part, int nwaves, int cin, int cout, float* __restrict__ grad)
{
  const int nw = 27 * cin * cout;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nw + cout) return;
  int src;
  if (p < nw) {
    const int t = p / (cin * cout), ci = (p / cout) % cin, co = p % cout;
    src = (t * 16 + ci) * 16 + co;
  } else {
    src = 27 * 256 + (p - nw);
  }
  float s = 0.0f;
  for (int w = 0; w < nwaves; ++w) s += part[(long long)w * (27 * 256 + 16) + src];
  grad[p] = s;
}

// Tap of chunk c for lane-group half h (0: lane groups 0, 1; 1: groups 2, 3)
// in the K order of k_conv16, or -1 for none; taps t = 9 (dx+1) + 3 (dy+1) + (dz+1).
__device__ inline int conv16_tap(int c, int h)
{
  if (c < 9) return 3 * c + (h ? 2 : 0);                  // (dx, dy, -1) | (dx, dy, +1)
  if (c < 12) return 9 * (c - 9) + (h ? 4 : 1);           // (dx, -1, 0) | (dx, 0, 0)
  if (c == 12) return h ? 16 : 7;                         // (-1, +1, 0) | (0, +1, 0)
  return h ? -1 : 25;                                     // (+1, +1, 0) | none
}

// fp32 master weights W[t][ci][co] -> forward B packs and data-gradient B
// packs.  cin = 8: the K order of k_conv8;
// cin = 16: the K order of k_conv16.  The data gradient is the forward
// convolution of the 16-channel output gradient with mirrored taps
// (26 - t) and transposed channels, in the K order of k_conv16.
__global__ void k_pack(const float* __restrict__ w, int cin, int cout, unsigned short* __restrict__ wf,
                       unsigned short* __restrict__ wb)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int c = blockIdx.x;  // chunk
  if (cin == 8) {
    // k_conv8: chunk c = 3 (dx+1) + (dy+1), lane group g -> tap dz = g - 1
    if (c < 9)
      for (int j = 0; j < 8; ++j) {
        const int tap = 3 * c + g;
        const float v = (g < 3 && col < cout) ? w[((long long)tap * 8 + j) * cout + col] : 0.0f;
        wf[((long long)c * 64 + lane) * 8 + j] = f2bf(v);
      }
    // k_conv1 / k_conv12r: the feature channel's taps by tap-row triples in
    // two chunks (element j of lane group g: operand slot j / 4, dz j % 4)
    if (c == kConv1Tri || c == kConv1Tri + 1)
      for (int j = 0; j < 8; ++j) {
        const int k = 2 * (c - kConv1Tri) + (j >> 2), dzi = j & 3;
        const bool used = dzi < 3 && (k < 2 || (g == 0 && k == 2));  // rows 0-7, then row 8 once
        const int tap = 3 * tri_row(g, k) + dzi;
        const float v = (used && col < cout) ? w[((long long)tap * 8) * cout + col] : 0.0f;
        wf[((long long)c * 64 + lane) * 8 + j] = f2bf(v);
      }
  } else {
    const int tap = conv16_tap(c, g >> 1);
    for (int j = 0; j < 8; ++j) {
      const int ci = 8 * (g & 1) + j;
      const float v = (tap >= 0 && col < cout) ? w[((long long)tap * 16 + ci) * cout + col] : 0.0f;
      wf[((long long)c * 64 + lane) * 8 + j] = f2bf(v);
    }
  }
  if (wb) {
    const int tap = conv16_tap(c, g >> 1);
    for (int j = 0; j < 8; ++j) {
      const int co = 8 * (g & 1) + j;
      const float v = (tap >= 0 && co < cout && col < cin) ? w[((long long)(26 - tap) * cin + col) * cout + co] : 0.0f;
      wb[((long long)c * 64 + lane) * 8 + j] = f2bf(v);
    }
  }
}

// fp32 master weights W[t][16][co] -> k_conv16r's forward A operands (the
// 15-chunk K order of roll_tap).
__global__ void k_pack_roll(const float* __restrict__ w, int cout, unsigned short* __restrict__ wr)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int c = blockIdx.x;
  const int tap = roll_tap(c, g >> 1);
  for (int j = 0; j < 8; ++j) {
    const int ci = 8 * (g & 1) + j;
    const float v = (tap >= 0 && col < cout) ? w[((long long)tap * 16 + ci) * cout + col] : 0.0f;
    wr[((long long)c * 64 + lane) * 8 + j] = f2bf(v);
  }
}

// Last-layer weights W[t][16][3] -> the last layer's A operands, three
// rotations R of five chunks (block 5 R + c): row 4 co + s of chunk c
// holds the taps roll_tap(5 dxi + c, h) of output channel co with
// dxi = (R - s) mod 3 (co = 3 and s = 3 zero), so the contribution of x
// offset dxi - 1 lands in accumulator element s = (R - dxi) mod 3 of output
// lane group co.
__global__ void k_pack_head(const float* __restrict__ w, unsigned short* __restrict__ wh)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int rot = blockIdx.x / kHeadChunks, c = blockIdx.x % kHeadChunks;
  const int co = col >> 2, sl = col & 3;
  const int dxi = (rot - sl + 3) % 3;
  const int tap = (co < 3 && sl < 3) ? roll_tap(5 * dxi + c, g >> 1) : -1;
  for (int j = 0; j < 8; ++j) {
    const int ci = 8 * (g & 1) + j;
    const float v = tap >= 0 ? w[((long long)tap * 16 + ci) * 3 + co] : 0.0f;
    wh[((long long)blockIdx.x * 64 + lane) * 8 + j] = f2bf(v);
  }
}

// Feature grid from a density grid: one workgroup per grid line (x, y),
// each thread's loads issued before its transcendental work.
__global__ void __launch_bounds__(kThreads)
k_features(const float* __restrict__ delta, long long row, int n, float scale, unsigned short* __restrict__ f)
{
  constexpr int kPer = 4;
  const long long line = blockIdx.x;
  for (int z0 = 0; z0 < n; z0 += kPer * kThreads) {
    float d[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int z = z0 + threadIdx.x + k * kThreads;
      d[k] = z < n ? delta[line * row + z] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int z = z0 + threadIdx.x + k * kThreads;
      if (z < n) f[line * n + z] = feature_of(d[k], scale);
    }
  }
}

// 8-channel first-layer input of a stored sample (the weight gradient's
// operand): channel 0 its feature (bf16 asinh(delta) * scale), channel 1 its
// a (bf16), rest 0.
__global__ void k_features_stored(const unsigned short* __restrict__ f, int n, float a, unsigned short* __restrict__ x)
{
  const long long cell = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (cell >= (long long)n * n * n) return;
  unsigned short v[8] = {f[cell], f2bf(a), 0, 0, 0, 0, 0, 0};
  *reinterpret_cast<uint4*>(x + cell * 8) = *reinterpret_cast<uint4*>(v);
}

// Loss = mean (y - t / out_scale)^2 over cells and 3 components; gradient
// G[cell][co] = 2 (y - t) / N for co < 3, 0 for the padding channels.
__global__ void __launch_bounds__(kThreads)
k_loss(const float* __restrict__ y, const unsigned short* __restrict__ tb, const float* __restrict__ tf, int n,
       float inv_scale, unsigned short* __restrict__ G, double* __restrict__ loss)
{
  __shared__ double part[kThreads / 64];
  const long long nc = (long long)n * n * n;
  const double inv_n = 1.0 / (3.0 * (double)nc);
  double s = 0.0;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long cell = (long long)blockIdx.x * kThreads + threadIdx.x; cell < nc; cell += stride) {
    unsigned short gv[16] = {0};
    for (int c = 0; c < 3; ++c) {
      const long long i = (long long)c * nc + cell;
      const float t = (tb ? bf2f(tb[i]) : tf[i]) * inv_scale;
      const float d = y[i] - t;
      s += (double)d * d;
      gv[c] = f2bf((float)(2.0 * inv_n * d));
    }
    uint4* o = reinterpret_cast<uint4*>(G + cell * 16);
    o[0] = *reinterpret_cast<uint4*>(gv);
    o[1] = *reinterpret_cast<uint4*>(gv + 8);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += part[w];
    unsafeAtomicAdd(loss, t * inv_n);
  }
}

__global__ void k_zero1(double* p)
{
  if (threadIdx.x == 0) *p = 0.0;
}

__global__ void k_zero3(double* p)
{
  if (threadIdx.x < 3) p[threadIdx.x] = 0.0;
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int np, float lr, float b1, float b2, float c1, float c2, float eps)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Fine force restricted to coarse cell (x, y, z): the average over the 3^3
// fine neighbours of the coincident fine cell (2x, 2y, 2z) with weights
// (ws, 1 - 2 ws, ws) per axis, periodic (ws = 0: the point value, 1/4: full
// weighting).
__device__ inline float restrict_fine(const float* __restrict__ f, long long frow, int nf, int x, int y, int z, float ws)
{
  if (ws == 0.0f) return f[((long long)(2 * x) * nf + 2 * y) * frow + 2 * z];
  const float w[3] = {ws, 1.0f - 2.0f * ws, ws};
  float s = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const int fx = wrapi(2 * x + a - 1, nf);
    for (int b = 0; b < 3; ++b) {
      const long long row = ((long long)fx * nf + wrapi(2 * y + b - 1, nf)) * frow;
      const float wab = w[a] * w[b];
      for (int c = 0; c < 3; ++c) s += wab * w[c] * f[row + wrapi(2 * z + c - 1, nf)];
    }
  }
  return s;
}

// Training pair: feature asinh(delta) * scale (bf16) and target = restricted
// fine force minus the coarse force (bf16).
__global__ void k_store_pair(const float* __restrict__ cd, long long crow, const float* __restrict__ c0,
                             const float* __restrict__ c1, const float* __restrict__ c2, const float* __restrict__ f0,
                             const float* __restrict__ f1, const float* __restrict__ f2, long long frow, int n,
                             float scale, float ws, unsigned short* __restrict__ feat, unsigned short* __restrict__ target)
{
  const long long cell = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long nc = (long long)n * n * n;
  if (cell >= nc) return;
  const int z = (int)(cell % n);
  const long long xy = cell / n;
  const int y = (int)(xy % n), x = (int)(xy / n);
  const long long ca = xy * crow + z;
  feat[cell] = feature_of(cd[ca], scale);
  target[cell] = f2bf(restrict_fine(f0, frow, 2 * n, x, y, z, ws) - c0[ca]);
  target[nc + cell] = f2bf(restrict_fine(f1, frow, 2 * n, x, y, z, ws) - c1[ca]);
  target[2 * nc + cell] = f2bf(restrict_fine(f2, frow, 2 * n, x, y, z, ws) - c2[ca]);
}

__global__ void __launch_bounds__(kThreads) k_sumsq(const unsigned short* __restrict__ t, long long n, double* out)
{
  __shared__ double part[kThreads / 64];
  double s = 0.0;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double v = bf2f(t[i]);
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) tot += part[w];
    unsafeAtomicAdd(out, tot);
  }
}

long long param_offset(int layer)
{
  long long off = 0;
  for (int l = 0; l < layer; ++l) off += 27LL * kNetCin[l] * kNetCout[l] + kNetCout[l];
  return off;
}

unsigned grid_cells(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

constexpr int kConvBlocks = 512;  // two workgroups per CU on 256 CUs
constexpr int kWgradBlocks = 512;

template <typename T>
int dev_alloc(T** p, long long count)
{
  if (hipMalloc(p, sizeof(T) * count) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    set_error("learned residual: unable to allocate device memory");
    return kErrAlloc;
  }
  return kOk;
}

}  // namespace

int conv_lds_attributes();

int resnet_init(ResNet* r, int n, unsigned seed)
{
  COLA_TRY(conv_lds_attributes());
  if (n < 32 || n % 32 != 0) {
    set_error("learned residual: the grid must be a multiple of 32 cells per axis");
    return kErrArg;
  }
  r->n = n;
  const long long nc = (long long)n * n * n;
  COLA_TRY(dev_alloc(&r->d_param, kNetParams));
  COLA_TRY(dev_alloc(&r->d_m, kNetParams));
  COLA_TRY(dev_alloc(&r->d_v, kNetParams));
  COLA_TRY(dev_alloc(&r->d_grad, kNetParams));
  for (int l = 0; l < kNetLayers; ++l) {
    COLA_TRY(dev_alloc(&r->d_wf[l], 14LL * 64 * 8));
    COLA_TRY(dev_alloc(&r->d_wb[l], 14LL * 64 * 8));
    COLA_TRY(dev_alloc(&r->d_wr[l], (long long)kRollChunks * 64 * 8));
  }
  COLA_TRY(dev_alloc(&r->d_x, nc * 8));
  COLA_TRY(dev_alloc(&r->d_f1, nc));
  for (int l = 0; l < 3; ++l) COLA_TRY(dev_alloc(&r->d_h[l], nc * 16));
  COLA_TRY(dev_alloc(&r->d_y, 3 * nc));
  COLA_TRY(dev_alloc(&r->d_loss, 1));
  COLA_TRY(dev_alloc(&r->d_fsum, 3));
  COLA_TRY(dev_alloc(&r->d_fi, 3 * nc));
  COLA_TRY(dev_alloc(&r->d_trash, 256));
  COLA_TRY(dev_alloc(&r->d_a, 1));
  r->wpart_waves = kWgradBlocks * (kThreads / 64);
  // He-normal initial weights (std sqrt(2 / fan_in)), zero biases and moments
  std::vector<float> p(kNetParams, 0.0f);
  std::mt19937 gen(seed);
  for (int l = 0; l < kNetLayers; ++l) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * kNetCin[l])) * (l == kNetLayers - 1 ? 0.1f : 1.0f));
    const long long off = param_offset(l);
    for (long long i = 0; i < 27LL * kNetCin[l] * kNetCout[l]; ++i) p[off + i] = nd(gen);
  }
  COLA_HIP(hipMemcpy(r->d_param, p.data(), sizeof(float) * kNetParams, hipMemcpyHostToDevice));
  COLA_HIP(hipMemset(r->d_m, 0, sizeof(float) * kNetParams));
  COLA_HIP(hipMemset(r->d_v, 0, sizeof(float) * kNetParams));
  return resnet_pack(r, nullptr);
}

void resnet_free(ResNet* r)
{
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  fr(r->d_param), fr(r->d_m), fr(r->d_v), fr(r->d_grad);
  for (int l = 0; l < kNetLayers; ++l) fr(r->d_wf[l]), fr(r->d_wb[l]), fr(r->d_wr[l]);
  fr(r->d_x), fr(r->d_f1), fr(r->d_y), fr(r->d_loss), fr(r->d_wpart), fr(r->d_a), fr(r->d_fsum), fr(r->d_fi), fr(r->d_trash);
  for (int l = 0; l < 3; ++l) fr(r->d_h[l]);
  for (int l = 0; l < 2; ++l) fr(r->d_g[l]);
  for (auto p : r->s_feat) fr(p);
  for (auto p : r->s_target) fr(p);
  *r = ResNet{};
}

int resnet_pack(ResNet* r, hipStream_t s)
{
  for (int l = 0; l < kNetLayers; ++l) {
    hipLaunchKernelGGL(k_pack, dim3(14), dim3(64), 0, s, r->d_param + param_offset(l), kNetCin[l], kNetCout[l],
                       r->d_wf[l], l > 0 ? r->d_wb[l] : nullptr);
    if (kNetCin[l] == 16 && kNetCout[l] == 16)
      hipLaunchKernelGGL(k_pack_roll, dim3(kRollChunks), dim3(64), 0, s, r->d_param + param_offset(l), kNetCout[l],
                         r->d_wr[l]);
    if (kNetCin[l] == 16 && kNetCout[l] == 3)
      hipLaunchKernelGGL(k_pack_head, dim3(3 * kHeadChunks), dim3(64), 0, s, r->d_param + param_offset(l), r->d_wr[l]);
  }
  COLA_HIP(hipGetLastError());
  return kOk;
}


// Persistent workgroups (two resident per CU: the 16-channel ring is 77 KiB)
// over units of `march` x planes of one (y, z) block column.  The march is
// the longest that still gives every workgroup about four units (load
// balance); its two halo planes are the only planes loaded twice.
struct ConvLaunch {
  unsigned blocks;
  int march;
};

ConvLaunch conv_launch(int n)
{
  const long long cols = (long long)(n / kBlockY) * (n / kBlockZ);
  const long long want = 4LL * kConvBlocks;
  const long long nseg = std::min<long long>(n, std::max<long long>(1, (want + cols - 1) / cols));
  ConvLaunch c;
  c.march = (int)((n + nseg - 1) / nseg);
  const long long units = (long long)((n + c.march - 1) / c.march) * cols;
  c.blocks = (unsigned)std::min<long long>(kConvBlocks, units);
  return c;
}

// k_conv16r: a march of m output planes costs about ceil((m + 2) / 3)
// rotations of three plane steps; the march is the one with the least
// work on the busiest workgroup (units per workgroup times steps per unit).
ConvLaunch conv_launch_roll(int n)
{
  const long long cols = (long long)(n / kBlockY) * (n / kBlockZ);
  long long best = -1;
  ConvLaunch c{1, n};
  for (int nseg = 1; nseg <= n; ++nseg) {
    const int march = (n + nseg - 1) / nseg;
    if ((long long)(nseg - 1) * march >= n) continue;  // the same march as a smaller nseg
    const long long units = (long long)((n + march - 1) / march) * cols;
    const long long blocks = std::min<long long>(kConvBlocks, units);
    const long long cost = (units + blocks - 1) / blocks * (3 * ((march + 4) / 3));
    if (best < 0 || cost < best) {
      best = cost;
      c.blocks = (unsigned)blocks;
      c.march = march;
    }
  }
  return c;
}

// k_conv3h: one workgroup per CU; a march of m output planes takes
// ceil((m + 4) / 3) rotations; the march with the least work on the busiest
// workgroup.
ConvLaunch conv_launch_3h(int n)
{
  const long long cols = (long long)(n / kBlockY) * ((n + kOut3Z - 1) / kOut3Z);
  const long long wgs = 256;
  long long best = -1;
  ConvLaunch c{1, n};
  for (int nseg = 1; nseg <= n; ++nseg) {
    const int march = (n + nseg - 1) / nseg;
    if ((long long)(nseg - 1) * march >= n) continue;
    const long long units = (long long)((n + march - 1) / march) * cols;
    const long long blocks = std::min<long long>(wgs, units);
    const long long cost = (units + blocks - 1) / blocks * (3 * ((march + 6) / 3));
    if (best < 0 || cost < best) {
      best = cost;
      c.blocks = (unsigned)blocks;
      c.march = march;
    }
  }
  return c;
}

constexpr size_t kRing16Bytes = sizeof(uint4) * 2 * kRingSlots * kPlaneCells;
constexpr size_t kRing8Bytes = sizeof(uint4) * kRingSlots * kPlaneCells;

// Dynamic LDS above 64 KiB must be allowed per kernel once.
int conv_lds_attributes()
{
  static bool done = false;
  if (done) return kOk;
  const void* k16[] = {reinterpret_cast<const void*>(&k_conv16<kFwdRelu>),
                       reinterpret_cast<const void*>(&k_conv16<kFwdLinear>),
                       reinterpret_cast<const void*>(&k_conv16<kBwdData>),
                       reinterpret_cast<const void*>(&k_conv16<kFwdAdd>)};
  for (const void* k : k16) COLA_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRing16Bytes));
  const void* kr[] = {reinterpret_cast<const void*>(&k_conv16r<kFwdRelu>),
                      reinterpret_cast<const void*>(&k_conv_head_r<kFwdAdd>),
                      reinterpret_cast<const void*>(&k_conv_head_r<kFwdLinear>)};
  for (const void* k : kr) COLA_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRollRingBytes));
  COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3h), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kConv3hBytes));
  COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv8), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kRing8Bytes));
  done = true;
  return kOk;
}

namespace {

// layers 2 and 3 (h1 -> h2 -> h3)
// hidden layer l (1 or 2: h_{l-1} -> h_l, relu)
void conv_hidden(ResNet* r, int l, hipStream_t s)
{
  const int n = r->n;
  const ConvLaunch cl = conv_launch(n);
  ConvOut o;
  o.bf = r->d_h[l];
  const float* bias = r->d_param + param_offset(l) + 27 * 16 * 16;
  if (r->roll) {
    const ConvLaunch cr = conv_launch_roll(n);
    hipLaunchKernelGGL((k_conv16r<kFwdRelu>), dim3(cr.blocks), dim3(kThreads), kRollRingBytes, s, r->d_h[l - 1],
                       r->d_wr[l], bias, 16, n, cr.march, o);
  } else
    hipLaunchKernelGGL((k_conv16<kFwdRelu>), dim3(cl.blocks), dim3(kThreads), kRing16Bytes, s, r->d_h[l - 1],
                       r->d_wf[l], bias, 16, n, cl.march, o);
}

int conv_layers_2_3(ResNet* r, hipStream_t s)
{
  conv_hidden(r, 1, s);
  conv_hidden(r, 2, s);
  COLA_HIP(hipGetLastError());
  return kOk;
}

// layer 4 (h3 -> output) into `o` by mode
int conv_layer_4(ResNet* r, int mode, const ConvOut& o, hipStream_t s)
{
  const int n = r->n;
  const ConvLaunch cl = conv_launch(n);
  const float* b4 = r->d_param + param_offset(3) + 27 * 16 * 3;
  if (r->roll) {
    const ConvLaunch cr = conv_launch_roll(n);
    if (mode == kFwdAdd)
      hipLaunchKernelGGL((k_conv_head_r<kFwdAdd>), dim3(cr.blocks), dim3(kThreads), kRollRingBytes, s, r->d_h[2],
                         r->d_wr[3], b4, n, cr.march, o);
    else
      hipLaunchKernelGGL((k_conv_head_r<kFwdLinear>), dim3(cr.blocks), dim3(kThreads), kRollRingBytes, s, r->d_h[2],
                         r->d_wr[3], b4, n, cr.march, o);
    COLA_HIP(hipGetLastError());
    return kOk;
  }
  if (mode == kFwdAdd)
    hipLaunchKernelGGL((k_conv16<kFwdAdd>), dim3(cl.blocks), dim3(kThreads), kRing16Bytes, s, r->d_h[2], r->d_wf[3],
                       b4, 3, n, cl.march, o);
  else
    hipLaunchKernelGGL((k_conv16<kFwdLinear>), dim3(cl.blocks), dim3(kThreads), kRing16Bytes, s, r->d_h[2],
                       r->d_wf[3], b4, 3, n, cl.march, o);
  COLA_HIP(hipGetLastError());
  return kOk;
}

// First layer from a bf16 feature grid.
int conv_layer_1(ResNet* r, const unsigned short* feat, const float* pa, float a, hipStream_t s)
{
  const ConvLaunch cl = conv_launch(r->n);
  const long long units = (long long)((r->n + cl.march - 1) / cl.march) * (r->n / kBlockY) * (r->n / kBlockZ);
  const unsigned blocks = (unsigned)std::min<long long>(256LL * kConv1Waves, units);
  hipLaunchKernelGGL(k_conv1, dim3(blocks), dim3(kThreads), 0, s, feat, r->d_wf[0], r->d_param + param_offset(0),
                     r->d_param + param_offset(0) + 27 * 8 * 16, pa, a, r->n, cl.march, r->d_h[0]);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int conv_layers_2_to_4(ResNet* r, int mode4, const ConvOut& o4, hipStream_t s)
{
  COLA_TRY(conv_layers_2_3(r, s));
  return conv_layer_4(r, mode4, o4, s);
}

}  // namespace

int resnet_forward(ResNet* r, hipStream_t s)
{
  const int n = r->n;
  COLA_TRY(conv_lds_attributes());
  const ConvLaunch cl = conv_launch(n);
  hipLaunchKernelGGL(k_conv8, dim3(cl.blocks), dim3(kThreads), kRing8Bytes, s, r->d_x, r->d_wf[0],
                     r->d_param + param_offset(0) + 27 * 8 * 16, n, cl.march, r->d_h[0]);
  ConvOut o;
  o.f = r->d_y;
  return conv_layers_2_to_4(r, kFwdLinear, o, s);
}

int resnet_forward_features(ResNet* r, const unsigned short* feat, const float* pa, float a, hipStream_t s)
{
  COLA_TRY(conv_lds_attributes());
  COLA_TRY(conv_layer_1(r, feat, pa, a, s));
  ConvOut o;
  o.f = r->d_y;
  return conv_layers_2_to_4(r, kFwdLinear, o, s);
}

int resnet_features(ResNet* r, const float* delta, long long row, hipStream_t s)
{
  hipLaunchKernelGGL(k_features, dim3((unsigned)((long long)r->n * r->n)), dim3(kThreads), 0, s, delta, row, r->n,
                     r->in_scale, r->d_f1);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int resnet_body(ResNet* r, const float* pa, hipStream_t s)
{
  COLA_TRY(conv_lds_attributes());
  COLA_TRY(conv_layer_1(r, r->d_f1, pa, 0.0f, s));
  return conv_layers_2_3(r, s);
}

int resnet_layer1(ResNet* r, const float* pa, hipStream_t s)
{
  COLA_TRY(conv_lds_attributes());
  return conv_layer_1(r, r->d_f1, pa, 0.0f, s);
}

int resnet_layers23(ResNet* r, hipStream_t s)
{
  COLA_TRY(conv_lds_attributes());
  return conv_layers_2_3(r, s);
}

int resnet_layers12(ResNet* r, const float* delta, long long row, const float* pa, hipStream_t s)
{
  if (r->roll) {
    const ConvLaunch cr = conv_launch_roll(r->n);
    hipLaunchKernelGGL(k_conv12r, dim3(cr.blocks), dim3(kThreads), kConv12rBytes, s, delta, row, r->in_scale,
                       r->d_wf[0], r->d_param + param_offset(0), r->d_param + param_offset(0) + 27 * 8 * 16, pa, 0.0f,
                       r->d_wr[1], r->d_param + param_offset(1) + 27 * 16 * 16, r->n, cr.march, r->d_h[1]);
    COLA_HIP(hipGetLastError());
    return kOk;
  }
  // by output planes there is no fused pass: features, layer 1, layer 2
  COLA_TRY(resnet_features(r, delta, row, s));
  COLA_TRY(conv_layer_1(r, r->d_f1, pa, 0.0f, s));
  conv_hidden(r, 1, s);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int resnet_layer3(ResNet* r, hipStream_t s)
{
  COLA_TRY(conv_lds_attributes());
  conv_hidden(r, 2, s);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int resnet_head_add(ResNet* r, float* const out[3], long long row, hipStream_t s)
{
  ConvOut o;
  for (int a = 0; a < 3; ++a) o.add[a] = out[a];
  o.row = row;
  o.scale = r->out_scale;
  hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, s, r->d_fsum);
  o.sum = r->d_fsum;
  o.fi = r->d_fi;
  return conv_layer_4(r, kFwdAdd, o, s);
}

int resnet_layer3_head(ResNet* r, float* const out[3], long long row, hipStream_t s)
{
  COLA_TRY(conv_lds_attributes());
  ConvOut o;
  for (int a = 0; a < 3; ++a) o.add[a] = out[a];
  o.row = row;
  o.scale = r->out_scale;
  hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, s, r->d_fsum);
  o.sum = r->d_fsum;
  o.fi = r->d_fi;
  o.trash = r->d_trash;
  const ConvLaunch cl = conv_launch_3h(r->n);
  hipLaunchKernelGGL(k_conv3h, dim3(cl.blocks), dim3(kThreads3), kConv3hBytes, s, r->d_h[1], r->d_wr[2],
                     r->d_param + param_offset(2) + 27 * 16 * 16, r->d_wr[3], r->d_param + param_offset(3) + 27 * 16 * 3,
                     r->n, cl.march, o);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int resnet_backward(ResNet* r, const unsigned short* tb, const float* tf, hipStream_t s)
{
  const int n = r->n;
  const long long nc = (long long)n * n * n;
  if (!r->d_g[0]) {
    COLA_TRY(dev_alloc(&r->d_g[0], nc * 16));
    COLA_TRY(dev_alloc(&r->d_g[1], nc * 16));
    COLA_TRY(dev_alloc(&r->d_wpart, (long long)r->wpart_waves * (27 * 256 + 16)));
  }
  COLA_TRY(conv_lds_attributes());
  hipLaunchKernelGGL(k_zero1, dim3(1), dim3(64), 0, s, r->d_loss);
  hipLaunchKernelGGL(k_loss, dim3(std::min<long long>(4096, grid_cells(nc))), dim3(kThreads), 0, s, r->d_y, tb, tf,
                     n, 1.0f / r->out_scale, r->d_g[0], r->d_loss);
  const unsigned short* X[kNetLayers] = {r->d_x, r->d_h[0], r->d_h[1], r->d_h[2]};
  unsigned short* G = r->d_g[0];
  unsigned short* Gp = r->d_g[1];
  for (int l = kNetLayers - 1; l >= 0; --l) {
    if (kNetCin[l] == 8)
      hipLaunchKernelGGL(k_conv_wgrad<8>, dim3(kWgradBlocks), dim3(kThreads), 0, s, X[l], G, n, r->d_wpart);
    else
      hipLaunchKernelGGL(k_conv_wgrad<16>, dim3(kWgradBlocks), dim3(kThreads), 0, s, X[l], G, n, r->d_wpart);
    const int np = 27 * kNetCin[l] * kNetCout[l] + kNetCout[l];
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((np + 255) / 256), dim3(256), 0, s, r->d_wpart, r->wpart_waves,
                       kNetCin[l], kNetCout[l], r->d_grad + param_offset(l));
    if (l == 0) break;
    // gradient at the output of layer l - 1: data gradient of layer l masked by relu'(h_{l-1})
    const ConvLaunch cl = conv_launch(n);
    ConvOut o;
    o.bf = Gp;
    o.hprev = r->d_h[l - 1];
    hipLaunchKernelGGL((k_conv16<kBwdData>), dim3(cl.blocks), dim3(kThreads), kRing16Bytes, s, G, r->d_wb[l],
                       nullptr, 0, n, cl.march, o);
    std::swap(G, Gp);
  }
  COLA_HIP(hipGetLastError());
  return kOk;
}

int resnet_adam(ResNet* r, float lr, hipStream_t s)
{
  const float b1 = 0.9f, b2 = 0.999f;
  r->adam_t += 1;
  const float c1 = 1.0f - std::pow(b1, (float)r->adam_t), c2 = 1.0f - std::pow(b2, (float)r->adam_t);
  hipLaunchKernelGGL(k_adam, dim3((kNetParams + 255) / 256), dim3(256), 0, s, r->d_param, r->d_m, r->d_v, r->d_grad,
                     kNetParams, lr, b1, b2, c1, c2, 1e-8f);
  COLA_HIP(hipGetLastError());
  return resnet_pack(r, s);
}

int resnet_store_pair(ResNet* r, const float* coarse_delta, long long crow, const float* const cf[3],
                      const float* const ff[3], long long frow, float a, hipStream_t s)
{
  const long long nc = (long long)r->n * r->n * r->n;
  unsigned short *feat = nullptr, *target = nullptr;
  COLA_TRY(dev_alloc(&feat, nc));
  if (dev_alloc(&target, 3 * nc) != kOk) {
    (void)hipFree(feat);
    return kErrAlloc;
  }
  r->s_feat.push_back(feat);
  r->s_target.push_back(target);
  r->s_a.push_back(a);
  hipLaunchKernelGGL(k_store_pair, dim3(grid_cells(nc)), dim3(kThreads), 0, s, coarse_delta, crow, cf[0], cf[1], cf[2],
                     ff[0], ff[1], ff[2], frow, r->n, r->in_scale, r->restrict_side, feat, target);
  COLA_HIP(hipGetLastError());
  return kOk;
}

namespace {

int run_sample(ResNet* r, int i, bool train, float lr, double* loss, hipStream_t s)
{
  const long long nc = (long long)r->n * r->n * r->n;
  // the 8-channel input is the first layer's weight-gradient operand
  hipLaunchKernelGGL(k_features_stored, dim3(grid_cells(nc)), dim3(kThreads), 0, s, r->s_feat[i], r->n, r->s_a[i],
                     r->d_x);
  COLA_TRY(resnet_forward_features(r, r->s_feat[i], nullptr, r->s_a[i], s));
  COLA_TRY(resnet_backward(r, r->s_target[i], nullptr, s));
  if (train) COLA_TRY(resnet_adam(r, lr, s));
  COLA_HIP(hipMemcpyAsync(loss, r->d_loss, sizeof(double), hipMemcpyDeviceToHost, s));
  COLA_HIP(hipStreamSynchronize(s));
  return kOk;
}

}  // namespace

// out_scale = RMS of the stored targets (so a zero prediction has loss 1)
int resnet_fit_scale(ResNet* r, int count, hipStream_t s)
{
  if (count < 1 || count > (int)r->s_target.size()) {
    set_error("learned residual: sample count out of range");
    return kErrArg;
  }
  const long long nc = (long long)r->n * r->n * r->n;
  hipLaunchKernelGGL(k_zero1, dim3(1), dim3(64), 0, s, r->d_loss);
  for (int i = 0; i < count; ++i)
    hipLaunchKernelGGL(k_sumsq, dim3(2048), dim3(kThreads), 0, s, r->s_target[i], 3 * nc, r->d_loss);
  double sq = 0.0;
  COLA_HIP(hipMemcpyAsync(&sq, r->d_loss, sizeof(double), hipMemcpyDeviceToHost, s));
  COLA_HIP(hipStreamSynchronize(s));
  const double rms = std::sqrt(sq / (3.0 * (double)nc * count));
  r->out_scale = rms > 0.0 ? (float)rms : 1.0f;
  r->scale_set = true;
  return kOk;
}

int resnet_eval(ResNet* r, int first, int count, double* loss, hipStream_t s)
{
  if (first < 0 || count < 1 || first + count > (int)r->s_feat.size()) {
    set_error("learned residual: sample range out of bounds");
    return kErrArg;
  }
  double tot = 0.0;
  for (int i = first; i < first + count; ++i) {
    double l = 0.0;
    COLA_TRY(run_sample(r, i, false, 0.0f, &l, s));
    tot += l;
  }
  *loss = tot / count;
  return kOk;
}

int resnet_train(ResNet* r, int ntrain, int epochs, float lr, unsigned seed, double* epoch_loss, hipStream_t s)
{
  if (ntrain < 1 || ntrain > (int)r->s_feat.size() || epochs < 1) {
    set_error("learned residual: need 1 <= ntrain <= stored samples and epochs >= 1");
    return kErrArg;
  }
  if (!r->scale_set) COLA_TRY(resnet_fit_scale(r, ntrain, s));
  std::vector<int> order(ntrain);
  for (int i = 0; i < ntrain; ++i) order[i] = i;
  std::mt19937 gen(seed);
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(order.begin(), order.end(), gen);
    double tot = 0.0;
    for (int i : order) {
      double l = 0.0;
      COLA_TRY(run_sample(r, i, true, lr, &l, s));
      tot += l;
    }
    if (epoch_loss) epoch_loss[e] = tot / ntrain;
  }
  return kOk;
}



}  // namespace cola
