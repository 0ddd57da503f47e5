// Stockham radix-2/3/4/5/8 FFT passes staged through LDS (see fft.h).
//
// Radix pass p with radix R and L = product of the earlier radices maps
// butterfly b (0 <= b < n/R) as
//   inputs  x[b + s*n/R], s < R,  twiddled by w_{RL}^{s*(b mod L)}
//   outputs y[(b div L)*R*L + t*L + (b mod L)], t < R,
// which leaves the result in natural order after the last pass (Stockham
// autosort), so no bit-reversal pass is needed.
#include "common.h"
#include "fft.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace cola {

namespace {

constexpr float kS3 = 0.866025403784438646763723170752936183f;   // sin(2pi/3)
constexpr float kC5a = 0.309016994374947424102293417182819059f;  // cos(2pi/5)
constexpr float kC5b = -0.809016994374947424102293417182819059f; // cos(4pi/5)
constexpr float kS5a = 0.951056516295153572116439333379382143f;  // sin(2pi/5)
constexpr float kS5b = 0.587785252292473129168705954639072769f;  // sin(4pi/5)
constexpr float kR2 = 0.707106781186547524400844362104849039f;   // sqrt(2)/2

// The butterflies and twiddle products contract their multiply-adds into
// fused multiply-adds (written out, so every kernel rounds them alike): the
// transforms are compared with the reference's FFTW by tolerance, not bit
// for bit, and the pass is VALU-bound.

// multiply by -i (DIR < 0) or +i (DIR > 0)
template <int DIR>
__device__ inline float2 rot(float2 z)
{
  return DIR < 0 ? make_float2(z.y, -z.x) : make_float2(-z.y, z.x);
}

// a * w (DIR < 0) or a * conj(w) (DIR > 0): two products, two fused adds
template <int DIR>
__device__ inline float2 twiddle(float2 a, float2 w)
{
  if (DIR < 0) return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x));
  return make_float2(fmaf(a.x, w.x, a.y * w.y), fmaf(-a.x, w.y, a.y * w.x));
}

template <int DIR>
__device__ inline void dft2(float2* v)
{
  const float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR>
__device__ inline void dft3(float2* v)
{
  const float2 t = cadd(v[1], v[2]);
  const float2 d = csub(v[1], v[2]);
  const float2 m = make_float2(fmaf(-0.5f, t.x, v[0].x), fmaf(-0.5f, t.y, v[0].y));
  // m +- rot(kS3 d)
  const float s = DIR < 0 ? kS3 : -kS3;
  v[0] = cadd(v[0], t);
  v[1] = make_float2(fmaf(s, d.y, m.x), fmaf(-s, d.x, m.y));
  v[2] = make_float2(fmaf(-s, d.y, m.x), fmaf(s, d.x, m.y));
}

template <int DIR>
__device__ inline void dft4(float2* v)
{
  const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const float2 t2 = cadd(v[1], v[3]), t3 = rot<DIR>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int DIR>
__device__ inline void dft5(float2* v)
{
  const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const float2 x0 = v[0];
  const float2 a1 = make_float2(fmaf(kC5b, t2.x, fmaf(kC5a, t1.x, x0.x)), fmaf(kC5b, t2.y, fmaf(kC5a, t1.y, x0.y)));
  const float2 a2 = make_float2(fmaf(kC5a, t2.x, fmaf(kC5b, t1.x, x0.x)), fmaf(kC5a, t2.y, fmaf(kC5b, t1.y, x0.y)));
  const float2 b1 = rot<DIR>(make_float2(fmaf(kS5b, t4.x, kS5a * t3.x), fmaf(kS5b, t4.y, kS5a * t3.y)));
  const float2 b2 = rot<DIR>(make_float2(fmaf(-kS5a, t4.x, kS5b * t3.x), fmaf(-kS5a, t4.y, kS5b * t3.y)));
  v[0] = make_float2(x0.x + t1.x + t2.x, x0.y + t1.y + t2.y);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <int DIR>
__device__ inline void dft8(float2* v)
{
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // o[k] *= exp(DIR * 2 pi i k / 8): o[1] = kR2 s1, o[3] = kR2 s3, the
  // factor kR2 taken into the final adds
  const float2 o1 = o[1], o3 = o[3];
  float2 s1, s3;
  if (DIR < 0) {
    s1 = make_float2(o1.x + o1.y, o1.y - o1.x);
    s3 = make_float2(o3.y - o3.x, -(o3.x + o3.y));
  } else {
    s1 = make_float2(o1.x - o1.y, o1.x + o1.y);
    s3 = make_float2(-(o3.x + o3.y), o3.x - o3.y);
  }
  const float2 o2 = rot<DIR>(o[2]);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[1] = make_float2(fmaf(kR2, s1.x, e[1].x), fmaf(kR2, s1.y, e[1].y));
  v[5] = make_float2(fmaf(-kR2, s1.x, e[1].x), fmaf(-kR2, s1.y, e[1].y));
  v[3] = make_float2(fmaf(kR2, s3.x, e[3].x), fmaf(kR2, s3.y, e[3].y));
  v[7] = make_float2(fmaf(-kR2, s3.x, e[3].x), fmaf(-kR2, s3.y, e[3].y));
}

template <int R, int DIR>
__device__ inline void dft(float2* v)
{
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 5) dft5<DIR>(v);
  else dft8<DIR>(v);
}

__device__ inline unsigned fdiv(unsigned x, unsigned d, unsigned mag) { return d == 1 ? x : __umulhi(x, mag); }

// Signed wave number of index i: the PM convention keeps the Nyquist index
// at +n/2 (pm.c:325,331), the LPT convention maps it to -n/2 (lpt.c:362-375).
__device__ inline int wave_pm(int i, int n) { return i <= n / 2 ? i : i - n; }

__device__ inline double kvec_lpt(int i, int n, double box)
{
  // lpt.c: i * 2 * M_PI / Box  or  -(Nmesh - i) * 2 * M_PI / Box
  if (i < n / 2) return (double)(i * 2) * kPiD / box;
  return (double)(-(n - i) * 2) * kPiD / box;
}

template <KOp OP>
__device__ inline float2 apply_kop(float2 v, int ix, int iy, int iz, int n, const KOpParams& kp)
{
  if constexpr (OP == KOp::PmPoisson) {
    const int I0 = wave_pm(ix, n), J0 = wave_pm(iy, n), K = iz;
    if (I0 == 0 && J0 == 0 && K == 0) return make_float2(0.0f, 0.0f);
    const float di = (float)(kp.axis == 0 ? I0 : (kp.axis == 1 ? J0 : K));
    const float f2 = kp.f1 / (float)(K * K + I0 * I0 + J0 * J0) * di;
    return make_float2(-f2 * v.y, f2 * v.x);
  } else if constexpr (OP == KOp::LptGrad) {
    const int idx = kp.axis == 0 ? ix : (kp.axis == 1 ? iy : iz);
    const double k = kvec_lpt(idx, n, kp.box);
    return make_float2((float)((double)(-v.y) * k), (float)((double)v.x * k));
  } else if constexpr (OP == KOp::Lpt2) {
    const double k0 = kvec_lpt(ix, n, kp.box), k1 = kvec_lpt(iy, n, kp.box), k2 = kvec_lpt(iz, n, kp.box);
    const double kmag2 = k0 * k0 + k1 * k1 + k2 * k2;
    if (!(kmag2 > 0.0)) return make_float2(0.0f, 0.0f);
    const double ka = kp.axis == 0 ? k0 : (kp.axis == 1 ? k1 : k2);
    return make_float2((float)((double)v.y * ka / kmag2), (float)((double)(-v.x) * ka / kmag2));
  } else {
    return v;
  }
}

enum Mode : int { kCols = 0, kR2C = 1, kC2R = 2 };

struct TileArgs {
  FftRadices P;
  int log2T;             // transforms per workgroup (columns or row pairs)
  int tpb;               // kCols: tiles per batch (grid is 1-D: batch-major)
  int tpb_pad;           // kCols: tpb rounded up to a multiple of the XCD group
  int xgroup;            // kCols: adjacent tiles per XCD group
  int ncols;             // valid columns (kCols) or row pairs (kR2C / kC2R)
  long long sr;          // kCols: stride between positions of one transform
  long long bs;          // kCols: batch stride (blockIdx.y)
  long long rs_in;       // kR2C: real row stride (floats); kC2R: complex row stride
  long long rs_out;      // kR2C: complex row stride; kC2R: real row stride (floats)
  int nz;                // x pass: columns are (y, kz) pairs with kz < nz (the row stride)
  unsigned magNz;
  int zero_pad = 0;      // kC2R: write zeros to the two padding floats of each real row
  int kz_valid = 0;      // x pass: tiles lying wholly in kz >= kz_valid (row padding) return at once
  KOpParams kp;
};

// An x-pass tile of 2^log2T columns starting at c0 that holds only padding
// columns (kz_valid <= kz < nz within one row).
__device__ __forceinline__ bool padding_tile(const TileArgs& A, int c0, int log2T)
{
  if (A.kz_valid <= 0) return false;
  const int kz = (int)((unsigned)c0 % (unsigned)A.nz);
  return kz >= A.kz_valid && kz + (1 << log2T) <= A.nz;
}

// One workgroup: T transforms of length n.  MODE kCols: column c of the tile
// is global column c0 + c, positions r at stride sr; LDS [r][c].  kR2C/kC2R:
// transform c packs real rows 2(c0+c), 2(c0+c)+1; LDS [c][r] padded by one.
// Radix sets compiled into a kernel: every radix, or only 3, 4 and 8 (the
// passes of n = 2^a 3^b with a != 1 mod 3 or a >= 4, 768 among them).  The
// smaller set needs fewer registers: the column pass then fits 8 waves per
// SIMD, two 1024-thread workgroups per CU.
constexpr int kRsAll = 0;
constexpr int kRs348 = 1;

template <int RS, class F>
__device__ inline void radix_dispatch(int R, F&& f)
{
  if constexpr (RS == kRs348) {
    switch (R) {
      case 3: f(std::integral_constant<int, 3>{}); break;
      case 4: f(std::integral_constant<int, 4>{}); break;
      default: f(std::integral_constant<int, 8>{}); break;
    }
  } else {
    switch (R) {
      case 2: f(std::integral_constant<int, 2>{}); break;
      case 3: f(std::integral_constant<int, 3>{}); break;
      case 4: f(std::integral_constant<int, 4>{}); break;
      case 5: f(std::integral_constant<int, 5>{}); break;
      default: f(std::integral_constant<int, 8>{}); break;
    }
  }
}

// Minimum waves per SIMD requested for the radix-{3,4,8} kernels with 8
// values per thread: 8 for the plain column pass (61 VGPRs), 6 for the row
// passes (<= 80 VGPRs); both without spills.
constexpr int tile_waves_per_eu(int mode, KOp op, int emax, int rs)
{
  if (emax != 8 || rs != kRs348) return 1;
  if (mode == 0) return op == KOp::None ? 8 : 1;
  return 6;
}

// Workgroup blk of a pass whose grid is nblk workgroups (the kernel's own
// grid, or one job's share of a batched launch).
template <int DIR, int MODE, KOp OP, int EMAX, int RS>
__device__ __forceinline__ void tile_run(const void* __restrict__ in_, void* __restrict__ out_,
                                         const float2* __restrict__ tw, const TileArgs& A, int blk, int nblk)
{
  extern __shared__ float2 smem[];
  const int n = A.P.n;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int log2T = A.log2T;
  const int T = 1 << log2T;
  int c0, batch = 0;
  if (MODE == kCols) {
    const int lb = xcd_group_remap(blk, nblk, A.xgroup);
    batch = lb / A.tpb_pad;
    const int tile = lb - batch * A.tpb_pad;
    if (tile >= A.tpb) return;  // padding workgroup (before any barrier)
    c0 = tile << log2T;
    if (padding_tile(A, c0, log2T)) return;
  } else {
    c0 = blk << log2T;
  }
  float2* tws = smem;
  float2* lds = smem + n;
  for (int k = tid; k < n; k += nthr) tws[k] = tw[k];

  auto lidx = [&](int r, int c) { return MODE == kCols ? (r << log2T) + c : c * (n + 1) + r; };
  auto split = [&](int u, const FftPass& p, int& c, int& b) {
    if (MODE == kCols) {
      c = u & (T - 1);
      b = u >> log2T;
    } else {
      c = (int)fdiv((unsigned)u, (unsigned)p.m, p.magM);
      b = u - c * p.m;
    }
  };

  using Tile = float2[EMAX];

  // ---- pass 0: inputs straight from global memory
  auto load_global = [&](auto Rc, Tile& reg) {
    constexpr int R = decltype(Rc)::value;
    const FftPass& p = A.P.pass[0];
    const int nb = p.m << log2T;
#pragma unroll
    for (int q = 0; q < EMAX / R; ++q) {
      const int u = tid + q * nthr;
      int c = 0, b = 0;
      split(u, p, c, b);
      const bool live = u < nb && c0 + c < A.ncols;
      if (MODE == kCols) {
        // uniform base per leg s (scalar registers) + 32-bit lane offset b*sr + c
        const float2* in = static_cast<const float2*>(in_) + (long long)batch * A.bs + c0;
        const unsigned lane = (unsigned)b * (unsigned)A.sr + (unsigned)c;
        int iy = 0, iz = 0;
        if constexpr (OP != KOp::None) {
          const int col = c0 + c;
          iy = (int)fdiv((unsigned)col, (unsigned)A.nz, A.magNz);
          iz = col - iy * A.nz;
        }
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const float2* leg = in + (long long)(s * p.m) * A.sr;
          reg[q * R + s] = live ? leg[lane] : make_float2(0.0f, 0.0f);
        }
        // the operator runs after all of this butterfly's loads are issued
        if constexpr (OP != KOp::None) {
#pragma unroll
          for (int s = 0; s < R; ++s) reg[q * R + s] = apply_kop<OP>(reg[q * R + s], b + s * p.m, iy, iz, n, A.kp);
        }
      } else if (MODE == kR2C) {
        const float* in = static_cast<const float*>(in_) + (2LL * c0) * A.rs_in;
        const unsigned lane = (unsigned)(2 * c) * (unsigned)A.rs_in + (unsigned)b;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const float* leg = in + s * p.m;
          reg[q * R + s] = live ? make_float2(leg[lane], leg[lane + (unsigned)A.rs_in]) : make_float2(0.0f, 0.0f);
        }
      } else {
        const float2* in = static_cast<const float2*>(in_) + (2LL * c0) * A.rs_in;
        const unsigned rowa = (unsigned)(2 * c) * (unsigned)A.rs_in;
        const unsigned rowb = rowa + (unsigned)A.rs_in;
        const int h = n / 2;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = b + s * p.m;
          float2 v = make_float2(0.0f, 0.0f);
          if (live) {
            // Hermitian extension of both half spectra; Im of kz = 0, n/2 dropped
            const unsigned rr = (unsigned)(r <= h ? r : n - r);
            float2 a = in[rowa + rr], bb = in[rowb + rr];
            if (r == 0 || r == h) {
              a.y = 0.0f;
              bb.y = 0.0f;
            } else if (r > h) {
              a.y = -a.y;
              bb.y = -bb.y;
            }
            v = make_float2(a.x - bb.y, a.y + bb.x);
          }
          reg[q * R + s] = v;
        }
      }
    }
  };

  auto compute = [&](auto Rc, int pi, Tile& reg) {
    constexpr int R = decltype(Rc)::value;
    const FftPass& p = A.P.pass[pi];
    const int nb = p.m << log2T;
#pragma unroll
    for (int q = 0; q < EMAX / R; ++q) {
      const int u = tid + q * nthr;
      if (u < nb) {
        int c, b;
        split(u, p, c, b);
        if (p.L > 1) {
          const int i = (int)fdiv((unsigned)b, (unsigned)p.L, p.magL);
          const int j = b - i * p.L;
#pragma unroll
          for (int s = 1; s < R; ++s) {
            reg[q * R + s] = twiddle<DIR>(reg[q * R + s], tws[s * j * p.twstep]);
          }
        }
        dft<R, DIR>(&reg[q * R]);
      }
    }
  };

  // outputs of pass pi to LDS at positions (b div L) R L + t L + (b mod L)
  auto store_lds = [&](auto Rc, int pi, Tile& reg) {
    constexpr int R = decltype(Rc)::value;
    const FftPass& p = A.P.pass[pi];
    const int nb = p.m << log2T;
#pragma unroll
    for (int q = 0; q < EMAX / R; ++q) {
      const int u = tid + q * nthr;
      if (u < nb) {
        int c, b;
        split(u, p, c, b);
        const int i = (int)fdiv((unsigned)b, (unsigned)p.L, p.magL);
        const int j = b - i * p.L;
        const int o = i * R * p.L + j;
#pragma unroll
        for (int t = 0; t < R; ++t) lds[lidx(o + t * p.L, c)] = reg[q * R + t];
      }
    }
  };

  auto load_lds = [&](auto Rc, int pi, Tile& reg) {
    constexpr int R = decltype(Rc)::value;
    const FftPass& p = A.P.pass[pi];
    const int nb = p.m << log2T;
#pragma unroll
    for (int q = 0; q < EMAX / R; ++q) {
      const int u = tid + q * nthr;
      if (u < nb) {
        int c, b;
        split(u, p, c, b);
#pragma unroll
        for (int s = 0; s < R; ++s) reg[q * R + s] = lds[lidx(b + s * p.m, c)];
      }
    }
  };

  auto store_global = [&](auto Rc, int pi, Tile& reg) {
    constexpr int R = decltype(Rc)::value;
    const FftPass& p = A.P.pass[pi];
    const int nb = p.m << log2T;
#pragma unroll
    for (int q = 0; q < EMAX / R; ++q) {
      const int u = tid + q * nthr;
      int c = 0, b = 0;
      split(u, p, c, b);
      if (u < nb && c0 + c < A.ncols) {
        // last pass: L = n / R = m, so b div L = 0 and the outputs sit at b + t*m
        if (MODE == kCols) {
          float2* out = static_cast<float2*>(out_) + (long long)batch * A.bs + c0;
          const unsigned lane = (unsigned)b * (unsigned)A.sr + (unsigned)c;
#pragma unroll
          for (int t = 0; t < R; ++t) out[(long long)(t * p.m) * A.sr + lane] = reg[q * R + t];
        } else {
          float* out = static_cast<float*>(out_) + (2LL * c0) * A.rs_out;
          const unsigned lane = (unsigned)(2 * c) * (unsigned)A.rs_out + (unsigned)b;
#pragma unroll
          for (int t = 0; t < R; ++t) {
            const float2 z = reg[q * R + t];
            float* leg = out + t * p.m;
            leg[lane] = z.x;
            leg[lane + (unsigned)A.rs_out] = z.y;
          }
        }
      }
    }
  };

  auto with_radix = [&](int R, auto&& f) { radix_dispatch<RS>(R, f); };

  // Each pass keeps its butterflies in a register tile local to its radix
  // case: inputs from HBM (first pass) or LDS, outputs to LDS or HBM (last).
  const int np_ = A.P.npass;
  for (int pi = 0; pi < np_; ++pi) {
    with_radix(A.P.pass[pi].R, [&](auto Rc) {
      Tile reg;
      const bool last = pi == np_ - 1;
      if (pi == 0) {
        load_global(Rc, reg);
      } else {
        load_lds(Rc, pi, reg);
        __syncthreads();
      }
      compute(Rc, pi, reg);
      if (!last || MODE == kR2C) {
        store_lds(Rc, pi, reg);
        __syncthreads();
      } else {
        store_global(Rc, pi, reg);
      }
    });
  }
  if (MODE == kR2C) {
    // unpack the two real rows: A = (Z_k + conj Z_{n-k}) / 2, B = (Z_k - conj Z_{n-k}) / (2i)
    float2* out = static_cast<float2*>(out_);
    const int h = n / 2;
    for (int c = 0; c < T; ++c) {
      if (c0 + c >= A.ncols) break;
      float2* oa = out + (2LL * (c0 + c)) * A.rs_out;
      float2* ob = oa + A.rs_out;
      for (int k = tid; k <= h; k += nthr) {
        const float2 z = lds[lidx(k, c)];
        const float2 zm = lds[lidx(k == 0 ? 0 : n - k, c)];
        oa[k] = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
        ob[k] = make_float2(0.5f * (z.y + zm.y), 0.5f * (zm.x - z.x));
      }
    }
  } else {
    if (MODE == kC2R && A.zero_pad && tid < T && c0 + tid < A.ncols) {
      // the two padding floats of each FFTW-layout real row are written as zero
      float* oa = static_cast<float*>(out_) + (2LL * (c0 + tid)) * A.rs_out;
      oa[n] = 0.0f;
      oa[n + 1] = 0.0f;
      oa[A.rs_out + n] = 0.0f;
      oa[A.rs_out + n + 1] = 0.0f;
    }
  }
}

template <int DIR, int MODE, KOp OP, int EMAX, int RS>
__global__ void __launch_bounds__(kFftMaxThreads) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu(MODE, OP, EMAX, RS))))
fft_tile(const void* __restrict__ in_, void* __restrict__ out_, const float2* __restrict__ tw, TileArgs A)
{
  tile_run<DIR, MODE, OP, EMAX, RS>(in_, out_, tw, A, blockIdx.x, gridDim.x);
}

// ---- fused x pass: forward x FFT, then per axis Poisson + inverse x FFT ----
// Column layout helpers (LDS [r][c], column fastest).
struct ColCtx {
  float2* lds;
  const float2* tws;
  int tid, nthr, log2T;
};

template <int R, int EMAX, int DIR>
__device__ inline void col_compute(float2 (&reg)[EMAX], const FftPass& p, const ColCtx& x)
{
  const int nb = p.m << x.log2T;
#pragma unroll
  for (int q = 0; q < EMAX / R; ++q) {
    const int u = x.tid + q * x.nthr;
    if (u < nb) {
      const int b = u >> x.log2T;
      if (p.L > 1) {
        const int i = (int)fdiv((unsigned)b, (unsigned)p.L, p.magL);
        const int j = b - i * p.L;
#pragma unroll
        for (int s = 1; s < R; ++s) {
          reg[q * R + s] = twiddle<DIR>(reg[q * R + s], x.tws[s * j * p.twstep]);
        }
      }
      dft<R, DIR>(&reg[q * R]);
    }
  }
}

template <int R, int EMAX>
__device__ inline void col_store_lds(const float2 (&reg)[EMAX], const FftPass& p, const ColCtx& x)
{
  const int nb = p.m << x.log2T;
#pragma unroll
  for (int q = 0; q < EMAX / R; ++q) {
    const int u = x.tid + q * x.nthr;
    if (u < nb) {
      const int c = u & ((1 << x.log2T) - 1), b = u >> x.log2T;
      const int i = (int)fdiv((unsigned)b, (unsigned)p.L, p.magL);
      const int j = b - i * p.L;
      const int o = i * R * p.L + j;
#pragma unroll
      for (int t = 0; t < R; ++t) x.lds[((o + t * p.L) << x.log2T) + c] = reg[q * R + t];
    }
  }
}

template <int R, int EMAX>
__device__ inline void col_load_lds(float2 (&reg)[EMAX], const FftPass& p, const ColCtx& x)
{
  const int nb = p.m << x.log2T;
#pragma unroll
  for (int q = 0; q < EMAX / R; ++q) {
    const int u = x.tid + q * x.nthr;
    if (u < nb) {
      const int c = u & ((1 << x.log2T) - 1), b = u >> x.log2T;
#pragma unroll
      for (int s = 0; s < R; ++s) reg[q * R + s] = x.lds[((b + s * p.m) << x.log2T) + c];
    }
  }
}


// One workgroup: T columns (y, kz) of the half spectrum after the forward z
// and y passes.  Forward x FFT (plan A.P, radix R_LAST last), then for each
// axis the PM Poisson operator (pm.c:322-346) and the inverse x FFT with the
// radices in reverse order (A.P2, radix R_LAST first): the forward's last
// butterflies are exactly the inverse's first, so delta(k) stays in registers
// and never goes to memory.  Outputs: the three x-inverted force spectra.
template <int R_LAST, int EMAX, int RS>
__device__ __forceinline__ void fused_run(const float2* __restrict__ in, float2* __restrict__ out0,
                                          float2* __restrict__ out1, float2* __restrict__ out2,
                                          const float2* __restrict__ tw, const TileArgs& A, const FftRadices& Pinv,
                                          int blk, int nblk)
{
  extern __shared__ float2 smem[];
  const int n = A.P.n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int log2T = A.log2T;
  const int lb = xcd_group_remap(blk, nblk, A.xgroup);
  if (lb >= A.tpb) return;  // padding workgroup (before any barrier)
  const int c0 = lb << log2T;
  if (padding_tile(A, c0, log2T)) return;
  ColCtx x{smem + n, smem, tid, nthr, log2T};
  for (int k = tid; k < n; k += nthr) smem[k] = tw[k];
  using Tile = float2[EMAX];
  const int np_ = A.P.npass;
  const int tmask = (1 << log2T) - 1;

  // forward pass 0 from global memory
  radix_dispatch<RS>(A.P.pass[0].R, [&](auto Rc) {
    constexpr int R = decltype(Rc)::value;
    const FftPass& p = A.P.pass[0];
    const int nb = p.m << log2T;
    Tile reg;
#pragma unroll
    for (int q = 0; q < EMAX / R; ++q) {
      const int u = tid + q * nthr;
      const int c = u & tmask, b = u >> log2T;
      const bool live = u < nb && c0 + c < A.ncols;
      const float2* base = in + c0;
      const unsigned lane = (unsigned)b * (unsigned)A.sr + (unsigned)c;
#pragma unroll
      for (int s = 0; s < R; ++s)
        reg[q * R + s] = live ? base[(long long)(s * p.m) * A.sr + lane] : make_float2(0.0f, 0.0f);
    }
    col_compute<R, EMAX, -1>(reg, p, x);
    col_store_lds<R, EMAX>(reg, p, x);
  });
  __syncthreads();
  for (int pi = 1; pi < np_ - 1; ++pi) {
    radix_dispatch<RS>(A.P.pass[pi].R, [&](auto Rc) {
      constexpr int R = decltype(Rc)::value;
      Tile reg;
      col_load_lds<R, EMAX>(reg, A.P.pass[pi], x);
      __syncthreads();
      col_compute<R, EMAX, -1>(reg, A.P.pass[pi], x);
      col_store_lds<R, EMAX>(reg, A.P.pass[pi], x);
    });
    __syncthreads();
  }
  // forward last pass: delta(k) of this tile stays in registers
  Tile keep;
  const FftPass& pl = A.P.pass[np_ - 1];
  col_load_lds<R_LAST, EMAX>(keep, pl, x);
  __syncthreads();
  col_compute<R_LAST, EMAX, -1>(keep, pl, x);

  // f1 / k^2 of each kept value (pm.c:339 evaluates (f1 / k^2) * d_axis
  // left to right, so the quotient is shared by the three axes bit for bit)
  float green[EMAX];
#pragma unroll
  for (int q = 0; q < EMAX / R_LAST; ++q) {
    const int u = tid + q * nthr;
    const int c = u & tmask, b = u >> log2T;
    const int col = c0 + c;
    const int iy = (int)fdiv((unsigned)col, (unsigned)A.nz, A.magNz);
    const int J0 = wave_pm(iy, n), K = col - iy * A.nz;
#pragma unroll
    for (int t = 0; t < R_LAST; ++t) {
      const int I0 = wave_pm(b + t * pl.m, n);
      const int k2 = K * K + I0 * I0 + J0 * J0;
      green[q * R_LAST + t] = k2 == 0 ? 0.0f : A.kp.f1 / (float)k2;
    }
  }

  // Two inverse transforms instead of three: along a column (fixed ky, kz)
  // the y and z force spectra are n_y and n_z times the same i f1 / k^2
  // delta (pm.c:336-341 with n_y, n_z constant in x), and the x transform is
  // linear, so transform 1 is that common spectrum and its outputs are
  // written twice, scaled by n_y and by n_z; transform 0 carries n_x.
#pragma unroll 1
  for (int axis = 0; axis < 2; ++axis) {
    Tile reg;
    // inverse pass 0 (L = 1): pm.c:322-346 operator on the forward outputs b + t*m
#pragma unroll
    for (int q = 0; q < EMAX / R_LAST; ++q) {
      const int u = tid + q * nthr;
      const int c = u & tmask, b = u >> log2T;
      const int col = c0 + c;
      const int iy = (int)fdiv((unsigned)col, (unsigned)A.nz, A.magNz);
      const int J0 = wave_pm(iy, n), K = col - iy * A.nz;
#pragma unroll
      for (int t = 0; t < R_LAST; ++t) {
        const int I0 = wave_pm(b + t * pl.m, n);
        const float f2 = axis == 0 ? green[q * R_LAST + t] * (float)I0 : green[q * R_LAST + t];
        const float2 v = keep[q * R_LAST + t];
        reg[q * R_LAST + t] = (I0 == 0 && J0 == 0 && K == 0) ? make_float2(0.0f, 0.0f)
                                                              : make_float2(-f2 * v.y, f2 * v.x);
      }
    }
    col_compute<R_LAST, EMAX, 1>(reg, Pinv.pass[0], x);
    col_store_lds<R_LAST, EMAX>(reg, Pinv.pass[0], x);
    __syncthreads();
    for (int pi = 1; pi < np_; ++pi) {
      const bool last = pi == np_ - 1;
      radix_dispatch<RS>(Pinv.pass[pi].R, [&](auto Rc) {
        constexpr int R = decltype(Rc)::value;
        const FftPass& p = Pinv.pass[pi];
        Tile r2;
        col_load_lds<R, EMAX>(r2, p, x);
        __syncthreads();
        col_compute<R, EMAX, 1>(r2, p, x);
        if (!last) {
          col_store_lds<R, EMAX>(r2, p, x);
        } else {
          // last pass: outputs at b + t*m
          const int nb = p.m << log2T;
#pragma unroll
          for (int q = 0; q < EMAX / R; ++q) {
            const int u = tid + q * nthr;
            const int c = u & tmask, b = u >> log2T;
            if (u < nb && c0 + c < A.ncols) {
              const unsigned lane = (unsigned)b * (unsigned)A.sr + (unsigned)c;
              if (axis == 0) {
#pragma unroll
                for (int t = 0; t < R; ++t) out0[c0 + (long long)(t * p.m) * A.sr + lane] = r2[q * R + t];
              } else {
                const int col = c0 + c;
                const int iy = (int)fdiv((unsigned)col, (unsigned)A.nz, A.magNz);
                const float fy = (float)wave_pm(iy, n), fz = (float)(col - iy * A.nz);
#pragma unroll
                for (int t = 0; t < R; ++t) {
                  const float2 v = r2[q * R + t];
                  const long long at = c0 + (long long)(t * p.m) * A.sr + lane;
                  out1[at] = make_float2(fy * v.x, fy * v.y);
                  out2[at] = make_float2(fz * v.x, fz * v.y);
                }
              }
            }
          }
        }
      });
      if (!last) __syncthreads();
    }
  }
}

template <int R_LAST, int EMAX, int RS>
__global__ void __launch_bounds__(kFftMaxThreads)
fft_x_fused(const float2* __restrict__ in, float2* __restrict__ out0, float2* __restrict__ out1,
            float2* __restrict__ out2, const float2* __restrict__ tw, TileArgs A, FftRadices Pinv)
{
  fused_run<R_LAST, EMAX, RS>(in, out0, out1, out2, tw, A, Pinv, blockIdx.x, gridDim.x);
}

unsigned magic_for(int d)
{
  if (d <= 1) return 0;
  return (unsigned)(((1ULL << 32) + (unsigned long long)d - 1) / (unsigned long long)d);
}

size_t tile_lds_bytes(int mode, int n, int log2T)
{
  const size_t data = mode == kCols ? ((size_t)n << log2T) : ((size_t)(n + 1) << log2T);
  return sizeof(float2) * (data + n);
}

template <int DIR, int MODE, KOp OP, int E, int RS>
int launch_e(dim3 grid, int threads, size_t lds, hipStream_t s, const void* in, void* out, const float2* tw,
             const TileArgs& A)
{
  static bool attr = false;
  if (!attr) {
    COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_tile<DIR, MODE, OP, E, RS>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL((fft_tile<DIR, MODE, OP, E, RS>), grid, dim3(threads), lds, s, in, out, tw, A);
  return kOk;
}

// Every pass of radix 3, 4 or 8.
bool radices_348(const FftRadices& r)
{
  for (int i = 0; i < r.npass; ++i)
    if (r.pass[i].R != 3 && r.pass[i].R != 4 && r.pass[i].R != 8) return false;
  return true;
}

template <int DIR, int MODE, KOp OP>
int launch_tile(const FftPassGeom& g, dim3 grid, hipStream_t s, const void* in, void* out, const float2* tw,
                const TileArgs& A)
{
  const size_t lds = tile_lds_bytes(MODE, A.P.n, g.log2T);
  if (g.emax <= 8) {
    if (radices_348(A.P)) return launch_e<DIR, MODE, OP, 8, kRs348>(grid, g.threads, lds, s, in, out, tw, A);
    return launch_e<DIR, MODE, OP, 8, kRsAll>(grid, g.threads, lds, s, in, out, tw, A);
  }
  if (g.emax <= 12) return launch_e<DIR, MODE, OP, 12, kRsAll>(grid, g.threads, lds, s, in, out, tw, A);
  if (g.emax <= 16) return launch_e<DIR, MODE, OP, 16, kRsAll>(grid, g.threads, lds, s, in, out, tw, A);
  return launch_e<DIR, MODE, OP, 24, kRsAll>(grid, g.threads, lds, s, in, out, tw, A);
}

int grid_x(long long ncols, int log2T) { return (int)((ncols + (1LL << log2T) - 1) >> log2T); }

void set_groups(TileArgs& A, int tpb, int group)
{
  A.tpb = tpb;
  A.xgroup = group;
  A.tpb_pad = (tpb + group - 1) / group * group;
}

// ---- batched passes: one launch over the workgroups of many grids ---------
// A job is one pass of one grid as its own launch runs it: that launch's
// arguments and workgroup count.  A batched launch lists the jobs'
// workgroups back to back (start[j], batch_find); each job starts on a
// multiple of 8 x its XCD group, so its workgroups land on the XCDs its own
// launch would put them on, and the tile arithmetic is the job's own, so the
// outputs are those of the separate launches bit for bit.
struct TileJob {
  const void* in;
  void* out;
  const float2* tw;
  TileArgs A;
  int blocks;
};

struct FusedJob {
  const float2* in;
  float2* out[3];
  const float2* tw;
  TileArgs A;
  FftRadices Pinv;
  int blocks;
};

template <int DIR, int MODE, int EMAX, int RS>
__global__ void __launch_bounds__(kFftMaxThreads)
    __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu(MODE, KOp::None, EMAX, RS))))
fft_tile_jobs(const TileJob* __restrict__ jobs, const int* __restrict__ start, int njobs)
{
  const int j = batch_find(start, njobs, blockIdx.x);
  const TileJob& J = jobs[j];
  const int lb = (int)blockIdx.x - start[j];
  if (lb >= J.blocks) return;  // alignment padding (before any barrier)
  tile_run<DIR, MODE, KOp::None, EMAX, RS>(J.in, J.out, J.tw, J.A, lb, J.blocks);
}

template <int EMAX, int RS>
__global__ void __launch_bounds__(kFftMaxThreads)
fft_x_fused_jobs(const FusedJob* __restrict__ jobs, const int* __restrict__ start, int njobs)
{
  const int j = batch_find(start, njobs, blockIdx.x);
  const FusedJob& J = jobs[j];
  const int lb = (int)blockIdx.x - start[j];
  if (lb >= J.blocks) return;
  // the last forward radix is a template constant of the fused pass
  // (FftPlan::can_fuse admits 3, 4, 5 and 8)
  auto run = [&](auto Rc) {
    fused_run<decltype(Rc)::value, EMAX, RS>(J.in, J.out[0], J.out[1], J.out[2], J.tw, J.A, J.Pinv, lb, J.blocks);
  };
  switch (J.A.P.pass[J.A.P.npass - 1].R) {
    case 3: run(std::integral_constant<int, 3>{}); break;
    case 4: run(std::integral_constant<int, 4>{}); break;
    case 5:
      if constexpr (RS == kRsAll) run(std::integral_constant<int, 5>{});
      break;
    default: run(std::integral_constant<int, 8>{}); break;
  }
}

}  // namespace

bool fft_factor(int n, FftRadices* r, int schedule)
{
  *r = FftRadices{};
  r->n = n;
  if (schedule < 0 || schedule >= kFftSchedules) return false;
  int m = n;
  int twos = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++twos;
  }
  std::vector<int> pow2;  // radices of the power of two
  if (schedule == 2) {
    // radix 4 throughout; an odd power of two takes one radix 8 (or a 2 for n = 2)
    int left = twos;
    if (left % 2 == 1) {
      pow2.push_back(left >= 3 ? 8 : 2);
      left -= left >= 3 ? 3 : 1;
    }
    for (; left > 0; left -= 2) pow2.push_back(4);
  } else {
    // radix 8 as long as it goes; a leftover 2 becomes 4 x 4 instead of 8 x 2
    int n8 = twos / 3, n4 = 0, n2 = 0;
    if (twos % 3 == 2) n4 = 1;
    if (twos % 3 == 1) {
      if (n8 > 0) {
        --n8;
        n4 = 2;
      } else {
        n2 = 1;
      }
    }
    for (int i = 0; i < n8; ++i) pow2.push_back(8);
    for (int i = 0; i < n4; ++i) pow2.push_back(4);
    for (int i = 0; i < n2; ++i) pow2.push_back(2);
    if (schedule == 3) std::reverse(pow2.begin(), pow2.end());
    if (schedule == 4 && pow2.size() >= 3) std::rotate(pow2.begin() + 1, pow2.end() - 1, pow2.end());
  }
  std::vector<int> odd;
  while (m % 3 == 0) {
    odd.push_back(3);
    m /= 3;
  }
  while (m % 5 == 0) {
    odd.push_back(5);
    m /= 5;
  }
  std::vector<int> rad = schedule == 1 ? odd : pow2;
  for (int v : schedule == 1 ? pow2 : odd) rad.push_back(v);
  if (m != 1 || rad.size() > 8 || n < 2) return false;
  r->npass = (int)rad.size();
  int L = 1;
  for (size_t i = 0; i < rad.size(); ++i) {
    FftPass& p = r->pass[i];
    p.R = rad[i];
    p.L = L;
    p.m = n / p.R;
    p.twstep = n / (p.R * L);
    p.magL = magic_for(L);
    p.magM = magic_for(p.m);
    L *= p.R;
  }
  return true;
}

int fft_emax(const FftRadices& r, int log2T, int threads)
{
  const long long T = 1LL << log2T;
  if (r.n * T > kFftMaxTile) return 0;
  int e = 0;
  for (int p = 0; p < r.npass; ++p) {
    const long long nb = (long long)r.pass[p].m * T;
    const int q = (int)((nb + threads - 1) / threads);
    e = std::max(e, q * r.pass[p].R);
  }
  return e <= 24 ? e : 0;
}

// Threads for a tile: the fewest (a multiple of 64, at most
// kFftMaxThreads) that keep every thread within 8 register slots, else within
// the smallest slot count reachable.
FftPassGeom pass_geom(const FftRadices& r, int log2T)
{
  FftPassGeom best;
  best.log2T = log2T;
  const long long elems = (long long)r.n << log2T;
  const int t0 = (int)std::max<long long>(64, std::min<long long>(kFftMaxThreads, ((elems / 8 + 63) / 64) * 64));
  for (int t = t0; t <= kFftMaxThreads; t += 64) {
    const int e = fft_emax(r, log2T, t);
    if (e == 0) continue;
    if (best.emax == 0 || e < best.emax) {
      best.threads = t;
      best.emax = e;
    }
    if (e <= 8) break;
  }
  return best;
}

// log2 of the widest tile with n*T <= 6144 complex values (48 KiB of LDS)
static int cap_log2(int n)
{
  int l = 0;
  while (((long long)n << (l + 1)) <= 6144) ++l;
  return l;
}

int FftPlan::init(int n_, int log2T_cols, int log2T_rows, bool pad_rows, int log2T_xcols, int schedule_)
{
  n = n_;
  schedule = schedule_;
  nzp = pad_rows ? (n / 2 + 1 + 15) / 16 * 16 : n / 2 + 1;
  if (n % 2 != 0 || !fft_factor(n, &rad, schedule)) return kErrUnsupported;
  // widest power-of-two tile with n*T <= 6144 (48 KiB of LDS), no wider than
  // the columns, and a register tile of at most 24 values per thread
  auto pick = [&](int want, long long ncols) -> FftPassGeom {
    int l = want;
    while (l > 0 && (1LL << (l - 1)) >= ncols) --l;
    while (l >= 0) {
      FftPassGeom g = pass_geom(rad, l);
      if (g.emax > 0 && tile_lds_bytes(kCols, n, l) <= 160 * 1024) return g;  // fits one CU's LDS
      --l;
    }
    return FftPassGeom{};
  };
  // columns: 8 adjacent complex values (64 B, 128 B per XCD-paired tiles);
  // rows: 2 row pairs per workgroup (measured fastest on MI355X at n = 768)
  const int xcols = log2T_xcols >= 0 ? log2T_xcols : log2T_cols;
  gx = pick(xcols >= 0 ? xcols : std::min(3, cap_log2(n)), (long long)n * nzp);
  gy = pick(log2T_cols >= 0 ? log2T_cols : std::min(3, cap_log2(n)), n / 2 + 1);
  grows = pick(log2T_rows >= 0 ? log2T_rows : 1, (long long)n * n / 2);
  if (gx.emax == 0 || gy.emax == 0 || grows.emax == 0) return kErrUnsupported;
  std::vector<float2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * kPiD * (double)k / (double)n;
    tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  COLA_HIP(hipMalloc(&d_tw, sizeof(float2) * n));
  COLA_HIP(hipMemcpy(d_tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return kOk;
}

void FftPlan::destroy()
{
  if (d_tw) (void)hipFree(d_tw);
  d_tw = nullptr;
}

int FftPlan::r2c(float* data, hipStream_t s) const
{
  const long long nz = n / 2 + 1, np_ = nzp;
  const int npairs = n * n / 2;
  float2* cdata = reinterpret_cast<float2*>(data);
  auto mark = [&](int pass, int begin) {
    if (hook) hook(hook_user, pass, begin);
  };
  TileArgs A;
  A.P = rad;
  // z: rows
  A.log2T = grows.log2T;
  A.ncols = npairs;
  A.rs_in = 2 * np_;
  A.rs_out = np_;
  A.kz_valid = 0;
  mark(0, 1);
  COLA_TRY((launch_tile<-1, kR2C, KOp::None>(grows, dim3(grid_x(npairs, grows.log2T)), s, data, cdata, d_tw, A)));
  mark(0, 0);
  // y: columns kz of each x plane
  A.log2T = gy.log2T;
  A.ncols = (int)nz;
  A.sr = np_;
  A.bs = (long long)n * np_;
  set_groups(A, grid_x(nz, gy.log2T), xgroup);
  mark(1, 1);
  COLA_TRY((launch_tile<-1, kCols, KOp::None>(gy, dim3(A.tpb_pad * n), s, cdata, cdata, d_tw, A)));
  mark(1, 0);
  // x: columns (y, kz), row padding included (its tiles return at once)
  const long long ncols = (long long)n * np_;
  A.log2T = gx.log2T;
  A.ncols = (int)ncols;
  A.sr = ncols;
  A.bs = 0;
  A.nz = (int)np_;
  A.kz_valid = np_ > nz ? (int)nz : 0;
  set_groups(A, grid_x(ncols, gx.log2T), xgroup);
  mark(2, 1);
  COLA_TRY((launch_tile<-1, kCols, KOp::None>(gx, dim3(A.tpb_pad), s, cdata, cdata, d_tw, A)));
  mark(2, 0);
  COLA_HIP(hipGetLastError());
  return kOk;
}

// The same radices in reverse order (L, m, twiddle stride recomputed).
static FftRadices reversed(const FftRadices& r)
{
  FftRadices v;
  v.n = r.n;
  v.npass = r.npass;
  int L = 1;
  for (int i = 0; i < r.npass; ++i) {
    FftPass& p = v.pass[i];
    p.R = r.pass[r.npass - 1 - i].R;
    p.L = L;
    p.m = r.n / p.R;
    p.twstep = r.n / (p.R * L);
    p.magL = magic_for(L);
    p.magM = magic_for(p.m);
    L *= p.R;
  }
  return v;
}

template <int RL, int E, int RS>
static int launch_fused_e(const FftPassGeom& g, int grid, hipStream_t s, const float2* in, float2* o0, float2* o1,
                          float2* o2, const float2* tw, const TileArgs& A, const FftRadices& Pinv)
{
  static bool attr = false;
  if (!attr) {
    COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_x_fused<RL, E, RS>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL((fft_x_fused<RL, E, RS>), dim3(grid), dim3(g.threads), tile_lds_bytes(kCols, A.P.n, g.log2T), s,
                     in, o0, o1, o2, tw, A, Pinv);
  return kOk;
}

template <int RL>
static int launch_fused_r(const FftPassGeom& g, int grid, hipStream_t s, const float2* in, float2* o0, float2* o1,
                          float2* o2, const float2* tw, const TileArgs& A, const FftRadices& Pinv)
{
  if (g.emax <= 8) {
    if constexpr (RL == 3 || RL == 4 || RL == 8)
      if (radices_348(A.P)) return launch_fused_e<RL, 8, kRs348>(g, grid, s, in, o0, o1, o2, tw, A, Pinv);
    return launch_fused_e<RL, 8, kRsAll>(g, grid, s, in, o0, o1, o2, tw, A, Pinv);
  }
  return launch_fused_e<RL, 12, kRsAll>(g, grid, s, in, o0, o1, o2, tw, A, Pinv);
}

bool FftPlan::can_fuse() const
{
  if (rad.npass < 2 || gx.emax > 12) return false;
  const int RL = rad.pass[rad.npass - 1].R;
  return RL == 3 || RL == 4 || RL == 5 || RL == 8;
}

namespace {

// The launches of FftPlan::forces (fused path) for one grid: z r2c, y
// forward, fused x, then per axis y inverse and z c2r.
struct ForceJobs {
  TileJob zr2c, yfwd, yinv[3], zc2r[3];
  FusedJob x;
};

void make_force_jobs(const FftPlan& P, float* density, float* const out[3], const KOpParams& kp, ForceJobs* J)
{
  const int n = P.n;
  const long long nz = n / 2 + 1, np_ = P.nzp;
  const int npairs = n * n / 2;
  float2* cdata = reinterpret_cast<float2*>(density);
  TileArgs A;
  A.P = P.rad;
  A.kp = kp;
  A.nz = (int)np_;
  A.magNz = magic_for((int)np_);
  A.kz_valid = 0;
  // z and y forward
  A.log2T = P.grows.log2T;
  A.ncols = npairs;
  A.rs_in = 2 * np_;
  A.rs_out = np_;
  J->zr2c = TileJob{density, cdata, P.d_tw, A, grid_x(npairs, P.grows.log2T)};
  A.log2T = P.gy.log2T;
  A.ncols = (int)nz;
  A.sr = np_;
  A.bs = (long long)n * np_;
  set_groups(A, grid_x(nz, P.gy.log2T), P.xgroup);
  J->yfwd = TileJob{cdata, cdata, P.d_tw, A, A.tpb_pad * n};
  // fused x: forward, Poisson per axis, inverse (row padding tiles return at once)
  const long long ncols = (long long)n * np_;
  A.log2T = P.gx.log2T;
  A.ncols = (int)ncols;
  A.sr = ncols;
  A.bs = 0;
  A.kz_valid = np_ > nz ? (int)nz : 0;
  set_groups(A, grid_x(ncols, P.gx.log2T), P.xgroup);
  FusedJob& x = J->x;
  x.in = cdata;
  for (int a = 0; a < 3; ++a) x.out[a] = reinterpret_cast<float2*>(out[a]);
  x.tw = P.d_tw;
  x.A = A;
  x.Pinv = reversed(P.rad);
  x.blocks = A.tpb_pad;
  // per axis: y inverse, z c2r
  for (int a = 0; a < 3; ++a) {
    TileArgs B;
    B.P = P.rad;
    B.kz_valid = 0;
    B.log2T = P.gy.log2T;
    B.ncols = (int)nz;
    B.sr = np_;
    B.bs = (long long)n * np_;
    set_groups(B, grid_x(nz, P.gy.log2T), P.xgroup);
    J->yinv[a] = TileJob{x.out[a], x.out[a], P.d_tw, B, B.tpb_pad * n};
    B.log2T = P.grows.log2T;
    B.ncols = npairs;
    B.rs_in = np_;
    B.rs_out = 2 * np_;
    B.zero_pad = np_ == nz;
    J->zc2r[a] = TileJob{x.out[a], out[a], P.d_tw, B, grid_x(npairs, P.grows.log2T)};
  }
}

template <int DIR, int MODE, int E, int RS>
int launch_jobs_e(const FftBatchStage& st, hipStream_t s)
{
  static bool attr = false;
  if (!attr) {
    COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_tile_jobs<DIR, MODE, E, RS>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL((fft_tile_jobs<DIR, MODE, E, RS>), dim3(st.blocks), dim3(st.threads), st.lds, s,
                     static_cast<const TileJob*>(st.jobs), st.start, st.njobs);
  return kOk;
}

template <int DIR, int MODE>
int launch_jobs(const FftBatchStage& st, hipStream_t s)
{
  if (st.emax <= 8) {
    if (st.rs == kRs348) return launch_jobs_e<DIR, MODE, 8, kRs348>(st, s);
    return launch_jobs_e<DIR, MODE, 8, kRsAll>(st, s);
  }
  if (st.emax <= 12) return launch_jobs_e<DIR, MODE, 12, kRsAll>(st, s);
  if (st.emax <= 16) return launch_jobs_e<DIR, MODE, 16, kRsAll>(st, s);
  return launch_jobs_e<DIR, MODE, 24, kRsAll>(st, s);
}

template <int E, int RS>
int launch_fused_jobs_e(const FftBatchStage& st, hipStream_t s)
{
  static bool attr = false;
  if (!attr) {
    COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_x_fused_jobs<E, RS>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL((fft_x_fused_jobs<E, RS>), dim3(st.blocks), dim3(st.threads), st.lds, s,
                     static_cast<const FusedJob*>(st.jobs), st.start, st.njobs);
  return kOk;
}

}  // namespace

int FftPlan::forces(float* density, float* const out[3], const KOpParams& kp, hipStream_t s) const
{
  if (!can_fuse()) {
    COLA_TRY(r2c(density, s));
    KOpParams k = kp;
    for (int a = 0; a < 3; ++a) {
      k.axis = a;
      COLA_TRY(c2r(reinterpret_cast<const float2*>(density), out[a], KOp::PmPoisson, k, s));
    }
    return kOk;
  }
  auto mark = [&](int pass, int begin) {
    if (hook) hook(hook_user, pass, begin);
  };
  ForceJobs J;
  make_force_jobs(*this, density, out, kp, &J);
  auto tile = [&](auto dir, auto mode, const FftPassGeom& g, const TileJob& t) {
    return launch_tile<decltype(dir)::value, decltype(mode)::value, KOp::None>(g, dim3(t.blocks), s, t.in, t.out,
                                                                              t.tw, t.A);
  };
  using Fwd = std::integral_constant<int, -1>;
  using Inv = std::integral_constant<int, 1>;
  mark(0, 1);
  COLA_TRY(tile(Fwd{}, std::integral_constant<int, kR2C>{}, grows, J.zr2c));
  mark(0, 0);
  mark(1, 1);
  COLA_TRY(tile(Fwd{}, std::integral_constant<int, kCols>{}, gy, J.yfwd));
  mark(1, 0);
  const FusedJob& x = J.x;
  mark(6, 1);
  switch (rad.pass[rad.npass - 1].R) {
    case 3: COLA_TRY(launch_fused_r<3>(gx, x.blocks, s, x.in, x.out[0], x.out[1], x.out[2], d_tw, x.A, x.Pinv)); break;
    case 4: COLA_TRY(launch_fused_r<4>(gx, x.blocks, s, x.in, x.out[0], x.out[1], x.out[2], d_tw, x.A, x.Pinv)); break;
    case 5: COLA_TRY(launch_fused_r<5>(gx, x.blocks, s, x.in, x.out[0], x.out[1], x.out[2], d_tw, x.A, x.Pinv)); break;
    default: COLA_TRY(launch_fused_r<8>(gx, x.blocks, s, x.in, x.out[0], x.out[1], x.out[2], d_tw, x.A, x.Pinv)); break;
  }
  mark(6, 0);
  for (int a = 0; a < 3; ++a) {
    mark(4, 1);
    COLA_TRY(tile(Inv{}, std::integral_constant<int, kCols>{}, gy, J.yinv[a]));
    mark(4, 0);
    mark(5, 1);
    COLA_TRY(tile(Inv{}, std::integral_constant<int, kC2R>{}, grows, J.zc2r[a]));
    mark(5, 0);
  }
  COLA_HIP(hipGetLastError());
  return kOk;
}

int FftForcesBatch::init(const FftPlan* const* plans, float* const* density, float* const* out, const KOpParams* kp,
                         int n, bool merge_all)
{
  destroy();
  if (n <= 0 || !plans || !density || !out || !kp) return kErrArg;
  std::vector<ForceJobs> fj(n);
  for (int i = 0; i < n; ++i) {
    if (!plans[i] || !plans[i]->can_fuse()) {
      set_error("batched force transforms need plans whose x pass fuses");
      return kErrUnsupported;
    }
    make_force_jobs(*plans[i], density[i], out + 3 * i, kp[i], &fj[i]);
  }
  // the jobs of the five passes and each job's own launch geometry
  struct Pass {
    int mode;
    std::vector<TileJob> tj;
    std::vector<FusedJob> xj;
    std::vector<const FftPassGeom*> g;
  };
  Pass ps[5] = {{kR2C, {}, {}, {}}, {kCols, {}, {}, {}}, {kCols, {}, {}, {}}, {kCols, {}, {}, {}}, {kC2R, {}, {}, {}}};
  for (int i = 0; i < n; ++i) {
    const FftPlan& P = *plans[i];
    ps[0].tj.push_back(fj[i].zr2c);
    ps[0].g.push_back(&P.grows);
    ps[1].tj.push_back(fj[i].yfwd);
    ps[1].g.push_back(&P.gy);
    ps[2].xj.push_back(fj[i].x);
    ps[2].g.push_back(&P.gx);
    for (int a = 0; a < 3; ++a) {
      ps[3].tj.push_back(fj[i].yinv[a]);
      ps[3].g.push_back(&P.gy);
      ps[4].tj.push_back(fj[i].zc2r[a]);
      ps[4].g.push_back(&P.grows);
    }
  }
  auto align_up = [](size_t v, size_t a) { return (v + a - 1) / a * a; };
  auto round_e = [](int e) { return e <= 8 ? 8 : e <= 12 ? 12 : e <= 16 ? 16 : 24; };
  std::vector<char> host;
  std::vector<std::pair<size_t, size_t>> offs;  // job table, start table
  // One stage: the jobs idx of pass k in one launch.
  auto build = [&](int k, const std::vector<int>& idx) -> int {
    Pass& p = ps[k];
    const bool fused = k == 2;
    const int nj = (int)idx.size();
    auto args = [&](int j) -> const TileArgs& { return fused ? p.xj[idx[j]].A : p.tj[idx[j]].A; };
    auto blocks = [&](int j) { return fused ? p.xj[idx[j]].blocks : p.tj[idx[j]].blocks; };
    FftBatchStage st;
    st.pass = k;
    st.njobs = nj;
    for (int j = 0; j < nj; ++j) st.threads = std::max(st.threads, p.g[idx[j]]->threads);
    int emax = 0;
    bool all348 = true;
    for (int j = 0; j < nj; ++j) {
      const TileArgs& A = args(j);
      const int e = fft_emax(A.P, A.log2T, st.threads);
      if (e == 0) {
        set_error("batched force transforms: a tile does not fit the common workgroup");
        return kErrUnsupported;
      }
      emax = std::max(emax, e);
      all348 = all348 && radices_348(A.P);
      st.lds = std::max(st.lds, tile_lds_bytes(p.mode, A.P.n, A.log2T));
    }
    st.emax = round_e(emax);
    if (fused && st.emax > 12) {
      set_error("batched force transforms: fused x tile wider than 12 values per thread");
      return kErrUnsupported;
    }
    st.rs = st.emax == 8 && all348 ? kRs348 : kRsAll;
    std::vector<int> start(nj + 1);
    long long at = 0;
    for (int j = 0; j < nj; ++j) {
      const long long al = p.mode == kCols ? 8LL * std::max(1, args(j).xgroup) : 1LL;
      at = (at + al - 1) / al * al;
      start[j] = (int)at;
      at += blocks(j);
      if (at > 0x7fffffffLL || at * st.threads > 0xffffffffLL) {
        set_error("batched force transforms: too many workgroups for one launch");
        return kErrUnsupported;
      }
    }
    start[nj] = (int)at;
    st.blocks = (int)at;
    const size_t jo = align_up(host.size(), 256);
    const size_t jb = fused ? sizeof(FusedJob) * nj : sizeof(TileJob) * nj;
    const size_t so = align_up(jo + jb, 256);
    host.resize(so + sizeof(int) * (nj + 1));
    for (int j = 0; j < nj; ++j) {
      if (fused)
        std::memcpy(host.data() + jo + sizeof(FusedJob) * j, &p.xj[idx[j]], sizeof(FusedJob));
      else
        std::memcpy(host.data() + jo + sizeof(TileJob) * j, &p.tj[idx[j]], sizeof(TileJob));
    }
    std::memcpy(host.data() + so, start.data(), sizeof(int) * (nj + 1));
    offs.push_back({jo, so});
    stages.push_back(st);
    return kOk;
  };
  for (int k = 0; k < 5; ++k) {
    Pass& p = ps[k];
    const int nj = k == 2 ? (int)p.xj.size() : (int)p.tj.size();
    if (merge_all) {
      std::vector<int> all(nj);
      for (int j = 0; j < nj; ++j) all[j] = j;
      COLA_TRY(build(k, all));
      continue;
    }
    // jobs grouped by the kernel their own launch runs (register tile and
    // radix set), first appearance first
    std::vector<std::pair<int, std::vector<int>>> groups;
    for (int j = 0; j < nj; ++j) {
      const TileArgs& A = k == 2 ? p.xj[j].A : p.tj[j].A;
      const int e = round_e(p.g[j]->emax);
      const int key = 2 * e + (e == 8 && radices_348(A.P) ? 1 : 0);
      auto it = std::find_if(groups.begin(), groups.end(), [&](const auto& g) { return g.first == key; });
      if (it == groups.end()) {
        groups.push_back({key, {}});
        it = groups.end() - 1;
      }
      it->second.push_back(j);
    }
    for (const auto& g : groups) COLA_TRY(build(k, g.second));
  }
  COLA_HIP(hipMalloc(&d_mem, host.size()));
  COLA_HIP(hipMemcpy(d_mem, host.data(), host.size(), hipMemcpyHostToDevice));
  char* base = static_cast<char*>(d_mem);
  for (size_t k = 0; k < stages.size(); ++k) {
    stages[k].jobs = base + offs[k].first;
    stages[k].start = reinterpret_cast<const int*>(base + offs[k].second);
  }
  return kOk;
}

int FftForcesBatch::run(hipStream_t s) const
{
  for (const FftBatchStage& st : stages) {
    switch (st.pass) {
      case 0: COLA_TRY((launch_jobs<-1, kR2C>(st, s))); break;
      case 1: COLA_TRY((launch_jobs<-1, kCols>(st, s))); break;
      case 2:
        if (st.emax <= 8 && st.rs == kRs348) {
          COLA_TRY((launch_fused_jobs_e<8, kRs348>(st, s)));
        } else if (st.emax <= 8) {
          COLA_TRY((launch_fused_jobs_e<8, kRsAll>(st, s)));
        } else {
          COLA_TRY((launch_fused_jobs_e<12, kRsAll>(st, s)));
        }
        break;
      case 3: COLA_TRY((launch_jobs<1, kCols>(st, s))); break;
      default: COLA_TRY((launch_jobs<1, kC2R>(st, s))); break;
    }
  }
  COLA_HIP(hipGetLastError());
  return kOk;
}

void FftForcesBatch::destroy()
{
  if (d_mem) (void)hipFree(d_mem);
  d_mem = nullptr;
  stages.clear();
}

int FftPlan::c2r(const float2* in_k, float* out, KOp op, const KOpParams& kp, hipStream_t s) const
{
  const long long nz = n / 2 + 1, np_ = nzp;
  const int npairs = n * n / 2;
  float2* cout = reinterpret_cast<float2*>(out);
  auto mark = [&](int pass, int begin) {
    if (hook) hook(hook_user, pass, begin);
  };
  TileArgs A;
  A.P = rad;
  A.kp = kp;
  A.nz = (int)np_;
  A.magNz = magic_for((int)np_);
  // x: columns (y, kz), k-space operator applied on load
  const long long ncols = (long long)n * np_;
  A.log2T = gx.log2T;
  A.ncols = (int)ncols;
  A.sr = ncols;
  A.bs = 0;
  A.kz_valid = np_ > nz ? (int)nz : 0;
  set_groups(A, grid_x(ncols, gx.log2T), xgroup);
  const float2* src = in_k ? in_k : cout;
  const dim3 grid(A.tpb_pad);
  mark(3, 1);
  switch (op) {
    case KOp::PmPoisson:
      COLA_TRY((launch_tile<1, kCols, KOp::PmPoisson>(gx, grid, s, src, cout, d_tw, A)));
      break;
    case KOp::LptGrad:
      COLA_TRY((launch_tile<1, kCols, KOp::LptGrad>(gx, grid, s, src, cout, d_tw, A)));
      break;
    case KOp::Lpt2:
      COLA_TRY((launch_tile<1, kCols, KOp::Lpt2>(gx, grid, s, src, cout, d_tw, A)));
      break;
    default:
      COLA_TRY((launch_tile<1, kCols, KOp::None>(gx, grid, s, src, cout, d_tw, A)));
      break;
  }
  mark(3, 0);
  // y
  A.log2T = gy.log2T;
  A.ncols = (int)nz;
  A.sr = np_;
  A.bs = (long long)n * np_;
  A.kz_valid = 0;
  set_groups(A, grid_x(nz, gy.log2T), xgroup);
  mark(4, 1);
  COLA_TRY((launch_tile<1, kCols, KOp::None>(gy, dim3(A.tpb_pad * n), s, cout, cout, d_tw, A)));
  mark(4, 0);
  // z: rows
  A.log2T = grows.log2T;
  A.ncols = npairs;
  A.rs_in = np_;
  A.rs_out = 2 * np_;
  A.zero_pad = np_ == nz;
  mark(5, 1);
  COLA_TRY((launch_tile<1, kC2R, KOp::None>(grows, dim3(grid_x(npairs, grows.log2T)), s, cout, out, d_tw, A)));
  mark(5, 0);
  COLA_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cola
