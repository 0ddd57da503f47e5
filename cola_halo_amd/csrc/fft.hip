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

#include <cmath>
#include <vector>

namespace cola {

namespace {

constexpr float kS3 = 0.866025403784438646763723170752936183f;   // sin(2pi/3)
constexpr float kC5a = 0.309016994374947424102293417182819059f;  // cos(2pi/5)
constexpr float kC5b = -0.809016994374947424102293417182819059f; // cos(4pi/5)
constexpr float kS5a = 0.951056516295153572116439333379382143f;  // sin(2pi/5)
constexpr float kS5b = 0.587785252292473129168705954639072769f;  // sin(4pi/5)
constexpr float kR2 = 0.707106781186547524400844362104849039f;   // sqrt(2)/2

// multiply by -i (DIR < 0) or +i (DIR > 0)
template <int DIR>
__device__ inline float2 rot(float2 z)
{
  return DIR < 0 ? make_float2(z.y, -z.x) : make_float2(-z.y, z.x);
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
  const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
  const float2 r = rot<DIR>(make_float2(kS3 * d.x, kS3 * d.y));
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
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
  const float2 a1 = make_float2(x0.x + kC5a * t1.x + kC5b * t2.x, x0.y + kC5a * t1.y + kC5b * t2.y);
  const float2 a2 = make_float2(x0.x + kC5b * t1.x + kC5a * t2.x, x0.y + kC5b * t1.y + kC5a * t2.y);
  const float2 b1 = rot<DIR>(make_float2(kS5a * t3.x + kS5b * t4.x, kS5a * t3.y + kS5b * t4.y));
  const float2 b2 = rot<DIR>(make_float2(kS5b * t3.x - kS5a * t4.x, kS5b * t3.y - kS5a * t4.y));
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
  // o[k] *= exp(DIR * 2 pi i k / 8)
  const float2 o1 = o[1], o3 = o[3];
  if (DIR < 0) {
    o[1] = make_float2(kR2 * (o1.x + o1.y), kR2 * (o1.y - o1.x));
    o[3] = make_float2(kR2 * (o3.y - o3.x), -kR2 * (o3.x + o3.y));
  } else {
    o[1] = make_float2(kR2 * (o1.x - o1.y), kR2 * (o1.x + o1.y));
    o[3] = make_float2(-kR2 * (o3.x + o3.y), kR2 * (o3.x - o3.y));
  }
  o[2] = rot<DIR>(o[2]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
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

// One radix-R Stockham pass over the LDS tile [n][T] (column fastest).
template <int R, int DIR>
__device__ inline void radix_pass(float2* lds, float2 (&reg)[kFftEmax], int n, int L, int log2T,
                                  const float2* __restrict__ tw, int tid)
{
  constexpr int Q = kFftEmax / R;
  const int m = n / R;
  const int nb = m << log2T;
  const int tmask = (1 << log2T) - 1;
  const int twstep = n / (R * L);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int u = tid + q * kFftThreads;
    if (u < nb) {
      const int c = u & tmask, b = u >> log2T;
#pragma unroll
      for (int s = 0; s < R; ++s) reg[q * R + s] = lds[((b + s * m) << log2T) + c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int u = tid + q * kFftThreads;
    if (u < nb) {
      const int c = u & tmask, b = u >> log2T;
      const int j = b % L, i = b / L;
      float2 v[R];
#pragma unroll
      for (int s = 0; s < R; ++s) v[s] = reg[q * R + s];
      if (L > 1) {
#pragma unroll
        for (int s = 1; s < R; ++s) {
          float2 w = tw[s * j * twstep];
          if (DIR > 0) w.y = -w.y;
          v[s] = cmul(v[s], w);
        }
      }
      dft<R, DIR>(v);
      const int o = i * R * L + j;
#pragma unroll
      for (int t = 0; t < R; ++t) lds[((o + t * L) << log2T) + c] = v[t];
    }
  }
  __syncthreads();
}

template <int DIR>
__device__ inline void run_passes(float2* lds, const FftRadices& P, const float2* tw, int log2T, int tid)
{
  float2 reg[kFftEmax];
  int L = 1;
  for (int p = 0; p < P.npass; ++p) {
    switch (P.radix[p]) {
      case 2: radix_pass<2, DIR>(lds, reg, P.n, L, log2T, tw, tid); break;
      case 3: radix_pass<3, DIR>(lds, reg, P.n, L, log2T, tw, tid); break;
      case 4: radix_pass<4, DIR>(lds, reg, P.n, L, log2T, tw, tid); break;
      case 5: radix_pass<5, DIR>(lds, reg, P.n, L, log2T, tw, tid); break;
      default: radix_pass<8, DIR>(lds, reg, P.n, L, log2T, tw, tid); break;
    }
    L *= P.radix[p];
  }
}

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

// Strided complex pass: element (r, c) of tile (blockIdx.x, batch blockIdx.y)
// lives at batch*bs + r*sr + (c0 + c); columns >= ncols are padding.
template <int DIR, KOp OP>
__global__ void __launch_bounds__(kFftThreads)
fft_cols_kernel(const float2* __restrict__ in, float2* __restrict__ out, FftRadices P,
                const float2* __restrict__ tw, long long sr, long long bs, int ncols, int log2T,
                int nz, KOpParams kp)
{
  extern __shared__ float2 lds[];
  const int n = P.n, tid = threadIdx.x;
  const int T = 1 << log2T;
  const int c = tid & (T - 1);
  const int col = (blockIdx.x << log2T) + c;
  const bool valid = col < ncols;
  const long long base = (long long)blockIdx.y * bs + col;
  const int rstep = kFftThreads >> log2T;
  // the x pass addresses columns as (y, kz) pairs of the half-spectrum
  int iy = 0, iz = 0;
  if constexpr (OP != KOp::None) {
    iy = col / nz;
    iz = col - iy * nz;
  }
  for (int r = tid >> log2T; r < n; r += rstep) {
    float2 v = make_float2(0.0f, 0.0f);
    if (valid) {
      v = in[base + r * sr];
      if constexpr (OP != KOp::None) v = apply_kop<OP>(v, r, iy, iz, n, kp);
    }
    lds[(r << log2T) + c] = v;
  }
  __syncthreads();
  run_passes<DIR>(lds, P, tw, log2T, tid);
  if (valid)
    for (int r = tid >> log2T; r < n; r += rstep) out[base + r * sr] = lds[(r << log2T) + c];
}

// z pass, forward: real rows 2q, 2q+1 (stride rs_in floats) packed as one
// complex transform, unpacked into half spectra (stride rs_out complex).
__global__ void __launch_bounds__(kFftThreads)
fft_rows_r2c_kernel(const float* __restrict__ in, float2* __restrict__ out, FftRadices P,
                    const float2* __restrict__ tw, long long rs_in, long long rs_out, int npairs, int log2T)
{
  extern __shared__ float2 lds[];
  const int n = P.n, tid = threadIdx.x;
  const int T = 1 << log2T;
  const int c = tid & (T - 1);
  const int q = (blockIdx.x << log2T) + c;
  const bool valid = q < npairs;
  const int rstep = kFftThreads >> log2T;
  const float* ra = in + (2LL * q) * rs_in;
  const float* rb = ra + rs_in;
  for (int r = tid >> log2T; r < n; r += rstep) {
    float2 v = make_float2(0.0f, 0.0f);
    if (valid) v = make_float2(ra[r], rb[r]);
    lds[(r << log2T) + c] = v;
  }
  __syncthreads();
  run_passes<-1>(lds, P, tw, log2T, tid);
  if (valid) {
    float2* oa = out + (2LL * q) * rs_out;
    float2* ob = oa + rs_out;
    for (int k = tid >> log2T; k <= n / 2; k += rstep) {
      const float2 z = lds[(k << log2T) + c];
      const int km = k == 0 ? 0 : n - k;
      const float2 zm = lds[(km << log2T) + c];
      // A = (Z_k + conj Z_{n-k}) / 2,  B = (Z_k - conj Z_{n-k}) / (2i)
      oa[k] = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
      ob[k] = make_float2(0.5f * (z.y + zm.y), 0.5f * (zm.x - z.x));
    }
  }
}

// z pass, inverse: half spectra of rows 2q, 2q+1 extended by Hermitian
// symmetry (imaginary parts of kz = 0 and n/2 dropped), Z = A + iB, one
// complex inverse transform, real part -> row 2q, imaginary part -> row 2q+1.
__global__ void __launch_bounds__(kFftThreads)
fft_rows_c2r_kernel(const float2* __restrict__ in, float* __restrict__ out, FftRadices P,
                    const float2* __restrict__ tw, long long rs_in, long long rs_out, int npairs, int log2T)
{
  extern __shared__ float2 lds[];
  const int n = P.n, tid = threadIdx.x, h = n / 2;
  const int T = 1 << log2T;
  const int c = tid & (T - 1);
  const int q = (blockIdx.x << log2T) + c;
  const bool valid = q < npairs;
  const int rstep = kFftThreads >> log2T;
  const float2* ia = in + (2LL * q) * rs_in;
  const float2* ib = ia + rs_in;
  for (int r = tid >> log2T; r < n; r += rstep) {
    float2 v = make_float2(0.0f, 0.0f);
    if (valid) {
      float2 a, b;
      if (r <= h) {
        a = ia[r];
        b = ib[r];
        if (r == 0 || r == h) {
          a.y = 0.0f;
          b.y = 0.0f;
        }
      } else {
        a = ia[n - r];
        b = ib[n - r];
        a.y = -a.y;
        b.y = -b.y;
      }
      v = make_float2(a.x - b.y, a.y + b.x);
    }
    lds[(r << log2T) + c] = v;
  }
  __syncthreads();
  run_passes<1>(lds, P, tw, log2T, tid);
  if (valid) {
    float* oa = out + (2LL * q) * rs_out;
    float* ob = oa + rs_out;
    for (int r = tid >> log2T; r < n; r += rstep) {
      const float2 z = lds[(r << log2T) + c];
      oa[r] = z.x;
      ob[r] = z.y;
    }
    // the two padding floats of each real row are written as zero
    if ((tid >> log2T) == 0) {
      oa[n] = 0.0f;
      oa[n + 1] = 0.0f;
      ob[n] = 0.0f;
      ob[n + 1] = 0.0f;
    }
  }
}

template <class K>
int set_lds_limit(K kernel)
{
  COLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               kFftMaxTile * (int)sizeof(float2)));
  return kOk;
}

int ensure_attributes()
{
  static int done = 0;
  if (done) return kOk;
  COLA_TRY(set_lds_limit(fft_cols_kernel<-1, KOp::None>));
  COLA_TRY(set_lds_limit(fft_cols_kernel<1, KOp::None>));
  COLA_TRY(set_lds_limit(fft_cols_kernel<1, KOp::PmPoisson>));
  COLA_TRY(set_lds_limit(fft_cols_kernel<1, KOp::LptGrad>));
  COLA_TRY(set_lds_limit(fft_cols_kernel<1, KOp::Lpt2>));
  COLA_TRY(set_lds_limit(fft_rows_r2c_kernel));
  COLA_TRY(set_lds_limit(fft_rows_c2r_kernel));
  done = 1;
  return kOk;
}

int grid_x(long long ncols, int log2T) { return (int)((ncols + (1LL << log2T) - 1) >> log2T); }

}  // namespace

bool fft_factor(int n, FftRadices* r)
{
  *r = FftRadices{};
  r->n = n;
  int m = n;
  std::vector<int> rad;
  while (m % 8 == 0) {
    rad.push_back(8);
    m /= 8;
  }
  while (m % 4 == 0) {
    rad.push_back(4);
    m /= 4;
  }
  while (m % 2 == 0) {
    rad.push_back(2);
    m /= 2;
  }
  while (m % 3 == 0) {
    rad.push_back(3);
    m /= 3;
  }
  while (m % 5 == 0) {
    rad.push_back(5);
    m /= 5;
  }
  if (m != 1 || rad.size() > 8 || n < 2) return false;
  r->npass = (int)rad.size();
  for (size_t i = 0; i < rad.size(); ++i) r->radix[i] = rad[i];
  return true;
}

int fft_tile_log2(const FftRadices& r, int cap_log2)
{
  for (int l = cap_log2; l >= 0; --l) {
    const long long T = 1LL << l;
    if (r.n * T > kFftMaxTile) continue;
    bool ok = true;
    for (int p = 0; p < r.npass; ++p) {
      const int R = r.radix[p];
      const long long nb = (r.n / R) * T;
      if ((nb + kFftThreads - 1) / kFftThreads > kFftEmax / R) ok = false;
    }
    if (ok) return l;
  }
  return -1;
}

int FftPlan::init(int n_, int force_log2T)
{
  n = n_;
  if (n % 2 != 0 || !fft_factor(n, &rad)) return kErrUnsupported;
  const int base = fft_tile_log2(rad, 5);
  if (base < 0) return kErrUnsupported;
  log2T_rows = log2T_y = log2T_x = base;
  if (force_log2T >= 0) {
    const int l = fft_tile_log2(rad, force_log2T);
    if (l < 0) return kErrUnsupported;
    log2T_rows = log2T_y = log2T_x = l;
  }
  // no wider tile than there are columns
  auto fit = [](int l, long long ncols) {
    while (l > 0 && (1LL << (l - 1)) >= ncols) --l;
    return l;
  };
  log2T_y = fit(log2T_y, n / 2 + 1);
  log2T_rows = fit(log2T_rows, (long long)n * n / 2);
  std::vector<float2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * kPiD * (double)k / (double)n;
    tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  COLA_HIP(hipMalloc(&d_tw, sizeof(float2) * n));
  COLA_HIP(hipMemcpy(d_tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return ensure_attributes();
}

void FftPlan::destroy()
{
  if (d_tw) (void)hipFree(d_tw);
  d_tw = nullptr;
}

int FftPlan::r2c(float* data, hipStream_t s) const
{
  const long long nz = n / 2 + 1;
  const int npairs = n * n / 2;
  float2* cdata = reinterpret_cast<float2*>(data);
  const dim3 blk(kFftThreads);
  auto mark = [&](int pass, int begin) {
    if (hook) hook(hook_user, pass, begin);
  };
  // z: rows
  {
    const size_t lds = sizeof(float2) * ((size_t)n << log2T_rows);
    mark(0, 1);
    hipLaunchKernelGGL(fft_rows_r2c_kernel, dim3(grid_x(npairs, log2T_rows)), blk, lds, s, data, cdata, rad, d_tw,
                       2 * nz, nz, npairs, log2T_rows);
    mark(0, 0);
  }
  // y: columns kz of each x plane
  {
    const size_t lds = sizeof(float2) * ((size_t)n << log2T_y);
    mark(1, 1);
    hipLaunchKernelGGL((fft_cols_kernel<-1, KOp::None>), dim3(grid_x(nz, log2T_y), n), blk, lds, s, cdata, cdata,
                       rad, d_tw, nz, (long long)n * nz, (int)nz, log2T_y, (int)nz, KOpParams{});
    mark(1, 0);
  }
  // x: columns (y, kz)
  {
    const size_t lds = sizeof(float2) * ((size_t)n << log2T_x);
    const long long ncols = (long long)n * nz;
    mark(2, 1);
    hipLaunchKernelGGL((fft_cols_kernel<-1, KOp::None>), dim3(grid_x(ncols, log2T_x), 1), blk, lds, s, cdata, cdata,
                       rad, d_tw, ncols, 0LL, (int)ncols, log2T_x, (int)nz, KOpParams{});
    mark(2, 0);
  }
  COLA_HIP(hipGetLastError());
  return kOk;
}

int FftPlan::c2r(const float2* in_k, float* out, KOp op, const KOpParams& kp, hipStream_t s) const
{
  const long long nz = n / 2 + 1;
  const int npairs = n * n / 2;
  float2* cout = reinterpret_cast<float2*>(out);
  const dim3 blk(kFftThreads);
  auto mark = [&](int pass, int begin) {
    if (hook) hook(hook_user, pass, begin);
  };
  {
    const size_t lds = sizeof(float2) * ((size_t)n << log2T_x);
    const long long ncols = (long long)n * nz;
    const dim3 grd(grid_x(ncols, log2T_x), 1);
    const float2* src = in_k ? in_k : cout;
    mark(3, 1);
    switch (op) {
      case KOp::PmPoisson:
        hipLaunchKernelGGL((fft_cols_kernel<1, KOp::PmPoisson>), grd, blk, lds, s, src, cout, rad, d_tw, ncols, 0LL,
                           (int)ncols, log2T_x, (int)nz, kp);
        break;
      case KOp::LptGrad:
        hipLaunchKernelGGL((fft_cols_kernel<1, KOp::LptGrad>), grd, blk, lds, s, src, cout, rad, d_tw, ncols, 0LL,
                           (int)ncols, log2T_x, (int)nz, kp);
        break;
      case KOp::Lpt2:
        hipLaunchKernelGGL((fft_cols_kernel<1, KOp::Lpt2>), grd, blk, lds, s, src, cout, rad, d_tw, ncols, 0LL,
                           (int)ncols, log2T_x, (int)nz, kp);
        break;
      default:
        hipLaunchKernelGGL((fft_cols_kernel<1, KOp::None>), grd, blk, lds, s, src, cout, rad, d_tw, ncols, 0LL,
                           (int)ncols, log2T_x, (int)nz, kp);
        break;
    }
    mark(3, 0);
  }
  {
    const size_t lds = sizeof(float2) * ((size_t)n << log2T_y);
    mark(4, 1);
    hipLaunchKernelGGL((fft_cols_kernel<1, KOp::None>), dim3(grid_x(nz, log2T_y), n), blk, lds, s, cout, cout, rad,
                       d_tw, nz, (long long)n * nz, (int)nz, log2T_y, (int)nz, KOpParams{});
    mark(4, 0);
  }
  {
    const size_t lds = sizeof(float2) * ((size_t)n << log2T_rows);
    mark(5, 1);
    hipLaunchKernelGGL(fft_rows_c2r_kernel, dim3(grid_x(npairs, log2T_rows)), blk, lds, s, cout, out, rad, d_tw, nz,
                       2 * nz, npairs, log2T_rows);
    mark(5, 0);
  }
  COLA_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cola
