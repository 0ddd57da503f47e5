// 2LPT initial conditions.
//
// White noise: the reference reseeds GSL's ranlxd1 generator for every
// (i, j) column from a seed table and draws (phase, amplitude) pairs along k
// (lpt.c:219-351).  The k = 0 plane receives, besides each column's own
// mode, the complex conjugate written by the mirrored column (N-i, N-j).  On
// the GPU each thread owns one column and writes every element of it; for a
// k = 0 element owned by the mirror column the thread regenerates that
// column's first draw pair from its seed, so all writes are column-local and
// the three spectra need no clearing pass.
#include "common.h"
#include "lpt.h"
#include "ranlux.h"

#include <cmath>

namespace cola {

namespace {

constexpr int kThreads = 256;

// RANLUX, 48-bit subtract-with-borrow x_n = x_{n-5} - x_{n-12} - c (mod 1),
// luxury 202: of each block of 202 iterates the last 12 are delivered
// (M. Luescher 1994; GSL's ranlxd1).  The window w[0..11] holds the 12 most
// recent iterates, oldest first, so every index below is static.
struct Ranlux48 {
  static constexpr double kOneBit = 1.0 / 281474976710656.0;  // 2^-48
  double w[12];
  double carry;
  int pos;  // next delivered entry of w; 12 = block exhausted

  __host__ __device__ void seed(unsigned long long s)
  {
    if (s == 0) s = 1;
    // 31 seed bits in a (31, 18) GF(2) shift register; every output bit is
    // the complement of the tapped bit.
    unsigned int reg = (unsigned int)(s & 0x7fffffffULL);
    int ib = 0, jb = 18;
    for (int k = 0; k < 12; ++k) {
      double x = 0.0;
      for (int l = 0; l < 48; ++l) {
        const unsigned bi = (reg >> ib) & 1u, bj = (reg >> jb) & 1u;
        x += x + (double)(bi ^ 1u);
        reg ^= bj << ib;
        ib = ib == 30 ? 0 : ib + 1;
        jb = jb == 30 ? 0 : jb + 1;
      }
      w[k] = kOneBit * x;
    }
    carry = 0.0;
    pos = 12;
  }

  __host__ __device__ inline double step(double a, double b, double& c)
  {
    double y = a - b - c;
    if (y < 0.0) {
      c = kOneBit;
      y += 1.0;
    } else {
      c = 0.0;
    }
    return y;
  }

  __host__ __device__ void block()
  {
    double c = carry;
    for (int round = 0; round < 16; ++round) {
      double y[12];
#pragma unroll
      for (int m = 0; m < 12; ++m) y[m] = step(m < 5 ? w[m + 7] : y[m - 5], w[m], c);
#pragma unroll
      for (int m = 0; m < 12; ++m) w[m] = y[m];
    }
    double y[10];
#pragma unroll
    for (int m = 0; m < 10; ++m) y[m] = step(m < 5 ? w[m + 7] : y[m - 5], w[m], c);
    w[0] = w[10];
    w[1] = w[11];
#pragma unroll
    for (int m = 0; m < 10; ++m) w[m + 2] = y[m];
    carry = c;
  }

  __host__ __device__ inline double uniform()
  {
    if (pos == 12) {
      block();
      pos = 0;
    }
    double r = w[0];
#pragma unroll
    for (int m = 1; m < 12; ++m)
      if (pos == m) r = w[m];
    ++pos;
    return r;
  }
};

// power.c:120-151 on the device table
__device__ inline double power_spec(const DevPower& pw, double k)
{
  const double lk = log10(k);
  if (lk < pw.logk[0] || lk > pw.logk[pw.n - 1]) return 0;
  int lo = 0, hi = pw.n - 1;
  while (hi - lo > 1) {
    const int mid = (hi + lo) / 2;
    if (lk < pw.logk[mid])
      hi = mid;
    else
      lo = mid;
  }
  const double u = (lk - pw.logk[lo]) / (pw.logk[hi] - pw.logk[lo]);
  const double ld = (1 - u) * pw.logD[lo] + u * pw.logD[hi];
  return pw.norm * pow(10.0, ld) / (4.0 * kPiD * k * k * k);
}

__device__ inline double kvec_lpt(int i, int n, double box)
{
  if (i < n / 2) return (double)(i * 2) * kPiD / box;
  return (double)(-(n - i) * 2) * kPiD / box;
}

// Mode (i, j, k) from one (phase, ampl) draw: lpt.c:240-286.  Returns false
// for the modes the cubic cut removes.  conj selects the mirrored value
// written for the k = 0 plane (lpt.c:305-310, 338-343).
__device__ inline bool mode_value(int i, int j, int k, int n, double phase, double ampl, const DevPower& pw,
                                  double box, double fac, bool conj, float2 out[3])
{
  double kv[3] = {kvec_lpt(i, n, box), kvec_lpt(j, n, box), kvec_lpt(k, n, box)};
  const double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
  const double kmag = sqrt(kmag2);
  const int half = n / 2;  // Nsample / 2
  for (int a = 0; a < 3; ++a)
    if (fabs(kv[a]) * box / (2 * kPiD) > half) return false;
  double p_of_k = power_spec(pw, kmag);
  p_of_k *= -log(ampl);
  const double delta = fac * sqrt(p_of_k);
  const double s = sin(phase), c = cos(phase);
  for (int a = 0; a < 3; ++a) {
    const double re = -kv[a] / kmag2 * delta * s;
    const double im = conj ? -kv[a] / kmag2 * delta * c : kv[a] / kmag2 * delta * c;
    out[a] = make_float2((float)re, (float)im);
  }
  return true;
}

__device__ inline void draw_pair(Ranlux48& r, double& phase, double& ampl)
{
  phase = r.uniform() * 2 * kPiD;
  do
    ampl = r.uniform();
  while (ampl == 0.0);
}

__global__ void __launch_bounds__(kThreads)
k_mode_fill(int n, const unsigned int* __restrict__ seedtable, DevPower pw, double box, double fac,
            float2* __restrict__ c0, float2* __restrict__ c1, float2* __restrict__ c2)
{
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  const int half = n / 2;
  const long long base = idx * (half + 1);
  float2* cd[3] = {c0, c1, c2};
  const float2 zero = make_float2(0.0f, 0.0f);
  float2 v[3];

  // k = n/2 plane and the Nyquist columns stay zero (lpt.c:235-236)
  for (int a = 0; a < 3; ++a) cd[a][base + half] = zero;
  const bool dead = (i == half || j == half);

  // own column, k >= 0
  Ranlux48 r;
  bool own_k0 = false;
  float2 v0[3] = {zero, zero, zero};
  if (!dead) {
    r.seed(seedtable[idx]);
    for (int k = 0; k < half; ++k) {
      double phase, ampl;
      draw_pair(r, phase, ampl);
      bool ok = !(i == 0 && j == 0 && k == 0) && mode_value(i, j, k, n, phase, ampl, pw, box, fac, false, v);
      if (k > 0) {
        for (int a = 0; a < 3; ++a) cd[a][base + k] = ok ? v[a] : zero;
      } else if (ok && ((i == 0 && j < half) || (i > 0 && i < half))) {
        own_k0 = true;
        for (int a = 0; a < 3; ++a) v0[a] = v[a];
      }
    }
  } else {
    for (int k = 1; k < half; ++k)
      for (int a = 0; a < 3; ++a) cd[a][base + k] = zero;
  }

  // k = 0 element owned by the mirrored column
  if (!own_k0) {
    int pi = -1, pj = -1;
    if (i == 0 && j > half) {
      pi = 0;
      pj = n - j;
    } else if (i > half && j != half) {
      pi = n - i;
      pj = (n - j) % n;
    }
    if (pi >= 0) {
      Ranlux48 rp;
      rp.seed(seedtable[(long long)pi * n + pj]);
      double phase, ampl;
      draw_pair(rp, phase, ampl);
      if (!mode_value(pi, pj, 0, n, phase, ampl, pw, box, fac, true, v0))
        for (int a = 0; a < 3; ++a) v0[a] = zero;
    }
  }
  for (int a = 0; a < 3; ++a) cd[a][base] = v0[a];
}

__global__ void __launch_bounds__(kThreads)
k_lpt_source(int n, const float* __restrict__ d0, const float* __restrict__ d1, const float* __restrict__ d2,
             float* __restrict__ d3, const float* __restrict__ d4, const float* __restrict__ d5)
{
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long total = (long long)n * n * n;
  if (idx >= total) return;
  const long long row = idx / n;
  const int k = (int)(idx - row * n);
  const long long c = row * (2LL * (n / 2 + 1)) + k;
  const float a0 = d0[c], a1 = d1[c], a2 = d2[c], a3 = d3[c], a4 = d4[c], a5 = d5[c];
  d3[c] = a0 * (a3 + a5) + a3 * a5 - a1 * a1 - a2 * a2 - a4 * a4;
}

struct DispPtrs {
  const float* d[3];
  const float* d2[3];
};

__global__ void __launch_bounds__(kThreads)
k_lpt_particles(int n, DispPtrs g, LptParticleFactors f, DevParticles p)
{
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long total = (long long)n * n * n;
  if (idx >= total) return;
  const int k = (int)(idx % n);
  const int j = (int)((idx / n) % n);
  const int i = (int)(idx / ((long long)n * n));
  const long long c = ((long long)i * n + j) * (2LL * (n / 2 + 1)) + k;
  float q[3];
  q[0] = (0 + i + 0.5f) * f.dx;
  q[1] = (j + 0.5f) * f.dx;
  q[2] = (k + 0.5f) * f.dx;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float dis = g.d[a][c];
    const float dis2 = (float)(f.nmesh3_inv * (double)g.d2[a][c]);
    p.x[a][idx] = (float)((double)(q[a] + dis * f.Dplus) - 3.0 / 7.0 * f.D20 * (double)dis2 * f.D2);
    p.dx1[a][idx] = dis;
    p.dx2[a][idx] = (float)(-3.0 / 7.0 * f.D20 * (double)dis2);
    p.v[a][idx] = 0.0f;
  }
}

}  // namespace

void lpt_seedtable(int N, int seed, unsigned int* t)
{
  Ranlux48 r;
  r.seed((unsigned long long)seed);
  auto U = [&]() { return (unsigned int)(0x7fffffff * r.uniform()); };
  // lpt.c:177-207 spiral fill order
  for (int i = 0; i < N / 2; i++) {
    for (int j = 0; j < i; j++) t[i * N + j] = U();
    for (int j = 0; j < i + 1; j++) t[j * N + i] = U();
    for (int j = 0; j < i; j++) t[(N - 1 - i) * N + j] = U();
    for (int j = 0; j < i + 1; j++) t[(N - 1 - j) * N + i] = U();
    for (int j = 0; j < i; j++) t[i * N + (N - 1 - j)] = U();
    for (int j = 0; j < i + 1; j++) t[j * N + (N - 1 - i)] = U();
    for (int j = 0; j < i; j++) t[(N - 1 - i) * N + (N - 1 - j)] = U();
    for (int j = 0; j < i + 1; j++) t[(N - 1 - j) * N + (N - 1 - i)] = U();
  }
}

struct RanluxStream {
  Ranlux48 r;
};

RanluxStream* ranlux_new(unsigned long long seed)
{
  RanluxStream* s = new RanluxStream;
  s->r.seed(seed);
  return s;
}

double ranlux_next(RanluxStream* s) { return s->r.uniform(); }

void ranlux_delete(RanluxStream* s) { delete s; }

void ranlux_uniforms(unsigned long long seed, int n, double* out)
{
  Ranlux48 r;
  r.seed(seed);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}

int launch_mode_fill(int n, const unsigned int* d_seedtable, const DevPower& pw, double box, double fac,
                     float2* cd0, float2* cd1, float2* cd2, hipStream_t s)
{
  const long long cols = (long long)n * n;
  hipLaunchKernelGGL(k_mode_fill, dim3((unsigned)((cols + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, n,
                     d_seedtable, pw, box, fac, cd0, cd1, cd2);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_lpt_source(int n, const float* g0, const float* g1, const float* g2, float* g3, const float* g4,
                      const float* g5, hipStream_t s)
{
  const long long total = (long long)n * n * n;
  hipLaunchKernelGGL(k_lpt_source, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, n, g0,
                     g1, g2, g3, g4, g5);
  COLA_HIP(hipGetLastError());
  return kOk;
}

int launch_lpt_particles(int n, const float* const disp[3], const float* const disp2[3],
                         const LptParticleFactors& f, const DevParticles& p, hipStream_t s)
{
  const long long total = (long long)n * n * n;
  DispPtrs g;
  for (int a = 0; a < 3; ++a) {
    g.d[a] = disp[a];
    g.d2[a] = disp2[a];
  }
  hipLaunchKernelGGL(k_lpt_particles, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, n,
                     g, f, p);
  COLA_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cola
