// fft_compare: device time of this library's Stockham FFT against rocFFT on
// one GPU, both on an n^3 float grid resident in HBM.
//
//     fft_compare [n] [reps]        (defaults 768, 10)
//
// ours:   cola_fft_time (in place on the padded layout): r2c, c2r and the PM
//         force sequence (r2c, then per axis the Poisson gradient and a c2r);
// rocFFT: out-of-place r2c and c2r plans on the packed layouts, and the same
//         force sequence as r2c + 3 x (gradient multiply kernel + c2r).
// Prints one JSON object with the mean milliseconds per call.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/cola_halo_amd.h"

#define CHECK_HIP(x)                                                        \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
      std::exit(1);                                                         \
    }                                                                       \
  } while (0)
#define CHECK_FFT(x)                                                        \
  do {                                                                      \
    rocfft_status s_ = (x);                                                 \
    if (s_ != rocfft_status_success) {                                      \
      std::fprintf(stderr, "%s: rocfft status %d\n", #x, (int)s_);          \
      std::exit(1);                                                         \
    }                                                                       \
  } while (0)

namespace {

// F_a = i m_a / |m|^2 delta (k = 0 -> 0) on the half spectrum [x][y][kz].
__global__ void k_gradient(const float2* __restrict__ in, float2* __restrict__ out, int n, int axis)
{
  const int nzc = n / 2 + 1;
  const long long total = (long long)n * n * nzc;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= total) return;
  const int kz = (int)(idx % nzc);
  const long long line = idx / nzc;
  const int iy = (int)(line % n), ix = (int)(line / n);
  const int mx = ix <= n / 2 ? ix : ix - n, my = iy <= n / 2 ? iy : iy - n;
  const float k2 = (float)(mx * mx + my * my + kz * kz);
  const float m = axis == 0 ? (float)mx : (axis == 1 ? (float)my : (float)kz);
  const float f = k2 > 0.0f ? m / k2 : 0.0f;
  const float2 v = in[idx];
  out[idx] = make_float2(-f * v.y, f * v.x);
}

struct Timer {
  hipEvent_t a, b;
  Timer()
  {
    CHECK_HIP(hipEventCreate(&a));
    CHECK_HIP(hipEventCreate(&b));
  }
  ~Timer()
  {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  template <typename F>
  double mean_ms(int reps, F&& fn)
  {
    fn();  // warm-up
    CHECK_HIP(hipDeviceSynchronize());
    double tot = 0;
    for (int r = 0; r < reps; ++r) {
      CHECK_HIP(hipEventRecord(a, nullptr));
      fn();
      CHECK_HIP(hipEventRecord(b, nullptr));
      CHECK_HIP(hipEventSynchronize(b));
      float ms = 0;
      CHECK_HIP(hipEventElapsedTime(&ms, a, b));
      tot += ms;
    }
    return tot / reps;
  }
};

}  // namespace

int main(int argc, char** argv)
{
  const int n = argc > 1 ? std::atoi(argv[1]) : 768;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 10;
  double ours[3];
  const int st = cola_fft_time(n, reps, ours);
  if (st != 0) {
    std::fprintf(stderr, "cola_fft_time: %d %s\n", st, cola_last_error());
    return 1;
  }

  CHECK_FFT(rocfft_setup());
  const size_t lengths[3] = {(size_t)n, (size_t)n, (size_t)n};
  rocfft_plan fwd = nullptr, inv = nullptr;
  CHECK_FFT(rocfft_plan_create(&fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                               rocfft_precision_single, 3, lengths, 1, nullptr));
  CHECK_FFT(rocfft_plan_create(&inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                               rocfft_precision_single, 3, lengths, 1, nullptr));
  size_t wf = 0, wi = 0;
  CHECK_FFT(rocfft_plan_get_work_buffer_size(fwd, &wf));
  CHECK_FFT(rocfft_plan_get_work_buffer_size(inv, &wi));
  const size_t wbytes = wf > wi ? wf : wi;
  void* work = nullptr;
  if (wbytes) CHECK_HIP(hipMalloc(&work, wbytes));
  rocfft_execution_info info = nullptr;
  CHECK_FFT(rocfft_execution_info_create(&info));
  if (wbytes) CHECK_FFT(rocfft_execution_info_set_work_buffer(info, work, wbytes));

  const long long ng = (long long)n * n * n, nc = (long long)n * n * (n / 2 + 1);
  float *x = nullptr, *out[3] = {nullptr, nullptr, nullptr};
  float2 *y = nullptr, *w = nullptr;
  CHECK_HIP(hipMalloc(&x, sizeof(float) * ng));
  for (int a = 0; a < 3; ++a) CHECK_HIP(hipMalloc(&out[a], sizeof(float) * ng));
  CHECK_HIP(hipMalloc(&y, sizeof(float2) * nc));
  CHECK_HIP(hipMalloc(&w, sizeof(float2) * nc));
  std::vector<float> host(ng);
  for (long long i = 0; i < ng; ++i) host[i] = (float)((i * 2654435761ULL) % 1000003) / 1000003.0f - 0.5f;
  CHECK_HIP(hipMemcpy(x, host.data(), sizeof(float) * ng, hipMemcpyHostToDevice));

  Timer t;
  const unsigned blocks = (unsigned)((nc + 255) / 256);
  const double r2c = t.mean_ms(reps, [&] {
    void* in[1] = {x};
    void* o[1] = {y};
    CHECK_FFT(rocfft_execute(fwd, in, o, info));
  });
  const double c2r = t.mean_ms(reps, [&] {
    CHECK_HIP(hipMemcpyAsync(w, y, sizeof(float2) * nc, hipMemcpyDeviceToDevice, nullptr));  // c2r consumes its input
    void* in[1] = {w};
    void* o[1] = {out[0]};
    CHECK_FFT(rocfft_execute(inv, in, o, info));
  });
  // the copy inside the c2r timing is measured separately and removed
  const double copy = t.mean_ms(reps, [&] {
    CHECK_HIP(hipMemcpyAsync(w, y, sizeof(float2) * nc, hipMemcpyDeviceToDevice, nullptr));
  });
  const double forces = t.mean_ms(reps, [&] {
    void* in[1] = {x};
    void* o[1] = {y};
    CHECK_FFT(rocfft_execute(fwd, in, o, info));
    for (int a = 0; a < 3; ++a) {
      hipLaunchKernelGGL(k_gradient, dim3(blocks), dim3(256), 0, nullptr, y, w, n, a);
      void* wi_[1] = {w};
      void* oa[1] = {out[a]};
      CHECK_FFT(rocfft_execute(inv, wi_, oa, info));
    }
  });
  std::printf("{\"n\": %d, \"reps\": %d, \"ours_ms\": {\"r2c\": %.4f, \"c2r\": %.4f, \"pm_forces\": %.4f}, "
              "\"rocfft_ms\": {\"r2c\": %.4f, \"c2r\": %.4f, \"pm_forces\": %.4f}, "
              "\"speedup\": {\"r2c\": %.3f, \"c2r\": %.3f, \"pm_forces\": %.3f}}\n",
              n, reps, ours[0], ours[1], ours[2], r2c, c2r - copy, forces, r2c / ours[0], (c2r - copy) / ours[1],
              forces / ours[2]);
  CHECK_FFT(rocfft_execution_info_destroy(info));
  CHECK_FFT(rocfft_plan_destroy(fwd));
  CHECK_FFT(rocfft_plan_destroy(inv));
  CHECK_FFT(rocfft_cleanup());
  (void)hipFree(work);
  (void)hipFree(x);
  (void)hipFree(y);
  (void)hipFree(w);
  for (float* o : out) (void)hipFree(o);
  return 0;
}
