// Learned force residual: a 4-layer 3-D convolution network on the coarse PM
// grid (16 channels, 3^3 kernels, periodic), input asinh(delta) and the scale
// factor a, output the 3-component force residual between a coarse PM grid
// and a grid twice as fine.  Activations and weights in bfloat16, products
// on MFMA (v_mfma_f32_16x16x32_bf16) with fp32 accumulation; master weights
// and Adam moments in fp32.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cola {

constexpr int kNetLayers = 4;
constexpr int kNetCin[kNetLayers] = {8, 16, 16, 16};
constexpr int kNetCout[kNetLayers] = {16, 16, 16, 3};
constexpr int kNetParams = 27 * 8 * 16 + 16 + 2 * (27 * 16 * 16 + 16) + 27 * 16 * 3 + 3;

struct ResNet {
  int n = 0;  // grid cells per axis (multiple of 16)
  // fp32 master parameters (W1 b1 W2 b2 W3 b3 W4 b4; W[27][cin][cout]) and Adam moments
  float* d_param = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_grad = nullptr;
  long long adam_t = 0;
  // MFMA B operands: forward (per layer) and data-gradient (layers 2-4) packs
  unsigned short* d_wf[kNetLayers] = {nullptr, nullptr, nullptr, nullptr};
  unsigned short* d_wb[kNetLayers] = {nullptr, nullptr, nullptr, nullptr};
  // forward A operands of the 16-channel-input layers in k_conv16r's K order
  unsigned short* d_wr[kNetLayers] = {nullptr, nullptr, nullptr, nullptr};
  // activations: input [n^3][8], h1..h3 [n^3][16] (bf16), output y [3][n^3] fp32
  unsigned short* d_x = nullptr;
  unsigned short* d_f1 = nullptr;  // feature grid [n^3] (bf16) of host inputs (timing, tests)
  unsigned short* d_h[3] = {nullptr, nullptr, nullptr};
  float* d_y = nullptr;
  // gradients at layer outputs [n^3][16] (bf16), two buffers
  unsigned short* d_g[2] = {nullptr, nullptr};
  float* d_wpart = nullptr;  // per-wave partial weight gradients
  int wpart_waves = 0;
  double* d_loss = nullptr;
  float* d_a = nullptr;      // scale factor of standalone evaluations
  double* d_fsum = nullptr;  // sums of the residual the last corrected force added, per component
  float* d_fi = nullptr;     // the last corrected force grids, interleaved [cell][3] (read by the readout)
  float* d_trash = nullptr;  // stores of lanes without an output cell (k_conv3h)
  float out_scale = 1.0f;    // residual = out_scale * y
  bool scale_set = false;    // out_scale fitted to training targets (or set by the caller)
  float in_scale = 1.0f;     // feature = asinh(delta) * in_scale
  float restrict_side = 0.125f;  // training targets: fine force averaged with weights (w, 1-2w, w) per axis
  bool fuse12 = true;            // inference: features and layers 1-2 in one launch (resnet_layers12)
  bool roll = true;              // 16-channel-input layers by input planes (k_conv12r, k_conv16r, k_conv_head_r)
  bool fuse34 = true;            // inference by input planes: layer 3 and the last layer in one launch (k_conv3h)

  // training set: per sample the feature grid (bf16 [n^3]), the target
  // residual (bf16 [3][n^3]) and the scale factor a
  std::vector<unsigned short*> s_feat, s_target;
  std::vector<float> s_a;
};

int resnet_init(ResNet* r, int n, unsigned seed);
void resnet_free(ResNet* r);
// fp32 master parameters -> bf16 MFMA packs
int resnet_pack(ResNet* r, hipStream_t s);
// stores a training pair: features of the coarse density, target = fine force
// at the coincident points (fine cell 2i) minus coarse force, and a
int resnet_store_pair(ResNet* r, const float* coarse_delta, long long crow, const float* const coarse_f[3],
                      const float* const fine_f[3], long long frow, float a, hipStream_t s);
// mean loss over stored samples [first, first + count); train: epochs over
// samples [0, ntrain) in shuffled order, one Adam step per sample
int resnet_eval(ResNet* r, int first, int count, double* loss, hipStream_t s);
// out_scale = RMS of the targets of samples [0, count)
int resnet_fit_scale(ResNet* r, int count, hipStream_t s);
int resnet_train(ResNet* r, int ntrain, int epochs, float lr, unsigned seed, double* epoch_loss, hipStream_t s);
// forward from the 8-channel input d_x; writes d_y (and keeps h1..h3)
int resnet_forward(ResNet* r, hipStream_t s);
// the same net from a feature grid and the scale factor (read from device
// memory at pa when pa is given, so a captured step reads each step's value;
// else a): writes d_y
int resnet_forward_features(ResNet* r, const unsigned short* feat, const float* pa, float a, hipStream_t s);
// inference in parts around the PM force FFTs: the feature grid d_f1 from
// the real-space density grid (rows of `row` floats) before the r2c, layers
// 1-3 (a read from pa), and layer 4 added to the force grids:
// out[a] = out[a] + out_scale * y[a]
int resnet_features(ResNet* r, const float* delta, long long row, hipStream_t s);
int resnet_body(ResNet* r, const float* pa, hipStream_t s);
// resnet_body in its two launches groups: layer 1 (from d_f1), layers 2-3
int resnet_layer1(ResNet* r, const float* pa, hipStream_t s);
int resnet_layers23(ResNet* r, hipStream_t s);
// out[a] + out_scale * y[a] into d_fi[cell][a] (the force grids keep the
// plain force; d_fsum gets the sums of the residual added)
int resnet_head_add(ResNet* r, float* const out[3], long long row, hipStream_t s);
// The inference path with the first two layers fused (fuse12): h2 straight
// from the real-space density grid (features, layer 1 and layer 2 in one
// launch, a read from pa), then layer 3; equal to resnet_features,
// resnet_layer1 and layer 2 bit for bit.
int resnet_layers12(ResNet* r, const float* delta, long long row, const float* pa, hipStream_t s);
int resnet_layer3(ResNet* r, hipStream_t s);
// Layer 3 and the last layer in one launch (from h2, after the force
// transforms): the corrected force into d_fi as resnet_layer3 followed by
// resnet_head_add, bit for bit, without writing h3.  By input planes only.
int resnet_layer3_head(ResNet* r, float* const out[3], long long row, hipStream_t s);
// loss against target (bf16 [3][n^3] or fp32 [3][n^3]) and gradient into d_grad
int resnet_backward(ResNet* r, const unsigned short* target_bf16, const float* target_f32, hipStream_t s);
int resnet_adam(ResNet* r, float lr, hipStream_t s);

}  // namespace cola
