// unet.hpp -- 3D conv U-net used as a learned preconditioner (unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ifem_gpu.h"

namespace ifem {

constexpr int kMaxLevels = 4;

struct UNet {
  struct Conv {
    int ci = 0, co = 0;
    bool relu = true;
    int64_t w = 0, b = 0; // offsets into the flat parameter buffer
  };
  ifem_gpu_unet_config cfg{};
  std::vector<Conv> convs;
  int64_t nparams = 0;
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr, *wtmp = nullptr;
  void* params16 = nullptr;
  double* part = nullptr;
  hipStream_t stream = nullptr;
  int step = 0;

  // activations (FP32) for the current capacity
  int cap_N = 0, cap_D = 0, cap_H = 0, cap_W = 0;
  int dims[kMaxLevels][3] = {};
  float *enc_a[kMaxLevels] = {}, *enc_b[kMaxLevels] = {}, *pooled[kMaxLevels] = {};
  float *dec_up[kMaxLevels] = {}, *dec_a[kMaxLevels] = {}, *dec_b[kMaxLevels] = {};
  float* outbuf = nullptr;
  float* gscratch = nullptr; // 4 slots of maxt floats
  int64_t maxt = 0;
  std::vector<void*> bufs;

  ~UNet();
  void release();
  int init(const ifem_gpu_unet_config& c, hipStream_t s);
  int ensure(int N, int D, int H, int W);
  const float* forward(const float* x, int N, int D, int H, int W);
  // gradient of 0.5*sum_n w_n*sum((y-t)^2)/N (w_n = 1 if wts is null), or
  // with dy != nullptr of sum(dy*y), w.r.t. the parameters into grads, after
  // forward(x); returns the loss when want_loss (synchronises)
  int backward(const float* x, const float* target, const float* wts, const float* dy, int N,
               int D, int H, int W, bool want_loss, float* loss);
  int adam(float lr);

private:
  void conv_backward(const Conv& cv, const float* in0, int c0, const float* in1, int c1,
                     const float* g, float* gin0, float* gin1, int N, int d, int h, int w);
};

} // namespace ifem
