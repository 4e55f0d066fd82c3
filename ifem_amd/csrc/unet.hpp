// unet.hpp -- 3D conv U-net used as a learned preconditioner (unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ifem_gpu.h"

namespace ifem {

constexpr int kMaxLevels = 4;

// Input of one convolution: in0 (c0 channels stored with stride c0p) read at
// the output resolution (mode 0), nearest-upsampled from half resolution
// (mode 1) or 2x2x2-average-pooled from double resolution (mode 2),
// followed by in1 (c1 channels, stride c1p, output resolution; may be null).
// Tensors are channels-last bf16 (raw bits): [n][z][y][x][c].
struct ConvIn {
  const uint16_t* in0 = nullptr;
  int c0 = 0, c0p = 0, mode0 = 0;
  const uint16_t* in1 = nullptr;
  int c1 = 0, c1p = 0;
};

struct UNet {
  struct Conv {
    int c0 = 0, c1 = 0;       // real input channels (in0, in1)
    int c0p = 0, c1p = 0;     // padded (multiples of 8)
    int co = 0, cop = 0;      // real and padded (8) output channels
    bool relu = true;
    int64_t w = 0, b = 0;     // offsets into the flat FP32 parameter buffer
    int kb = 0, kbt = 0;      // 32-wide K blocks of the packed forward / data-gradient weights
    int64_t wp = 0, wt = 0;   // offsets (bf16 elements) into the packed weight buffers
    int ci() const { return c0 + c1; }
  };
  ifem_gpu_unet_config cfg{};
  std::vector<Conv> convs;
  int64_t nparams = 0;
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t *wpack = nullptr, *wtpack = nullptr; // bf16 MFMA operands
  int64_t wpack_n = 0, wtpack_n = 0;
  float* wpart = nullptr;     // weight-gradient partials
  int64_t wpart_n = 0;
  double* part = nullptr;
  hipStream_t stream = nullptr;
  int step = 0;
  bool packed = false;        // wpack/wtpack hold the current params

  // activations (bf16 channels-last) for the current capacity
  int cap_N = 0, cap_D = 0, cap_H = 0, cap_W = 0;
  int dims[kMaxLevels][3] = {};
  int cinp = 0, coutp = 0;
  uint16_t* xin = nullptr;
  uint16_t *enc_a[kMaxLevels] = {}, *enc_b[kMaxLevels] = {};
  uint16_t *dec_a[kMaxLevels] = {}, *dec_b[kMaxLevels] = {};
  uint16_t *g_skip[kMaxLevels] = {};
  uint16_t *gb1 = nullptr, *gb2 = nullptr, *gcat = nullptr, *gout = nullptr;
  float* outbuf = nullptr;
  std::vector<void*> bufs;

  ~UNet();
  void release();
  int init(const ifem_gpu_unet_config& c, hipStream_t s);
  int ensure(int N, int D, int H, int W);
  // repacks the bf16 MFMA weights from the FP32 parameters when they changed
  int pack();
  // skip (device, may be null): the kernels of this forward pass return at
  // once when *skip != 0 -- a solver's convergence flag, so that the
  // applications launched after convergence inside a check window cost only
  // their launches
  const float* forward(const float* x, int N, int D, int H, int W,
                       const double* skip = nullptr);
  // gradient of 0.5*sum_n w_n*sum((y-t)^2)/N (w_n = 1 if wts is null), or
  // with dy != nullptr of sum(dy*y), w.r.t. the parameters into grads, after
  // forward(x); returns the loss when want_loss (synchronises)
  int backward(const float* target, const float* wts, const float* dy, int N, bool want_loss,
               float* loss);
  int adam(float lr);

private:
  void conv_fwd(const Conv& cv, const ConvIn& in, int N, int d, int h, int w, uint16_t* out,
                float* out_f32, const double* skip = nullptr);
  void conv_dgrad(const Conv& cv, const uint16_t* g, int N, int d, int h, int w, uint16_t* gin);
  void conv_wgrad(const Conv& cv, const ConvIn& in, const uint16_t* g, int N, int d, int h, int w);
};

} // namespace ifem
