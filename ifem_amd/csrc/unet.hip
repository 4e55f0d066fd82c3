// unet.hip -- a small 3D convolutional U-net used as a learned
// preconditioner, with its own forward, backward and Adam kernels.
//
// Layout: activations NCDHW with FP32 accumulation everywhere; the storage
// type of activations is float (training) or bf16 (inference path of the
// preconditioner).  Convolutions are 3x3x3, zero padding 1, stride 1, with
// bias and optional ReLU; down-sampling is 2x2x2 average pooling,
// up-sampling is nearest-neighbour x2; skip connections concatenate
// channels (the consuming convolution reads two input tensors).
//
// Architecture (levels L, channels c[0..L-1], in/out channels Ci, Co):
//   enc l: conv(in_l -> c_l) relu, conv(c_l -> c_l) relu  [pool to l+1]
//   dec l (l = L-2..0): up(dec_{l+1}) ++ enc_l -> conv(c_{l+1}+c_l -> c_l)
//                       relu, conv(c_l -> c_l) relu
//   out: conv(c_0 -> Co), no ReLU
// Parameters live in one flat FP32 buffer (weights [co][ci][27], bias [co]
// per conv); gradients and Adam moments mirror it.
//
// Weight loads use block-uniform addresses, so they become scalar loads
// and the inner loop is LDS read + FMAs with a scalar operand.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "unet.hpp"

namespace ifem {

typedef __hip_bfloat16 bf16;

__device__ inline float ld(const float* p, int64_t i) { return p[i]; }
__device__ inline float ld(const bf16* p, int64_t i) { return __bfloat162float(p[i]); }
__device__ inline void st(float* p, int64_t i, float v) { p[i] = v; }
__device__ inline void st(bf16* p, int64_t i, float v) { p[i] = __float2bfloat16(v); }

// ---------------------------------------------------------------------------
// Convolution forward: out[n,co,v] = b[co] + sum_{ci,k} w[co,ci,k] in[n,ci,v+k-1]
// Inputs: in0 (c0 channels) followed by in1 (c1 channels, may be null).
// Block: 8x8x4 output voxels, COT output channels per thread.
constexpr int TX = 8, TY = 8, TZ = 4;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;

template <typename TI, typename TO, int COT>
__global__ __launch_bounds__(TX* TY* TZ) void k_conv3d(
    const TI* __restrict__ in0, int c0, const TI* __restrict__ in1, int c1,
    const float* __restrict__ w, const float* __restrict__ bias, TO* __restrict__ out, int co_tot,
    int D, int H, int W, int relu, int flip)
{
  __shared__ float tile[HZ][HY][HX];
  const int tx = threadIdx.x % TX, ty = (threadIdx.x / TX) % TY, tz = threadIdx.x / (TX * TY);
  const int ntx = (W + TX - 1) / TX, nty = (H + TY - 1) / TY;
  const int bt = blockIdx.x;
  const int bx = bt % ntx, by = (bt / ntx) % nty, bz = bt / (ntx * nty);
  const int x = bx * TX + tx, y = by * TY + ty, z = bz * TZ + tz;
  const int cog = blockIdx.y * COT;
  const int n = blockIdx.z;
  const int cin = c0 + c1;
  const int64_t vol = (int64_t)D * H * W;
  float acc[COT];
#pragma unroll
  for (int o = 0; o < COT; o++) acc[o] = (bias && cog + o < co_tot) ? bias[cog + o] : 0.f;
  for (int ci = 0; ci < cin; ci++) {
    const TI* src = ci < c0 ? in0 + ((int64_t)n * c0 + ci) * vol
                            : in1 + ((int64_t)n * c1 + (ci - c0)) * vol;
    __syncthreads();
    for (int t = threadIdx.x; t < HZ * HY * HX; t += TX * TY * TZ) {
      const int hx = t % HX, hy = (t / HX) % HY, hz = t / (HX * HY);
      const int gx = bx * TX + hx - 1, gy = by * TY + hy - 1, gz = bz * TZ + hz - 1;
      float v = 0.f;
      if (gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D)
        v = ld(src, ((int64_t)gz * H + gy) * W + gx);
      tile[hz][hy][hx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 27; k++) {
      const int kx = k % 3, ky = (k / 3) % 3, kz = k / 9;
      const float v = tile[tz + kz][ty + ky][tx + kx];
      const int kk = flip ? 26 - k : k;
#pragma unroll
      for (int o = 0; o < COT; o++) {
        const int co = cog + o;
        if (co < co_tot) acc[o] += w[((int64_t)co * cin + ci) * 27 + kk] * v;
      }
    }
  }
  if (x < W && y < H && z < D) {
    const int64_t vi = ((int64_t)z * H + y) * W + x;
#pragma unroll
    for (int o = 0; o < COT; o++) {
      const int co = cog + o;
      if (co < co_tot) {
        float r = acc[o];
        if (relu && r < 0.f) r = 0.f;
        st(out, ((int64_t)n * co_tot + co) * vol + vi, r);
      }
    }
  }
}

// Weight/bias gradients: gw[co,ci,k] = sum_{n,v} g[n,co,v] in[n,ci,v+k-1],
// gb[co] = sum_{n,v} g[n,co,v].  One block per (co, ci); threads stride over
// voxels, fixed-order block reduction (deterministic).  ci == cin computes gb.
template <typename TI>
__global__ __launch_bounds__(256) void k_conv3d_wgrad(
    const TI* __restrict__ in0, int c0, const TI* __restrict__ in1, int c1,
    const float* __restrict__ g, int co_tot, int N, int D, int H, int W,
    float* __restrict__ gw, float* __restrict__ gb)
{
  __shared__ float red[27][256 / 64];
  const int co = blockIdx.x, ci = blockIdx.y;
  const int cin = c0 + c1;
  const int64_t vol = (int64_t)D * H * W;
  float acc[27];
#pragma unroll
  for (int k = 0; k < 27; k++) acc[k] = 0.f;
  const bool is_bias = ci == cin;
  for (int n = 0; n < N; n++) {
    const float* gp = g + ((int64_t)n * co_tot + co) * vol;
    const TI* src = nullptr;
    if (!is_bias)
      src = ci < c0 ? in0 + ((int64_t)n * c0 + ci) * vol : in1 + ((int64_t)n * c1 + (ci - c0)) * vol;
    for (int64_t v = threadIdx.x; v < vol; v += blockDim.x) {
      const float gv = gp[v];
      if (is_bias) {
        acc[0] += gv;
        continue;
      }
      const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
#pragma unroll
      for (int k = 0; k < 27; k++) {
        const int gx = x + k % 3 - 1, gy = y + (k / 3) % 3 - 1, gz = z + k / 9 - 1;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D)
          acc[k] += gv * ld(src, ((int64_t)gz * H + gy) * W + gx);
      }
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nk = is_bias ? 1 : 27;
  for (int k = 0; k < nk; k++) {
    float s = acc[k];
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < nk) {
    float s = 0.f;
    for (int wv = 0; wv < (int)(blockDim.x / 64); wv++) s += red[threadIdx.x][wv];
    if (is_bias) gb[co] = s;
    else gw[((int64_t)co * cin + ci) * 27 + threadIdx.x] = s;
  }
}

// Transposed weights for the data gradient: wt[ci][co][k] = w[co][ci][k]
// (the spatial flip is applied by the conv kernel's flip flag).
__global__ void k_transpose_w(const float* __restrict__ w, float* __restrict__ wt, int co_tot,
                              int cin)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)co_tot * cin * 27;
  if (i >= tot) return;
  const int k = (int)(i % 27);
  const int64_t r = i / 27;
  const int ci = (int)(r % cin), co = (int)(r / cin);
  wt[((int64_t)ci * co_tot + co) * 27 + k] = w[i];
}

// ReLU backward in place: g *= (y > 0)
template <typename TY>
__global__ void k_relu_bwd(float* __restrict__ g, const TY* __restrict__ y, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(ld(y, i) > 0.f)) g[i] = 0.f;
}

// 2x2x2 average pooling (D, H, W even) and its adjoint.
template <typename T>
__global__ void k_pool(const T* __restrict__ in, T* __restrict__ out, int NC, int D, int H, int W)
{
  const int d = D / 2, h = H / 2, w = W / 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)NC * d * h * w;
  if (i >= tot) return;
  const int x = (int)(i % w), y = (int)((i / w) % h), z = (int)((i / ((int64_t)w * h)) % d);
  const int64_t c = i / ((int64_t)w * h * d);
  const T* s = in + c * (int64_t)D * H * W;
  float acc = 0.f;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++)
        acc += ld(s, ((int64_t)(2 * z + dz) * H + 2 * y + dy) * W + 2 * x + dx);
  st(out, i, 0.125f * acc);
}

__global__ void k_pool_bwd(const float* __restrict__ gout, float* __restrict__ gin, int NC, int D,
                           int H, int W)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)NC * D * H * W;
  if (i >= tot) return;
  const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)((i / ((int64_t)W * H)) % D);
  const int64_t c = i / ((int64_t)W * H * D);
  const int d = D / 2, h = H / 2, w = W / 2;
  gin[i] = 0.125f * gout[c * (int64_t)d * h * w + ((int64_t)(z / 2) * h + y / 2) * w + x / 2];
}

// nearest up-sampling x2 and its adjoint (sum of the 8 children)
template <typename T>
__global__ void k_up(const T* __restrict__ in, T* __restrict__ out, int NC, int D, int H, int W)
{
  // D, H, W: output (fine) sizes
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)NC * D * H * W;
  if (i >= tot) return;
  const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)((i / ((int64_t)W * H)) % D);
  const int64_t c = i / ((int64_t)W * H * D);
  const int d = D / 2, h = H / 2, w = W / 2;
  st(out, i, ld(in, c * (int64_t)d * h * w + ((int64_t)(z / 2) * h + y / 2) * w + x / 2));
}

__global__ void k_up_bwd(const float* __restrict__ gout, float* __restrict__ gin, int NC, int D,
                         int H, int W)
{
  // D, H, W: fine sizes; gin is coarse
  const int d = D / 2, h = H / 2, w = W / 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)NC * d * h * w;
  if (i >= tot) return;
  const int x = (int)(i % w), y = (int)((i / w) % h), z = (int)((i / ((int64_t)w * h)) % d);
  const int64_t c = i / ((int64_t)w * h * d);
  const float* s = gout + c * (int64_t)D * H * W;
  float acc = 0.f;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++)
        acc += s[((int64_t)(2 * z + dz) * H + 2 * y + dy) * W + 2 * x + dx];
  gin[i] = acc;
}

// split a concatenated gradient [c0 | c1] per sample into two tensors
__global__ void k_split(const float* __restrict__ g, float* __restrict__ g0, int c0,
                        float* __restrict__ g1, int c1, int N, int64_t vol)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)N * (c0 + c1) * vol;
  if (i >= tot) return;
  const int64_t v = i % vol;
  const int64_t r = i / vol;
  const int c = (int)(r % (c0 + c1)), n = (int)(r / (c0 + c1));
  if (c < c0) g0[((int64_t)n * c0 + c) * vol + v] = g[i];
  else g1[((int64_t)n * c1 + (c - c0)) * vol + v] = g[i];
}

__global__ void k_add_inplace(float* __restrict__ a, const float* __restrict__ b, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

// Weighted MSE loss 0.5*sum_n w_n sum_v (y-t)^2 / N (w_n = 1 without
// weights), gradient g = w_n (y - t)/N; per-block partials of the loss.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ g, int64_t n, float inv_n,
                                             const float* __restrict__ wts, int64_t per_sample,
                                             double* __restrict__ part)
{
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float d = y[i] - t[i];
    const float wn = wts ? wts[i / per_sample] * inv_n : inv_n;
    g[i] = d * wn;
    s += 0.5 * (double)d * d * wn;
  }
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

template <typename T>
__global__ void k_convert(const float* __restrict__ a, T* __restrict__ b, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st(b, i, a[i]);
}

// ---------------------------------------------------------------------------
static inline unsigned grid1(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <typename TI, typename TO>
static void conv_launch(hipStream_t s, const TI* in0, int c0, const TI* in1, int c1,
                        const float* w, const float* b, TO* out, int co, int N, int D, int H,
                        int W, bool relu, bool flip)
{
  const int ntile = ((W + TX - 1) / TX) * ((H + TY - 1) / TY) * ((D + TZ - 1) / TZ);
  if (co >= 8) {
    dim3 g(ntile, (co + 7) / 8, N);
    k_conv3d<TI, TO, 8><<<g, TX * TY * TZ, 0, s>>>(in0, c0, in1, c1, w, b, out, co, D, H, W,
                                                   relu ? 1 : 0, flip ? 1 : 0);
  } else {
    dim3 g(ntile, (co + 1) / 2, N);
    k_conv3d<TI, TO, 2><<<g, TX * TY * TZ, 0, s>>>(in0, c0, in1, c1, w, b, out, co, D, H, W,
                                                   relu ? 1 : 0, flip ? 1 : 0);
  }
}

// ---------------------------------------------------------------------------
UNet::~UNet() { release(); }

void UNet::release()
{
  for (void* p : bufs) hipFree(p);
  bufs.clear();
  hipFree(params); hipFree(grads); hipFree(m1); hipFree(m2); hipFree(wtmp); hipFree(params16);
  hipFree(part); hipFree(gscratch);
  params = grads = m1 = m2 = wtmp = nullptr;
  params16 = nullptr;
  part = nullptr;
  gscratch = nullptr;
  cap_N = cap_D = cap_H = cap_W = 0;
}

int UNet::init(const ifem_gpu_unet_config& c, hipStream_t s)
{
  cfg = c;
  if (c.levels < 1 || c.levels > kMaxLevels) return fail(IFEM_GPU_EINVAL, "unet: levels must be 1..4");
  if (c.in_channels < 1 || c.out_channels < 1) return fail(IFEM_GPU_EINVAL, "unet: channel counts");
  for (int l = 0; l < c.levels; l++)
    if (c.channels[l] < 1) return fail(IFEM_GPU_EINVAL, "unet: channel counts");
  convs.clear();
  int64_t off = 0;
  auto add = [&](int ci, int co, bool relu) {
    Conv cv;
    cv.ci = ci; cv.co = co; cv.relu = relu;
    cv.w = off; off += (int64_t)co * ci * 27;
    cv.b = off; off += co;
    convs.push_back(cv);
  };
  const int L = c.levels;
  int prev = c.in_channels;
  for (int l = 0; l < L; l++) {
    add(prev, c.channels[l], true);
    add(c.channels[l], c.channels[l], true);
    prev = c.channels[l];
  }
  for (int l = L - 2; l >= 0; l--) {
    add(c.channels[l + 1] + c.channels[l], c.channels[l], true);
    add(c.channels[l], c.channels[l], true);
  }
  add(c.channels[0], c.out_channels, false);
  nparams = off;
  stream = s;
  IFEM_HIP(hipMalloc(&params, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&grads, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&m1, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&m2, sizeof(float) * nparams));
  IFEM_HIP(hipMemsetAsync(m1, 0, sizeof(float) * nparams, s));
  IFEM_HIP(hipMemsetAsync(m2, 0, sizeof(float) * nparams, s));
  int64_t maxw = 0;
  for (auto& cv : convs) maxw = std::max<int64_t>(maxw, (int64_t)cv.co * cv.ci * 27);
  IFEM_HIP(hipMalloc(&wtmp, sizeof(float) * maxw));
  IFEM_HIP(hipMalloc(&part, sizeof(double) * 1024));
  // He-normal initialisation from splitmix64 (seeded), biases 0
  std::vector<float> h(nparams, 0.f);
  uint64_t st = c.seed;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  for (size_t i = 0; i < convs.size(); i++) {
    const Conv& cv = convs[i];
    const double sd = std::sqrt(2.0 / (cv.ci * 27.0)) * (i + 1 == convs.size() ? 0.1 : 1.0);
    for (int64_t k = 0; k < (int64_t)cv.co * cv.ci * 27; k += 2) {
      const double u1 = unif(), u2 = unif();
      const double r = std::sqrt(-2.0 * std::log(u1));
      h[cv.w + k] = (float)(sd * r * std::cos(2 * M_PI * u2));
      if (k + 1 < (int64_t)cv.co * cv.ci * 27) h[cv.w + k + 1] = (float)(sd * r * std::sin(2 * M_PI * u2));
    }
  }
  IFEM_HIP(hipMemcpyAsync(params, h.data(), sizeof(float) * nparams, hipMemcpyHostToDevice, s));
  IFEM_HIP(hipStreamSynchronize(s));
  step = 0;
  return IFEM_GPU_OK;
}

int UNet::ensure(int N, int D, int H, int W)
{
  if (N <= cap_N && D == cap_D && H == cap_H && W == cap_W) return IFEM_GPU_OK;
  const int L = cfg.levels;
  for (int l = 1; l < L; l++)
    if ((D >> (l - 1)) % 2 || (H >> (l - 1)) % 2 || (W >> (l - 1)) % 2)
      return fail(IFEM_GPU_EUNSUPPORTED, "unet: grid sizes must be divisible by 2^(levels-1)");
  for (void* p : bufs) hipFree(p);
  bufs.clear();
  hipFree(gscratch);
  gscratch = nullptr;
  bool ok = true;
  auto alloc = [&](int64_t n) -> float* {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(float) * (n > 0 ? n : 1)) != hipSuccess) { ok = false; return nullptr; }
    bufs.push_back(p);
    return (float*)p;
  };
  maxt = 0;
  for (int l = 0; l < L; l++) {
    dims[l][0] = D >> l; dims[l][1] = H >> l; dims[l][2] = W >> l;
    const int64_t vol = (int64_t)dims[l][0] * dims[l][1] * dims[l][2];
    enc_a[l] = alloc(N * cfg.channels[l] * vol);
    enc_b[l] = alloc(N * cfg.channels[l] * vol);
    pooled[l] = l > 0 ? alloc(N * cfg.channels[l - 1] * vol) : nullptr;
    dec_up[l] = l < L - 1 ? alloc(N * cfg.channels[l + 1] * vol) : nullptr;
    dec_a[l] = l < L - 1 ? alloc(N * cfg.channels[l] * vol) : nullptr;
    dec_b[l] = l < L - 1 ? alloc(N * cfg.channels[l] * vol) : nullptr;
    int cmax = cfg.channels[l] + (l + 1 < L ? cfg.channels[l + 1] : 0);
    if (l == 0) cmax = std::max(cmax, std::max(cfg.in_channels, cfg.out_channels));
    maxt = std::max<int64_t>(maxt, (int64_t)N * cmax * vol);
  }
  outbuf = alloc((int64_t)N * cfg.out_channels * D * H * W);
  if (!ok) return fail(IFEM_GPU_ENOMEM, "unet: out of device memory");
  IFEM_HIP(hipMalloc(&gscratch, sizeof(float) * 4 * maxt));
  cap_N = N; cap_D = D; cap_H = H; cap_W = W;
  return IFEM_GPU_OK;
}

// Forward pass (FP32 storage, FP32 accumulation); returns the output buffer.
const float* UNet::forward(const float* x, int N, int D, int H, int W)
{
  (void)D; (void)H; (void)W;
  const int L = cfg.levels;
  hipStream_t s = stream;
  const bool act = cfg.activation == 0;
  int ic = 0;
  for (int l = 0; l < L; l++) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    const float* in = x;
    int cin = cfg.in_channels;
    if (l > 0) {
      k_pool<float><<<grid1((int64_t)N * cfg.channels[l - 1] * d * h * w), 256, 0, s>>>(
          enc_b[l - 1], pooled[l], N * cfg.channels[l - 1], 2 * d, 2 * h, 2 * w);
      in = pooled[l];
      cin = cfg.channels[l - 1];
    }
    const Conv& c1 = convs[ic++];
    conv_launch<float, float>(s, in, cin, (const float*)nullptr, 0, params + c1.w, params + c1.b,
                              enc_a[l], c1.co, N, d, h, w, act, false);
    const Conv& c2 = convs[ic++];
    conv_launch<float, float>(s, enc_a[l], c1.co, (const float*)nullptr, 0, params + c2.w,
                              params + c2.b, enc_b[l], c2.co, N, d, h, w, act, false);
  }
  const float* cur = enc_b[L - 1];
  for (int l = L - 2; l >= 0; l--) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    k_up<float><<<grid1((int64_t)N * cfg.channels[l + 1] * d * h * w), 256, 0, s>>>(
        cur, dec_up[l], N * cfg.channels[l + 1], d, h, w);
    const Conv& c1 = convs[ic++];
    conv_launch<float, float>(s, dec_up[l], cfg.channels[l + 1], enc_b[l], cfg.channels[l],
                              params + c1.w, params + c1.b, dec_a[l], c1.co, N, d, h, w, act,
                              false);
    const Conv& c2 = convs[ic++];
    conv_launch<float, float>(s, dec_a[l], c1.co, (const float*)nullptr, 0, params + c2.w,
                              params + c2.b, dec_b[l], c2.co, N, d, h, w, act, false);
    cur = dec_b[l];
  }
  const Conv& co = convs[ic++];
  conv_launch<float, float>(s, cur, cfg.channels[0], (const float*)nullptr, 0, params + co.w,
                            params + co.b, outbuf, co.co, N, dims[0][0], dims[0][1], dims[0][2],
                            false, false);
  return outbuf;
}

// Backward of one convolution given g = dLoss/d(pre-activation output):
// weight and bias gradients into grads, input gradient(s) into gin0 (c0
// channels) and gin1 (c1 channels) when requested.
void UNet::conv_backward(const Conv& cv, const float* in0, int c0, const float* in1, int c1,
                         const float* g, float* gin0, float* gin1, int N, int d, int h, int w)
{
  hipStream_t s = stream;
  dim3 gg(cv.co, cv.ci + 1);
  k_conv3d_wgrad<float><<<gg, 256, 0, s>>>(in0, c0, in1, c1, g, cv.co, N, d, h, w, grads + cv.w,
                                           grads + cv.b);
  if (!gin0) return;
  const int64_t nw = (int64_t)cv.co * cv.ci * 27;
  k_transpose_w<<<grid1(nw), 256, 0, s>>>(params + cv.w, wtmp, cv.co, cv.ci);
  if (!gin1) {
    conv_launch<float, float>(s, g, cv.co, (const float*)nullptr, 0, wtmp, nullptr, gin0, cv.ci, N,
                              d, h, w, false, true);
  } else {
    float* gcat = gscratch; // slot 0
    conv_launch<float, float>(s, g, cv.co, (const float*)nullptr, 0, wtmp, nullptr, gcat, cv.ci, N,
                              d, h, w, false, true);
    const int64_t vol = (int64_t)d * h * w;
    k_split<<<grid1((int64_t)N * cv.ci * vol), 256, 0, s>>>(gcat, gin0, c0, gin1, c1, N, vol);
  }
}

int UNet::backward(const float* x, const float* target, const float* wts, const float* dy, int N,
                   int D, int H, int W, bool want_loss, float* loss)
{
  const int L = cfg.levels;
  hipStream_t s = stream;
  const int64_t nout = (int64_t)N * cfg.out_channels * D * H * W;
  float* gA = gscratch + maxt;
  float* gB = gscratch + 2 * maxt;
  const int nb = 256;
  if (dy) {
    // the caller's loss: dL/dy given
    IFEM_HIP(hipMemcpyAsync(gA, dy, sizeof(float) * nout, hipMemcpyDeviceToDevice, s));
    want_loss = false;
  } else {
    k_mse<<<nb, 256, 0, s>>>(outbuf, target, gA, nout, 1.0f / (float)N, wts, nout / N, part);
  }
  int ic = (int)convs.size() - 1;
  const bool act = cfg.activation == 0;
  auto relu_bwd = [&](float* g, const float* y, int64_t n) {
    if (act) k_relu_bwd<float><<<grid1(n), 256, 0, s>>>(g, y, n);
  };
  // output convolution (no ReLU)
  {
    const Conv& cv = convs[ic--];
    const float* cur = L > 1 ? dec_b[0] : enc_b[0];
    conv_backward(cv, cur, cfg.channels[0], nullptr, 0, gA, gB, nullptr, N, D, H, W);
    std::swap(gA, gB);
  }
  // decoder stages in reverse of the forward order (l = 0 .. L-2)
  int icd = ic; // index of the last decoder conv
  for (int l = 0; l <= L - 2; l++) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    const int64_t n_l = (int64_t)N * cfg.channels[l] * d * h * w;
    const Conv& c2 = convs[icd - 2 * l];
    const Conv& c1 = convs[icd - 2 * l - 1];
    relu_bwd(gA, dec_b[l], n_l);
    conv_backward(c2, dec_a[l], c2.ci, nullptr, 0, gA, gB, nullptr, N, d, h, w);
    std::swap(gA, gB);
    relu_bwd(gA, dec_a[l], n_l);
    // inputs were [up(cur) (c_{l+1}), enc_b[l] (c_l)]; the skip gradient goes
    // to dec_a[l], which is no longer needed
    conv_backward(c1, dec_up[l], cfg.channels[l + 1], enc_b[l], cfg.channels[l], gA, gB, dec_a[l],
                  N, d, h, w);
    k_up_bwd<<<grid1((int64_t)N * cfg.channels[l + 1] * d * h * w / 8), 256, 0, s>>>(
        gB, gA, N * cfg.channels[l + 1], d, h, w);
  }
  ic = icd - 2 * (L - 1);
  // encoder levels from the coarsest; gA = gradient w.r.t. enc_b[L-1]
  for (int l = L - 1; l >= 0; l--) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    const int64_t n_l = (int64_t)N * cfg.channels[l] * d * h * w;
    if (l < L - 1) k_add_inplace<<<grid1(n_l), 256, 0, s>>>(gA, dec_a[l], n_l);
    const Conv& c2 = convs[2 * l + 1];
    const Conv& c1 = convs[2 * l];
    relu_bwd(gA, enc_b[l], n_l);
    conv_backward(c2, enc_a[l], c2.ci, nullptr, 0, gA, gB, nullptr, N, d, h, w);
    std::swap(gA, gB);
    relu_bwd(gA, enc_a[l], n_l);
    const float* in = l > 0 ? pooled[l] : x;
    conv_backward(c1, in, c1.ci, nullptr, 0, gA, l > 0 ? gB : nullptr, nullptr, N, d, h, w);
    if (l > 0) {
      const int64_t nf = (int64_t)N * cfg.channels[l - 1] * dims[l - 1][0] * dims[l - 1][1] *
                         dims[l - 1][2];
      k_pool_bwd<<<grid1(nf), 256, 0, s>>>(gB, gA, N * cfg.channels[l - 1], dims[l - 1][0],
                                           dims[l - 1][1], dims[l - 1][2]);
    }
  }
  (void)ic;
  IFEM_HIP(hipGetLastError());
  if (want_loss && loss) {
    double hp[nb];
    IFEM_HIP(hipMemcpyAsync(hp, part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    IFEM_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (int i = 0; i < nb; i++) t += hp[i];
    *loss = (float)t;
  }
  return IFEM_GPU_OK;
}

int UNet::adam(float lr)
{
  step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  k_adam<<<grid1(nparams), 256, 0, stream>>>(params, grads, m1, m2, nparams, lr, b1, b2, 1e-8f, c1,
                                              c2, 0.f);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

} // namespace ifem

// ---------------------------------------------------------------------------
// C ABI
using namespace ifem;

extern "C" {

int ifem_gpu_unet_create(ifem_gpu_ctx* ctx, const ifem_gpu_unet_config* cfg, int64_t* nparams)
{
  if (!ctx || !cfg) return fail(IFEM_GPU_EINVAL, "unet_create: null argument");
  IFEM_HIP(hipSetDevice(ctx->c.device));
  ctx->c.net.reset(new UNet());
  int rc = ctx->c.net->init(*cfg, ctx->c.stream);
  if (rc) {
    ctx->c.net.reset();
    return rc;
  }
  if (nparams) *nparams = ctx->c.net->nparams;
  return IFEM_GPU_OK;
}

int ifem_gpu_unet_get_params(ifem_gpu_ctx* ctx, float* host)
{
  if (!ctx || !ctx->c.net || !host) return fail(IFEM_GPU_EINVAL, "unet_get_params: no network");
  UNet& n = *ctx->c.net;
  IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
  IFEM_HIP(hipMemcpy(host, n.params, sizeof(float) * n.nparams, hipMemcpyDeviceToHost));
  return IFEM_GPU_OK;
}

int ifem_gpu_unet_set_params(ifem_gpu_ctx* ctx, const float* host)
{
  if (!ctx || !ctx->c.net || !host) return fail(IFEM_GPU_EINVAL, "unet_set_params: no network");
  UNet& n = *ctx->c.net;
  IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
  IFEM_HIP(hipMemcpy(n.params, host, sizeof(float) * n.nparams, hipMemcpyHostToDevice));
  return IFEM_GPU_OK;
}

int ifem_gpu_unet_forward(ifem_gpu_ctx* ctx, const float* x, float* y, int N, int D, int H, int W)
{
  if (!ctx || !ctx->c.net || !x || !y) return fail(IFEM_GPU_EINVAL, "unet_forward: bad argument");
  IFEM_HIP(hipSetDevice(ctx->c.device));
  UNet& n = *ctx->c.net;
  int rc = n.ensure(N, D, H, W);
  if (rc) return rc;
  const float* o = n.forward(x, N, D, H, W);
  IFEM_HIP(hipMemcpyAsync(y, o, sizeof(float) * (int64_t)N * n.cfg.out_channels * D * H * W,
                          hipMemcpyDeviceToDevice, ctx->c.stream));
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

int ifem_gpu_unet_gradient(ifem_gpu_ctx* ctx, const float* x, const float* t, const float* w,
                           int N, int D, int H, int W, float* grad_host, float* loss)
{
  if (!ctx || !ctx->c.net || !x || !t) return fail(IFEM_GPU_EINVAL, "unet_gradient: bad argument");
  IFEM_HIP(hipSetDevice(ctx->c.device));
  UNet& n = *ctx->c.net;
  int rc = n.ensure(N, D, H, W);
  if (rc) return rc;
  n.forward(x, N, D, H, W);
  rc = n.backward(x, t, w, nullptr, N, D, H, W, true, loss);
  if (rc) return rc;
  if (grad_host) {
    IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
    IFEM_HIP(hipMemcpy(grad_host, n.grads, sizeof(float) * n.nparams, hipMemcpyDeviceToHost));
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_unet_train_step(ifem_gpu_ctx* ctx, const float* x, const float* t, const float* w,
                             int N, int D, int H, int W, float lr, float* loss)
{
  if (!ctx || !ctx->c.net || !x || !t) return fail(IFEM_GPU_EINVAL, "unet_train_step: bad argument");
  IFEM_HIP(hipSetDevice(ctx->c.device));
  UNet& n = *ctx->c.net;
  int rc = n.ensure(N, D, H, W);
  if (rc) return rc;
  n.forward(x, N, D, H, W);
  rc = n.backward(x, t, w, nullptr, N, D, H, W, loss != nullptr, loss);
  if (rc) return rc;
  return n.adam(lr);
}

int ifem_gpu_unet_backward(ifem_gpu_ctx* ctx, const float* x, const float* dy, int N, int D, int H,
                           int W, float lr, float* grad_host)
{
  if (!ctx || !ctx->c.net || !x || !dy) return fail(IFEM_GPU_EINVAL, "unet_backward: bad argument");
  if (!(lr >= 0.f)) return fail(IFEM_GPU_EINVAL, "unet_backward: lr must be >= 0");
  IFEM_HIP(hipSetDevice(ctx->c.device));
  UNet& n = *ctx->c.net;
  int rc = n.ensure(N, D, H, W);
  if (rc) return rc;
  n.forward(x, N, D, H, W);
  rc = n.backward(x, nullptr, nullptr, dy, N, D, H, W, false, nullptr);
  if (rc) return rc;
  if (grad_host) {
    IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
    IFEM_HIP(hipMemcpy(grad_host, n.grads, sizeof(float) * n.nparams, hipMemcpyDeviceToHost));
  }
  if (lr > 0.f) return n.adam(lr);
  return IFEM_GPU_OK;
}

} // extern "C"
