// unet.hip -- a small 3D convolutional U-net used as a learned
// preconditioner: forward, data-gradient and weight-gradient kernels on BF16
// matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), and Adam.
//
// Storage: activations and their gradients are channels-last bf16,
// [n][z][y][x][c], with c padded to a multiple of 8 by zero channels, so that
// 8 channels of one voxel are one 16-byte load.  The parameters are FP32 in
// one flat buffer (per convolution the weights [co][ci][27], tap k = kz*9 +
// ky*3 + kx, then the biases [co]); gradients and Adam moments mirror it.
// Before use they are repacked into bf16 MFMA operands (pack()).
//
// Convolutions are 3x3x3, zero padding 1, stride 1, with bias and optional
// ReLU; down-sampling is 2x2x2 average pooling, up-sampling nearest x2, skip
// connections concatenate channels.  Pooling, up-sampling and concatenation
// are folded into the halo load of the consuming convolution (ConvIn), so
// they never exist as tensors.
//
// Architecture (levels L, channels c[0..L-1], in/out channels Ci, Co):
//   enc l: conv(in_l -> c_l) relu, conv(c_l -> c_l) relu   (in_l = pool(enc_{l-1}))
//   dec l (l = L-2..0): conv([up(dec_{l+1}), enc_l] -> c_l) relu, conv(c_l -> c_l) relu
//   out: conv(c_0 -> Co), no ReLU, FP32 NCDHW output
//
// Forward convolution = implicit GEMM: M = output voxels (16 along x per
// MFMA row block), N = output channels (16 per column block), K = (tap,
// input channel) in chunks of at most 32 channels whose 3D halo is staged in
// LDS.  A workgroup owns a 16x4x4 brick of output voxels; wave w the z-slice
// w (four 16-voxel rows).  The data gradient is the same kernel with the
// weights transposed and flipped.  The weight gradient is a GEMM over voxels
// (K) with M = output channels and N = (tap, input channel): the brick's
// output gradients are staged transposed in LDS (voxels contiguous), the
// input halo in three x-shifted transposed copies, so that every MFMA operand
// is one aligned 16-byte LDS read; per-workgroup partial sums are reduced in
// a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "unet.hpp"

namespace ifem {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ inline float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
// round to nearest even
__device__ inline uint32_t f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
__device__ inline void unpack8(const uint4 v, float* f)
{
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f[2 * i] = bf2f(w[i] & 0xffffu);
    f[2 * i + 1] = bf2f(w[i] >> 16);
  }
}
__device__ inline uint4 pack8(const float* f)
{
  uint4 v;
  v.x = f2bf(f[0]) | (f2bf(f[1]) << 16);
  v.y = f2bf(f[2]) | (f2bf(f[3]) << 16);
  v.z = f2bf(f[4]) | (f2bf(f[5]) << 16);
  v.w = f2bf(f[6]) | (f2bf(f[7]) << 16);
  return v;
}

// 8 channels [c, c+8) of a convolution input at voxel (z, y, x) of sample n
// (output resolution D x H x W); zero outside the volume.
__device__ inline uint4 load8(const ConvIn& in, int n, int D, int H, int W, int z, int y, int x,
                              int c)
{
  uint4 r = make_uint4(0u, 0u, 0u, 0u);
  if (z < 0 || z >= D || y < 0 || y >= H || x < 0 || x >= W) return r;
  if (c >= in.c0p)
    return *(const uint4*)(in.in1 + ((((int64_t)n * D + z) * H + y) * W + x) * in.c1p +
                           (c - in.c0p));
  if (in.mode0 == 0)
    return *(const uint4*)(in.in0 + ((((int64_t)n * D + z) * H + y) * W + x) * in.c0p + c);
  if (in.mode0 == 1) {
    const int d = D >> 1, h = H >> 1, w = W >> 1;
    return *(const uint4*)(in.in0 + ((((int64_t)n * d + (z >> 1)) * h + (y >> 1)) * w + (x >> 1)) *
                                        in.c0p + c);
  }
  const int D2 = 2 * D, H2 = 2 * H, W2 = 2 * W;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, f[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
    unpack8(*(const uint4*)(in.in0 + ((((int64_t)n * D2 + zz) * H2 + yy) * W2 + xx) * in.c0p + c), f);
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] += f[i];
  }
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] *= 0.125f;
  return pack8(s);
}

// ---------------------------------------------------------------------------
// Convolution (forward, and data gradient with transposed weights).
constexpr int BX = 16, BY = 4, BZ = 4;           // output brick
constexpr int HXs = BX + 2, HYs = BY + 2, HZs = BZ + 2;
constexpr int HVOX = HXs * HYs * HZs;             // 648 halo voxels
constexpr int CCMAX = 32;                         // input channels per K chunk

struct ConvArgs {
  ConvIn x;
  int D, H, W;
  const uint16_t* wp; // packed weights [NT*16][kb_total*32]
  int kb_total;
  const float* bias;  // cout entries, or null
  int cout;           // channels written with values (the rest of cout_p8: 0)
  int cout_p8;        // channel stride of out
  uint16_t* out;      // bf16 channels-last, or null
  float* out_f32;     // FP32 NCDHW (cout channels), or null
  int relu;
  const double* skip = nullptr; // return at once when *skip != 0 (a solver's done flag)
};

__host__ __device__ inline int conv_bricks(int D, int H, int W)
{
  return ((W + BX - 1) / BX) * ((H + BY - 1) / BY) * ((D + BZ - 1) / BZ);
}

template <int NT>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a)
{
  if (a.skip && *a.skip != 0.0) return;
  // halo tile of up to CCMAX input channels, sized by the launch to the
  // channels this convolution has (fewer channels: more blocks per CU)
  extern __shared__ uint4 tile[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tb = blockIdx.z * NT; // first 16-channel output tile of this block
  const int row = lane & 15, grp = lane >> 4;
  const int nbx = (a.W + BX - 1) / BX, nby = (a.H + BY - 1) / BY;
  const int b = blockIdx.x;
  const int x0 = (b % nbx) * BX, y0 = ((b / nbx) % nby) * BY, z0 = (b / (nbx * nby)) * BZ;
  const int n = blockIdx.y;
  const int cin_p = a.x.c0p + a.x.c1p;
  f32x4 acc[4][NT];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t wrow = (int64_t)a.kb_total * 32;
  int kb_base = 0;
  for (int c_lo = 0; c_lo < cin_p; c_lo += CCMAX) {
    const int cc = min(CCMAX, cin_p - c_lo);
    const int ng8 = cc >> 3;
    __syncthreads();
    for (int t = threadIdx.x; t < HVOX * ng8; t += 256) {
      const int g8 = t % ng8, v = t / ng8;
      const int hx = v % HXs, hy = (v / HXs) % HYs, hz = v / (HXs * HYs);
      tile[t] = load8(a.x, n, a.D, a.H, a.W, z0 + hz - 1, y0 + hy - 1, x0 + hx - 1, c_lo + 8 * g8);
    }
    __syncthreads();
    const int nkb = (27 * cc + 31) / 32;
    for (int kb = 0; kb < nkb; kb++) {
      const int kl = kb * 32 + grp * 8;
      const int tap0 = kl / cc;
      const int cl = kl - tap0 * cc;
      const int tap = tap0 > 26 ? 26 : tap0; // padding of the last block: zero weights
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      bf16x8 bw[NT];
#pragma unroll
      for (int t = 0; t < NT; t++)
        bw[t] = __builtin_bit_cast(
            bf16x8, *(const uint4*)(a.wp + (int64_t)((tb + t) * 16 + row) * wrow + (kb_base + kb) * 32 + grp * 8));
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int hv = ((wave + dz) * HYs + (m + dy)) * HXs + (row + dx);
        const bf16x8 av = __builtin_bit_cast(bf16x8, tile[hv * ng8 + (cl >> 3)]);
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[t], acc[m][t], 0, 0, 0);
      }
    }
    kb_base += nkb;
  }
  // D fragment: column (output channel) t*16 + row, rows (x) 4*grp + i.
  // bf16 outputs go through LDS ([brick voxel][NT*16 channels]) so that each
  // voxel's channels leave as 16-byte stores; FP32 NCDHW outputs (the last
  // convolution) are written directly (x contiguous per channel).
  const int z = z0 + wave;
  const int64_t V = (int64_t)a.D * a.H * a.W;
  constexpr int CST = NT * 16; // channels per staged voxel
  uint16_t* ot = reinterpret_cast<uint16_t*>(tile);
  if (a.out) __syncthreads(); // every wave is done with the halo tile
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int y = y0 + m;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int co = (tb + t) * 16 + row;
      const float bias = (a.bias && co < a.cout) ? a.bias[co] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int x = x0 + 4 * grp + i;
        float v = acc[m][t][i] + bias;
        if (a.relu && v < 0.f) v = 0.f;
        if (a.out) ot[((wave * BY + m) * BX + 4 * grp + i) * CST + t * 16 + row] =
            (uint16_t)(co < a.cout ? f2bf(v) : 0u);
        if (x >= a.W || y >= a.H || z >= a.D) continue;
        const int64_t vox = ((int64_t)z * a.H + y) * a.W + x;
        if (a.out_f32 && co < a.cout) a.out_f32[((int64_t)n * a.cout + co) * V + vox] = v;
      }
    }
  }
  if (!a.out) return;
  __syncthreads();
  // channels [tb*16, tb*16 + CST) of the brick, clipped to cout_p8, in
  // groups of 8 (cout_p8 and tb*16 are multiples of 8)
  const int c0 = tb * 16;
  const int ncg = (a.cout_p8 - c0 < CST ? a.cout_p8 - c0 : CST) / 8;
  for (int idx = threadIdx.x; idx < BX * BY * BZ * ncg; idx += 256) {
    const int g = idx % ncg, lv = idx / ncg;
    const int x = x0 + lv % BX, y = y0 + (lv / BX) % BY, zz = z0 + lv / (BX * BY);
    if (x >= a.W || y >= a.H || zz >= a.D) continue;
    const int64_t vox = ((int64_t)zz * a.H + y) * a.W + x;
    *(uint4*)(a.out + ((int64_t)n * V + vox) * a.cout_p8 + c0 + 8 * g) =
        *(const uint4*)(ot + lv * CST + 8 * g);
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: gw[co][ci][tap] = sum_{n,v} g[n,v,co] * in[n, v+off(tap), ci]
struct WgradArgs {
  ConvIn x;
  int D, H, W, N;
  const uint16_t* g; // output gradient (pre-activation), channels-last, stride coutp
  int coutp;
  int cinp;
  float* part;       // [gridDim.x][gridDim.y][MT*16][27][16]
  float* partb;      // [gridDim.x][MT*16] (blockIdx.y == 0)
};

template <int MT>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a)
{
  __shared__ __attribute__((aligned(16))) uint16_t gT[MT * 16][256];
  __shared__ __attribute__((aligned(16))) uint16_t xT[3][16][HZs][HYs][BX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const int cig = blockIdx.y;
  const int nbx = (a.W + BX - 1) / BX, nby = (a.H + BY - 1) / BY;
  const int nbr = conv_bricks(a.D, a.H, a.W);
  const int64_t total = (int64_t)a.N * nbr;
  const int ng = a.coutp >> 3;
  f32x4 acc[MT][7];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int j = 0; j < 7; j++) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int64_t br = blockIdx.x; br < total; br += gridDim.x) {
    const int n = (int)(br / nbr), b = (int)(br % nbr);
    const int x0 = (b % nbx) * BX, y0 = ((b / nbx) % nby) * BY, z0 = (b / (nbx * nby)) * BZ;
    __syncthreads();
    for (int t = threadIdx.x; t < 256 * ng; t += 256) {
      const int c8 = t % ng, v = t / ng;
      const int x = x0 + (v & 15), y = y0 + ((v >> 4) & 3), z = z0 + (v >> 6);
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (x < a.W && y < a.H && z < a.D)
        val = *(const uint4*)(a.g + ((((int64_t)n * a.D + z) * a.H + y) * a.W + x) * a.coutp + 8 * c8);
      const uint32_t w4[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        gT[8 * c8 + 2 * j][v] = (uint16_t)(w4[j] & 0xffffu);
        gT[8 * c8 + 2 * j + 1][v] = (uint16_t)(w4[j] >> 16);
      }
    }
    for (int t = threadIdx.x; t < (MT * 16 - a.coutp) * 256; t += 256)
      gT[a.coutp + (t >> 8)][t & 255] = 0;
    for (int t = threadIdx.x; t < HVOX * 2; t += 256) {
      const int c8 = t & 1, v = t >> 1;
      const int hx = v % HXs, hy = (v / HXs) % HYs, hz = v / (HXs * HYs);
      const int c = cig * 16 + 8 * c8;
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (c < a.cinp) val = load8(a.x, n, a.D, a.H, a.W, z0 + hz - 1, y0 + hy - 1, x0 + hx - 1, c);
      const uint32_t w4[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        const int xs = hx - dx;
        if (xs < 0 || xs >= BX) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          xT[dx][8 * c8 + 2 * j][hz][hy][xs] = (uint16_t)(w4[j] & 0xffffu);
          xT[dx][8 * c8 + 2 * j + 1][hz][hy][xs] = (uint16_t)(w4[j] >> 16);
        }
      }
    }
    __syncthreads();
    if (cig == 0 && threadIdx.x < a.coutp)
      for (int v = 0; v < 256; v++) bsum += bf2f(gT[threadIdx.x][v]);
#pragma unroll
    for (int kb = 0; kb < 8; kb++) {
      // voxels 32 kb + 8 grp + j: row r = 2 kb + grp/2 (y = r % 4, z = r / 4), x = 8 (grp % 2) + j
      const int r = 2 * kb + (grp >> 1), xo = 8 * (grp & 1);
      const int y = r & 3, z = r >> 2;
      bf16x8 av[MT];
#pragma unroll
      for (int m = 0; m < MT; m++)
        av[m] = *(const bf16x8*)&gT[m * 16 + row][32 * kb + 8 * grp];
#pragma unroll
      for (int j = 0; j < 7; j++) {
        const int tap = wave + 4 * j;
        if (tap < 27) {
          const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
          const bf16x8 bv = *(const bf16x8*)&xT[dx][row][z + dz][y + dy][xo];
#pragma unroll
          for (int m = 0; m < MT; m++)
            acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv, acc[m][j], 0, 0, 0);
        }
      }
    }
  }
  float* P = a.part + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * (MT * 16 * 27 * 16);
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int j = 0; j < 7; j++) {
      const int tap = wave + 4 * j;
      if (tap < 27)
#pragma unroll
        for (int i = 0; i < 4; i++) P[((m * 16 + 4 * grp + i) * 27 + tap) * 16 + row] = acc[m][j][i];
    }
  if (cig == 0 && threadIdx.x < MT * 16) a.partb[(int64_t)blockIdx.x * MT * 16 + threadIdx.x] = bsum;
}

// fixed-order sum of the partials into the FP32 gradient [co][ci][27], [co]
__global__ void k_wgrad_reduce(const float* __restrict__ part, const float* __restrict__ partb,
                               int nb, int ncig, int mt16, int co_n, int c0, int c0p, int c1,
                               float* __restrict__ gw, float* __restrict__ gb)
{
  const int ci_n = c0 + c1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = (int64_t)co_n * ci_n * 27;
  if (i < nw) {
    const int tap = (int)(i % 27);
    const int ci = (int)((i / 27) % ci_n), co = (int)(i / (27 * (int64_t)ci_n));
    const int pci = ci < c0 ? ci : c0p + (ci - c0);
    const int cg = pci >> 4, cl = pci & 15;
    float s = 0.f;
    for (int b = 0; b < nb; b++)
      s += part[(((int64_t)b * ncig + cg) * mt16 + co) * 27 * 16 + tap * 16 + cl];
    gw[i] = s;
  } else if (i < nw + co_n) {
    const int co = (int)(i - nw);
    float s = 0.f;
    for (int b = 0; b < nb; b++) s += partb[(int64_t)b * mt16 + co];
    gb[co] = s;
  }
}

// ---------------------------------------------------------------------------
// bf16 MFMA operands from the FP32 weights of one convolution.
// Forward: row co, K index (chunk q of <= 32 input channels: block 27 q +
// (tap*cc + ci - 32 q)/32).  Data gradient: row = padded input channel,
// K over the output channels with the tap flipped (26 - tap).
__global__ void k_pack_w(const float* __restrict__ w, int co_n, int c0, int c0p, int c1, int c1p,
                         int cop, uint16_t* __restrict__ wp, int kb, uint16_t* __restrict__ wt,
                         int kbt)
{
  const int ci_n = c0 + c1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)co_n * ci_n * 27) return;
  const int tap = (int)(i % 27);
  const int ci = (int)((i / 27) % ci_n), co = (int)(i / (27 * (int64_t)ci_n));
  const int pci = ci < c0 ? ci : c0p + (ci - c0);
  const int cinp = c0p + c1p;
  const uint16_t v = (uint16_t)f2bf(w[i]);
  {
    const int q = pci >> 5, cc = min(32, cinp - 32 * q);
    const int kl = tap * cc + (pci - 32 * q);
    wp[(int64_t)co * kb * 32 + (27 * q + kl / 32) * 32 + (kl & 31)] = v;
  }
  {
    const int q = co >> 5, cc = min(32, cop - 32 * q);
    const int kl = (26 - tap) * cc + (co - 32 * q);
    wt[(int64_t)pci * kbt * 32 + (27 * q + kl / 32) * 32 + (kl & 31)] = v;
  }
}

__global__ void k_fill_u16(uint16_t* __restrict__ p, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// element-wise kernels on channels-last bf16 tensors (8 channels per thread)

// FP32 NCDHW (c channels) -> bf16 channels-last (cp channels, zero padded)
__global__ void k_pack_in(const float* __restrict__ x, int c, int cp, int64_t V, int64_t nvox,
                          uint16_t* __restrict__ out, const double* __restrict__ skip)
{
  if (skip && *skip != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = cp >> 3;
  if (i >= nvox * g) return;
  const int c8 = (int)(i % g);
  const int64_t nv = i / g, n = nv / V, v = nv % V;
  float f[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int ch = 8 * c8 + j;
    f[j] = ch < c ? x[(n * c + ch) * V + v] : 0.f;
  }
  *(uint4*)(out + nv * cp + 8 * c8) = pack8(f);
}

// g = 0 where the layer's output y (ReLU) is not positive
__global__ void k_relu_mask(uint16_t* __restrict__ g, const uint16_t* __restrict__ y, int64_t n8)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const uint4 gv = ((const uint4*)g)[i];
  const uint4 yv = ((const uint4*)y)[i];
  const uint32_t gw[4] = {gv.x, gv.y, gv.z, gv.w}, yw[4] = {yv.x, yv.y, yv.z, yv.w};
  uint32_t o[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    uint32_t m = 0;
    const uint32_t lo = yw[j] & 0xffffu, hi = yw[j] >> 16;
    if (lo != 0 && !(lo & 0x8000u)) m |= 0xffffu;
    if (hi != 0 && !(hi & 0x8000u)) m |= 0xffff0000u;
    o[j] = gw[j] & m;
  }
  ((uint4*)g)[i] = make_uint4(o[0], o[1], o[2], o[3]);
}

// out (coarse, stride cn) = sum of the 8 children of channels [0, cn) of gcat
// (fine, stride cs); D, H, W coarse
__global__ void k_up_bwd(const uint16_t* __restrict__ gcat, int cs, int cn, int N, int D, int H,
                         int W, uint16_t* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = cn >> 3;
  const int64_t V = (int64_t)D * H * W;
  if (i >= (int64_t)N * V * g) return;
  const int c8 = (int)(i % g);
  const int64_t nv = i / g, n = nv / V, v = nv % V;
  const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, f[8];
  for (int q = 0; q < 8; q++) {
    const int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
    unpack8(*(const uint4*)(gcat + (((n * 2 * D + zz) * 2 * H + yy) * 2 * W + xx) * cs + 8 * c8), f);
    for (int j = 0; j < 8; j++) s[j] += f[j];
  }
  *(uint4*)(out + nv * cn + 8 * c8) = pack8(s);
}

// out[v][0..cn) = src[v][off..off+cn)  (src stride cs)
__global__ void k_take(const uint16_t* __restrict__ src, int cs, int off, int cn, int64_t nvox,
                       uint16_t* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = cn >> 3;
  if (i >= nvox * g) return;
  const int c8 = (int)(i % g);
  const int64_t v = i / g;
  *(uint4*)(out + v * cn + 8 * c8) = *(const uint4*)(src + v * cs + off + 8 * c8);
}

// g (fine) += gpool[parent] / 8  (stride cp; D, H, W fine)
__global__ void k_pool_bwd_add(const uint16_t* __restrict__ gpool, int cp, int N, int D, int H,
                               int W, uint16_t* __restrict__ g)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int gg = cp >> 3;
  const int64_t V = (int64_t)D * H * W;
  if (i >= (int64_t)N * V * gg) return;
  const int c8 = (int)(i % gg);
  const int64_t nv = i / gg, n = nv / V, v = nv % V;
  const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
  const int d = D >> 1, h = H >> 1, w = W >> 1;
  float a[8], b[8];
  unpack8(*(const uint4*)(g + nv * cp + 8 * c8), a);
  unpack8(*(const uint4*)(gpool + (((n * d + (z >> 1)) * h + (y >> 1)) * w + (x >> 1)) * cp + 8 * c8), b);
  for (int j = 0; j < 8; j++) a[j] += 0.125f * b[j];
  *(uint4*)(g + nv * cp + 8 * c8) = pack8(a);
}

// Weighted MSE loss 0.5*sum_n w_n sum_v (y-t)^2 / N (w_n = 1 without
// weights): gradient (bf16 channels-last, cop channels) and per-block loss
// partials; y, t FP32 NCDHW with co channels.  With dy != null the gradient
// is dy itself (the caller's loss).
__global__ __launch_bounds__(256) void k_out_grad(const float* __restrict__ y,
                                                  const float* __restrict__ t,
                                                  const float* __restrict__ wts,
                                                  const float* __restrict__ dy, int co, int cop,
                                                  int64_t V, int64_t nvox, float inv_n,
                                                  uint16_t* __restrict__ g, double* __restrict__ part)
{
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = i / V, v = i % V;
    for (int c8 = 0; c8 < cop; c8 += 8) {
      float f[8];
      for (int j = 0; j < 8; j++) {
        const int ch = c8 + j;
        f[j] = 0.f;
        if (ch >= co) continue;
        const int64_t k = (n * co + ch) * V + v;
        if (dy) {
          f[j] = dy[k];
        } else {
          const float d = y[k] - t[k];
          const float wn = wts ? wts[n] * inv_n : inv_n;
          f[j] = d * wn;
          s += 0.5 * (double)d * d * wn;
        }
      }
      *(uint4*)(g + i * cop + c8) = pack8(f);
    }
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

__global__ void k_zero_f(float* __restrict__ p, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// ---------------------------------------------------------------------------
static inline unsigned grid1(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
static inline int pad8(int c) { return (c + 7) & ~7; }
static inline int pad16(int c) { return (c + 15) & ~15; }
// 32-wide K blocks of an input of cp channels (chunks of 32, 27 taps each)
static inline int kblocks(int cp) { return 27 * (cp / 32) + (27 * (cp % 32) + 31) / 32; }

constexpr int kMaxConvCh = 64;      // output channels of a convolution
constexpr int kMaxConvIn = 128;     // padded input channels of a convolution
constexpr int kWgradBlocks = 512;   // weight-gradient workgroups (all channel groups)

template <int NT>
static void launch_conv(const ConvArgs& a, int N, hipStream_t s, int ztiles = 1)
{
  const int cin_p = a.x.c0p + a.x.c1p;
  const size_t halo = sizeof(uint4) * HVOX * ((cin_p < CCMAX ? cin_p : CCMAX) / 8);
  const size_t stage = a.out ? sizeof(uint16_t) * BX * BY * BZ * NT * 16 : 0;
  k_conv<NT><<<dim3(conv_bricks(a.D, a.H, a.W), N, ztiles), 256, halo > stage ? halo : stage, s>>>(
      a);
}

// nt output tiles of 16 channels: one block runs all of them over its brick,
// unless the grid of bricks is too small to fill the GPU (the coarse levels),
// where each block takes one tile (blockIdx.z) and the bricks are loaded once
// per tile instead
static void conv_dispatch(const ConvArgs& a, int nt, int N, hipStream_t s)
{
  if (nt > 1 && (int64_t)conv_bricks(a.D, a.H, a.W) * N < 512) {
    launch_conv<1>(a, N, s, nt);
    return;
  }
  switch (nt) {
  case 1: launch_conv<1>(a, N, s); break;
  case 2: launch_conv<2>(a, N, s); break;
  case 3: launch_conv<3>(a, N, s); break;
  case 4: launch_conv<4>(a, N, s); break;
  case 5: launch_conv<5>(a, N, s); break;
  case 6: launch_conv<6>(a, N, s); break;
  case 7: launch_conv<7>(a, N, s); break;
  default: launch_conv<8>(a, N, s); break;
  }
}

// ---------------------------------------------------------------------------
UNet::~UNet() { release(); }

void UNet::release()
{
  for (void* p : bufs) hipFree(p);
  bufs.clear();
  hipFree(params); hipFree(grads); hipFree(m1); hipFree(m2);
  hipFree(wpack); hipFree(wtpack); hipFree(wpart); hipFree(part);
  params = grads = m1 = m2 = nullptr;
  wpack = wtpack = nullptr;
  wpart = nullptr;
  part = nullptr;
  cap_N = cap_D = cap_H = cap_W = 0;
}

int UNet::init(const ifem_gpu_unet_config& c, hipStream_t s)
{
  cfg = c;
  if (c.levels < 1 || c.levels > kMaxLevels) return fail(IFEM_GPU_EINVAL, "unet: levels must be 1..4");
  if (c.in_channels < 1 || c.out_channels < 1 || c.in_channels > kMaxConvCh ||
      c.out_channels > kMaxConvCh)
    return fail(IFEM_GPU_EINVAL, "unet: in/out channels must be 1..64");
  for (int l = 0; l < c.levels; l++)
    if (c.channels[l] < 1 || c.channels[l] > kMaxConvCh)
      return fail(IFEM_GPU_EINVAL, "unet: channels per level must be 1..64");
  convs.clear();
  int64_t off = 0, wpo = 0, wto = 0;
  auto add = [&](int c0, int c1, int co, bool relu) {
    Conv cv;
    cv.c0 = c0; cv.c1 = c1; cv.c0p = pad8(c0); cv.c1p = c1 ? pad8(c1) : 0;
    cv.co = co; cv.cop = pad8(co); cv.relu = relu;
    cv.w = off; off += (int64_t)co * (c0 + c1) * 27;
    cv.b = off; off += co;
    cv.kb = kblocks(cv.c0p + cv.c1p);
    cv.kbt = kblocks(cv.cop);
    cv.wp = wpo; wpo += (int64_t)pad16(co) * cv.kb * 32;
    cv.wt = wto; wto += (int64_t)pad16(cv.c0p + cv.c1p) * cv.kbt * 32;
    convs.push_back(cv);
  };
  const int L = c.levels;
  int prev = c.in_channels;
  for (int l = 0; l < L; l++) {
    add(prev, 0, c.channels[l], true);
    add(c.channels[l], 0, c.channels[l], true);
    prev = c.channels[l];
  }
  for (int l = L - 2; l >= 0; l--) {
    add(c.channels[l + 1], c.channels[l], c.channels[l], true);
    add(c.channels[l], 0, c.channels[l], true);
  }
  add(c.channels[0], 0, c.out_channels, false);
  for (const Conv& cv : convs)
    if (cv.c0p + cv.c1p > kMaxConvIn) return fail(IFEM_GPU_EINVAL, "unet: too many input channels");
  nparams = off;
  wpack_n = wpo;
  wtpack_n = wto;
  cinp = pad8(c.in_channels);
  coutp = pad8(c.out_channels);
  stream = s;
  int64_t wpart_max = 0;
  for (const Conv& cv : convs)
    wpart_max = std::max<int64_t>(wpart_max, (int64_t)kWgradBlocks * pad16(cv.co) * 27 * 16 +
                                                 (int64_t)kWgradBlocks * pad16(cv.co));
  wpart_n = wpart_max;
  IFEM_HIP(hipMalloc(&params, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&grads, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&m1, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&m2, sizeof(float) * nparams));
  IFEM_HIP(hipMalloc(&wpack, sizeof(uint16_t) * wpack_n));
  IFEM_HIP(hipMalloc(&wtpack, sizeof(uint16_t) * wtpack_n));
  IFEM_HIP(hipMalloc(&wpart, sizeof(float) * wpart_n));
  IFEM_HIP(hipMalloc(&part, sizeof(double) * 1024));
  k_zero_f<<<grid1(nparams), 256, 0, s>>>(m1, nparams);
  k_zero_f<<<grid1(nparams), 256, 0, s>>>(m2, nparams);
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
    const int64_t nw = (int64_t)cv.co * cv.ci() * 27;
    const double sd = std::sqrt(2.0 / (cv.ci() * 27.0)) * (i + 1 == convs.size() ? 0.1 : 1.0);
    for (int64_t k = 0; k < nw; k += 2) {
      const double u1 = unif(), u2 = unif();
      const double r = std::sqrt(-2.0 * std::log(u1));
      h[cv.w + k] = (float)(sd * r * std::cos(2 * M_PI * u2));
      if (k + 1 < nw) h[cv.w + k + 1] = (float)(sd * r * std::sin(2 * M_PI * u2));
    }
  }
  IFEM_HIP(hipMemcpyAsync(params, h.data(), sizeof(float) * nparams, hipMemcpyHostToDevice, s));
  IFEM_HIP(hipStreamSynchronize(s));
  step = 0;
  packed = false;
  return IFEM_GPU_OK;
}

int UNet::pack()
{
  if (packed) return IFEM_GPU_OK;
  k_fill_u16<<<grid1(wpack_n), 256, 0, stream>>>(wpack, wpack_n);
  k_fill_u16<<<grid1(wtpack_n), 256, 0, stream>>>(wtpack, wtpack_n);
  for (const Conv& cv : convs) {
    const int64_t nw = (int64_t)cv.co * cv.ci() * 27;
    k_pack_w<<<grid1(nw), 256, 0, stream>>>(params + cv.w, cv.co, cv.c0, cv.c0p, cv.c1, cv.c1p,
                                             cv.cop, wpack + cv.wp, cv.kb, wtpack + cv.wt, cv.kbt);
  }
  IFEM_HIP(hipGetLastError());
  packed = true;
  return IFEM_GPU_OK;
}

int UNet::ensure(int N, int D, int H, int W)
{
  if (N <= cap_N && D == cap_D && H == cap_H && W == cap_W) return IFEM_GPU_OK;
  if (N < 1 || D < 1 || H < 1 || W < 1) return fail(IFEM_GPU_EINVAL, "unet: empty lattice");
  const int L = cfg.levels;
  for (int l = 1; l < L; l++)
    if ((D >> (l - 1)) % 2 || (H >> (l - 1)) % 2 || (W >> (l - 1)) % 2)
      return fail(IFEM_GPU_EUNSUPPORTED, "unet: grid sizes must be divisible by 2^(levels-1)");
  for (void* p : bufs) hipFree(p);
  bufs.clear();
  bool ok = true;
  auto alloc = [&](int64_t n) -> uint16_t* {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(uint16_t) * (n > 0 ? n : 1)) != hipSuccess) { ok = false; return nullptr; }
    bufs.push_back(p);
    return (uint16_t*)p;
  };
  int64_t gmax = 0, catmax = 0;
  for (int l = 0; l < L; l++) {
    dims[l][0] = D >> l; dims[l][1] = H >> l; dims[l][2] = W >> l;
    const int64_t nv = (int64_t)N * dims[l][0] * dims[l][1] * dims[l][2];
    const int cp = pad8(cfg.channels[l]);
    const int cpn = l + 1 < L ? pad8(cfg.channels[l + 1]) : 0;
    const int cpp = l > 0 ? pad8(cfg.channels[l - 1]) : 0;
    enc_a[l] = alloc(nv * cp);
    enc_b[l] = alloc(nv * cp);
    dec_a[l] = l < L - 1 ? alloc(nv * cp) : nullptr;
    dec_b[l] = l < L - 1 ? alloc(nv * cp) : nullptr;
    g_skip[l] = l < L - 1 ? alloc(nv * cp) : nullptr;
    int cm = std::max(std::max(cp, cpn), cpp);
    if (l == 0) cm = std::max(cm, std::max(cinp, coutp));
    gmax = std::max<int64_t>(gmax, nv * cm);
    catmax = std::max<int64_t>(catmax, nv * (cp + cpn));
  }
  const int64_t nv0 = (int64_t)N * D * H * W;
  xin = alloc(nv0 * cinp);
  gb1 = alloc(gmax);
  gb2 = alloc(gmax);
  gcat = alloc(catmax);
  gout = alloc(nv0 * coutp);
  void* o = nullptr;
  if (ok && hipMalloc(&o, sizeof(float) * nv0 * cfg.out_channels) == hipSuccess) bufs.push_back(o);
  else ok = false;
  outbuf = (float*)o;
  if (!ok) return fail(IFEM_GPU_ENOMEM, "unet: out of device memory");
  cap_N = N; cap_D = D; cap_H = H; cap_W = W;
  return IFEM_GPU_OK;
}

void UNet::conv_fwd(const Conv& cv, const ConvIn& in, int N, int d, int h, int w, uint16_t* out,
                    float* out_f32, const double* skip)
{
  ConvArgs a;
  a.skip = skip;
  a.x = in;
  a.D = d; a.H = h; a.W = w;
  a.wp = wpack + cv.wp;
  a.kb_total = cv.kb;
  a.bias = params + cv.b;
  a.cout = cv.co;
  a.cout_p8 = cv.cop;
  a.out = out;
  a.out_f32 = out_f32;
  a.relu = cv.relu && cfg.activation == 0 ? 1 : 0;
  conv_dispatch(a, pad16(cv.co) / 16, N, stream);
}

void UNet::conv_dgrad(const Conv& cv, const uint16_t* g, int N, int d, int h, int w, uint16_t* gin)
{
  ConvArgs a;
  a.x.in0 = g; a.x.c0 = cv.co; a.x.c0p = cv.cop; a.x.mode0 = 0;
  a.x.in1 = nullptr; a.x.c1 = 0; a.x.c1p = 0;
  a.D = d; a.H = h; a.W = w;
  a.wp = wtpack + cv.wt;
  a.kb_total = cv.kbt;
  a.bias = nullptr;
  const int cinp_ = cv.c0p + cv.c1p;
  a.cout = cinp_;
  a.cout_p8 = cinp_;
  a.out = gin;
  a.out_f32 = nullptr;
  a.relu = 0;
  conv_dispatch(a, pad16(cinp_) / 16, N, stream);
}

void UNet::conv_wgrad(const Conv& cv, const ConvIn& in, const uint16_t* g, int N, int d, int h,
                      int w)
{
  const int mt = pad16(cv.co) / 16;
  const int cinp_ = cv.c0p + cv.c1p;
  const int ncig = (cinp_ + 15) / 16;
  const int64_t bricks = (int64_t)N * conv_bricks(d, h, w);
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(bricks, kWgradBlocks / ncig));
  WgradArgs a;
  a.x = in;
  a.D = d; a.H = h; a.W = w; a.N = N;
  a.g = g;
  a.coutp = cv.cop;
  a.cinp = cinp_;
  a.part = wpart;
  a.partb = wpart + (int64_t)nb * ncig * mt * 16 * 27 * 16;
  const dim3 grid(nb, ncig);
  switch (mt) {
  case 1: k_wgrad<1><<<grid, 256, 0, stream>>>(a); break;
  case 2: k_wgrad<2><<<grid, 256, 0, stream>>>(a); break;
  case 3: k_wgrad<3><<<grid, 256, 0, stream>>>(a); break;
  default: k_wgrad<4><<<grid, 256, 0, stream>>>(a); break;
  }
  const int64_t nw = (int64_t)cv.co * cv.ci() * 27 + cv.co;
  k_wgrad_reduce<<<grid1(nw), 256, 0, stream>>>(a.part, a.partb, nb, ncig, mt * 16, cv.co, cv.c0,
                                                cv.c0p, cv.c1, grads + cv.w, grads + cv.b);
}

static ConvIn cin_of(const uint16_t* p, int c, int mode = 0, const uint16_t* p1 = nullptr, int c1 = 0)
{
  ConvIn in;
  in.in0 = p; in.c0 = c; in.c0p = pad8(c); in.mode0 = mode;
  in.in1 = p1; in.c1 = c1; in.c1p = p1 ? pad8(c1) : 0;
  return in;
}

// Forward pass (bf16 activations, FP32 accumulation); x FP32 NCDHW; returns
// the FP32 NCDHW output.
const float* UNet::forward(const float* x, int N, int D, int H, int W, const double* skip)
{
  const int L = cfg.levels;
  hipStream_t s = stream;
  pack();
  const int64_t V0 = (int64_t)D * H * W;
  k_pack_in<<<grid1(N * V0 * (cinp / 8)), 256, 0, s>>>(x, cfg.in_channels, cinp, V0, N * V0, xin,
                                                       skip);
  int ic = 0;
  for (int l = 0; l < L; l++) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    const ConvIn in = l == 0 ? cin_of(xin, cfg.in_channels) : cin_of(enc_b[l - 1], cfg.channels[l - 1], 2);
    conv_fwd(convs[ic++], in, N, d, h, w, enc_a[l], nullptr, skip);
    conv_fwd(convs[ic++], cin_of(enc_a[l], cfg.channels[l]), N, d, h, w, enc_b[l], nullptr, skip);
  }
  const uint16_t* cur = enc_b[L - 1];
  for (int l = L - 2; l >= 0; l--) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    conv_fwd(convs[ic++], cin_of(cur, cfg.channels[l + 1], 1, enc_b[l], cfg.channels[l]), N, d, h,
             w, dec_a[l], nullptr, skip);
    conv_fwd(convs[ic++], cin_of(dec_a[l], cfg.channels[l]), N, d, h, w, dec_b[l], nullptr, skip);
    cur = dec_b[l];
  }
  conv_fwd(convs[ic], cin_of(cur, cfg.channels[0]), N, D, H, W, nullptr, outbuf, skip);
  return outbuf;
}

int UNet::backward(const float* target, const float* wts, const float* dy, int N, bool want_loss,
                   float* loss)
{
  const int L = cfg.levels;
  hipStream_t s = stream;
  const int D = dims[0][0], H = dims[0][1], W = dims[0][2];
  const int64_t V0 = (int64_t)D * H * W;
  const int nb = 256;
  if (dy) want_loss = false;
  k_out_grad<<<nb, 256, 0, s>>>(outbuf, target, wts, dy, cfg.out_channels, coutp, V0, N * V0,
                                1.0f / (float)N, gout, part);
  auto nvox = [&](int l) { return (int64_t)N * dims[l][0] * dims[l][1] * dims[l][2]; };
  auto mask = [&](uint16_t* g, const uint16_t* y, int l, int c) {
    if (cfg.activation != 0) return;
    const int64_t n8 = nvox(l) * pad8(c) / 8;
    k_relu_mask<<<grid1(n8), 256, 0, s>>>(g, y, n8);
  };
  const int nconv = (int)convs.size();
  // output convolution (no ReLU)
  const uint16_t* last = L > 1 ? dec_b[0] : enc_b[0];
  conv_wgrad(convs[nconv - 1], cin_of(last, cfg.channels[0]), gout, N, D, H, W);
  conv_dgrad(convs[nconv - 1], gout, N, D, H, W, gb1);
  // decoder stages in reverse of the forward order (l = 0 .. L-2)
  for (int l = 0; l <= L - 2; l++) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    const int cl = cfg.channels[l], cn = cfg.channels[l + 1];
    const int i1 = 2 * L + 2 * (L - 2 - l); // first convolution of decoder level l
    const Conv& c1 = convs[i1];
    const Conv& c2 = convs[i1 + 1];
    mask(gb1, dec_b[l], l, cl);
    conv_wgrad(c2, cin_of(dec_a[l], cl), gb1, N, d, h, w);
    conv_dgrad(c2, gb1, N, d, h, w, gb2);
    mask(gb2, dec_a[l], l, cl);
    const uint16_t* up_src = l + 1 <= L - 2 ? dec_b[l + 1] : enc_b[L - 1];
    conv_wgrad(c1, cin_of(up_src, cn, 1, enc_b[l], cl), gb2, N, d, h, w);
    conv_dgrad(c1, gb2, N, d, h, w, gcat);
    const int cs = pad8(cn) + pad8(cl);
    k_take<<<grid1(nvox(l) * pad8(cl) / 8), 256, 0, s>>>(gcat, cs, pad8(cn), pad8(cl), nvox(l),
                                                          g_skip[l]);
    k_up_bwd<<<grid1(nvox(l + 1) * pad8(cn) / 8), 256, 0, s>>>(
        gcat, cs, pad8(cn), N, dims[l + 1][0], dims[l + 1][1], dims[l + 1][2], gb1);
  }
  // encoder levels from the coarsest; g = gradient w.r.t. enc_b[l]
  uint16_t* g = gb1;
  for (int l = L - 1; l >= 0; l--) {
    const int d = dims[l][0], h = dims[l][1], w = dims[l][2];
    const int cl = cfg.channels[l];
    mask(g, enc_b[l], l, cl);
    conv_wgrad(convs[2 * l + 1], cin_of(enc_a[l], cl), g, N, d, h, w);
    conv_dgrad(convs[2 * l + 1], g, N, d, h, w, gb2);
    mask(gb2, enc_a[l], l, cl);
    const ConvIn in = l == 0 ? cin_of(xin, cfg.in_channels) : cin_of(enc_b[l - 1], cfg.channels[l - 1], 2);
    conv_wgrad(convs[2 * l], in, gb2, N, d, h, w);
    if (l > 0) {
      conv_dgrad(convs[2 * l], gb2, N, d, h, w, gb1);
      const int cp = pad8(cfg.channels[l - 1]);
      k_pool_bwd_add<<<grid1(nvox(l - 1) * cp / 8), 256, 0, s>>>(
          gb1, cp, N, dims[l - 1][0], dims[l - 1][1], dims[l - 1][2], g_skip[l - 1]);
      g = g_skip[l - 1];
    }
  }
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
  packed = false;
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
  n.packed = false;
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
  rc = n.backward(t, w, nullptr, N, true, loss);
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
  rc = n.backward(t, w, nullptr, N, loss != nullptr, loss);
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
  rc = n.backward(nullptr, nullptr, dy, N, false, nullptr);
  if (rc) return rc;
  if (grad_host) {
    IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
    IFEM_HIP(hipMemcpy(grad_host, n.grads, sizeof(float) * n.nparams, hipMemcpyDeviceToHost));
  }
  if (lr > 0.f) return n.adam(lr);
  return IFEM_GPU_OK;
}

} // extern "C"
