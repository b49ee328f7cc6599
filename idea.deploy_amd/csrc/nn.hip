// nn.hip -- kernels of the warm-start network (a small U-Net) for gfx950:
// 3x3 convolutions as implicit GEMMs on BF16 MFMA (v_mfma_f32_16x16x16_bf16)
// with FP32 accumulation, forward, backward-data and backward-weight, and
// the elementwise pieces (ReLU backward, 2x2 average pool, nearest 2x
// upsample, channel concat/split, MSE loss, Adam, casts, bilinear resize to
// the lattice). No reference counterpart: idpy has no learned component.
//
// Tensors are NHWC. Activations and weights used by the MFMA are bf16
// (stored as uint16 bit patterns); master weights, gradients and the loss
// are FP32. Conv weights are [Cout][9*Cin] with k = tap*Cin + ci,
// tap = 3*ky + kx. When Cin is a multiple of 4 a lane's MFMA K slice is 4
// channels of one tap (vector loads); other channel counts are handled
// element by element.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "internal.h"

namespace idmx {

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(uint16_t b) {
  return __uint_as_float(((uint32_t)b) << 16);
}
// round to nearest even
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct ConvShape {
  int N, H, W, Cin, Cout;
};

// ---------------------------------------------------------------------------
// Forward 3x3 convolution, stride 1, zero padding 1:
//   out[p][co] = act(bias[co] + sum_k A[p][k] * Wt[co][k]),
//   A[p][k] = in[n, y+ky-1, x+kx-1, ci] (im2col, never materialised).
// One wave computes a 16 (pixels) x 16 (output channels) tile; K advances
// 16 per MFMA (4 per lane: 4 consecutive input channels of one tap).
// MFMA 16x16x16 lane layout: A row = lane%16, k = 4*(lane/16)+j; B column =
// lane%16, same k; D column = lane%16, rows 4*(lane/16)+j.
// out_f32 (optional) receives the pre-activation FP32 values.
__global__ void __launch_bounds__(256)
k_conv3x3_fwd(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
              const float* __restrict__ bias, uint16_t* __restrict__ out,
              float* __restrict__ out_f32, ConvShape s, int relu) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int P = s.N * s.H * s.W;
  const int ptiles = (P + 15) / 16;
  const int ctiles = (s.Cout + 15) / 16;
  if (wave >= ptiles * ctiles) return;
  const int pt = wave % ptiles, ct = wave / ptiles;
  const int r = lane & 15, kq = lane >> 4;
  const int p = pt * 16 + r;  // A row (pixel)
  const bool pvalid = p < P;
  const int n = pvalid ? p / (s.H * s.W) : 0;
  const int yx = pvalid ? p % (s.H * s.W) : 0;
  const int y = yx / s.W, x = yx % s.W;
  const int co = ct * 16 + r;  // B column
  const bool cvalid = co < s.Cout;
  const int K = 9 * s.Cin;
  const bool vec4 = (s.Cin & 3) == 0;
  v4f32 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int k = k0 + 4 * kq;
    v4i16 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
    if (vec4) {
      // Cin % 4 == 0: the 4 k values of a lane are 4 channels of one tap
      if (k < K) {
        const int tap = k / s.Cin, ci = k % s.Cin;
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        if (pvalid && yy >= 0 && yy < s.H && xx >= 0 && xx < s.W) {
          const uint16_t* src = in + (((int64_t)n * s.H + yy) * s.W + xx) * s.Cin + ci;
#pragma unroll
          for (int j = 0; j < 4; j++) a[j] = (short)src[j];
        }
        if (cvalid) {
          const uint16_t* wsrc = wt + (int64_t)co * K + k;
#pragma unroll
          for (int j = 0; j < 4; j++) b[j] = (short)wsrc[j];
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int kk = k + j;
        if (kk < K) {
          const int tap = kk / s.Cin, ci = kk % s.Cin;
          const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
          if (pvalid && yy >= 0 && yy < s.H && xx >= 0 && xx < s.W)
            a[j] = (short)in[(((int64_t)n * s.H + yy) * s.W + xx) * s.Cin + ci];
          if (cvalid) b[j] = (short)wt[(int64_t)co * K + kk];
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // D: column = lane%16 (output channel), rows 4*(lane/16)+j (pixels)
  const int oc = ct * 16 + (lane & 15);
  if (oc < s.Cout) {
    const float bv = bias ? bias[oc] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int op = pt * 16 + 4 * kq + j;
      if (op < P) {
        float v = acc[j] + bv;
        if (out_f32) out_f32[(int64_t)op * s.Cout + oc] = v;
        if (relu) v = fmaxf(v, 0.f);
        out[(int64_t)op * s.Cout + oc] = f2bf(v);
      }
    }
  }
}

// Backward-weight: dW[co][k] += sum_p dY[p][co] * A[p][k] over a chunk of
// pixels per wave (split over pixels, FP32 atomics into dW). MFMA rows are
// output channels, columns are k, the reduction runs over 16 pixels per
// instruction (4 per lane).
__global__ void __launch_bounds__(256)
k_conv3x3_bwd_w(const uint16_t* __restrict__ in, const uint16_t* __restrict__ dy,
                float* __restrict__ dw, ConvShape s, int pchunk) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int P = s.N * s.H * s.W;
  const int K = 9 * s.Cin;
  const int ctiles = (s.Cout + 15) / 16, ktiles = (K + 15) / 16;
  const int nchunks = (P + pchunk - 1) / pchunk;
  if (wave >= ctiles * ktiles * nchunks) return;
  int t = wave;
  const int kt = t % ktiles;
  t /= ktiles;
  const int ct = t % ctiles;
  const int ch = t / ctiles;
  const int r = lane & 15, kq = lane >> 4;
  const int co = ct * 16 + r;  // A row
  const int k = kt * 16 + r;   // B column
  const bool cvalid = co < s.Cout, kvalid = k < K;
  const int tap = kvalid ? k / s.Cin : 0, ci = kvalid ? k % s.Cin : 0;
  const int dy_ = tap / 3 - 1, dx_ = tap % 3 - 1;
  v4f32 acc = {0.f, 0.f, 0.f, 0.f};
  const int pbeg = ch * pchunk, pend = min(P, pbeg + pchunk);
  for (int p0 = pbeg; p0 < pend; p0 += 16) {
    v4i16 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int p = p0 + 4 * kq + j;
      if (p < pend) {
        if (cvalid) a[j] = (short)dy[(int64_t)p * s.Cout + co];
        if (kvalid) {
          const int n = p / (s.H * s.W), yx = p % (s.H * s.W);
          const int yy = yx / s.W + dy_, xx = yx % s.W + dx_;
          if (yy >= 0 && yy < s.H && xx >= 0 && xx < s.W)
            b[j] = (short)in[(((int64_t)n * s.H + yy) * s.W + xx) * s.Cin + ci];
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int kc = kt * 16 + (lane & 15);
  if (kc < K) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int oc = ct * 16 + 4 * kq + j;
      if (oc < s.Cout) atomicAdd(dw + (int64_t)oc * K + kc, acc[j]);
    }
  }
}

// Weights for the backward-data pass: the transposed, tap-flipped kernel
// Wd[ci][tap'*Cout + co] = Wt[co][(8 - tap')*Cin + ci], in bf16.
__global__ void k_conv_flip_weights(const float* __restrict__ w, uint16_t* __restrict__ wd,
                                    int Cin, int Cout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = Cin * 9 * Cout;
  if (i >= total) return;
  const int ci = i / (9 * Cout), rem = i % (9 * Cout);
  const int tp = rem / Cout, co = rem % Cout;
  wd[i] = f2bf(w[(int64_t)co * 9 * Cin + (8 - tp) * Cin + ci]);
}

// dX = dY * (Y > 0) on bf16 tensors (ReLU backward from the stored output).
__global__ void k_relu_bwd(const uint16_t* __restrict__ y, const uint16_t* __restrict__ dy,
                           uint16_t* __restrict__ dx, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dx[i] = bf2f(y[i]) > 0.f ? dy[i] : (uint16_t)0;
}

// db[c] = sum_p dY[p][c] (FP32)
__global__ void k_bias_grad(const uint16_t* __restrict__ dy, float* __restrict__ db, int P,
                            int C) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (int p = threadIdx.x; p < P; p += blockDim.x) s += bf2f(dy[(int64_t)p * C + c]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ float part[16];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += part[k];
    db[c] += t;
  }
}

// 2x2 average pool (H, W even) and its backward.
__global__ void k_pool2_fwd(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int N,
                            int H, int W, int C) {
  const int64_t total = (int64_t)N * (H / 2) * (W / 2) * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    int64_t t = i / C;
    const int xo = t % (W / 2);
    t /= (W / 2);
    const int yo = t % (H / 2);
    const int n = t / (H / 2);
    float s = 0.f;
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++)
        s += bf2f(x[(((int64_t)n * H + 2 * yo + dy) * W + 2 * xo + dx) * C + c]);
    y[i] = f2bf(0.25f * s);
  }
}

__global__ void k_pool2_bwd(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int N,
                            int H, int W, int C) {
  const int64_t total = (int64_t)N * H * W * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    int64_t t = i / C;
    const int x = t % W;
    t /= W;
    const int y = t % H;
    const int n = t / H;
    dx[i] = f2bf(0.25f * bf2f(dy[(((int64_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]));
  }
}

// Nearest 2x upsample of x (N, H, W, C) into channels [c0, c0 + C) of y
// (N, 2H, 2W, Cy), and its backward (sum of the 4 children).
__global__ void k_up2_fwd(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int N,
                          int H, int W, int C, int Cy, int c0) {
  const int64_t total = (int64_t)N * 2 * H * 2 * W * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    int64_t t = i / C;
    const int xo = t % (2 * W);
    t /= (2 * W);
    const int yo = t % (2 * H);
    const int n = t / (2 * H);
    y[(((int64_t)n * 2 * H + yo) * 2 * W + xo) * Cy + c0 + c] =
        x[(((int64_t)n * H + yo / 2) * W + xo / 2) * C + c];
  }
}

__global__ void k_up2_bwd(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int N,
                          int H, int W, int C, int Cy, int c0) {
  const int64_t total = (int64_t)N * H * W * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    int64_t t = i / C;
    const int x = t % W;
    t /= W;
    const int y = t % H;
    const int n = t / H;
    float s = 0.f;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        s += bf2f(dy[(((int64_t)n * 2 * H + 2 * y + a) * 2 * W + 2 * x + b) * Cy + c0 + c]);
    dx[i] = f2bf(s);
  }
}

// Copy channels: y[..., cy0 + c] = x[..., cx0 + c] for c < C (concat / split).
__global__ void k_copy_channels(const uint16_t* __restrict__ x, uint16_t* __restrict__ y,
                                int64_t P, int Cx, int cx0, int Cy, int cy0, int C) {
  const int64_t total = P * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    const int64_t p = i / C;
    y[p * Cy + cy0 + c] = x[p * Cx + cx0 + c];
  }
}

// Elementwise add of bf16 tensors into FP32 accumulation and back:
// y = bf16(a + b).
__global__ void k_add_bf16(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                           uint16_t* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// MSE loss of FP32 prediction vs FP32 target; grad = 2 (pred - target) / n
// in bf16; per-block partial losses in part.
__global__ void __launch_bounds__(256)
k_mse(const float* __restrict__ pred, const float* __restrict__ tgt, uint16_t* __restrict__ grad,
      float* __restrict__ part, int64_t n) {
  float s = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float d = pred[i] - tgt[i];
    s += d * d;
    grad[i] = f2bf(2.f * d / (float)n);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ float ps[4];
  if ((threadIdx.x & 63) == 0) ps[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (ps[0] + ps[1] + ps[2] + ps[3]) / (float)n;
}

// Adam on FP32 master weights; refreshes the bf16 copy; zeroes the gradient.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
    g[i] = 0.f;
  }
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

__global__ void k_bf16_to_f32(const uint16_t* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = bf2f(x[i]);
}

// Bilinear resize of one S x S FP32 image (pixel centres at (i + 0.5) S/L
// ... in lattice units) to an L x L FP64 lattice field, with a constant
// added (mass correction).
__global__ void k_resize_to_lattice(const float* __restrict__ img, int S, double* __restrict__ n,
                                    int L, double add) {
  const int64_t V = (int64_t)L * L;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int x = g % L, y = g / L;
    const float fx = ((x + 0.5f) * S) / L - 0.5f, fy = ((y + 0.5f) * S) / L - 0.5f;
    int x0 = (int)floorf(fx), y0 = (int)floorf(fy);
    const float ax = fx - x0, ay = fy - y0;
    const int x1 = min(max(x0 + 1, 0), S - 1), y1 = min(max(y0 + 1, 0), S - 1);
    x0 = min(max(x0, 0), S - 1);
    y0 = min(max(y0, 0), S - 1);
    const float v = (1 - ay) * ((1 - ax) * img[y0 * S + x0] + ax * img[y0 * S + x1]) +
                    ay * ((1 - ax) * img[y1 * S + x0] + ax * img[y1 * S + x1]);
    n[g] = (double)v + add;
  }
}

inline unsigned g1(int64_t n, int bs = 256) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + bs - 1) / bs, 65536));
}

}  // namespace idmx

using namespace idmx;

#define NCHECK(cond, msg)   \
  do {                      \
    if (!(cond)) {          \
      set_error(msg);       \
      return IDMX_E_ARG;    \
    }                       \
  } while (0)

extern "C" {

int idmx_nn_conv3x3_fwd(idmx_ctx_t c, const void* in, const void* wt, const float* bias,
                        void* out, float* out_f32, int N, int H, int W, int Cin, int Cout,
                        int relu) {
  NCHECK(c && in && wt && out && N > 0 && H > 0 && W > 0 && Cin > 0 && Cout > 0,
         "idmx_nn_conv3x3_fwd: invalid argument");
  ConvShape s{N, H, W, Cin, Cout};
  const int64_t waves = (int64_t)((N * H * W + 15) / 16) * ((Cout + 15) / 16);
  hipLaunchKernelGGL(k_conv3x3_fwd, dim3(g1(waves * 64)), dim3(256), 0, c->stream,
                     (const uint16_t*)in, (const uint16_t*)wt, bias, (uint16_t*)out, out_f32, s,
                     relu);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_conv3x3_bwd_data(idmx_ctx_t c, const void* dy, const float* w, void* wd_scratch,
                             void* dx, int N, int H, int W, int Cin, int Cout) {
  NCHECK(c && dy && w && wd_scratch && dx && N > 0 && H > 0 && W > 0 && Cin > 0 && Cout > 0,
         "idmx_nn_conv3x3_bwd_data: invalid argument");
  const int tot = Cin * 9 * Cout;
  hipLaunchKernelGGL(k_conv_flip_weights, dim3(g1(tot)), dim3(256), 0, c->stream, w,
                     (uint16_t*)wd_scratch, Cin, Cout);
  ConvShape s{N, H, W, Cout, Cin};
  const int64_t waves = (int64_t)((N * H * W + 15) / 16) * ((Cin + 15) / 16);
  hipLaunchKernelGGL(k_conv3x3_fwd, dim3(g1(waves * 64)), dim3(256), 0, c->stream,
                     (const uint16_t*)dy, (const uint16_t*)wd_scratch, (const float*)nullptr,
                     (uint16_t*)dx, (float*)nullptr, s, 0);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_conv3x3_bwd_weight(idmx_ctx_t c, const void* in, const void* dy, float* dw,
                               float* db, int N, int H, int W, int Cin, int Cout) {
  NCHECK(c && in && dy && dw, "idmx_nn_conv3x3_bwd_weight: invalid argument");
  ConvShape s{N, H, W, Cin, Cout};
  const int P = N * H * W;
  const int pchunk = 1024;
  const int64_t waves = (int64_t)((Cout + 15) / 16) * ((9 * Cin + 15) / 16) *
                        ((P + pchunk - 1) / pchunk);
  hipLaunchKernelGGL(k_conv3x3_bwd_w, dim3(g1(waves * 64)), dim3(256), 0, c->stream,
                     (const uint16_t*)in, (const uint16_t*)dy, dw, s, pchunk);
  if (db)
    hipLaunchKernelGGL(k_bias_grad, dim3(Cout), dim3(256), 0, c->stream, (const uint16_t*)dy,
                       db, P, Cout);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_relu_bwd(idmx_ctx_t c, const void* y, const void* dy, void* dx, int64_t n) {
  NCHECK(c && y && dy && dx && n >= 0, "idmx_nn_relu_bwd: invalid argument");
  hipLaunchKernelGGL(k_relu_bwd, dim3(g1(n)), dim3(256), 0, c->stream, (const uint16_t*)y,
                     (const uint16_t*)dy, (uint16_t*)dx, n);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_pool2(idmx_ctx_t c, const void* x, void* y, int N, int H, int W, int C, int bwd) {
  NCHECK(c && x && y && H % 2 == 0 && W % 2 == 0, "idmx_nn_pool2: invalid argument");
  const int64_t tot = bwd ? (int64_t)N * H * W * C : (int64_t)N * (H / 2) * (W / 2) * C;
  if (bwd)
    hipLaunchKernelGGL(k_pool2_bwd, dim3(g1(tot)), dim3(256), 0, c->stream, (const uint16_t*)x,
                       (uint16_t*)y, N, H, W, C);
  else
    hipLaunchKernelGGL(k_pool2_fwd, dim3(g1(tot)), dim3(256), 0, c->stream, (const uint16_t*)x,
                       (uint16_t*)y, N, H, W, C);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_up2(idmx_ctx_t c, const void* x, void* y, int N, int H, int W, int C, int Cy, int c0,
                int bwd) {
  NCHECK(c && x && y && c0 >= 0 && c0 + C <= Cy, "idmx_nn_up2: invalid argument");
  const int64_t tot = bwd ? (int64_t)N * H * W * C : (int64_t)N * 4 * H * W * C;
  if (bwd)
    hipLaunchKernelGGL(k_up2_bwd, dim3(g1(tot)), dim3(256), 0, c->stream, (const uint16_t*)x,
                       (uint16_t*)y, N, H, W, C, Cy, c0);
  else
    hipLaunchKernelGGL(k_up2_fwd, dim3(g1(tot)), dim3(256), 0, c->stream, (const uint16_t*)x,
                       (uint16_t*)y, N, H, W, C, Cy, c0);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_copy_channels(idmx_ctx_t c, const void* x, void* y, int64_t P, int Cx, int cx0,
                          int Cy, int cy0, int C) {
  NCHECK(c && x && y && cx0 + C <= Cx && cy0 + C <= Cy, "idmx_nn_copy_channels: invalid argument");
  hipLaunchKernelGGL(k_copy_channels, dim3(g1(P * C)), dim3(256), 0, c->stream,
                     (const uint16_t*)x, (uint16_t*)y, P, Cx, cx0, Cy, cy0, C);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_add(idmx_ctx_t c, const void* a, const void* b, void* y, int64_t n) {
  NCHECK(c && a && b && y, "idmx_nn_add: invalid argument");
  hipLaunchKernelGGL(k_add_bf16, dim3(g1(n)), dim3(256), 0, c->stream, (const uint16_t*)a,
                     (const uint16_t*)b, (uint16_t*)y, n);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_mse(idmx_ctx_t c, const float* pred, const float* tgt, void* grad, int64_t n,
                double* loss) {
  NCHECK(c && pred && tgt && grad && loss && n > 0, "idmx_nn_mse: invalid argument");
  const unsigned grid = std::min<unsigned>(g1(n), 1024);
  hipLaunchKernelGGL(k_mse, dim3(grid), dim3(256), 0, c->stream, pred, tgt, (uint16_t*)grad,
                     (float*)c->partials, n);
  IDMX_HIP(hipGetLastError());
  float h[1024];
  IDMX_HIP(hipMemcpyAsync(h, c->partials, sizeof(float) * grid, hipMemcpyDeviceToHost, c->stream));
  IDMX_HIP(hipStreamSynchronize(c->stream));
  double s = 0;
  for (unsigned k = 0; k < grid; k++) s += h[k];
  *loss = s;
  return IDMX_OK;
}

int idmx_nn_adam(idmx_ctx_t c, float* w, float* g, float* m, float* v, void* wb, int64_t n,
                 float lr, float b1, float b2, float eps, int t) {
  NCHECK(c && w && g && m && v && n > 0 && t >= 1, "idmx_nn_adam: invalid argument");
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(g1(n)), dim3(256), 0, c->stream, w, g, m, v, (uint16_t*)wb, n,
                     lr, b1, b2, eps, c1, c2);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_cast(idmx_ctx_t c, const void* x, void* y, int64_t n, int to_bf16) {
  NCHECK(c && x && y && n >= 0, "idmx_nn_cast: invalid argument");
  if (to_bf16)
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(g1(n)), dim3(256), 0, c->stream, (const float*)x,
                       (uint16_t*)y, n);
  else
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(g1(n)), dim3(256), 0, c->stream, (const uint16_t*)x,
                       (float*)y, n);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_nn_resize_to_lattice(idmx_ctx_t c, const float* img, int S, double* n, int L,
                              double add) {
  NCHECK(c && img && n && S > 0 && L > 0, "idmx_nn_resize_to_lattice: invalid argument");
  hipLaunchKernelGGL(k_resize_to_lattice, dim3(g1((int64_t)L * L)), dim3(256), 0, c->stream, img,
                     S, n, L, add);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

}  // extern "C"
