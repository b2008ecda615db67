// Learned per-sub-aperture centroider: MLP 64 -> 128 -> 128 -> 2 with GELU
// (tanh form), bf16 operands on the matrix cores (v_mfma_f32_16x16x32_bf16),
// FP32 accumulation, FP32 master weights and Adam.  No counterpart in the
// reference (a new centroider type next to libsutra's CoG family).
//
// One workgroup = 4 waves = a tile of 32 samples.  Forward:
//   Z1 = X W1^T + b1 (MFMA), H1 = gelu(Z1); Z2 = H1 W2^T + b2 (MFMA),
//   H2 = gelu(Z2); Y = H2 W3^T + b3 (VALU, 2 outputs)
// Backward of the mean squared error (dY = 2 (Y - T) / (2 n)):
//   dW3 = dY^T H2 (VALU); dZ2 = (dY W3) * gelu'(Z2); dW2 = dZ2^T H1 (MFMA);
//   dH1 = dZ2 W2 (MFMA, W2^T copy); dZ1 = dH1 * gelu'(Z1); dW1 = dZ1^T X (MFMA)
// Weight gradients of the tiles are summed with float atomics.
#include <cmath>

#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"
#include "shwfs_mfma.hpp"
#include "shwfs_mlp.hpp"

namespace shwfs {
namespace {

using namespace mlp;
static_assert(NPARAM == SHWFS_MLP_NPARAM, "parameter layout");
constexpr int TS = 32;          // samples per tile
constexpr int TPB = 256;        // 4 waves
using Lds = mlp::Lds<TS>;
constexpr size_t LDS_BYTES = Lds::bytes();
// work buffer: bf16 copies W1 (128x64), W2 (128x128), W2^T (128x128)
constexpr size_t WORK_ELEMS = (size_t)NH * NIN + 2 * (size_t)NH * NH;

// X tile of samples row0 .. row0+TS-1 (rows >= rows_valid zero) -> bf16 in LDS.
__device__ void load_rows(const float* __restrict__ x, const int* __restrict__ index, long row0,
                          int rows_valid, float x_scale, const Lds& L) {
  for (int i = threadIdx.x; i < TS * NIN; i += TPB) {
    const int s = i / NIN, c = i % NIN;
    float v = 0.0f;
    if (s < rows_valid) {
      const long r = index ? (long)index[row0 + s] : row0 + s;
      v = x[r * NIN + c] * x_scale;
    }
    L.x[s * SX + c] = (__bf16)v;
  }
}

__global__ __launch_bounds__(TPB) void mlp_prepare_kernel(const float* __restrict__ P,
                                                          __bf16* __restrict__ work) {
  __bf16* W1b = work;
  __bf16* W2b = W1b + NH * NIN;
  __bf16* W2t = W2b + NH * NH;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < NH * NH; i += gridDim.x * TPB) {
    if (i < NH * NIN) W1b[i] = (__bf16)P[OW1 + i];
    const int r = i / NH, c = i % NH;
    const __bf16 v = (__bf16)P[OW2 + i];
    W2b[i] = v;
    W2t[c * NH + r] = v;
  }
}

template <bool TRAIN>
__global__ __launch_bounds__(TPB) void mlp_tile_kernel(const float* __restrict__ P,
                                                       const __bf16* __restrict__ work,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ t,
                                                       const int* __restrict__ index, long n,
                                                       float x_scale, float t_scale,
                                                       float* __restrict__ y,
                                                       float* __restrict__ grad,
                                                       float* __restrict__ loss) {
  extern __shared__ char smem_raw[];
  const Lds L(smem_raw);
  const __bf16* W1b = work;
  const __bf16* W2b = W1b + NH * NIN;
  const __bf16* W2t = W2b + NH * NH;
  const long row0 = (long)blockIdx.x * TS;
  const int rows = (int)((n - row0) < TS ? (n - row0) : TS);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  load_rows(x, index, row0, rows, x_scale, L);
  mlp::tile_forward<TS, Bf16Weights, TRAIN>(P, Bf16Weights{W1b, W2b}, L, (int)threadIdx.x);
  if constexpr (!TRAIN) {
    if (tid < rows * NOUT) y[(row0 + tid / NOUT) * NOUT + tid % NOUT] = L.y[tid];
    return;
  } else {
    // ---- loss and dY = 2 (Y - T) / (n * NOUT)
    float part = 0.0f;
    if (tid < TS * NOUT) {
      const int s = tid / NOUT, o = tid % NOUT;
      float d = 0.0f;
      if (s < rows) {
        const long r = index ? (long)index[row0 + s] : row0 + s;
        d = L.y[tid] - t[r * NOUT + o] * t_scale;
      }
      part = d * d;
      L.dy[tid] = 2.0f * d / (float)(n * NOUT);
    }
    part = wave_sum(part);
    if (lane == 0 && w == 0) atomicAdd(loss, part / (float)(n * NOUT));
    __syncthreads();
    // ---- dW3, db3 (thread = (o, k)), and dZ2 = (dY W3) * gelu'(Z2) in place
    {
      const int o = tid / NH, k = tid % NH;  // TPB == NOUT * NH
      float g = 0.0f, gb = 0.0f;
      for (int s = 0; s < TS; ++s) {
        g += L.dy[s * NOUT + o] * gelu(L.z2[s * SZ + k]);
        gb += L.dy[s * NOUT + o];
      }
      atomicAdd(&grad[OW3 + o * NH + k], g);
      if (k == 0) atomicAdd(&grad[OB3 + o], gb);
    }
    __syncthreads();
    for (int i = tid; i < TS * NH; i += TPB) {
      const int s = i / NH, k = i % NH;
      float dh = 0.0f;
#pragma unroll
      for (int o = 0; o < NOUT; ++o) dh += L.dy[s * NOUT + o] * P[OW3 + o * NH + k];
      const float dz = dh * gelu_grad(L.z2[s * SZ + k]);
      L.z2[s * SZ + k] = dz;
      L.h2[s * SH + k] = (__bf16)dz;
    }
    __syncthreads();
    // db2
    if (tid < NH) {
      float gb = 0.0f;
      for (int s = 0; s < TS; ++s) gb += L.z2[s * SZ + tid];
      atomicAdd(&grad[OB2 + tid], gb);
    }
    // ---- dW2 (n2 x n1) = dZ2^T H1: m-tiles (n2) 2w, 2w+1; all 8 n-tiles (n1); K = TS
    {
      bf16x8 bfr[NH / 16];
#pragma unroll
      for (int nj = 0; nj < NH / 16; ++nj)
        bfr[nj] = ld8s(L.h1 + (8 * (lane >> 4)) * SH + 16 * nj + (lane & 15), SH);
#pragma unroll
      for (int mm = 0; mm < 2; ++mm) {
        const int mi = 2 * w + mm;
        const bf16x8 a = ld8s(L.h2 + (8 * (lane >> 4)) * SH + 16 * mi + (lane & 15), SH);
#pragma unroll
        for (int nj = 0; nj < NH / 16; ++nj) {
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          acc = mfma_bf16(a, bfr[nj], acc);
          const int col = 16 * nj + (lane & 15);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = 16 * mi + 4 * (lane >> 4) + v;
            atomicAdd(&grad[OW2 + row * NH + col], acc[v]);
          }
        }
      }
    }
    __syncthreads();  // h1 is overwritten below
    // ---- dH1 = dZ2 W2 (B from the W2^T copy); dZ1 = dH1 * gelu'(Z1) in place
#pragma unroll
    for (int mi = 0; mi < TS / 16; ++mi)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const int nj = 2 * w + jj;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < NH / 32; ++ks) {
          const bf16x8 a = ld8(L.h2 + (16 * mi + (lane & 15)) * SH + 32 * ks + 8 * (lane >> 4));
          const bf16x8 b = ld8(W2t + (16 * nj + (lane & 15)) * NH + 32 * ks + 8 * (lane >> 4));
          acc = mfma_bf16(a, b, acc);
        }
        const int col = 16 * nj + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * mi + 4 * (lane >> 4) + v;
          const float dz = acc[v] * gelu_grad(L.z1[row * SZ + col]);
          L.z1[row * SZ + col] = dz;
          L.h1[row * SH + col] = (__bf16)dz;
        }
      }
    __syncthreads();
    if (tid < NH) {
      float gb = 0.0f;
      for (int s = 0; s < TS; ++s) gb += L.z1[s * SZ + tid];
      atomicAdd(&grad[OB1 + tid], gb);
    }
    // ---- dW1 (n1 x in) = dZ1^T X: m-tiles 2w, 2w+1 (n1), 4 n-tiles (in), K = TS
    {
      bf16x8 bfr[NIN / 16];
#pragma unroll
      for (int nj = 0; nj < NIN / 16; ++nj)
        bfr[nj] = ld8s(L.x + (8 * (lane >> 4)) * SX + 16 * nj + (lane & 15), SX);
#pragma unroll
      for (int mm = 0; mm < 2; ++mm) {
        const int mi = 2 * w + mm;
        const bf16x8 a = ld8s(L.h1 + (8 * (lane >> 4)) * SH + 16 * mi + (lane & 15), SH);
#pragma unroll
        for (int nj = 0; nj < NIN / 16; ++nj) {
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          acc = mfma_bf16(a, bfr[nj], acc);
          const int col = 16 * nj + (lane & 15);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = 16 * mi + 4 * (lane >> 4) + v;
            atomicAdd(&grad[OW1 + row * NIN + col], acc[v]);
          }
        }
      }
    }
  }
}

__global__ __launch_bounds__(TPB) void mlp_zero_kernel(float* __restrict__ a, int n) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) a[i] = 0.0f;
}

__global__ __launch_bounds__(TPB) void mlp_adam_kernel(float* __restrict__ P, float* __restrict__ G,
                                                       float* __restrict__ M, float* __restrict__ V,
                                                       float lr_t, float b1, float b2, float eps) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < NPARAM; i += gridDim.x * TPB) {
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    // lr_t carries the bias corrections: lr * sqrt(1 - b2^t) / (1 - b1^t)
    P[i] -= lr_t * m / (sqrtf(v) + eps);
    G[i] = 0.0f;
  }
}

__global__ __launch_bounds__(TPB) void mlp_init_kernel(float* __restrict__ P,
                                                       unsigned long long seed) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < NPARAM; i += gridDim.x * TPB) {
    float std = 0.0f;
    if (i < OB1) std = rsqrtf((float)NIN);
    else if (i >= OW2 && i < OB2) std = rsqrtf((float)NH);
    else if (i >= OW3 && i < OB3) std = rsqrtf((float)NH);
    P[i] = std * (float)rng_normal((unsigned)i, 0u, 4u, seed);
  }
}

// Inference on 8x8-pixel spots with the CoG fallback: sub-aperture s uses
// the MLP unless its flux is outside [flux_lo, flux_hi] (out of the training
// distribution), then the CoG of the same pixels.  Spots are read from the
// detector image windows at (vx[s], vy[s]) (row length `size`), or from a
// spot cube when vx is NULL.
__global__ __launch_bounds__(TPB) void mlp_centroid_kernel(
    const float* __restrict__ P, const __bf16* __restrict__ work, const float* __restrict__ img,
    int size, const int* __restrict__ vx, const int* __restrict__ vy, int nvalid, float x_scale,
    float flux_lo, float flux_hi, float offset, float scale, float y_scale,
    const float* __restrict__ ref, int slope_order, float* __restrict__ slopes,
    float* __restrict__ intens, int* __restrict__ fallback) {
  extern __shared__ char smem_raw[];
  const Lds L(smem_raw);
  const __bf16* W1b = work;
  const __bf16* W2b = W1b + NH * NIN;
  const long row0 = (long)blockIdx.x * TS;
  const int rows = (int)((nvalid - row0) < TS ? (nvalid - row0) : TS);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // spot pixels (px, py) = (c & 7, c >> 3)
  auto pixel = [&](long s, int c) -> float {
    if (vx == nullptr) return img[s * NIN + c];
    return img[(long)(vx[s] + (c & 7)) + (long)(vy[s] + (c >> 3)) * size];
  };
  for (int i = tid; i < TS * NIN; i += TPB) {
    const int sl = i / NIN, c = i % NIN;
    L.x[sl * SX + c] = (__bf16)(sl < rows ? pixel(row0 + sl, c) * x_scale : 0.0f);
  }
  mlp::tile_forward<TS, Bf16Weights, false>(P, Bf16Weights{W1b, W2b}, L, (int)threadIdx.x);
  // flux and CoG from the raw pixels: one wave per 8 sub-apertures
  for (int k = 0; k < TS / 4; ++k) {
    const int sl = w * (TS / 4) + k;
    if (sl >= rows) break;  // uniform over the wave
    const long s = row0 + sl;
    mlp::spot_slopes(pixel(s, lane), lane, (int)s, nvalid, L.y[sl * NOUT + 0], L.y[sl * NOUT + 1],
                     flux_lo, flux_hi, offset, scale, y_scale, ref, slope_order, slopes, intens,
                     fallback);
  }
}

__global__ __launch_bounds__(TPB) void phase_window_kernel(const float* __restrict__ src,
                                                           int side, int ox, int oy, float gain,
                                                           float* __restrict__ dst, int n) {
  const long total = (long)n * n;
  for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < total; i += (long)gridDim.x * TPB) {
    const int x = (int)(i % n), y = (int)(i / n);
    dst[i] = gain * src[(long)(ox + x) + (long)(oy + y) * side];
  }
}

int blocks_for(long n) { return (int)((n + TS - 1) / TS); }

int32_t prepare(const float* P, void* work, hipStream_t st) {
  hipLaunchKernelGGL(mlp_prepare_kernel, dim3(64), dim3(TPB), 0, st, P, (__bf16*)work);
  SHWFS_LAUNCHED("mlp_prepare_kernel");
  return SHWFS_OK;
}

}  // namespace
}  // namespace shwfs

using namespace shwfs;

extern "C" {

size_t shwfs_mlp_work_bytes(void) { return WORK_ELEMS * sizeof(__bf16); }

int32_t shwfs_phase_window(const float* d_src, int32_t src_side, int32_t ox, int32_t oy,
                           float gain, float* d_dst, int32_t n, void* stream) {
  if (!d_src || !d_dst) return fail(SHWFS_ERR_ARG, "phase_window: NULL argument");
  if (n <= 0 || src_side <= 0 || ox < 0 || oy < 0 || ox + n > src_side || oy + n > src_side)
    return fail(SHWFS_ERR_ARG, "phase_window: window outside the source screen");
  const long total = (long)n * n;
  long blocks = (total + TPB - 1) / TPB;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(phase_window_kernel, dim3((unsigned)blocks), dim3(TPB), 0,
                     (hipStream_t)stream, d_src, src_side, ox, oy, gain, d_dst, n);
  SHWFS_LAUNCHED("phase_window_kernel");
  return SHWFS_OK;
}

int32_t shwfs_mlp_init(float* d_params, uint64_t seed, void* stream) {
  if (!d_params) return fail(SHWFS_ERR_ARG, "mlp_init: NULL parameters");
  hipLaunchKernelGGL(mlp_init_kernel, dim3(64), dim3(TPB), 0, (hipStream_t)stream, d_params,
                     (unsigned long long)seed);
  SHWFS_LAUNCHED("mlp_init_kernel");
  return SHWFS_OK;
}

int32_t shwfs_mlp_forward(const float* d_params, void* d_work, const float* d_x, int64_t n,
                          float x_scale, float* d_y, void* stream) {
  if (!d_params || !d_work || !d_x || !d_y) return fail(SHWFS_ERR_ARG, "mlp_forward: NULL argument");
  if (n <= 0) return fail(SHWFS_ERR_ARG, "mlp_forward: n must be positive");
  if ((n + TS - 1) / TS > 0x7fffffff) return fail(SHWFS_ERR_ARG, "mlp_forward: n too large");
  hipStream_t st = (hipStream_t)stream;
  int32_t rc = prepare(d_params, d_work, st);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_tile_kernel<false>, dim3(blocks_for(n)), dim3(TPB), LDS_BYTES, st,
                     d_params, (const __bf16*)d_work, d_x, nullptr, nullptr, (long)n, x_scale, 1.0f,
                     d_y, nullptr, nullptr);
  SHWFS_LAUNCHED("mlp_tile_kernel");
  return SHWFS_OK;
}

int32_t shwfs_mlp_loss_grad(const float* d_params, void* d_work, const float* d_x,
                            const float* d_t, const int32_t* d_index, int64_t n, float x_scale,
                            float t_scale, float* d_grad, float* d_loss, void* stream) {
  if (!d_params || !d_work || !d_x || !d_t || !d_grad || !d_loss)
    return fail(SHWFS_ERR_ARG, "mlp_loss_grad: NULL argument");
  if (n <= 0) return fail(SHWFS_ERR_ARG, "mlp_loss_grad: n must be positive");
  hipStream_t st = (hipStream_t)stream;
  int32_t rc = prepare(d_params, d_work, st);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_zero_kernel, dim3(32), dim3(TPB), 0, st, d_grad, NPARAM);
  hipLaunchKernelGGL(mlp_zero_kernel, dim3(1), dim3(TPB), 0, st, d_loss, 1);
  SHWFS_LAUNCHED("mlp_zero_kernel");
  hipLaunchKernelGGL(mlp_tile_kernel<true>, dim3(blocks_for(n)), dim3(TPB), LDS_BYTES, st,
                     d_params, (const __bf16*)d_work, d_x, d_t, d_index, (long)n, x_scale, t_scale,
                     nullptr, d_grad, d_loss);
  SHWFS_LAUNCHED("mlp_tile_kernel");
  return SHWFS_OK;
}

int32_t shwfs_mlp_adam(float* d_params, float* d_grad, float* d_m, float* d_v, int32_t step,
                       float lr, float beta1, float beta2, float eps, void* stream) {
  if (!d_params || !d_grad || !d_m || !d_v) return fail(SHWFS_ERR_ARG, "mlp_adam: NULL argument");
  if (step < 1) return fail(SHWFS_ERR_ARG, "mlp_adam: step counts from 1");
  const double c1 = 1.0 - std::pow((double)beta1, step);
  const double c2 = 1.0 - std::pow((double)beta2, step);
  const float lr_t = (float)(lr * std::sqrt(c2) / c1);
  const float eps_t = (float)(eps * std::sqrt(c2));
  hipLaunchKernelGGL(mlp_adam_kernel, dim3(64), dim3(TPB), 0, (hipStream_t)stream, d_params,
                     d_grad, d_m, d_v, lr_t, beta1, beta2, eps_t);
  SHWFS_LAUNCHED("mlp_adam_kernel");
  return SHWFS_OK;
}

int32_t shwfs_mlp_centroids(const float* d_params, void* d_work, const float* d_img,
                            int32_t size, const int32_t* d_validx, const int32_t* d_validy,
                            int32_t nvalid, int32_t npix, float x_scale, float flux_lo,
                            float flux_hi, float offset, float scale, float y_scale,
                            const float* d_ref, int32_t slope_order, float* d_slopes,
                            float* d_intensities, int32_t* d_fallback, void* stream) {
  if (!d_params || !d_work || !d_img || !d_slopes)
    return fail(SHWFS_ERR_ARG, "mlp_centroids: NULL argument");
  if (npix * npix != NIN) return fail(SHWFS_ERR_UNSUPPORTED, "mlp_centroids: needs 8x8-pixel spots");
  if (nvalid <= 0) return fail(SHWFS_ERR_ARG, "mlp_centroids: nvalid must be positive");
  if ((d_validx == nullptr) != (d_validy == nullptr))
    return fail(SHWFS_ERR_ARG, "mlp_centroids: give both valid position arrays or neither");
  if (d_validx && size < npix) return fail(SHWFS_ERR_ARG, "mlp_centroids: image smaller than a spot");
  hipStream_t st = (hipStream_t)stream;
  int32_t rc = prepare(d_params, d_work, st);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_centroid_kernel, dim3(blocks_for(nvalid)), dim3(TPB), LDS_BYTES, st,
                     d_params, (const __bf16*)d_work, d_img, size, d_validx, d_validy, nvalid,
                     x_scale, flux_lo, flux_hi, offset, scale, y_scale, d_ref, slope_order,
                     d_slopes, d_intensities, d_fallback);
  SHWFS_LAUNCHED("mlp_centroid_kernel");
  return SHWFS_OK;
}

}  // extern "C"
