// Device pieces of the learned centroider (MLP 64 -> 128 -> 128 -> 2, GELU)
// shared by its own kernels (shwfs_mlp.hip: training, inference over spot
// cubes or detector windows) and by the spot kernel that runs the inference
// after binning (shwfs_image_fft.hip).  No counterpart in the reference.
//
// Parameters: flat fp32 [NPARAM] = W1 (128x64 row-major), b1, W2 (128x128),
// b2, W3 (2x128), b3.  Matrix products on v_mfma_f32_16x16x32_bf16 with bf16
// operands (the weights rounded to bf16 once, by the same conversion in every
// kernel) and fp32 accumulation; the 2-output layer in fp32 on the VALU.
#pragma once

#include <cmath>

#include "shwfs_device.hpp"
#include "shwfs_mfma.hpp"

namespace shwfs {
namespace mlp {

constexpr int NIN = 64, NH = 128, NOUT = 2;
constexpr int OW1 = 0, OB1 = OW1 + NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH,
              OW3 = OB2 + NH, OB3 = OW3 + NOUT * NH, NPARAM = OB3 + NOUT;
// LDS row strides (elements), padded against bank conflicts
constexpr int SX = NIN + 8;  // bf16 X
constexpr int SZ = NH + 4;   // fp32 Z1 / Z2
constexpr int SH = NH + 8;   // bf16 H1 / H2

// LDS of a tile of TS samples.
template <int TS>
struct Lds {
  __bf16* x;   // [TS][SX]
  float* z1;   // [TS][SZ]   z1, later dZ1
  float* z2;   // [TS][SZ]   z2, later dZ2
  __bf16* h1;  // [TS][SH]   gelu(z1), later dZ1 (bf16)
  __bf16* h2;  // [TS][SH]   gelu(z2), later dZ2 (bf16)
  float* y;    // [TS][NOUT]
  float* dy;   // [TS][NOUT]
  static constexpr size_t bytes() {
    return (size_t)TS * SX * 2 + 2 * (size_t)TS * SZ * 4 + 2 * (size_t)TS * SH * 2 +
           (size_t)TS * NOUT * 4 * 2;
  }
  __device__ explicit Lds(char* base) {
    x = (__bf16*)base;
    base += (size_t)TS * SX * 2;
    z1 = (float*)base;
    base += (size_t)TS * SZ * 4;
    z2 = (float*)base;
    base += (size_t)TS * SZ * 4;
    h1 = (__bf16*)base;
    base += (size_t)TS * SH * 2;
    h2 = (__bf16*)base;
    base += (size_t)TS * SH * 2;
    y = (float*)base;
    base += (size_t)TS * NOUT * 4;
    dy = (float*)base;
  }
};

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + tanhf(kGeluC * (z + 0.044715f * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float t = tanhf(kGeluC * (z + 0.044715f * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * 0.044715f * z * z);
}

// 8 consecutive bf16 of a row (16-byte aligned)
__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
// 8 bf16 at stride `st` (column of a row-major array)
__device__ __forceinline__ bf16x8 ld8s(const __bf16* p, int st) {
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = p[e * st];
  return v;
}

// Weight operands from bf16 copies of W1 and W2 (the training and
// stand-alone inference kernels, which convert them once per launch).
struct Bf16Weights {
  const __bf16* W1b;
  const __bf16* W2b;
  __device__ bf16x8 w1(int row, int k) const { return ld8(W1b + row * NIN + k); }
  __device__ bf16x8 w2(int row, int k) const { return ld8(W2b + row * NH + k); }
};
// The same operands converted from the fp32 parameters as they are read
// (the conversion is the one the bf16 copies are made with, so both give
// the same operands).
struct Fp32Weights {
  const float* P;
  __device__ static bf16x8 cvt8(const float* p) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    bf16x8 v;
    v[0] = (__bf16)a.x; v[1] = (__bf16)a.y; v[2] = (__bf16)a.z; v[3] = (__bf16)a.w;
    v[4] = (__bf16)b.x; v[5] = (__bf16)b.y; v[6] = (__bf16)b.z; v[7] = (__bf16)b.w;
    return v;
  }
  __device__ bf16x8 w1(int row, int k) const { return cvt8(P + OW1 + row * NIN + k); }
  __device__ bf16x8 w2(int row, int k) const { return cvt8(P + OW2 + row * NH + k); }
};

// Forward of the TS samples in L.x (bf16, scaled inputs) into LDS (z1, h1,
// z2, h2, y).  Every thread of the workgroup calls it (it contains
// barriers); waves 0-3 compute the products (wave w: output tiles 2w and
// 2w+1 of each layer), four lanes per output the last layer.  The
// value of a sample does not depend on TS or on the other samples of the
// tile.  tid: the thread index.
// KEEP_Z: z1 and z2 keep the pre-activations (the backward pass reads them);
// otherwise z2 receives gelu(z2) in fp32, which the last layer then reads
// without computing it again (the same values either way).
template <int TS, class WS, bool KEEP_Z = true>
__device__ __forceinline__ void tile_forward(const float* __restrict__ P, const WS& ws,
                                             const Lds<TS>& L, int tid) {
  const int lane = tid & 63, w = tid >> 6;
  __syncthreads();
  if (w < 4) {
    // Z1 = X W1^T + b1
#pragma unroll
    for (int mi = 0; mi < TS / 16; ++mi)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const int nj = 2 * w + jj;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < NIN / 32; ++ks) {
          const bf16x8 a = ld8(L.x + (16 * mi + (lane & 15)) * SX + 32 * ks + 8 * (lane >> 4));
          const bf16x8 b = ws.w1(16 * nj + (lane & 15), 32 * ks + 8 * (lane >> 4));
          acc = mfma_bf16(a, b, acc);
        }
        const int col = 16 * nj + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * mi + 4 * (lane >> 4) + v;
          const float z = acc[v] + P[OB1 + col];
          if constexpr (KEEP_Z) L.z1[row * SZ + col] = z;
          L.h1[row * SH + col] = (__bf16)gelu(z);
        }
      }
  }
  __syncthreads();
  if (w < 4) {
    // Z2 = H1 W2^T + b2
#pragma unroll
    for (int mi = 0; mi < TS / 16; ++mi)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const int nj = 2 * w + jj;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < NH / 32; ++ks) {
          const bf16x8 a = ld8(L.h1 + (16 * mi + (lane & 15)) * SH + 32 * ks + 8 * (lane >> 4));
          const bf16x8 b = ws.w2(16 * nj + (lane & 15), 32 * ks + 8 * (lane >> 4));
          acc = mfma_bf16(a, b, acc);
        }
        const int col = 16 * nj + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * mi + 4 * (lane >> 4) + v;
          const float z = acc[v] + P[OB2 + col];
          const float g = gelu(z);
          L.z2[row * SZ + col] = KEEP_Z ? z : g;
          L.h2[row * SH + col] = (__bf16)g;
        }
      }
  }
  __syncthreads();
  // Y = gelu(Z2) W3^T + b3 (fp32): each 128-term dot product (s, o) is
  // shared by PARTS consecutive lanes (terms k = part + PARTS j), whose
  // partial sums meet in a fixed butterfly order; the bias is added last.
  // PARTS does not depend on TS: a sample's value stays the same in every
  // tile size.
  constexpr int PARTS = 4;
  static_assert(TS * NOUT * PARTS <= 256 && NH % PARTS == 0,
                "tile_forward: four-wave workgroups, whole dot products per lane group");
  if (tid < TS * NOUT * PARTS) {
    const int pair = tid / PARTS, part = tid % PARTS;
    const int s = pair / NOUT, o = pair % NOUT;
    float acc = 0.0f;
#pragma unroll 4
    for (int j = 0; j < NH / PARTS; ++j) {
      const int k = part + PARTS * j;
      const float z = L.z2[s * SZ + k];
      acc += (KEEP_Z ? gelu(z) : z) * P[OW3 + o * NH + k];
    }
#pragma unroll
    for (int m = 1; m < PARTS; m <<= 1) acc += __shfl_xor(acc, m, PARTS);
    if (part == 0) L.y[s * NOUT + o] = acc + P[OB3 + o];
  }
  __syncthreads();
}

// Slopes of one 8x8 spot from the network output (y0, y1, pixels) or, when
// the spot flux I is outside [flux_lo, flux_hi], its CoG
// (Sx/(I+1e-6) - offset) * scale.  b: the lane's pixel (px, py) = (lane & 7,
// lane >> 3); every lane of the wave calls it; lane 0 writes.
__device__ __forceinline__ void spot_slopes(float b, int lane, int s, int nvalid, float y0,
                                            float y1, float flux_lo, float flux_hi, float offset,
                                            float scale, float y_scale, const float* ref,
                                            int slope_order, float* slopes, float* intens,
                                            int* fallback) {
  const float I = wave_sum(b);
  const float Sx = wave_sum(b * (float)(lane & 7));
  const float Sy = wave_sum(b * (float)(lane >> 3));
  if (lane == 0) {
    const bool ood = !(I >= flux_lo && I <= flux_hi);
    const int ix = slope_order == 0 ? s : 2 * s;
    const int iy = slope_order == 0 ? s + nvalid : 2 * s + 1;
    float cx, cy;
    if (ood) {
      cx = ((float)((double)Sx / ((double)I + 1.e-6)) - offset) * scale;
      cy = ((float)((double)Sy / ((double)I + 1.e-6)) - offset) * scale;
    } else {
      cx = y0 * y_scale;
      cy = y1 * y_scale;
    }
    slopes[ix] = cx - (ref ? ref[ix] : 0.0f);
    slopes[iy] = cy - (ref ? ref[iy] : 0.0f);
    if (intens) intens[s] = I;
    if (fallback) fallback[s] = ood ? 1 : 0;
  }
}

}  // namespace mlp
}  // namespace shwfs
