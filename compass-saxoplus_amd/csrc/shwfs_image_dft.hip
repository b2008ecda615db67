// MFMA spot-image kernels: the zero-padded 2-D FFT of each sub-aperture field
// becomes two dense products with the pruned DFT matrix on the matrix cores,
//     G = W E      (stage A, along x)      F^T = W G^T   (stage B, along y)
// with W[k][n] = exp(-2 pi i k n / N), k < N, n < pdiam (the zero rows of the
// padded field are never multiplied).  Complex products run as real ones on
// the stacked matrices
//     A_W = [[Wr, -Wi], [Wi, Wr]]  (2N x 2*P16)       B = [Re; Im]
// in v_mfma_f32_16x16x32 tiles, FP32 accumulation, one wave per
// sub-aperture.  Two operand precisions:
//   * SHWFS_METHOD_DFT_BF16X3: every operand x = hi + lo (bf16 each),
//     product = hi*hi + hi*lo + lo*hi (lo*lo dropped), 3 MFMA per tile;
//   * SHWFS_METHOD_DFT_FP16: fp16 operands, 1 MFMA per tile; G is scaled per
//     sub-aperture by a power of two (max |G| -> [2^14, 2^15)) before the
//     fp16 conversion, the scale is divided out of the flux.
// The field stage and the tail (binning ... centroids) are those of the FFT
// kernel (shwfs_tail.hpp).
#include "shwfs_mfma.hpp"
#include "shwfs_tail.hpp"

namespace shwfs {

template <int N, int Q>
struct DftLayout {
  static constexpr int P16 = 16 * Q;  // padded phase side (pdiam <= P16 <= N)
  static constexpr int MT = N / 8;    // M tiles of 16 rows: (re/im, kx or ky) = 2N rows
  static constexpr int KS = Q;        // K steps of 32: (re/im, x or y) = 2*P16
  static constexpr int NTA = Q;       // stage A column tiles (y)
  static constexpr int NTB = N / 16;  // stage B column tiles (kx)
  static constexpr int PE = P16 + 4;  // planar row stride (floats), 16-byte aligned rows
  static constexpr int PG = P16 + 4;
  static constexpr int EPLANE = P16 * PE;
  static constexpr int GPLANE = N * PG;
  static constexpr int EPL = (P16 * P16 + 63) / 64;  // field elements per lane
  static __host__ __device__ int wave_floats(int npix) {
    int m = 2 * EPLANE;
    if (2 * GPLANE > m) m = 2 * GPLANE;
    if (N * N > m) m = N * N;
    return ((m + 3) & ~3) + ((npix * npix + 3) & ~3);
  }
};

template <int PREC>
struct Frag;
template <>
struct Frag<SHWFS_METHOD_DFT_FP16> {
  f16x8 v;
  __device__ void set(int e, float x) { v[e] = (_Float16)x; }
};
template <>
struct Frag<SHWFS_METHOD_DFT_BF16X3> {
  bf16x8 hi, lo;
  __device__ void set(int e, float x) {
    __bf16 h, l;
    split_bf16(x, h, l);
    hi[e] = h;
    lo[e] = l;
  }
};

__device__ __forceinline__ f32x4 mma(const Frag<SHWFS_METHOD_DFT_FP16>& a,
                                     const Frag<SHWFS_METHOD_DFT_FP16>& b, f32x4 c) {
  return mfma_f16(a.v, b.v, c);
}
__device__ __forceinline__ f32x4 mma(const Frag<SHWFS_METHOD_DFT_BF16X3>& a,
                                     const Frag<SHWFS_METHOD_DFT_BF16X3>& b, f32x4 c) {
  c = mfma_bf16(a.lo, b.hi, c);
  c = mfma_bf16(a.hi, b.lo, c);
  return mfma_bf16(a.hi, b.hi, c);
}

// Element e of lane l of the A_W fragment (mt, ks).
template <int N, int Q>
__device__ __forceinline__ float aw_value(int mt, int ks, int lane, int e) {
  constexpr int P16 = 16 * Q;
  const int m = 16 * mt + (lane & 15);
  const int k = 32 * ks + 8 * (lane >> 4) + e;
  const bool row_im = m >= N, col_im = k >= P16;
  const int kx = row_im ? m - N : m;
  const int x = col_im ? k - P16 : k;
  const int idx = (kx * x) & (N - 1);
  const float c = kTw<N>.c[idx];   // cos(2 pi idx / N)
  const float s = kTw<N>.s[idx];   // -sin(2 pi idx / N)
  if (!row_im) return col_im ? -s : c;  // [Wr, -Wi] with Wi = -sin
  return col_im ? c : s;                // [Wi,  Wr]
}

template <int N, int Q, int PREC, int WPB, bool NOISE>
__global__ __launch_bounds__(64 * WPB) void sh_image_dft_kernel(ImageParams P) {
  using L = DftLayout<N, Q>;
  using F = Frag<PREC>;
  constexpr int TPB = 64 * WPB;
  extern __shared__ float smem[];

  const int pdiam = P.pdiam;
  const int wave_f = L::wave_floats(P.npix);
  int* s_binmap = (int*)(smem + WPB * wave_f);
  if (!P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += TPB) s_binmap[i] = P.binmap[i];
  }
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * WPB + w;
  const bool valid = s < P.nvalid;
  float* base = smem + w * wave_f;
  float* E = base;   // Er[y*PE + x], Ei[EPLANE + y*PE + x]
  float* G = base;   // Gr[kx*PG + y], Gi[GPLANE + kx*PG + y]   (after a barrier)
  float* hr = base;  // hr[kx + ky*N]                             (after a barrier)
  float* bin = base + (wave_f - ((P.npix * P.npix + 3) & ~3));

  // constant DFT operand (the same matrix serves both stages)
  F aw[L::MT][L::KS];
#pragma unroll
  for (int mt = 0; mt < L::MT; ++mt)
#pragma unroll
    for (int ks = 0; ks < L::KS; ++ks)
#pragma unroll
      for (int e = 0; e < 8; ++e) aw[mt][ks].set(e, aw_value<N, Q>(mt, ks, lane, e));

  // ---- 0. fields
  load_fields<64, L::EPL, true>(P, s, valid, lane, E, L::PE, L::EPLANE);
  __syncthreads();

  // ---- stage A operands: B[k][y] = E{r,i}[x][y], k = (re/im, x)
  const bool full = (pdiam == L::P16);
  F be[L::KS][L::NTA];
#pragma unroll
  for (int ks = 0; ks < L::KS; ++ks) {
#pragma unroll
    for (int nt = 0; nt < L::NTA; ++nt) {
      const int k0 = 32 * ks + 8 * (lane >> 4);
      const int im = k0 >= L::P16 ? 1 : 0;
      const int x0 = k0 - im * L::P16;
      const int y = 16 * nt + (lane & 15);
      const float* src = E + im * L::EPLANE + y * L::PE + x0;
      const float4 a = *reinterpret_cast<const float4*>(src);
      const float4 b = *reinterpret_cast<const float4*>(src + 4);
      float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const bool in = full || (x0 + e < pdiam && y < pdiam);
        be[ks][nt].set(e, (valid && in) ? v[e] : 0.0f);
      }
    }
  }
  __syncthreads();

  // ---- stage A: G = W E
  f32x4 ga[L::MT][L::NTA];
#pragma unroll
  for (int mt = 0; mt < L::MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < L::NTA; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < L::KS; ++ks) acc = mma(aw[mt][ks], be[ks][nt], acc);
      ga[mt][nt] = acc;
    }

  float gscale = 1.0f;
  if constexpr (PREC == SHWFS_METHOD_DFT_FP16) {
    float gmax = 0.0f;
#pragma unroll
    for (int mt = 0; mt < L::MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < L::NTA; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) gmax = fmaxf(gmax, fabsf(ga[mt][nt][v]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) gmax = fmaxf(gmax, __shfl_xor(gmax, o, 64));
    if (gmax > 0.0f) gscale = ldexpf(1.0f, 14 - ilogbf(gmax));
  }
#pragma unroll
  for (int mt = 0; mt < L::MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < L::NTA; ++nt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = 16 * mt + 4 * (lane >> 4) + v;
        const int im = m >= N ? 1 : 0;
        const int kx = m - im * N;
        const int y = 16 * nt + (lane & 15);
        G[im * L::GPLANE + kx * L::PG + y] = ga[mt][nt][v] * gscale;
      }
  __syncthreads();

  // ---- stage B operands: B[k][kx] = G{r,i}[kx][y], k = (re/im, y)
  F bg[L::KS][L::NTB];
#pragma unroll
  for (int ks = 0; ks < L::KS; ++ks)
#pragma unroll
    for (int nt = 0; nt < L::NTB; ++nt) {
      const int k0 = 32 * ks + 8 * (lane >> 4);
      const int im = k0 >= L::P16 ? 1 : 0;
      const int y0 = k0 - im * L::P16;
      const int kx = 16 * nt + (lane & 15);
      const float* src = G + im * L::GPLANE + kx * L::PG + y0;
      const float4 a = *reinterpret_cast<const float4*>(src);
      const float4 b = *reinterpret_cast<const float4*>(src + 4);
      float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int e = 0; e < 8; ++e) bg[ks][nt].set(e, v[e]);
    }
  __syncthreads();

  // ---- stage B: F^T = W G^T, |F|^2 -> hr
#pragma unroll
  for (int nt = 0; nt < L::NTB; ++nt) {
    f32x4 fb[L::MT];
#pragma unroll
    for (int mt = 0; mt < L::MT; ++mt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < L::KS; ++ks) acc = mma(aw[mt][ks], bg[ks][nt], acc);
      fb[mt] = acc;
    }
    const int kx = 16 * nt + (lane & 15);
#pragma unroll
    for (int mt = 0; mt < L::MT / 2; ++mt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ky = 16 * mt + 4 * (lane >> 4) + v;
        const float fr = fb[mt][v], fi = fb[mt + L::MT / 2][v];
        hr[kx + ky * N] = fr * fr + fi * fi;
      }
  }
  __syncthreads();

  // ---- tail
  spot_tail<64, N, NOISE>(P, s, valid, lane, hr, bin, s_binmap, 1.0f / (gscale * gscale));
}

template <int N, int Q, int PREC, int WPB>
static int32_t launch_dft_t(const ImageParams& p, hipStream_t st) {
  using L = DftLayout<N, Q>;
  const int nbm = p.bin_block ? 0 : p.nrebin * p.nrebin * p.npix * p.npix;
  const size_t lds = sizeof(float) * ((size_t)WPB * L::wave_floats(p.npix) + nbm);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "DFT spot kernel: LDS footprint exceeds the per-workgroup limit");
  const int grid = (p.nvalid + WPB - 1) / WPB;
  if (p.noise > -1.0f)
    hipLaunchKernelGGL((sh_image_dft_kernel<N, Q, PREC, WPB, true>), dim3(grid), dim3(64 * WPB),
                       lds, st, p);
  else
    hipLaunchKernelGGL((sh_image_dft_kernel<N, Q, PREC, WPB, false>), dim3(grid), dim3(64 * WPB),
                       lds, st, p);
  SHWFS_LAUNCHED("sh_image_dft_kernel");
  return SHWFS_OK;
}

template <int N, int Q, int PREC>
static int32_t launch_dft_w(const ImageParams& p, hipStream_t st) {
  if (p.tpb == 64) return launch_dft_t<N, Q, PREC, 1>(p, st);
  if (p.tpb == 128) return launch_dft_t<N, Q, PREC, 2>(p, st);
  return launch_dft_t<N, Q, PREC, 4>(p, st);
}

template <int PREC>
static int32_t launch_dft_p(const ImageParams& p, int nfft, hipStream_t st) {
  if (p.pdiam <= 16) {
    switch (nfft) {
      case 16: return launch_dft_w<16, 1, PREC>(p, st);
      case 32: return launch_dft_w<32, 1, PREC>(p, st);
      case 64: return launch_dft_w<64, 1, PREC>(p, st);
      default: break;
    }
  } else if (p.pdiam <= 32) {
    switch (nfft) {
      case 32: return launch_dft_w<32, 2, PREC>(p, st);
      case 64: return launch_dft_w<64, 2, PREC>(p, st);
      default: break;
    }
  }
  return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path: needs nfft in {16, 32, 64} and pdiam <= min(32, nfft)");
}

int32_t launch_image_dft(const ImageParams& p, int nfft, int method, hipStream_t st) {
  if (method == SHWFS_METHOD_DFT_FP16) return launch_dft_p<SHWFS_METHOD_DFT_FP16>(p, nfft, st);
  if (method == SHWFS_METHOD_DFT_BF16X3) return launch_dft_p<SHWFS_METHOD_DFT_BF16X3>(p, nfft, st);
  return fail(SHWFS_ERR_ARG, "unknown DFT method");
}

// ---- layout probe: D = A B for one 16x32 by 32x16 tile, operands given in
// float (row-major), rounded to the MFMA input type, D row-major float.
template <int PREC>
__global__ void mfma_probe_kernel(const float* A, const float* B, float* D) {
  const int lane = threadIdx.x;
  Frag<PREC> a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * (lane >> 4) + e;
    a.set(e, A[(lane & 15) * 32 + k]);
    b.set(e, B[k * 16 + (lane & 15)]);
  }
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = mma(a, b, d);
#pragma unroll
  for (int v = 0; v < 4; ++v) D[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = d[v];
}

int32_t launch_mfma_probe(const float* A, const float* B, float* D, int method, hipStream_t st) {
  if (method == SHWFS_METHOD_DFT_FP16)
    hipLaunchKernelGGL(mfma_probe_kernel<SHWFS_METHOD_DFT_FP16>, dim3(1), dim3(64), 0, st, A, B, D);
  else if (method == SHWFS_METHOD_DFT_BF16X3)
    hipLaunchKernelGGL(mfma_probe_kernel<SHWFS_METHOD_DFT_BF16X3>, dim3(1), dim3(64), 0, st, A, B, D);
  else
    return fail(SHWFS_ERR_ARG, "mfma_probe: method must be a DFT method");
  SHWFS_LAUNCHED("mfma_probe_kernel");
  return SHWFS_OK;
}

}  // namespace shwfs
