// Plain FP32 spot-image kernel (stages in shwfs_image_fft.hpp) and the
// background readout-noise kernel of the detector image.
#include "shwfs_image_fft.hpp"
#include "shwfs_mlp.hpp"

namespace shwfs {

// Workgroups [0, bg.spot_blocks) compute spots; with noise, those past them
// draw the fused background (bg).
template <int N, int R, int TPB, bool NOISE>
__global__ __launch_bounds__(TPB) void sh_image_fft_kernel(ImageParams P, BgParams bg) {
  using L = FftLayout<N, R>;
  constexpr int S = TPB / N;
  if constexpr (NOISE) {
    if ((int)blockIdx.x >= bg.spot_blocks) {
      draw_background<TPB>(P, bg, (int)blockIdx.x - bg.spot_blocks);
      return;
    }
  }
  extern __shared__ float smem[];
  int* s_binmap = (int*)(smem + S * L::slot_floats(P.pdiam, P.npix));
  if (!P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += TPB) s_binmap[i] = P.binmap[i];
  }
  fft_tile<N, R, TPB, NOISE>(P, xcd_remap(blockIdx.x, bg.spot_blocks) * S, smem, s_binmap);
}

template <int N, int R, int TPB>
static int32_t launch_fft_nrt(const ImageParams& p, hipStream_t st, BgParams bg) {
  using L = FftLayout<N, R>;
  constexpr int S = TPB / N;
  const int nbm = p.bin_block ? 0 : p.nrebin * p.nrebin * p.npix * p.npix;
  const size_t lds = sizeof(float) * ((size_t)S * L::slot_floats(p.pdiam, p.npix) + nbm);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "spot kernel: LDS footprint exceeds the per-workgroup limit");
  bg.spot_blocks = (p.nvalid + S - 1) / S;
  const bool noisy = p.noise > -1.0f;
  const int grid = bg.spot_blocks + (noisy ? (bg.n + TPB - 1) / TPB : 0);
  if (noisy)
    hipLaunchKernelGGL((sh_image_fft_kernel<N, R, TPB, true>), dim3(grid), dim3(TPB), lds, st, p, bg);
  else
    hipLaunchKernelGGL((sh_image_fft_kernel<N, R, TPB, false>), dim3(grid), dim3(TPB), lds, st, p, bg);
  SHWFS_LAUNCHED("sh_image_fft_kernel");
  return SHWFS_OK;
}

template <int N, int R>
static int32_t launch_fft_nr(const ImageParams& p, hipStream_t st, const BgParams& bg) {
  // workgroup size: 64 lanes by default (more, smaller workgroups spread the
  // sub-apertures evenly over the 256 CUs); 256 on request (tuning)
  if (p.tpb == 256) return launch_fft_nrt<N, R, 256>(p, st, bg);
  if (p.tpb == 128 && N <= 128) return launch_fft_nrt<N, R, 128>(p, st, bg);
  return launch_fft_nrt<N, R, 64>(p, st, bg);
}

int32_t launch_image_fft(const ImageParams& p, int nfft, hipStream_t st, const BgParams& bg) {
  if (bg.n > 0 && !(p.noise > -1.0f))
    return fail(SHWFS_ERR_ARG, "spot kernel: the fused background needs a noisy frame");
  const bool half = 2 * p.pdiam <= nfft;
  switch (nfft) {
    case 8: return half ? launch_fft_nr<8, 2>(p, st, bg) : launch_fft_nr<8, 1>(p, st, bg);
    case 16: return half ? launch_fft_nr<16, 2>(p, st, bg) : launch_fft_nr<16, 1>(p, st, bg);
    case 32: return half ? launch_fft_nr<32, 2>(p, st, bg) : launch_fft_nr<32, 1>(p, st, bg);
    case 64: return half ? launch_fft_nr<64, 2>(p, st, bg) : launch_fft_nr<64, 1>(p, st, bg);
    default: return fail(SHWFS_ERR_UNSUPPORTED, "nfft must be 8, 16, 32 or 64");
  }
}

// ---- spots with the guide-star raytrace in the field stage ------------------
// The FFT spot kernel (one-wave workgroups) whose window phases are traced
// through the layers (ray_phase) instead of read from a rendered screen: the
// npup^2 screen is neither written nor read back.
template <int N, int R, bool NOISE>
__global__ __launch_bounds__(64) void sh_image_fft_ray_kernel(ImageParams P, RayParams Ry) {
  using L = FftLayout<N, R>;
  constexpr int S = 64 / N;
  extern __shared__ float smem[];
  int* s_binmap = (int*)(smem + S * L::slot_floats(P.pdiam, P.npix));
  if (!P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += 64) s_binmap[i] = P.binmap[i];
  }
  fft_tile<N, R, 64, NOISE, false, true>(P, xcd_remap(blockIdx.x, gridDim.x) * S, smem, s_binmap,
                                         0u, (int)threadIdx.x, &Ry);
}

template <int N, int R>
static int32_t launch_fft_ray_nr(const ImageParams& p, const RayParams& r, hipStream_t st) {
  using L = FftLayout<N, R>;
  constexpr int S = 64 / N;
  const int nbm = p.bin_block ? 0 : p.nrebin * p.nrebin * p.npix * p.npix;
  const size_t lds = sizeof(float) * ((size_t)S * L::slot_floats(p.pdiam, p.npix) + nbm);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "spot kernel: LDS footprint exceeds the per-workgroup limit");
  const int grid = (p.nvalid + S - 1) / S;
  if (p.noise > -1.0f)
    hipLaunchKernelGGL((sh_image_fft_ray_kernel<N, R, true>), dim3(grid), dim3(64), lds, st, p, r);
  else
    hipLaunchKernelGGL((sh_image_fft_ray_kernel<N, R, false>), dim3(grid), dim3(64), lds, st, p, r);
  SHWFS_LAUNCHED("sh_image_fft_ray_kernel");
  return SHWFS_OK;
}

int32_t launch_image_fft_ray(const ImageParams& p, const RayParams& r, int nfft, hipStream_t st) {
  if (p.nvalid <= 0) return SHWFS_OK;
  const bool half = 2 * p.pdiam <= nfft;
  switch (nfft) {
    case 8: return half ? launch_fft_ray_nr<8, 2>(p, r, st) : launch_fft_ray_nr<8, 1>(p, r, st);
    case 16: return half ? launch_fft_ray_nr<16, 2>(p, r, st) : launch_fft_ray_nr<16, 1>(p, r, st);
    case 32: return half ? launch_fft_ray_nr<32, 2>(p, r, st) : launch_fft_ray_nr<32, 1>(p, r, st);
    case 64: return half ? launch_fft_ray_nr<64, 2>(p, r, st) : launch_fft_ray_nr<64, 1>(p, r, st);
    default: return fail(SHWFS_ERR_UNSUPPORTED, "nfft must be 8, 16, 32 or 64");
  }
}

// ---- spots + learned centroider in one launch ------------------------------
// 256 / N sub-apertures per workgroup of four waves (N lanes each, npix 8):
// 8 at N = 32, so the 1,256 sub-apertures of a 40x40 WFS spread over 157
// workgroups (16 per workgroup left most CUs idle during the spot stage).
// The tile's tail leaves each spot's spot-cube values in its slot
// (fft_tile<..., KEEP>); the spots then go through one MLP forward on the
// matrix cores in a 16-row tile (rows past the workgroup's spots are zero;
// weights converted to bf16 as they are read) and each wave turns network
// outputs into slopes, or takes the CoG of the same pixels where the spot
// flux is outside the network's range.  Same arithmetic as
// shwfs_mlp_centroids over the spot cube, so both give the same slopes.
constexpr int kMlpThreads = 256;  // tile_forward computes with four waves
constexpr int kMlpTile = 16;      // MFMA rows of the MLP tile

template <int N>
__host__ __device__ constexpr int mlp_subaps() {
  return kMlpThreads / N;
}

template <int N, int R>
__host__ __device__ constexpr int mlp_lds_floats_min() {
  // spots (S x 64 floats) + the MLP tile, laid over the dead slots
  return mlp_subaps<N>() * 64 + (int)(mlp::Lds<kMlpTile>::bytes() / sizeof(float));
}

// Workgroups past bg.spot_blocks draw the fused background (noisy frames).
template <int N, int R, bool NOISE>
__global__ __launch_bounds__(kMlpThreads) void sh_image_fft_mlp_kernel(ImageParams P,
                                                                       MlpParams M, BgParams bg) {
  using L = FftLayout<N, R>;
  constexpr int S = mlp_subaps<N>(), TPB = kMlpThreads, NW = TPB / 64;
  static_assert(S <= kMlpTile && S * N == TPB, "fused MLP: one tile of spots per workgroup");
  if constexpr (NOISE) {
    if ((int)blockIdx.x >= bg.spot_blocks) {
      draw_background<TPB>(P, bg, (int)blockIdx.x - bg.spot_blocks);
      return;
    }
  }
  extern __shared__ float smem[];
  const int slot_f = L::slot_floats(P.pdiam, 8);
  int* s_binmap = (int*)(smem + S * slot_f);
  if (!P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * 64;
    for (int i = threadIdx.x; i < nbm; i += TPB) s_binmap[i] = P.binmap[i];
  }
  const int s0 = xcd_remap(blockIdx.x, bg.spot_blocks) * S;
  fft_tile<N, R, TPB, NOISE, true>(P, s0, smem, s_binmap);  // ends with a barrier
  const int tid = threadIdx.x;
  const int rows = P.nvalid - s0 < S ? P.nvalid - s0 : S;
  // the S spots (pixel c of spot k at slot k's spot area) into registers,
  // then, past a barrier, to their places in the MLP tile
  constexpr int PER = S * 64 / TPB;
  float v[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = tid + e * TPB, k = i >> 6, c = i & 63;
    v[e] = k < rows ? smem[k * slot_f + slot_f - 64 + c] : 0.0f;
  }
  __syncthreads();
  float* spots = smem;  // [S][64]
  const mlp::Lds<kMlpTile> T(reinterpret_cast<char*>(smem + S * 64));
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = tid + e * TPB, k = i >> 6, c = i & 63;
    spots[i] = v[e];
    T.x[k * mlp::SX + c] = (__bf16)(v[e] * M.xs);
  }
  // tile rows without a spot of this workgroup: zeros
  for (int i = S * 64 + tid; i < kMlpTile * 64; i += TPB)
    T.x[(i >> 6) * mlp::SX + (i & 63)] = (__bf16)0.0f;
  // starts with a barrier
  mlp::tile_forward<kMlpTile, mlp::Fp32Weights, false>(M.params, mlp::Fp32Weights{M.params}, T, tid);
  const int lane = tid & 63, w = tid >> 6;
  for (int k = w; k < rows; k += NW)
    mlp::spot_slopes(spots[k * 64 + lane], lane, s0 + k, P.nvalid, T.y[k * mlp::NOUT],
                     T.y[k * mlp::NOUT + 1], M.lo, M.hi, P.coffset, P.cscale, M.ys, P.ref,
                     P.slope_order, P.slopes, P.cintens, M.fallback);
}

template <int N, int R>
static int32_t launch_fft_mlp_nr(const ImageParams& p, const MlpParams& m, hipStream_t st,
                                 BgParams bg) {
  using L = FftLayout<N, R>;
  constexpr int S = mlp_subaps<N>();
  const int nbm = p.bin_block ? 0 : p.nrebin * p.nrebin * 64;
  size_t floats = (size_t)S * L::slot_floats(p.pdiam, 8) + nbm;
  if (floats < (size_t)mlp_lds_floats_min<N, R>()) floats = mlp_lds_floats_min<N, R>();
  const size_t lds = sizeof(float) * floats;
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "spot + MLP kernel: LDS footprint exceeds the per-workgroup limit");
  bg.spot_blocks = (p.nvalid + S - 1) / S;
  const bool noisy = p.noise > -1.0f;
  const int grid = bg.spot_blocks + (noisy ? (bg.n + kMlpThreads - 1) / kMlpThreads : 0);
  if (noisy)
    hipLaunchKernelGGL((sh_image_fft_mlp_kernel<N, R, true>), dim3(grid), dim3(kMlpThreads), lds,
                       st, p, m, bg);
  else
    hipLaunchKernelGGL((sh_image_fft_mlp_kernel<N, R, false>), dim3(grid), dim3(kMlpThreads), lds,
                       st, p, m, bg);
  SHWFS_LAUNCHED("sh_image_fft_mlp_kernel");
  return SHWFS_OK;
}

int32_t launch_image_fft_mlp(const ImageParams& p, const MlpParams& m, int nfft, hipStream_t st,
                             const BgParams& bg) {
  if (p.npix != 8) return fail(SHWFS_ERR_UNSUPPORTED, "spot + MLP kernel: needs 8x8-pixel spots");
  if (bg.n > 0 && !(p.noise > -1.0f))
    return fail(SHWFS_ERR_ARG, "spot + MLP kernel: the fused background needs a noisy frame");
  if (p.nvalid <= 0) return SHWFS_OK;
  const bool half = 2 * p.pdiam <= nfft;
  switch (nfft) {
    case 16: return half ? launch_fft_mlp_nr<16, 2>(p, m, st, bg) : launch_fft_mlp_nr<16, 1>(p, m, st, bg);
    case 32: return half ? launch_fft_mlp_nr<32, 2>(p, m, st, bg) : launch_fft_mlp_nr<32, 1>(p, m, st, bg);
    default: return fail(SHWFS_ERR_UNSUPPORTED, "spot + MLP kernel: nfft must be 16 or 32");
  }
}

// ---- background readout noise of the detector image: img = sigma * N(0,1)
__global__ __launch_bounds__(256) void background_kernel(float* img, long n, float sigma,
                                                         unsigned long long seed, unsigned call,
                                                         const unsigned* call_base,
                                                         unsigned domain) {
  if (call_base) call += call_base[0];
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L)
    img[i] = sigma * (float)rng_normal((unsigned)i, call, domain, seed);
}

int32_t launch_background(float* img, long n, float sigma, unsigned long long seed, unsigned call,
                          hipStream_t st, const unsigned* call_base, unsigned domain) {
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(background_kernel, dim3((unsigned)blocks), dim3(256), 0, st, img, n, sigma,
                     seed, call, call_base, domain);
  SHWFS_LAUNCHED("background_kernel");
  return SHWFS_OK;
}

}  // namespace shwfs
