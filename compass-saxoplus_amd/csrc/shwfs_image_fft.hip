// Plain FP32 spot-image kernel (stages in shwfs_image_fft.hpp) and the
// background readout-noise kernel of the detector image.
#include "shwfs_image_fft.hpp"

namespace shwfs {

template <int N, int R, int TPB, bool NOISE>
__global__ __launch_bounds__(TPB) void sh_image_fft_kernel(ImageParams P) {
  using L = FftLayout<N, R>;
  constexpr int S = TPB / N;
  extern __shared__ float smem[];
  int* s_binmap = (int*)(smem + S * L::slot_floats(P.pdiam, P.npix));
  if (!P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += TPB) s_binmap[i] = P.binmap[i];
  }
  fft_tile<N, R, TPB, NOISE>(P, xcd_remap(blockIdx.x, gridDim.x) * S, smem, s_binmap);
}

template <int N, int R, int TPB>
static int32_t launch_fft_nrt(const ImageParams& p, hipStream_t st) {
  using L = FftLayout<N, R>;
  constexpr int S = TPB / N;
  const int nbm = p.bin_block ? 0 : p.nrebin * p.nrebin * p.npix * p.npix;
  const size_t lds = sizeof(float) * ((size_t)S * L::slot_floats(p.pdiam, p.npix) + nbm);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "spot kernel: LDS footprint exceeds the per-workgroup limit");
  const int grid = (p.nvalid + S - 1) / S;
  if (p.noise > -1.0f)
    hipLaunchKernelGGL((sh_image_fft_kernel<N, R, TPB, true>), dim3(grid), dim3(TPB), lds, st, p);
  else
    hipLaunchKernelGGL((sh_image_fft_kernel<N, R, TPB, false>), dim3(grid), dim3(TPB), lds, st, p);
  SHWFS_LAUNCHED("sh_image_fft_kernel");
  return SHWFS_OK;
}

template <int N, int R>
static int32_t launch_fft_nr(const ImageParams& p, hipStream_t st) {
  // workgroup size: 64 lanes by default (more, smaller workgroups spread the
  // sub-apertures evenly over the 256 CUs); 256 on request (tuning)
  if (p.tpb == 256) return launch_fft_nrt<N, R, 256>(p, st);
  if (p.tpb == 128 && N <= 128) return launch_fft_nrt<N, R, 128>(p, st);
  return launch_fft_nrt<N, R, 64>(p, st);
}

int32_t launch_image_fft(const ImageParams& p, int nfft, hipStream_t st) {
  const bool half = 2 * p.pdiam <= nfft;
  switch (nfft) {
    case 8: return half ? launch_fft_nr<8, 2>(p, st) : launch_fft_nr<8, 1>(p, st);
    case 16: return half ? launch_fft_nr<16, 2>(p, st) : launch_fft_nr<16, 1>(p, st);
    case 32: return half ? launch_fft_nr<32, 2>(p, st) : launch_fft_nr<32, 1>(p, st);
    case 64: return half ? launch_fft_nr<64, 2>(p, st) : launch_fft_nr<64, 1>(p, st);
    default: return fail(SHWFS_ERR_UNSUPPORTED, "nfft must be 8, 16, 32 or 64");
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
