// Plain FP32 spot-image kernel.  A workgroup of TPB lanes handles TPB/N
// sub-apertures (N = FFT support); one group of N consecutive lanes owns one
// sub-aperture and everything below stays on chip:
//
//   0. the pdiam x pdiam phase / pupil window is loaded coalesced, every load
//      of a lane issued before the first use; compensated phase reduction +
//      sincos; field E(x, y) -> LDS
//   1. row transforms along x: lane (y, r) computes the M = N/R point FFT of
//      E(., y) * W_N^(x r), i.e. the outputs kx = R k' + r (R = 2 when
//      pdiam <= N/2: the zero padding halves every transform) -> LDS G(kx, y)
//   2. column transforms along y: lane kx computes all N outputs ky of its
//      column in registers, |F|^2 -> LDS hr(kx + ky N)
//   3. the shared tail (binning, flux, normalisation, noise, bincube /
//      binimg writes, optional fused centroider).
//
// Semantics: libsutra/src.cu/sutra_wfs.cu:27-53 (camplipup_krnl), cuFFT
// forward C2C (libcarma/src.cpp/carma_fft.cpp:164-180), sutra_utils.cu:144-193,
// sutra_wfs.cu:183-229,369-414,556-590.
#include "shwfs_tail.hpp"

namespace shwfs {

template <int N, int R>
struct FftLayout {
  static constexpr int M = N / R;    // row transform size
  static constexpr int PS = M + 1;   // G row stride in float2 (odd: fewer bank conflicts)
  static constexpr int EPL = (M * M + N - 1) / N;  // field elements per lane (pdiam <= M)
  static __host__ __device__ int pe(int pdiam) { return pdiam + 1; }  // E row stride
  static __host__ __device__ int slot_floats(int pdiam, int npix) {
    int e = pdiam * pe(pdiam) * 2;
    int g = N * PS * 2;
    int h = N * N;
    int m = e > g ? e : g;
    m = m > h ? m : h;
    m = (m + 3) & ~3;
    return m + ((npix * npix + 3) & ~3);
  }
};

template <int N, int R, int TPB, bool NOISE>
__global__ __launch_bounds__(TPB) void sh_image_fft_kernel(ImageParams P) {
  using L = FftLayout<N, R>;
  constexpr int M = L::M;
  constexpr int S = TPB / N;
  constexpr int PS = L::PS;
  extern __shared__ float smem[];

  const int pdiam = P.pdiam;
  const int PE = L::pe(pdiam);
  const int slot_f = L::slot_floats(pdiam, P.npix);
  int* s_binmap = (int*)(smem + S * slot_f);
  if (!P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += TPB) s_binmap[i] = P.binmap[i];
  }

  const int slot = threadIdx.x / N;
  const int lane = threadIdx.x % N;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * S + slot;
  const bool valid = s < P.nvalid;
  float* base = smem + slot * slot_f;
  float2* E = (float2*)base;  // E[y*PE + x]
  float2* G = (float2*)base;  // G[kx*PS + y]   (aliases E after a barrier)
  float* hr = base;           // hr[u + v*N]    (aliases G after a barrier)
  float* bin = base + (slot_f - ((P.npix * P.npix + 3) & ~3));

  // ---- 0. fields
  load_fields<N, L::EPL, false>(P, s, valid, lane, (float*)E, PE, 0);
  __syncthreads();

  // ---- 1. row transforms
  float re[M], im[M];
  const int y = lane / R;
  const int r = lane % R;
  const bool row_active = valid && (y < pdiam);
#pragma unroll
  for (int x = 0; x < M; ++x) {
    float2 v = make_float2(0.f, 0.f);
    if (row_active && x < pdiam) v = E[y * PE + x];
    re[x] = v.x;
    im[x] = v.y;
  }
  if constexpr (R == 2) {
    if (r) PreTw<N, M, 0>::run(re, im);
  }
  fft_reg<M>(re, im);
  __syncthreads();
  if (row_active) {
#pragma unroll
    for (int k = 0; k < M; ++k) G[(R * k + r) * PS + y] = make_float2(re[k], im[k]);
  }
  __syncthreads();

  // ---- 2. column transforms (lane = kx)
  float gr[M], gi[M];
#pragma unroll
  for (int yy = 0; yy < M; ++yy) {
    float2 v = make_float2(0.f, 0.f);
    if (valid && yy < pdiam) v = G[lane * PS + yy];
    gr[yy] = v.x;
    gi[yy] = v.y;
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < R; ++rr) {
    float zr[M], zi[M];
#pragma unroll
    for (int yy = 0; yy < M; ++yy) {
      zr[yy] = gr[yy];
      zi[yy] = gi[yy];
    }
    if (rr == 1) PreTw<N, M, 0>::run(zr, zi);
    fft_reg<M>(zr, zi);
    if (valid) {
#pragma unroll
      for (int k = 0; k < M; ++k) hr[lane + (R * k + rr) * N] = zr[k] * zr[k] + zi[k] * zi[k];
    }
  }
  __syncthreads();

  // ---- 3. tail
  spot_tail<N, N, NOISE>(P, s, valid, lane, hr, bin, s_binmap, 1.0f);
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
                                                         unsigned long long seed, unsigned call) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L)
    img[i] = sigma * (float)rng_normal((unsigned)i, call, 2u, seed);
}

int32_t launch_background(float* img, long n, float sigma, unsigned long long seed, unsigned call,
                          hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(background_kernel, dim3((unsigned)blocks), dim3(256), 0, st, img, n, sigma,
                     seed, call);
  SHWFS_LAUNCHED("background_kernel");
  return SHWFS_OK;
}

}  // namespace shwfs
