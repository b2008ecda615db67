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
#include <map>
#include <mutex>

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

// The A_W operand as per-lane fragments, built once per device and kept in a
// small device table: a wave loads its MT*KS fragments with coalesced vector
// loads instead of rebuilding them.
template <int N, int Q, int PREC>
__global__ void build_wfrag_kernel(Frag<PREC>* out) {
  using L = DftLayout<N, Q>;
  const int lane = threadIdx.x;
  for (int mt = 0; mt < L::MT; ++mt)
    for (int ks = 0; ks < L::KS; ++ks) {
      Frag<PREC> f;
#pragma unroll
      for (int e = 0; e < 8; ++e) f.set(e, aw_value<N, Q>(mt, ks, lane, e));
      out[(mt * L::KS + ks) * 64 + lane] = f;
    }
}

template <int N, int Q, int PREC>
static const Frag<PREC>* wfrag_table(hipStream_t st) {
  static std::mutex mu;
  static std::map<int, Frag<PREC>*> tables;  // device -> table
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = tables.find(dev);
  if (it != tables.end()) return it->second;
  using L = DftLayout<N, Q>;
  Frag<PREC>* t = nullptr;
  if (hipMalloc((void**)&t, sizeof(Frag<PREC>) * L::MT * L::KS * 64) != hipSuccess) return nullptr;
  hipLaunchKernelGGL((build_wfrag_kernel<N, Q, PREC>), dim3(1), dim3(64), 0, st, t);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    (void)hipFree(t);
    return nullptr;
  }
  tables[dev] = t;
  return t;
}

// Register binning of the |F|^2 values a lane holds after stage B (kx of the
// lane, four consecutive ky per accumulator): the NR x NR block of a binned
// pixel is NR consecutive v of one lane times NR consecutive lanes (NR in
// {1, 2, 4}, block origin a multiple of NR).  Summation: within the lane
// over ky, then across lanes over kx.
template <int N, int NR>
struct RegBin {
  static constexpr int MT = N / 8, NTB = N / 16;
  static constexpr int VPL = (MT / 2) * NTB * (4 / NR);  // binned values a lane may own
};

template <int N, int Q, int PREC, int WPB, bool NOISE, int NR>
__global__ __launch_bounds__(64 * WPB) void sh_image_dft_kernel(ImageParams P, int iters,
                                                                const Frag<PREC>* wtab) {
  using L = DftLayout<N, Q>;
  using F = Frag<PREC>;
  constexpr int TPB = 64 * WPB;
  extern __shared__ float smem[];

  const int pdiam = P.pdiam;
  const int wave_f = L::wave_floats(P.npix);
  int* s_binmap = (int*)(smem + WPB * wave_f);
  if (NR == 0 && !P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += TPB) s_binmap[i] = P.binmap[i];
  }
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // persistent waves: wave gw handles sub-apertures gw, gw + TW, ... ; blocks
  // of one XCD hold contiguous gw
  const int gw = xcd_remap(blockIdx.x, gridDim.x) * WPB + w;
  const int TW = gridDim.x * WPB;
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
    for (int ks = 0; ks < L::KS; ++ks) aw[mt][ks] = wtab[(mt * L::KS + ks) * 64 + lane];
  const bool full = (pdiam == L::P16);

  // every wave of the workgroup runs `iters` iterations (barriers stay uniform)
  for (int it = 0; it < iters; ++it) {
    const int s = gw + it * TW;
    const bool valid = s < P.nvalid;

    // ---- 0. fields
    load_fields<64, L::EPL, true>(P, s, valid, lane, E, L::PE, L::EPLANE);
    __syncthreads();

    // ---- stage A operands: B[k][y] = E{r,i}[x][y], k = (re/im, x)
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
      gmax = wave_max(gmax);
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
    const float inten_scale = 1.0f / (gscale * gscale);

    if constexpr (NR == 0) {
      // ---- stage B: F^T = W G^T, |F|^2 -> hr, shared tail (any binmap)
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
      spot_tail<64, N, NOISE>(P, s, valid, lane, hr, bin, s_binmap, inten_scale);
      __syncthreads();
    } else {
      // ---- stage B with the binning done on the accumulators
      using RB = RegBin<N, NR>;
      float bv[RB::VPL];
      int bp[RB::VPL];
      const int o = P.bin_origin, npix = P.npix;
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
        const int u = (kx - o) & (N - 1);
#pragma unroll
        for (int mt = 0; mt < L::MT / 2; ++mt) {
          float iv[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float fr = fb[mt][v], fi = fb[mt + L::MT / 2][v];
            iv[v] = fr * fr + fi * fi;
          }
#pragma unroll
          for (int g = 0; g < 4 / NR; ++g) {
            float b = 0.0f;
#pragma unroll
            for (int v = 0; v < NR; ++v) b += iv[g * NR + v];
            if constexpr (NR >= 2) b += __shfl_xor(b, 1, 64);
            if constexpr (NR >= 4) b += __shfl_xor(b, 2, 64);
            const int wv = (16 * mt + 4 * (lane >> 4) + g * NR - o) & (N - 1);
            const int px = u / NR, py = wv / NR;
            const bool own = ((lane & (NR - 1)) == 0) && px < npix && py < npix;
            const int k = (nt * (L::MT / 2) + mt) * (4 / NR) + g;
            bv[k] = b;
            bp[k] = own ? px + py * npix : -1;
          }
        }
      }
      // ---- tail on the register-held spot
      float part = 0.0f;
#pragma unroll
      for (int k = 0; k < RB::VPL; ++k) part += (bp[k] >= 0) ? bv[k] : 0.0f;
      const float I = wave_sum(part);
      const int npix2 = npix * npix;
      const float fl = valid ? P.flux[s] : 0.0f;
      const int vx = valid ? P.validsubsx[s] : 0;
      const int vy = valid ? P.validsubsy[s] : 0;
      float cI = 0.f, cX = 0.f, cY = 0.f;
      double dI = 0.0, dX = 0.0, dY = 0.0;
      if (valid) {
        if (lane == 0) P.intensities[s] = I * inten_scale;
#pragma unroll
        for (int k = 0; k < RB::VPL; ++k) {
          if (bp[k] < 0) continue;
          const int p = bp[k];
          float b = bv[k];
          if (I != 0.0f) b = b * fl / I * P.nphot;
          const unsigned e = (unsigned)(p + s * npix2);
          if constexpr (NOISE) {
            if (P.noise > -1.0f) b = (float)rng_poisson((double)b, e, P.call, 0u, P.seed);
            if (P.noise > 0.0f) b = 1.0f * b + P.noise * (float)rng_normal(e, P.call, 1u, P.seed);
          }
          P.bincube[e] = b;
          const int px = p % npix, py = p / npix;
          float out = b;
          if (P.ckind != SHWFS_CENTROID_NONE) {
            float d = b;
            if (P.ckind != SHWFS_CENTROID_COG) d = (b > P.threshold) ? b - P.threshold : 1.e-6f;
            if (P.ckind == SHWFS_CENTROID_TCOG) out = d;
            const float fx = (float)px, fy = (float)py;
            const float wgt = (P.ckind == SHWFS_CENTROID_WCOG) ? P.weights[e] : 1.0f;
            if (P.fp64) {
              dI += (double)d * wgt;
              dX += (double)d * fx * wgt;
              dY += (double)d * fy * wgt;
            } else {
              cI += d * wgt;
              cX += d * fx * wgt;
              cY += d * fy * wgt;
            }
          }
          if (P.binimg) P.binimg[(vx + px) + (size_t)(vy + py) * P.nsub] = out;
        }
      }
      if (P.ckind != SHWFS_CENTROID_NONE) {
        if (P.fp64)
          centroid_finish<64, double>(P, s, valid, lane, dI, dX, dY);
        else
          centroid_finish<64, float>(P, s, valid, lane, cI, cX, cY);
      }
    }
  }
}

// Resident waves per CU the persistent grid is sized for.
constexpr int kDftWavesPerCU = 16;

template <int N, int Q, int PREC, int WPB, int NR>
static int32_t launch_dft_t(const ImageParams& p, hipStream_t st) {
  using L = DftLayout<N, Q>;
  const int nbm = (NR == 0 && !p.bin_block) ? p.nrebin * p.nrebin * p.npix * p.npix : 0;
  const size_t lds = sizeof(float) * ((size_t)WPB * L::wave_floats(p.npix) + nbm);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "DFT spot kernel: LDS footprint exceeds the per-workgroup limit");
  // grid: as many workgroups as fit on the device at once (capped by the
  // work), so a frame normally runs in a single pass of resident waves
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const bool noisy = p.noise > -1.0f;
  auto kern = noisy ? sh_image_dft_kernel<N, Q, PREC, WPB, true, NR>
                    : sh_image_dft_kernel<N, Q, PREC, WPB, false, NR>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WPB, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = kDftWavesPerCU / WPB > 0 ? kDftWavesPerCU / WPB : 1;
  const int needed = (p.nvalid + WPB - 1) / WPB;
  int grid = cus * per_cu;
  if (grid > needed) grid = needed;
  if (grid < 1) grid = 1;
  const int iters = (p.nvalid + grid * WPB - 1) / (grid * WPB);
  const Frag<PREC>* wtab = wfrag_table<N, Q, PREC>(st);
  if (!wtab) return fail(SHWFS_ERR_HIP, "DFT spot kernel: building the DFT operand table failed");
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WPB), lds, st, p, iters, wtab);
  SHWFS_LAUNCHED("sh_image_dft_kernel");
  return SHWFS_OK;
}

template <int N, int Q, int PREC, int WPB>
static int32_t launch_dft_nr(const ImageParams& p, hipStream_t st) {
  // register binning when the binmap is the NGS block form with a block
  // origin aligned on the rebin factor; otherwise the LDS gather
  const bool blk = p.bin_block && (p.bin_origin % (p.nrebin > 0 ? p.nrebin : 1)) == 0;
  if (blk && p.nrebin == 4) return launch_dft_t<N, Q, PREC, WPB, 4>(p, st);
  if (blk && p.nrebin == 2) return launch_dft_t<N, Q, PREC, WPB, 2>(p, st);
  if (blk && p.nrebin == 1) return launch_dft_t<N, Q, PREC, WPB, 1>(p, st);
  return launch_dft_t<N, Q, PREC, WPB, 0>(p, st);
}

template <int N, int Q, int PREC>
static int32_t launch_dft_w(const ImageParams& p, hipStream_t st) {
  // default: one wave per workgroup (its barriers cost nothing)
  if (p.tpb == 128) return launch_dft_nr<N, Q, PREC, 2>(p, st);
  if (p.tpb == 256) return launch_dft_nr<N, Q, PREC, 4>(p, st);
  return launch_dft_nr<N, Q, PREC, 1>(p, st);
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
