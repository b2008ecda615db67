// Spot kernel of the enlarged high-resolution grid and of the kernel
// convolution: the comp_generic branches with Ntot != Nfft and / or kernconv
// (libsutra/src.cpp/sutra_wfs_sh.cpp:409-481 and 514-525).  One workgroup of
// 256 lanes per sub-aperture; everything stays in LDS:
//
//   0. field E (NF x NF, zero-padded) -> LDS
//   1. 2-D forward FFT on NF (rows, then columns; one lane per line)
//   2. |F|^2 placed on the NT x NT grid: element k = u + v*NF goes to flat
//      index hrmap[k] of a zero grid (indexfill_krnl,
//      libsutra/src.cu/sutra_wfs.cu:462-479); the identity when NT == NF
//   3. CONV: forward FFT on NT, product with ftkernel (conv_krnl,
//      sutra_wfs.cu:491-513) and / or with this sub-aperture's LGS kernel
//      (convolve, sutra_utils.cu:553-564), unnormalised inverse FFT
//      (CarmaFFT dir -1), real part
//   4. binning through binmap (indices on the NT grid, fillbincube_krnl,
//      sutra_wfs.cu:369-414), then the shared tail (flux, normalisation,
//      noise, bincube / binimg, fused centroider) by the first wave.
// The reference runs these over chunks of nmaxhr sub-apertures to bound the
// memory of its NT^2 complex cube; nothing is materialised here, so there
// are no chunks (the chunking does not change the values).
#include "shwfs_image_fft.hpp"

namespace shwfs {

constexpr int kExtThreads = 256;

// One NxN complex array in LDS, row y at a[y * (N + 1)] (odd stride: the
// column passes are free of bank conflicts).  Every line of the given
// direction is transformed by one lane: rows (x) when ROWS, else columns (y).
// INV: unnormalised inverse transform, as conj(FFT(conj(.))).
template <int N, bool ROWS, bool INV>
__device__ __forceinline__ void lds_fft_lines(float2* a, int tid) {
  constexpr int S = N + 1;
  if (tid < N) {
    float re[N], im[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const float2 v = ROWS ? a[tid * S + k] : a[k * S + tid];
      re[k] = v.x;
      im[k] = INV ? -v.y : v.y;
    }
    fft_reg<N>(re, im);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const float2 v = make_float2(re[k], INV ? -im[k] : im[k]);
      if (ROWS)
        a[tid * S + k] = v;
      else
        a[k * S + tid] = v;
    }
  }
  __syncthreads();
}

template <int NF, int NT, bool CONV, bool NOISE>
__global__ __launch_bounds__(kExtThreads) void sh_image_ext_kernel(ImageParams P, ExtParams X) {
  extern __shared__ float smem[];
  constexpr int SF = NF + 1, ST = NT + 1;
  float2* A = reinterpret_cast<float2*>(smem);            // NF x (NF+1)
  float2* B = A + NF * SF;                                 // NT x (NT+1) (CONV or NT != NF)
  float* bin = reinterpret_cast<float*>(B + NT * ST);      // npix^2
  const int tid = threadIdx.x;
  const int s = xcd_remap(blockIdx.x, gridDim.x);
  const bool valid = s < P.nvalid;  // always (one workgroup per sub-aperture)

  // ---- 0. field, zero-padded
  for (int i = tid; i < NF * SF; i += kExtThreads) A[i] = make_float2(0.f, 0.f);
  for (int i = tid; i < NT * ST; i += kExtThreads) B[i] = make_float2(0.f, 0.f);
  __syncthreads();
  constexpr int EPL = (NF * NF + kExtThreads - 1) / kExtThreads;
  load_fields<kExtThreads, EPL, false>(P, s, valid, tid, reinterpret_cast<float*>(A), SF, 0);
  __syncthreads();

  // ---- 1. forward FFT on the NF grid: rows (along x), then columns (along y)
  lds_fft_lines<NF, true, false>(A, tid);
  lds_fft_lines<NF, false, false>(A, tid);

  // ---- 2. |F|^2 onto the NT grid (element (u, v) of F at A[v * SF + u])
  for (int k = tid; k < NF * NF; k += kExtThreads) {
    const int u = k % NF, v = k / NF;
    const float2 f = A[v * SF + u];
    const float h = f.x * f.x + f.y * f.y;
    const int t = (NT == NF) ? k : X.hrmap[k];
    B[(t / NT) * ST + (t % NT)] = make_float2(h, 0.f);
  }
  __syncthreads();

  // ---- 3. kernel convolution on the NT grid
  if constexpr (CONV) {
    lds_fft_lines<NT, true, false>(B, tid);
    lds_fft_lines<NT, false, false>(B, tid);
    const float2* lk = X.ftlgs ? X.ftlgs + (size_t)s * NT * NT : nullptr;
    for (int k = tid; k < NT * NT; k += kExtThreads) {
      const int u = k % NT, v = k / NT;
      const float2 a = B[v * ST + u];
      float2 w;
      if (lk && X.ftkernel) {  // both kernels: one product, times kscale
        const float2 l = lk[k], c = X.ftkernel[k];
        w = make_float2((l.x * c.x - l.y * c.y) * X.kscale, (l.x * c.y + l.y * c.x) * X.kscale);
      } else {
        w = lk ? lk[k] : X.ftkernel[k];
      }
      B[v * ST + u] = make_float2(w.x * a.x - w.y * a.y, w.y * a.x + w.x * a.y);
    }
    __syncthreads();
    lds_fft_lines<NT, true, true>(B, tid);
    lds_fft_lines<NT, false, true>(B, tid);
  }

  // ---- 4. binning (real part) and the tail (first wave)
  const int npix2 = P.npix * P.npix, nreb2 = P.nrebin * P.nrebin;
  for (int p = tid; p < npix2; p += kExtThreads) {
    const int* bm = P.binmap + p * nreb2;
    float b = 0.0f;
    for (int i = 0; i < nreb2; ++i) {
      const int t = bm[i];
      b += B[(t / NT) * ST + (t % NT)].x;
    }
    bin[p] = b;
  }
  __syncthreads();
  if (tid < 64) {
    float part = 0.0f;
    for (int p = tid; p < npix2; p += 64) part += bin[p];
    const float I = lane_group_sum<64>(part);
    finish_spot<64, NOISE>(P, s, valid, tid, bin, I, 1.0f);
  }
}

template <int NF, int NT>
static int32_t launch_ext_nt(const ImageParams& p, const ExtParams& x, hipStream_t st) {
  const size_t lds = sizeof(float2) * ((size_t)NF * (NF + 1) + (size_t)NT * (NT + 1)) +
                     sizeof(float) * (size_t)((p.npix * p.npix + 3) & ~3);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "enlarged-grid spot kernel: LDS footprint exceeds the per-workgroup limit");
  const bool noisy = p.noise > -1.0f;
  auto k = x.conv ? (noisy ? sh_image_ext_kernel<NF, NT, true, true> : sh_image_ext_kernel<NF, NT, true, false>)
                  : (noisy ? sh_image_ext_kernel<NF, NT, false, true> : sh_image_ext_kernel<NF, NT, false, false>);
  hipLaunchKernelGGL(k, dim3(p.nvalid), dim3(kExtThreads), lds, st, p, x);
  SHWFS_LAUNCHED("sh_image_ext_kernel");
  return SHWFS_OK;
}

int32_t launch_image_ext(const ImageParams& p, const ExtParams& x, int nfft, int ntot,
                         hipStream_t st) {
  if (p.nvalid <= 0) return SHWFS_OK;
#define SHWFS_EXT(NF, NT) \
  if (nfft == NF && ntot == NT) return launch_ext_nt<NF, NT>(p, x, st);
  SHWFS_EXT(16, 16) SHWFS_EXT(16, 32) SHWFS_EXT(16, 64)
  SHWFS_EXT(32, 32) SHWFS_EXT(32, 64)
  SHWFS_EXT(64, 64)
#undef SHWFS_EXT
  return fail(SHWFS_ERR_UNSUPPORTED,
              "enlarged grid / kernel convolution: needs nfft in {16, 32, 64}, ntot in {nfft, ..., 64}");
}

// ---- LGS elongated kernels: image, rotation, forward transform -------------
// One workgroup per sub-aperture (SutraLGS::lgs_makespot, libsutra/src.cpp/
// sutra_lgs.cpp:180-212): rollbeamexp_krnl (sutra_lgs.cu:106-129) then
// rotate_krnl (155-226) into an NT x NT image, 2-D forward transform.

struct LgsKernArgs {
  float2* ftlgs;         // [NT^2 * nvalid]
  float* lgskern;        // [NT^2 * nvalid] or NULL
  const float2* prof2d;  // [NT * nvalid] or NULL (transform lgskern as given)
  const float* beam;     // [NT]
  const float* azimuth;  // [nvalid]
};

// bilinear sample index and weight along one axis, as rotate_krnl clamps it
__device__ __forceinline__ void lgs_axis(float t, int n, int& ref, float& w) {
  if (t < 0.0f) {
    ref = 0;
    w = 0.0f;
  } else if (t < (float)(n - 2)) {
    ref = (int)t;
    w = t - (float)ref;
  } else {
    ref = n - 2;
    w = 1.0f;
  }
}

template <int NT>
__global__ __launch_bounds__(kExtThreads) void lgs_kernels_kernel(LgsKernArgs a) {
  constexpr int ST = NT + 1;
  __shared__ float img[NT * NT];
  __shared__ float2 B[NT * ST];
  const int s = blockIdx.x, tid = threadIdx.x;
  const size_t base = (size_t)s * NT * NT;
  if (a.prof2d) {
    const float2* pr = a.prof2d + (size_t)s * NT;
    for (int k = tid; k < NT * NT; k += kExtThreads) {
      const int x = k % NT, y = k / NT;
      img[k] = fabsf(pr[(x + NT / 2) % NT].x) * a.beam[y];
    }
    __syncthreads();
    const float th = a.azimuth[s];
    const float cs = cosf(th), sn = sinf(th);
    const float cent = (float)(NT / 2);  // width / 2 - center, center 0
    for (int k = tid; k < NT * NT; k += kExtThreads) {
      const int x = k % NT, y = k / NT;
      const float u = (float)x - cent, v = (float)y - cent;
      const float tu = u * cs + v * sn + cent;
      const float tv = -u * sn + v * cs + cent;
      int ur, vr;
      float wu, wv;
      lgs_axis(tu, NT, ur, wu);
      lgs_axis(tv, NT, vr, wv);
      const float i1 = img[ur + vr * NT], i2 = img[ur + 1 + vr * NT];
      const float i3 = img[ur + (vr + 1) * NT], i4 = img[ur + 1 + (vr + 1) * NT];
      const float val = (i1 * (1 - wu) + i2 * wu) * (1 - wv) + (i3 * (1 - wu) + i4 * wu) * wv;
      B[y * ST + x] = make_float2(val, 0.0f);
      if (a.lgskern) a.lgskern[base + k] = val;
    }
  } else {
    for (int k = tid; k < NT * NT; k += kExtThreads)
      B[(k / NT) * ST + (k % NT)] = make_float2(a.lgskern[base + k], 0.0f);
  }
  __syncthreads();
  lds_fft_lines<NT, true, false>(B, tid);
  lds_fft_lines<NT, false, false>(B, tid);
  for (int k = tid; k < NT * NT; k += kExtThreads) a.ftlgs[base + k] = B[(k / NT) * ST + (k % NT)];
}

int32_t launch_lgs_kernels(float2* ftlgs, float* lgskern, const float2* prof2d, const float* beam,
                           const float* azimuth, int ntot, int nvalid, hipStream_t st) {
  const LgsKernArgs a{ftlgs, lgskern, prof2d, beam, azimuth};
  if (nvalid <= 0) return SHWFS_OK;
  if (ntot == 16)
    hipLaunchKernelGGL(lgs_kernels_kernel<16>, dim3(nvalid), dim3(kExtThreads), 0, st, a);
  else if (ntot == 32)
    hipLaunchKernelGGL(lgs_kernels_kernel<32>, dim3(nvalid), dim3(kExtThreads), 0, st, a);
  else if (ntot == 64)
    hipLaunchKernelGGL(lgs_kernels_kernel<64>, dim3(nvalid), dim3(kExtThreads), 0, st, a);
  else
    return fail(SHWFS_ERR_UNSUPPORTED, "lgs_kernels: ntot must be 16, 32 or 64");
  SHWFS_LAUNCHED("lgs_kernels_kernel");
  return SHWFS_OK;
}

}  // namespace shwfs
