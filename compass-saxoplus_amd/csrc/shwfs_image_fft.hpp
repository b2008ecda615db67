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
#pragma once

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

// LDS hand-over between the lanes of one tile: a workgroup barrier, or for a
// tile of one wave (TPB == 64) the wave's own ordering (other waves of the
// workgroup, working on other tiles, are not waited for).
template <int TPB>
__device__ __forceinline__ void tile_sync() {
  if constexpr (TPB == 64)
    wave_lds_sync();
  else
    __syncthreads();
}

// One tile: the TPB / N sub-apertures s0, s0 + 1, ... of TPB lanes (group
// `slot` of N lanes owns s0 + slot); tid: the lane's index in the tile (the
// thread index, or the lane in the wave for one-wave tiles).  smem: the
// tile's LDS (S slots of slot_floats, then the binmap if bm is not the global
// table).  Contains tile barriers (tile_sync): every lane of the tile calls
// it.  call_add: added to the Philox call counter of the noise.
// KEEP: the tail leaves the spot-cube values of each sub-aperture in its
// slot's spot area (the last npix^2 floats, 4-aligned, of the slot) for a
// centroider that runs after the tile.  RAY: the window phases are traced
// through the layers of *ray (guide-star raytrace in the field stage).
template <int N, int R, int TPB, bool NOISE, bool KEEP = false, bool RAY = false>
__device__ __forceinline__ void fft_tile(const ImageParams& P, int s0, float* smem,
                                         const int* bm, unsigned call_add = 0u,
                                         int tid = (int)threadIdx.x,
                                         const RayParams* ray = nullptr) {
  using L = FftLayout<N, R>;
  constexpr int M = L::M;
  constexpr int PS = L::PS;
  const int pdiam = P.pdiam;
  const int PE = L::pe(pdiam);
  const int slot_f = L::slot_floats(pdiam, P.npix);
  const int slot = tid / N;
  const int lane = tid % N;
  const int s = s0 + slot;
  const bool valid = s < P.nvalid;
  float* base = smem + slot * slot_f;
  float2* E = (float2*)base;  // E[y*PE + x]
  float2* G = (float2*)base;  // G[kx*PS + y]   (aliases E after a barrier)
  float* hr = base;           // hr[u + v*N]    (aliases G after a barrier)
  float* bin = base + (slot_f - ((P.npix * P.npix + 3) & ~3));

  // ---- 0. fields
  load_fields<N, L::EPL, false, false, 0, RAY>(P, s, valid, lane, (float*)E, PE, 0, ray);
  tile_sync<TPB>();

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
  tile_sync<TPB>();
  if (row_active) {
#pragma unroll
    for (int k = 0; k < M; ++k) G[(R * k + r) * PS + y] = make_float2(re[k], im[k]);
  }
  tile_sync<TPB>();

  // ---- 2. column transforms (lane = kx)
  float gr[M], gi[M];
#pragma unroll
  for (int yy = 0; yy < M; ++yy) {
    float2 v = make_float2(0.f, 0.f);
    if (valid && yy < pdiam) v = G[lane * PS + yy];
    gr[yy] = v.x;
    gi[yy] = v.y;
  }
  tile_sync<TPB>();
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
  tile_sync<TPB>();

  // ---- 3. tail
  spot_tail<N, N, NOISE, KEEP>(P, s, valid, lane, hr, bin, bm, 1.0f, call_add);
  tile_sync<TPB>();
}

}  // namespace shwfs
