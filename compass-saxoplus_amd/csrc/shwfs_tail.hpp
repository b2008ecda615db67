// Per-subaperture tail shared by the FFT and the MFMA spot kernels:
// binning through binmap (LDS gather), per-subap flux, photometric
// normalisation, Philox noise, bincube / binimg writes and an optional fused
// CoG-family centroider.  One group of N consecutive lanes owns one
// sub-aperture; N is a power of two <= 64 so a group never spans two waves.
#pragma once

#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"

namespace shwfs {

template <int N, typename Acc>
__device__ __forceinline__ void centroid_finish(const ImageParams& P, int s, bool valid, int lane,
                                                Acc I, Acc Sx, Acc Sy) {
  I = group_sum<N>(I);
  Sx = group_sum<N>(Sx);
  Sy = group_sum<N>(Sy);
  if (valid && lane == 0) {
    const int ix = P.slope_order == 0 ? s : 2 * s;
    const int iy = P.slope_order == 0 ? s + P.nvalid : 2 * s + 1;
    const float rx = P.ref ? P.ref[ix] : 0.0f;
    const float ry = P.ref ? P.ref[iy] : 0.0f;
    // As the reference: the division is done in double (1.e-6 is a double
    // literal), the rest in float.  With float64 moments the sums enter exact.
    P.slopes[ix] = ((float)((double)Sx / ((double)I + 1.e-6)) - P.coffset) * P.cscale - rx;
    P.slopes[iy] = ((float)((double)Sy / ((double)I + 1.e-6)) - P.coffset) * P.cscale - ry;
    P.cintens[s] = (float)I;
  }
}

// hr: the N x N |E|^2 image of this sub-aperture in LDS, flat index u + v*N.
// bin: npix^2 floats of LDS private to this sub-aperture.
template <int N, bool NOISE>
__device__ __forceinline__ void spot_tail(const ImageParams& P, int s, bool valid, int lane,
                                          const float* hr, float* bin, const int* s_binmap) {
  const int npix2 = P.npix * P.npix;
  const int nreb2 = P.nrebin * P.nrebin;
  float part = 0.0f;
  if (valid) {
    for (int p = lane; p < npix2; p += N) {
      const int* bm = s_binmap + p * nreb2;
      float b = 0.0f;
      for (int i = 0; i < nreb2; ++i) b += hr[bm[i]];
      bin[p] = b;
      part += b;
    }
  }
  const float I = group_sum<N>(part);
  const float fl = valid ? P.flux[s] : 0.0f;
  const int vx = valid ? P.validsubsx[s] : 0;
  const int vy = valid ? P.validsubsy[s] : 0;
  float cI = 0.0f, cX = 0.0f, cY = 0.0f;
  double dI = 0.0, dX = 0.0, dY = 0.0;
  if (valid) {
    if (lane == 0) P.intensities[s] = I;
    for (int p = lane; p < npix2; p += N) {
      float b = bin[p];
      if (I != 0.0f) b = b * fl / I * P.nphot;
      const unsigned e = (unsigned)(p + s * npix2);
      if constexpr (NOISE) {
        if (P.noise > -1.0f) b = (float)rng_poisson((double)b, e, P.call, 0u, P.seed);
        if (P.noise > 0.0f) b = 1.0f * b + P.noise * (float)rng_normal(e, P.call, 1u, P.seed);
      }
      P.bincube[e] = b;
      const int px = p % P.npix, py = p / P.npix;
      float out = b;
      if (P.ckind != SHWFS_CENTROID_NONE) {
        float d = b;
        if (P.ckind != SHWFS_CENTROID_COG) d = (b > P.threshold) ? b - P.threshold : 1.e-6f;
        if (P.ckind == SHWFS_CENTROID_TCOG) out = d;
        const float fx = (float)px, fy = (float)py;
        if (P.ckind == SHWFS_CENTROID_WCOG) {
          const float w = P.weights[e];
          if (P.fp64) {
            dI += (double)d * w;
            dX += (double)d * fx * w;
            dY += (double)d * fy * w;
          } else {
            cI += d * w;
            cX += d * fx * w;
            cY += d * fy * w;
          }
        } else {
          if (P.fp64) {
            dI += (double)d;
            dX += (double)d * fx;
            dY += (double)d * fy;
          } else {
            cI += d;
            cX += d * fx;
            cY += d * fy;
          }
        }
      }
      if (P.binimg) P.binimg[(vx + px) + (size_t)(vy + py) * P.nsub] = out;
    }
  }
  if (P.ckind != SHWFS_CENTROID_NONE) {
    if (P.fp64)
      centroid_finish<N, double>(P, s, valid, lane, dI, dX, dY);
    else
      centroid_finish<N, float>(P, s, valid, lane, cI, cX, cY);
  }
}

}  // namespace shwfs
