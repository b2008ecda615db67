// Stages shared by the FFT and the MFMA spot kernels.  One group of LANES
// consecutive lanes owns one sub-aperture (LANES a power of two <= 64, so a
// group never spans two waves); NF is the FFT support.
//
//  * load_fields: phase / pupil window -> complex field in LDS
//  * spot_tail:   binning, per-subap flux, photometric normalisation, Philox
//                 noise, bincube / binimg writes, optional fused centroider
#pragma once

#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"

namespace shwfs {

// Logical block index with consecutive blocks kept on one XCD: the hardware
// hands block b to XCD b % 8; remapping gives each XCD a contiguous range of
// sub-apertures, so screen rows shared by neighbouring sub-apertures are read
// once into that XCD's L2 instead of once per XCD.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int kXcd = 8;
  const int q = nblocks / kXcd, r = nblocks % kXcd;
  const int x = b % kXcd;
  return x * q + (x < r ? x : r) + b / kXcd;
}

// Field E(x, y) = mask * exp(i (phase*scale - offset)) of one sub-aperture.
// PLANAR: Er at E[y*pe + x], Ei at E[plane + y*pe + x] (floats); otherwise
// interleaved float2 at E2[y*pe + x].  All global loads of a chunk are issued
// before the first use.
template <int LANES, int EPL, bool PLANAR>
__device__ __forceinline__ void load_fields(const ImageParams& P, int s, bool valid, int lane,
                                            float* E, int pe, int plane) {
  const int pdiam = P.pdiam;
  const int np2 = pdiam * pdiam;
  constexpr int CH = EPL < 16 ? EPL : 16;  // elements in flight per lane (register bound)
  if (valid) {
    const int i0 = P.istart[s], j0 = P.jstart[s];
    int x = lane % pdiam, y = lane / pdiam;
    const int dx = LANES % pdiam, dy = LANES / pdiam;
    for (int c0 = 0; c0 < EPL; c0 += CH) {
      float ph[CH], pu[CH], of[CH];
      int xs[CH], ys[CH];
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        const int idx = lane + (c0 + e) * LANES;
        xs[e] = x;
        ys[e] = y;
        if (idx < np2) {
          const size_t g = (size_t)(i0 + x) + (size_t)(j0 + y) * P.npup;
          ph[e] = P.phase[g];
          pu[e] = P.pupil[g];
          of[e] = P.offsets[idx];
        }
        x += dx;
        y += dy;
        if (x >= pdiam) {
          x -= pdiam;
          y += 1;
        }
      }
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        if (lane + (c0 + e) * LANES < np2) {
          float sn, cs;
          phase_sincos(ph[e], P.scale, of[e], sn, cs);
          if constexpr (PLANAR) {
            E[ys[e] * pe + xs[e]] = cs * pu[e];
            E[plane + ys[e] * pe + xs[e]] = sn * pu[e];
          } else {
            reinterpret_cast<float2*>(E)[ys[e] * pe + xs[e]] = make_float2(cs * pu[e], sn * pu[e]);
          }
        }
      }
    }
  }
}

template <int LANES, typename Acc>
__device__ __forceinline__ void centroid_finish(const ImageParams& P, int s, bool valid, int lane,
                                                Acc I, Acc Sx, Acc Sy) {
  I = lane_group_sum<LANES>(I);
  Sx = lane_group_sum<LANES>(Sx);
  Sy = lane_group_sum<LANES>(Sy);
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

// hr: the NF x NF |E|^2 image of this sub-aperture in LDS, flat index u + v*NF,
// holding inten_scale^-1 times the true intensities (a power of two).
// bin: npix^2 floats of LDS private to this sub-aperture.
template <int LANES, int NF, bool NOISE>
__device__ __forceinline__ void spot_tail(const ImageParams& P, int s, bool valid, int lane,
                                          const float* hr, float* bin, const int* s_binmap,
                                          float inten_scale) {
  const int npix2 = P.npix * P.npix;
  const int nreb2 = P.nrebin * P.nrebin;
  float part = 0.0f;
  if (valid) {
    if (P.bin_block) {
      // binned pixel (px, py) sums the nrebin x nrebin HR block starting at
      // ((o + px*nrebin) mod NF, (o + py*nrebin) mod NF), row (v) slow, column fast
      const int nr = P.nrebin;
      for (int p = lane; p < npix2; p += LANES) {
        const int u0 = P.bin_origin + (p % P.npix) * nr;
        const int v0 = P.bin_origin + (p / P.npix) * nr;
        float b = 0.0f;
        for (int bb = 0; bb < nr; ++bb) {
          const float* row = hr + ((v0 + bb) & (NF - 1)) * NF;
          for (int a = 0; a < nr; ++a) b += row[(u0 + a) & (NF - 1)];
        }
        bin[p] = b;
        part += b;
      }
    } else {
      for (int p = lane; p < npix2; p += LANES) {
        const int* bm = s_binmap + p * nreb2;
        float b = 0.0f;
        for (int i = 0; i < nreb2; ++i) b += hr[bm[i]];
        bin[p] = b;
        part += b;
      }
    }
  }
  const float I = lane_group_sum<LANES>(part);
  const float fl = valid ? P.flux[s] : 0.0f;
  const int vx = valid ? P.validsubsx[s] : 0;
  const int vy = valid ? P.validsubsy[s] : 0;
  float cI = 0.0f, cX = 0.0f, cY = 0.0f;
  double dI = 0.0, dX = 0.0, dY = 0.0;
  if (valid) {
    if (lane == 0) P.intensities[s] = I * inten_scale;
    for (int p = lane; p < npix2; p += LANES) {
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
      centroid_finish<LANES, double>(P, s, valid, lane, dI, dX, dY);
    else
      centroid_finish<LANES, float>(P, s, valid, lane, cI, cX, cY);
  }
}

}  // namespace shwfs
