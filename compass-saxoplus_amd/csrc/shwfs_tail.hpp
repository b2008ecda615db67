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
#include "shwfs_raytrace.hpp"

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

// exp(i theta) for theta = phase*scale - offset.  FAST: the hardware sine and
// cosine on the angle in turns (theta / 2 pi reduced to [-1/2, 1/2]), about
// 1e-6 absolute error for the phase ranges of a sub-aperture; otherwise the
// float-exact reduction of phase_sincos.
template <bool FAST>
__device__ __forceinline__ void field_sincos(float ph, float scale, float turn_scale, float off,
                                             float& sn, float& cs) {
  if constexpr (FAST) {
    constexpr float kInv2Pi = 0.15915494309189533577f;
    float r = fmaf(ph, turn_scale, -off * kInv2Pi);
    r -= rintf(r);
    sn = __builtin_amdgcn_sinf(r);
    cs = __builtin_amdgcn_cosf(r);
  } else {
    phase_sincos(ph, scale, off, sn, cs);
  }
}

// Field E(x, y) = mask * exp(i (phase*scale - offset)) of one sub-aperture.
// PLANAR: Er at E[y*pe + x], Ei at E[plane + y*pe + x] (floats); otherwise
// interleaved float2 at E2[y*pe + x].  All global loads of a chunk are issued
// unconditionally (addresses clamped into the window) before the first use,
// so their latencies overlap.  RAY: the phase of each window point is traced
// through the layers of *ray (ray_phase) instead of read from P.phase.
template <int LANES, int EPL, bool PLANAR, bool FAST = false, int PD = 0, bool RAY = false>
__device__ __forceinline__ void load_fields(const ImageParams& P, int s, bool valid, int lane,
                                            float* E, int pe, int plane,
                                            const RayParams* ray = nullptr) {
  const int pdiam = PD > 0 ? PD : P.pdiam;  // PD: pdiam known at compile time
  const int np2 = pdiam * pdiam;
  if (lane >= LANES) __builtin_unreachable();
  constexpr int CH = EPL < 16 ? EPL : 16;  // elements in flight per lane (register bound)
  if (valid) {
    const int npup = P.npup;
    const float* phase = P.phase + P.istart[s] + (size_t)P.jstart[s] * npup;
    const float* pupil = P.pupil + P.istart[s] + (size_t)P.jstart[s] * npup;
    const bool unit = P.pupil_unit != nullptr && P.pupil_unit[s] != 0;  // window all ones
    const float turn_scale = P.scale * 0.15915494309189533577f;
    int x = lane % pdiam, y = lane / pdiam;
    const int dx = LANES % pdiam, dy = LANES / pdiam;
    for (int c0 = 0; c0 < EPL; c0 += CH) {
      float ph[CH], pu[CH], of[CH];
      int xs[CH], ys[CH];
      unsigned gs[CH];  // non-negative window offsets: 32-bit vector offsets on a scalar base
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        const int idx = lane + (c0 + e) * LANES;
        const bool in = idx < np2;
        xs[e] = x;
        ys[e] = y;
        const int yc = in ? y : pdiam - 1;
        const int xc = in ? x : pdiam - 1;
        gs[e] = (unsigned)(xc + yc * npup);
        if constexpr (RAY)
          ph[e] = ray_phase(*ray, P.istart[s] + xc, P.jstart[s] + yc);
        else
          ph[e] = phase[gs[e]];
        of[e] = P.offsets[(unsigned)(in ? idx : np2 - 1)];
        x += dx;
        y += dy;
        if (x >= pdiam) {
          x -= pdiam;
          y += 1;
        }
      }
      if (unit) {
#pragma unroll
        for (int e = 0; e < CH; ++e) pu[e] = 1.0f;
      } else {
#pragma unroll
        for (int e = 0; e < CH; ++e) pu[e] = pupil[gs[e]];
      }
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        if (lane + (c0 + e) * LANES < np2) {
          float sn, cs;
          field_sincos<FAST>(ph[e], P.scale, turn_scale, of[e], sn, cs);
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
                                                Acc I, Acc Sx, Acc Sy, float rx, float ry) {
  I = lane_group_sum<LANES>(I);
  Sx = lane_group_sum<LANES>(Sx);
  Sy = lane_group_sum<LANES>(Sy);
  // every lane of the group holds the totals: lane 0 writes the x slope,
  // lane 1 the y slope (one division for both)
  if (valid && lane < 2) {
    const int iy = lane;
    const int ix = P.slope_order == 0 ? s + iy * P.nvalid : 2 * s + iy;
    const Acc num = iy ? Sy : Sx;
    const float r = iy ? ry : rx;
    // As the reference: the division is done in double (1.e-6 is a double
    // literal), the rest in float.  With float64 moments the sums enter exact.
    P.slopes[ix] = ((float)((double)num / ((double)I + 1.e-6)) - P.coffset) * P.cscale - r;
    if (iy == 0) P.cintens[s] = (float)I;
  }
}

// Per-sub-aperture scalars of the tail (flux, detector offsets, reference
// slopes), loaded when a wave starts so that their latency overlaps the spot
// computation instead of adding to the end of it.
struct SubapScalars {
  float fl, rx, ry;
  int vx, vy;
  __device__ __forceinline__ void load(const ImageParams& P, int s) {
    fl = P.flux[s];
    vx = P.validsubsx[s];
    vy = P.validsubsy[s];
    rx = 0.0f;
    ry = 0.0f;
    if (P.ref) {
      rx = P.ref[P.slope_order == 0 ? s : 2 * s];
      ry = P.ref[P.slope_order == 0 ? s + P.nvalid : 2 * s + 1];
    }
  }
};

// One wave (64 lanes) per sub-aperture: the three moment sums reach their
// totals through LDS (wave_sum3); lanes 0 and 1 divide and write the x and y
// slopes.  red: 192 entries of Acc of LDS private to the workgroup; pre: the
// prefetched reference slopes, or nullptr.
template <typename Acc>
__device__ __forceinline__ void centroid_finish64(const ImageParams& P, int s, bool valid,
                                                  int lane, Acc I, Acc Sx, Acc Sy, Acc* red,
                                                  const SubapScalars* pre = nullptr) {
  wave_sum3(red, lane, I, Sx, Sy);
  if (valid && lane < 2) {
    const int iy = lane;  // lane 0: x slope, lane 1: y slope
    const int ix = P.slope_order == 0 ? s + iy * P.nvalid : 2 * s + iy;
    const float r = pre ? (iy ? pre->ry : pre->rx) : (P.ref ? P.ref[ix] : 0.0f);
    const Acc num = iy ? Sy : Sx;
    P.slopes[ix] = ((float)((double)num / ((double)I + 1.e-6)) - P.coffset) * P.cscale - r;
    if (iy == 0) P.cintens[s] = (float)I;
  }
}

// Four sub-apertures per wave, 16 lanes each (group q = wave lane / 16),
// float64 moments: the 12 x 16 partials go through LDS (red: 192 doubles) and
// are summed by sum16_quads, the order of the one-sub-aperture tail
// (wave_sum3); lanes 0 and 1 of each group divide and write the x and y
// slopes.  r: the lane within its group.
__device__ __forceinline__ void centroid_finish16x4(const ImageParams& P, int s, bool valid, int r,
                                                    double I, double Sx, double Sy, double* red,
                                                    float rx, float ry) {
  const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const int q = lane >> 4;
  red[(3 * q) * 16 + r] = I;
  red[(3 * q + 1) * 16 + r] = Sx;
  red[(3 * q + 2) * 16 + r] = Sy;
  wave_lds_sync();
  const double t = sum16_quads(red, lane, 12);
  wave_lds_sync();
  if (lane < 48 && (lane & 3) == 0) red[lane >> 2] = t;  // group g = 3 q + moment
  wave_lds_sync();
  I = red[3 * q];
  Sx = red[3 * q + 1];
  Sy = red[3 * q + 2];
  if (valid && r < 2) {
    const int iy = r;  // lane 0: x slope, lane 1: y slope
    const int ix = P.slope_order == 0 ? s + iy * P.nvalid : 2 * s + iy;
    const double num = iy ? Sy : Sx;
    P.slopes[ix] = ((float)(num / (I + 1.e-6)) - P.coffset) * P.cscale - (iy ? ry : rx);
    if (iy == 0) P.cintens[s] = (float)I;
  }
}

// a[i] = v through a 32-bit byte offset from a base held in scalar registers
// (the arrays written here are under 4 GiB: i < 2^30).
__device__ __forceinline__ void store_at(float* a, unsigned i, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(a) + (i << 2)) = v;
}

// Workgroup b (of TPB threads) of the fused background part of a spot
// launch: one listed pixel per thread, the draw of the background kernel
// (image index as the element, domain 2, the frame's call).
template <int TPB>
__device__ __forceinline__ void draw_background(const ImageParams& P, const BgParams& bg, int b) {
  const int i = b * TPB + (int)threadIdx.x;
  if (i < bg.n) {
    const unsigned e = (unsigned)bg.pix[i];
    if (e < (unsigned)P.nsub * (unsigned)P.nsub)
      P.binimg[e] = bg.sigma * (float)rng_normal(e, P.call, 2u, P.seed);
  }
}

// Row of pixel p in a spot of npix columns: p / npix through the float
// reciprocal of npix (exact for p < 2^20: (p + 1/2) / npix stays at least
// 1/(2 npix) away from an integer, far beyond the few-ulp product error).
__device__ __forceinline__ int pixel_row(int p, float inv_npix) {
  return (int)(((float)p + 0.5f) * inv_npix);
}

// Moment accumulators of the fused centroider (float or float64 sums).
struct Moments {
  float cI = 0.0f, cX = 0.0f, cY = 0.0f;
  double dI = 0.0, dX = 0.0, dY = 0.0;
};

// One detector pixel (px, py) of sub-aperture s, flux-normalised value b,
// flat bincube index e, detector image index img: noise, bincube, moments
// and binimg.  KIND (centroider) and FP64 (float64 moments) are fixed per
// instantiation so the pixel loops carry no per-pixel tests of them.
template <bool NOISE, int KIND, bool FP64>
__device__ __forceinline__ float tail_pixel(const ImageParams& P, unsigned e, float b, float fx,
                                            float fy, unsigned img, unsigned call_add,
                                            Moments& m) {
  // no fused multiply-add anywhere in the pixel: the thresholded value is
  // taken from the stored (rounded) pixel, and the moment sums depend only on
  // the order of their terms
#pragma clang fp contract(off)
  if constexpr (NOISE) {
    const unsigned call = P.call + call_add;
    if (P.noise > -1.0f) b = rng_poisson_f(b, e, call, 0u, P.seed);
    if (P.noise > 0.0f) b = 1.0f * b + P.noise * (float)rng_normal(e, call, 1u, P.seed);
  }
  store_at(P.bincube, e, b);
  float out = b;
  if constexpr (KIND != SHWFS_CENTROID_NONE) {
    float d = b;
    if constexpr (KIND != SHWFS_CENTROID_COG) d = (b > P.threshold) ? b - P.threshold : 1.e-6f;
    if constexpr (KIND == SHWFS_CENTROID_TCOG) out = d;
    {
      if constexpr (KIND == SHWFS_CENTROID_WCOG) {
        const float w = P.weights[e];
        if constexpr (FP64) {
          m.dI += (double)d * w;
          m.dX += (double)d * fx * w;
          m.dY += (double)d * fy * w;
        } else {
          m.cI += d * w;
          m.cX += d * fx * w;
          m.cY += d * fy * w;
        }
      } else if constexpr (FP64) {
        // d * fx is exact in double (a float times an integer below 2^29),
        // so the fused form rounds once, as the product-then-sum does
        const double dd = d;
        m.dI += dd;
        m.dX = fma(dd, (double)fx, m.dX);
        m.dY = fma(dd, (double)fy, m.dY);
      } else {
        m.cI += d;
        m.cX += d * fx;
        m.cY += d * fy;
      }
    }
  }
  if (P.binimg) store_at(P.binimg, img, out);
  return b;  // the spot-cube value
}

// The pixels p = lane, lane + LANES, ... of sub-aperture s (tail_pixel), in
// that order.  NPIX > 0: npix == NPIX is known and, LANES being a multiple
// of it, the lane's pixel column is fixed and its row advances by
// LANES / NPIX from one pixel to the next (no per-pixel index work).
// KEEP: pixel p's spot-cube value replaces bin[p] (the spot stays in LDS for
// a centroider that runs after the tail).
template <int LANES, bool NOISE, int KIND, bool FP64, int NPIX = 0, bool KEEP = false>
__device__ __forceinline__ void tail_pixels(const ImageParams& P, int s, int lane, float* bin,
                                            float fnorm, int vx, int vy, unsigned call_add,
                                            Moments& m) {
  if constexpr (NPIX > 0 && LANES % NPIX == 0) {
    constexpr int NP2 = NPIX * NPIX, DY = LANES / NPIX;
    const int px = lane % NPIX, py0 = lane / NPIX;
    const unsigned e0 = (unsigned)(lane + s * NP2);
    unsigned img0 = (unsigned)(vx + px) + __umul24(vy + py0, P.nsub);
    // held as computed: the row step stays one add per pixel instead of being
    // refolded into a 32-bit multiply (a quarter-rate instruction) per pixel
    asm volatile("" : "+v"(img0));
    const unsigned dimg = __umul24(DY, P.nsub);
#pragma unroll
    for (int j = 0; j * LANES < NP2; ++j)
      if (lane + j * LANES < NP2) {
        const float v = tail_pixel<NOISE, KIND, FP64>(P, e0 + j * LANES,
                                                      bin[lane + j * LANES] * fnorm, (float)px,
                                                      (float)(py0 + j * DY), img0 + j * dimg,
                                                      call_add, m);
        if constexpr (KEEP) bin[lane + j * LANES] = v;
      }
  } else {
    const int npix = P.npix;
    const int npix2 = npix * npix;
    const float inv_npix = __builtin_amdgcn_rcpf((float)npix);
    for (int p = lane; p < npix2; p += LANES) {
      const int py = pixel_row(p, inv_npix), px = p - (int)__umul24(py, npix);
      const float v = tail_pixel<NOISE, KIND, FP64>(P, (unsigned)(p + s * npix2), bin[p] * fnorm,
                                                    (float)px, (float)py,
                                                    (unsigned)(vx + px) + __umul24(vy + py, P.nsub),
                                                    call_add, m);
      if constexpr (KEEP) bin[p] = v;
    }
  }
}

// Sum of a spot's npix2 pixels in LDS, for one group of G lanes (G = 16 or
// 64; lane = the lane's index within the group), held by every lane of the
// group.  Lanes r < 16 of the group add pixels r, r + 16, ... in turn, then
// one 16-lane row sum: the same float sum for either group size.
template <int G>
__device__ __forceinline__ float spot_sum(const float* sp, int npix2, bool valid, int lane) {
  static_assert(G == 16 || G == 64, "spot_sum: groups of 16 or 64 lanes");
  float part = 0.0f;
  if (valid && lane < 16) {
    if (npix2 == 64) {
#pragma unroll
      for (int j = 0; j < 4; ++j) part += sp[lane + 16 * j];  // 8x8 spots, same order
    } else {
      for (int p = lane; p < npix2; p += 16) part += sp[p];
    }
  }
  part = row_sum(part);
  if constexpr (G == 64) part = lane_value(part, 0);
  return part;
}

// Binned spot -> flux normalisation, noise, bincube / binimg, fused
// centroid.  bin: the npix^2 binned pixels of sub-aperture s in LDS (x
// fastest), in units of 1/inten_scale; I: their sum (the same units).
// red: for LANES == 64, LDS scratch of 192 doubles for the moment totals
// (wave_sum3; must not overlap bin), or nullptr for the DPP reductions.
// pre: the sub-aperture's scalars loaded beforehand, or nullptr.
// KEEP: no centroider here; the spot-cube values replace bin (for the
// learned centroider, which runs after the tail over several spots).
template <int LANES, bool NOISE, bool KEEP = false>
__device__ __forceinline__ void finish_spot(const ImageParams& P, int s, bool valid, int lane,
                                            float* bin, float I, float inten_scale,
                                            unsigned call_add = 0u, void* red = nullptr,
                                            const SubapScalars* pre = nullptr) {
  Moments m;
  if constexpr (KEEP) {
    if (valid) {
      const float fl = pre ? pre->fl : P.flux[s];
      if (lane == 0) P.intensities[s] = I * inten_scale;
      const float fnorm = (I != 0.0f) ? fl * __builtin_amdgcn_rcpf(I) * P.nphot : 1.0f;
      const int vx = pre ? pre->vx : P.validsubsx[s];
      const int vy = pre ? pre->vy : P.validsubsy[s];
      if (!NOISE && P.npix == 8)
        tail_pixels<LANES, NOISE, SHWFS_CENTROID_NONE, false, NOISE ? 0 : 8, true>(
            P, s, lane, bin, fnorm, vx, vy, call_add, m);
      else
        tail_pixels<LANES, NOISE, SHWFS_CENTROID_NONE, false, 0, true>(P, s, lane, bin, fnorm, vx,
                                                                       vy, call_add, m);
    }
    return;
  }
  if (valid) {
    const float fl = pre ? pre->fl : P.flux[s];
    const int vx = pre ? pre->vx : P.validsubsx[s];
    const int vy = pre ? pre->vy : P.validsubsy[s];
    if (lane == 0) P.intensities[s] = I * inten_scale;
    // photometric normalisation b * flux / I * nphot as one factor
    const float fnorm = (I != 0.0f) ? fl * __builtin_amdgcn_rcpf(I) * P.nphot : 1.0f;
    // noise-free 8x8-pixel spots (the usual SH detector) take the fixed-size
    // form; with noise the samplers dominate the pixel work
#define SHWFS_TAIL(K, F)                                                                      \
  ((!NOISE && P.npix == 8)                                                                    \
       ? tail_pixels<LANES, NOISE, K, F, NOISE ? 0 : 8>(P, s, lane, bin, fnorm, vx, vy, call_add, m) \
       : tail_pixels<LANES, NOISE, K, F>(P, s, lane, bin, fnorm, vx, vy, call_add, m))
    switch (P.ckind) {
      case SHWFS_CENTROID_NONE: SHWFS_TAIL(SHWFS_CENTROID_NONE, false); break;
      case SHWFS_CENTROID_COG:
        if (P.fp64) SHWFS_TAIL(SHWFS_CENTROID_COG, true); else SHWFS_TAIL(SHWFS_CENTROID_COG, false);
        break;
      case SHWFS_CENTROID_WCOG:
        if (P.fp64) SHWFS_TAIL(SHWFS_CENTROID_WCOG, true); else SHWFS_TAIL(SHWFS_CENTROID_WCOG, false);
        break;
      default:
        if (P.fp64) SHWFS_TAIL(SHWFS_CENTROID_TCOG, true); else SHWFS_TAIL(SHWFS_CENTROID_TCOG, false);
        break;
    }
#undef SHWFS_TAIL
  }
  double dI = m.dI, dX = m.dX, dY = m.dY;
  float cI = m.cI, cX = m.cX, cY = m.cY;
  if (P.ckind != SHWFS_CENTROID_NONE) {
    if constexpr (LANES == 64) {
      if (red != nullptr) {
        if (P.fp64)
          centroid_finish64<double>(P, s, valid, lane, dI, dX, dY, static_cast<double*>(red),
                                    pre);
        else
          centroid_finish64<float>(P, s, valid, lane, cI, cX, cY, static_cast<float*>(red), pre);
        return;
      }
    }
    float rx = 0.0f, ry = 0.0f;
    if (pre) {
      rx = pre->rx;
      ry = pre->ry;
    } else if (valid && P.ref && lane < 2) {
      rx = P.ref[P.slope_order == 0 ? s : 2 * s];
      ry = P.ref[P.slope_order == 0 ? s + P.nvalid : 2 * s + 1];
    }
    if constexpr (LANES == 16) {
      if (red != nullptr && P.fp64) {
        centroid_finish16x4(P, s, valid, lane, dI, dX, dY, static_cast<double*>(red), rx, ry);
        return;
      }
    }
    if (P.fp64)
      centroid_finish<LANES, double>(P, s, valid, lane, dI, dX, dY, rx, ry);
    else
      centroid_finish<LANES, float>(P, s, valid, lane, cI, cX, cY, rx, ry);
  }
}

// Binned spot from the NF x NF |E|^2 image hr in LDS (flat index u + v*NF):
// block form or the general binmap gather; returns this lane's share of the
// spot's sum.  bin: npix^2 floats of LDS.
template <int LANES, int NF>
__device__ __forceinline__ float bin_gather(const ImageParams& P, bool valid, int lane,
                                            const float* hr, float* bin, const int* s_binmap) {
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
  return part;
}

// hr: the NF x NF |E|^2 image of this sub-aperture in LDS, flat index u + v*NF,
// holding inten_scale^-1 times the true intensities (a power of two).
// bin: npix^2 floats of LDS private to this sub-aperture.
template <int LANES, int NF, bool NOISE, bool KEEP = false>
__device__ __forceinline__ void spot_tail(const ImageParams& P, int s, bool valid, int lane,
                                          const float* hr, float* bin, const int* s_binmap,
                                          float inten_scale, unsigned call_add = 0u,
                                          void* red = nullptr,
                                          const SubapScalars* pre = nullptr) {
  const float part = bin_gather<LANES, NF>(P, valid, lane, hr, bin, s_binmap);
  const float I = lane_group_sum<LANES>(part);
  finish_spot<LANES, NOISE, KEEP>(P, s, valid, lane, bin, I, inten_scale, call_add, red, pre);
}

}  // namespace shwfs
