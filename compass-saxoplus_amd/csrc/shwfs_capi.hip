// C ABI entry points of libshwfs.so, the standalone centroid kernels, the
// counter-based noise kernel and the weight broadcast.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"
#include "shwfs_raytrace.hpp"

namespace shwfs {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- centroiders
// One group of G lanes per sub-aperture, 256/G sub-apertures per workgroup.
// Pixel (x, y) of sub-aperture s is img[(x + validx[s]) + (y + validy[s]) * size],
// guarded by idim < size*size as in the reference kernels.
template <int G, int KIND, typename Acc>
__global__ __launch_bounds__(256) void centroid_kernel(const float* img_in, float* img_out,
                                                       shwfs_centroid_args a) {
  constexpr int S = 256 / G;
  const int slot = threadIdx.x / G, lane = threadIdx.x % G;
  const int s = blockIdx.x * S + slot;
  const bool valid = s < a.nvalid;
  const int npix2 = a.npix * a.npix;
  const long lim = (long)a.size * a.size;
  Acc I = 0, Sx = 0, Sy = 0;
  if (valid) {
    const int xv = a.d_validx[s], yv = a.d_validy[s];
    for (int p = lane; p < npix2; p += G) {
      const int x = p % a.npix, y = p / a.npix;
      const long idim = (long)(x + xv) + (long)(y + yv) * a.size;
      if (idim < lim) {
        const float v = img_in[idim];
        float d = v;
        if constexpr (KIND != SHWFS_CENTROID_COG) d = (v > a.threshold) ? v - a.threshold : 1.e-6f;
        if constexpr (KIND == SHWFS_CENTROID_TCOG) {
          if (img_out) img_out[idim] = d;
        }
        const float fx = (float)x, fy = (float)y;
        if constexpr (KIND == SHWFS_CENTROID_WCOG) {
          const float w = a.d_weights[p + s * npix2];
          I += (Acc)d * w;
          Sx += (Acc)d * fx * w;
          Sy += (Acc)d * fy * w;
        } else {
          I += (Acc)d;
          Sx += (Acc)d * fx;
          Sy += (Acc)d * fy;
        }
      }
    }
  }
  I = group_sum<G>(I);
  Sx = group_sum<G>(Sx);
  Sy = group_sum<G>(Sy);
  if (valid && lane == 0) {
    const int ix = a.slope_order == 0 ? s : 2 * s;
    const int iy = a.slope_order == 0 ? s + a.nvalid : 2 * s + 1;
    const float rx = a.d_ref ? a.d_ref[ix] : 0.0f;
    const float ry = a.d_ref ? a.d_ref[iy] : 0.0f;
    a.d_slopes[ix] = ((float)((double)Sx / ((double)I + 1.e-6)) - a.offset) * a.scale - rx;
    a.d_slopes[iy] = ((float)((double)Sy / ((double)I + 1.e-6)) - a.offset) * a.scale - ry;
    a.d_intensities[s] = (float)I;
  }
}

template <int G, int KIND>
static int32_t launch_centroid_g(const float* in, float* out, const shwfs_centroid_args& a,
                                 hipStream_t st) {
  constexpr int S = 256 / G;
  const int grid = (a.nvalid + S - 1) / S;
  if (a.fp64_moments)
    hipLaunchKernelGGL((centroid_kernel<G, KIND, double>), dim3(grid), dim3(256), 0, st, in, out, a);
  else
    hipLaunchKernelGGL((centroid_kernel<G, KIND, float>), dim3(grid), dim3(256), 0, st, in, out, a);
  SHWFS_LAUNCHED("centroid_kernel");
  return SHWFS_OK;
}

template <int KIND>
static int32_t launch_centroid_k(const float* in, float* out, const shwfs_centroid_args& a,
                                 hipStream_t st) {
  const int n = a.npix * a.npix;
  if (n <= 4) return launch_centroid_g<4, KIND>(in, out, a, st);
  if (n <= 8) return launch_centroid_g<8, KIND>(in, out, a, st);
  if (n <= 16) return launch_centroid_g<16, KIND>(in, out, a, st);
  if (n <= 32) return launch_centroid_g<32, KIND>(in, out, a, st);
  return launch_centroid_g<64, KIND>(in, out, a, st);
}

// Brightest-pixel CoG, one wave per sub-aperture, four per workgroup.  The
// window (npix^2 pixels, 0 outside the image) goes to the wave's LDS slot.
// Pixel p has rank #{q: v_q > v_p} in decreasing order; the value with
// #{q: v_q > v} <= nbpix < #{q: v_q >= v} is element nbpix of the window
// sorted in decreasing order (ties included), the threshold.  Pixels above
// it keep v - threshold, the others 0 (written back into the image); then the
// float CoG moments and the double division of the other centroiders.
__global__ __launch_bounds__(256) void bpcog_kernel(float* img, int nbpix, shwfs_centroid_args a) {
  extern __shared__ float sm[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + wave;
  const bool valid = s < a.nvalid;  // uniform over the wave
  const int npix = a.npix, n2 = npix * npix;
  float* v = sm + wave * (n2 + 1);
  const long lim = (long)a.size * a.size;
  int xv = 0, yv = 0;
  if (valid) {
    xv = a.d_validx[s];
    yv = a.d_validy[s];
  }
  for (int p = lane; p < n2; p += 64) {
    const long idim = (long)(p % npix + xv) + (long)(p / npix + yv) * a.size;
    v[p] = (valid && idim >= 0 && idim < lim) ? img[idim] : 0.0f;
  }
  if (lane == 0) v[n2] = 0.0f;  // threshold when no value qualifies (NaN in the window)
  wave_lds_sync();
  for (int p = lane; p < n2; p += 64) {
    const float vp = v[p];
    int gt = 0, ge = 0;
    for (int q = 0; q < n2; ++q) {
      const float w = v[q];  // same address in every lane: an LDS broadcast
      gt += w > vp ? 1 : 0;
      ge += w >= vp ? 1 : 0;
    }
    if (gt <= nbpix && nbpix < ge) v[n2] = vp;  // every such lane writes the same value
  }
  wave_lds_sync();
  const float thr = v[n2];
  float I = 0.0f, Sx = 0.0f, Sy = 0.0f;
  if (valid) {
    for (int p = lane; p < n2; p += 64) {
      const int x = p % npix, y = p / npix;
      const long idim = (long)(x + xv) + (long)(y + yv) * a.size;
      if (idim >= 0 && idim < lim) {
        const float d = v[p] > thr ? v[p] - thr : 0.0f;
        img[idim] = d;
        I += d;
        Sx += d * (float)x;
        Sy += d * (float)y;
      }
    }
  }
  I = wave_sum(I);
  Sx = wave_sum(Sx);
  Sy = wave_sum(Sy);
  if (valid && lane == 0) {
    const int ix = a.slope_order == 0 ? s : 2 * s;
    const int iy = a.slope_order == 0 ? s + a.nvalid : 2 * s + 1;
    const float rx = a.d_ref ? a.d_ref[ix] : 0.0f;
    const float ry = a.d_ref ? a.d_ref[iy] : 0.0f;
    a.d_slopes[ix] = ((float)((double)Sx / ((double)I + 1.e-6)) - a.offset) * a.scale - rx;
    a.d_slopes[iy] = ((float)((double)Sy / ((double)I + 1.e-6)) - a.offset) * a.scale - ry;
    a.d_intensities[s] = I;
  }
}

// Spots of an external detector image (acquisim), one thread per element.
__global__ __launch_bounds__(256) void fill_bincube_kernel(const float* image, float* bcube,
                                                           const int* vx, const int* vy, int npix,
                                                           long n, int size) {
  const int npix2 = npix * npix;
  const long lim = (long)size * size;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256L) {
    const int s = (int)(t / npix2), p = (int)(t - (long)s * npix2);
    const long idim = (long)(p % npix + vx[s]) + (long)(p / npix + vy[s]) * size;
    bcube[t] = (idim >= 0 && idim < lim) ? image[idim] : 0.0f;
  }
}

// The same from a sub-aperture number map: element (x, y) of grid cell
// (i, j) of the image goes to spot num_ssp[i + nxsub j] - 1.
__global__ __launch_bounds__(256) void fill_bincube_2d_kernel(const float* image, float* bcube,
                                                              const int* num, int npix, int nxsub,
                                                              int nvalid) {
  const int npix2 = npix * npix;
  const long n = (long)nxsub * nxsub * npix2;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256L) {
    const int cell = (int)(t / npix2), p = (int)(t - (long)cell * npix2);
    const int i = cell % nxsub, j = cell / nxsub, x = p % npix, y = p / npix;
    const int k = num[cell];
    if (k > 0 && k <= nvalid)
      bcube[(long)(k - 1) * npix2 + p] =
          image[(long)(x + npix * i) + (long)nxsub * npix * (y + npix * j)];
  }
}

// ---------------------------------------------------------------- noise on an array
__global__ __launch_bounds__(256) void prng_kernel(float* d, long n, char gtype, float alpha,
                                                   float beta, unsigned long long seed,
                                                   unsigned call, unsigned domain) {
  // no fused multiply-add: beta * d + alpha * N(0, 1) rounds like the spot
  // kernels' readout-noise sum, so the two agree bit for bit
#pragma clang fp contract(off)
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const unsigned e = (unsigned)i;
    if (gtype == 'P') {
      d[i] = rng_poisson_f(d[i], e, call, domain, seed);
    } else if (gtype == 'N') {
      d[i] = beta * d[i] + alpha * (float)rng_normal(e, call, domain, seed);
    } else {
      u32x4 w = rng_block(e, 0, call, domain, seed);
      d[i] = beta * d[i] + (float)u01(w.x);
    }
  }
}

// one wave per sub-aperture: is every pupil value of its window exactly 1?
__global__ __launch_bounds__(64) void unit_pupil_kernel(const float* pupil, int npup,
                                                        const int* istart, const int* jstart,
                                                        int pdiam, int nvalid, int* flags) {
  const int s = blockIdx.x;
  if (s >= nvalid) return;
  bool all = true;
  for (int i = threadIdx.x; i < pdiam * pdiam; i += 64) {
    const int x = i % pdiam, y = i / pdiam;
    all = all && pupil[(size_t)(istart[s] + x) + (size_t)(jstart[s] + y) * npup] == 1.0f;
  }
  const bool wave_all = __all(all);
  if (threadIdx.x == 0) flags[s] = wave_all ? 1 : 0;
}

__global__ __launch_bounds__(256) void fill_weights_kernel(float* out, const float* in, int npix2,
                                                           long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L)
    out[i] = in[i % npix2];
}

template <typename Tin>
__global__ __launch_bounds__(256) void calibrate_kernel(const Tin* raw, float* cal,
                                                        const float* dark, const float* flat,
                                                        const int* lut, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const int j = lut[i];
    cal[i] = (j >= 0 && j < n) ? ((float)raw[j] - dark[i]) * flat[i] : 0.0f;
  }
}

// Camera image: u16(min(fma(img, flat, dark) / mfpp, 1) * mpv), converted
// toward zero and saturated to [0, 65535] (NaN to 0).
__global__ __launch_bounds__(256) void digitalize_kernel(uint16_t* cam, const float* img,
                                                         const float* dark, const float* flat,
                                                         float mfpp, float mpv, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    float t = fmaf(img[i], flat[i], dark[i]) / mfpp;
    t = t < 1.0f ? t : 1.0f;
    const float v = truncf(t * mpv);
    // !(v > 0) also catches NaN
    cam[i] = !(v > 0.0f) ? (uint16_t)0 : (v >= 65535.0f ? (uint16_t)65535 : (uint16_t)v);
  }
}

// Geometric slopes, one group of 64 lanes per sub-aperture; phi(x, y) =
// phase[pmap[x + y n + s n^2]].
// type 0: slope_x = alpha/n * sum_y (phi(n-1, y) - phi(0, y)), same along y.
// type 1: slope_x = alpha/(2 n flux[s]) * sum_{x,y} D_x phi(x, y) * pupil(x, y),
//         D_x the central difference, one-sided (not halved) at x = 0, n-1.
// type 2: slope_x = sum_k ttprojmat[k, s] * phi(k), slope_y with column s + nvalid.
__global__ __launch_bounds__(256) void slopes_geom_kernel(int type, const float* phase,
                                                          const float* pupil, const float* flux,
                                                          float* slopes, const int* pmap,
                                                          const float* proj, int n, int nvalid,
                                                          float alpha) {
  const int slot = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int s = blockIdx.x * 4 + slot;
  const bool valid = s < nvalid;
  const int n2 = n * n;
  float ax = 0.f, ay = 0.f;
  if (valid) {
    const int* m = pmap + (size_t)s * n2;
    if (type == 0) {
      for (int t = lane; t < n; t += 64) {
        ax += phase[m[t * n + n - 1]] - phase[m[t * n]];
        ay += phase[m[t + (n - 1) * n]] - phase[m[t]];
      }
    } else if (type == 1) {
      for (int k = lane; k < n2; k += 64) {
        const int x = k % n, y = k / n;
        const float w = pupil[m[k]];
        const int xp = x + 1 < n ? k + 1 : k, xm = x > 0 ? k - 1 : k;
        const int yp = y + 1 < n ? k + n : k, ym = y > 0 ? k - n : k;
        ax += (phase[m[xp]] - phase[m[xm]]) * w;
        ay += (phase[m[yp]] - phase[m[ym]]) * w;
      }
    } else {
      const float* px = proj + (size_t)s * n2;
      const float* py = proj + (size_t)(s + nvalid) * n2;
      for (int k = lane; k < n2; k += 64) {
        const float v = phase[m[k]];
        ax += px[k] * v;
        ay += py[k] * v;
      }
    }
  }
  ax = group_sum<64>(ax);
  ay = group_sum<64>(ay);
  if (valid && lane == 0) {
    if (type == 0) {
      slopes[s] = ax / n * alpha;
      slopes[s + nvalid] = ay / n * alpha;
    } else if (type == 1) {
      slopes[s] = ax / n * alpha / flux[s] / 2;
      slopes[s + nvalid] = ay / n / 2 * alpha / flux[s];
    } else {
      slopes[s] = ax;
      slopes[s + nvalid] = ay;
    }
  }
}

// Tip/tilt removal of a slope vector by one workgroup: both projections are
// taken on the unfiltered slopes, then both are subtracted.
constexpr int kTtThreads = 1024;
__global__ __launch_bounds__(kTtThreads) void tt_filter_kernel(float* c, const float* tip,
                                                               const float* tilt, float* tt,
                                                               int n) {
  __shared__ float red[2][kTtThreads / 64];
  float a = 0.f, b = 0.f;
  for (int i = threadIdx.x; i < n; i += kTtThreads) {
    const float v = c[i];
    a += v * tip[i];
    b += v * tilt[i];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  a = 0.f;
  b = 0.f;
  for (int k = 0; k < kTtThreads / 64; ++k) {
    a += red[0][k];
    b += red[1][k];
  }
  for (int i = threadIdx.x; i < n; i += kTtThreads) c[i] = (c[i] - a * tip[i]) - b * tilt[i];
  if (threadIdx.x == 0 && tt) {
    tt[0] = a;
    tt[1] = b;
  }
}

// Spots into the detector image: img[(vx[s] + px) + (vy[s] + py) * nsub] =
// bincube[px + py npix + s npix^2], one thread per spot pixel.
__global__ __launch_bounds__(256) void fill_binimg_kernel(const float* bcube, float* img,
                                                          const int* vx, const int* vy, int npix,
                                                          long n, int nsub) {
  const int npix2 = npix * npix;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256L) {
    const int s = (int)(t / npix2), p = (int)(t - (long)s * npix2);
    const int px = p % npix, py = p / npix;
    img[(long)(vx[s] + px) + (long)(vy[s] + py) * nsub] = bcube[t];
  }
}

bool ext_needed(const shwfs_sh_geom* g) {
  return g->ntot != g->nfft || g->kernconv || g->d_ftlgskern != nullptr;
}

// comp_generic's order (libsutra/src.cpp/sutra_wfs_sh.cpp:409-525): with
// ntot != nfft the LGS kernels and then the kernel; with ntot == nfft the LGS
// kernels alone, or the kernel alone.  Two unnormalised transform pairs
// multiply the result by ntot^2 once more than one pair over the product of
// both kernels: kscale keeps the reference's raw flux.
ExtParams ext_params(const shwfs_sh_geom* g) {
  const bool lgs = g->d_ftlgskern != nullptr;
  const bool kern = g->kernconv && (g->ntot != g->nfft || !lgs);
  ExtParams x{};
  x.hrmap = g->ntot != g->nfft ? g->d_hrmap : nullptr;
  x.ftkernel = kern ? reinterpret_cast<const float2*>(g->d_ftkernel) : nullptr;
  x.ftlgs = lgs ? reinterpret_cast<const float2*>(g->d_ftlgskern) : nullptr;
  x.conv = (kern || lgs) ? 1 : 0;
  x.kscale = (kern && lgs) ? (float)g->ntot * (float)g->ntot : 1.0f;
  return x;
}

int32_t check_geom(const shwfs_sh_geom* g, int method) {
  if (!g) return fail(SHWFS_ERR_ARG, "geometry is NULL");
  if (g->nvalid <= 0 || g->nphase <= 0 || g->npup <= 0 || g->npix <= 0 || g->nrebin <= 0 ||
      g->nxsub <= 0)
    return fail(SHWFS_ERR_ARG, "geometry sizes must be positive");
  if (!(g->nfft == 8 || g->nfft == 16 || g->nfft == 32 || g->nfft == 64))
    return fail(SHWFS_ERR_UNSUPPORTED, "nfft must be 8, 16, 32 or 64");
  const bool ext = ext_needed(g);
  if (g->d_ftlgskern && !(g->ntot == 16 || g->ntot == 32 || g->ntot == 64))
    return fail(SHWFS_ERR_UNSUPPORTED, "LGS kernels: needs ntot in {16, 32, 64}");
  if (g->ntot != g->nfft) {
    if (!(g->ntot == 16 || g->ntot == 32 || g->ntot == 64) || g->ntot < g->nfft || g->nfft < 16)
      return fail(SHWFS_ERR_UNSUPPORTED, "ntot != nfft: needs nfft in {16, 32, 64} and ntot a "
                                         "power of two in (nfft, 64]");
    if (!g->d_hrmap) return fail(SHWFS_ERR_STATE, "ntot != nfft: hrmap not loaded");
  }
  if (g->kernconv && !g->d_ftkernel) return fail(SHWFS_ERR_STATE, "kernconv: ftkernel not loaded");
  if (ext && method != SHWFS_METHOD_FFT_FP32)
    return fail(SHWFS_ERR_UNSUPPORTED, "ntot != nfft / kernconv / LGS: FFT method only");
  if (ext && g->nfft < 16)
    return fail(SHWFS_ERR_UNSUPPORTED, "ntot != nfft / kernconv / LGS: needs nfft in {16, 32, 64}");
  if (g->nphase > g->nfft) return fail(SHWFS_ERR_ARG, "nphase must be <= nfft");
  if (g->npix * g->nrebin > g->ntot) return fail(SHWFS_ERR_ARG, "npix*nrebin must be <= ntot");
  // the kernels index the spot cube, the detector image and the screen with
  // 32-bit offsets (and 24-bit products)
  const long long nsub = (long long)g->npix * g->nxsub;
  if ((long long)g->nvalid * g->npix * g->npix >= (1LL << 30) || nsub * nsub >= (1LL << 30) ||
      (long long)g->npup * g->npup >= (1LL << 30))
    return fail(SHWFS_ERR_UNSUPPORTED, "geometry too large (spot cube, image or screen >= 2^30 elements)");
  if (g->d_bgpix && (g->nbgpix < 0 || g->nbgpix > nsub * nsub))
    return fail(SHWFS_ERR_ARG, "background pixel list: need 0 <= nbgpix <= (npix * nxsub)^2");
  if (!g->d_istart || !g->d_jstart || !g->d_validsubsx || !g->d_validsubsy || !g->d_offsets ||
      !g->d_fluxPerSub || !g->d_binmap)
    return fail(SHWFS_ERR_STATE, "geometry tables not loaded (load_arrays not called?)");
  if (method == SHWFS_METHOD_DFT_BF16X3 || method == SHWFS_METHOD_DFT_FP16) {
    if (g->nfft < 16) return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path needs nfft >= 16");
  } else if (method != SHWFS_METHOD_FFT_FP32) {
    return fail(SHWFS_ERR_ARG, "unknown method");
  }
  return SHWFS_OK;
}

int32_t image_params(const shwfs_sh_geom* g, const shwfs_frame_args* f, ImageParams* out) {
  int32_t rc = check_geom(g, f ? f->method : 0);
  if (rc) return rc;
  if (!f || (!f->d_phase && f->nlayers <= 0) || !f->d_pupil || !f->d_bincube ||
      !f->d_intensities)
    return fail(SHWFS_ERR_ARG, "frame: phase (or layers), pupil, bincube and intensities are required");
  if (ext_needed(g) &&
      (f->nlayers > 0 || (f->centroid && f->centroid->kind == SHWFS_CENTROID_MLP)))
    return fail(SHWFS_ERR_UNSUPPORTED,
                "ntot != nfft / kernconv / LGS: no layers and no fused learned centroider");
  if (f->nlayers > 0) {
    if (f->method != SHWFS_METHOD_FFT_FP32)
      return fail(SHWFS_ERR_UNSUPPORTED, "frame: the raytrace in the field stage runs in the FFT method only");
    if (f->wg_size != 0 && f->wg_size != 64)
      return fail(SHWFS_ERR_UNSUPPORTED, "frame: the raytrace in the field stage uses 64-lane workgroups");
    if (f->centroid && f->centroid->kind == SHWFS_CENTROID_MLP)
      return fail(SHWFS_ERR_UNSUPPORTED, "frame: layers and the fused learned centroider are exclusive");
  }
  // every sub-aperture window must lie inside the screen: checked by the caller
  // when the tables are loaded (the tables live on the device).
  ImageParams p{};
  p.phase = f->d_phase;
  p.pupil = f->d_pupil;
  p.offsets = g->d_offsets;
  p.istart = g->d_istart;
  p.jstart = g->d_jstart;
  p.validsubsx = g->d_validsubsx;
  p.validsubsy = g->d_validsubsy;
  p.flux = g->d_fluxPerSub;
  p.binmap = g->d_binmap;
  p.pupil_unit = f->d_pupil_unit;
  p.npup = g->npup;
  p.nvalid = g->nvalid;
  p.pdiam = g->nphase;
  p.npix = g->npix;
  p.nrebin = g->nrebin;
  p.nsub = g->npix * g->nxsub;
  p.bin_block = g->binmap_block ? 1 : 0;
  p.bin_origin = g->bin_origin;
  if (p.bin_block && (g->bin_origin < 0 || g->bin_origin >= g->nfft))
    return fail(SHWFS_ERR_ARG, "bin_origin outside [0, nfft)");
  if (f->wg_size != 0 && f->wg_size != 64 && f->wg_size != 128 && f->wg_size != 256)
    return fail(SHWFS_ERR_ARG, "wg_size must be 0, 64, 128 or 256");
  p.tpb = f->wg_size;
  p.scale = f->scale;
  p.nphot = f->nphot;
  p.noise = f->noise;
  p.seed = f->seed;
  p.call = f->call;
  p.bincube = f->d_bincube;
  p.binimg = f->d_binimg;
  p.intensities = f->d_intensities;
  p.ckind = SHWFS_CENTROID_NONE;
  if (f->centroid && f->centroid->kind != SHWFS_CENTROID_NONE) {
    const shwfs_centroid_args* c = f->centroid;
    if (c->kind < SHWFS_CENTROID_COG || c->kind > SHWFS_CENTROID_MLP)
      return fail(SHWFS_ERR_ARG, "unknown centroider kind");
    if (c->kind == SHWFS_CENTROID_MLP) {
      if (f->method != SHWFS_METHOD_FFT_FP32 &&
          !(g->nfft == 32 && g->nphase == 16 && g->npix == 8 && g->nrebin == 4 &&
            g->binmap_block && g->bin_origin == 16))
        return fail(SHWFS_ERR_UNSUPPORTED,
                    "fused learned centroider: an MFMA method needs nfft 32, 16-point windows "
                    "and 8x8 pixels binned 4x4 about the grid centre (else the FFT method)");
      if (g->npix != 8 || (g->nfft != 16 && g->nfft != 32))
        return fail(SHWFS_ERR_UNSUPPORTED,
                    "fused learned centroider: needs 8x8-pixel spots and nfft 16 or 32");
      if (!c->d_mlp_params) return fail(SHWFS_ERR_ARG, "fused learned centroider: NULL parameters");
    }
    if (!c->d_slopes || !c->d_intensities) return fail(SHWFS_ERR_ARG, "centroid outputs are NULL");
    if (c->kind == SHWFS_CENTROID_WCOG && !c->d_weights)
      return fail(SHWFS_ERR_ARG, "WCOG needs weights");
    if (c->nvalid != g->nvalid || c->npix != g->npix)
      return fail(SHWFS_ERR_ARG, "fused centroider sizes differ from the WFS");
    p.ckind = c->kind;
    p.fp64 = c->fp64_moments;
    p.slope_order = c->slope_order;
    p.threshold = c->threshold;
    p.coffset = c->offset;
    p.cscale = c->scale;
    p.weights = c->d_weights;
    p.ref = c->d_ref;
    p.slopes = c->d_slopes;
    p.cintens = c->d_intensities;
  }
  *out = p;
  return SHWFS_OK;
}

// ---- device scratch of shwfs_sh_comp_image_noalloc --------------------------
// The MFMA DFT operand table of the frame's shape (the largest, nfft 64 with
// a 32-point padded window in split-BF16x3, is 32 KiB), rebuilt into this
// array by every call of that entry point.  Each device holds its own copy
// for the life of the process.  Host code uses the address that
// hipGetSymbolAddress returns per device (scratch_address below), never the
// array's name.
constexpr size_t kScratchBytes = SHWFS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, looked up once per
// device with that device current; NULL with the error set on failure.
static void* scratch_address() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("comp_image_noalloc: no current device (or device index above %d)", kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("comp_image_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                hipGetErrorString(e));
      return nullptr;
    }
    addr[dev] = ptr;
  }
  return addr[dev];
}

}  // namespace shwfs

using namespace shwfs;

extern "C" {

int32_t shwfs_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return SHWFS_ERR_ARG;
  strncpy(buf, g_err, len - 1);
  buf[len - 1] = 0;
  return SHWFS_OK;
}

const char* shwfs_version(void) { return "shwfs-mi355x 0.1 (gfx950)"; }

int32_t shwfs_mfma_probe(const float* d_A, const float* d_B, float* d_D, int32_t method,
                         void* stream) {
  if (!d_A || !d_B || !d_D) return fail(SHWFS_ERR_ARG, "mfma_probe: NULL argument");
  return launch_mfma_probe(d_A, d_B, d_D, method, (hipStream_t)stream);
}

int32_t shwfs_sh_check_geom(const shwfs_sh_geom* g, int32_t method) { return check_geom(g, method); }

int32_t shwfs_sh_unit_pupil(const shwfs_sh_geom* g, const float* d_pupil, int32_t* d_flags,
                            void* stream) {
  int32_t rc = check_geom(g, SHWFS_METHOD_FFT_FP32);
  if (rc) return rc;
  if (!d_pupil || !d_flags) return fail(SHWFS_ERR_ARG, "unit_pupil: NULL argument");
  hipLaunchKernelGGL(unit_pupil_kernel, dim3(g->nvalid), dim3(64), 0, (hipStream_t)stream, d_pupil,
                     g->npup, g->d_istart, g->d_jstart, g->nphase, g->nvalid, d_flags);
  SHWFS_LAUNCHED("unit_pupil_kernel");
  return SHWFS_OK;
}

int32_t shwfs_binmap_block_origin(const int32_t* h_binmap, int32_t nfft, int32_t npix,
                                  int32_t nrebin, int32_t* origin) {
  if (!h_binmap || !origin || nfft <= 0 || npix <= 0 || nrebin <= 0 || npix * nrebin > nfft)
    return fail(SHWFS_ERR_ARG, "binmap_block_origin: bad arguments");
  if (!is_pow2(nfft)) return 0;
  const int nr2 = nrebin * nrebin;
  std::vector<int32_t> want(nr2), have(nr2);
  for (int o = 0; o < nfft; ++o) {
    bool ok = true;
    for (int p = 0; p < npix * npix && ok; ++p) {
      const int px = p % npix, py = p / npix;
      for (int b = 0; b < nrebin; ++b)
        for (int a = 0; a < nrebin; ++a)
          want[a + b * nrebin] = ((o + px * nrebin + a) % nfft) + ((o + py * nrebin + b) % nfft) * nfft;
      for (int i = 0; i < nr2; ++i) have[i] = h_binmap[i + p * nr2];
      std::sort(want.begin(), want.end());
      std::sort(have.begin(), have.end());
      ok = (want == have);
    }
    if (ok) {
      *origin = o;
      return 1;
    }
  }
  return 0;
}

// One frame of one WFS on stream st: the readout background of binimg, then
// the spot kernel of the frame's method.  A noisy frame of the FFT, DFT or
// fused learned-centroider kernel whose geometry lists the pixels outside the
// spot windows draws the background in the spot launch (extra workgroups on
// those pixels); any other frame with a background launches the background
// kernel over the whole image first, which the spots then overwrite.  wtab: the DFT operand table,
// or NULL for the library's own.
static int32_t comp_image_on(const shwfs_sh_geom* g, const shwfs_frame_args* f, hipStream_t st,
                             const void* wtab) {
  ImageParams p;
  int32_t rc = image_params(g, f, &p);
  if (rc) return rc;
  const bool ext = ext_needed(g);
  const bool mlp = p.ckind == SHWFS_CENTROID_MLP;
  BgParams bg;
  if (f->d_binimg && f->noise_background > 0.0f) {
    if (g->d_bgpix && p.noise > -1.0f && !ext && f->nlayers <= 0) {
      bg.pix = g->d_bgpix;
      bg.n = g->nbgpix;
      bg.sigma = f->noise_background;
    } else {
      rc = launch_background(f->d_binimg, (long)p.nsub * p.nsub, f->noise_background, f->seed,
                             f->call, st);
      if (rc) return rc;
    }
  }
  if (ext) return launch_image_ext(p, ext_params(g), g->nfft, g->ntot, st);
  if (f->nlayers > 0) {
    RayParams r;
    rc = ray_params(f->layers, f->nlayers, g->npup, g->npup, &r);
    if (rc) return rc;
    return launch_image_fft_ray(p, r, g->nfft, st);
  }
  if (mlp) {
    const shwfs_centroid_args* c = f->centroid;
    const MlpParams m{c->d_mlp_params, c->d_mlp_fallback, c->mlp_x_scale, c->mlp_y_scale,
                      c->mlp_flux_lo, c->mlp_flux_hi};
    if (f->method != SHWFS_METHOD_FFT_FP32)
      return launch_image_dft_mlp(p, m, g->nfft, f->method, st, wtab, bg);
    return launch_image_fft_mlp(p, m, g->nfft, st, bg);
  }
  if (f->method == SHWFS_METHOD_FFT_FP32) return launch_image_fft(p, g->nfft, st, bg);
  return launch_image_dft(p, g->nfft, f->method, st, wtab, bg);
}

int32_t shwfs_sh_comp_image(const shwfs_sh_geom* g, const shwfs_frame_args* f, void* stream) {
  return comp_image_on(g, f, (hipStream_t)stream, nullptr);
}

int32_t shwfs_sh_comp_image_noalloc(const shwfs_sh_geom* g, const shwfs_frame_args* f,
                                    void* stream) {
  ImageParams p;
  int32_t rc = image_params(g, f, &p);  // arguments checked before the table is written
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const void* wtab = nullptr;
  if (f->method != SHWFS_METHOD_FFT_FP32) {
    void* buf = scratch_address();
    if (!buf) return SHWFS_ERR_HIP;
    // the operand table of this shape, written into the scratch array ahead
    // of the spot kernel on the same stream (a size beyond the array is
    // rejected before any launch)
    size_t bytes = 0;
    rc = build_dft_operand_table(g->nfft, g->nphase, f->method, buf, kScratchBytes, st, &bytes);
    if (rc) return rc;
    wtab = buf;
  }
  return comp_image_on(g, f, st, wtab);
}

int32_t shwfs_centroids(const float* d_img_in, float* d_img_out, const shwfs_centroid_args* a,
                        void* stream) {
  if (!a || !d_img_in) return fail(SHWFS_ERR_ARG, "centroids: NULL argument");
  if (a->nvalid <= 0 || a->npix <= 0 || a->size <= 0)
    return fail(SHWFS_ERR_ARG, "centroids: sizes must be positive");
  if (!a->d_validx || !a->d_validy || !a->d_slopes || !a->d_intensities)
    return fail(SHWFS_ERR_STATE, "centroids: valid positions or outputs not set");
  hipStream_t st = (hipStream_t)stream;
  switch (a->kind) {
    case SHWFS_CENTROID_COG: return launch_centroid_k<SHWFS_CENTROID_COG>(d_img_in, nullptr, *a, st);
    case SHWFS_CENTROID_TCOG:
      return launch_centroid_k<SHWFS_CENTROID_TCOG>(d_img_in, d_img_out, *a, st);
    case SHWFS_CENTROID_WCOG:
      if (!a->d_weights) return fail(SHWFS_ERR_STATE, "WCOG: weights not initialised");
      return launch_centroid_k<SHWFS_CENTROID_WCOG>(d_img_in, nullptr, *a, st);
    default: return fail(SHWFS_ERR_ARG, "centroids: unknown kind");
  }
}

int32_t shwfs_centroids_bpcog(float* d_img, int32_t nbpix, const shwfs_centroid_args* a,
                              void* stream) {
  if (!a || !d_img) return fail(SHWFS_ERR_ARG, "centroids_bpcog: NULL argument");
  if (a->nvalid <= 0 || a->npix <= 0 || a->size <= 0)
    return fail(SHWFS_ERR_ARG, "centroids_bpcog: sizes must be positive");
  const int n2 = a->npix * a->npix;
  if (n2 > 1024) return fail(SHWFS_ERR_UNSUPPORTED, "centroids_bpcog: npix^2 above 1024");
  if (nbpix < 0 || nbpix >= n2)
    return fail(SHWFS_ERR_ARG, "centroids_bpcog: nbpix must lie in [0, npix^2)");
  if (!a->d_validx || !a->d_validy || !a->d_slopes || !a->d_intensities)
    return fail(SHWFS_ERR_STATE, "centroids_bpcog: valid positions or outputs not set");
  const int grid = (a->nvalid + 3) / 4;
  const size_t lds = sizeof(float) * 4 * (size_t)(n2 + 1);
  hipLaunchKernelGGL(bpcog_kernel, dim3(grid), dim3(256), lds, (hipStream_t)stream, d_img, nbpix,
                     *a);
  SHWFS_LAUNCHED("bpcog_kernel");
  return SHWFS_OK;
}

int32_t shwfs_fill_bincube(const float* d_image, float* d_bincube, const int32_t* d_validx,
                           const int32_t* d_validy, int32_t npix, int32_t nvalid, int32_t size,
                           void* stream) {
  if (!d_image || !d_bincube || !d_validx || !d_validy || npix <= 0 || nvalid < 0 || size <= 0)
    return fail(SHWFS_ERR_ARG, "fill_bincube: bad args");
  const long n = (long)nvalid * npix * npix;
  if (n == 0) return SHWFS_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_bincube_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, d_image, d_bincube, d_validx, d_validy, npix, n, size);
  SHWFS_LAUNCHED("fill_bincube_kernel");
  return SHWFS_OK;
}

int32_t shwfs_fill_bincube_2d(const float* d_image, float* d_bincube, const int32_t* d_num_ssp,
                              int32_t npix, int32_t nxsub, int32_t nvalid, void* stream) {
  if (!d_image || !d_bincube || !d_num_ssp || npix <= 0 || nxsub <= 0 || nvalid < 0)
    return fail(SHWFS_ERR_ARG, "fill_bincube_2d: bad args");
  const long n = (long)nxsub * nxsub * npix * npix;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_bincube_2d_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, d_image, d_bincube, d_num_ssp, npix, nxsub, nvalid);
  SHWFS_LAUNCHED("fill_bincube_2d_kernel");
  return SHWFS_OK;
}

int32_t shwfs_fill_weights(float* d_out, const float* d_in, int32_t npix, int32_t n_total,
                           void* stream) {
  if (!d_out || !d_in || npix <= 0 || n_total < 0) return fail(SHWFS_ERR_ARG, "fill_weights: bad args");
  if (n_total == 0) return SHWFS_OK;
  long blocks = (n_total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_weights_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     d_out, d_in, npix * npix, (long)n_total);
  SHWFS_LAUNCHED("fill_weights_kernel");
  return SHWFS_OK;
}

int32_t shwfs_calibrate(const float* d_raw, float* d_cal, const float* d_dark, const float* d_flat,
                        const int32_t* d_lutPix, int64_t n, void* stream) {
  if (!d_raw || !d_cal || !d_dark || !d_flat || !d_lutPix || n < 0)
    return fail(SHWFS_ERR_ARG, "calibrate: bad args");
  if (n == 0) return SHWFS_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(calibrate_kernel<float>, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, d_raw, d_cal, d_dark, d_flat, d_lutPix, (long)n);
  SHWFS_LAUNCHED("calibrate_kernel");
  return SHWFS_OK;
}

int32_t shwfs_calibrate_u16(const uint16_t* d_raw, float* d_cal, const float* d_dark,
                            const float* d_flat, const int32_t* d_lutPix, int64_t n, void* stream) {
  if (!d_raw || !d_cal || !d_dark || !d_flat || !d_lutPix || n < 0)
    return fail(SHWFS_ERR_ARG, "calibrate_u16: bad args");
  if (n == 0) return SHWFS_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(calibrate_kernel<uint16_t>, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, d_raw, d_cal, d_dark, d_flat, d_lutPix, (long)n);
  SHWFS_LAUNCHED("calibrate_kernel<uint16_t>");
  return SHWFS_OK;
}

int32_t shwfs_digitalize(uint16_t* d_cam, const float* d_img, const float* d_dark,
                         const float* d_flat, int32_t max_flux_per_pix, int32_t max_pix_value,
                         int64_t n, void* stream) {
  if (!d_cam || !d_img || !d_dark || !d_flat || n < 0)
    return fail(SHWFS_ERR_ARG, "digitalize: bad args");
  if (max_flux_per_pix <= 0) return fail(SHWFS_ERR_ARG, "digitalize: max_flux_per_pix must be > 0");
  if (n == 0) return SHWFS_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(digitalize_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     d_cam, d_img, d_dark, d_flat, (float)max_flux_per_pix, (float)max_pix_value,
                     (long)n);
  SHWFS_LAUNCHED("digitalize_kernel");
  return SHWFS_OK;
}

int32_t shwfs_slopes_geom(int32_t type, const float* d_phase, const float* d_pupil,
                          const float* d_fluxPerSub, float* d_slopes, const int32_t* d_phasemap,
                          const float* d_ttprojmat, int32_t nphase, int32_t nvalid, float alpha,
                          void* stream) {
  if (!d_phase || !d_slopes || !d_phasemap || nphase < 2 || nvalid <= 0)
    return fail(SHWFS_ERR_ARG, "slopes_geom: bad args");
  if (type < 0 || type > 2) return fail(SHWFS_ERR_ARG, "slopes_geom: type must be 0, 1 or 2");
  if (type == 1 && (!d_pupil || !d_fluxPerSub))
    return fail(SHWFS_ERR_ARG, "slopes_geom: type 1 needs the pupil and fluxPerSub");
  if (type == 2 && !d_ttprojmat) return fail(SHWFS_ERR_ARG, "slopes_geom: type 2 needs ttprojmat");
  hipLaunchKernelGGL(slopes_geom_kernel, dim3((nvalid + 3) / 4), dim3(256), 0, (hipStream_t)stream,
                     type, d_phase, d_pupil, d_fluxPerSub, d_slopes, d_phasemap, d_ttprojmat,
                     nphase, nvalid, alpha);
  SHWFS_LAUNCHED("slopes_geom_kernel");
  return SHWFS_OK;
}

int32_t shwfs_tt_filter(float* d_slopes, const float* d_ref_tip, const float* d_ref_tilt,
                        float* d_tt, int32_t nslopes, void* stream) {
  if (!d_slopes || !d_ref_tip || !d_ref_tilt || nslopes < 0)
    return fail(SHWFS_ERR_ARG, "tt_filter: bad args");
  if (nslopes == 0) return SHWFS_OK;
  hipLaunchKernelGGL(tt_filter_kernel, dim3(1), dim3(kTtThreads), 0, (hipStream_t)stream, d_slopes,
                     d_ref_tip, d_ref_tilt, d_tt, (int)nslopes);
  SHWFS_LAUNCHED("tt_filter_kernel");
  return SHWFS_OK;
}

int32_t shwfs_fill_binimg(const float* d_bincube, float* d_binimg, const int32_t* d_validsubsx,
                          const int32_t* d_validsubsy, int32_t npix, int32_t nvalid, int32_t nsub,
                          float noise_background, uint64_t seed, uint32_t call, uint32_t domain,
                          void* stream) {
  if (!d_bincube || !d_binimg || !d_validsubsx || !d_validsubsy || npix <= 0 || nvalid < 0 ||
      nsub <= 0)
    return fail(SHWFS_ERR_ARG, "fill_binimg: bad args");
  if ((long long)nsub * nsub >= (1LL << 31)) return fail(SHWFS_ERR_UNSUPPORTED, "fill_binimg: image too large");
  hipStream_t st = (hipStream_t)stream;
  if (noise_background > 0.0f) {
    int32_t rc = launch_background(d_binimg, (long)nsub * nsub, noise_background, seed, call, st,
                                   nullptr, domain);
    if (rc) return rc;
  }
  const long n = (long)nvalid * npix * npix;
  if (n == 0) return SHWFS_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_binimg_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d_bincube,
                     d_binimg, d_validsubsx, d_validsubsy, npix, n, nsub);
  SHWFS_LAUNCHED("fill_binimg_kernel");
  return SHWFS_OK;
}

int32_t shwfs_upload(void* d_dst, int64_t dst_elems, const void* h_src, int64_t src_elems,
                     int32_t elem_size, void* stream) {
  if (src_elems != dst_elems)
    return (set_error("upload: %lld elements given, the device array holds %lld",
                      (long long)src_elems, (long long)dst_elems),
            SHWFS_ERR_ARG);
  if (!d_dst || !h_src || dst_elems < 0 || elem_size <= 0)
    return fail(SHWFS_ERR_ARG, "upload: bad args");
  hipStream_t st = (hipStream_t)stream;
  SHWFS_HIP(hipMemcpyAsync(d_dst, h_src, (size_t)dst_elems * (size_t)elem_size,
                           hipMemcpyHostToDevice, st));
  SHWFS_HIP(hipStreamSynchronize(st));
  return SHWFS_OK;
}

int32_t shwfs_prng(float* d_data, int64_t n, char gtype, float alpha, float beta, uint64_t seed,
                   uint32_t call, uint32_t domain, void* stream) {
  if (!d_data || n < 0) return fail(SHWFS_ERR_ARG, "prng: bad args");
  if (gtype != 'P' && gtype != 'N' && gtype != 'U') return fail(SHWFS_ERR_ARG, "prng: gtype must be P, N or U");
  if (n > 0xFFFFFFFFLL) return fail(SHWFS_ERR_UNSUPPORTED, "prng: more than 2^32 elements");
  if (n == 0) return SHWFS_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(prng_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, d_data,
                     (long)n, gtype, alpha, beta, (unsigned long long)seed, call, domain);
  SHWFS_LAUNCHED("prng_kernel");
  return SHWFS_OK;
}

}  // extern "C"
