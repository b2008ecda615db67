// Correlation centroider: per sub-aperture, the cross-correlation of the
// spot with a reference function, its maximum and a parabola fit around it.
// Semantics of SutraCentroiderCorr::get_cog (libsutra/src.cpp/
// sutra_centroider_corr.cpp:190-262; kernels fillcorr, correl, roll2real,
// corr_norm, sortmaxi, interp_parab in libsutra/src.cu/
// sutra_centroider_corr.cu).  The reference correlates through a zero-padded
// 2n x 2n FFT; for the lags it keeps (|k| < n) that circular correlation is
// the linear one, which is summed here directly (n^2 terms per lag, (2n-1)^2
// lags) and scaled by (2n)^2 as the unnormalised inverse transform leaves it:
// one workgroup per sub-aperture, spot and reference in LDS.
#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"

namespace shwfs {
namespace {

constexpr int kCorrThreads = 256;
constexpr int kCorrMaxNpix = 16;   // (2n-1)^2 = 961 lags
constexpr int kCorrMaxInterp = 49; // sizex * sizey

struct CorrArgs {
  const float* img;
  int size;
  const int* validx;
  const int* validy;
  int nvalid, npix;
  const float* ref_fn;  // [npix^2] shared, or [npix^2 * nvalid]
  int ref_per_subap;
  const float* corrnorm;  // [(2n-1)^2] or NULL
  const float* interpmat; // [sizex*sizey + 6 columns]: element (t, c) at t + c * sizex * sizey
  int sizex, sizey;
  float offset, scale;
  const float* ref;       // [2 nvalid] or NULL
  int slope_order;
  float* slopes;
  float* corr;            // [(2n-1)^2 * nvalid] or NULL
  int* corrmax;           // [nvalid] or NULL
};

__global__ __launch_bounds__(kCorrThreads) void corr_kernel(CorrArgs a) {
  __shared__ float b[kCorrMaxNpix * kCorrMaxNpix];
  __shared__ float r[kCorrMaxNpix * kCorrMaxNpix];
  __shared__ float c[(2 * kCorrMaxNpix - 1) * (2 * kCorrMaxNpix - 1)];
  __shared__ float best_v[kCorrThreads];
  __shared__ int best_i[kCorrThreads];
  __shared__ float patch[kCorrMaxInterp];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int n = a.npix, n2 = n * n, N = 2 * n - 1;
  const int vx = a.validx[s], vy = a.validy[s];
  const float* rf = a.ref_fn + (a.ref_per_subap ? (size_t)s * n2 : 0);
  for (int i = tid; i < n2; i += kCorrThreads) {
    const int x = i % n, y = i / n;
    b[i] = a.img[(size_t)(vx + x) + (size_t)(vy + y) * a.size];
    r[i] = rf[i];
  }
  __syncthreads();
  // correlation at lag (kx, ky): sum over m of b[m + k] r[m]
  for (int l = tid; l < N * N; l += kCorrThreads) {
    const int kx = l % N - (n - 1), ky = l / N - (n - 1);
    const int x0 = kx < 0 ? -kx : 0, x1 = kx > 0 ? n - kx : n;
    const int y0 = ky < 0 ? -ky : 0, y1 = ky > 0 ? n - ky : n;
    float acc = 0.0f;
    for (int my = y0; my < y1; ++my)
      for (int mx = x0; mx < x1; ++mx) acc = fmaf(b[(mx + kx) + (my + ky) * n], r[mx + my * n], acc);
    acc *= (float)(4 * n2);  // the reference's unnormalised inverse transform on 2n x 2n
    if (a.corrnorm) acc = acc / a.corrnorm[l];
    c[l] = acc;
    if (a.corr) a.corr[(size_t)s * N * N + l] = acc;
  }
  __syncthreads();
  // maximum over the central nbmax x nbmax window, lowest window index on ties
  const int nbmax = N > 20 ? n : N;
  const int off = N / 2 - nbmax / 2;
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  for (int w = tid; w < nbmax * nbmax; w += kCorrThreads) {
    const int ix = off + w % nbmax, iy = off + w / nbmax;
    const float v = c[ix + iy * N];
    if (v > bv) {  // w increases: the first maximum of this thread is kept
      bv = v;
      bi = w;
    }
  }
  best_v[tid] = bv;
  best_i[tid] = bi;
  __syncthreads();
  for (int h = kCorrThreads / 2; h > 0; h >>= 1) {
    if (tid < h) {
      const float v2 = best_v[tid + h];
      const int i2 = best_i[tid + h];
      if (v2 > best_v[tid] || (v2 == best_v[tid] && i2 < best_i[tid])) {
        best_v[tid] = v2;
        best_i[tid] = i2;
      }
    }
    __syncthreads();
  }
  const int w = best_i[0];
  int ix = off + w % nbmax, iy = off + w / nbmax;
  if (tid == 0 && a.corrmax) a.corrmax[s] = ix + iy * N;
  // the patch centred on the maximum, kept inside the map
  const int hx = a.sizex / 2, hy = a.sizey / 2;
  ix = min(max(ix, hx), N - 1 - hx);
  iy = min(max(iy, hy), N - 1 - hy);
  const int ox = ix - hx, oy = iy - hy, np = a.sizex * a.sizey;
  for (int t = tid; t < np; t += kCorrThreads)
    patch[t] = c[(ox + t % a.sizex) + (oy + t / a.sizex) * N];
  __syncthreads();
  if (tid < 2) {
    float k[6];
    for (int q = 0; q < 6; ++q) {
      float acc = 0.0f;
      for (int t = 0; t < np; ++t) acc += a.interpmat[t + q * np] * patch[t];
      k[q] = acc;
    }
    const float denom = k[2] * k[2] - 4.0f * k[1] * k[0];
    float v = 0.0f;
    if (denom != 0.0f) {
      const float ctr = tid == 0 ? (2.0f * k[1] * k[3] - k[4] * k[2]) / denom
                                 : (2.0f * k[0] * k[4] - k[3] * k[2]) / denom;
      int m = tid == 0 ? 2 * ox + a.sizex : 2 * oy + a.sizey;
      m = (m & 1) == 0 ? m / 2 : m / 2 + 1;
      v = ((ctr + (float)m - 0.5f - (float)((N + 1) / 4)) - a.offset) * a.scale;
    }
    const int io = a.slope_order == 0 ? s + tid * a.nvalid : 2 * s + tid;
    a.slopes[io] = v - (a.ref ? a.ref[io] : 0.0f);
  }
}

}  // namespace
}  // namespace shwfs

using namespace shwfs;

extern "C" int32_t shwfs_centroids_corr(const float* d_img, int32_t size, const int32_t* d_validx,
                                        const int32_t* d_validy, int32_t nvalid, int32_t npix,
                                        const float* d_ref_fn, int32_t ref_per_subap,
                                        const float* d_corrnorm, const float* d_interpmat,
                                        int32_t sizex, int32_t sizey, float offset, float scale,
                                        const float* d_ref, int32_t slope_order, float* d_slopes,
                                        float* d_corr, int32_t* d_corrmax, void* stream) {
  if (!d_img || !d_validx || !d_validy || !d_ref_fn || !d_interpmat || !d_slopes)
    return fail(SHWFS_ERR_ARG, "centroids_corr: NULL argument");
  if (nvalid <= 0 || npix <= 0 || size < npix)
    return fail(SHWFS_ERR_ARG, "centroids_corr: sizes must be positive, image >= one spot");
  if (npix > kCorrMaxNpix) return fail(SHWFS_ERR_UNSUPPORTED, "centroids_corr: npix above 16");
  const int N = 2 * npix - 1;
  if (sizex <= 0 || sizey <= 0 || sizex * sizey > kCorrMaxInterp || sizex > N || sizey > N)
    return fail(SHWFS_ERR_ARG, "centroids_corr: interpolation patch empty, above 49 points or "
                               "larger than the map");
  if (sizex * sizey < 6) return fail(SHWFS_ERR_ARG, "centroids_corr: a parabola fit needs 6 points");
  CorrArgs a{d_img, size, d_validx, d_validy, nvalid, npix, d_ref_fn, ref_per_subap ? 1 : 0,
             d_corrnorm, d_interpmat, sizex, sizey, offset, scale, d_ref, slope_order, d_slopes,
             d_corr, d_corrmax};
  hipLaunchKernelGGL(corr_kernel, dim3(nvalid), dim3(kCorrThreads), 0, (hipStream_t)stream, a);
  SHWFS_LAUNCHED("corr_kernel");
  return SHWFS_OK;
}
