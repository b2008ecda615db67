// LGS elongation: the sodium profile of every sub-aperture along its
// elongation, convolved with the laser beam (SutraLGS::lgs_update,
// libsutra/src.cpp/sutra_lgs.cpp:160-178), and the C ABI of the kernel
// builder (SutraLGS::lgs_makespot / load_kernels, whose kernel is in
// shwfs_image_ext.hip next to the transform it shares with the spot kernel).
//
// One workgroup of 64 lanes per sub-aperture, lane x = HR pixel x: the
// interpolated profile (iprof_krnl, libsutra/src.cu/sutra_lgs.cu:25-84), its
// length-ntot DFT, the product with ftbeam (tftbeam_krnl, 86-104) and the
// unnormalised inverse DFT, each output summed directly over ntot <= 64 terms
// with exact integer twiddle indices.
#include <cmath>

#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"

namespace shwfs {
namespace {

constexpr int kLgsLanes = 64;

struct ProfArgs {
  float2* prof2d;
  const float* prof1d;
  const float* profcum;
  int nprof;
  const float* doffaxis;
  float hg, pixsize, h0, deltah;
  const float2* ftbeam;
  int ntot;
};

// interp_transf: linear interpolation of a profile sampled every deltah
// metres at the HR pixel position t (sutra_lgs.cu:25-50)
__device__ float interp_transf(const float* idata, float t, int npix, float pixsize, float norm,
                               float shiftx, float deltah, int hmax) {
  if (npix % 2 == 0)
    t -= ((float)npix / 2.0f + 0.5f);
  else
    t -= ((float)npix / 2.0f);
  t *= pixsize;
  t /= norm;
  t += shiftx;
  t /= deltah;
  int xref;
  float w;
  if (t < 0.0f) {
    xref = 0;
    w = 0.0f;
  } else if (t < (float)(hmax - 2)) {
    xref = (int)t;
    w = t - (float)xref;
  } else {
    xref = hmax - 2;
    w = 1.0f;
  }
  return idata[xref] * (1.0f - w) + idata[xref + 1] * w;
}

__global__ __launch_bounds__(kLgsLanes) void lgs_prof_kernel(ProfArgs a) {
  __shared__ float2 tw[kLgsLanes];  // exp(-2 pi i m / n)
  __shared__ float2 p[kLgsLanes];
  const int s = blockIdx.x, x = threadIdx.x, n = a.ntot;
  if (x < n) {
    float sn, cs;
    sincospif(-2.0f * (float)x / (float)n, &sn, &cs);
    tw[x] = make_float2(cs, sn);
    const float norm = 206265.0f * a.doffaxis[s] / (a.hg * a.hg);
    const float shift = a.hg - a.h0;
    float v;
    if (a.pixsize > a.deltah * norm)
      v = interp_transf(a.profcum, (float)x + 1.0f, n, a.pixsize, norm, shift, a.deltah, a.nprof) -
          interp_transf(a.profcum, (float)x, n, a.pixsize, norm, shift, a.deltah, a.nprof);
    else
      v = interp_transf(a.prof1d, (float)x, n, a.pixsize, norm, shift, a.deltah, a.nprof);
    p[x] = make_float2(v, 0.0f);
  }
  __syncthreads();
  // forward DFT (imaginary input parts are zero), times ftbeam
  float2 F = make_float2(0.f, 0.f);
  if (x < n) {
    for (int m = 0; m < n; ++m) {
      const float2 w = tw[(x * m) & (n - 1)];
      F.x = fmaf(p[m].x, w.x, F.x);
      F.y = fmaf(p[m].x, w.y, F.y);
    }
    const float2 b = a.ftbeam[x];
    F = make_float2(F.x * b.x - F.y * b.y, F.y * b.x + F.x * b.y);
  }
  __syncthreads();
  if (x < n) p[x] = F;
  __syncthreads();
  // unnormalised inverse DFT
  if (x < n) {
    float2 o = make_float2(0.f, 0.f);
    for (int k = 0; k < n; ++k) {
      const float2 w = tw[(x * k) & (n - 1)];  // conjugated below
      const float2 q = p[k];
      o.x += q.x * w.x + q.y * w.y;
      o.y += q.y * w.x - q.x * w.y;
    }
    a.prof2d[(size_t)s * n + x] = o;
  }
}

}  // namespace
}  // namespace shwfs

using namespace shwfs;

extern "C" int32_t shwfs_lgs_prof2d(float* d_prof2d, const float* d_prof1d, const float* d_profcum,
                                    int32_t nprof, const float* d_doffaxis, float hg,
                                    float pixsize, float h0, float deltah, const float* d_ftbeam,
                                    int32_t ntot, int32_t nvalid, void* stream) {
  if (!d_prof2d || !d_prof1d || !d_profcum || !d_doffaxis || !d_ftbeam)
    return fail(SHWFS_ERR_ARG, "lgs_prof2d: NULL argument");
  if (nvalid <= 0 || nprof < 2) return fail(SHWFS_ERR_ARG, "lgs_prof2d: nvalid > 0, nprof >= 2");
  if (!(ntot == 16 || ntot == 32 || ntot == 64))
    return fail(SHWFS_ERR_UNSUPPORTED, "lgs_prof2d: ntot must be 16, 32 or 64");
  const ProfArgs a{reinterpret_cast<float2*>(d_prof2d), d_prof1d, d_profcum, nprof, d_doffaxis, hg,
                   pixsize, h0, deltah, reinterpret_cast<const float2*>(d_ftbeam), ntot};
  hipLaunchKernelGGL(lgs_prof_kernel, dim3(nvalid), dim3(kLgsLanes), 0, (hipStream_t)stream, a);
  SHWFS_LAUNCHED("lgs_prof_kernel");
  return SHWFS_OK;
}

extern "C" int32_t shwfs_lgs_kernels(float* d_ftlgskern, float* d_lgskern, const float* d_prof2d,
                                     const float* d_beam, const float* d_azimuth, int32_t ntot,
                                     int32_t nvalid, void* stream) {
  if (!d_ftlgskern) return fail(SHWFS_ERR_ARG, "lgs_kernels: NULL output");
  if (d_prof2d ? (!d_beam || !d_azimuth) : !d_lgskern)
    return fail(SHWFS_ERR_ARG, "lgs_kernels: beam and azimuth with a profile, else the kernels");
  if (nvalid <= 0) return fail(SHWFS_ERR_ARG, "lgs_kernels: nvalid must be positive");
  return launch_lgs_kernels(reinterpret_cast<float2*>(d_ftlgskern), d_lgskern,
                            reinterpret_cast<const float2*>(d_prof2d), d_beam, d_azimuth, ntot,
                            nvalid, (hipStream_t)stream);
}
