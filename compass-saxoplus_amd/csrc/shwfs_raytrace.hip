// Stand-alone guide-star raytrace (shwfs_raytrace) and the host form of the
// layers shared with the spot kernel that traces its windows on the fly.
#include <cmath>

#include "shwfs_internal.hpp"
#include "shwfs_raytrace.hpp"

namespace shwfs {

int32_t ray_params(const shwfs_layer* layers, int32_t nlayers, int32_t nx, int32_t ny,
                   RayParams* out) {
  if (nlayers < 0 || nlayers > SHWFS_MAX_LAYERS)
    return fail(SHWFS_ERR_ARG, "raytrace: between 0 and SHWFS_MAX_LAYERS layers");
  if (nlayers > 0 && !layers) return fail(SHWFS_ERR_ARG, "raytrace: NULL layer array");
  if (nx <= 0 || ny <= 0) return fail(SHWFS_ERR_ARG, "raytrace: screen sides must be positive");
  RayParams r{};
  r.n = nlayers;
  r.nx = nx;
  r.ny = ny;
  for (int i = 0; i < nlayers; ++i) {
    const shwfs_layer& a = layers[i];
    if (!a.d_screen || a.dim < 2) return fail(SHWFS_ERR_ARG, "raytrace: layer screen NULL or dim < 2");
    if (!(a.delta != 0.0f) || !std::isfinite(a.delta))
      return fail(SHWFS_ERR_ARG, "raytrace: delta must be finite and non-zero");
    RayLayer& l = r.L[i];
    l.screen = a.d_screen;
    l.dim = a.dim;
    l.G = a.G;
    l.c = std::cos(a.thetaML);  // float cos / sin of the float angle
    l.s = std::sin(a.thetaML);
    l.xoff_d = a.xoff / a.delta;
    l.yoff_d = a.yoff / a.delta;
    l.delta = a.delta;
    l.dx = a.dx;
    l.dy = a.dy;
  }
  *out = r;
  return SHWFS_OK;
}

__global__ __launch_bounds__(256) void raytrace_kernel(float* __restrict__ out, RayParams R,
                                                       int accumulate) {
  const long n = (long)R.nx * R.ny;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int x = (int)(i % R.nx), y = (int)(i / R.nx);
    const float v = ray_phase(R, x, y);
    out[i] = accumulate ? out[i] + v : v;
  }
}

}  // namespace shwfs

using namespace shwfs;

extern "C" int32_t shwfs_raytrace(float* d_out, int32_t nx, int32_t ny, const shwfs_layer* layers,
                                  int32_t nlayers, int32_t accumulate, void* stream) {
  if (!d_out) return fail(SHWFS_ERR_ARG, "raytrace: NULL output");
  RayParams r;
  const int32_t rc = ray_params(layers, nlayers, nx, ny, &r);
  if (rc) return rc;
  const long n = (long)nx * ny;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(raytrace_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     d_out, r, accumulate ? 1 : 0);
  SHWFS_LAUNCHED("raytrace_kernel");
  return SHWFS_OK;
}
