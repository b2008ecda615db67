// Guide-star raytrace: the phase a source sees through a set of layers,
// sum over the layers of a bilinear read of each layer at the source point
// mapped into it (magnification G, rotation thetaML, shifts, cone
// compression delta).  Semantics of generic_raytrace
// (libsutra/src.cu/sutra_target.cu:25-66) as SutraSource::raytrace calls it
// (libsutra/src.cpp/sutra_source.cpp:222-267): the screen starts at zero and
// each layer adds its value in turn; a point whose interpolation square
// leaves the layer adds nothing.  Used by the stand-alone raytrace kernel and
// by the spot kernel that computes the phase of its windows on the fly.
#pragma once

#include <hip/hip_runtime.h>

#include "shwfs.h"

namespace shwfs {

// One layer as the kernels read it (built on the host from shwfs_layer).
struct RayLayer {
  const float* screen;  // [dim * dim], x fastest
  int dim;
  float G, c, s;        // magnification, cos / sin of the rotation
  float xoff_d, yoff_d; // xoff / delta, yoff / delta
  float delta, dx, dy;
};

// Kernel argument: the layers and the side of the source screen.
struct RayParams {
  RayLayer L[SHWFS_MAX_LAYERS];
  int n;       // layers in use
  int nx, ny;  // source screen
};

// Value layer l adds at source point (x, y).  No contraction into fused
// multiply-adds: the stand-alone and the fused kernels give the same bits.
__device__ __forceinline__ float ray_layer(const RayLayer& l, int nx, int ny, int x, int y) {
#pragma clang fp contract(off)
  const float hx = (float)nx / 2.0f, hy = (float)ny / 2.0f;
  const float x1 = ((float)x - hx) * l.G;
  const float y1 = ((float)y - hy) * l.G;
  const float x2 = ((l.c * x1 - l.s * y1) + hx) + l.xoff_d;
  const float y2 = ((l.s * x1 + l.c * y1) + hy) + l.yoff_d;
  const float xref = x2 * l.delta + l.dx;
  const float yref = y2 * l.delta + l.dy;
  const int i = (int)xref, j = (int)yref;  // truncation, as the reference
  if (i < 0 || j < 0 || i + 1 >= l.dim || j + 1 >= l.dim) return 0.0f;
  const float u = xref - (float)i, v = yref - (float)j;
  const float wx1 = 1.0f - u, wx2 = u, wy1 = 1.0f - v, wy2 = v;
  const float* p = l.screen + (size_t)i + (size_t)j * (size_t)l.dim;
  const size_t d = (size_t)l.dim;
  return wx1 * wy1 * p[0] + wx2 * wy1 * p[1] + wx1 * wy2 * p[d] + wx2 * wy2 * p[d + 1];
}

// Phase of the source screen at (x, y): 0, then each layer added in order.
__device__ __forceinline__ float ray_phase(const RayParams& R, int x, int y) {
#pragma clang fp contract(off)
  float acc = 0.0f;
  for (int l = 0; l < R.n; ++l) acc += ray_layer(R.L[l], R.nx, R.ny, x, y);
  return acc;
}

// Host: kernel form of the caller's layers (checked: non-NULL screens,
// dim >= 2, delta != 0, at most SHWFS_MAX_LAYERS).
int32_t ray_params(const shwfs_layer* layers, int32_t nlayers, int32_t nx, int32_t ny,
                   RayParams* out);

}  // namespace shwfs
