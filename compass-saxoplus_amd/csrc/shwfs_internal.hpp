// Host-side helpers shared by the translation units of libshwfs.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "shwfs.h"

namespace shwfs {

void set_error(const char* fmt, ...);

inline int32_t fail(int32_t code, const char* msg) {
  set_error("%s", msg);
  return code;
}

#define SHWFS_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::shwfs::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                         __FILE__, __LINE__);                                        \
      return SHWFS_ERR_HIP;                                                          \
    }                                                                                \
  } while (0)

// Launch-error check right after a kernel launch.
#define SHWFS_LAUNCHED(name)                                                         \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      ::shwfs::set_error("launch of %s failed: %s", name, hipGetErrorString(e_));    \
      return SHWFS_ERR_HIP;                                                          \
    }                                                                                \
  } while (0)

inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Dynamic LDS a workgroup may use on the current device (64 KiB if unknown).
inline size_t max_lds_per_block() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    return 65536;
  return (size_t)v;
}

// Kernel-side parameter block of the fused spot + centroid kernels.
struct ImageParams {
  const float* phase;
  const float* pupil;
  const float* offsets;
  const int* istart;
  const int* jstart;
  const int* validsubsx;
  const int* validsubsy;
  const float* flux;
  const int* binmap;
  const int* pupil_unit;  // optional [nvalid]: 1 = pupil window all ones (not read)
  int npup, nvalid, pdiam, npix, nrebin, nsub;
  int bin_block;   // 1: binmap is the (shifted) block form, indices computed from bin_origin
  int bin_origin;
  int tpb;         // threads per workgroup of the spot kernel (0 = default)
  float scale, nphot, noise;
  unsigned long long seed;
  unsigned call;
  float* bincube;
  float* binimg;
  float* intensities;
  // fused centroider
  int ckind, fp64, slope_order;
  float threshold, coffset, cscale;
  const float* weights;
  const float* ref;
  float* slopes;
  float* cintens;
};

// Readout background of the detector image drawn by extra workgroups of a
// noisy FFT or DFT spot kernel, after its spot workgroups: binimg[pix[i]] =
// sigma N(0, 1) (element = image index, Philox domain 2, the frame's call),
// i < n.  The pixels lie outside every spot window, so the two parts of the
// launch write disjoint pixels.  A kernel argument of its own (ImageParams is
// also the per-task record of the batches).
struct BgParams {
  const int* pix = nullptr;
  int n = 0;
  float sigma = 0.0f;
  int spot_blocks = 0;  // workgroups of the spot part (set by the launcher)
};

// Learned centroider run after binning by the FFT spot kernel (a kernel
// argument of its own, not part of ImageParams, which batches pass by value).
struct MlpParams {
  const float* params;  // [SHWFS_MLP_NPARAM]
  int* fallback;        // [nvalid] or NULL
  float xs, ys, lo, hi; // input scale, output scale, flux range of the network
};

// One WFS of a batch (device task table): its parameters, stage kind and
// range of global tiles.
struct BatchTask {
  ImageParams P;
  int32_t kind;        // BatchKind
  int32_t first_tile;  // global index of the task's first tile
  int32_t ntiles;
  int32_t per_tile;    // sub-apertures per tile
  int32_t nr;          // DFT register rebin factor (0: gather)
  int32_t pad_;
  const void* wtab;    // DFT operand table
};

// First tile of each task (ascending), passed by value as a kernel argument:
// a workgroup finds its task with scalar arithmetic instead of a chain of
// dependent loads from the task table; tasks past kTaskStarts are found in
// the table.  Entries past the last task hold kNoTask.
constexpr int kTaskStarts = 16;
constexpr int32_t kNoTask = 0x7fffffff;
struct TaskStarts {
  int32_t first[kTaskStarts];
};

// Parameters of the first kTaskArgs tasks as a kernel argument (read through
// the constant cache with the other arguments instead of from the task table
// in global memory).  Same-kind batches of at most kTaskArgs tasks use it.
constexpr int kTaskArgs = 16;
struct TaskArgs {
  ImageParams P[kTaskArgs];
};

#if defined(__HIPCC__)
// Sign bit of a wave-uniform x, computed in a scalar register (written as a
// scalar shift so that the compiler does not turn it into a per-lane compare
// and select).
__device__ __forceinline__ int uniform_sign_bit(int x) {
  int r;
  asm("s_lshr_b32 %0, %1, 31" : "=s"(r) : "s"(x));
  return r;
}

// Number of entries first[1 .. K-1] <= t (t >= 0, wave-uniform):
// t - first[k] >= 0 exactly when its sign bit is clear, and kNoTask entries
// never count.
template <int K>
__device__ __forceinline__ int count_starts_le(const TaskStarts& ts, int t) {
  int n = K - 1;
#pragma unroll
  for (int k = 1; k < K; ++k) n -= uniform_sign_bit(t - ts.first[k]);
  return n;
}

__device__ __forceinline__ int find_task(const TaskStarts& ts, const BatchTask* tasks, int ntasks,
                                         int t) {
  int i = count_starts_le<kTaskStarts>(ts, t);
  if (ntasks > kTaskStarts && i == kTaskStarts - 1) {  // past the argument table
    while (i + 1 < ntasks && tasks[i + 1].first_tile <= t) ++i;
  }
  return i;
}

// The same for at most kTaskArgs tasks (no table lookup).
__device__ __forceinline__ int find_task_args(const TaskStarts& ts, int ntasks, int t) {
  (void)ntasks;  // entries past the last task hold kNoTask
  return count_starts_le<kTaskArgs>(ts, t);
}
#endif

// Validated kernel parameters of one frame of one WFS (no launch).
int32_t check_geom(const shwfs_sh_geom* g, int method);
int32_t image_params(const shwfs_sh_geom* g, const shwfs_frame_args* f, ImageParams* p);

// bg: the fused background (noisy frames; bg.n == 0 for none).
int32_t launch_image_fft(const ImageParams& p, int nfft, hipStream_t s, const BgParams& bg = {});
// The enlarged HR grid (ntot != nfft) and the kernel convolution: one
// workgroup per sub-aperture (shwfs_image_ext.hip).
struct ExtParams {
  const int* hrmap;        // [nfft^2] or NULL when ntot == nfft
  const float2* ftkernel;  // [ntot^2] or NULL without kernel convolution
  int conv;                // 1: convolve (kernel, LGS kernels or both)
  const float2* ftlgs;     // [ntot^2 * nvalid] per-sub-aperture LGS kernels, or NULL
  float kscale;            // factor of the product of both kernels (ntot^2, else 1)
};
// The parameters of the enlarged-grid / convolution kernel for a geometry,
// and whether the geometry needs that kernel.
bool ext_needed(const shwfs_sh_geom* g);
ExtParams ext_params(const shwfs_sh_geom* g);
int32_t launch_image_ext(const ImageParams& p, const ExtParams& x, int nfft, int ntot,
                         hipStream_t s);
// The LGS kernels' rotation and 2-D transform (shwfs_image_ext.hip).
int32_t launch_lgs_kernels(float2* ftlgs, float* lgskern, const float2* prof2d, const float* beam,
                           const float* azimuth, int ntot, int nvalid, hipStream_t s);
// The FFT spot kernel tracing its window phases through layers (RayParams,
// shwfs_raytrace.hpp) instead of reading p.phase.
struct RayParams;
int32_t launch_image_fft_ray(const ImageParams& p, const RayParams& r, int nfft, hipStream_t s);
// The FFT spot kernel with the learned centroider after binning (p.ckind ==
// SHWFS_CENTROID_MLP): 16 sub-apertures per workgroup, npix 8, nfft 16 or 32.
int32_t launch_image_fft_mlp(const ImageParams& p, const MlpParams& m, int nfft, hipStream_t s,
                             const BgParams& bg = {});
// The same after MFMA DFT spots (shwfs_dft_mlp.hip) for the shapes
// dft_mlp_fits accepts (nfft 32, pdiam 16, 8x8 px, half-grid 4x4 binning).
bool dft_mlp_fits(const ImageParams& p, int nfft);
int32_t launch_image_dft_mlp(const ImageParams& p, const MlpParams& m, int nfft, int method,
                             hipStream_t s, const void* wtab, const BgParams& bg);
// wtab: NULL (the library's operand table of the shape, built and cached on
// first use) or a table built by build_dft_operand_table.
int32_t launch_image_dft(const ImageParams& p, int nfft, int method, hipStream_t s,
                         const void* wtab = nullptr, const BgParams& bg = {});
int32_t launch_mfma_probe(const float* A, const float* B, float* D, int method, hipStream_t s);
// img = sigma * N(0,1) (Philox domain `domain`, element = image index); the
// call is `call` plus *call_base when call_base is not NULL (device frame
// counter of a batch).  The spot path draws its background with domain 2.
int32_t launch_background(float* img, long n, float sigma, unsigned long long seed, unsigned call,
                          hipStream_t s, const unsigned* call_base = nullptr,
                          unsigned domain = 2u);
// Same-kind batches of the MFMA stage: every task uses one (nfft, pdiam
// class, method, full window, rebin) variant; one single-wave workgroup per
// sub-aperture of the flattened task range.
int32_t launch_dft_tasks(int method, int nfft, int pdiam, bool full, int nr, bool noisy,
                         const BatchTask* tasks, const TaskStarts& starts, const TaskArgs& args,
                         int ntasks, int total, int grid, size_t lds,
                         const uint32_t* frame_base, uint32_t frame_off, int remap,
                         hipStream_t st);
// Device table of the DFT operand fragments for (nfft, pdiam, method), built
// on first use on the current device; NULL (and the error set) otherwise.
const void* dft_operand_table(int nfft, int pdiam, int method, hipStream_t s);
// The same table built into caller memory dst of cap bytes by one launch on
// s (no allocation, no wait); *bytes (may be NULL) receives its size.
int32_t build_dft_operand_table(int nfft, int pdiam, int method, void* dst, size_t cap,
                                hipStream_t s, size_t* bytes);

}  // namespace shwfs
