// Field stop of an SH WFS: the pupil field pupil * exp(i scale phase), zero-
// padded to n x n, transformed to the focal plane, multiplied by the stop
// mask, and transformed back (reference: SutraWfsSH::comp_generic's field-
// stop branch, libsutra/src.cpp/sutra_wfs_sh.cpp:366-391: fillfsamplipup,
// CarmaFFT forward, apply_submask, CarmaFFT inverse, then fillcamplipup reads
// the result divided by n^2; kernels fsamplipup_krnl and submask_krnl,
// libsutra/src.cu/sutra_wfs.cu:130-181, 1868-1888).
//
// The filtered field F / n^2 is handed to the spot kernels as an amplitude
// |F| / n^2 and a phase arg(F) / scale on the npup x npup support, which the
// spot field stage turns back into amplitude * exp(i (scale phase - offset))
// = F / n^2 * exp(-i offset): the same field the reference's camplipup
// builds from d_fsamplifoc, so every spot method runs unchanged on it.
//
// Three passes over an n x npup work array (rows y >= npup of the padded
// field are zero and are never stored):
//   rows:    the field of rows y < npup computed on the fly, forward transform;
//   columns: forward transform, product with the mask, inverse transform;
//   rows:    inverse transform, amplitude and phase of x, y < npup.
// Each workgroup transforms several lines held in LDS (radix-2, decimation in
// time from bit-reversed order for the forward transforms, decimation in
// frequency into bit-reversed order for the inverse ones), so the column
// pass reads and writes rows of adjacent columns.
#include <algorithm>
#include <cmath>

#include "shwfs_device.hpp"
#include "shwfs_internal.hpp"

namespace shwfs {
namespace {

constexpr int kFsThreads = 256;

struct FsArgs {
  const float* phase;
  const float* pupil;
  int npup;
  float scale;
  const float* mask;  // [n * n], element kx + ky * n
  int n, logn;
  float2* work;       // [n * npup], element x + y * n
  float* phase_out;   // [npup * npup]
  float* ampl_out;    // [npup * npup]
  int lines;          // lines per workgroup
  int ld;             // LDS stride of a line (n + 1)
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ int bitrev(int i, int logn) {
  return (int)(__builtin_bitreverse32((unsigned)i) >> (32 - logn));
}

// tw[k] = exp(-2 pi i k / n), k < n / 2
__device__ void init_twiddles(float2* tw, int n, int tid) {
  for (int k = tid; k < n / 2; k += kFsThreads) {
    float s, c;
    sincospif(-2.0f * (float)k / (float)n, &s, &c);
    tw[k] = make_float2(c, s);
  }
}

// Decimation in time: bit-reversed input, natural output.  INV: conjugate
// twiddles (the unnormalised inverse transform).
template <bool INV>
__device__ void fft_dit(float2* buf, int lines, int ld, int n, int logn, const float2* tw, int tid) {
  const int half = n >> 1;
  for (int s = 1; s <= logn; ++s) {
    const int h = 1 << (s - 1);
    const int step = n >> s;  // twiddle index stride
    for (int b = tid; b < lines * half; b += kFsThreads) {
      const int l = b >> (logn - 1), j = b & (half - 1);
      const int k = j & (h - 1);
      const int i0 = ((j >> (s - 1)) << s) + k;
      float2 w = tw[k * step];
      if (INV) w.y = -w.y;
      float2* p = buf + l * ld;
      const float2 a = p[i0];
      const float2 t = cmul(w, p[i0 + h]);
      p[i0] = make_float2(a.x + t.x, a.y + t.y);
      p[i0 + h] = make_float2(a.x - t.x, a.y - t.y);
    }
    __syncthreads();
  }
}

// Decimation in frequency: natural input, bit-reversed output.
template <bool INV>
__device__ void fft_dif(float2* buf, int lines, int ld, int n, int logn, const float2* tw, int tid) {
  const int half = n >> 1;
  for (int s = logn; s >= 1; --s) {
    const int h = 1 << (s - 1);
    const int step = n >> s;
    for (int b = tid; b < lines * half; b += kFsThreads) {
      const int l = b >> (logn - 1), j = b & (half - 1);
      const int k = j & (h - 1);
      const int i0 = ((j >> (s - 1)) << s) + k;
      float2 w = tw[k * step];
      if (INV) w.y = -w.y;
      float2* p = buf + l * ld;
      const float2 a = p[i0], c = p[i0 + h];
      p[i0] = make_float2(a.x + c.x, a.y + c.y);
      p[i0 + h] = cmul(make_float2(a.x - c.x, a.y - c.y), w);
    }
    __syncthreads();
  }
}

// Rows y < npup: field, forward transform along x.
__global__ __launch_bounds__(kFsThreads) void fs_rows_forward(FsArgs a) {
  extern __shared__ float2 sm[];
  float2* tw = sm;
  float2* buf = sm + a.n / 2;
  const int tid = threadIdx.x, n = a.n;
  const int y0 = blockIdx.x * a.lines;
  const int lines = min(a.lines, a.npup - y0);
  init_twiddles(tw, n, tid);
  for (int e = tid; e < lines * n; e += kFsThreads) {
    const int l = e >> a.logn, x = e & (n - 1);
    const int y = y0 + l;
    float2 v = make_float2(0.f, 0.f);
    if (x < a.npup) {
      const size_t p = (size_t)x + (size_t)y * a.npup;
      float s, c;
      phase_sincos(a.phase[p], a.scale, 0.0f, s, c);
      const float m = a.pupil[p];
      v = make_float2(c * m, s * m);
    }
    buf[l * a.ld + bitrev(x, a.logn)] = v;
  }
  __syncthreads();
  fft_dit<false>(buf, lines, a.ld, n, a.logn, tw, tid);
  for (int e = tid; e < lines * n; e += kFsThreads) {
    const int l = e >> a.logn, x = e & (n - 1);
    a.work[(size_t)x + (size_t)(y0 + l) * n] = buf[l * a.ld + x];
  }
}

// Columns: forward transform along y (rows y >= npup are zero), mask,
// inverse transform; rows y < npup stored back.
__global__ __launch_bounds__(kFsThreads) void fs_columns(FsArgs a) {
  extern __shared__ float2 sm[];
  float2* tw = sm;
  float2* buf = sm + a.n / 2;
  const int tid = threadIdx.x, n = a.n, C = a.lines;
  const int kx0 = blockIdx.x * C;
  const int cols = min(C, n - kx0);
  init_twiddles(tw, n, tid);
  for (int e = tid; e < cols * n; e += kFsThreads) {
    const int c = e % cols, y = e / cols;
    float2 v = make_float2(0.f, 0.f);
    if (y < a.npup) v = a.work[(size_t)(kx0 + c) + (size_t)y * n];
    buf[c * a.ld + bitrev(y, a.logn)] = v;
  }
  __syncthreads();
  fft_dit<false>(buf, cols, a.ld, n, a.logn, tw, tid);
  for (int e = tid; e < cols * n; e += kFsThreads) {
    const int c = e % cols, ky = e / cols;
    const float m = a.mask[(size_t)(kx0 + c) + (size_t)ky * n];
    float2& v = buf[c * a.ld + ky];
    v = make_float2(v.x * m, v.y * m);
  }
  __syncthreads();
  fft_dif<true>(buf, cols, a.ld, n, a.logn, tw, tid);
  for (int e = tid; e < cols * a.npup; e += kFsThreads) {
    const int c = e % cols, y = e / cols;
    a.work[(size_t)(kx0 + c) + (size_t)y * n] = buf[c * a.ld + bitrev(y, a.logn)];
  }
}

// Rows y < npup: inverse transform along x, then amplitude and phase of the
// field divided by n^2 on the npup x npup support.
__global__ __launch_bounds__(kFsThreads) void fs_rows_inverse(FsArgs a) {
  extern __shared__ float2 sm[];
  float2* tw = sm;
  float2* buf = sm + a.n / 2;
  const int tid = threadIdx.x, n = a.n;
  const int y0 = blockIdx.x * a.lines;
  const int lines = min(a.lines, a.npup - y0);
  init_twiddles(tw, n, tid);
  for (int e = tid; e < lines * n; e += kFsThreads) {
    const int l = e >> a.logn, x = e & (n - 1);
    buf[l * a.ld + x] = a.work[(size_t)x + (size_t)(y0 + l) * n];
  }
  __syncthreads();
  fft_dif<true>(buf, lines, a.ld, n, a.logn, tw, tid);
  const float inv = 1.0f / ((float)n * (float)n);  // a power of two: exact
  for (int e = tid; e < lines * a.npup; e += kFsThreads) {
    const int l = e / a.npup, x = e % a.npup;
    const float2 v = buf[l * a.ld + bitrev(x, a.logn)];
    const float re = v.x * inv, im = v.y * inv;
    const size_t o = (size_t)x + (size_t)(y0 + l) * a.npup;
    a.ampl_out[o] = sqrtf(re * re + im * im);
    a.phase_out[o] = atan2f(im, re) / a.scale;
  }
}

}  // namespace
}  // namespace shwfs

using namespace shwfs;

extern "C" int32_t shwfs_field_stop(const float* d_phase, const float* d_pupil, int32_t npup,
                                    float scale, const float* d_submask, int32_t n, void* d_work,
                                    float* d_phase_out, float* d_ampl_out, void* stream) {
  if (!d_phase || !d_pupil || !d_submask || !d_work || !d_phase_out || !d_ampl_out)
    return fail(SHWFS_ERR_ARG, "field_stop: NULL argument");
  if (n < 16 || n > 4096 || (n & (n - 1)) != 0)
    return fail(SHWFS_ERR_UNSUPPORTED, "field_stop: n must be a power of two in [16, 4096]");
  if (npup <= 0 || npup > n) return fail(SHWFS_ERR_ARG, "field_stop: need 0 < npup <= n");
  if (!(scale != 0.0f) || !std::isfinite(scale))
    return fail(SHWFS_ERR_ARG, "field_stop: scale must be finite and non-zero");
  int logn = 0;
  while ((1 << logn) < n) ++logn;
  const size_t lds_cap = max_lds_per_block();
  const size_t tw_bytes = (size_t)(n / 2) * sizeof(float2);
  const size_t line_bytes = (size_t)(n + 1) * sizeof(float2);
  if (lds_cap < tw_bytes + line_bytes)
    return fail(SHWFS_ERR_UNSUPPORTED, "field_stop: one line does not fit in LDS");
  // lines per workgroup: as many as fit, at most 16 (the column pass reads
  // 16 adjacent columns, 128 contiguous bytes per row)
  const int lines = (int)std::min<size_t>(16, (lds_cap - tw_bytes) / line_bytes);
  FsArgs a{d_phase, d_pupil, npup, scale, d_submask, n, logn, static_cast<float2*>(d_work),
           d_phase_out, d_ampl_out, lines, n + 1};
  const size_t lds = tw_bytes + (size_t)lines * line_bytes;
  const hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fs_rows_forward, dim3((npup + lines - 1) / lines), dim3(kFsThreads), lds, st, a);
  SHWFS_LAUNCHED("fs_rows_forward");
  hipLaunchKernelGGL(fs_columns, dim3((n + lines - 1) / lines), dim3(kFsThreads), lds, st, a);
  SHWFS_LAUNCHED("fs_columns");
  hipLaunchKernelGGL(fs_rows_inverse, dim3((npup + lines - 1) / lines), dim3(kFsThreads), lds, st, a);
  SHWFS_LAUNCHED("fs_rows_inverse");
  return SHWFS_OK;
}
