// Dispatch of the MFMA spot kernels (instantiated per precision in
// shwfs_dft_*.hip) and the MFMA operand-layout probe.
#include "shwfs_image_dft.hpp"

namespace shwfs {

template <int PREC>
static int32_t launch_dft_p(const ImageParams& p, int nfft, hipStream_t st, const void* wtab,
                            const BgParams& bg) {
  if (p.pdiam <= 16) return launch_dft_pq<PREC, 1>(p, nfft, st, wtab, bg);
  if (p.pdiam <= 32) return launch_dft_pq<PREC, 2>(p, nfft, st, wtab, bg);
  return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path: needs nfft in {16, 32, 64} and pdiam <= min(32, nfft)");
}

const void* dft_operand_table(int nfft, int pdiam, int method, hipStream_t st) {
  const void* t = nullptr;
  const bool q1 = pdiam <= 16;
  if (pdiam > 32) {
    t = nullptr;
  } else if (method == SHWFS_METHOD_DFT_FP16) {
    t = q1 ? dft_table_pq<SHWFS_METHOD_DFT_FP16, 1>(nfft, st)
           : dft_table_pq<SHWFS_METHOD_DFT_FP16, 2>(nfft, st);
  } else if (method == SHWFS_METHOD_DFT_BF16X3) {
    t = q1 ? dft_table_pq<SHWFS_METHOD_DFT_BF16X3, 1>(nfft, st)
           : dft_table_pq<SHWFS_METHOD_DFT_BF16X3, 2>(nfft, st);
  }
  if (!t) set_error("no DFT operand table for nfft %d, pdiam %d, method %d", nfft, pdiam, method);
  return t;
}

int32_t launch_dft_tasks(int method, int nfft, int pdiam, bool full, int nr, bool noisy,
                         const BatchTask* tasks, const TaskStarts& starts, const TaskArgs& args,
                         int ntasks, int total, int grid, size_t lds,
                         const uint32_t* frame_base, uint32_t frame_off, int remap,
                         hipStream_t st) {
  const bool q1 = pdiam <= 16;
  if (pdiam > 32)
    return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT batch: pdiam must be <= 32");
#define SHWFS_ARGS \
  nfft, full, nr, noisy, tasks, starts, args, ntasks, total, grid, lds, frame_base, frame_off, remap, st
  if (method == SHWFS_METHOD_DFT_FP16)
    return q1 ? launch_dft_tasks_pq<SHWFS_METHOD_DFT_FP16, 1>(SHWFS_ARGS)
              : launch_dft_tasks_pq<SHWFS_METHOD_DFT_FP16, 2>(SHWFS_ARGS);
  if (method == SHWFS_METHOD_DFT_BF16X3)
    return q1 ? launch_dft_tasks_pq<SHWFS_METHOD_DFT_BF16X3, 1>(SHWFS_ARGS)
              : launch_dft_tasks_pq<SHWFS_METHOD_DFT_BF16X3, 2>(SHWFS_ARGS);
#undef SHWFS_ARGS
  return fail(SHWFS_ERR_ARG, "unknown DFT method");
}

int32_t launch_image_dft(const ImageParams& p, int nfft, int method, hipStream_t st,
                         const void* wtab, const BgParams& bg) {
  if (bg.n > 0 && !(p.noise > -1.0f))
    return fail(SHWFS_ERR_ARG, "DFT spot kernel: the fused background needs a noisy frame");
  if (method == SHWFS_METHOD_DFT_FP16)
    return launch_dft_p<SHWFS_METHOD_DFT_FP16>(p, nfft, st, wtab, bg);
  if (method == SHWFS_METHOD_DFT_BF16X3)
    return launch_dft_p<SHWFS_METHOD_DFT_BF16X3>(p, nfft, st, wtab, bg);
  return fail(SHWFS_ERR_ARG, "unknown DFT method");
}

int32_t build_dft_operand_table(int nfft, int pdiam, int method, void* dst, size_t cap,
                                hipStream_t st, size_t* bytes) {
  if (!dst) return fail(SHWFS_ERR_ARG, "DFT operand table: NULL destination");
  if (pdiam <= 0 || pdiam > 32)
    return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path: needs pdiam <= min(32, nfft)");
  const bool q1 = pdiam <= 16;
  if (method == SHWFS_METHOD_DFT_FP16)
    return q1 ? build_dft_table_pq<SHWFS_METHOD_DFT_FP16, 1>(nfft, dst, cap, st, bytes)
              : build_dft_table_pq<SHWFS_METHOD_DFT_FP16, 2>(nfft, dst, cap, st, bytes);
  if (method == SHWFS_METHOD_DFT_BF16X3)
    return q1 ? build_dft_table_pq<SHWFS_METHOD_DFT_BF16X3, 1>(nfft, dst, cap, st, bytes)
              : build_dft_table_pq<SHWFS_METHOD_DFT_BF16X3, 2>(nfft, dst, cap, st, bytes);
  return fail(SHWFS_ERR_ARG, "unknown DFT method");
}

// ---- layout probe: D = A B for one 16x32 by 32x16 tile, operands given in
// float (row-major), rounded to the MFMA input type, D row-major float.
template <int PREC>
__global__ void mfma_probe_kernel(const float* A, const float* B, float* D) {
  const int lane = threadIdx.x;
  Frag<PREC> a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * (lane >> 4) + e;
    a.set(e, A[(lane & 15) * 32 + k]);
    b.set(e, B[k * 16 + (lane & 15)]);
  }
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = mma(a, b, d);
#pragma unroll
  for (int v = 0; v < 4; ++v) D[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = d[v];
}

int32_t launch_mfma_probe(const float* A, const float* B, float* D, int method, hipStream_t st) {
  if (method == SHWFS_METHOD_DFT_FP16)
    hipLaunchKernelGGL(mfma_probe_kernel<SHWFS_METHOD_DFT_FP16>, dim3(1), dim3(64), 0, st, A, B, D);
  else if (method == SHWFS_METHOD_DFT_BF16X3)
    hipLaunchKernelGGL(mfma_probe_kernel<SHWFS_METHOD_DFT_BF16X3>, dim3(1), dim3(64), 0, st, A, B, D);
  else
    return fail(SHWFS_ERR_ARG, "mfma_probe: method must be a DFT method");
  SHWFS_LAUNCHED("mfma_probe_kernel");
  return SHWFS_OK;
}

}  // namespace shwfs
