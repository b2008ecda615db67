// One launch per frame over every sub-aperture of several SH WFS of uneven
// size.  The work list is a table of tasks (one per WFS) in device memory;
// task i covers the global tiles [first_tile, first_tile + ntiles); a tile is
// one single-wave workgroup's worth of work: 64/nfft sub-apertures for the
// FFT stages, kDftTaskSubaps sub-apertures for the MFMA DFT stages (one
// shared tail).  Workgroup b handles
// tiles b, b + grid, ... (grid = number of tiles: the dispatcher balances the
// workgroups as they drain; smaller grids make the launch persistent).
//
// The Philox call counter of every WFS is its base call plus a device-side
// frame counter (read by the kernel, advanced by a one-thread kernel after
// the frame), so a captured hipGraph replays successive frames.
#pragma once

#include "shwfs_image_dft.hpp"
#include "shwfs_image_fft.hpp"

namespace shwfs {

enum BatchKind : int32_t {
  BK_FFT16_1 = 0, BK_FFT16_2, BK_FFT32_1, BK_FFT32_2, BK_FFT64_1, BK_FFT64_2,
  BK_BF16_16, BK_BF16_32, BK_BF16_64, BK_BF16_32Q2, BK_BF16_64Q2,
  BK_FP16_16, BK_FP16_32, BK_FP16_64, BK_FP16_32Q2, BK_FP16_64Q2,
  BK_COUNT
};


// One device function per stage kind, not inlined into the batch kernel:
// each is register-allocated on its own (inlined into one switch, the kinds
// share one allocation and spill).  Arguments arrive in VGPRs; readfirstlane
// restores their wave uniformity so the task is read with scalar loads.
__device__ __forceinline__ const BatchTask* uniform_task(const BatchTask* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<const BatchTask*>(((uint64_t)hi << 32) | lo);
}

template <int N, int R, bool NOISE>
__device__ __noinline__ void batch_fft(const BatchTask* task, int s0, unsigned call) {
  extern __shared__ float smem[];
  const BatchTask* T = uniform_task(task);
  s0 = __builtin_amdgcn_readfirstlane(s0);
  call = __builtin_amdgcn_readfirstlane(call);
  fft_tile<N, R, 64, NOISE>(T->P, s0, smem, T->P.binmap, call);
}

// A DFT tile: kDftTaskSubaps sub-apertures from s0 by one wave (spots one
// after the other, one shared tail), rebin factor read from the task.
template <int N, int Q, int PREC, bool NOISE>
__device__ __noinline__ void batch_dft(const BatchTask* task, int s0, unsigned call) {
  extern __shared__ float smem[];
  const BatchTask* T = uniform_task(task);
  s0 = __builtin_amdgcn_readfirstlane(s0);
  call = __builtin_amdgcn_readfirstlane(call);
  const AwOperand<N, Q, PREC> aw(static_cast<const Frag<PREC>*>(T->wtab), (int)threadIdx.x);
  if (T->P.pdiam == DftLayout<N, Q>::P16)
    dft_subaps<N, Q, PREC, NOISE, true, -1, kDftTaskSubaps>(T->P, s0, (int)threadIdx.x, smem,
                                                            T->P.binmap, T->nr, aw, call);
  else
    dft_subaps<N, Q, PREC, NOISE, false, -1, kDftTaskSubaps>(T->P, s0, (int)threadIdx.x, smem,
                                                             T->P.binmap, T->nr, aw, call);
}

// Kind sets compiled into one batch kernel: the host launches the smallest
// instantiation that covers the kinds of its tasks (a kernel holding every
// kind pays for the union of their registers and code size).
enum BatchKindSet : int { BKS_FFT = 1, BKS_BF16 = 2, BKS_FP16 = 4, BKS_ALL = 7 };

__host__ __device__ constexpr int batch_kind_set(int kind) {
  return kind < BK_BF16_16 ? BKS_FFT : (kind < BK_FP16_16 ? BKS_BF16 : BKS_FP16);
}

template <bool NOISE, int SET>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(SET == BKS_FP16 && !NOISE ? 5 : 4)))
void sh_batch_kernel(const BatchTask* __restrict__ tasks, const TaskStarts starts, int ntasks,
                     int total_tiles, const uint32_t* __restrict__ frame_base, uint32_t frame_off,
                     int remap) {
  extern __shared__ float smem[];
  const uint32_t fcall = frame_base[0] + frame_off;
  const int G = (int)gridDim.x;
  const int b0 = remap ? xcd_remap((int)blockIdx.x, G) : (int)blockIdx.x;
  for (int t = b0; t < total_tiles; t += G) {
    const int i = find_task(starts, tasks, ntasks, t);
    // the task's parameters are read in place (uniform addresses: scalar
    // loads); a private copy would live in scratch memory
    const BatchTask& T = tasks[i];
    const int s0 = (t - T.first_tile) * T.per_tile;
    switch (T.kind) {
      case BK_FFT16_1:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<16, 1, NOISE>(&T, s0, fcall);
        break;
      case BK_FFT16_2:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<16, 2, NOISE>(&T, s0, fcall);
        break;
      case BK_FFT32_1:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<32, 1, NOISE>(&T, s0, fcall);
        break;
      case BK_FFT32_2:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<32, 2, NOISE>(&T, s0, fcall);
        break;
      case BK_FFT64_1:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<64, 1, NOISE>(&T, s0, fcall);
        break;
      case BK_FFT64_2:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<64, 2, NOISE>(&T, s0, fcall);
        break;
      case BK_BF16_16:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<16, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall);
        break;
      case BK_BF16_32:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<32, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall);
        break;
      case BK_BF16_64:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<64, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall);
        break;
      case BK_BF16_32Q2:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<32, 2, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall);
        break;
      case BK_BF16_64Q2:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<64, 2, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall);
        break;
      case BK_FP16_16:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<16, 1, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall);
        break;
      case BK_FP16_32:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<32, 1, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall);
        break;
      case BK_FP16_64:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<64, 1, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall);
        break;
      case BK_FP16_32Q2:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<32, 2, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall);
        break;
      case BK_FP16_64Q2:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<64, 2, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall);
        break;
      default: break;
    }
  }
}

// Launchers of the instantiations (noise-free / noisy in translation units of
// their own); `set` is a BatchKindSet.
int32_t launch_batch_quiet(int set, const BatchTask* tasks, const TaskStarts& starts, int ntasks,
                           int total_tiles, int grid,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st);
int32_t launch_batch_noisy(int set, const BatchTask* tasks, const TaskStarts& starts, int ntasks,
                           int total_tiles, int grid,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st);

// Resident single-wave workgroups per CU for a given dynamic LDS size.
int batch_blocks_per_cu_quiet(int set, size_t lds);
int batch_blocks_per_cu_noisy(int set, size_t lds);

}  // namespace shwfs
