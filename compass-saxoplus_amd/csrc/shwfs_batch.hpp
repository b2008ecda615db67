// One launch per frame over every sub-aperture of several SH WFS of uneven
// size.  The work list is a table of tasks (one per WFS) in device memory;
// task i covers the global tiles [first_tile, first_tile + ntiles); a tile is
// one single-wave workgroup's worth of work: 64/nfft sub-apertures for the
// FFT stages, one sub-aperture for the MFMA DFT stages.  Workgroup b handles
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

template <int N, int Q, int PREC, bool NOISE>
__device__ __forceinline__ void batch_dft(const BatchTask& T, const ImageParams& P, int s,
                                          float* smem) {
  const AwOperand<N, Q, PREC> aw(static_cast<const Frag<PREC>*>(T.wtab), (int)threadIdx.x);
  dft_subap<N, Q, PREC, NOISE, false>(P, s, (int)threadIdx.x, smem, P.binmap, T.nr, aw);
}

template <bool NOISE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
void sh_batch_kernel(const BatchTask* __restrict__ tasks, int ntasks, int total_tiles,
                     const uint32_t* __restrict__ frame_base, uint32_t frame_off, int remap) {
  extern __shared__ float smem[];
  const uint32_t fcall = frame_base[0] + frame_off;
  const int G = (int)gridDim.x;
  const int b0 = remap ? xcd_remap((int)blockIdx.x, G) : (int)blockIdx.x;
  for (int t = b0; t < total_tiles; t += G) {
    int i = 0;
    while (i + 1 < ntasks && tasks[i + 1].first_tile <= t) ++i;
    const BatchTask& T = tasks[i];
    ImageParams P = T.P;
    P.call += fcall;
    const int s0 = (t - T.first_tile) * T.per_tile;
    switch (T.kind) {
      case BK_FFT16_1: fft_tile<16, 1, 64, NOISE>(P, s0, smem, P.binmap); break;
      case BK_FFT16_2: fft_tile<16, 2, 64, NOISE>(P, s0, smem, P.binmap); break;
      case BK_FFT32_1: fft_tile<32, 1, 64, NOISE>(P, s0, smem, P.binmap); break;
      case BK_FFT32_2: fft_tile<32, 2, 64, NOISE>(P, s0, smem, P.binmap); break;
      case BK_FFT64_1: fft_tile<64, 1, 64, NOISE>(P, s0, smem, P.binmap); break;
      case BK_FFT64_2: fft_tile<64, 2, 64, NOISE>(P, s0, smem, P.binmap); break;
      case BK_BF16_16: batch_dft<16, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(T, P, s0, smem); break;
      case BK_BF16_32: batch_dft<32, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(T, P, s0, smem); break;
      case BK_BF16_64: batch_dft<64, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(T, P, s0, smem); break;
      case BK_BF16_32Q2: batch_dft<32, 2, SHWFS_METHOD_DFT_BF16X3, NOISE>(T, P, s0, smem); break;
      case BK_BF16_64Q2: batch_dft<64, 2, SHWFS_METHOD_DFT_BF16X3, NOISE>(T, P, s0, smem); break;
      case BK_FP16_16: batch_dft<16, 1, SHWFS_METHOD_DFT_FP16, NOISE>(T, P, s0, smem); break;
      case BK_FP16_32: batch_dft<32, 1, SHWFS_METHOD_DFT_FP16, NOISE>(T, P, s0, smem); break;
      case BK_FP16_64: batch_dft<64, 1, SHWFS_METHOD_DFT_FP16, NOISE>(T, P, s0, smem); break;
      case BK_FP16_32Q2: batch_dft<32, 2, SHWFS_METHOD_DFT_FP16, NOISE>(T, P, s0, smem); break;
      case BK_FP16_64Q2: batch_dft<64, 2, SHWFS_METHOD_DFT_FP16, NOISE>(T, P, s0, smem); break;
      default: break;
    }
  }
}

// Launchers of the two instantiations (each in a translation unit of its own).
int32_t launch_batch_quiet(const BatchTask* tasks, int ntasks, int total_tiles, int grid,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st);
int32_t launch_batch_noisy(const BatchTask* tasks, int ntasks, int total_tiles, int grid,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st);

// Resident single-wave workgroups per CU for a given dynamic LDS size.
int batch_blocks_per_cu_quiet(size_t lds);
int batch_blocks_per_cu_noisy(size_t lds);

}  // namespace shwfs
