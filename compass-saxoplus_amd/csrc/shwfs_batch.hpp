// One launch per frame over every sub-aperture of several SH WFS of uneven
// size.  The work list is a table of tasks (one per WFS) in device memory;
// task i covers the global tiles [first_tile, first_tile + ntiles); a tile is
// one wave's worth of work: 64/nfft sub-apertures for the FFT stages,
// kDftTaskSubaps sub-apertures for the MFMA DFT stages (one shared tail).  A
// workgroup holds kBatchWaves independent waves, each with its own LDS slot
// and its own tiles (no workgroup barrier): wave w of workgroup b handles
// tiles kBatchWaves b + w, then that plus kBatchWaves grid, ... (grid = the
// tiles / kBatchWaves: the dispatcher balances the workgroups as they drain;
// smaller grids make the launch persistent).
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

// Independent waves per workgroup of the mixed batch kernel.
constexpr int kBatchWaves = 4;

// The wave's LDS slot of `slot` floats.
__device__ __forceinline__ float* wave_slot(int slot) {
  extern __shared__ float smem[];
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  return smem + wave * slot;
}

template <int N, int R, bool NOISE>
__device__ __noinline__ void batch_fft(const BatchTask* task, int s0, unsigned call, int slot) {
  const BatchTask* T = uniform_task(task);
  s0 = __builtin_amdgcn_readfirstlane(s0);
  call = __builtin_amdgcn_readfirstlane(call);
  slot = __builtin_amdgcn_readfirstlane(slot);
  fft_tile<N, R, 64, NOISE>(T->P, s0, wave_slot(slot), T->P.binmap, call,
                            (int)threadIdx.x & 63);
}

// A DFT tile: kDftTaskSubaps sub-apertures from s0 by one wave (spots one
// after the other, one shared tail), rebin factor read from the task.
template <int N, int Q, int PREC, bool NOISE>
__device__ __noinline__ void batch_dft(const BatchTask* task, int s0, unsigned call, int slot) {
  const BatchTask* T = uniform_task(task);
  s0 = __builtin_amdgcn_readfirstlane(s0);
  call = __builtin_amdgcn_readfirstlane(call);
  slot = __builtin_amdgcn_readfirstlane(slot);
  const int lane = (int)threadIdx.x & 63;
  float* smem = wave_slot(slot);
  const AwOperand<N, Q, PREC> aw(static_cast<const Frag<PREC>*>(T->wtab), lane);
  if (T->P.pdiam == DftLayout<N, Q>::P16)
    dft_subaps<N, Q, PREC, NOISE, true, -1, kDftTaskSubaps>(T->P, s0, lane, smem, T->P.binmap,
                                                            T->nr, aw, call);
  else
    dft_subaps<N, Q, PREC, NOISE, false, -1, kDftTaskSubaps>(T->P, s0, lane, smem, T->P.binmap,
                                                             T->nr, aw, call);
}

// Kind sets compiled into one batch kernel: the host launches the smallest
// instantiation that covers the kinds of its tasks (a kernel holding every
// kind pays for the union of their registers and code size).
enum BatchKindSet : int { BKS_FFT = 1, BKS_BF16 = 2, BKS_FP16 = 4, BKS_ALL = 7 };

__host__ __device__ constexpr int batch_kind_set(int kind) {
  return kind < BK_BF16_16 ? BKS_FFT : (kind < BK_FP16_16 ? BKS_BF16 : BKS_FP16);
}

// slot: LDS floats per wave (the largest slot of the tasks); launched with
// kBatchWaves waves per workgroup, or fewer (blockDim.x / 64) when that many
// slots exceed the workgroup's LDS.
template <bool NOISE, int SET>
__global__ __launch_bounds__(64 * kBatchWaves)
__attribute__((amdgpu_waves_per_eu(SET == BKS_FP16 && !NOISE ? 5 : 4)))
void sh_batch_kernel(const BatchTask* __restrict__ tasks, const TaskStarts starts, int ntasks,
                     int total_tiles, const uint32_t* __restrict__ frame_base, uint32_t frame_off,
                     int remap, int slot) {
  const uint32_t fcall = frame_base[0] + frame_off;
  const int G = (int)gridDim.x;
  const int W = (int)blockDim.x >> 6;  // kBatchWaves, fewer when the LDS slots do not fit
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int b0 = remap ? xcd_remap((int)blockIdx.x, G) : (int)blockIdx.x;
  for (int t = b0 * W + wave; t < total_tiles; t += G * W) {
    const int i = find_task(starts, tasks, ntasks, t);
    // the task's parameters are read in place (uniform addresses: scalar
    // loads); a private copy would live in scratch memory
    const BatchTask& T = tasks[i];
    const int s0 = (t - T.first_tile) * T.per_tile;
    switch (T.kind) {
      case BK_FFT16_1:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<16, 1, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FFT16_2:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<16, 2, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FFT32_1:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<32, 1, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FFT32_2:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<32, 2, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FFT64_1:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<64, 1, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FFT64_2:
        if constexpr ((SET & BKS_FFT) != 0) batch_fft<64, 2, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_BF16_16:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<16, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_BF16_32:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<32, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_BF16_64:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<64, 1, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_BF16_32Q2:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<32, 2, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_BF16_64Q2:
        if constexpr ((SET & BKS_BF16) != 0) batch_dft<64, 2, SHWFS_METHOD_DFT_BF16X3, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FP16_16:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<16, 1, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FP16_32:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<32, 1, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FP16_64:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<64, 1, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FP16_32Q2:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<32, 2, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall, slot);
        break;
      case BK_FP16_64Q2:
        if constexpr ((SET & BKS_FP16) != 0) batch_dft<64, 2, SHWFS_METHOD_DFT_FP16, NOISE>(&T, s0, fcall, slot);
        break;
      default: break;
    }
  }
}

// ---- stage-group kernels (shwfs_batch_groups.hip) ---------------------------
// Code of a stage: the FFT (prec < 0, qr = row split) or an MFMA stage (prec
// = method, qr = padded-window class, full window, register rebin code).
__host__ __device__ constexpr int stage_code(int prec, int n, int qr, bool full, int nr) {
  return ((((prec + 1) * 128 + n) * 4 + qr) * 2 + (full ? 1 : 0)) * 32 + nr;
}
// Sub-apertures per wave of an MFMA stage in a group kernel: one for the
// 64-point transform, kDftTaskSubaps otherwise.
__host__ __device__ constexpr int dft_stage_subaps(int n) { return n == 64 ? 1 : kDftTaskSubaps; }
// Groups of one launch: group g covers the tiles [first[g], first[g + 1])
// (first[3] implied by the tile count), its MFMA operand table wtab[g].
struct GroupArgs {
  int first[3];
  const void* wtab[3];
};
// Index of the group kernel compiled for the three stage codes (in group
// order), or -1; the codes of a kernel by index.
int group_kernel_index(const int codes[3]);
const int* group_kernel_codes(int index);
// lds: bytes of one wave's slot; one wave per tile, kDftWaves per workgroup.
int32_t launch_batch_groups(int index, bool noisy, int wpe, const TaskStarts& starts,
                            const TaskArgs& args, int ntasks, int total, const GroupArgs& ga,
                            size_t lds, const uint32_t* frame_base, uint32_t frame_off,
                            hipStream_t st);

// Launchers of the instantiations (noise-free / noisy in translation units of
// their own); `set` is a BatchKindSet.
// grid: workgroups of `waves` waves; lds: bytes of one wave's slot.
int32_t launch_batch_quiet(int set, const BatchTask* tasks, const TaskStarts& starts, int ntasks,
                           int total_tiles, int grid, int waves,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st);
int32_t launch_batch_noisy(int set, const BatchTask* tasks, const TaskStarts& starts, int ntasks,
                           int total_tiles, int grid, int waves,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st);

// Resident workgroups (kBatchWaves waves each) per CU for a given dynamic
// LDS size per workgroup.
int batch_blocks_per_cu_quiet(int set, int waves, size_t lds);
int batch_blocks_per_cu_noisy(int set, int waves, size_t lds);

}  // namespace shwfs
