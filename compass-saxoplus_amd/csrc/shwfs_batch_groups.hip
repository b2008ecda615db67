// One launch per frame over uneven WFS whose tiles fall into at most three
// stage kinds ("groups"), each kind's code compiled into the kernel without
// function calls.  The general batch kernel (shwfs_batch.hpp) dispatches to
// one non-inlined device function per kind; the calls save and restore
// registers through scratch memory on every tile, and the MFMA kinds there
// read their rebin factor at run time.  Here the stage of every group (method,
// transform size, window class, register rebin) is a template argument, the
// task parameters travel as kernel arguments, and each wave takes exactly one
// tile (no loop), as in the same-kind kernel.
//
// The kernels exist for the stage combinations listed in kGroupKernels (the
// uneven-batch workloads: 74x74 WFS at Nfft 16 / pdiam 8 / 2x2 binning,
// 40x40 at Nfft 32 / pdiam 16 / 4x4 binning to 8x8 pixels, 8x8 at Nfft 64 /
// pdiam 32 / 4x4 binning to 16x16 pixels, on the FFT, split-BF16x3 and FP16
// stages); any other combination runs the general kernel.
#include <algorithm>

#include "shwfs_batch.hpp"

namespace shwfs {

// Stage of a group: PREC < 0 for the FFT (QR = row split R), else the MFMA
// precision (QR = padded-window class Q, FULL, NR = register rebin code).
template <int PREC_, int N_, int QR_, bool FULL_, int NR_>
struct GroupStage {
  static constexpr int PREC = PREC_, N = N_, QR = QR_, NR = NR_;
  static constexpr bool FULL = FULL_;
  static constexpr int code = stage_code(PREC_, N_, QR_, FULL_, NR_);
};

// Sub-apertures per tile (one wave): 64 / N for the FFT; for the MFMA stages
// kDftTaskSubaps (spots one after the other, one shared tail).  The 64-point
// MFMA stage's sub-apertures are few and long (the frame's longest tiles):
// each takes a whole workgroup, one kx quarter per wave (dft_subap_parts),
// so its tiles are quarter sub-apertures and the stage must be the first
// group (workgroups then never straddle two stages).
template <class S>
__host__ __device__ constexpr int stage_subaps() {
  return S::PREC < 0 ? 64 / S::N : dft_stage_subaps(S::N);
}
template <class S>
__host__ __device__ constexpr bool stage_split() {
  return S::PREC >= 0 && S::N == 64;
}

// tt: the tile within its task
template <class S, bool NOISE>
__device__ __forceinline__ void run_stage(const ImageParams& P, int tt, int lane, int wave,
                                          float* smem, int slot, const void* wtab,
                                          unsigned call) {
  float* ws = smem + wave * slot;
  if constexpr (S::PREC < 0) {
    fft_tile<S::N, S::QR, 64, NOISE>(P, tt * stage_subaps<S>(), ws, P.binmap, call, lane);
  } else if constexpr (stage_split<S>()) {
    const AwOperand<S::N, S::QR, S::PREC> aw(static_cast<const Frag<S::PREC>*>(wtab), lane);
    dft_subap_parts<S::N, S::QR, S::PREC, NOISE, S::FULL, S::NR, kDftWaves>(
        P, min(tt / kDftWaves, P.nvalid - 1), lane, tt % kDftWaves, smem, aw, call);
  } else {
    const AwOperand<S::N, S::QR, S::PREC> aw(static_cast<const Frag<S::PREC>*>(wtab), lane);
    dft_subaps<S::N, S::QR, S::PREC, NOISE, S::FULL, S::NR, kDftTaskSubaps>(
        P, tt * stage_subaps<S>(), lane, ws, P.binmap, S::NR & 7, aw, call);
  }
}

// WPE: the register budget as resident waves per SIMD.  At 4 (128 VGPRs)
// the noise-free kernels spill 180-268 bytes per lane to scratch; at 3 (168)
// they spill nothing.  Which is faster depends on the stage mix, so both are
// compiled for the noise-free frames and the autotuner picks.
template <class S0, class S1, class S2, bool NOISE, int WPE>
__global__ __launch_bounds__(64 * kDftWaves) __attribute__((amdgpu_waves_per_eu(WPE)))
void sh_batch_groups_kernel(const TaskStarts starts, const TaskArgs args, int ntasks, int total,
                            GroupArgs ga, int slot, const uint32_t* __restrict__ frame_base,
                            uint32_t frame_off) {
  extern __shared__ float smem[];
  const int lane = (int)threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const uint32_t call = frame_base[0] + frame_off;
  // spare waves of the last workgroup recompute the last tile (same values)
  const int t = min((int)blockIdx.x * kDftWaves + wave, total - 1);
  const int i = find_task_args(starts, ntasks, t);
  static_assert(!stage_split<S1>() && !stage_split<S2>(), "the split stage must be the first");
  const ImageParams& P = args.P[i];
  const int tt = t - starts.first[i];  // tile within its task
  if (t < ga.first[1])
    run_stage<S0, NOISE>(P, tt, lane, wave, smem, slot, ga.wtab[0], call);
  else if (t < ga.first[2])
    run_stage<S1, NOISE>(P, tt, lane, wave, smem, slot, ga.wtab[1], call);
  else
    run_stage<S2, NOISE>(P, tt, lane, wave, smem, slot, ga.wtab[2], call);
}

template <class S0, class S1, class S2>
static int32_t launch_groups(bool noisy, int wpe, const TaskStarts& starts, const TaskArgs& args,
                             int ntasks, int total, const GroupArgs& ga, size_t lds,
                             const uint32_t* frame_base, uint32_t frame_off, hipStream_t st) {
  if (total <= 0) return SHWFS_OK;
  const int slot = (int)(lds / sizeof(float));
  const dim3 grid((total + kDftWaves - 1) / kDftWaves), block(64 * kDftWaves);
  if (noisy)  // the noisy kernels keep the 4-wave budget
    hipLaunchKernelGGL((sh_batch_groups_kernel<S0, S1, S2, true, 4>), grid, block,
                       lds * kDftWaves, st, starts, args, ntasks, total, ga, slot, frame_base,
                       frame_off);
  else if (wpe == 3)
    hipLaunchKernelGGL((sh_batch_groups_kernel<S0, S1, S2, false, 3>), grid, block,
                       lds * kDftWaves, st, starts, args, ntasks, total, ga, slot, frame_base,
                       frame_off);
  else
    hipLaunchKernelGGL((sh_batch_groups_kernel<S0, S1, S2, false, 4>), grid, block,
                       lds * kDftWaves, st, starts, args, ntasks, total, ga, slot, frame_base,
                       frame_off);
  SHWFS_LAUNCHED("sh_batch_groups_kernel");
  return SHWFS_OK;
}

namespace {

constexpr int kFft = -1, kBf = SHWFS_METHOD_DFT_BF16X3, kFp = SHWFS_METHOD_DFT_FP16;
constexpr int kHalf4 = 4 | kHalfGrid;
// the stages of the uneven-batch WFS
using Fft16 = GroupStage<kFft, 16, 2, false, 0>;   // 74x74: pdiam 8 <= 16/2
using Fft32 = GroupStage<kFft, 32, 2, false, 0>;   // 40x40: pdiam 16 <= 32/2
using Fft64 = GroupStage<kFft, 64, 2, false, 0>;   // 8x8: pdiam 32 <= 64/2
using Bf16a = GroupStage<kBf, 16, 1, false, 2>;    // 74x74: 8-point window, 2x2 register binning
using Bf32a = GroupStage<kBf, 32, 1, true, kHalf4>;
using Bf64b = GroupStage<kBf, 64, 2, true, kHalf4>;
using Fp16a = GroupStage<kFp, 16, 1, false, 2>;
using Fp32a = GroupStage<kFp, 32, 1, true, kHalf4>;
using Fp64b = GroupStage<kFp, 64, 2, true, kHalf4>;

using Launcher = int32_t (*)(bool, int, const TaskStarts&, const TaskArgs&, int, int,
                             const GroupArgs&, size_t, const uint32_t*, uint32_t, hipStream_t);
struct GroupKernel {
  int code[3];
  Launcher launch;
};

#define SHWFS_GK(A, B, C) {{A::code, B::code, C::code}, &launch_groups<A, B, C>}
const GroupKernel kGroupKernels[] = {
    SHWFS_GK(Fft64, Fft16, Fft32),  // every WFS on the FFT stage
    SHWFS_GK(Fft64, Bf16a, Bf32a),  // FP32-accurate: MFMA split-BF16x3, FFT for the 64-point WFS
    SHWFS_GK(Fft64, Fp16a, Fp32a),
    SHWFS_GK(Fft64, Fft16, Bf32a),
    SHWFS_GK(Fft64, Fft16, Fp32a),
    SHWFS_GK(Bf64b, Bf16a, Bf32a),  // every WFS on split-BF16x3
    SHWFS_GK(Fp64b, Fp16a, Fp32a),  // every WFS on FP16
    SHWFS_GK(Fp64b, Bf16a, Bf32a),
    SHWFS_GK(Fp64b, Fft16, Fft32),
};
#undef SHWFS_GK

}  // namespace

int group_kernel_index(const int codes[3]) {
  for (size_t k = 0; k < sizeof(kGroupKernels) / sizeof(kGroupKernels[0]); ++k)
    if (std::equal(codes, codes + 3, kGroupKernels[k].code)) return (int)k;
  return -1;
}

const int* group_kernel_codes(int index) { return kGroupKernels[index].code; }

int32_t launch_batch_groups(int index, bool noisy, int wpe, const TaskStarts& starts,
                            const TaskArgs& args, int ntasks, int total, const GroupArgs& ga,
                            size_t lds, const uint32_t* frame_base, uint32_t frame_off,
                            hipStream_t st) {
  if (index < 0 || index >= (int)(sizeof(kGroupKernels) / sizeof(kGroupKernels[0])))
    return fail(SHWFS_ERR_ARG, "batch: no group kernel of that index");
  if (wpe != 3 && wpe != 4) return fail(SHWFS_ERR_ARG, "batch: group kernels run 3 or 4 waves per SIMD");
  if (lds * kDftWaves > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "batch: LDS footprint exceeds the per-workgroup limit");
  return kGroupKernels[index].launch(noisy, wpe, starts, args, ntasks, total, ga, lds, frame_base,
                                     frame_off, st);
}

}  // namespace shwfs
