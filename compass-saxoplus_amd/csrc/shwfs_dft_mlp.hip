// The learned centroider after MFMA DFT spots, in one launch: a workgroup of
// eight waves computes eight sub-apertures' spots (one per wave, as the
// one-frame DFT kernel: every lane draws the noise of one pixel; the tail
// writes bincube, binimg and intensities and keeps the spot-cube values in
// LDS), then runs the network over them as rows 0-7 of a 16-row MFMA tile
// (rows 8-15 zero; waves 0-3 compute the products), and writes the slopes
// (the CoG of the same pixels outside the network's flux range).  The
// network part is the FFT-spot fused kernel's (shwfs_image_fft.hip), so the
// slopes equal the stand-alone network's over the same spot cube.  512 lanes
// per workgroup: two waves per SIMD, the noise samplers' spill-free budget.
// Shape: nfft 32, a full 16-point window, 8x8-pixel spots binned 4x4 in
// registers (the half-grid form); the FFT method covers the other shapes.
#include "shwfs_image_dft.hpp"
#include "shwfs_mlp.hpp"

namespace shwfs {
namespace {

constexpr int kDmWaves = 8;                 // one sub-aperture each
constexpr int kDmThreads = 64 * kDmWaves;
constexpr int kDmTile = 16;                 // MFMA rows of the MLP tile
constexpr int kDmN = 32, kDmNrc = 4 | kHalfGrid;

// Workgroups past bg.spot_blocks draw the fused background (noisy frames).
template <int PREC, bool NOISE>
__global__ __launch_bounds__(kDmThreads) void sh_image_dft_mlp_kernel(ImageParams P, MlpParams M,
                                                                      const Frag<PREC>* wtab,
                                                                      BgParams bg, int slot) {
  if constexpr (NOISE) {
    if ((int)blockIdx.x >= bg.spot_blocks) {
      draw_background<kDmThreads>(P, bg, (int)blockIdx.x - bg.spot_blocks);
      return;
    }
  }
  extern __shared__ float smem[];
  const int tid = (int)threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s0 = xcd_remap(blockIdx.x, bg.spot_blocks) * kDmWaves;
  {
    // spare waves of the last workgroup recompute the last sub-aperture (the
    // same values to the same addresses); their rows are not used
    const AwOperand<kDmN, 1, PREC> aw(wtab, lane);
    dft_subap<kDmN, 1, PREC, NOISE, true, kDmNrc, true>(P, min(s0 + wave, P.nvalid - 1), lane,
                                                       smem + wave * slot, P.binmap, kDmNrc & 7,
                                                       aw);
  }
  __syncthreads();
  // the spots (wave w's at the start of its slot) into registers, then, past
  // a barrier, to their rows of the MLP tile (rows past them zero)
  const int rows = P.nvalid - s0 < kDmWaves ? P.nvalid - s0 : kDmWaves;
  constexpr int PER = kDmTile * 64 / kDmThreads;
  float v[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = tid + e * kDmThreads, k = i >> 6, c = i & 63;
    v[e] = k < rows ? smem[k * slot + c] : 0.0f;
  }
  __syncthreads();
  float* spots = smem;  // [16][64]
  const mlp::Lds<kDmTile> T(reinterpret_cast<char*>(smem + kDmTile * 64));
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = tid + e * kDmThreads, k = i >> 6, c = i & 63;
    spots[i] = v[e];
    T.x[k * mlp::SX + c] = (__bf16)(v[e] * M.xs);
  }
  // starts with a barrier
  mlp::tile_forward<kDmTile, mlp::Fp32Weights, false>(M.params, mlp::Fp32Weights{M.params}, T,
                                                      tid);
  for (int k = wave; k < rows; k += kDmWaves)
    mlp::spot_slopes(spots[k * 64 + lane], lane, s0 + k, P.nvalid, T.y[k * mlp::NOUT],
                     T.y[k * mlp::NOUT + 1], M.lo, M.hi, P.coffset, P.cscale, M.ys, P.ref,
                     P.slope_order, P.slopes, P.cintens, M.fallback);
}

template <int PREC>
int32_t launch_dft_mlp_p(const ImageParams& p, const MlpParams& m, hipStream_t st,
                         const void* wtab_in, BgParams bg) {
  const Frag<PREC>* wtab = wtab_in ? static_cast<const Frag<PREC>*>(wtab_in)
                                   : wfrag_table<kDmN, 1, PREC>(st);
  if (!wtab) return fail(SHWFS_ERR_HIP, "DFT + MLP kernel: building the DFT operand table failed");
  const int slot = DftLayout<kDmN, 1>::wave_floats(p.npix, false);
  size_t floats = (size_t)kDmWaves * slot;
  const size_t mlp_floats = (size_t)kDmTile * 64 + mlp::Lds<kDmTile>::bytes() / sizeof(float);
  if (floats < mlp_floats) floats = mlp_floats;
  const size_t lds = floats * sizeof(float);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "DFT + MLP kernel: LDS footprint exceeds the per-workgroup limit");
  bg.spot_blocks = (p.nvalid + kDmWaves - 1) / kDmWaves;
  const bool noisy = p.noise > -1.0f;
  const int grid = bg.spot_blocks + (noisy ? (bg.n + kDmThreads - 1) / kDmThreads : 0);
  if (noisy)
    hipLaunchKernelGGL((sh_image_dft_mlp_kernel<PREC, true>), dim3(grid), dim3(kDmThreads), lds,
                       st, p, m, wtab, bg, slot);
  else
    hipLaunchKernelGGL((sh_image_dft_mlp_kernel<PREC, false>), dim3(grid), dim3(kDmThreads), lds,
                       st, p, m, wtab, bg, slot);
  SHWFS_LAUNCHED("sh_image_dft_mlp_kernel");
  return SHWFS_OK;
}

}  // namespace

bool dft_mlp_fits(const ImageParams& p, int nfft) {
  return nfft == kDmN && p.pdiam == 16 && p.npix == 8 && dft_half_grid(p, nfft);
}

int32_t launch_image_dft_mlp(const ImageParams& p, const MlpParams& m, int nfft, int method,
                             hipStream_t st, const void* wtab, const BgParams& bg) {
  if (!dft_mlp_fits(p, nfft))
    return fail(SHWFS_ERR_UNSUPPORTED, "DFT + MLP kernel: needs nfft 32, pdiam 16, 8x8 px, 4x4 binning");
  if (bg.n > 0 && !(p.noise > -1.0f))
    return fail(SHWFS_ERR_ARG, "DFT + MLP kernel: the fused background needs a noisy frame");
  if (p.nvalid <= 0) return SHWFS_OK;
  if (method == SHWFS_METHOD_DFT_BF16X3)
    return launch_dft_mlp_p<SHWFS_METHOD_DFT_BF16X3>(p, m, st, wtab, bg);
  if (method == SHWFS_METHOD_DFT_FP16)
    return launch_dft_mlp_p<SHWFS_METHOD_DFT_FP16>(p, m, st, wtab, bg);
  return fail(SHWFS_ERR_ARG, "DFT + MLP kernel: an MFMA method");
}

}  // namespace shwfs
