// MFMA pruned-DFT spot kernels (split-BF16x3 and FP16 with per-subap scale).
#include "shwfs_tail.hpp"

namespace shwfs {

int32_t launch_image_dft(const ImageParams& p, int nfft, int method, hipStream_t st) {
  (void)p; (void)nfft; (void)method; (void)st;
  return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path not built yet");
}

}  // namespace shwfs
