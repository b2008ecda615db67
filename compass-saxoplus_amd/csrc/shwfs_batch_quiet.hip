// Batch kernel instantiation: noise-free frames.
#include "shwfs_batch.hpp"

namespace shwfs {

int32_t launch_batch_quiet(const BatchTask* tasks, int ntasks, int total_tiles, int grid,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st) {
  hipLaunchKernelGGL(sh_batch_kernel<false>, dim3(grid), dim3(64), lds, st, tasks, ntasks,
                     total_tiles, frame_base, frame_off, remap);
  SHWFS_LAUNCHED("sh_batch_kernel");
  return SHWFS_OK;
}

int batch_blocks_per_cu_quiet(size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sh_batch_kernel<false>, 64, lds) != hipSuccess)
    return 0;
  return n;
}

}  // namespace shwfs
