// Batch kernel instantiations: frames with spot noise.
#include "shwfs_batch.hpp"

namespace shwfs {

template <int SET>
static int32_t launch_set(const BatchTask* tasks, const TaskStarts& starts, int ntasks,
                          int total_tiles, int grid, int waves, size_t lds,
                          const uint32_t* frame_base, uint32_t frame_off, int remap,
                          hipStream_t st) {
  // lds: one wave's slot; grid: workgroups of `waves` waves
  hipLaunchKernelGGL((sh_batch_kernel<true, SET>), dim3(grid), dim3(64 * waves), lds * waves, st,
                     tasks, starts, ntasks, total_tiles, frame_base, frame_off, remap,
                     (int)(lds / sizeof(float)));
  SHWFS_LAUNCHED("sh_batch_kernel");
  return SHWFS_OK;
}

int32_t launch_batch_noisy(int set, const BatchTask* tasks, const TaskStarts& starts, int ntasks,
                           int total_tiles, int grid, int waves,
                           size_t lds, const uint32_t* frame_base, uint32_t frame_off, int remap,
                           hipStream_t st) {
  switch (set) {
    case BKS_FFT: return launch_set<BKS_FFT>(tasks, starts, ntasks, total_tiles, grid, waves, lds, frame_base, frame_off, remap, st);
    case BKS_BF16: return launch_set<BKS_BF16>(tasks, starts, ntasks, total_tiles, grid, waves, lds, frame_base, frame_off, remap, st);
    case BKS_FP16: return launch_set<BKS_FP16>(tasks, starts, ntasks, total_tiles, grid, waves, lds, frame_base, frame_off, remap, st);
    default: return launch_set<BKS_ALL>(tasks, starts, ntasks, total_tiles, grid, waves, lds, frame_base, frame_off, remap, st);
  }
}

template <int SET>
static int occupancy(int waves, size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sh_batch_kernel<true, SET>, 64 * waves,
                                                   lds * waves) != hipSuccess)
    return 0;
  return n;
}

int batch_blocks_per_cu_noisy(int set, int waves, size_t lds) {
  switch (set) {
    case BKS_FFT: return occupancy<BKS_FFT>(waves, lds);
    case BKS_BF16: return occupancy<BKS_BF16>(waves, lds);
    case BKS_FP16: return occupancy<BKS_FP16>(waves, lds);
    default: return occupancy<BKS_ALL>(waves, lds);
  }
}

}  // namespace shwfs
