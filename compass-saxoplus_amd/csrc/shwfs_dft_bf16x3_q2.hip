// MFMA spot kernels, bf16x3 operands, padded phase side 32.
#include "shwfs_image_dft.hpp"

namespace shwfs {
template int32_t launch_dft_pq<SHWFS_METHOD_DFT_BF16X3, 2>(const ImageParams&, int, hipStream_t,
                                                   const void*, const BgParams&);
template int32_t build_dft_table_pq<SHWFS_METHOD_DFT_BF16X3, 2>(int, void*, size_t, hipStream_t,
                                                        size_t*);
template const void* dft_table_pq<SHWFS_METHOD_DFT_BF16X3, 2>(int, hipStream_t);
template int32_t launch_dft_tasks_pq<SHWFS_METHOD_DFT_BF16X3, 2>(int, bool, int, bool, const BatchTask*,
                                                   const TaskStarts&, const TaskArgs&, int, int,
                                                   int, size_t, const uint32_t*, uint32_t, int,
                                                   hipStream_t);
}  // namespace shwfs
