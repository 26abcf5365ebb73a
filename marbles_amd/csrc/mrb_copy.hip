// Streaming device-to-device copy: the measured HBM ceiling of a plain
// read + write kernel on this GPU, against which the step kernels' achieved
// bandwidth is compared.  No counterpart in the reference.
//
// A grid-stride loop of 16-byte vectors on a fixed number of blocks
// (default 1024: four per CU of a 256-CU device).  Measured on MI355X this
// plain form outran an unrolled one that keeps four vectors of distant
// addresses in flight per thread.
#include "mrb_internal.h"

namespace mrb {
namespace {

constexpr int kCopyBlock = 256;

__global__ void __launch_bounds__(kCopyBlock) k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                            int64_t n)
{
    const int64_t stride = int64_t(gridDim.x) * kCopyBlock;
    for (int64_t i = int64_t(blockIdx.x) * kCopyBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace
}  // namespace mrb

using namespace mrb;

extern "C" {

int mrb_stream_copy(void *dst, const void *src, int64_t bytes, int blocks, void *stream)
{
    if (!dst || !src) return set_error(MRB_E_ARG, "null buffer");
    if (bytes < 0 || bytes % 16 != 0) return set_error(MRB_E_ARG, "bytes must be a non-negative multiple of 16");
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 != 0)
        return set_error(MRB_E_ARG, "buffers must be 16-byte aligned");
    if (blocks < 0 || blocks > (1 << 20)) return set_error(MRB_E_ARG, "blocks must be in 0..2^20");
    if (bytes == 0) return MRB_OK;
    const int nb = blocks > 0 ? blocks : 1024;
    k_stream_copy<<<nb, kCopyBlock, 0, as_stream(stream)>>>(static_cast<const uint4 *>(src),
                                                           static_cast<uint4 *>(dst), bytes / 16);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

}  // extern "C"
