// Conservation sums without any allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops).
//
// mrb_reduce_conservation keeps its per-block partial sums in the level's
// own buffer and adds them on the host after a synchronisation.  The
// variant here keeps them in a file-scope device array (one copy per GPU,
// held until the process ends) and adds them on the device, in the same
// block order, into a caller-provided device array: the five sums are
// bit-identical, nothing is allocated and nothing waits for the stream.
#include <mutex>

#include "mrb_internal.h"

namespace mrb {

namespace {

// the block partial sums: 5 doubles per block, kConservationBlocks blocks
// (the conservation kernel strides over any level size with at most that
// many blocks), 160 KiB per GPU
constexpr int kScratchDoubles = 5 * int(kConservationBlocks);
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (the array's name in host code is not a device
// address)
int scratch_on_current_device(double **out)
{
    int dev = -1;
    MRB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(MRB_E_UNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return MRB_OK;
}

// out[n] = sum over blocks b = 0..nb-1 of part[5 b + n], added in block
// order as mrb_reduce_conservation adds them on the host
__global__ void __launch_bounds__(64) k_sum_partials(const double *__restrict__ part, unsigned nb,
                                                    double *__restrict__ out)
{
    const int n = int(threadIdx.x);
    if (n >= 5) return;
    double acc = 0.0;
    for (unsigned b = 0; b < nb; ++b) acc += part[5 * size_t(b) + n];
    out[n] = acc;
}

}  // namespace

}  // namespace mrb

using namespace mrb;

extern "C" {

int mrb_reduce_conservation_noalloc(const mrb_level *L, double *out_device, void *stream)
{
    if (!L || !out_device) return set_error(MRB_E_ARG, "null argument");
    const unsigned nb = conservation_blocks(L);
    if (5 * size_t(nb) > size_t(kScratchDoubles))
        return set_error(MRB_E_ARG, "conservation sums need more partial sums than the scratch array holds");
    double *part = nullptr;
    int rc = scratch_on_current_device(&part);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((rc = conservation_partials(L, part, s))) return rc;
    k_sum_partials<<<1, 64, 0, s>>>(part, nb, out_device);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

}  // extern "C"
