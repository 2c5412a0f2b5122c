// Allocation-free entry point: glsc3 on device vectors with its block
// partials in a static device array, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// nekbone_glsc3_noalloc_d computes what glsc3_ (src/math.f:603-617;
// legacy entry point in nb_legacy.hip) and nekbone_glsc3 compute, with the
// same two kernels and the same number of block partials, so its value is
// bit-identical to theirs.  It calls no hipMalloc, hipFree or allocating
// library routine and does not synchronise the host: the sum lands in the
// caller's device memory in stream order.
#include "nb_device.h"

// Block partials of one call: reduce_blocks(n) <= kMaxRedBlocks of them
// (8 KiB).  Each GPU holds its own copy until the process ends.  Every
// call on a GPU shares that GPU's copy; all calls go through the library
// stream and so run one after another, but two replays of a captured call
// launched on different streams at the same time race on it.  No entry
// needs clearing: each call writes every partial it reads.
static __device__ __align__(256) double scratch[nb::kMaxRedBlocks];

namespace nb {
namespace {
constexpr int64_t kNoallocMaxN = (int64_t(1) << 31) - 1;  // the library's vector limit
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
int g_device_now = -1;
}  // namespace

// At nekbone_init, with `device` current: the device address of this
// device's copy of the array (the array's name is a host-side stand-in,
// not a device address; it is never passed to a kernel).
int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(NEKBONE_ERR_UNSUPPORTED, "noalloc: device index beyond 63");
    g_device_now = device;
    if (g_scratch[device]) return NEKBONE_OK;
    void *ptr = nullptr;
    NB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double *>(ptr);
    return NEKBONE_OK;
}

}  // namespace nb

using namespace nb;

extern "C" int nekbone_glsc3_noalloc_d(const double *a, const double *b, const double *mult,
                                       int64_t n, double *result)
{
    if (!stream() || g_device_now < 0 || !g_scratch[g_device_now])
        return fail(NEKBONE_ERR_STATE, "glsc3_noalloc: call nekbone_init first");
    if (!a || !b || !mult || !result) return fail(NEKBONE_ERR_ARG, "glsc3_noalloc: null pointer");
    if (n < 0 || n > kNoallocMaxN)
        return fail(NEKBONE_ERR_ARG, "glsc3_noalloc: need 0 <= n <= 2^31 - 1");
    static_assert(kMaxRedBlocks <= int64_t(sizeof(scratch) / sizeof(double)),
                  "the block partials must fit the scratch array");
    return launch_glsc3_raw(a, b, mult, n, g_scratch[g_device_now], result);
}
