// The no-allocation field entry point (ccoil_fields_noalloc) and the device scratch it works in.
//
// scratch is a file-scope device array: each device the library runs on holds its own copy, at its own address,
// from the time the code object is loaded until the process ends. In host code its name is only a stand-in for
// the symbol, so host code never passes it to a kernel or a copy; it uses the device address that
// hipGetSymbolAddress returns, looked up once per device (with that device current) when the first context on the
// device is created, and kept per device.
#include "context.hpp"

#include <mutex>

namespace ccoil {

constexpr size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mutex;
void *g_scratch_ptr[64] = {};   // per device ordinal
}  // namespace

// Device address of the current device's scratch (device must be current), looked up on the first call for it.
hipError_t scratch_address(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

size_t scratch_bytes() { return kScratchBytes; }

}  // namespace ccoil

using namespace ccoil;

extern "C" ccoil_status ccoil_fields_noalloc(ccoil_ctx ctx, const ccoil_coils *coils, const double *points,
                                             int64_t n_points, uint32_t quantities, int32_t precision,
                                             uint32_t flags, double *A, double *B, double *E, double *G,
                                             ccoil_report *report)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_fields_noalloc: ctx is NULL");
    if (!(flags & CCOIL_DEVICE_POINTERS))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "ccoil_fields_noalloc takes device pointers (CCOIL_DEVICE_POINTERS)");
    if (flags & CCOIL_AUTOTUNE)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "ccoil_fields_noalloc does not tune (tuning runs allocating calls)");
    if (!ctx->scratch) return fail(ctx, CCOIL_ERR_DEVICE, "ccoil_fields_noalloc: the device scratch is not set up");
    Arena arena;
    arena.base = (char *)ctx->scratch;
    arena.cap = kScratchBytes;
    arena.fixed = true;
    return fields_call(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G, report, &arena,
                       nullptr);
}
