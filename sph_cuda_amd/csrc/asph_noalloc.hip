// asph_noalloc.hip - the coarse dump without device allocation.
//
// asph_coarse_dump stages the block means in a device buffer that it allocates and frees on
// every call. asph_coarse_dump_noalloc runs the same kernel (k_coarse_dump) into a fixed
// file-scope device array instead, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid it). Every GPU the library runs on holds
// its own copy of the array (ASPH_NOALLOC_MAX_FLOATS floats, 4 MiB) until the process ends.
// Calls that overlap on one GPU share that array and race; calls are ordered on the handle's
// stream and the entry point synchronises it before returning.
#include <mutex>
#include <vector>

#include "asph_handle.h"

namespace {

// The staging buffer. In host code `scratch` names a host-side stand-in, not device memory:
// its device address is looked up once per device (scratch_ptr) and only that is used.
static __device__ __align__(256) float scratch[ASPH_NOALLOC_MAX_FLOATS];

std::mutex g_mu;
std::vector<float*> g_ptr;   // per device id: the device address of `scratch`, or null

// Device address of `scratch` on the handle's device, which must be the current one (looked up on
// first use per device).
int scratch_ptr(const asph_t* h, float** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != h->device)
        return fail("coarse dump (no alloc): device %d is current, the handle belongs to device %d", dev,
                    h->device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev >= (int)g_ptr.size()) g_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_ptr[(size_t)dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail("coarse dump (no alloc): hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_ptr[(size_t)dev] = (float*)p;
    }
    *out = g_ptr[(size_t)dev];
    return ASPH_OK;
}

}  // namespace

extern "C" int asph_coarse_dump_noalloc(asph_t* h, int32_t sx, int32_t sy, int32_t sz, float* out) {
    if (!h || !out) return fail("null argument");
    if (!h->sorted) return fail("no cell table: run asph_sort first");
    if (sx < 1 || sy < 1 || sz < 1) return fail("dump strides must be >= 1");
    const DevParams& P = h->mem_h[0].P;
    int64_t cx = (P.nx + sx - 1) / sx, cy = (P.ny + sy - 1) / sy, cz = (P.nz + sz - 1) / sz;
    int64_t n = cx * cy * cz * 6;
    if (n > (int64_t)ASPH_NOALLOC_MAX_FLOATS)
        return fail("coarse dump (no alloc): %lld x %lld x %lld blocks need %lld floats, more than "
                    "the %d of the fixed buffer; use larger strides or asph_coarse_dump",
                    (long long)cx, (long long)cy, (long long)cz, (long long)n, ASPH_NOALLOC_MAX_FLOATS);
    float* d = nullptr;
    if (scratch_ptr(h, &d) != ASPH_OK) return ASPH_ERR;
    // every block's six means are written, empty blocks as zeros: nothing needs clearing first
    launch_coarse_dump(h->L, 0, h->stream, h->cstart, h->vel, h->thq, d, sx, sy, sz, (int)cx,
                       (int)cy, (int)cz);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("coarse dump (no alloc): %s", hipGetErrorString(e));
    return ASPH_OK;
}
