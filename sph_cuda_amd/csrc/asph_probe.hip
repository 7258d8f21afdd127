// asph_probe.hip - issue-rate probe for the VALU roofline of the neighbour kernels (measurement
// only; no reference counterpart). Independent chains of FP32 fused multiply-adds, non-packed as
// the library's kernels are compiled, in both encodings, over eight rounds of the resident
// wavefronts of every SIMD (a launch of one round lost up to 20% to its ramp and tail).
#include "asph_internal.h"

namespace asph {
namespace {

constexpr int PK = 16;        // independent chains per lane
constexpr int PTPB = 256;
constexpr int PBLK = 64;      // workgroups per CU: eight rounds of the resident waves

// VOP2 = false: a = fma(a, b, c), the three-operand v_fma_f32; true: a = fma(b, c, a), the
// two-operand v_fmac_f32 that most of the kernels' arithmetic compiles to.
template <bool VOP2>
__global__ __launch_bounds__(PTPB) void k_valu_probe(const float* __restrict__ in, float* __restrict__ out,
                                                     int iters) {
    const int t = blockIdx.x * PTPB + threadIdx.x;
    const float b = in[t & 255], c = in[(t + 7) & 255];
    float a[PK];
#pragma unroll
    for (int k = 0; k < PK; ++k) a[k] = in[(t + k) & 255];
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < PK; ++k) a[k] = VOP2 ? __builtin_fmaf(b, c, a[k]) : __builtin_fmaf(a[k], b, c);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < PK; ++k) s += a[k];
    out[t] = s;
}

}  // namespace
}  // namespace asph

extern "C" int asph_probe_valu(int32_t iters, int32_t reps, double* wave_fma_per_s) {
    using namespace asph;
    if (!wave_fma_per_s || iters < 1 || reps < 1) return ASPH_ERR;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return ASPH_ERR;
    const int nblk = cus * PBLK;
    float *in = nullptr, *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ASPH_ERR;
    double best = 0.0;
    if (hipMalloc(&in, 256 * sizeof(float)) == hipSuccess &&
        hipMalloc(&out, (size_t)nblk * PTPB * sizeof(float)) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        float h[256];
        for (int i = 0; i < 256; ++i) h[i] = 1.0f + 1e-6f * (float)i;
        if (hipMemcpy(in, h, sizeof(h), hipMemcpyHostToDevice) == hipSuccess) {
            rc = ASPH_OK;
            for (int r = 0; r <= 2 * reps + 1 && rc == ASPH_OK; ++r) {   // r = 0, 1: warm-up
                (void)hipEventRecord(e0, nullptr);
                if (r & 1) k_valu_probe<true><<<nblk, PTPB, 0, nullptr>>>(in, out, iters);
                else k_valu_probe<false><<<nblk, PTPB, 0, nullptr>>>(in, out, iters);
                (void)hipEventRecord(e1, nullptr);
                float ms = 0.f;
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
                    hipGetLastError() != hipSuccess) {
                    rc = ASPH_ERR;
                    break;
                }
                const double w = (double)nblk * (PTPB / 64) * PK * (double)iters / (ms * 1e-3);
                if (r > 1 && w > best) best = w;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    *wave_fma_per_s = best;
    return rc;
}
