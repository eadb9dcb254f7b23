// fa_probe.hip -- the VALU issue peak of the device, measured: a kernel of
// independent 32-bit integer add chains (v_add_u32), enough waves to fill
// every SIMD, timed with HIP events.  The frame kernels are bound by
// instruction issue, so their issue rates are quoted as fractions of this
// figure (wave64 VALU instructions per second over the whole chip).
#include <hip/hip_runtime.h>

#include "fa_internal.hpp"

namespace {

using fa::fail;

constexpr uint32_t kChains = 8;      // independent dependency chains per lane
constexpr uint32_t kUnroll = 16;     // adds per chain per loop iteration

// Each iteration issues kChains * kUnroll v_add_u32 (written as inline asm so
// the compiler cannot fold the adds); the chains are independent, so a SIMD
// with several resident waves can issue one every cycle it has a slot.
__global__ void __launch_bounds__(256) fa_valu_probe_kernel(uint32_t iters, uint32_t *sink) {
    uint32_t v[kChains];
    for (uint32_t c = 0; c < kChains; ++c) v[c] = threadIdx.x + c;
    const uint32_t inc = blockIdx.x | 1u;
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (uint32_t c = 0; c < kChains; ++c) __asm__ volatile("v_add_u32 %0, %1, %0" : "+v"(v[c]) : "v"(inc));
        }
    }
    uint32_t x = 0;
    for (uint32_t c = 0; c < kChains; ++c) x ^= v[c];
    if (x == 0x12345678u) sink[0] = x;   // keeps the chains live; practically never stored
}

}  // namespace

extern "C" {

// Runs the probe `reps` times (after one warm-up) with `iters` loop
// iterations per wave; returns the mean kernel time and the wave64 VALU adds
// per launch (waves x iters x 128).
int fa_probe_valu_peak(uint32_t iters, uint32_t reps, double *ms_per_launch, uint64_t *adds_per_launch) {
    if (!ms_per_launch || !adds_per_launch || iters == 0 || reps == 0) return fail(FA_EINVAL, "bad arguments");
    int dev = 0, cus = 0;
    FA_HIP_TRY(hipGetDevice(&dev));
    FA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 8 waves per SIMD: 4 SIMDs x 8 waves = 32 waves = 8 workgroups of 256 per CU
    const uint32_t blocks = (uint32_t)cus * 8u;
    uint32_t *sink = nullptr;
    FA_HIP_TRY(hipMalloc(&sink, 4));
    hipEvent_t a, b;
    FA_HIP_TRY(hipEventCreate(&a));
    FA_HIP_TRY(hipEventCreate(&b));
    hipLaunchKernelGGL(fa_valu_probe_kernel, dim3(blocks), dim3(256), 0, 0, iters, sink);
    FA_HIP_TRY(hipGetLastError());
    FA_HIP_TRY(hipEventRecord(a, 0));
    for (uint32_t r = 0; r < reps; ++r) hipLaunchKernelGGL(fa_valu_probe_kernel, dim3(blocks), dim3(256), 0, 0, iters, sink);
    FA_HIP_TRY(hipGetLastError());
    FA_HIP_TRY(hipEventRecord(b, 0));
    FA_HIP_TRY(hipEventSynchronize(b));
    float ms = 0.f;
    FA_HIP_TRY(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(sink);
    *ms_per_launch = (double)ms / reps;
    *adds_per_launch = (uint64_t)blocks * 4u * iters * kChains * kUnroll;
    return FA_OK;
}

}  // extern "C"
