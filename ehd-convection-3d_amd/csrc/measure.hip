// measure.hip -- the practical HBM ceiling of the GPU the library runs on
// (ehd_copy_ceiling): a float4 device-to-device copy, plain and with
// nontemporal loads and stores (the population streams of the LBM kernels
// use the nontemporal form).  Its known bytes per launch also calibrate the
// FETCH_SIZE / WRITE_SIZE counters of rocprofv3.
#include <hip/hip_runtime.h>

#include "box.h"

namespace {

template <bool NT>
__global__ void __launch_bounds__(256) copy_ceiling_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                           size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        if constexpr (NT) {
            float4 v;
            v.x = __builtin_nontemporal_load(&a[i].x);
            v.y = __builtin_nontemporal_load(&a[i].y);
            v.z = __builtin_nontemporal_load(&a[i].z);
            v.w = __builtin_nontemporal_load(&a[i].w);
            __builtin_nontemporal_store(v.x, &b[i].x);
            __builtin_nontemporal_store(v.y, &b[i].y);
            __builtin_nontemporal_store(v.z, &b[i].z);
            __builtin_nontemporal_store(v.w, &b[i].w);
        } else {
            b[i] = a[i];
        }
    }
}

struct CopyRun {
    void* a = nullptr;
    void* b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~CopyRun()
    {
        if (st) (void)hipStreamSynchronize(st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (st) (void)hipStreamDestroy(st);
    }
};

}  // namespace

extern "C" int ehd_copy_ceiling(size_t bytes, int reps, double* out2)
{
    if (!out2) return ehd::set_error(EHD_EINVAL, "null output");
    if (bytes < 16 || bytes % 16 != 0 || reps < 1)
        return ehd::set_error(EHD_EINVAL, "copy ceiling: bytes must be a positive multiple of 16, reps >= 1");
    CopyRun r;
    if (hipMalloc(&r.a, bytes) != hipSuccess || hipMalloc(&r.b, bytes) != hipSuccess)
        return ehd::set_error(EHD_ENOMEM, "copy ceiling: hipMalloc of 2 x %zu bytes failed", bytes);
    HIPCHK(hipStreamCreateWithFlags(&r.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&r.e0));
    HIPCHK(hipEventCreate(&r.e1));
    HIPCHK(hipMemsetAsync(r.a, 0, bytes, r.st));
    HIPCHK(hipMemsetAsync(r.b, 0, bytes, r.st));
    const size_t n4 = bytes / 16;
    const unsigned grid = 256u * 1024u;
    for (int mode = 0; mode < 2; ++mode) {
        auto launch = [&]() {
            if (mode == 0) copy_ceiling_kernel<false><<<grid, 256, 0, r.st>>>((const float4*)r.a, (float4*)r.b, n4);
            else copy_ceiling_kernel<true><<<grid, 256, 0, r.st>>>((const float4*)r.a, (float4*)r.b, n4);
        };
        for (int w = 0; w < 3; ++w) launch();
        HIPCHK(hipEventRecord(r.e0, r.st));
        for (int i = 0; i < reps; ++i) launch();
        HIPCHK(hipEventRecord(r.e1, r.st));
        HIPCHK(hipEventSynchronize(r.e1));
        HIPCHK(hipGetLastError());
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.e0, r.e1));
        out2[mode] = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    }
    return EHD_OK;
}
