// copybw -- the practical HBM ceiling: a float4 device-to-device copy over n
// floats (read n*4 bytes, write n*4 bytes per launch), plain and with
// nontemporal loads/stores, timed with HIP events.  Also the calibration
// kernel of the FETCH_SIZE / WRITE_SIZE counters (known bytes per launch).
//   usage: copybw [n_floats (default 2^29)] [reps (default 20)]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_float4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void __launch_bounds__(256) copy_float4_nt(const float4* __restrict__ a, float4* __restrict__ b, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        float4 v;
        v.x = __builtin_nontemporal_load(&a[i].x);
        v.y = __builtin_nontemporal_load(&a[i].y);
        v.z = __builtin_nontemporal_load(&a[i].z);
        v.w = __builtin_nontemporal_load(&a[i].w);
        __builtin_nontemporal_store(v.x, &b[i].x);
        __builtin_nontemporal_store(v.y, &b[i].y);
        __builtin_nontemporal_store(v.z, &b[i].z);
        __builtin_nontemporal_store(v.w, &b[i].w);
    }
}

#define CHK(x)                                                                    \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::printf("%s failed: %s\n", #x, hipGetErrorString(e_));            \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main(int argc, char** argv)
{
    const size_t n = (argc > 1) ? std::strtoull(argv[1], nullptr, 10) : ((size_t)1 << 29);
    const int reps = (argc > 2) ? std::atoi(argv[2]) : 20;
    if (n < 4 || n % 4 != 0 || reps < 1) {
        std::printf("n must be a positive multiple of 4, reps >= 1\n");
        return 2;
    }
    float *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, n * sizeof(float)));
    CHK(hipMalloc(&b, n * sizeof(float)));
    CHK(hipMemset(a, 0, n * sizeof(float)));
    CHK(hipMemset(b, 0, n * sizeof(float)));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    const size_t n4 = n / 4;
    const unsigned grid = 256u * 1024u;
    for (int mode = 0; mode < 2; ++mode) {
        auto launch = [&]() {
            if (mode == 0) copy_float4<<<grid, 256>>>((const float4*)a, (float4*)b, n4);
            else copy_float4_nt<<<grid, 256>>>((const float4*)a, (float4*)b, n4);
        };
        for (int w = 0; w < 3; ++w) launch();
        CHK(hipEventRecord(e0));
        for (int r = 0; r < reps; ++r) launch();
        CHK(hipEventRecord(e1));
        CHK(hipEventSynchronize(e1));
        CHK(hipGetLastError());
        float ms = 0.f;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        const double bytes = 2.0 * (double)n * sizeof(float);
        std::printf("%s: %.0f bytes read + written per launch, %.4f ms/launch, %.1f GB/s\n",
                    mode == 0 ? "copy_float4" : "copy_float4_nt", bytes, ms / reps, bytes * reps / (ms * 1e-3) / 1e9);
    }
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
