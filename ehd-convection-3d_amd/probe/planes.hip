// Copy through NP separate source planes and NP destination planes, the way
// the LBM kernel streams its population planes (54 read, 54 written), against
// the same bytes through 2 and 8 planes.  Total per launch: 54 x 8,388,608
// floats read and as many written (the populations of one 256x256x128 step).
// Nontemporal loads and stores, one thread per element of a plane.
#include <hip/hip_runtime.h>

#include <cstdio>

template <int NP>
__global__ void __launch_bounds__(256) plane_copy(const float* __restrict__ src, float* __restrict__ dst, unsigned S)
{
    const unsigned s = blockIdx.x * 256u + threadIdx.x;
    if (s >= S) return;
    float v[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) v[k] = __builtin_nontemporal_load(src + (size_t)k * S + s);
#pragma unroll
    for (int k = 0; k < NP; ++k) __builtin_nontemporal_store(v[k] + 1.0f, dst + (size_t)k * S + s);
}

int main()
{
    const size_t total = (size_t)54 * 8388608;   // floats per side
    float *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, total * 4) != hipSuccess || hipMalloc(&b, total * 4) != hipSuccess) {
        printf("allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, total * 4) != hipSuccess || hipMemset(b, 0, total * 4) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    auto timeit = [&](const char* name, auto launch) {
        for (int w = 0; w < 3; ++w) launch();
        (void)hipEventRecord(e0);
        for (int r = 0; r < 10; ++r) launch();
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        printf("%-10s %.4f ms  %.1f GB/s (%s)\n", name, ms / 10, 2.0 * total * 4 * 10 / (ms * 1e-3) / 1e9,
               hipGetErrorString(hipGetLastError()));
    };
    timeit("2 planes", [&] { const unsigned S = total / 2; plane_copy<2><<<(S + 255) / 256, 256>>>(a, b, S); });
    timeit("8 planes", [&] { const unsigned S = total / 8; plane_copy<8><<<(S + 255) / 256, 256>>>(a, b, S); });
    timeit("27 planes", [&] { const unsigned S = total / 27; plane_copy<27><<<(S + 255) / 256, 256>>>(a, b, S); });
    timeit("54 planes", [&] { const unsigned S = total / 54; plane_copy<54><<<(S + 255) / 256, 256>>>(a, b, S); });
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
