// Plain device-to-device copy of a known size: the calibration kernel for HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) and the achievable-bandwidth reference next to the lattice kernels.
//   copy_bw [MiB] [iterations] [blocks]   prints one JSON line
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

// Each thread moves four 16-byte vectors per iteration, loads issued before stores.
__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += 4 * stride) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (t + u * stride < n) v[u] = src[t + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (t + u * stride < n) dst[t + u * stride] = v[u];
    }
}

#define CHK(x)                                                                      \
    do {                                                                            \
        hipError_t e = (x);                                                         \
        if (e != hipSuccess) {                                                      \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
            return 1;                                                               \
        }                                                                           \
    } while (0)

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 2048;
    const int iters = argc > 2 ? atoi(argv[2]) : 20;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 0, bytes));
    CHK(hipMemset(b, 0, bytes));
    hipDeviceProp_t prop;
    CHK(hipGetDeviceProperties(&prop, 0));
    const int blocks = argc > 3 ? atoi(argv[3]) : prop.multiProcessorCount * 16;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / iters;
    printf("{\"kernel\": \"copy_kernel\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"ms_per_launch\": %.5f, "
           "\"GBps\": %.1f, \"cus\": %d}\n", bytes, bytes, per, 2.0 * bytes / (per * 1e-3) / 1e9, prop.multiProcessorCount);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
