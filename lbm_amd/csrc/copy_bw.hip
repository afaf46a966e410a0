// Plain device-to-device copy of a known size: the calibration kernel for the HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) and the achievable-bandwidth reference next to the lattice kernels.
// A measurement tool built beside the library, not part of it.
//   copy_bw [MiB] [iterations] [blocks] [mode]   prints one JSON line
//   mode 0: grid-stride copy, four 16-byte vectors per thread per iteration (the calibration copy;
//           blocks default to 16 per CU); mode 4: one 16-byte vector per thread in a one-shot grid
//           (the ceiling of a single-pass stream; the grid covers the data)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                               \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            exit(1);                                                                         \
        }                                                                                    \
    } while (0)

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

// One float4 per thread, no loop (the one-shot grid shape of the lattice kernels).
__global__ void __launch_bounds__(256) copy_once_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 2048;
    const int iters = argc > 2 ? atoi(argv[2]) : 20;
    const int mode = argc > 4 ? atoi(argv[4]) : 0;
    if (mib == 0 || iters < 1 || (mode != 0 && mode != 4)) {
        fprintf(stderr, "usage: copy_bw [MiB > 0] [iterations >= 1] [blocks] [mode 0|4]\n");
        return 2;
    }
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 0, bytes));
    CHK(hipMemset(b, 0, bytes));
    hipDeviceProp_t prop;
    CHK(hipGetDeviceProperties(&prop, 0));
    const size_t want = mode == 4 ? (n + 255) / 256 : (argc > 3 ? strtoull(argv[3], nullptr, 10) : prop.multiProcessorCount * 16);
    if (want < 1 || want > 0x7fffffffull) {
        fprintf(stderr, "copy_bw: grid of %zu blocks out of range\n", want);
        return 2;
    }
    const int blocks = (int)want;
    auto launch = [&]() {
        if (mode == 4) hipLaunchKernelGGL(copy_once_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
        else hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
    };
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    launch();
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; ++it) launch();
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / iters, moved = (double)bytes;
    printf("{\"kernel\": \"%s\", \"bytes_read\": %.0f, \"bytes_written\": %.0f, \"ms_per_launch\": %.5f, "
           "\"GBps\": %.1f, \"blocks\": %d, \"cus\": %d}\n", mode == 4 ? "copy_once_kernel" : "copy_kernel", moved,
           moved, per, 2 * moved / (per * 1e-3) / 1e9, blocks, prop.multiProcessorCount);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
