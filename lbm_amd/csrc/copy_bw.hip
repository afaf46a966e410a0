// Plain device-to-device copy of a known size: the calibration kernel for HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) and the achievable-bandwidth reference next to the lattice kernels.
//   copy_bw [MiB] [iterations] [blocks] [mode]   prints one JSON line (mode 0 copy, 1 non-temporal
//   copy, 2 read only, 3 write only)
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

// Variants for characterising the memory system: non-temporal copy, read-only, write-only.
typedef float fvec4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy_nt_kernel(const float4* __restrict__ srcv, float4* __restrict__ dstv, size_t n) {
    const fvec4* src = reinterpret_cast<const fvec4*>(srcv);
    fvec4* dst = reinterpret_cast<fvec4*>(dstv);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += 4 * stride) {
        fvec4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (t + u * stride < n) v[u] = __builtin_nontemporal_load(src + t + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (t + u * stride < n) __builtin_nontemporal_store(v[u], dst + t + u * stride);
    }
}

__global__ void __launch_bounds__(256) read_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    float acc = 0.f;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const float4 v = src[t];
        acc += v.x + v.y + v.z + v.w;
    }
    if (acc == 12345.f) dst[0] = make_float4(acc, 0, 0, 0);   // keeps the loads; never true for zeros
}

__global__ void __launch_bounds__(256) write_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) dst[t] = make_float4(1.f, 2.f, 3.f, 4.f);
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
    const int mode = argc > 4 ? atoi(argv[4]) : 0;   // 0 copy, 1 non-temporal copy, 2 read, 3 write
    void (*kern)(const float4*, float4*, size_t) = mode == 1 ? copy_nt_kernel : mode == 2 ? read_kernel : mode == 3 ? write_kernel : copy_kernel;
    const char* names[4] = {"copy_kernel", "copy_nt_kernel", "read_kernel", "write_kernel"};
    const double rd = (mode == 3) ? 0.0 : 1.0, wr = (mode == 2) ? 0.0 : 1.0;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / iters;
    printf("{\"kernel\": \"%s\", \"bytes_read\": %.0f, \"bytes_written\": %.0f, \"ms_per_launch\": %.5f, "
           "\"GBps\": %.1f, \"blocks\": %d, \"cus\": %d}\n", names[mode], rd * bytes, wr * bytes, per,
           (rd + wr) * bytes / (per * 1e-3) / 1e9, blocks, prop.multiProcessorCount);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
