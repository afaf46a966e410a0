// Plain device-to-device copy of a known size: the calibration kernel for HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) and the achievable-bandwidth reference next to the lattice kernels.
//   copy_bw [MiB] [iterations] [blocks] [mode]   prints one JSON line (mode 0 copy, 1 non-temporal
//   copy, 2 read only, 3 write only, 4 one float4 per thread, 5 27 planes per thread as the
//   lattice kernels move them, 6/7/8 the same with 1/2/4 cells per thread and the third argument
//   as dynamic LDS bytes per block to cap occupancy, 9/10 the row-blocked layout with rows of
//   256/64 cells; modes 4-10 size their own grid; 11/12 the D3Q27 pull-stream copy of 4-byte
//   values over the 384x192x128 workload in the plane / row-blocked layout, 13/14 the same with
//   2-byte values)
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

// One float4 per thread, no loop (the one-shot grid shape of the lattice kernels).
__global__ void __launch_bounds__(256) copy_once_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

// The lattice kernels' stream shape: 27 planes of 4-byte values read and 27 written per thread,
// one cell per thread (loads issued before stores), non-temporal stores.
__global__ void __launch_bounds__(256) planes_kernel(const float* __restrict__ src, float* __restrict__ dst, size_t ncell) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float v[27];
#pragma unroll
    for (int l = 0; l < 27; ++l) v[l] = src[(size_t)l * ncell + c];
#pragma unroll
    for (int l = 0; l < 27; ++l) __builtin_nontemporal_store(v[l], dst + (size_t)l * ncell + c);
}

// The same with W consecutive cells per thread (W-wide vector per plane), and dynamic LDS to
// cap the number of resident waves.
template <int W>
__global__ void __launch_bounds__(256) planes_vec_kernel(const float* __restrict__ src, float* __restrict__ dst, size_t ncell) {
    typedef float vec __attribute__((ext_vector_type(W)));
    extern __shared__ float pad[];
    const size_t c = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
    if (c + W > ncell) return;
    vec v[27];
#pragma unroll
    for (int l = 0; l < 27; ++l) v[l] = *reinterpret_cast<const vec*>(src + (size_t)l * ncell + c);
#pragma unroll
    for (int l = 0; l < 27; ++l) __builtin_nontemporal_store(v[l], reinterpret_cast<vec*>(dst + (size_t)l * ncell + c));
    if (c == (size_t)-1) pad[0] = 0.f;
}

// Row-blocked layout [row][population][R cells]: the 27 values of a thread's cell lie R*4 bytes
// apart inside one 27*R*4-byte block.
template <int R>
__global__ void __launch_bounds__(256) rows_kernel(const float* __restrict__ src, float* __restrict__ dst, size_t ncell) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const size_t row = c / R, w = c % R;
    const size_t base = row * 27 * R + w;
    float v[27];
#pragma unroll
    for (int l = 0; l < 27; ++l) v[l] = src[base + (size_t)l * R];
#pragma unroll
    for (int l = 0; l < 27; ++l) __builtin_nontemporal_store(v[l], dst + base + (size_t)l * R);
}

// The step kernel's memory traffic without its arithmetic: a D3Q27 pull-stream copy of FP32
// populations over the 384x192x128 workload (contiguous axis a0 = 192 cells, padded rows of 256
// with the first interior cell 128-byte aligned), one cell per thread, blocks of one row, raw
// buffer loads with the population offset in the scalar operand, non-temporal stores.  The
// element of population l at (a0, a1, a2) is at  l*PL + a2*S2 + a1*S1 + a0 + 31:
//   planes (the library's layout)          PL = plane, S1 = 256,      S2 = 256*(n1+2)
//   row-blocked [row][population][256]     PL = 256,   S1 = 27*256,   S2 = 27*256*(n1+2)
struct LatShape {
    int n0, n1, n2;
    unsigned PL, S1, S2;
    unsigned elems;   // per buffer
};
// B: bytes per value (4: FP32 storage, 2: 16-bit storage)
template <int B>
__global__ void __launch_bounds__(1024) lattice_copy_kernel(const void* __restrict__ src, void* __restrict__ dst, LatShape g) {
    extern __shared__ float pad[];
    const int a0 = threadIdx.x + 1, a1 = blockIdx.y + 1, a2 = blockIdx.z + 1;
    if (a0 > g.n0) return;
    if (a0 == 0) pad[0] = 0.f;   // never true: keeps the dynamic LDS allocation
    const unsigned own = (unsigned)a2 * g.S2 + (unsigned)a1 * g.S1 + (unsigned)a0 + 31u;
    const int nbytes = (int)(g.elems * (unsigned)B);
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(src), 0, nbytes, 0x00020000);
    __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dst, 0, nbytes, 0x00020000);
    unsigned v[27];
#pragma unroll
    for (int l = 0; l < 27; ++l) {
        const int c0 = l % 3 - 1, c1 = (l / 3) % 3 - 1, c2 = l / 9 - 1;
        const int off = -(c2 * (int)g.S2 + c1 * (int)g.S1 + c0);
        if constexpr (B == 4) v[l] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)((own + off) * 4u), (int)(l * g.PL * 4u), 0);
        else v[l] = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)((own + off) * 2u), (int)(l * g.PL * 2u), 0);
    }
#pragma unroll
    for (int l = 0; l < 27; ++l) {
        if constexpr (B == 4) __builtin_amdgcn_raw_buffer_store_b32(v[l], rd, (int)(own * 4u), (int)(l * g.PL * 4u), 2);
        else __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v[l], rd, (int)(own * 2u), (int)(l * g.PL * 2u), 2);
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
    const int mode = argc > 4 ? atoi(argv[4]) : 0;
    // modes 4 and 5 cover the data with one grid of one-shot threads
    const size_t ncell = bytes / (27 * sizeof(float));
    // modes 6, 7, 8: planes with 1, 2, 4 cells per thread; argument 3 is then the dynamic LDS
    // bytes per block (occupancy cap)
    const int W = mode == 7 ? 2 : mode == 8 ? 4 : 1;
    const size_t nrow = ncell / 256 * 256;   // rows_kernel: whole rows only
    // modes 6-8 and 11-14: argument 3 is dynamic LDS bytes per block (occupancy cap)
    const int lds = mode >= 6 && argc > 3 ? atoi(argv[3]) : 0;
    const int blocks = mode == 4 ? (int)((n + 255) / 256)
                       : mode == 5 ? (int)((ncell + 255) / 256)
                       : mode >= 9 ? (int)(nrow / 256)
                       : mode >= 6 ? (int)((ncell / W + 255) / 256)
                                   : (argc > 3 ? atoi(argv[3]) : prop.multiProcessorCount * 16);
    if (mode >= 11 && mode <= 14) {
        const int B = mode <= 12 ? 4 : 2;
        // lattice_copy_kernel over two buffers of its own (the first argument is ignored)
        LatShape g;
        g.n0 = 192; g.n1 = 384; g.n2 = 128;
        const unsigned rowl = 256, nrows = (unsigned)(g.n1 + 2) * (g.n2 + 2);
        g.elems = 27u * rowl * nrows + 256u;
        if (mode == 11 || mode == 13) { g.PL = rowl * nrows; g.S1 = rowl; g.S2 = rowl * (g.n1 + 2); }
        else { g.PL = rowl; g.S1 = 27u * rowl; g.S2 = 27u * rowl * (g.n1 + 2); }
        float *la = nullptr, *lb = nullptr;
        CHK(hipMalloc(&la, (size_t)g.elems * B));
        CHK(hipMalloc(&lb, (size_t)g.elems * B));
        CHK(hipMemset(la, 0, (size_t)g.elems * B));
        CHK(hipMemset(lb, 0, (size_t)g.elems * B));
        auto lat = [&](const float* s_, float* d_) {
            if (B == 4) hipLaunchKernelGGL(lattice_copy_kernel<4>, dim3(1, g.n1, g.n2), dim3(192), lds, 0, s_, d_, g);
            else hipLaunchKernelGGL(lattice_copy_kernel<2>, dim3(1, g.n1, g.n2), dim3(192), lds, 0, s_, d_, g);
        };
        hipEvent_t e0, e1;
        CHK(hipEventCreate(&e0));
        CHK(hipEventCreate(&e1));
        lat(la, lb);
        CHK(hipDeviceSynchronize());
        CHK(hipEventRecord(e0, 0));
        for (int it = 0; it < iters; ++it) lat((it & 1) ? lb : la, (it & 1) ? la : lb);
        CHK(hipEventRecord(e1, 0));
        CHK(hipEventSynchronize(e1));
        float ms = 0;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / iters, moved = 2.0 * 27 * B * (double)g.n0 * g.n1 * g.n2;
        const char* ln[4] = {"lattice_copy_planes_b32", "lattice_copy_rows_b32", "lattice_copy_planes_b16",
                             "lattice_copy_rows_b16"};
        printf("{\"kernel\": \"%s\", \"bytes_moved\": %.0f, \"ms_per_launch\": %.5f, \"GBps\": %.1f}\n",
               ln[mode - 11], moved, per, moved / (per * 1e-3) / 1e9);
        CHK(hipFree(la));
        CHK(hipFree(lb));
        return 0;
    }
    const char* names[11] = {"copy_kernel", "copy_nt_kernel", "read_kernel", "write_kernel", "copy_once_kernel",
                             "planes_kernel", "planes_vec1", "planes_vec2", "planes_vec4", "rows_256", "rows_64"};
    const double rd = (mode == 3) ? 0.0 : 1.0, wr = (mode == 2) ? 0.0 : 1.0;
    const double moved = mode >= 9 ? (double)nrow * 27 * sizeof(float)
                         : mode >= 5 ? (double)(ncell / W * W) * 27 * sizeof(float) : (double)bytes;
    auto launch = [&]() {
        const float* fa = reinterpret_cast<const float*>(a);
        float* fb = reinterpret_cast<float*>(b);
        if (mode == 9) hipLaunchKernelGGL(rows_kernel<256>, dim3(blocks), dim3(256), 0, 0, fa, fb, nrow);
        else if (mode == 10) hipLaunchKernelGGL(rows_kernel<64>, dim3(blocks), dim3(256), 0, 0, fa, fb, nrow);
        else if (mode == 6) hipLaunchKernelGGL(planes_vec_kernel<1>, dim3(blocks), dim3(256), lds, 0, fa, fb, ncell);
        else if (mode == 7) hipLaunchKernelGGL(planes_vec_kernel<2>, dim3(blocks), dim3(256), lds, 0, fa, fb, ncell);
        else if (mode == 8) hipLaunchKernelGGL(planes_vec_kernel<4>, dim3(blocks), dim3(256), lds, 0, fa, fb, ncell);
        else if (mode == 5)
            hipLaunchKernelGGL(planes_kernel, dim3(blocks), dim3(256), 0, 0, reinterpret_cast<const float*>(a),
                               reinterpret_cast<float*>(b), ncell);
        else if (mode == 4)
            hipLaunchKernelGGL(copy_once_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
        else {
            void (*kern)(const float4*, float4*, size_t) =
                mode == 1 ? copy_nt_kernel : mode == 2 ? read_kernel : mode == 3 ? write_kernel : copy_kernel;
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, a, b, n);
        }
    };
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    launch();
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; ++it) launch();
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / iters;
    printf("{\"kernel\": \"%s\", \"bytes_read\": %.0f, \"bytes_written\": %.0f, \"ms_per_launch\": %.5f, "
           "\"GBps\": %.1f, \"blocks\": %d, \"cus\": %d}\n", names[mode], rd * moved, wr * moved, per,
           (rd + wr) * moved / (per * 1e-3) / 1e9, blocks, prop.multiProcessorCount);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
