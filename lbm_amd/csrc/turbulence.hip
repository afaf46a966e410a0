// Inflow-turbulence fields generated on the device (lbm_generate_turbulence): the reference's
// inflow_turbulence_compute (src/m_inflow_turbulence_compute.F90:3-62) over pseudo2D
// (src/m_pseudo2D.F90:3-167), for the four variables uu, vv, ww, rr and slices 0..nrturb:
//   1. spectral amplitudes  X(l,p) = e(l,p) sqrt(dk) c (cos phi, sin phi), l = 0..n1/2,
//      p = -n2/2..n2/2-1, with e = exp(-(a11 kappa^2 l^2 + a22 lambda^2 p^2)) and the mean mode
//      zeroed (lmean = .false.);  the phases phi = 2 pi U come from SplitMix64(seed) keyed by
//      (variable, slice, l, p) where the reference draws random_number;
//   2. inverse real 2-D DFT (unnormalized, as FFTW's c2r): a complex DFT along p, then the
//      real transform along l, y(i,m) = Re Z(0,m) + (-1)^i Re Z(n1/2,m)
//                                       + 2 sum_{l=1}^{n1/2-1} Re(Z(l,m) e^{2 pi i l i/n1});
//   3. for slices 1..nrturb: u = 0.98 u(prev) + sqrt(1 - 0.98^2) u, then the slice's mean is
//      removed and it is divided by its sample deviation sqrt(sum u^2 / (N-1)).
// The transforms are direct sums in FP64 (n1 x n2 <= a few hundred squared per slice).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "impl.h"

namespace lbm {
namespace {

__device__ inline double u01(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// X[((v * ns + s) * n2 + pp) * nh + l], pp = p mod n2, nh = n1/2 + 1
__global__ void spectrum_kernel(double2* X, int ns, int n1, int n2, uint64_t seed, double ak, double al, double amp) {
    const int nh = n1 / 2 + 1;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int pp = blockIdx.y;
    const int vs = blockIdx.z;   // v * ns + s
    if (l >= nh) return;
    const int p = pp < n2 / 2 ? pp : pp - n2;
    const uint64_t counter = ((uint64_t)vs * nh + l) * n2 + pp;
    const double phi = 2.0 * M_PI * u01(seed, counter);
    const double e = exp(-(ak * (double)l * l + al * (double)p * p));
    double2 x;
    x.x = e * cos(phi) * amp;
    x.y = e * sin(phi) * amp;
    if (l == 0 && p == 0) x.x = x.y = 0.0;
    X[((size_t)vs * n2 + pp) * nh + l] = x;
}

// Z(l, m) = sum_pp X(l, pp) e^{+2 pi i pp m / n2},  m = 0..nz-1
__global__ void dft_p_kernel(const double2* X, double2* Z, int n1, int n2, int nz) {
    const int nh = n1 / 2 + 1;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    const int vs = blockIdx.z;
    if (l >= nh) return;
    const double2* x = X + (size_t)vs * n2 * nh;
    double re = 0.0, im = 0.0;
    for (int pp = 0; pp < n2; ++pp) {
        double sn, cs;
        sincospi(2.0 * (double)((long long)pp * m % n2) / n2, &sn, &cs);
        const double2 a = x[(size_t)pp * nh + l];
        re += a.x * cs - a.y * sn;
        im += a.x * sn + a.y * cs;
    }
    Z[((size_t)vs * nz + m) * nh + l] = make_double2(re, im);
}

// y(i, m) for i = 0..ny-1 into out[(v, s, m, i)] in the compute type
template <typename R>
__global__ void dft_l_kernel(const double2* Z, R* out, int n1, int ny, int nz) {
    const int nh = n1 / 2 + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    const int vs = blockIdx.z;
    if (i >= ny) return;
    const double2* z = Z + ((size_t)vs * nz + m) * nh;
    double y = z[0].x + ((i & 1) ? -z[n1 / 2].x : z[n1 / 2].x);
    for (int l = 1; l < n1 / 2; ++l) {
        double sn, cs;
        sincospi(2.0 * (double)((long long)l * i % n1) / n1, &sn, &cs);
        y += 2.0 * (z[l].x * cs - z[l].y * sn);
    }
    out[((size_t)vs * nz + m) * ny + i] = (R)y;
}

// AR(1) in slices and per-slice normalization; one block per variable.
template <typename R>
__global__ void __launch_bounds__(1024) ar1_kernel(R* out, int ns, int n) {
    __shared__ double red[1024];
    R* f = out + (size_t)blockIdx.x * ns * n;
    const R a = R(0.98);
    const R b = sqrt(R(1) - a * a);
    const int t = threadIdx.x;
    auto reduce = [&](double v) {
        red[t] = v;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
        }
        const double r = red[0];
        __syncthreads();
        return r;
    };
    for (int s = 1; s < ns; ++s) {
        R* cur = f + (size_t)s * n;
        const R* prev = f + (size_t)(s - 1) * n;
        double sum = 0.0;
        for (int e = t; e < n; e += blockDim.x) {
            const R u = a * prev[e] + b * cur[e];
            cur[e] = u;
            sum += (double)u;
        }
        const R mean = (R)(reduce(sum) / n);
        double sq = 0.0;
        for (int e = t; e < n; e += blockDim.x) {
            const R u = cur[e] - mean;
            cur[e] = u;
            sq += (double)u * u;
        }
        const R dev = (R)sqrt(reduce(sq) / (n - 1));
        for (int e = t; e < n; e += blockDim.x) cur[e] = cur[e] / dev;
        __syncthreads();
    }
}

}  // namespace

// Scratch: X and Z need 4*ns*n2*(n1/2+1) and 4*ns*nz*(n1/2+1) complex doubles; out receives
// 4 * ns * nz * ny reals of rb bytes (variables uu, vv, ww, rr in turn).
size_t turbulence_scratch_bytes(int ny, int nz, int nrturb) {
    const int n1 = ny + (ny & 1), n2 = nz + (nz & 1);
    const size_t nh = n1 / 2 + 1, ns = nrturb + 1;
    return 4 * ns * nh * (size_t)(n2 + nz) * sizeof(double2);
}

hipError_t turbulence_generate(int rb, int ny, int nz, int nrturb, uint64_t seed, void* scratch, void* out,
                               hipStream_t st) {
    const int n1 = ny + (ny & 1), n2 = nz + (nz & 1);   // even FFT sizes (src/m_pseudo2D.F90:56-57)
    const int nh = n1 / 2 + 1, ns = nrturb + 1;
    // src/m_pseudo2D.F90:66-112 with dx = dy = 1, theta = 0, rx = ry = 10/sqrt(3)
    // (src/m_inflow_turbulence_compute.F90:19-20); r2 is formed from rx as the reference does
    const double pi2 = 2.0 * 3.141592653589;
    const double rx = 10.0 / std::sqrt(3.0);
    const double dk = pi2 * pi2 / ((double)n1 * n2);
    const double kappa = pi2 / n1, lambda = pi2 / n2;
    const double r1 = std::sqrt(8.0) / rx, r2 = std::sqrt(8.0) / rx;
    double summ = 0.0;
    for (int p = -n2 / 2 + 1; p <= n2 / 2; ++p)
        for (int l = -n1 / 2 + 1; l <= n1 / 2; ++l)
            summ += std::exp(-2.0 * (kappa * kappa * l * l / (r1 * r1) + lambda * lambda * p * p / (r2 * r2)));
    summ -= 1.0;   // mean mode not sampled
    const double c = std::sqrt(1.0 / (dk * summ));
    const double ak = kappa * kappa / (r1 * r1), al = lambda * lambda / (r2 * r2);
    double2* X = static_cast<double2*>(scratch);
    double2* Z = X + (size_t)4 * ns * n2 * nh;
    hipLaunchKernelGGL(spectrum_kernel, dim3((nh + 63) / 64, n2, 4 * ns), dim3(64), 0, st, X, ns, n1, n2, seed, ak, al,
                       std::sqrt(dk) * c);
    hipLaunchKernelGGL(dft_p_kernel, dim3((nh + 63) / 64, nz, 4 * ns), dim3(64), 0, st, X, Z, n1, n2, nz);
    if (rb == 8) {
        hipLaunchKernelGGL(dft_l_kernel<double>, dim3((ny + 63) / 64, nz, 4 * ns), dim3(64), 0, st, Z,
                           static_cast<double*>(out), n1, ny, nz);
        hipLaunchKernelGGL(ar1_kernel<double>, dim3(4), dim3(1024), 0, st, static_cast<double*>(out), ns, ny * nz);
    } else {
        hipLaunchKernelGGL(dft_l_kernel<float>, dim3((ny + 63) / 64, nz, 4 * ns), dim3(64), 0, st, Z,
                           static_cast<float*>(out), n1, ny, nz);
        hipLaunchKernelGGL(ar1_kernel<float>, dim3(4), dim3(1024), 0, st, static_cast<float*>(out), ns, ny * nz);
    }
    return hipGetLastError();
}

}  // namespace lbm
