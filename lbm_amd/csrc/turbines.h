// Actuator-line turbines on the device: the producer of the external forcing that
// forcings_apply consumes (src/main.F90:291-306), replacing turbine_forcing
// (src/m_turbine_forcing.F90:8-140) and the routines it calls.
//
// Per step, before the step kernel, two launches on the context's stream:
//   turbine_point_kernel   one block per actuator point (rotor, blade, section): the azimuth
//                          advanced by this step, the point's position, the moments of the step's
//                          input interpolated there, angle of attack, lift/drag table lookup and the
//                          blade-section force; the block also sums the point's Gaussian stencil
//                          weights (the normalisation of the deposit);
//   turbine_deposit_kernel one thread per cell of the rotors' deposit box: the forces of every point
//                          whose stencil covers the cell, gathered in point order (the order of the
//                          reference's scatter, so each cell sums the same terms in the same order,
//                          with no atomics); block 0 also stores the advanced azimuths.
// The step kernel reads the force planes inside that box only (StepArgs::flo/fhi).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "codec.h"
#include "lattice.h"
#include "impl.h"
#include "layout.h"

namespace lbm {

template <typename R>
struct TurbArgs {
    Layout lay;
    const void* src;      // the step's input buffer (as StepArgs::src)
    int stream;           // 1: the input is P(n-1), moments of its pull-streamed populations; 0: as stored
    const R* ext[4];      // moments given by an equilibrium start (first step only), else nullptr
    int nturb, nblades, nchords, nang_max, krad;
    R sigma2;             // 2 sigma^2 of the Gaussian stencil
    const R* rot;         // [nturb][kTurbRot]
    R* theta;             // [nturb] rotor azimuths [rad] of the state's step
    const R* blade;       // [nchords][4]: relm, dc, chord [cells], twist [deg]
    const int* nfoil;     // [nchords] airfoil table, 0-based
    const int* nang;      // [nfoils] rows of each table
    const R* foil;        // [nfoils][nang_max][3]: angle of attack [deg], cl, cd
    R* pts;               // [points][kTurbPt]
    const int* tbox;      // [nturb][6]: i, j, k bounds (lo, hi) of each rotor's deposit box
    int lo[3], hi[3];     // their union (the launch's box)
    R* F[3];              // force planes (per-cell layout)
};

template <typename R>
__device__ __forceinline__ R t_cos(R x) { if constexpr (std::is_same<R, float>::value) return cosf(x); else return cos(x); }
template <typename R>
__device__ __forceinline__ R t_sin(R x) { if constexpr (std::is_same<R, float>::value) return sinf(x); else return sin(x); }
template <typename R>
__device__ __forceinline__ R t_exp(R x) { if constexpr (std::is_same<R, float>::value) return expf(x); else return exp(x); }
template <typename R>
__device__ __forceinline__ R t_floor(R x) { if constexpr (std::is_same<R, float>::value) return floorf(x); else return floor(x); }
template <typename R>
__device__ __forceinline__ R t_atan2(R y, R x) { if constexpr (std::is_same<R, float>::value) return atan2f(y, x); else return atan2(y, x); }

// macrovars of one interior cell (src/m_macrovars_kernel.F90:35-49): rho = sum f, u = (sum c f) / rho
// as a product with the reciprocal, of the populations the step collides (pulled when stream).
template <int Q, class C>
__device__ __forceinline__ void turbine_cell_moments(const TurbArgs<typename C::real>& a, int i, int j, int k,
                                                     typename C::real m[4]) {
    using R = typename C::real;
    using Lt = Lattice<Q>;
    const Layout& L = a.lay;
    if (a.ext[0] != nullptr) {
        const int o = L.cidx(i, j, k);
#pragma unroll
        for (int q = 0; q < 4; ++q) m[q] = a.ext[q][o];
        return;
    }
    const typename C::raw* f = static_cast<const typename C::raw*>(a.src);
    const int s = a.stream;
    R r = 0, u = 0, v = 0, w = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const R x = C::template decode<Q>(f[(size_t)l * L.plane + L.cell_w(i - s * Lt::cx(l), j - s * Lt::cy(l), k - s * Lt::cz(l))], l);
        r = r + x;
        u = u + R(Lt::cx(l)) * x;
        v = v + R(Lt::cy(l)) * x;
        w = w + R(Lt::cz(l)) * x;
    }
    const R inv = R(1) / r;
    m[0] = r;
    m[1] = u * inv;
    m[2] = v * inv;
    m[3] = w * inv;
}

// src/m_turbine_interpolation.F90:29-33, terms and products in that order
template <typename R>
__device__ __forceinline__ R turbine_trilinear(const R c[8], R fx, R fy, R fz) {
    const R one = R(1);
    return c[0] * (one - fx) * (one - fy) * (one - fz) + c[1] * fx * (one - fy) * (one - fz) +
           c[2] * (one - fx) * fy * (one - fz) + c[3] * fx * fy * (one - fz) +
           c[4] * (one - fx) * (one - fy) * fz + c[5] * fx * (one - fy) * fz +
           c[6] * (one - fx) * fy * fz + c[7] * fx * fy * fz;
}

template <int Q, class C>
__global__ void __launch_bounds__(256) turbine_point_kernel(const TurbArgs<typename C::real> a) {
    using R = typename C::real;
    __shared__ R corner[4][8];
    __shared__ R red[256];
    const int p = blockIdx.x, tid = threadIdx.x;
    const int per = a.nblades * a.nchords;
    const int t = p / per, b = (p - t * per) / a.nchords, c = p - t * per - b * a.nchords;
    const R* rot = a.rot + t * kTurbRot;
    const R pi = R(3.14159265358979323846);
    const R pi2 = R(2) * pi;
    const R omegand = rot[12];
    // azimuth advanced by this step (src/m_turbine_forcing.F90:72), blade b's offset
    // (src/m_turbine_distribute_points.F90:38)
    const R th = (a.theta[t] + omegand) + R(b) * pi2 / R(a.nblades);
    const R ct = t_cos(th), st = t_sin(th);
    const R* bl = a.blade + 4 * c;
    const R relm = bl[0];
    const Layout& L = a.lay;
    const int n[3] = {L.nx, L.ny, L.nz};
    R x[3], fr[3];
    int g[3], cl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = rot[d] + relm * (ct * rot[6 + d] + st * rot[9 + d]);
        g[d] = (int)t_floor(x[d]);
        fr[d] = x[d] - R(g[d]);
        cl[d] = g[d] < 1 ? 1 : (g[d] > n[d] - 1 ? n[d] - 1 : g[d]);
    }
    // corners c000, c100, c010, c110, c001, c101, c011, c111 of the clamped cell, one per thread
    if (tid < 8) {
        R m[4];
        turbine_cell_moments<Q, C>(a, cl[0] + (tid & 1), cl[1] + ((tid >> 1) & 1), cl[2] + (tid >> 2), m);
#pragma unroll
        for (int q = 0; q < 4; ++q) corner[q][tid] = m[q];
    }
    // sum of the stencil's weights (src/m_turbine_deposit.F90:51-60) over the clipped box
    int lo[3], ext[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = g[d] - a.krad < 1 ? 1 : g[d] - a.krad;
        const int hi = g[d] + a.krad > n[d] ? n[d] : g[d] + a.krad;
        ext[d] = hi - lo[d] + 1;
    }
    const int ntot = ext[0] > 0 && ext[1] > 0 && ext[2] > 0 ? ext[0] * ext[1] * ext[2] : 0;
    R s = 0;
    for (int e = tid; e < ntot; e += blockDim.x) {
        const int ii = lo[0] + e % ext[0], jj = lo[1] + (e / ext[0]) % ext[1], kk = lo[2] + e / (ext[0] * ext[1]);
        const R dx = R(ii) - x[0], dy = R(jj) - x[1], dz = R(kk) - x[2];
        s += t_exp(-(dx * dx + dy * dy + dz * dz) / a.sigma2);
    }
    red[tid] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid != 0) return;
    const R dens = turbine_trilinear(corner[0], fr[0], fr[1], fr[2]);
    const R ux = turbine_trilinear(corner[1], fr[0], fr[1], fr[2]);
    const R uy = turbine_trilinear(corner[2], fr[0], fr[1], fr[2]);
    const R uz = turbine_trilinear(corner[3], fr[0], fr[1], fr[2]);
    // src/m_turbine_point_forces_kernel.F90:57-66
    const R utheta = omegand * relm - uz * ct - uy * st;
    const R phi = t_atan2(ux, utheta);
    const R ang = phi * R(180) / pi - bl[3] - rot[13];
    // src/m_nrelliftdrag.F90:17-33: linear in the first interval deg(i) < ang <= deg(i+1)
    R clift = 0, cdrag = 0;
    const int kf = a.nfoil[c];
    const R* tab = a.foil + (size_t)kf * a.nang_max * 3;
    for (int i = 0; i + 1 < a.nang[kf]; ++i) {
        if (tab[3 * i] < ang && ang <= tab[3 * i + 3]) {
            R slope = (tab[3 * i + 4] - tab[3 * i + 1]) / (tab[3 * i + 3] - tab[3 * i]);
            clift = tab[3 * i + 1] + slope * (ang - tab[3 * i]);
            slope = (tab[3 * i + 5] - tab[3 * i + 2]) / (tab[3 * i + 3] - tab[3 * i]);
            cdrag = tab[3 * i + 2] + slope * (ang - tab[3 * i]);
            break;
        }
    }
    // src/m_turbine_compute_bladeforce.F90:41-64 (tip-loss factor 1)
    const R speed2 = ux * ux + utheta * utheta;
    const R area = bl[2] * bl[1];
    const R lift = R(0.5) * dens * speed2 * clift * area * R(1);
    const R drag = R(0.5) * dens * speed2 * cdrag * area * R(1);
    const R sphi = t_sin(phi), cphi = t_cos(phi);
    R* out = a.pts + (size_t)p * kTurbPt;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const R ea = rot[3 + d];
        const R et = -st * rot[6 + d] + ct * rot[9 + d];
        out[3 + d] = lift * (cphi * ea + sphi * et) - drag * (sphi * ea - cphi * et);
        out[d] = x[d];
    }
    out[6] = red[0];
    out[7] = R(0);
}

// One thread per cell of the union box, threads along the contiguous storage axis.
template <typename R>
__global__ void __launch_bounds__(256) turbine_deposit_kernel(const TurbArgs<R> a) {
    if (blockIdx.x == 0)
        for (int t = threadIdx.x; t < a.nturb; t += blockDim.x) a.theta[t] = a.theta[t] + a.rot[t * kTurbRot + 12];
    const Layout& L = a.lay;
    int e[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) e[d] = a.hi[d] - a.lo[d] + 1;
    const int ax[3] = {L.ax0, L.ax1, L.ax2};
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n0 = e[ax[0]], n1 = e[ax[1]];
    if (idx >= n0 * n1 * e[ax[2]]) return;
    int x[3];
    x[ax[0]] = a.lo[ax[0]] + (int)(idx % n0);
    x[ax[1]] = a.lo[ax[1]] + (int)((idx / n0) % n1);
    x[ax[2]] = a.lo[ax[2]] + (int)(idx / (n0 * n1));
    const int per = a.nblades * a.nchords;
    R f0 = 0, f1 = 0, f2 = 0;
    for (int t = 0; t < a.nturb; ++t) {
        const int* bx = a.tbox + 6 * t;
        if (x[0] < bx[0] || x[0] > bx[1] || x[1] < bx[2] || x[1] > bx[3] || x[2] < bx[4] || x[2] > bx[5]) continue;
        for (int p = t * per; p < (t + 1) * per; ++p) {
            const R* P = a.pts + (size_t)p * kTurbPt;
            // src/m_turbine_deposit.F90:45-46 and :62
            if (P[3] == R(0) && P[4] == R(0) && P[5] == R(0)) continue;
            if (P[6] < R(1.0e-6)) continue;
            bool in = true;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int g = (int)t_floor(P[d]);
                if (x[d] < g - a.krad || x[d] > g + a.krad) in = false;
            }
            if (!in) continue;
            const R dx = R(x[0]) - P[0], dy = R(x[1]) - P[1], dz = R(x[2]) - P[2];
            const R w = t_exp(-(dx * dx + dy * dy + dz * dz) / a.sigma2) / P[6];
            f0 = f0 + w * P[3];
            f1 = f1 + w * P[4];
            f2 = f2 + w * P[5];
        }
    }
    const int o = L.cidx(x[0], x[1], x[2]);
    a.F[0][o] = f0;
    a.F[1][o] = f1;
    a.F[2][o] = f2;
}

}  // namespace lbm
