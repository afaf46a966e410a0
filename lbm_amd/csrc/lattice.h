// Compile-time D3Q19 / D3Q27 lattices.
//
// Direction order, weights and opposite pairs follow src/mod_D3Q27setup.F90:28-48; the order is
// part of the restart-file layout f(nl,0:nx+1,0:ny+1,0:nz+1), so it is kept.  D3Q19 is the
// first 19 directions of the D3Q27 list.  Every per-direction quantity here is a constexpr
// function of the direction index so that fully unrolled loops fold it into immediates.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

// Each velocity component of the 27 directions, packed two bits per direction (c + 1) into one
// 64-bit constant: a lookup with a run-time (wave-uniform) index is then a scalar shift and mask,
// not a load from a constant table (the rolled direction loops of the adjoint kernel).
namespace detail {
struct Dir27 {
    int t[27];
};
constexpr unsigned long long pack2(const Dir27& d) {
    unsigned long long p = 0;
    for (int l = 0; l < 27; ++l) p |= (unsigned long long)(d.t[l] + 1) << (2 * l);
    return p;
}
constexpr unsigned long long kCx = pack2({{0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1}});
constexpr unsigned long long kCy = pack2({{0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1}});
constexpr unsigned long long kCz = pack2({{0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1}});
}  // namespace detail

__host__ __device__ constexpr int cx27(int l) { return (int)((detail::kCx >> (2 * l)) & 3ull) - 1; }
__host__ __device__ constexpr int cy27(int l) { return (int)((detail::kCy >> (2 * l)) & 3ull) - 1; }
__host__ __device__ constexpr int cz27(int l) { return (int)((detail::kCz >> (2 * l)) & 3ull) - 1; }

template <int Q>
struct Lattice {
    static_assert(Q == 19 || Q == 27, "D3Q19 or D3Q27");
    static constexpr int q = Q;
    __host__ __device__ static constexpr int cx(int l) { return cx27(l); }
    __host__ __device__ static constexpr int cy(int l) { return cy27(l); }
    __host__ __device__ static constexpr int cz(int l) { return cz27(l); }
    __host__ __device__ static constexpr int c(int l, int d) { return d == 0 ? cx27(l) : d == 1 ? cy27(l) : cz27(l); }
    __host__ __device__ static constexpr int norm2(int l) { return cx(l) * cx(l) + cy(l) * cy(l) + cz(l) * cz(l); }
    // Weights as exact rationals; w_num(l)/w_den(l).
    __host__ __device__ static constexpr double w(int l) {
        return Q == 19 ? (norm2(l) == 0 ? 1.0 / 3.0 : norm2(l) == 1 ? 1.0 / 18.0 : 1.0 / 36.0)
                       : (norm2(l) == 0 ? 8.0 / 27.0 : norm2(l) == 1 ? 2.0 / 27.0 : norm2(l) == 2 ? 1.0 / 54.0 : 1.0 / 216.0);
    }
    // Weight rounded as the reference's default real rounds it (1.0/3.0 evaluated in FP32).
    __host__ __device__ static constexpr float wf(int l) {
        return Q == 19 ? (norm2(l) == 0 ? 1.0f / 3.0f : norm2(l) == 1 ? 1.0f / 18.0f : 1.0f / 36.0f)
                       : (norm2(l) == 0 ? 8.0f / 27.0f : norm2(l) == 1 ? 2.0f / 27.0f : norm2(l) == 2 ? 1.0f / 54.0f : 1.0f / 216.0f);
    }
    // Direction with velocity (x,y,z), or -1.
    __host__ __device__ static constexpr int find(int x, int y, int z) {
        for (int m = 0; m < Q; ++m)
            if (cx(m) == x && cy(m) == y && cz(m) == z) return m;
        return -1;
    }
    // Opposite directions are the adjacent pairs (1,2), (3,4), ... (0-based), checked below.
    __host__ __device__ static constexpr int opposite(int l) { return l == 0 ? 0 : ((l & 1) ? l + 1 : l - 1); }
    static constexpr bool pairs_ok() {
        for (int l = 0; l < Q; ++l)
            if (find(-cx(l), -cy(l), -cz(l)) != opposite(l)) return false;
        return true;
    }
};
static_assert(Lattice<19>::pairs_ok() && Lattice<27>::pairs_ok(), "opposite directions are adjacent pairs");
static_assert(cx27(0) == 0 && cx27(2) == -1 && cy27(3) == 1 && cz27(26) == 1 && cy27(26) == -1, "packed tables");

// Reals of the compute type in the reference's rounding: w evaluated in that precision.
template <int Q, typename R>
__host__ __device__ constexpr R weight(int l) {
    if constexpr (sizeof(R) == 4) return Lattice<Q>::wf(l);
    else return (R)Lattice<Q>::w(l);
}

}  // namespace lbm
