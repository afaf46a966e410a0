// Device layout of one population buffer (structure of arrays).
//
// A buffer holds Q planes; plane l is the padded grid (0:nx+1, 0:ny+1, 0:nz+1) of population l
// with i contiguous.  Rows are padded so that interior cell i = 1 starts on a 128-byte
// boundary: storage column of i is xoff + i with xoff = 128/sizeof(raw) - 1, and the row pitch is
// a multiple of 128 bytes.  A 64-lane wave along i therefore reads and writes whole cache lines
// for the c_x = 0 populations and straddles one extra line for c_x = +-1.
//
// Ghost layers in a periodic direction are not maintained: every reader wraps that index onto
// the interior instead (ghost 0 -> n, n+1 -> 1), which is what the reference's periodic passes
// (src/m_boundarycond.F90:69-159 and :373-463) make the ghost hold when it is read.  Ghost layers
// in a closed or inflow/outflow direction are persistent state, written by the boundary kernels
// exactly as the reference writes them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

struct Layout {
    int nx, ny, nz;
    int xoff;          // storage column of ghost i = 0 is xoff, of cell i is xoff + i
    int pitch;         // elements per (j,k) row
    int ystride;       // elements per k slab = pitch*(ny+2)
    long long plane;   // elements per population plane (ystride*(nz+2), rounded to 256 B)
    int per_i, per_j, per_k;

    __host__ __device__ inline int cell(int i, int j, int k) const { return k * ystride + j * pitch + xoff + i; }
    __host__ __device__ static inline int wrap(int x, int n, int per) {
        return per ? (x == 0 ? n : (x == n + 1 ? 1 : x)) : x;
    }
    // Cell offset with periodic directions folded onto the interior.
    __host__ __device__ inline int cell_w(int i, int j, int k) const {
        return cell(wrap(i, nx, per_i), wrap(j, ny, per_j), wrap(k, nz, per_k));
    }
};

inline Layout make_layout(int nx, int ny, int nz, int raw_bytes, int per_i, int per_j, int per_k) {
    Layout L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    int a = 128 / raw_bytes;                 // elements per 128 B
    L.xoff = a - 1;
    L.pitch = ((a + nx + 1) + a - 1) / a * a;
    L.ystride = L.pitch * (ny + 2);
    long long pl = (long long)L.ystride * (nz + 2);
    long long ea = 256 / raw_bytes;
    L.plane = (pl + ea - 1) / ea * ea;
    L.per_i = per_i; L.per_j = per_j; L.per_k = per_k;
    return L;
}

}  // namespace lbm
