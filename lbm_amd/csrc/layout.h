// Device layout of one population buffer.
//
// Cells of the padded grid (0:nx+1, 0:ny+1, 0:nz+1) lie in rows along one axis (the "fast" axis);
// rows are padded so that interior index 1 starts on a 128-byte boundary (storage column of index
// x is xoff + x with xoff = 128/sizeof(raw) - 1) and the row pitch is a multiple of 128 bytes.
// The two other axes follow in natural order.  A 64-lane wave along the fast axis reads and
// writes whole cache lines for populations with no velocity along that axis.
//
// Two arrangements of the Q populations, both addressed as  l * plane + cell(i, j, k):
//   row-blocked (the default; needs element offsets below 2^31 and one slab of the slowest axis
//     below 4 GiB): [row][population][pitch] - plane is the row pitch and the row strides are Q
//     times larger; the Q segments of a row are adjacent;
//   planes: [population][row][pitch] - plane is the whole padded grid.
// Per-cell arrays (tau, moments, forcing, averages) always use the plane arrangement of one
// population: cidx(i, j, k), cplane elements.
//
// The fast axis is chosen per domain (make_layout): a periodic axis if there is one (i first),
// so that the ghost planes the boundary kernels touch (closed faces, the inflow/outflow planes
// of i) are rows along the fast axis and their accesses coalesce.
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
    int per_i, per_j, per_k;
    int ax0, ax1, ax2;      // storage order: ax0 contiguous, ax2 slowest (0 = i, 1 = j, 2 = k)
    int si, sj, sk;         // element strides of i, j, k in a population buffer
    int base;               // storage offset of cell (0,0,0)
    int pitch;              // elements per row along ax0
    long long plane;        // element stride between populations (row pitch when row-blocked)
    long long total;        // elements per population buffer (all Q populations)
    int ci, cj, ck;         // element strides of i, j, k in per-cell arrays
    long long cplane;       // elements per per-cell array (the padded grid, rounded to 256 B)
    int rowblock;           // 1: row-blocked buffers

    __host__ __device__ inline int n(int a) const { return a == 0 ? nx : a == 1 ? ny : nz; }
    __host__ __device__ inline int cell(int i, int j, int k) const { return base + i * si + j * sj + k * sk; }
    // index of cell (i, j, k) in a per-cell array
    __host__ __device__ inline int cidx(int i, int j, int k) const { return base + i * ci + j * cj + k * ck; }
    __host__ __device__ static inline int wrap(int x, int n, int per) {
        return per ? (x == 0 ? n : (x == n + 1 ? 1 : x)) : x;
    }
    // Cell offset with periodic directions folded onto the interior.
    __host__ __device__ inline int cell_w(int i, int j, int k) const {
        return cell(wrap(i, nx, per_i), wrap(j, ny, per_j), wrap(k, nz, per_k));
    }
    // Map coordinates given in storage-axis order (a along ax0, b along ax1, c along ax2) to i,j,k.
    __host__ __device__ inline void ijk(int a, int b, int c, int& i, int& j, int& k) const {
        i = ax0 == 0 ? a : (ax1 == 0 ? b : c);
        j = ax0 == 1 ? a : (ax1 == 1 ? b : c);
        k = ax0 == 2 ? a : (ax1 == 2 ? b : c);
    }
};

// Split steps along the halo axis, e edge rows at each end (n >= 2e + 1 cells): launch coordinate
// c (0-based) of part 1, which covers 2e launch rows (cells 1..e and n-e+1..n), or of part 2, which
// covers n - 2e (cells e+1 .. n-e) -> the 0-based coordinate of its cell, or n (cell n + 1, which
// the step kernel skips) past the part's rows.
__host__ __device__ inline int part_row(int c, int part, int n, int e) {
    return part == 1 ? (c < e ? c : c < 2 * e ? n - 2 * e + c : n) : (c < n - 2 * e ? c + e : n);
}

// fast < 0: choose (first periodic axis, else i).  q > 0 with rowblock: row-blocked buffers when every
// element offset fits 31 bits and one slab of the slowest axis (Q rows per ax1 index) stays below
// 4 GiB (the step kernel's 32-bit offsets within a slab), planes otherwise.
inline Layout make_layout(int nx, int ny, int nz, int raw_bytes, int per_i, int per_j, int per_k, int fast = -1,
                          int q = 0, bool rowblock = false) {
    Layout L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.per_i = per_i; L.per_j = per_j; L.per_k = per_k;
    if (fast < 0) fast = per_i ? 0 : per_j ? 1 : per_k ? 2 : 0;
    L.ax0 = fast;
    L.ax1 = fast == 0 ? 1 : 0;
    L.ax2 = fast == 2 ? 1 : 2;
    const int a = 128 / raw_bytes;           // elements per 128 B
    const int n0 = L.n(L.ax0), n1 = L.n(L.ax1), n2 = L.n(L.ax2);
    L.pitch = ((a + n0 + 1) + a - 1) / a * a;
    int st[3];
    st[L.ax0] = 1;
    st[L.ax1] = L.pitch;
    st[L.ax2] = L.pitch * (n1 + 2);
    L.ci = st[0]; L.cj = st[1]; L.ck = st[2];
    L.base = a - 1;
    const long long pl = (long long)L.pitch * (n1 + 2) * (n2 + 2);
    const long long ea = 256 / raw_bytes;
    L.cplane = (pl + ea - 1) / ea * ea;
    const int nq = q > 0 ? q : 1;
    L.rowblock = rowblock && q > 0 && (double)nq * (double)L.cplane < 2147483648.0 &&
                 (double)nq * L.pitch * (n1 + 2) * raw_bytes < 4294967296.0;
    if (L.rowblock) {
        st[L.ax1] = nq * L.pitch;
        st[L.ax2] = nq * L.pitch * (n1 + 2);
        L.plane = L.pitch;
        L.total = (nq * pl + ea - 1) / ea * ea;
    } else {
        L.plane = L.cplane;
        L.total = nq * L.cplane;
    }
    L.si = st[0]; L.sj = st[1]; L.sk = st[2];
    return L;
}

}  // namespace lbm
