// Launch wrappers, one set per (lattice, storage codec).  Instantiated in kernels_<codec>.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace lbm {

struct EdgeSpec;

struct EdgeList {
    int n;
    int len[20];
    int kind[20];
    int t[20];
    int g[20][3];
    int out[20][3];
    int mir[20][3];
};

// Ghost cells that a closed-face pass reads while they hold a copy made by the reference's first
// periodic pass (src/m_boundarycond.F90:69-159) of a cell that an earlier closed pass of the same
// call has since rewritten.  g holds the ghost coordinates as read; the snapshot of the copied
// cell is taken before the closed passes.
struct SnapList {
    int n;
    int g[8][3];
};

// Everything the host passes to one step-kernel launch, in double; converted to the compute
// type by the wrapper.
struct StepHost {
    const void* src;
    void* dst;
    int stream;
    double fac0, kinevisc, vconst;
    const void* ablvisc;
    const void* ext[4];
    const uint32_t* solid;
    int solid_wpr;
    int ip;
    const void* turb[3];
    double ampl;
    const void* F[3];
    void* tau_out;
    int fold_dir, fold_opt_lo, fold_opt_hi;   // closed face performed by the step kernel
    int bgk3, hrr, vre;
    int variant;      // step-kernel variant: bit 0 non-temporal stores, bit 2 two cells per lane
    int block_x, block_y;
};

struct Ops {
    hipError_t (*step)(const Layout&, const StepHost&, hipStream_t);
    hipError_t (*closed)(const Layout&, void* f1, const void* f2, int dir, int opt_lo, int opt_hi, const void* snap,
                         const SnapList&, hipStream_t);
    hipError_t (*snapshot)(const Layout&, const void* f1, void* snap, const SnapList&, hipStream_t);
    hipError_t (*edges)(const Layout&, void* f1, const void* f2, const EdgeList&, hipStream_t);
    hipError_t (*inflow)(const Layout&, void* f1, const void* shear, double utime, double cosd, double sind,
                         double rho0, hipStream_t);
    hipError_t (*aos_in)(const Layout&, const void* aos, int realkind, void* dstA, void* dstB, hipStream_t);
    hipError_t (*state_out)(const Layout&, const void* cur, const void* gsrc, int stream_interior, int ghost_raw,
                            void* aos, void* rho, void* u, void* v, void* w, int realkind, hipStream_t);
    hipError_t (*feq_init)(const Layout&, int bgk3, const void* rho, const void* u, const void* v, const void* w,
                           int realkind, void* dst, void* mr, void* mu, void* mv, void* mw, hipStream_t);
    hipError_t (*totals)(const Layout&, const void* cur, int stream_interior, double* partial, int nblocks,
                         hipStream_t);
    int real_bytes;   // compute type size
    int raw_bytes;    // storage size
};

// Defined in the per-codec translation units.
const Ops* ops_f32(int q);
const Ops* ops_f64(int q);
const Ops* ops_f16s(int q);
const Ops* ops_bf16s(int q);

}  // namespace lbm
