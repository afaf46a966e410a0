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
    int flo[3], fhi[3];   // cells (i, j, k; inclusive) whose force is read; the others have none
    const void* theta;    // potential temperature (cell layout) of the buoyancy forcing, or nullptr
    double theta0, bscale;
    void* avg;            // averaging sums (6 cell planes of the compute type) or nullptr
    void* tau_out;
    int fold_dir, fold_opt_lo, fold_opt_hi;   // closed face performed by the step kernel
    int inflow_fold;                          // inflow/outflow planes performed by the step kernel
    const void* shear;                        // uvel_shear(nz) (compute type)
    double utime, cosd, sind, rho0;           // inflow values of the step
    int heavy_first;                          // bit 0 / 1: cells 1 and n of ax1 / ax2 first
    int halo_dir;                             // halo axis (-1: none)
    int part, part_axis, part_edge;           // split step along the halo axis (StepArgs::part)
    int bgk3, hrr, vre;
    int variant;      // step-kernel variant: bit 0 non-temporal stores, bit 1 the >= 4 GiB buffer form,
                      // bit 2 the one-base form with run-time axis mapping
    int block_x, block_y;
    int lds_bytes;    // dynamic LDS reserved per step-kernel block (caps the resident blocks per CU; 0: none)
};

// One ensemble member for the batched kernels: its step parameters (src, dst and the turbulence
// slice are chosen on the device per step) and the buffers.
struct MemberHost {
    Layout lay;
    StepHost h;
    void* buf[2];
    int cur0;
    int nrturb;
    const void* turb[3];   // slice 0 of uu, vv, ww (nullptr: off)
    const void* shear;     // uvel_shear(nz)
    double rho0;
    int inflow;            // inflow/outflow planes of i (folded into the batched step)
};

// One batched step: the step kernel over every member's tiles (faces and inflow planes folded),
// the clock tick.
struct EnsLaunch {
    const void* tab;                 // device MemberTab<R>[nmem]
    const int* start;                // device tile starts [nmem]
    int nmem, total, block, grid;
    const void* vals;                // device (utime, cos, sin)[steps of the table][nmem]
    int* clock;                      // device {it, steps since set, table's first step}
    int bgk3, hrr, vre;
};

// Actuator-line turbines (turbines.h): device tables of the compute type and the deposit boxes.
constexpr int kTurbRot = 16;   // per-rotor table: hub x,y,z | e_axis | e1 | e2 | omegand, pitch | pad
constexpr int kTurbPt = 8;     // per-point record: xg, yg, zg, F(3), sum of stencil weights, pad
struct TurbHost {
    int nturb, nblades, nchords, nang_max, npts, krad;
    double sigma2;
    const void* rot;      // [nturb][16] (compute type)
    void* theta;          // [nturb]
    const void* blade;    // [nchords][4]
    const int* nfoil;     // [nchords], 0-based
    const int* nang;      // [nfoils]
    const void* foil;     // [nfoils][nang_max][3]
    void* pts;            // [npts][8]
    const int* tbox;      // [nturb][6]
    int lo[3], hi[3];     // union of the rotors' deposit boxes (i, j, k; 1-based, inclusive)
    void* F[3];           // force planes
};

// Totals (totals_kernel): partial results per block, and the block count for a grid of n interior
// cells (one 256-thread block per 256 cells, at most kTotalsMaxBlocks, grid-stride beyond).
constexpr int kTotalsSlots = 6;
constexpr int kTotalsMaxBlocks = 1024;
inline int totals_blocks(long long ncell) {
    const long long b = (ncell + 255) / 256;
    return (int)(b < 1 ? 1 : (b > kTotalsMaxBlocks ? kTotalsMaxBlocks : b));
}

// Objective box (inclusive, 1-based interior indices i, j, k).
struct AdjBox {
    int lo[3], hi[3];
};
// Blocks of the objective kernel for a box; each owns two FP64 accumulator slots.
inline int objective_blocks(const AdjBox& b) {
    const long long nb = (long long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
    const long long g = (nb + 255) / 256;
    return (int)(g < 1024 ? g : 1024);
}

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
    hipError_t (*average)(const Layout&, const void* cur, int stream_interior, void* avg, hipStream_t);
    // ensembles
    size_t member_bytes;   // sizeof(MemberTab<R>)
    void (*ens_pack)(const MemberHost&, int block, void* out, int* tiles);
    hipError_t (*ens_step)(const EnsLaunch&, hipStream_t);
    // adjoint (nullptr for the 16-bit storage codecs)
    hipError_t (*adj_step)(const Layout&, const StepHost&, const void* lam_y, void* lam_x, const AdjBox&, double kappa,
                           int coll, double* gF, hipStream_t);
    hipError_t (*adj_inflow)(const Layout&, void* lam, double utime, double cosd, double sind, double rho0, double* gS,
                             hipStream_t);
    hipError_t (*adj_face)(const Layout&, void* lam_y, void* lam_x, int dir, int opt_lo, int opt_hi, hipStream_t);
    hipError_t (*objective)(const Layout&, const void* f, const AdjBox&, double* acc, hipStream_t);
    // halo planes of j (side 0: j = 1 / ghost 0, side 1: j = ny / ghost ny+1)
    hipError_t (*halo_pack)(const Layout&, const void* f, int side, void* buf, hipStream_t);
    hipError_t (*halo_unpack)(const Layout&, void* f, int side, const void* buf, int fold_dir, hipStream_t);
    hipError_t (*ghost_copy)(const Layout&, const void* src, void* dst, hipStream_t);
    // turbine forcing of one step from the step's input (src, stream as the step kernel's; ext: the
    // moments of an equilibrium start or nullptr)
    hipError_t (*turb_step)(const Layout&, const TurbHost&, const void* src, int stream, const void* const ext[4],
                            hipStream_t);
    int real_bytes;   // compute type size
    int raw_bytes;    // storage size
};

// Inflow-turbulence generation (turbulence.hip): scratch size and the generator writing the
// four fields uu, vv, ww, rr (each (nrturb+1) x nz x ny, j fastest) of the compute type (rb bytes).
size_t turbulence_scratch_bytes(int ny, int nz, int nrturb);
hipError_t turbulence_generate(int rb, int ny, int nz, int nrturb, uint64_t seed, void* scratch, void* out,
                               hipStream_t stream);

// Defined in the per-codec translation units.
const Ops* ops_f32(int q);
const Ops* ops_f64(int q);
const Ops* ops_f16s(int q);
const Ops* ops_bf16s(int q);
const Ops* ops_f64t(int q);

}  // namespace lbm
