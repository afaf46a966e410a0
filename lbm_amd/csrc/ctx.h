// Internal: the context behind lbm_ctx* and the host helpers shared by the C-ABI translation
// units (api.cpp: single domains, ensemble.cpp: ensembles).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "impl.h"
#include "layout.h"

namespace lbm_host {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

// Tag in the first field of every handle, checked by each export before it reads anything else: a
// handle of the other kind (an lbm_ens* passed where an lbm_ctx* is expected, or the reverse) or a
// destroyed one (the tag is cleared on destroy) fails with LBM_EINVAL instead of being used.
constexpr uint32_t kCtxTag = 0x4c424d43u;   // "CMBL"
constexpr uint32_t kEnsTag = 0x4c424d45u;   // "EMBL"

}  // namespace lbm_host

struct lbm_ctx {
    uint32_t tag = lbm_host::kCtxTag;
    lbm_desc d{};
    lbm::Layout lay{};
    const lbm::Ops* ops = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    // population buffers
    void* buf[2] = {nullptr, nullptr};
    size_t buf_bytes = 0;
    int cur = 0;
    bool have_state = false;
    bool pending = false;          // buf[cur] holds an uncollided state (next step collides in place)
    bool ghost_raw[2] = {true, true};
    bool use_ext = false;          // first collision uses the moments given at initialisation
    // inputs
    lbm_host::DevBuf ablvisc, shear, solid, turb[3], force[3], tau, ext[4], stage, partial, snap;
    lbm::SnapList snap_k{}, snap_j{};
    int solid_wpr = 0;
    bool has_solid = false, has_turb = false, has_force = false;
    int nrturb = 0;
    double ampl = 0.0;
    double uini = 0.0, udir = 0.0;
    int series_lo = 0, series_hi = -1;
    std::vector<double> ut_series, ud_series;
    double tau0 = 0.0, fac0 = 0.0;
    double obj_tangent = 0.0;      // directional derivative of the last objective (tangent storage)
    // actuator-line turbines (lbm_set_turbines): their forcing is written into force[] each step
    bool turb_on = false;
    lbm::TurbHost th{};
    lbm_host::DevBuf t_rot, t_theta, t_blade, t_nfoil, t_nang, t_foil, t_pts, t_box;
    // adjoint workspace (lbm_adjoint_run): checkpoints, recomputed states, adjoint buffers, the final
    // state, gradient and objective accumulators; kept between runs, grown when a run needs more
    lbm_host::DevBuf adj_ck, adj_snaps, adj_lam, adj_gF, adj_gS, adj_acc;
    // buoyancy forcing (lbm_set_buoyancy): potential temperature, reference temperature, scaling
    lbm_host::DevBuf theta;
    bool buoy_on = false;
    double btheta0 = 0.0, bscale = 0.0;
    // time averaging: sums of u, v, w, u^2, v^2, w^2 (6 cell planes of the compute type)
    lbm_host::DevBuf avg;
    bool avg_on = false;
    bool avg_pending = false;      // the current state's moments are not in the sums yet
    int avg_count = 0;
    int fold_dir = -1;             // the only closed axis (-1: none or several): its face pass can
                                   // run inside the step kernel (always in ensembles)
    bool fold_on = false;          // single-domain steps fold it
    bool inflow_fold_on = false;   // single-domain steps write the inflow/outflow planes of i in the
                                   // step kernel (with fold_on, or with no closed axis)
    int variant = 1;               // step-kernel variant (LBM_STEP_VARIANT), see impl.h; non-temporal
    int step_lds = 0;              // dynamic LDS bytes per step-kernel block (residency cap), see lbm_create
                                   // stores measured fastest on MI355X (5.32 vs 5.11 TB/s)
    // a split step (lbm_step_part) whose first part has run: its step number and arguments
    bool part_open = false;
    int part_it = 0;
    lbm::StepHost part_h{};
    // timing
    bool timing = false;
    int timing_every = 1;          // events around every timing_every-th step
    long long timing_seq = 0;
    std::vector<hipEvent_t> ev_step, ev_bc;   // pairs
    size_t n_step_ev = 0, n_bc_ev = 0;
    std::string err;
};

namespace lbm_host {

int fail(lbm_ctx* c, int code, const std::string& msg);
// LBM_OK for a live context handle; otherwise LBM_EINVAL with the message kept for
// lbm_last_error(NULL)
int check_ctx(const lbm_ctx* c);
// mass, momentum and the instability test from the totals kernel's nb per-block partial results
// (LBM_EUNSTABLE on a non-finite population or a cell density <= 0; the outputs are set either way)
int finish_totals(lbm_ctx* c, const double* partial, int nb, double* mass, double* mom);
// LBM_ESTATE while a split step (lbm_step_part) is open
int check_no_part(lbm_ctx* c);
int hipfail(lbm_ctx* c, hipError_t e, const char* where);
int dev_alloc(lbm_ctx* c, DevBuf& b, size_t n);
void dev_free(DevBuf& b);
double round_real(const lbm_ctx* c, double v);
// element q of a host image of a device array of the compute type (rb bytes per value)
void put_real(void* img, size_t q, double v, int rb);
double get_real(const void* img, size_t q, int rb);
// host reals of either kind (realkind 4 or 8) -> device array of the compute type
int upload_reals(lbm_ctx* c, DevBuf& b, const void* host, size_t n, int realkind);
void opts_of(int code, int* o1, int* oN);
double utime_at(const lbm_ctx* c, int it);
double udir_at(const lbm_ctx* c, int it);
// (utime, cos(udir), sin(udir)) of step it in the compute type's rounding
void inflow_values(const lbm_ctx* c, int it, double* ut, double* cs, double* sn);
// the step-kernel parameters of c that do not change from step to step; fold: the step kernel
// performs the face pass of c->fold_dir
lbm::StepHost step_host(const lbm_ctx* c, bool fold);
// boundarycond(f1, f2) for the non-periodic directions; skip_dir: face done by the step kernel;
// skip_inflow: inflow/outflow planes done by the step kernel
int boundary(lbm_ctx* c, void* f1, const void* f2, int it, int skip_dir = -1, bool skip_inflow = false,
             bool timed = false);
// one step (iteration it) from buffer src into buffer dst, which holds the ghosts of the state
// before src; the context's own buffers are not touched
int step_buffers(lbm_ctx* c, const void* src, void* dst, int it);
// single domains fold the closed face and the inflow planes into the step kernel unless LBM_FOLD=0
// (measured on MI355X, D3Q27 384x192x128: one launch per step instead of three, +3 % for FP16
// storage, +0.5 % for FP32)
constexpr bool kFoldDefault = true;

}  // namespace lbm_host

#define HIPCHK(c, call)                                                   \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return lbm_host::hipfail((c), e_, #call);   \
    } while (0)
