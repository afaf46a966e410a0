// C ABI of liblbm_hip.so (declared in include/lbm_hip.h): context, device memory, and the
// per-step launch sequence that stands in for the reference's time loop body
// (src/main.F90:290-329).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "ctx.h"
#include "impl.h"
#include "layout.h"
#include "dual.h"

using namespace lbm;
using namespace lbm_host;

static thread_local std::string g_create_error;

// Scratch of lbm_get_totals_noalloc: the totals kernel's per-block partial results, at most
// kTotalsMaxBlocks blocks of kTotalsSlots doubles (48 KiB on each GPU, held until the process
// ends).  In host code this name is only a placeholder: kernels and copies get each device's
// own address, resolved by scratch_resolve() when a context is created on that device.
static __device__ __align__(256) double scratch[kTotalsMaxBlocks * kTotalsSlots];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};         // device address of scratch, per device
hipError_t g_scratch_err[kMaxDevices] = {};  // result of the lookup when it failed

// With device dev current: look up that device's copy of scratch once.
void scratch_resolve(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
        g_scratch[dev] = static_cast<double*>(p);
        g_scratch_err[dev] = hipSuccess;
    } else {
        g_scratch_err[dev] = e == hipSuccess ? hipErrorInvalidSymbol : e;
    }
}

}  // namespace

namespace lbm_host {

int fail(lbm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int check_ctx(const lbm_ctx* c) {
    if (!c) return fail(nullptr, LBM_EINVAL, "null context");
    if (c->tag != kCtxTag)
        return fail(nullptr, LBM_EINVAL, "not a live lbm_ctx handle (an ensemble handle, or a destroyed context)");
    return LBM_OK;
}

int hipfail(lbm_ctx* c, hipError_t e, const char* where) {
    return fail(c, LBM_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}


int dev_alloc(lbm_ctx* c, DevBuf& b, size_t n) {
    if (b.n >= n && b.p) return LBM_OK;
    if (b.p) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(b.p));
        b.p = nullptr;
        b.n = 0;
    }
    HIPCHK(c, hipMalloc(&b.p, n));
    b.n = n;
    return LBM_OK;
}

void dev_free(DevBuf& b) {
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

size_t ncell_padded(const lbm_ctx* c) {
    return (size_t)(c->d.nx + 2) * (c->d.ny + 2) * (c->d.nz + 2);
}

int check_kind(lbm_ctx* c, int realkind) {
    if (realkind != 4 && realkind != 8) return fail(c, LBM_EINVAL, "realkind must be 4 or 8");
    return LBM_OK;
}

// Element q of a host image of a device array of the compute type (rb = 4 float, 8 double,
// 16 value + tangent; stored tangents are zero).
void put_real(void* img, size_t q, double v, int rb) {
    if (rb == 16) reinterpret_cast<Dual*>(img)[q] = Dual(v, 0.0);
    else if (rb == 8) reinterpret_cast<double*>(img)[q] = v;
    else reinterpret_cast<float*>(img)[q] = (float)v;
}
double get_real(const void* img, size_t q, int rb) {
    if (rb == 16) return reinterpret_cast<const Dual*>(img)[q].v;
    if (rb == 8) return reinterpret_cast<const double*>(img)[q];
    return reinterpret_cast<const float*>(img)[q];
}

// Host reals of either kind -> device array of the compute type.
int upload_reals(lbm_ctx* c, DevBuf& b, const void* host, size_t n, int realkind) {
    const int rb = c->ops->real_bytes;
    std::vector<unsigned char> tmp(n * rb);
    for (size_t q = 0; q < n; ++q) {
        double v = realkind == 8 ? static_cast<const double*>(host)[q] : static_cast<const float*>(host)[q];
        put_real(tmp.data(), q, v, rb);
    }
    int rc = dev_alloc(c, b, n * rb);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(b.p, tmp.data(), n * rb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// A value in the compute type's rounding.
double round_real(const lbm_ctx* c, double v) {
    return c->ops->real_bytes == 4 ? (double)(float)v : v;
}

void opts_of(int code, int* o1, int* oN) {
    // src/m_boundarycond.F90:166-172
    switch (code) {
        case 11: *o1 = -1; *oN = -1; break;
        case 12: *o1 = -1; *oN = +1; break;
        case 21: *o1 = +1; *oN = -1; break;
        case 22: *o1 = +1; *oN = +1; break;
        default: *o1 = 0; *oN = 0;
    }
}

// Ghost edges and corners in the reference's order (src/m_boundarycond.F90:270-310).
EdgeList edge_list(const lbm_desc& d) {
    EdgeList e{};
    e.n = 0;
    const int n[3] = {d.nx, d.ny, d.nz};
    auto add_edge = [&](int d1, int s1, int d2, int s2, int m0, int m1, int m2) {
        const int t = 3 - d1 - d2;
        const int q = e.n++;
        e.kind[q] = 0;
        e.t[q] = t;
        e.len[q] = n[t];
        e.g[q][t] = 0;
        e.g[q][d1] = s1 ? n[d1] + 1 : 0;
        e.g[q][d2] = s2 ? n[d2] + 1 : 0;
        e.out[q][0] = e.out[q][1] = e.out[q][2] = 0;
        e.out[q][d1] = s1 ? 1 : -1;
        e.out[q][d2] = s2 ? 1 : -1;
        e.mir[q][0] = m0; e.mir[q][1] = m1; e.mir[q][2] = m2;
    };
    if (d.jbnd > 10 && d.kbnd > 10) {
        // src/m_boundary_j_k_edges.F90; the (ny+1,nz+1) edge keeps cx (line 129)
        const int opt = (d.jbnd == 22 && d.kbnd == 22) ? 1 : -1;
        add_edge(1, 0, 2, 0, opt, -1, -1);
        add_edge(1, 1, 2, 0, opt, -1, -1);
        add_edge(1, 0, 2, 1, opt, -1, -1);
        add_edge(1, 1, 2, 1, 1, -1, -1);
    }
    if (d.ibnd > 10 && d.jbnd > 10) {
        const int opt = (d.ibnd == 22 && d.jbnd == 22) ? 1 : -1;   // src/m_boundary_i_j_edges.F90
        add_edge(0, 0, 1, 0, -1, -1, opt);
        add_edge(0, 1, 1, 0, -1, -1, opt);
        add_edge(0, 0, 1, 1, -1, -1, opt);
        add_edge(0, 1, 1, 1, -1, -1, opt);
    }
    if (d.ibnd > 10 && d.kbnd > 10) {
        const int opt = (d.ibnd == 22 && d.kbnd == 22) ? 1 : -1;   // src/m_boundary_i_k_edges.F90
        add_edge(0, 0, 2, 0, -1, opt, -1);
        add_edge(0, 1, 2, 0, -1, opt, -1);
        add_edge(0, 0, 2, 1, -1, opt, -1);
        add_edge(0, 1, 2, 1, -1, opt, -1);
    }
    if (d.ibnd > 10 && d.jbnd > 10 && d.kbnd > 10) {
        // src/m_boundary_closed_corners.F90:177-186 builds its table with reshape(list,[8,3]),
        // which fills column major: corner m = (list[m], list[8+m], list[16+m]).  Entries whose
        // indices are not all ghost indices of their own axis do nothing there; repeated
        // corners are idempotent.  Keep the distinct real corners of that table.
        const int nx = d.nx, ny = d.ny, nz = d.nz;
        const int list[24] = {0, 0, 0, 0, 0, nz + 1, 0, ny + 1, 0, 0, ny + 1, nz + 1,
                              nx + 1, 0, 0, nx + 1, 0, nz + 1, nx + 1, ny + 1, 0, nx + 1, ny + 1, nz + 1};
        for (int m = 0; m < 8; ++m) {
            const int ic = list[m], jc = list[8 + m], kc = list[16 + m];
            const bool ok = (ic == 0 || ic == nx + 1) && (jc == 0 || jc == ny + 1) && (kc == 0 || kc == nz + 1);
            if (!ok) continue;
            bool dup = false;
            for (int q = 0; q < e.n; ++q)
                if (e.kind[q] == 1 && e.g[q][0] == ic && e.g[q][1] == jc && e.g[q][2] == kc) dup = true;
            if (dup) continue;
            const int q = e.n++;
            e.kind[q] = 1;
            e.t[q] = 0;
            e.len[q] = 1;
            e.g[q][0] = ic; e.g[q][1] = jc; e.g[q][2] = kc;
            e.out[q][0] = ic == 0 ? -1 : 1;
            e.out[q][1] = jc == 0 ? -1 : 1;
            e.out[q][2] = kc == 0 ? -1 : 1;
            e.mir[q][0] = e.mir[q][1] = e.mir[q][2] = -1;
        }
    }
    return e;
}

// Ghost cells read by a closed-face pass while holding the first periodic pass's copy of a cell
// that an earlier closed pass of the same call has rewritten (pass order i, j, k;
// src/m_boundarycond.F90:69-255).  For the k faces: tangential (i, j) ghosts where one of i, j is
// periodic and the other closed; for the j faces: (i, k) ghosts with k periodic and i closed.
void snap_lists(const lbm_desc& d, SnapList& sk, SnapList& sj) {
    sk = SnapList{};
    sj = SnapList{};
    const bool ci = d.ibnd > 10, cj = d.jbnd > 10, ck = d.kbnd > 10;
    const bool pi = d.ibnd == 0, pj = d.jbnd == 0, pk = d.kbnd == 0;
    if (ck && ((pi && cj) || (ci && pj)))
        for (int a : {0, d.nx + 1})
            for (int b : {0, d.ny + 1})
                for (int c : {1, d.nz}) {
                    sk.g[sk.n][0] = a; sk.g[sk.n][1] = b; sk.g[sk.n][2] = c;
                    ++sk.n;
                }
    if (cj && pk && ci)
        for (int a : {0, d.nx + 1})
            for (int b : {1, d.ny})
                for (int c : {0, d.nz + 1}) {
                    sj.g[sj.n][0] = a; sj.g[sj.n][1] = b; sj.g[sj.n][2] = c;
                    ++sj.n;
                }
}

double utime_at(const lbm_ctx* c, int it) {
    if (!c->ut_series.empty() && it >= c->series_lo && it <= c->series_hi) return c->ut_series[it - c->series_lo];
    return c->uini;
}
double udir_at(const lbm_ctx* c, int it) {
    if (!c->ud_series.empty() && it >= c->series_lo && it <= c->series_hi) return c->ud_series[it - c->series_lo];
    return c->udir;
}

hipEvent_t next_event(lbm_ctx* c, std::vector<hipEvent_t>& pool, size_t& used) {
    if (used == pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        pool.push_back(e);
    }
    return pool[used++];
}

void inflow_values(const lbm_ctx* c, int it, double* ut, double* cs, double* sn) {
    // cos(udir*pi/180.0) in the reference's REAL precision (src/m_boundary_i_inflow_kernel.F90:71-73)
    *ut = round_real(c, utime_at(c, it));
    if (c->ops->real_bytes == 4) {
        const float ang = (float)udir_at(c, it) * 3.1415927410125732f / 180.0f;
        *cs = cosf(ang);
        *sn = sinf(ang);
    } else {
        const double ang = udir_at(c, it) * 3.1415927410125732 / 180.0;
        *cs = cos(ang);
        *sn = sin(ang);
    }
}

StepHost step_host(const lbm_ctx* c, bool fold) {
    const lbm_desc& d = c->d;
    StepHost h{};
    h.fac0 = c->fac0;
    h.kinevisc = d.kinevisc;
    h.vconst = round_real(c, 2.5 * round_real(c, round_real(c, d.smagorinsky) * round_real(c, d.smagorinsky)));
    h.ablvisc = c->ablvisc.p;
    h.solid = c->has_solid ? static_cast<const uint32_t*>(c->solid.p) : nullptr;
    h.solid_wpr = c->solid_wpr;
    h.ip = d.iturb_pos;
    h.ampl = c->ampl;
    // forcing: the caller's field (every interior cell) or the turbines' (their deposit box)
    for (int q = 0; q < 3; ++q) h.F[q] = (c->has_force || c->turb_on) ? c->force[q].p : nullptr;
    const int n[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; ++a) {
        h.flo[a] = c->turb_on ? c->th.lo[a] : 1;
        h.fhi[a] = c->turb_on ? c->th.hi[a] : n[a];
    }
    h.theta = c->buoy_on ? c->theta.p : nullptr;
    h.theta0 = c->btheta0;
    h.bscale = c->bscale;
    h.tau_out = (d.store_tau && d.ivreman) ? c->tau.p : nullptr;
    h.bgk3 = d.ibgk == 3;
    h.hrr = d.ihrr == 1;
    h.vre = d.ivreman == 1;
    h.block_x = d.block_x;
    h.block_y = d.block_y;
    h.variant = c->variant;
    h.lds_bytes = c->step_lds;
    h.fold_dir = fold ? c->fold_dir : -1;
    h.fold_opt_lo = h.fold_opt_hi = 0;
    if (h.fold_dir >= 0) {
        const int code = c->fold_dir == 0 ? d.ibnd : c->fold_dir == 1 ? d.jbnd : d.kbnd;
        opts_of(code, &h.fold_opt_lo, &h.fold_opt_hi);
    }
    h.heavy_first = h.fold_dir < 0 ? 0 : h.fold_dir == c->lay.ax1 ? 1 : h.fold_dir == c->lay.ax2 ? 2 : 0;
    h.halo_dir = d.jbnd == LBM_BND_HALO ? 1 : -1;
    return h;
}

// One step from src into dst (both buffers of the context's layout), as lbm_step performs a step
// after the first: dst must hold the ghosts of the state before src.  Used by the adjoint's
// recompute, which writes each step into a snapshot slot of its own.  No turbulence forcing, no
// averaging, no external moments (the adjoint's preconditions).
int step_buffers(lbm_ctx* c, const void* src, void* dst, int it) {
    StepHost h = step_host(c, c->fold_on);
    if (c->inflow_fold_on) {
        h.inflow_fold = 1;
        h.shear = c->shear.p;
        h.rho0 = c->d.rho0;
        inflow_values(c, it, &h.utime, &h.cosd, &h.sind);
    }
    h.src = src;
    h.dst = dst;
    h.stream = 1;
    HIPCHK(c, c->ops->step(c->lay, h, c->stream));
    return boundary(c, dst, src, it, c->fold_on ? c->fold_dir : -1, c->inflow_fold_on);
}

// boundarycond(f1, f2) for the non-periodic directions (src/m_boundarycond.F90:166-365).
// skip_dir: a closed axis whose face pass the step kernel has already performed.
int boundary(lbm_ctx* c, void* f1, const void* f2, int it, int skip_dir, bool skip_inflow, bool timed) {
    const lbm_desc& d = c->d;
    const Layout& L = c->lay;
    int a1, a2, b1, b2, k1, k2;
    opts_of(d.ibnd, &a1, &a2);
    opts_of(d.jbnd, &b1, &b2);
    opts_of(d.kbnd, &k1, &k2);
    // timed: bracket the boundary kernels with events, if this step launches any
    timed = timed && (c->snap_k.n || c->snap_j.n || (a1 && skip_dir != 0) || (b1 && skip_dir != 1) ||
                      (k1 && skip_dir != 2) || (d.ibnd == 1 && !skip_inflow));
    hipEvent_t e0 = nullptr;
    if (timed) {
        e0 = next_event(c, c->ev_bc, c->n_bc_ev);
        if (e0) hipEventRecord(e0, c->stream);
    }
    const SnapList none{};
    char* snap = static_cast<char*>(c->snap.p);
    const size_t half = (size_t)8 * d.q * c->ops->raw_bytes;
    if (c->snap_k.n) HIPCHK(c, c->ops->snapshot(L, f1, snap, c->snap_k, c->stream));
    if (c->snap_j.n) HIPCHK(c, c->ops->snapshot(L, f1, snap + half, c->snap_j, c->stream));
    if (a1 && skip_dir != 0) HIPCHK(c, c->ops->closed(L, f1, f2, 0, a1, a2, nullptr, none, c->stream));
    if (b1 && skip_dir != 1) HIPCHK(c, c->ops->closed(L, f1, f2, 1, b1, b2, snap + half, c->snap_j, c->stream));
    if (k1 && skip_dir != 2) HIPCHK(c, c->ops->closed(L, f1, f2, 2, k1, k2, snap, c->snap_k, c->stream));
    const EdgeList e = edge_list(d);
    if (e.n) HIPCHK(c, c->ops->edges(L, f1, f2, e, c->stream));
    if (d.ibnd == 1 && !skip_inflow) {
        double ut, cs, sn;
        inflow_values(c, it, &ut, &cs, &sn);
        HIPCHK(c, c->ops->inflow(L, f1, c->shear.p, ut, cs, sn, d.rho0, c->stream));
    }
    if (timed && e0) {
        hipEvent_t e1 = next_event(c, c->ev_bc, c->n_bc_ev);
        if (e1) hipEventRecord(e1, c->stream);
    }
    return LBM_OK;
}

int finish_totals(lbm_ctx* c, const double* h, int nb, double* mass, double* mom) {
    double t[4] = {0, 0, 0, 0}, rmin = INFINITY, bad = 0;
    for (int b = 0; b < nb; ++b) {
        const double* p = h + (size_t)kTotalsSlots * b;
        for (int q = 0; q < 4; ++q) t[q] += p[q];
        rmin = std::min(rmin, p[4]);
        bad += p[5];
    }
    if (mass) *mass = t[0];
    if (mom) { mom[0] = t[1]; mom[1] = t[2]; mom[2] = t[3]; }
    // src/m_diag.F90:169-173 stops on a negative density
    if (bad > 0 || !(rmin > 0.0)) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "unstable state: %.0f non-finite populations, smallest cell density %g", bad, rmin);
        return fail(c, LBM_EUNSTABLE, msg);
    }
    return LBM_OK;
}

int check_no_part(lbm_ctx* c) {
    if (c->part_open)
        return fail(c, LBM_ESTATE, "a split step is open: finish it with lbm_step_part(ctx, it, LBM_PART_INNER)");
    return LBM_OK;
}

}  // namespace

extern "C" {

const char* lbm_last_error(const lbm_ctx* ctx) {
    return ctx && ctx->tag == kCtxTag ? ctx->err.c_str() : g_create_error.c_str();
}

int lbm_create(const lbm_desc* desc, lbm_ctx** out) {
    if (!desc || !out) return fail(nullptr, LBM_EINVAL, "null argument");
    *out = nullptr;
    const lbm_desc& d = *desc;
    if (d.nx < 2 || d.ny < 2 || d.nz < 2) return fail(nullptr, LBM_EINVAL, "nx, ny, nz must be >= 2");
    if ((long long)(d.nx + 2) * (d.ny + 2) * (d.nz + 2) > (1LL << 30))
        return fail(nullptr, LBM_EINVAL, "grid too large for 32-bit cell offsets");
    if (d.q != 19 && d.q != 27) return fail(nullptr, LBM_EINVAL, "q must be 19 or 27");
    if (d.ibgk != 2 && d.ibgk != 3) return fail(nullptr, LBM_EINVAL, "ibgk must be 2 or 3");
    auto okb = [](int b) { return b == 0 || b == 11 || b == 12 || b == 21 || b == 22; };
    if (!(okb(d.ibnd) || d.ibnd == 1) || !(okb(d.jbnd) || d.jbnd == LBM_BND_HALO) || !okb(d.kbnd))
        return fail(nullptr, LBM_EINVAL, "boundary codes: 0, 11, 12, 21, 22 (and 1 for ibnd, LBM_BND_HALO for jbnd)");
    if (d.jbnd == LBM_BND_HALO && d.ibnd > 10 && d.kbnd > 10)
        return fail(nullptr, LBM_EINVAL, "a halo domain takes at most one closed axis");
    if (d.ibnd == 1 && (d.iturb_pos < 1 || d.iturb_pos > d.nx) && d.iturb_pos != 0)
        return fail(nullptr, LBM_EINVAL, "iturb_pos outside 1..nx");
    const Ops* ops = nullptr;
    switch (d.storage) {
        case LBM_STORE_F32: ops = ops_f32(d.q); break;
        case LBM_STORE_F64: ops = ops_f64(d.q); break;
        case LBM_STORE_F16S: ops = ops_f16s(d.q); break;
        case LBM_STORE_BF16S: ops = ops_bf16s(d.q); break;
        case LBM_STORE_F64T: ops = ops_f64t(d.q); break;
        default: return fail(nullptr, LBM_EINVAL, "unknown storage kind");
    }
    // contiguous storage axis: the first periodic axis (i, else j, else k), else i; LBM_FAST_AXIS
    // (0, 1, 2 for i, j, k) overrides it
    int fast = -1;
    if (const char* fa = getenv("LBM_FAST_AXIS")) {
        fast = atoi(fa);
        if (fast < 0 || fast > 2) return fail(nullptr, LBM_EINVAL, "LBM_FAST_AXIS must be 0, 1 or 2");
    }
    // population buffers row-blocked ([row][population][pitch]) unless LBM_LAYOUT=planes: measured on
    // MI355X, D3Q27 384x192x128: FP32 25.5 k vs 24.4 k MLUPS, FP16 47.4 k vs 45.0 k
    const char* lv = getenv("LBM_LAYOUT");
    const bool rowblock = !(lv && std::string(lv) == "planes");
    const Layout lay = make_layout(d.nx, d.ny, d.nz, ops->raw_bytes, d.ibnd == 0, d.jbnd == 0, d.kbnd == 0, fast,
                                   d.q, rowblock);
    // the step kernel of 8-byte (and tangent) storage is compiled for at most 256 threads per block
    // (more registers per lane, no spills)
    const int max_threads = ops->raw_bytes >= 8 ? 256 : 1024;
    int bx = d.block_x, by = d.block_y;
    if (bx == 0) {
        // Measured on MI355X (step kernel, bench.py):
        // 4-byte storage: blocks of 64 cells x 4 rows (D3Q27 reg+Vreman 384x192x128: 352.3 us against
        //   353-361 us for 64 x 2/3/8 and one 192-cell row, 384 us for 128 x 2);
        // 2-byte storage: one whole row per block, rows over 512 cells in blocks of 256 (D3Q27
        //   768x256x192: 732 us against 749 us for 64 x 1; D3Q19 BGK 768x256x192: 484 us against
        //   487 / 497 us for 128 / 64 cells; on the 192-cell rows of 384x192x128, 64 x 1 is 1 % faster);
        // 8-byte storage: blocks of one wave (D3Q27 reg+Vreman 384x192x128 FP64: 727 us against 758 us
        //   for one 192-cell row per block); the 16-byte tangent storage keeps whole rows.
        const int rows = (lay.n(lay.ax0) + 63) / 64;
        if (ops->raw_bytes == 4) {
            bx = 64;
        } else if (ops->raw_bytes == 8) {
            bx = 64;
            if (by == 0) by = 1;
        } else {
            bx = rows * 64 <= std::min(512, max_threads) ? rows * 64 : 256;
        }
    }
    if (by == 0) by = bx >= 256 ? 1 : 256 / bx;
    if (bx % 64 != 0 || bx * by > max_threads || by < 1)
        return fail(nullptr, LBM_EINVAL, "block_x must be a multiple of 64, block_x*block_y <= " +
                                             std::to_string(max_threads) + " for this storage");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(nullptr, LBM_EHIP, std::string("no HIP device: ") + hipGetErrorString(e));

    lbm_ctx* c = new lbm_ctx();
    c->d = d;
    c->d.block_x = bx;
    c->d.block_y = by;
    if (c->d.iturb_pos == 0) c->d.iturb_pos = 10;
    c->ops = ops;
    // LBM_STEP_VARIANT overrides the step-kernel variant: bit 0 non-temporal stores (the default),
    // bit 1 the form for population buffers of 4 GiB and more (64-bit plane offsets) whatever the
    // buffer's size, which lets small grids exercise that path, bit 2 the one-base form with the
    // storage axes mapped at run time.
    if (const char* v = getenv("LBM_STEP_VARIANT")) c->variant = atoi(v) & 7;
    {
        // With one closed axis (no edges, corners or snapshots then) the step kernel can perform
        // that axis's face pass itself, and with at most that one closed axis the inflow/outflow
        // planes of i as well: a step is then one launch.  Ensembles always fold the face (a
        // member's face must not need a launch of its own).  LBM_FOLD=0 selects the separate
        // boundary kernels for a single domain.
        const int codes[3] = {d.ibnd, d.jbnd, d.kbnd};
        int nclosed = 0, cl = -1;
        for (int a = 0; a < 3; ++a)
            if (codes[a] > 10) { ++nclosed; cl = a; }
        if (nclosed == 1) c->fold_dir = cl;
        const char* fv = getenv("LBM_FOLD");
        // halo domains always fold: the separate boundary kernels do not know halo planes
        const bool fold = d.jbnd == LBM_BND_HALO || (fv ? atoi(fv) != 0 : kFoldDefault);
        c->fold_on = c->fold_dir >= 0 && fold;
        c->inflow_fold_on = d.ibnd == 1 && fold && (nclosed == 0 || c->fold_on);
        // LBM_INFLOW_FOLD=0 keeps the face folded but writes the inflow/outflow planes with their own
        // kernel (measured slower: tile 384x24x128 14.0 k vs 16.2 k MLUPS, whole domain -1 %)
        if (const char* iv = getenv("LBM_INFLOW_FOLD")) c->inflow_fold_on = c->inflow_fold_on && atoi(iv) != 0;
    }
    hipGetDevice(&c->device);
    // lbm_get_totals_noalloc's scratch on this device (a failed lookup is reported by that call)
    scratch_resolve(c->device);
    // Resident step-kernel blocks per CU, capped by reserving 1/n of the CU's LDS per block (the
    // kernel itself uses none).  D3Q27 with 4-byte storage: 5 waves per SIMD (20 per CU; 5 blocks of
    // 64 x 4 cells) where its 78 VGPRs would allow 6 (measured on MI355X, reg+Vreman 384x192x128:
    // 351.1 us per step kernel against 358.9 us; 768x256x192: 1373 us against 1444 us); the 16-bit
    // and D3Q19 kernels run best uncapped.  LBM_STEP_BLOCKS_PER_CU=n overrides (0: no cap).
    {
        const int waves_per_block = std::max(1, bx * by / 64);
        int n = ops->raw_bytes == 4 && d.q == 27 ? std::max(1, 20 / waves_per_block) : 0;
        if (const char* v = getenv("LBM_STEP_BLOCKS_PER_CU")) n = atoi(v);
        int lds_cu = 0;
        if (n >= 1 && hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device) == hipSuccess) {
            int per_block = 0;
            if (hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess) per_block = 0;
            c->step_lds = std::min(lds_cu / n, per_block);
        }
    }
    c->lay = lay;
    c->tau0 = round_real(c, round_real(c, 3.0 * round_real(c, d.kinevisc)) + 0.5);
    c->fac0 = round_real(c, 1.0 - round_real(c, 1.0 / c->tau0));
    auto bail = [&](hipError_t err, const char* where) {
        std::string m = std::string(where) + ": " + hipGetErrorString(err);
        lbm_destroy(c);
        return fail(nullptr, LBM_EHIP, m);
    };
    if ((e = hipStreamCreate(&c->stream)) != hipSuccess) return bail(e, "hipStreamCreate");
    c->buf_bytes = (size_t)c->lay.total * ops->raw_bytes;
    for (int b = 0; b < 2; ++b)
        if ((e = hipMalloc(&c->buf[b], c->buf_bytes)) != hipSuccess) return bail(e, "hipMalloc(populations)");
    for (int b = 0; b < 2; ++b)
        if ((e = hipMemsetAsync(c->buf[b], 0, c->buf_bytes, c->stream)) != hipSuccess) return bail(e, "hipMemset");
    snap_lists(d, c->snap_k, c->snap_j);
    if (c->snap_k.n || c->snap_j.n) {
        const size_t sb = (size_t)16 * d.q * ops->raw_bytes;
        if ((e = hipMalloc(&c->snap.p, sb)) != hipSuccess) return bail(e, "hipMalloc(snapshot)");
        c->snap.n = sb;
    }
    if (d.store_tau && d.ivreman) {
        const size_t tb = (size_t)c->lay.cplane * ops->real_bytes;
        if ((e = hipMalloc(&c->tau.p, tb)) != hipSuccess) return bail(e, "hipMalloc(tau)");
        c->tau.n = tb;
        if ((e = hipMemsetAsync(c->tau.p, 0, tb, c->stream)) != hipSuccess) return bail(e, "hipMemset(tau)");
    }
    // defaults: ablvisc = 0, uvel_shear = 1
    std::vector<double> zeros(d.nz, 0.0), ones(d.nz, 1.0);
    if (upload_reals(c, c->ablvisc, zeros.data(), d.nz, 8) || upload_reals(c, c->shear, ones.data(), d.nz, 8)) {
        std::string m = c->err;
        lbm_destroy(c);
        return fail(nullptr, LBM_EHIP, m);
    }
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx* c) {
    if (!c) return LBM_OK;
    if (int rc = check_ctx(c)) return rc;
    c->tag = 0;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int b = 0; b < 2; ++b)
        if (c->buf[b]) hipFree(c->buf[b]);
    DevBuf* all[] = {&c->ablvisc, &c->shear, &c->solid, &c->turb[0], &c->turb[1], &c->turb[2], &c->force[0],
                     &c->force[1], &c->force[2], &c->tau, &c->ext[0], &c->ext[1], &c->ext[2], &c->ext[3],
                     &c->stage, &c->partial, &c->snap, &c->t_rot, &c->t_theta, &c->t_blade, &c->t_nfoil,
                     &c->t_nang, &c->t_foil, &c->t_pts, &c->t_box, &c->theta, &c->adj_ck, &c->adj_snaps,
                     &c->adj_lam, &c->adj_gF, &c->adj_gS, &c->adj_acc};
    for (DevBuf* b : all) dev_free(*b);
    for (hipEvent_t ev : c->ev_step) hipEventDestroy(ev);
    for (hipEvent_t ev : c->ev_bc) hipEventDestroy(ev);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return LBM_OK;
}

int lbm_set_state(lbm_ctx* c, const void* f, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!f) return fail(c, LBM_EINVAL, "null argument");
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    const size_t n = ncell_padded(c) * c->d.q;
    if (int rc = dev_alloc(c, c->stage, n * realkind)) return rc;
    HIPCHK(c, hipMemcpyAsync(c->stage.p, f, n * realkind, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->ops->aos_in(c->lay, c->stage.p, realkind, c->buf[0], c->buf[1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->have_state = true;
    c->pending = true;
    c->use_ext = false;
    c->ghost_raw[0] = c->ghost_raw[1] = true;
    return LBM_OK;
}

int lbm_init_equilibrium(lbm_ctx* c, const void* rho, const void* u, const void* v, const void* w, int realkind,
                         int it0) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!rho || !u || !v || !w) return fail(c, LBM_EINVAL, "null argument");
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    if (c->d.jbnd == LBM_BND_HALO)
        return fail(c, LBM_EINVAL, "a halo domain starts from lbm_set_state (its ghost planes are the neighbours')");
    const size_t nc = ncell_padded(c);
    const size_t cb = nc * realkind;
    if (int rc = dev_alloc(c, c->stage, 4 * cb)) return rc;
    const size_t mb = (size_t)c->lay.cplane * c->ops->real_bytes;
    for (int q = 0; q < 4; ++q)
        if (int rc = dev_alloc(c, c->ext[q], mb)) return rc;
    char* st = static_cast<char*>(c->stage.p);
    HIPCHK(c, hipMemcpyAsync(st, rho, cb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(st + cb, u, cb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(st + 2 * cb, v, cb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(st + 3 * cb, w, cb, hipMemcpyHostToDevice, c->stream));
    // fB = feq with clamped ghosts; fA = fB; boundarycond(fB,fA); boundarycond(fA,fB)
    HIPCHK(c, c->ops->feq_init(c->lay, c->d.ibgk == 3, st, st + cb, st + 2 * cb, st + 3 * cb, realkind, c->buf[1],
                               c->ext[0].p, c->ext[1].p, c->ext[2].p, c->ext[3].p, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->buf[0], c->buf[1], c->buf_bytes, hipMemcpyDeviceToDevice, c->stream));
    if (int rc = boundary(c, c->buf[1], c->buf[0], it0)) return rc;
    if (int rc = boundary(c, c->buf[0], c->buf[1], it0)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_bc_ev = 0;
    c->cur = 0;
    c->have_state = true;
    c->pending = true;
    c->use_ext = true;
    c->ghost_raw[0] = c->ghost_raw[1] = false;
    return LBM_OK;
}

namespace {
// The step arguments that stay the same from step to step.
StepHost step_base(lbm_ctx* c) {
    StepHost h = step_host(c, c->fold_on);
    if (c->inflow_fold_on) {
        h.inflow_fold = 1;
        h.shear = c->shear.p;
        h.rho0 = c->d.rho0;
        h.heavy_first |= c->lay.ax1 == 0 ? 1 : c->lay.ax2 == 0 ? 2 : 0;
    }
    return h;
}

// Buffers and per-step inputs of step it.
void step_inputs(lbm_ctx* c, StepHost& h, int it) {
    const lbm_desc& d = c->d;
    if (c->pending) {
        // collide the uploaded state in place: read the copy in the other buffer
        h.src = c->buf[c->cur ^ 1];
        h.dst = c->buf[c->cur];
        h.stream = 0;
    } else {
        h.src = c->buf[c->cur];
        h.dst = c->buf[c->cur ^ 1];
        h.stream = 1;
    }
    for (int q = 0; q < 4; ++q) h.ext[q] = c->use_ext ? c->ext[q].p : nullptr;
    // the step kernel adds the moments of its input (the state after the previous step)
    h.avg = (c->avg_on && c->avg_pending) ? c->avg.p : nullptr;
    if (h.avg) ++c->avg_count;
    if (c->has_turb) {
        const size_t slice = (size_t)d.ny * d.nz * c->ops->real_bytes;
        int lit = it % c->nrturb;           // src/m_inflow_turbulence_forcing.F90:45-46
        if (lit == 0) lit = c->nrturb;
        for (int q = 0; q < 3; ++q) h.turb[q] = static_cast<const char*>(c->turb[q].p) + (size_t)lit * slice;
    } else {
        h.turb[0] = h.turb[1] = h.turb[2] = nullptr;
    }
    if (h.inflow_fold) inflow_values(c, it, &h.utime, &h.cosd, &h.sind);
}

// After the step kernel: the new state is current.
void step_advance(lbm_ctx* c) {
    if (!c->pending) c->cur ^= 1;
    c->pending = false;
    c->use_ext = false;
}

// The boundary passes that the step kernel did not perform, and the bookkeeping of a finished step.
int step_complete(lbm_ctx* c, int it, bool timed) {
    if (int rc = boundary(c, c->buf[c->cur], c->buf[c->cur ^ 1], it, c->fold_on ? c->fold_dir : -1,
                          c->inflow_fold_on, timed))
        return rc;
    c->ghost_raw[c->cur] = false;
    c->avg_pending = c->avg_on;
    return LBM_OK;
}

}  // namespace

int lbm_step(lbm_ctx* c, int it0, int nsteps) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state: call lbm_set_state or lbm_init_equilibrium first");
    if (nsteps < 0) return fail(c, LBM_EINVAL, "nsteps < 0");
    if (int rc = check_no_part(c)) return rc;
    StepHost h = step_base(c);
    for (int it = it0 + 1; it <= it0 + nsteps; ++it) {
        step_inputs(c, h, it);
        // events on every timing_every-th step only: each record costs the stream a few us
        const bool timed = c->timing && c->timing_seq++ % c->timing_every == 0;
        // turbine forcing of this step from its input's moments (src/main.F90:291-292)
        if (c->turb_on) HIPCHK(c, c->ops->turb_step(c->lay, c->th, h.src, h.stream, h.ext, c->stream));
        hipEvent_t e0 = nullptr;
        if (timed && (e0 = next_event(c, c->ev_step, c->n_step_ev))) hipEventRecord(e0, c->stream);
        HIPCHK(c, c->ops->step(c->lay, h, c->stream));
        if (timed && e0) {
            hipEvent_t e1 = next_event(c, c->ev_step, c->n_step_ev);
            if (e1) hipEventRecord(e1, c->stream);
        }
        step_advance(c);
        if (int rc = step_complete(c, it, timed)) return rc;
    }
    return LBM_OK;
}

int lbm_step_part(lbm_ctx* c, int it, int part) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (part != LBM_PART_EDGE && part != LBM_PART_INNER)
        return fail(c, LBM_EINVAL, "part must be LBM_PART_EDGE or LBM_PART_INNER");
    if (part == LBM_PART_INNER) {
        if (!c->part_open) return fail(c, LBM_ESTATE, "no split step is open: run its LBM_PART_EDGE first");
        if (it != c->part_it) return fail(c, LBM_EINVAL, "it differs from the open split step's");
        c->part_open = false;
        if (c->part_h.part) {
            StepHost h = c->part_h;
            h.part = LBM_PART_INNER;
            HIPCHK(c, c->ops->step(c->lay, h, c->stream));
        }
        return step_complete(c, it, false);
    }
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state: call lbm_set_state first");
    if (int rc = check_no_part(c)) return rc;
    if (c->d.jbnd != LBM_BND_HALO) return fail(c, LBM_EINVAL, "split steps need a halo domain (jbnd = LBM_BND_HALO)");
    if (c->turb_on) return fail(c, LBM_EINVAL, "split steps take no turbines");
    if (c->d.ibnd == 1 && !c->inflow_fold_on)
        return fail(c, LBM_EINVAL, "split steps need the inflow/outflow planes in the step kernel (LBM_INFLOW_FOLD)");
    const int axis = c->lay.ax1 == 1 ? 1 : c->lay.ax2 == 1 ? 2 : 0;
    if (!axis) return fail(c, LBM_EINVAL, "split steps need j off the contiguous axis (LBM_FAST_AXIS)");
    StepHost h = step_base(c);
    step_inputs(c, h, it);
    // Edge rows per side: 2 with a folded face (the unpack completes face ghosts in columns 1
    // and ny, which rows 2 and ny-1 pull; rows 2 and ny-1 push face ghosts into the columns the
    // pack reads), else 1.  Tiles without inner rows: the first part is the whole step.
    h.part_edge = c->fold_on && (c->fold_dir == 0 || c->fold_dir == 2) ? 2 : 1;
    h.part = c->d.ny > 2 * h.part_edge ? LBM_PART_EDGE : 0;
    h.part_axis = axis;
    HIPCHK(c, c->ops->step(c->lay, h, c->stream));
    step_advance(c);
    c->part_h = h;
    c->part_it = it;
    c->part_open = true;
    return LBM_OK;
}

namespace {
// Adds the current state's moments to the averaging sums if they are not there yet.
int average_flush(lbm_ctx* c) {
    if (!c->avg_on || !c->avg_pending) return LBM_OK;
    HIPCHK(c, c->ops->average(c->lay, c->buf[c->cur], c->pending ? 0 : 1, c->avg.p, c->stream));
    ++c->avg_count;
    c->avg_pending = false;
    return LBM_OK;
}
}  // namespace

int lbm_set_averaging(lbm_ctx* c, int on) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (on) {
        const size_t n = 6 * (size_t)c->lay.cplane * c->ops->real_bytes;
        if (int rc = dev_alloc(c, c->avg, n)) return rc;
        HIPCHK(c, hipMemsetAsync(c->avg.p, 0, n, c->stream));
        c->avg_on = true;
        c->avg_pending = false;
        c->avg_count = 0;
        return LBM_OK;
    }
    if (int rc = average_flush(c)) return rc;
    c->avg_on = false;
    return LBM_OK;
}

int lbm_get_averages(lbm_ctx* c, void* uave, void* vave, void* wave, void* ti, int* count, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    if (!c->avg.p) return fail(c, LBM_ESTATE, "no averages: call lbm_set_averaging first");
    if (int rc = average_flush(c)) return rc;
    const Layout& L = c->lay;
    const int rb = c->ops->real_bytes;
    const size_t P = (size_t)L.cplane;
    std::vector<unsigned char> img(6 * P * rb);
    HIPCHK(c, hipMemcpyAsync(img.data(), c->avg.p, img.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (count) *count = c->avg_count;
    const size_t nc = ncell_padded(c);
    void* out[4] = {uave, vave, wave, ti};
    for (int q = 0; q < 4; ++q)
        if (out[q]) std::memset(out[q], 0, nc * realkind);
    if (c->avg_count == 0) return LBM_OK;
    auto put = [&](void* o, size_t idx, double v) {
        if (!o) return;
        if (realkind == 8) static_cast<double*>(o)[idx] = v;
        else static_cast<float*>(o)[idx] = (float)v;
    };
    const bool f32 = rb == 4;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i) {
                const size_t o = (size_t)L.cidx(i, j, k);
                const size_t hc = ((size_t)k * (L.ny + 2) + j) * (L.nx + 2) + i;
                double m[3], m2[3];
                for (int q = 0; q < 3; ++q) {
                    m[q] = get_real(img.data(), q * P + o, rb) / c->avg_count;
                    m2[q] = get_real(img.data(), (3 + q) * P + o, rb) / c->avg_count;
                    if (f32) {
                        m[q] = (float)m[q];
                        m2[q] = (float)m2[q];
                    }
                }
                // src/m_averaging_full_kernelfin.F90:46-60
                double t = 0.0;
                for (int q = 0; q < 3; ++q) {
                    const double var = f32 ? (double)((float)m2[q] - (float)m[q] * (float)m[q]) : m2[q] - m[q] * m[q];
                    t += var > 0.0 ? var : 0.0;
                }
                t = f32 ? (double)((float)t / 3.0f) : t / 3.0;
                put(uave, hc, m[0]);
                put(vave, hc, m[1]);
                put(wave, hc, m[2]);
                put(ti, hc, f32 ? (double)std::sqrt((float)(t > 0.0 ? t : 0.0)) : std::sqrt(t > 0.0 ? t : 0.0));
            }
    return LBM_OK;
}

int lbm_get_state(lbm_ctx* c, void* f, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!f) return fail(c, LBM_EINVAL, "null argument");
    if (int rc = check_no_part(c)) return rc;
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state");
    if (int rc = check_kind(c, realkind)) return rc;
    const size_t n = ncell_padded(c) * c->d.q;
    if (int rc = dev_alloc(c, c->stage, n * realkind)) return rc;
    const void* cur = c->buf[c->cur];
    const int g = c->pending ? c->cur : c->cur ^ 1;
    HIPCHK(c, c->ops->state_out(c->lay, cur, c->buf[g], c->pending ? 0 : 1, c->ghost_raw[g] ? 1 : 0, c->stage.p,
                                nullptr, nullptr, nullptr, nullptr, realkind, c->stream));
    HIPCHK(c, hipMemcpyAsync(f, c->stage.p, n * realkind, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_macros(lbm_ctx* c, void* rho, void* u, void* v, void* w, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!rho || !u || !v || !w) return fail(c, LBM_EINVAL, "null argument");
    if (int rc = check_no_part(c)) return rc;
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state");
    if (int rc = check_kind(c, realkind)) return rc;
    const size_t nc = ncell_padded(c);
    const size_t cb = nc * realkind;
    if (int rc = dev_alloc(c, c->stage, 4 * cb)) return rc;
    char* st = static_cast<char*>(c->stage.p);
    const int g = c->pending ? c->cur : c->cur ^ 1;
    HIPCHK(c, c->ops->state_out(c->lay, c->buf[c->cur], c->buf[g], c->pending ? 0 : 1, c->ghost_raw[g] ? 1 : 0,
                                nullptr, st, st + cb, st + 2 * cb, st + 3 * cb, realkind, c->stream));
    HIPCHK(c, hipMemcpyAsync(rho, st, cb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(u, st + cb, cb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(v, st + 2 * cb, cb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(w, st + 3 * cb, cb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the instability test of src/m_diag.F90:169-173 on the moments just returned (all cells)
    size_t bad = 0, nonpos = 0;
    double rmin = INFINITY;
    const void* arr[4] = {rho, u, v, w};
    for (int q = 0; q < 4; ++q)
        for (size_t e = 0; e < nc; ++e) {
            const double x = realkind == 8 ? static_cast<const double*>(arr[q])[e] : static_cast<const float*>(arr[q])[e];
            if (!std::isfinite(x)) ++bad;
            else if (q == 0) {
                rmin = std::min(rmin, x);
                if (x <= 0.0) ++nonpos;
            }
        }
    if (bad || nonpos) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "unstable state: %zu non-finite moments, %zu cells with density <= 0 (smallest %g)",
                      bad, nonpos, rmin);
        return fail(c, LBM_EUNSTABLE, msg);
    }
    return LBM_OK;
}

int lbm_get_tau(lbm_ctx* c, void* tau, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!tau) return fail(c, LBM_EINVAL, "null argument");
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    const Layout& L = c->lay;
    const size_t nc = ncell_padded(c);
    std::vector<double> out(nc, c->tau0);
    if (c->d.store_tau && c->d.ivreman && c->tau.p) {
        const int rb = c->ops->real_bytes;
        std::vector<unsigned char> dev((size_t)L.cplane * rb);
        HIPCHK(c, hipMemcpyAsync(dev.data(), c->tau.p, dev.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int k = 1; k <= L.nz; ++k)
            for (int j = 1; j <= L.ny; ++j)
                for (int i = 1; i <= L.nx; ++i) {
                    const size_t o = (size_t)L.cidx(i, j, k);
                    out[((size_t)k * (L.ny + 2) + j) * (L.nx + 2) + i] = get_real(dev.data(), o, rb);
                }
    }
    for (size_t q = 0; q < nc; ++q) {
        if (realkind == 8) static_cast<double*>(tau)[q] = out[q];
        else static_cast<float*>(tau)[q] = (float)out[q];
    }
    return LBM_OK;
}

int lbm_set_blanking(lbm_ctx* c, const uint8_t* mask) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (!mask) {
        c->has_solid = false;
        return LBM_OK;
    }
    const lbm_desc& d = c->d;
    const Layout& L = c->lay;
    // one bit per interior cell, rows of bits along the contiguous axis ax0
    const int n0 = L.n(L.ax0), n1 = L.n(L.ax1), n2 = L.n(L.ax2);
    const int wpr = (n0 + 31) / 32;
    std::vector<uint32_t> bits((size_t)wpr * n1 * n2, 0u);
    bool any = false;
    for (int cc = 1; cc <= n2; ++cc)
        for (int cb = 1; cb <= n1; ++cb)
            for (int ca = 1; ca <= n0; ++ca) {
                int i, j, k;
                L.ijk(ca, cb, cc, i, j, k);
                if (mask[((size_t)k * (d.ny + 2) + j) * (d.nx + 2) + i]) {
                    bits[((size_t)(cc - 1) * n1 + (cb - 1)) * wpr + (ca - 1) / 32] |= 1u << ((ca - 1) % 32);
                    any = true;
                }
            }
    if (int rc = dev_alloc(c, c->solid, bits.size() * 4)) return rc;
    HIPCHK(c, hipMemcpyAsync(c->solid.p, bits.data(), bits.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->solid_wpr = wpr;
    c->has_solid = any;   // the reference's lsolids (src/main.F90:310)
    return LBM_OK;
}

int lbm_set_ablvisc(lbm_ctx* c, const void* abl, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!abl) return fail(c, LBM_EINVAL, "null argument");
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    return upload_reals(c, c->ablvisc, abl, c->d.nz, realkind);
}

int lbm_set_inflow(lbm_ctx* c, const void* shear, double uini, double udir, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    c->uini = uini;
    c->udir = udir;
    if (shear) return upload_reals(c, c->shear, shear, c->d.nz, realkind);
    std::vector<double> ones(c->d.nz, 1.0);
    return upload_reals(c, c->shear, ones.data(), c->d.nz, 8);
}

int lbm_set_inflow_series(lbm_ctx* c, int it_lo, int it_hi, const void* ut, const void* ud, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (int rc = check_kind(c, realkind)) return rc;
    c->ut_series.clear();
    c->ud_series.clear();
    if (!ut || !ud || it_hi < it_lo) return LBM_OK;
    const int n = it_hi - it_lo + 1;
    c->series_lo = it_lo;
    c->series_hi = it_hi;
    for (int q = 0; q < n; ++q) {
        c->ut_series.push_back(realkind == 8 ? static_cast<const double*>(ut)[q] : static_cast<const float*>(ut)[q]);
        c->ud_series.push_back(realkind == 8 ? static_cast<const double*>(ud)[q] : static_cast<const float*>(ud)[q]);
    }
    return LBM_OK;
}

int lbm_set_turbulence(lbm_ctx* c, const void* uu, const void* vv, const void* ww, int nrturb, double ampl,
                       int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (!uu) {
        c->has_turb = false;
        return LBM_OK;
    }
    if (!vv || !ww || nrturb < 1) return fail(c, LBM_EINVAL, "turbulence fields and nrturb >= 1 required");
    if (int rc = check_kind(c, realkind)) return rc;
    if (c->d.iturb_pos < 1 || c->d.iturb_pos > c->d.nx) return fail(c, LBM_EINVAL, "iturb_pos outside the grid");
    const size_t n = (size_t)c->d.ny * c->d.nz * (nrturb + 1);
    const void* src[3] = {uu, vv, ww};
    for (int q = 0; q < 3; ++q)
        if (int rc = upload_reals(c, c->turb[q], src[q], n, realkind)) return rc;
    c->nrturb = nrturb;
    c->ampl = round_real(c, ampl);
    c->has_turb = true;
    return LBM_OK;
}

int lbm_generate_turbulence(lbm_ctx* c, int nrturb, double ampl, uint64_t seed, void* uu, void* vv, void* ww, void* rr,
                            int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (nrturb < 1) return fail(c, LBM_EINVAL, "nrturb must be >= 1");
    if ((uu || vv || ww || rr) && check_kind(c, realkind)) return LBM_EINVAL;
    const int rb = c->ops->real_bytes;
    if (rb != 4 && rb != 8) return fail(c, LBM_EINVAL, "turbulence generation needs FP32/FP64 compute (not the tangent storage)");
    if (c->d.iturb_pos < 1 || c->d.iturb_pos > c->d.nx) return fail(c, LBM_EINVAL, "iturb_pos outside the grid");
    const lbm_desc& d = c->d;
    const size_t n = (size_t)d.ny * d.nz * (nrturb + 1);
    DevBuf scratch, out;
    int rc = LBM_OK;
    if ((rc = dev_alloc(c, scratch, turbulence_scratch_bytes(d.ny, d.nz, nrturb))) ||
        (rc = dev_alloc(c, out, 4 * n * rb))) {
        dev_free(scratch);
        dev_free(out);
        return rc;
    }
    auto run = [&]() -> int {
        HIPCHK(c, turbulence_generate(rb, d.ny, d.nz, nrturb, seed, scratch.p, out.p, c->stream));
        for (int q = 0; q < 3; ++q) {
            if (int r = dev_alloc(c, c->turb[q], n * rb)) return r;
            HIPCHK(c, hipMemcpyAsync(c->turb[q].p, static_cast<char*>(out.p) + q * n * rb, n * rb,
                                     hipMemcpyDeviceToDevice, c->stream));
        }
        void* host[4] = {uu, vv, ww, rr};
        std::vector<unsigned char> img(n * rb);
        for (int q = 0; q < 4; ++q) {
            if (!host[q]) continue;
            HIPCHK(c, hipMemcpyAsync(img.data(), static_cast<char*>(out.p) + q * n * rb, n * rb, hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (size_t e = 0; e < n; ++e) {
                const double v = get_real(img.data(), e, rb);
                if (realkind == 8) static_cast<double*>(host[q])[e] = v;
                else static_cast<float*>(host[q])[e] = (float)v;
            }
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return LBM_OK;
    };
    rc = run();
    dev_free(scratch);
    dev_free(out);
    if (rc) return rc;
    c->nrturb = nrturb;
    c->ampl = round_real(c, ampl);
    c->has_turb = true;
    return LBM_OK;
}

int lbm_set_external_forcing(lbm_ctx* c, const void* force, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (c->turb_on) return fail(c, LBM_EINVAL, "the turbines write the external forcing: remove them first");
    if (!force) {
        c->has_force = false;
        return LBM_OK;
    }
    if (int rc = check_kind(c, realkind)) return rc;
    const Layout& L = c->lay;
    const size_t nc = ncell_padded(c);
    const int rb = c->ops->real_bytes;
    for (int q = 0; q < 3; ++q) {
        std::vector<unsigned char> plane((size_t)L.cplane * rb, 0);
        for (int k = 0; k <= L.nz + 1; ++k)
            for (int j = 0; j <= L.ny + 1; ++j)
                for (int i = 0; i <= L.nx + 1; ++i) {
                    const size_t hc = ((size_t)k * (L.ny + 2) + j) * (L.nx + 2) + i;
                    const double v = realkind == 8 ? static_cast<const double*>(force)[3 * hc + q]
                                                   : static_cast<const float*>(force)[3 * hc + q];
                    put_real(plane.data(), (size_t)L.cidx(i, j, k), v, rb);
                }
        if (int rc = dev_alloc(c, c->force[q], plane.size())) return rc;
        HIPCHK(c, hipMemcpyAsync(c->force[q].p, plane.data(), plane.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    (void)nc;
    c->has_force = true;
    return LBM_OK;
}

int lbm_set_buoyancy(lbm_ctx* c, const void* pottemp, double theta0, double scaling, int realkind) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (!pottemp) {
        c->buoy_on = false;
        return LBM_OK;
    }
    if (int rc = check_kind(c, realkind)) return rc;
    if (!(theta0 != 0.0) || !std::isfinite(theta0) || !std::isfinite(scaling))
        return fail(c, LBM_EINVAL, "buoyancy: theta0 must be finite and non-zero, scaling finite");
    // the reference layout (0:nx+1, 0:ny+1, 0:nz+1) -> the cell layout of the compute type
    const Layout& L = c->lay;
    const int rb = c->ops->real_bytes;
    std::vector<unsigned char> img((size_t)L.cplane * rb, 0);
    for (int k = 0; k <= L.nz + 1; ++k)
        for (int j = 0; j <= L.ny + 1; ++j)
            for (int i = 0; i <= L.nx + 1; ++i) {
                const size_t hc = ((size_t)k * (L.ny + 2) + j) * (L.nx + 2) + i;
                const double v = realkind == 8 ? static_cast<const double*>(pottemp)[hc]
                                               : static_cast<const float*>(pottemp)[hc];
                put_real(img.data(), (size_t)L.cidx(i, j, k), v, rb);
            }
    if (int rc = dev_alloc(c, c->theta, img.size())) return rc;
    HIPCHK(c, hipMemcpyAsync(c->theta.p, img.data(), img.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->btheta0 = round_real(c, theta0);
    c->bscale = round_real(c, scaling);
    c->buoy_on = true;
    return LBM_OK;
}

int lbm_synchronize(lbm_ctx* c) {
    if (int rc_ = check_ctx(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

void* lbm_stream(lbm_ctx* c) { return check_ctx(c) == LBM_OK ? (void*)c->stream : nullptr; }

int lbm_halo_bytes(lbm_ctx* c, size_t* bytes) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!bytes) return fail(c, LBM_EINVAL, "null argument");
    if (c->d.jbnd != LBM_BND_HALO) return fail(c, LBM_EINVAL, "not a halo domain (jbnd != LBM_BND_HALO)");
    *bytes = (size_t)(c->d.q == 27 ? 9 : 5) * (c->d.nx + 2) * (c->d.nz + 2) * c->ops->raw_bytes;
    return LBM_OK;
}

namespace {
int halo_check(lbm_ctx* c, int side, const void* buf) {
    if (!buf) return fail(c, LBM_EINVAL, "null argument");
    if (c->d.jbnd != LBM_BND_HALO) return fail(c, LBM_EINVAL, "not a halo domain (jbnd != LBM_BND_HALO)");
    if (side < 0 || side > 2) return fail(c, LBM_EINVAL, "side must be 0, 1 or 2 (both)");
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state: call lbm_set_state first");
    return LBM_OK;
}
}  // namespace

int lbm_halo_pack(lbm_ctx* c, int side, void* dev_buf) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = halo_check(c, side, dev_buf)) return rc;
    // the state of the last step (or the uploaded state, which both buffers hold)
    HIPCHK(c, c->ops->halo_pack(c->lay, c->buf[c->cur], side, dev_buf, c->stream));
    return LBM_OK;
}

int lbm_halo_unpack(lbm_ctx* c, int side, const void* dev_buf) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = halo_check(c, side, dev_buf)) return rc;
    HIPCHK(c, c->ops->halo_unpack(c->lay, c->buf[c->cur], side, dev_buf, c->fold_on ? c->fold_dir : -1, c->stream));
    return LBM_OK;
}

int lbm_set_timing(lbm_ctx* c, int on) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (on < 0) return fail(c, LBM_EINVAL, "timing interval < 0");
    c->timing = on != 0;
    c->timing_every = on > 1 ? on : 1;
    c->timing_seq = 0;
    c->n_step_ev = c->n_bc_ev = 0;
    return LBM_OK;
}

int lbm_timing_read(lbm_ctx* c, double* step_ms, int* n_step, double* bc_ms, int* n_bc) {
    if (int rc_ = check_ctx(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s = 0, b = 0;
    for (size_t q = 0; q + 1 < c->n_step_ev; q += 2) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_step[q], c->ev_step[q + 1]));
        s += ms;
    }
    for (size_t q = 0; q + 1 < c->n_bc_ev; q += 2) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_bc[q], c->ev_bc[q + 1]));
        b += ms;
    }
    if (step_ms) *step_ms = s;
    if (n_step) *n_step = (int)(c->n_step_ev / 2);
    if (bc_ms) *bc_ms = b;
    if (n_bc) *n_bc = (int)(c->n_bc_ev / 2);
    c->n_step_ev = c->n_bc_ev = 0;
    return LBM_OK;
}

int lbm_get_totals(lbm_ctx* c, double* mass, double* mom) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state");
    const int nb = totals_blocks((long long)c->d.nx * c->d.ny * c->d.nz);
    if (int rc = dev_alloc(c, c->partial, (size_t)nb * kTotalsSlots * sizeof(double))) return rc;
    HIPCHK(c, c->ops->totals(c->lay, c->buf[c->cur], c->pending ? 0 : 1, static_cast<double*>(c->partial.p), nb,
                             c->stream));
    std::vector<double> h((size_t)nb * kTotalsSlots);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return finish_totals(c, h.data(), nb, mass, mom);
}

int lbm_get_totals_noalloc(lbm_ctx* c, double* mass, double* mom) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state");
    // the totals kernel grid-strides, so every grid needs at most kTotalsMaxBlocks partial results
    const int nb = totals_blocks((long long)c->d.nx * c->d.ny * c->d.nz);
    if (nb < 1 || nb > kTotalsMaxBlocks)
        return fail(c, LBM_EINVAL, "totals: " + std::to_string(nb) + " partial results exceed the scratch of " +
                                       std::to_string(kTotalsMaxBlocks));
    double* p = nullptr;
    hipError_t pe = hipErrorInvalidDevice;
    if (c->device >= 0 && c->device < kMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        p = g_scratch[c->device];
        pe = g_scratch_err[c->device];
    }
    if (!p) return hipfail(c, pe, "hipGetSymbolAddress(scratch) for this context's device");
    // every block writes all of its slots, so nothing left by an earlier call is read
    HIPCHK(c, c->ops->totals(c->lay, c->buf[c->cur], c->pending ? 0 : 1, p, nb, c->stream));
    double h[kTotalsMaxBlocks * kTotalsSlots];
    HIPCHK(c, hipMemcpyAsync(h, p, (size_t)nb * kTotalsSlots * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return finish_totals(c, h, nb, mass, mom);
}

int lbm_get_info(lbm_ctx* c, int64_t* bytes, int32_t* pitch, int64_t* plane) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (bytes) *bytes = (int64_t)(2 * c->buf_bytes);
    if (pitch) *pitch = c->lay.pitch;
    if (plane) *plane = c->lay.cplane;
    return LBM_OK;
}

}  // extern "C"
