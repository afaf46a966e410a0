// Objective and adjoint entry points (lbm_objective_run, lbm_adjoint_run in include/lbm_hip.h).
// No counterpart in the reference: they differentiate its time loop.
//
// The adjoint run advances the context nsteps steps from its current state, storing every `every`
// steps a checkpoint (the state and the ghost shell of the state before it).  The reverse sweep
// walks the segments from last to first: it recomputes the segment's states from its checkpoint
// (kept as full buffers, ghosts included: they are the inputs of the steps to reverse) and reverses
// the steps with the kernels of adjoint.h.  The sweep leaves the context's buffers alone, so the
// context continues from the forward state after nsteps as if lbm_step had been called.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "ctx.h"
#include "dual.h"
#include "impl.h"
#include "layout.h"

using namespace lbm;
using namespace lbm_host;

namespace {

int check_box(lbm_ctx* c, const int32_t* box, AdjBox* b) {
    if (!box) return fail(c, LBM_EINVAL, "null box");
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    for (int d = 0; d < 3; ++d) {
        b->lo[d] = box[2 * d];
        b->hi[d] = box[2 * d + 1];
        if (b->lo[d] < 1 || b->hi[d] > n[d] || b->lo[d] > b->hi[d])
            return fail(c, LBM_EINVAL, "box must satisfy 1 <= lo <= hi <= n on every axis");
    }
    return LBM_OK;
}

double box_cells(const AdjBox& b) {
    return (double)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}

int copy_dev(lbm_ctx* c, void* dst, const void* src, size_t n) {
    HIPCHK(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, c->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_objective_run(lbm_ctx* c, int it0, int nsteps, const int32_t* box, double* J) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!J) return fail(c, LBM_EINVAL, "null argument");
    if (nsteps < 1) return fail(c, LBM_EINVAL, "nsteps must be >= 1");
    if (c->turb_on || c->buoy_on) return fail(c, LBM_EINVAL, "objective runs take no turbines or buoyancy");
    AdjBox b;
    if (int rc = check_box(c, box, &b)) return rc;
    DevBuf acc;
    const int nblk = objective_blocks(b);
    const size_t accb = 2 * (size_t)nblk * sizeof(double);
    if (int rc = dev_alloc(c, acc, accb)) return rc;
    int rc = LBM_OK;
    if (hipMemsetAsync(acc.p, 0, accb, c->stream) != hipSuccess) rc = fail(c, LBM_EHIP, "hipMemsetAsync");
    for (int n = 1; n <= nsteps && rc == LBM_OK; ++n) {
        if ((rc = lbm_step(c, it0 + n - 1, 1))) break;
        const hipError_t e = c->ops->objective(c->lay, c->buf[c->cur], b, static_cast<double*>(acc.p), c->stream);
        if (e != hipSuccess) rc = hipfail(c, e, "objective kernel");
    }
    double s[2] = {0.0, 0.0};
    if (rc == LBM_OK) {
        std::vector<double> h(2 * (size_t)nblk);
        if (hipMemcpyAsync(h.data(), acc.p, accb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(c, LBM_EHIP, "objective read-back");
        for (int q = 0; q < nblk; ++q) {
            s[0] += h[2 * q];
            s[1] += h[2 * q + 1];
        }
    }
    dev_free(acc);
    if (rc == LBM_OK) {
        const double w = 1.0 / ((double)nsteps * box_cells(b));
        *J = s[0] * w;
        c->obj_tangent = s[1] * w;
    }
    return rc;
}

int lbm_objective_tangent(lbm_ctx* c, double* dJ) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!dJ) return fail(c, LBM_EINVAL, "null argument");
    if (c->d.storage != LBM_STORE_F64T) return fail(c, LBM_EINVAL, "tangents need LBM_STORE_F64T storage");
    *dJ = c->obj_tangent;
    return LBM_OK;
}

int lbm_set_tangent(lbm_ctx* c, const double* dforce, const double* dshear) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    const lbm_desc& d = c->d;
    if (d.storage != LBM_STORE_F64T) return fail(c, LBM_EINVAL, "tangents need LBM_STORE_F64T storage");
    if (dforce && !c->has_force) return fail(c, LBM_ESTATE, "set the external forcing before its tangent");
    const Layout& L = c->lay;
    // the value parts stay; only the derivative parts are replaced
    if (c->has_force)
        for (int q = 0; q < 3; ++q) {
            std::vector<Dual> img((size_t)L.cplane);
            HIPCHK(c, hipMemcpyAsync(img.data(), c->force[q].p, img.size() * sizeof(Dual), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (auto& x : img) x.d = 0.0;
            if (dforce)
                for (int k = 0; k <= L.nz + 1; ++k)
                    for (int j = 0; j <= L.ny + 1; ++j)
                        for (int i = 0; i <= L.nx + 1; ++i)
                            img[(size_t)L.cidx(i, j, k)].d = dforce[3 * (((size_t)k * (L.ny + 2) + j) * (L.nx + 2) + i) + q];
            HIPCHK(c, hipMemcpyAsync(c->force[q].p, img.data(), img.size() * sizeof(Dual), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
    std::vector<Dual> sh((size_t)d.nz);
    HIPCHK(c, hipMemcpyAsync(sh.data(), c->shear.p, sh.size() * sizeof(Dual), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < d.nz; ++k) sh[k].d = dshear ? dshear[k] : 0.0;
    HIPCHK(c, hipMemcpyAsync(c->shear.p, sh.data(), sh.size() * sizeof(Dual), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_adjoint_run(lbm_ctx* c, int it0, int nsteps, int every, const int32_t* box, double* J, double* grad_force,
                    double* grad_shear) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!J) return fail(c, LBM_EINVAL, "null argument");
    const lbm_desc& d = c->d;
    if (!c->have_state) return fail(c, LBM_ESTATE, "no state");
    if (c->pending || c->use_ext)
        return fail(c, LBM_ESTATE, "the adjoint starts from a stepped state: call lbm_step once after set_state / init_equilibrium");
    if (!c->ops->adj_step) return fail(c, LBM_EINVAL, "the adjoint needs FP32 or FP64 storage");
    if (c->has_turb) return fail(c, LBM_EINVAL, "the adjoint does not differentiate the inflow-turbulence forcing: switch it off");
    if (c->turb_on) return fail(c, LBM_EINVAL, "the adjoint does not differentiate the turbines: remove them");
    if (c->buoy_on) return fail(c, LBM_EINVAL, "the adjoint does not differentiate the buoyancy forcing: switch it off");
    if (c->avg_on) return fail(c, LBM_ESTATE, "switch time averaging off before an adjoint run");
    const int closed = (d.ibnd > 10) + (d.jbnd > 10) + (d.kbnd > 10);
    if (c->d.jbnd == LBM_BND_HALO) return fail(c, LBM_EINVAL, "the adjoint takes untiled domains (no halo planes)");
    if (closed > 1 || (closed == 1 && c->fold_dir < 0))
        return fail(c, LBM_EINVAL, "the adjoint supports at most one closed axis (its face pass runs inside the step kernel)");
    if (nsteps < 1 || every < 1) return fail(c, LBM_EINVAL, "nsteps and every must be >= 1");
    // the reverse kernel splits its flat index over a plane of the two fastest storage axes in FP32
    if ((long long)c->lay.n(c->lay.ax0) * c->lay.n(c->lay.ax1) >= (1LL << 24))
        return fail(c, LBM_EINVAL, "the adjoint takes at most 2^24 cells in a plane of the two fastest storage axes");
    AdjBox b;
    if (int rc = check_box(c, box, &b)) return rc;

    const size_t bb = c->buf_bytes;
    const int nck = (nsteps + every - 1) / every;
    const size_t plane = (size_t)c->lay.cplane;   // per-cell arrays (force gradient)
    // the workspace stays with the context (a gradient loop calls this repeatedly); a failed
    // allocation frees it all
    DevBuf &ck = c->adj_ck, &snaps = c->adj_snaps, &lam = c->adj_lam, &gF = c->adj_gF, &gS = c->adj_gS,
           &acc = c->adj_acc;
    auto release = [&]() {
        for (DevBuf* x : {&ck, &snaps, &lam, &gF, &gS, &acc}) dev_free(*x);
    };
    int rc = LBM_OK;
    const bool want_gF = grad_force && c->has_force;
    // a checkpoint: X(n0) and the ghost shell of X(n0-1); the segment's states X(n0+1) .. X(n1-1) in
    // `every - 1` slots (X(n0) is read from the checkpoint)
    if ((rc = dev_alloc(c, ck, (size_t)nck * 2 * bb)) || (rc = dev_alloc(c, snaps, (size_t)std::max(every - 1, 1) * bb)) ||
        (rc = dev_alloc(c, lam, 2 * bb)) ||
        (rc = dev_alloc(c, acc, 2 * (size_t)objective_blocks(b) * sizeof(double))) ||
        (rc = dev_alloc(c, gS, (size_t)d.ny * d.nz * sizeof(double))) ||
        (want_gF && (rc = dev_alloc(c, gF, 3 * plane * sizeof(double))))) {
        release();
        return rc;
    }
    double* gFp = want_gF ? static_cast<double*>(gF.p) : nullptr;
    char* ckp = static_cast<char*>(ck.p);
    char* snp = static_cast<char*>(snaps.p);
    char* lmp = static_cast<char*>(lam.p);
    std::vector<int> ck_cur(nck, 0);
    auto run = [&]() -> int {
        const int nblk = objective_blocks(b);
        HIPCHK(c, hipMemsetAsync(acc.p, 0, 2 * (size_t)nblk * sizeof(double), c->stream));
        // ---- forward sweep with checkpoints and the objective
        for (int n = 1; n <= nsteps; ++n) {
            if ((n - 1) % every == 0) {
                const int q = (n - 1) / every;
                ck_cur[q] = c->cur;   // the ping-pong parity of X(n0), which the adjoint buffers follow
                if (int r = copy_dev(c, ckp + (size_t)q * 2 * bb, c->buf[c->cur], bb)) return r;
                HIPCHK(c, c->ops->ghost_copy(c->lay, c->buf[c->cur ^ 1], ckp + (size_t)q * 2 * bb + bb, c->stream));
            }
            if (int r = lbm_step(c, it0 + n - 1, 1)) return r;
            HIPCHK(c, c->ops->objective(c->lay, c->buf[c->cur], b, static_cast<double*>(acc.p), c->stream));
        }
        // the reverse sweep reads the context's buffers but writes only the workspace, so the forward
        // state after nsteps stays where lbm_step left it
        const int fin_cur = c->cur;

        // ---- reverse sweep
        HIPCHK(c, hipMemsetAsync(lam.p, 0, 2 * bb, c->stream));
        HIPCHK(c, hipMemsetAsync(gS.p, 0, (size_t)d.ny * d.nz * sizeof(double), c->stream));
        if (gFp) HIPCHK(c, hipMemsetAsync(gFp, 0, 3 * plane * sizeof(double), c->stream));
        const double kappa = 1.0 / ((double)nsteps * box_cells(b));
        StepHost h = step_host(c, false);
        h.stream = 1;
        h.tau_out = nullptr;
        for (int q = 0; q < 4; ++q) h.ext[q] = nullptr;
        h.turb[0] = h.turb[1] = h.turb[2] = nullptr;
        // the objective's last term: moments of X(N)
        h.src = c->buf[fin_cur];
        HIPCHK(c, c->ops->adj_step(c->lay, h, nullptr, lmp + (size_t)fin_cur * bb, b, kappa, 0, nullptr, c->stream));
        int opt_lo = 0, opt_hi = 0;
        if (c->fold_dir >= 0) {
            const int code = c->fold_dir == 0 ? d.ibnd : c->fold_dir == 1 ? d.jbnd : d.kbnd;
            opts_of(code, &opt_lo, &opt_hi);
        }
        for (int seg = nck - 1; seg >= 0; --seg) {
            const int n0 = seg * every, n1 = std::min(n0 + every, nsteps);
            // the inputs X(n0) .. X(n1-1) of the segment's steps, slot m holding X(n0+m): slot 0 is
            // the checkpoint's X(n0) (read in place); step n0+m is written straight into slot m after
            // that slot received the ghost shell of X(n0+m-2) (slot m-2, or the checkpoint's shell of
            // X(n0-1)), which is what the buffer receiving that step holds in the forward run's
            // ping-pong
            char* ckx = ckp + (size_t)seg * 2 * bb;
            char* x_before = ckx + bb;
            auto slot = [&](int m) { return m == 0 ? ckx : snp + (size_t)(m - 1) * bb; };
            for (int n = n0 + 1; n < n1; ++n) {
                const int m = n - n0;
                char* dst = slot(m);
                HIPCHK(c, c->ops->ghost_copy(c->lay, m >= 2 ? slot(m - 2) : x_before, dst, c->stream));
                if (int r = step_buffers(c, slot(m - 1), dst, it0 + n)) return r;
            }
            for (int n = n1; n > n0; --n) {
                const int bd = ck_cur[seg] ^ ((n - n0) & 1), bs = bd ^ 1;
                char* ly = lmp + (size_t)bd * bb;
                char* lx = lmp + (size_t)bs * bb;
                if (d.ibnd == 1) {
                    double ut, cs, sn;
                    inflow_values(c, it0 + n, &ut, &cs, &sn);
                    HIPCHK(c, c->ops->adj_inflow(c->lay, ly, ut, cs, sn, d.rho0, static_cast<double*>(gS.p), c->stream));
                }
                if (c->fold_dir >= 0) HIPCHK(c, c->ops->adj_face(c->lay, ly, lx, c->fold_dir, opt_lo, opt_hi, c->stream));
                h.src = slot(n - 1 - n0);
                HIPCHK(c, c->ops->adj_step(c->lay, h, ly, lx, b, n - 1 >= 1 ? kappa : 0.0, 1, gFp,
                                           c->stream));
            }
        }
        // results
        double s = 0.0;
        std::vector<double> hs((size_t)d.ny * d.nz), hf, ha(2 * (size_t)nblk);
        HIPCHK(c, hipMemcpyAsync(ha.data(), acc.p, ha.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(hs.data(), gS.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (gFp) {
            hf.resize(3 * plane);
            HIPCHK(c, hipMemcpyAsync(hf.data(), gFp, hf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int q = 0; q < nblk; ++q) s += ha[2 * q];
        *J = s * kappa;
        if (grad_shear)
            for (int k = 0; k < d.nz; ++k) {
                double t = 0.0;
                for (int j = 0; j < d.ny; ++j) t += hs[(size_t)k * d.ny + j];
                grad_shear[k] = t;
            }
        if (grad_force) {
            const size_t nxp = d.nx + 2, nyp = d.ny + 2, nzp = d.nz + 2;
            std::fill(grad_force, grad_force + 3 * nxp * nyp * nzp, 0.0);
            if (gFp)
                for (int k = 1; k <= d.nz; ++k)
                    for (int j = 1; j <= d.ny; ++j)
                        for (int i = 1; i <= d.nx; ++i) {
                            const size_t o = c->lay.cidx(i, j, k);
                            double* g = grad_force + 3 * ((k * nyp + j) * nxp + i);
                            g[0] = hf[o];
                            g[1] = hf[plane + o];
                            g[2] = hf[2 * plane + o];
                        }
        }
        return LBM_OK;
    };
    return run();
}

}  // extern "C"
