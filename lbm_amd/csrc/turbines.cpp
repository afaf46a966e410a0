// Actuator-line turbines (lbm_set_turbines / lbm_get_turbine_theta in include/lbm_hip.h): the
// host side of turbines.h.  The reference builds its actuator points on the host every step
// (src/m_turbine_forcing.F90:72-76); here the tables are uploaded once and each step's points,
// forces and deposit are computed on the device from the step's input (lbm_step launches them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "ctx.h"
#include "impl.h"

using namespace lbm;
using namespace lbm_host;

namespace {

// Host image of n compute-type values.
std::vector<unsigned char> image(const std::vector<double>& v, int rb) {
    std::vector<unsigned char> img(v.size() * rb);
    for (size_t q = 0; q < v.size(); ++q) put_real(img.data(), q, v[q], rb);
    return img;
}

int upload(lbm_ctx* c, DevBuf& b, const void* host, size_t bytes) {
    if (int rc = dev_alloc(c, b, bytes > 0 ? bytes : 1)) return rc;
    HIPCHK(c, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_set_turbines(lbm_ctx* c, const lbm_turbines* t) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (int rc = check_no_part(c)) return rc;
    if (!t) {
        c->turb_on = false;
        return LBM_OK;
    }
    const lbm_desc& d = c->d;
    const int rb = c->ops->real_bytes;
    if (rb != 4 && rb != 8) return fail(c, LBM_EINVAL, "turbines need FP32 or FP64 compute (not the tangent storage)");
    if (d.jbnd == LBM_BND_HALO) return fail(c, LBM_EINVAL, "turbines need the whole domain (no halo tiles)");
    if (c->has_force) return fail(c, LBM_EINVAL, "the turbines write the external forcing: remove the caller's field first");
    if (t->nturbines < 1 || t->nblades < 1 || t->nchords < 1 || t->nfoils < 1 || t->nang_max < 2)
        return fail(c, LBM_EINVAL, "turbines: nturbines, nblades, nchords, nfoils >= 1 and nang_max >= 2 required");
    if (!t->hub || !t->yaw || !t->tilt || !t->omegand || !t->pitch || !t->theta || !t->relm || !t->dc || !t->chord ||
        !t->twist || !t->nfoil || !t->nang || !t->foil)
        return fail(c, LBM_EINVAL, "turbines: null table");
    const long long npts = (long long)t->nturbines * t->nblades * t->nchords;
    if (npts > (1 << 20)) return fail(c, LBM_EINVAL, "turbines: more than 2^20 actuator points");
    for (int m = 0; m < 3 * t->nturbines; ++m)
        if (!(std::fabs(t->hub[m]) < 1e6)) return fail(c, LBM_EINVAL, "turbines: hub position not finite or beyond 1e6 cells");
    for (int k = 0; k < t->nfoils; ++k)
        if (t->nang[k] < 0 || t->nang[k] > t->nang_max) return fail(c, LBM_EINVAL, "turbines: nang[k] outside 0..nang_max");
    for (int q = 0; q < t->nchords; ++q)
        if (t->nfoil[q] < 1 || t->nfoil[q] > t->nfoils) return fail(c, LBM_EINVAL, "turbines: nfoil outside 1..nfoils");
    const int n[3] = {d.nx, d.ny, d.nz};
    auto R = [&](double v) { return round_real(c, v); };
    // Rotor table: hub, rotor basis (src/m_turbine_rotor_basis.F90:19-21, evaluated once here in the
    // compute precision), omegand, pitch.  Deposit box of each rotor: every point lies within
    // relm_max * |(e1_d, e2_d)| of the hub along axis d; its stencil reaches krad cells beyond.
    const double sigma = R(R(2.0) / R(std::sqrt(R(2.0))));
    const int krad = std::min(5, (int)std::ceil(R(R(3.0) * sigma)));   // src/m_turbine_deposit.F90:36-38
    double relm_max = 0;
    for (int q = 0; q < t->nchords; ++q) relm_max = std::max(relm_max, std::fabs(R(t->relm[q])));
    std::vector<double> rot((size_t)t->nturbines * kTurbRot, 0.0);
    std::vector<int> box((size_t)t->nturbines * 6);
    int lo[3] = {n[0], n[1], n[2]}, hi[3] = {1, 1, 1};
    for (int m = 0; m < t->nturbines; ++m) {
        double* r = &rot[(size_t)m * kTurbRot];
        double cy, sy, ctl, stl;
        if (rb == 4) {
            const float yaw = (float)t->yaw[m], tilt = (float)t->tilt[m];
            cy = cosf(yaw); sy = sinf(yaw); ctl = cosf(tilt); stl = sinf(tilt);
        } else {
            cy = std::cos(t->yaw[m]); sy = std::sin(t->yaw[m]); ctl = std::cos(t->tilt[m]); stl = std::sin(t->tilt[m]);
        }
        const double e_axis[3] = {R(cy * ctl), R(sy * ctl), -stl};
        const double e1[3] = {-sy, cy, 0.0};
        const double e2[3] = {R(-cy * stl), R(-sy * stl), -ctl};
        for (int a = 0; a < 3; ++a) {
            r[a] = R(t->hub[3 * m + a]);
            r[3 + a] = e_axis[a];
            r[6 + a] = e1[a];
            r[9 + a] = e2[a];
            const double reach = relm_max * std::sqrt(e1[a] * e1[a] + e2[a] * e2[a]);
            const int blo = std::max(1, (int)std::floor(r[a] - reach) - krad - 1);
            const int bhi = std::min(n[a], (int)std::floor(r[a] + reach) + krad + 1);
            box[6 * m + 2 * a] = blo;
            box[6 * m + 2 * a + 1] = bhi;
            lo[a] = std::min(lo[a], blo);
            hi[a] = std::max(hi[a], bhi);
        }
        r[12] = R(t->omegand[m]);
        r[13] = R(t->pitch[m]);
    }
    for (int a = 0; a < 3; ++a)
        if (lo[a] > hi[a]) { lo[a] = 1; hi[a] = 0; }
    std::vector<double> blade((size_t)t->nchords * 4), theta(t->theta, t->theta + t->nturbines);
    std::vector<int> nfoil(t->nchords), nang(t->nang, t->nang + t->nfoils);
    for (int q = 0; q < t->nchords; ++q) {
        blade[4 * q] = t->relm[q];
        blade[4 * q + 1] = t->dc[q];
        blade[4 * q + 2] = t->chord[q];
        blade[4 * q + 3] = t->twist[q];
        nfoil[q] = t->nfoil[q] - 1;
    }
    std::vector<double> foil(t->foil, t->foil + (size_t)t->nfoils * t->nang_max * 3);
    int rc = LBM_OK;
    auto img_rot = image(rot, rb), img_blade = image(blade, rb), img_theta = image(theta, rb), img_foil = image(foil, rb);
    if ((rc = upload(c, c->t_rot, img_rot.data(), img_rot.size())) || (rc = upload(c, c->t_theta, img_theta.data(), img_theta.size())) ||
        (rc = upload(c, c->t_blade, img_blade.data(), img_blade.size())) ||
        (rc = upload(c, c->t_foil, img_foil.data(), img_foil.size())) ||
        (rc = upload(c, c->t_nfoil, nfoil.data(), nfoil.size() * sizeof(int))) ||
        (rc = upload(c, c->t_nang, nang.data(), nang.size() * sizeof(int))) ||
        (rc = upload(c, c->t_box, box.data(), box.size() * sizeof(int))) ||
        (rc = dev_alloc(c, c->t_pts, (size_t)npts * kTurbPt * rb)))
        return rc;
    // force planes: zero outside the deposit box, which the deposit rewrites every step
    const size_t pb = (size_t)c->lay.cplane * rb;
    for (int q = 0; q < 3; ++q) {
        if ((rc = dev_alloc(c, c->force[q], pb))) return rc;
        HIPCHK(c, hipMemsetAsync(c->force[q].p, 0, pb, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    TurbHost& h = c->th;
    h.nturb = t->nturbines;
    h.nblades = t->nblades;
    h.nchords = t->nchords;
    h.nang_max = t->nang_max;
    h.npts = (int)npts;
    h.krad = krad;
    h.sigma2 = R(R(R(2.0) * sigma) * sigma);
    h.rot = c->t_rot.p;
    h.theta = c->t_theta.p;
    h.blade = c->t_blade.p;
    h.nfoil = static_cast<const int*>(c->t_nfoil.p);
    h.nang = static_cast<const int*>(c->t_nang.p);
    h.foil = c->t_foil.p;
    h.pts = c->t_pts.p;
    h.tbox = static_cast<const int*>(c->t_box.p);
    for (int a = 0; a < 3; ++a) {
        h.lo[a] = lo[a];
        h.hi[a] = hi[a];
    }
    for (int q = 0; q < 3; ++q) h.F[q] = c->force[q].p;
    c->turb_on = true;
    return LBM_OK;
}

int lbm_get_turbine_theta(lbm_ctx* c, double* theta) {
    if (int rc_ = check_ctx(c)) return rc_;
    if (!theta) return fail(c, LBM_EINVAL, "null argument");
    if (!c->turb_on) return fail(c, LBM_ESTATE, "no turbines: call lbm_set_turbines first");
    const int rb = c->ops->real_bytes;
    std::vector<unsigned char> img((size_t)c->th.nturb * rb);
    HIPCHK(c, hipMemcpyAsync(img.data(), c->t_theta.p, img.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < c->th.nturb; ++m) theta[m] = get_real(img.data(), m, rb);
    return LBM_OK;
}

}  // extern "C"
