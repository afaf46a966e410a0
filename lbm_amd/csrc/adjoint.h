// Hand-written adjoint of the time step (reverse mode), for gradients of a time-mean kinetic
// energy functional with respect to the external body force and the inflow shear profile.
//
// The forward step n reads buffer X(n-1) (interior and ghosts) and writes buffer X(n): the
// interior by pull-streaming + collision (step_cell), then its ghosts by the folded closed-face
// pass and the inflow/outflow pass.  Ghosts that no pass rewrites keep the value the buffer held
// two steps earlier.  The adjoint keeps one adjoint buffer per forward buffer (same layout) and
// reverses step n as the sequence of writes in reverse order:
//   1. inflow/outflow pass  -> adjoint_inflow_kernel
//   2. closed-face pass     -> adjoint_face_kernel
//   3. stream + collide     -> adjoint_step_kernel
// Every write "y = f(x)" becomes "lambda_x += (df/dx)^T lambda_y; lambda_y = 0" for the slots it
// overwrote.  The pull gather is a permutation of the source slots it reads, so its reverse is a
// conflict-free scatter: interior slots of the source buffer are assigned, ghost slots (which may
// already hold adjoint from the face pass) accumulate, and interior slots that no cell pulls are
// cleared by their own cell.
//
// collide_adjoint() is the reverse of collide_cell() (kernels.h) for one cell: the transposed
// Jacobian of moments -> equilibrium -> regularized non-equilibrium (Malaspinas recursion) ->
// Vreman eddy viscosity -> relaxation -> Kupershtokh body-force increment -> solid swap,
// applied to the adjoint of the post-collision populations.  It is host-callable so that it can
// be checked against central differences of collide_cell on the CPU.
#pragma once

#include <cmath>

#include "kernels.h"

namespace lbm {

// Adjoint of one cell, the reverse of collide_cell (kernels.h) in the same pair-symmetric form.
// D: departures g_l - w_l of the streamed (pre-collision) populations; Ly: adjoint of the
// post-collision populations; kappa: weight of the kinetic energy 0.5*rho*|u|^2 of g in the
// objective (0: not in the objective box); coll = false: objective term only (Ly unused).  lg
// receives the adjoint of g, lF the adjoint of the cell's body force.
//
// Directions come in opposite pairs (l, l+1), c_{l+1} = -c_l.  Every per-direction factor of the
// forward is even or odd in c, so each pair contributes  ls*(even) + ld*(odd)  with
// ls = Ly_l + Ly_{l+1}, ld = Ly_l - Ly_{l+1} (the solid swap exchanges the pair: ld changes sign),
// and the outputs are lg_l = even + odd, lg_{l+1} = even - odd.  The pair loops are unrolled, so
// every lattice constant is an immediate.
template <int Q, bool BGK3, bool HRR, bool VRE, typename R>
__host__ __device__ __forceinline__ void collide_adjoint(const StepArgs<R>& a, const R (&D)[Q], int k, int oc, bool solid,
                                                         const R (&Ly)[Q], R kappa, bool coll, R (&lg)[Q], R (&lF)[3]) {
    using Lt = Lattice<Q>;
    constexpr int NP = (Q - 1) / 2;
    // per-cell inputs read first (their latency overlaps the moments instead of stalling later)
    R abl = R(0);
    if (VRE && coll) abl = a.ablvisc[k - 1];
    R fx = R(0), fy = R(0), fz = R(0);
    if (a.Fx != nullptr && coll) {
        fx = a.Fx[oc];
        fy = a.Fy[oc];
        fz = a.Fz[oc];
    }
    // ---- forward recompute (collide_cell's moments)
    R dr = D[0], jx = 0, jy = 0, jz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int l = 2 * p + 1;
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        const R sp = D[l] + D[l + 1], dp = D[l] - D[l + 1];
        dr += sp;
        if (cx > 0) jx += dp; else if (cx < 0) jx -= dp;
        if (cy > 0) jy += dp; else if (cy < 0) jy -= dp;
        if (cz > 0) jz += dp; else if (cz < 0) jz -= dp;
        if (cx != 0) pxx += sp;
        if (cy != 0) pyy += sp;
        if (cz != 0) pzz += sp;
        if (cx * cy > 0) pxy += sp; else if (cx * cy < 0) pxy -= sp;
        if (cx * cz > 0) pxz += sp; else if (cx * cz < 0) pxz -= sp;
        if (cy * cz > 0) pyz += sp; else if (cy * cz < 0) pyz -= sp;
    }
    const R rho = R(1) + dr;
    const R inv = recip(rho);
    const R ux = jx * inv, uy = jy * inv, uz = jz * inv;
    const R usq = ux * ux + uy * uy + uz * uz;
    lF[0] = lF[1] = lF[2] = R(0);
    if (!coll) {
        // objective only: d(0.5 |j|^2 / rho)/dg_l = c.u - |u|^2/2
        const R e = R(-0.5) * kappa * usq;
        lg[0] = e;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int l = 2 * p + 1;
            const R o = kappa * sdot<R>(Lt::cx(l), Lt::cy(l), Lt::cz(l), ux, uy, uz);
            lg[l] = e + o;
            lg[l + 1] = e - o;
        }
        return;
    }
    const R third = R(1) / R(3);
    const R axx = pxx - rho * ux * ux - dr * third, ayy = pyy - rho * uy * uy - dr * third,
            azz = pzz - rho * uz * uz - dr * third;
    const R axy = pxy - rho * ux * uy, axz = pxz - rho * ux * uz, ayz = pyz - rho * uy * uz;
    R fac = a.fac0, tau = 0, amag = 0, amag_raw = 0, t = 0, b11 = 0, b22 = 0, b33 = 0, b12 = 0, b13 = 0, b23 = 0;
    if (VRE) {
        amag_raw = axx * axx + ayy * ayy + azz * azz + R(2) * (axy * axy + axz * axz + ayz * ayz);
        amag = amag_raw > a.tiny ? amag_raw : a.tiny;
        b11 = a.eps + (axx * axx + axy * axy + axz * axz);
        b22 = a.eps + (axy * axy + ayy * ayy + ayz * ayz);
        b33 = a.eps + (axz * axz + ayz * ayz + azz * azz);
        b12 = a.eps + (axx * axy + axy * ayy + axz * ayz);
        b13 = a.eps + (axx * axz + axy * ayz + axz * azz);
        b23 = a.eps + (axy * axz + ayy * ayz + ayz * azz);
        const R bb = b11 * b22 - b12 * b12 + b11 * b33 - b13 * b13 + b22 * b33 - b23 * b23;
        t = bb * recip(amag);
        const R eddy = t > a.eps ? a.vconst * sqrt_r(t) : R(0);
        tau = R(3) * (a.kinevisc + eddy + abl) + R(0.5);
        fac = R(1) - recip(tau);
    }
    const R T = axx + ayy + azz;
    const R gx = axx * ux + axy * uy + axz * uz, gy = axy * ux + ayy * uy + ayz * uz, gz = axz * ux + ayz * uy + azz * uz;

    // ---- reverse
    // pair sums / differences of the adjoint before the solid swap
    const R sgn = solid ? R(-1) : R(1);
    R lrho = 0, lux = 0, luy = 0, luz = 0;
    R lxx = 0, lyy = 0, lzz = 0, lxy = 0, lxz = 0, lyz = 0;   // adjoint of a2
    R lfac = 0;
    R su = 0;   // sum of the coefficients of u in the adjoint of u (added as su * u at the end)
    const R L0 = Ly[0];
    const R w0 = weight<Q, R>(0);

    // body force: y += feq2(rho, v) - feq2(rho, u), v = u - F/rho (skipped when rho <= 0).
    // Per pair: P2(v) - P2(u) = [-1.5 (v.v - u.u) + 4.5 (cv^2 - cu^2)] + [3 (cv - cu)];
    // grad_u: -3 (v - u) + 9 (cv - cu) c (even); grad_v P2(v) = [-3 v + 9 cv c] + [3 c].
    if (a.Fx != nullptr && rho > R(0)) {
        const R Fx = fx, Fy = fy, Fz = fz;
        const R vx = ux - Fx * inv, vy = uy - Fy * inv, vz = uz - Fz * inv;
        const R vsq = vx * vx + vy * vy + vz * vz;
        const R dsq = R(-1.5) * (vsq - usq);
        // coefficients of (v - u) and v in lu and lv; the c-terms go straight to lu / lv
        R sdv = w0 * L0 * R(-3), sv = w0 * L0 * R(-3);
        R lvx = 0, lvy = 0, lvz = 0, lgx = 0, lgy = 0, lgz = 0;
        lrho += w0 * L0 * dsq;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int l = 2 * p + 1;
            const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
            const R w = weight<Q, R>(l);
            const R ls = w * (Ly[l] + Ly[l + 1]);
            const R ld = w * sgn * (Ly[l] - Ly[l + 1]);
            const R cv = sdot<R>(cx, cy, cz, vx, vy, vz), cu = sdot<R>(cx, cy, cz, ux, uy, uz);
            lrho += ls * (dsq + R(4.5) * (cv * cv - cu * cu)) + ld * (R(3) * (cv - cu));
            sdv += R(-3) * ls;
            sv += R(-3) * ls;
            const R gu = R(9) * (cv - cu) * ls;            // coefficient of c in lu
            const R gv = R(9) * cv * ls + R(3) * ld;        // coefficient of c in lv
            if (cx > 0) { lgx += gu; lvx += gv; } else if (cx < 0) { lgx -= gu; lvx -= gv; }
            if (cy > 0) { lgy += gu; lvy += gv; } else if (cy < 0) { lgy -= gu; lvy -= gv; }
            if (cz > 0) { lgz += gu; lvz += gv; } else if (cz < 0) { lgz -= gu; lvz -= gv; }
        }
        lvx = rho * (lvx + sv * vx);
        lvy = rho * (lvy + sv * vy);
        lvz = rho * (lvz + sv * vz);
        lux += rho * (lgx + sdv * (vx - ux));
        luy += rho * (lgy + sdv * (vy - uy));
        luz += rho * (lgz + sdv * (vz - uz));
        // dv/drho = F / rho^2, dv/dF = -1/rho (the du/du parts are in lu above)
        lrho += (lvx * Fx + lvy * Fy + lvz * Fz) * inv * inv;
        lF[0] = -lvx * inv;
        lF[1] = -lvy * inv;
        lF[2] = -lvz * inv;
    }

    // relaxation y = feq + fac * fneq.  Per pair (cu = c.u, cu2 = cu^2):
    //   P = [1 - 1.5 u.u + 4.5 cu2] + [3 cu (+ 4.5 cu (cu2 - u.u))]
    //   grad_u P = [-3 u + 9 cu c] + [(3 (+ 4.5 (3 cu2 - u.u))) c (- 9 cu u)]
    //   fneq = 4.5 w ([cAc - T/3] + [cu (3 cAc - T) - 2 c.g])
    // and d fneq / d a2, d fneq / d u (odd: c (3 cAc - T) - 2 a2 c).
    const R P0 = R(1) - R(1.5) * usq;
    const R lw0 = HRR ? L0 : (R(1) - fac) * L0;
    lrho += lw0 * w0 * P0;
    su += lw0 * w0 * rho * R(-3);
    if (HRR) {
        lfac += L0 * (R(-1.5) * w0 * T);
        const R phi0 = fac * L0 * (R(4.5) * w0) * (-third);
        lxx += phi0;
        lyy += phi0;
        lzz += phi0;
    } else {
        lfac += L0 * ((w0 + D[0]) - w0 * rho * P0);
    }
    R lcx = 0, lcy = 0, lcz = 0;   // coefficients of c in the adjoint of u
    // pair sums of the regularized part (see the loop)
    R sr = 0, rx = 0, ry = 0, rz = 0, oxy = 0, oxz = 0, oyz = 0, dx = 0, dy = 0, dz = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int l = 2 * p + 1;
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        const R w = weight<Q, R>(l);
        const R ly0 = solid ? Ly[l + 1] : Ly[l], ly1 = solid ? Ly[l] : Ly[l + 1];
        const R cu = sdot<R>(cx, cy, cz, ux, uy, uz);
        const R cu2 = cu * cu;
        const R Pe = P0 + R(4.5) * cu2;
        R Po = R(3) * cu, gq = R(3), gr = R(0);
        if (BGK3) {
            Po += R(4.5) * cu * (cu2 - usq);
            gq += R(4.5) * (R(3) * cu2 - usq);
            gr = R(-9) * cu;
        }
        const R ls0 = ly0 + ly1, ld0 = ly0 - ly1;
        // feq part (weight w, times (1 - fac) for BGK relaxation)
        const R ls = HRR ? ls0 : (R(1) - fac) * ls0;
        const R ld = HRR ? ld0 : (R(1) - fac) * ld0;
        lrho += w * (ls * Pe + ld * Po);
        const R wr = w * rho;
        su += wr * (R(-3) * ls + gr * ld);
        const R cc = wr * (R(9) * cu * ls + gq * ld);
        if (cx > 0) lcx += cc; else if (cx < 0) lcx -= cc;
        if (cy > 0) lcy += cc; else if (cy < 0) lcy -= cc;
        if (cz > 0) lcz += cc; else if (cz < 0) lcz -= cc;
        if (HRR) {
            R cac = sdot<R>(cx * cy, cx * cz, cy * cz, R(2) * axy, R(2) * axz, R(2) * ayz);
            if (cx != 0) cac += axx;
            if (cy != 0) cac += ayy;
            if (cz != 0) cac += azz;
            const R cg = sdot<R>(cx, cy, cz, gx, gy, gz);
            const R w45 = R(4.5) * w;
            const R s3 = R(3) * cac - T;
            lfac += w45 * (ls0 * (cac - T * third) + ld0 * (cu * s3 - R(2) * cg));
            const R kf = fac * w45;
            const R ps = kf * ls0, pd = kf * ld0;
            // d fneq / d a_pp = [c_p^2 - 1/3] + [cu (3 c_p^2 - 1) - 2 c_p u_p],
            // d fneq / d a_pq = [2 c_p c_q] + [6 cu c_p c_q - 2 (c_p u_q + c_q u_p)]  (p != q),
            // d fneq / d u = c (3 cAc - T) - 2 a2 c: with r = ps + 3 pd cu they are sums over the pairs
            // of r (all pairs, and the pairs with c_p != 0), of the signed c_p c_q r and c_p pd, and of
            // the signed c_p pd s3, formed into the adjoints of a2 and u after the loop
            const R r = ps + R(3) * (pd * cu);
            sr += r;
            if (cx != 0) rx += r;
            if (cy != 0) ry += r;
            if (cz != 0) rz += r;
            if (cx * cy > 0) oxy += r; else if (cx * cy < 0) oxy -= r;
            if (cx * cz > 0) oxz += r; else if (cx * cz < 0) oxz -= r;
            if (cy * cz > 0) oyz += r; else if (cy * cz < 0) oyz -= r;
            if (cx > 0) dx += pd; else if (cx < 0) dx -= pd;
            if (cy > 0) dy += pd; else if (cy < 0) dy -= pd;
            if (cz > 0) dz += pd; else if (cz < 0) dz -= pd;
            const R pds = pd * s3;
            if (cx > 0) lcx += pds; else if (cx < 0) lcx -= pds;
            if (cy > 0) lcy += pds; else if (cy < 0) lcy -= pds;
            if (cz > 0) lcz += pds; else if (cz < 0) lcz -= pds;
        } else {
            // y = feq + fac (g - feq): d/dfac = g - feq
            const R wrP = wr;
            lfac += ls0 * (w - wrP * Pe) - ld0 * (wrP * Po) + (ly0 * D[l] + ly1 * D[l + 1]);
        }
    }
    lux += su * ux + lcx;
    luy += su * uy + lcy;
    luz += su * uz + lcz;
    if (HRR) {
        const R s3r = sr * third;
        lxx += rx - s3r - R(2) * ux * dx;
        lyy += ry - s3r - R(2) * uy * dy;
        lzz += rz - s3r - R(2) * uz * dz;
        lxy += R(2) * (oxy - (uy * dx + ux * dy));
        lxz += R(2) * (oxz - (uz * dx + ux * dz));
        lyz += R(2) * (oyz - (uz * dy + uy * dz));
        lux -= R(2) * (axx * dx + axy * dy + axz * dz);
        luy -= R(2) * (axy * dx + ayy * dy + ayz * dz);
        luz -= R(2) * (axz * dx + ayz * dy + azz * dz);
    }

    if (VRE) {
        // fac = 1 - 1/tau, tau = 3 (nu + nu_t + abl) + 1/2, nu_t = c sqrt(t) (t > eps)
        const R itau = recip(tau), iam = recip(amag);
        const R ltau = lfac * (itau * itau);
        const R lnut = R(3) * ltau;
        const R lt = t > a.eps ? lnut * a.vconst * R(0.5) * recip(sqrt_r(t)) : R(0);
        const R lbb = lt * iam;
        const R lam = amag_raw > a.tiny ? -lt * t * iam : R(0);
        const R l11 = lbb * (b22 + b33), l22 = lbb * (b11 + b33), l33 = lbb * (b11 + b22);
        const R l12 = R(-2) * b12 * lbb, l13 = R(-2) * b13 * lbb, l23 = R(-2) * b23 * lbb;
        lxx += R(2) * axx * l11 + axy * l12 + axz * l13 + R(2) * axx * lam;
        lyy += R(2) * ayy * l22 + axy * l12 + ayz * l23 + R(2) * ayy * lam;
        lzz += R(2) * azz * l33 + axz * l13 + ayz * l23 + R(2) * azz * lam;
        lxy += R(2) * axy * (l11 + l22) + (axx + ayy) * l12 + ayz * l13 + axz * l23 + R(4) * axy * lam;
        lxz += R(2) * axz * (l11 + l33) + ayz * l12 + (axx + azz) * l13 + axy * l23 + R(4) * axz * lam;
        lyz += R(2) * ayz * (l22 + l33) + axz * l12 + axy * l13 + (ayy + azz) * l23 + R(4) * ayz * lam;
    }

    // a2 = Pi' - rho u u - (rho - 1) I/3
    lrho -= lxx * ux * ux + lyy * uy * uy + lzz * uz * uz + lxy * ux * uy + lxz * ux * uz + lyz * uy * uz;
    lrho -= (lxx + lyy + lzz) * third;
    lux -= rho * (R(2) * ux * lxx + uy * lxy + uz * lxz);
    luy -= rho * (R(2) * uy * lyy + ux * lxy + uz * lyz);
    luz -= rho * (R(2) * uz * lzz + ux * lxz + uy * lyz);
    // u = j / rho
    const R ljx = lux * inv + kappa * ux, ljy = luy * inv + kappa * uy, ljz = luz * inv + kappa * uz;
    lrho -= (lux * ux + luy * uy + luz * uz) * inv;
    const R k0 = lrho - R(0.5) * kappa * usq;
    lg[0] = HRR ? k0 : k0 + fac * L0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int l = 2 * p + 1;
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        R e = k0;
        if (cx != 0) e += lxx;
        if (cy != 0) e += lyy;
        if (cz != 0) e += lzz;
        e += sdot<R>(cx * cy, cx * cz, cy * cz, lxy, lxz, lyz);
        const R o = sdot<R>(cx, cy, cz, ljx, ljy, ljz);
        R v0 = e + o, v1 = e - o;
        if (!HRR) {
            v0 += fac * (solid ? Ly[l + 1] : Ly[l]);
            v1 += fac * (solid ? Ly[l] : Ly[l + 1]);
        }
        lg[l] = v0;
        lg[l + 1] = v1;
    }
}

// ------------------------------------------------------------------------------------------
// Reverse of the stream + collide kernel for step n.  src: forward buffer X(n-1) (the step's
// input); lam_y: adjoint buffer of X(n) (read, interior; unused when coll == 0); lam_x: adjoint
// buffer of X(n-1) (interior slots assigned, ghost slots accumulated).  box: objective cells
// (inclusive, 1-based); kappa: objective weight of the moments of X(n-1) streamed, i.e. this
// step's gathered input.  gF: FP64 accumulators of the body-force adjoint (3 cell planes) or
// nullptr.  The gather offsets are formed as in step_cell (uniform strides, wrap distances of the
// periodic axes); SMALL: the buffer is below 4 GiB (32-bit byte offsets from one base).
// ------------------------------------------------------------------------------------------
template <int Q, class C, bool BGK3, bool HRR, bool VRE, bool SMALL>
__global__ void __launch_bounds__(256) adjoint_step_kernel(const StepArgs<typename C::real> a, const typename C::raw* lam_y,
                                                           typename C::raw* lam_x, AdjBox box, typename C::real kappa,
                                                           int coll, double* gF) {
    using R = typename C::real;
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    const Layout& L = a.lay;
    // flat over the cells of one plane of ax0 x ax1 (no idle lanes when n0 is not a multiple of 64),
    // blockIdx.y along ax2; t / n0 through the FP32 reciprocal with one correction step (exact for
    // t < 2^24)
    const int n0 = L.n(L.ax0);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n0 * L.n(L.ax1)) return;
    int q = (int)((float)t * __builtin_amdgcn_rcpf((float)n0));
    int rem = t - q * n0;
    if (rem < 0) { --q; rem += n0; } else if (rem >= n0) { ++q; rem -= n0; }
    const int ca = rem + 1, cb = q + 1, cc = blockIdx.y + 1;
    int i, j, k;
    L.ijk(ca, cb, cc, i, j, k);
    const int x[3] = {i, j, k};
    const int n[3] = {L.nx, L.ny, L.nz};
    const int per[3] = {L.per_i, L.per_j, L.per_k};
    // element offsets of the sources i - cx, j - cy, k - cz (index c + 1)
    const int oc = L.cell(i, j, k);
    int ox[3], oy[3], oz[3];
    auto offsets = [&](int xx, int nn, int pp, int st, int (&o)[3]) {
        const int wrapd = pp ? nn * st : 0;
        const int lo = (pp && xx == nn) ? -wrapd : 0;
        const int hi = (pp && xx == 1) ? wrapd : 0;
        o[0] = st + lo;
        o[1] = 0;
        o[2] = -(st - hi);
    };
    offsets(i, L.nx, L.per_i, L.si, ox);
    offsets(j, L.ny, L.per_j, L.sj, oy);
    offsets(k, L.nz, L.per_k, L.sk, oz);
    auto src_off = [&](int l) { return oc + oz[Lt::cz(l) + 1] + oy[Lt::cy(l) + 1] + ox[Lt::cx(l) + 1]; };
    // population l of element offset o
    const size_t plane = (size_t)L.plane;
    auto at = [&](auto* base, int l, int o) {
        if constexpr (SMALL) {
            using B = std::conditional_t<std::is_const<std::remove_pointer_t<decltype(base)>>::value, const char*, char*>;
            return reinterpret_cast<decltype(base)>(reinterpret_cast<B>(base) +
                                                    ((uint32_t)l * (uint32_t)(plane * sizeof(raw)) + (uint32_t)o * (uint32_t)sizeof(raw)));
        } else {
            return base + ((size_t)l * plane + (size_t)o);
        }
    };
    const raw* __restrict__ src = static_cast<const raw*>(a.src);
    // 4-byte storage below 4 GiB: raw buffer loads and stores with the population's plane in the
    // scalar offset (no per-access address arithmetic), as the forward step kernel
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr bool BUF = SMALL && sizeof(raw) == 4;
#else
    constexpr bool BUF = false;
#endif
    const uint32_t pb32 = (uint32_t)(plane * sizeof(raw));
    const size_t tbytes = (size_t)L.total * sizeof(raw);
    const int nbytes = (int)(uint32_t)(tbytes < 0xffffffffull ? tbytes : 0xffffffffull);
    R D[Q], Ly[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if constexpr (BUF) {
#if defined(__HIP_DEVICE_COMPILE__)
            D[l] = C::template dep<Q>(buffer_load<raw>(src, nbytes, (uint32_t)src_off(l) * (uint32_t)sizeof(raw), (uint32_t)l * pb32), l);
#endif
        } else {
            D[l] = C::template dep<Q>(*at(src, l, src_off(l)), l);
        }
    }
    if (coll) {
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            if constexpr (BUF) {
#if defined(__HIP_DEVICE_COMPILE__)
                Ly[l] = C::template decode<Q>(buffer_load<raw>(lam_y, nbytes, (uint32_t)oc * (uint32_t)sizeof(raw), (uint32_t)l * pb32), l);
#endif
            } else {
                Ly[l] = C::template decode<Q>(*at(lam_y, l, oc), l);
            }
        }
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) Ly[l] = R(0);
    }
    // the solid word is read behind the populations and used after the gradient loads are issued
    uint32_t sword = 0;
    if (a.solid != nullptr) sword = a.solid[((cc - 1) * L.n(L.ax1) + (cb - 1)) * a.solid_wpr + ((ca - 1) >> 5)];
    const bool inbox = i >= box.lo[0] && i <= box.hi[0] && j >= box.lo[1] && j <= box.hi[1] && k >= box.lo[2] && k <= box.hi[2];
    const int ocx = L.cidx(i, j, k);   // per-cell arrays (force, its gradient)
    // the force-gradient accumulators are read here, with the populations, so that their latency
    // overlaps the collision instead of following it
    const bool want_gF = gF != nullptr && a.Fx != nullptr;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (want_gF) {
        g0 = gF[ocx];
        g1 = gF[(size_t)L.cplane + ocx];
        g2 = gF[2 * (size_t)L.cplane + ocx];
    }
    const bool solid = a.solid != nullptr && ((sword >> ((ca - 1) & 31)) & 1u);
    R lg[Q], lF[3];
    collide_adjoint<Q, BGK3, HRR, VRE, R>(a, D, k, ocx, solid, Ly, inbox ? kappa : R(0), coll != 0, lg, lF);
    // a cell next to a non-periodic face: some sources are ghosts, some own slots are unhit
    bool edge = false;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (!per[d] && (x[d] == 1 || x[d] == n[d])) edge = true;
    raw* __restrict__ lx = lam_x;
    if (!edge) {
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            if constexpr (BUF) {
#if defined(__HIP_DEVICE_COMPILE__)
                buffer_store<raw, true>(lx, nbytes, (uint32_t)src_off(l) * (uint32_t)sizeof(raw), (uint32_t)l * pb32,
                                        C::template encode<Q>(lg[l], l));
#endif
            } else {
                *at(lx, l, src_off(l)) = C::template encode<Q>(lg[l], l);
            }
        }
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            // the source slot: interior -> assign; ghost -> accumulate
            bool ghost = false, unhit = false;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (per[d]) continue;
                const int cd = Lt::c(l, d);
                const int s = x[d] - cd, t = x[d] + cd;
                if (s < 1 || s > n[d]) ghost = true;
                if (t < 1 || t > n[d]) unhit = true;
            }
            raw* p = at(lx, l, src_off(l));
            if (ghost) *p = C::template encode<Q>(C::template decode<Q>(*p, l) + lg[l], l);
            else *p = C::template encode<Q>(lg[l], l);
            // this cell's own slot of direction l is pulled by no interior cell: clear it
            if (unhit) *at(lx, l, oc) = C::template encode<Q>(R(0), l);
        }
    }
    if (want_gF) {
        gF[ocx] = g0 + (double)lF[0];
        gF[(size_t)L.cplane + ocx] = g1 + (double)lF[1];
        gF[2 * (size_t)L.cplane + ocx] = g2 + (double)lF[2];
    }
}

// Reverse of the inflow/outflow pass (inflow_cell) of step n on the adjoint buffer lam of X(n).
// gS: FP64 accumulators of d/d uvel_shear per (j,k) (j fastest).
template <int Q, class C>
__global__ void __launch_bounds__(64) adjoint_inflow_kernel(Layout L, typename C::raw* lam, typename C::real utime, typename C::real cosd,
                                      typename C::real sind, typename C::real rho0, double* gS) {
    using R = typename C::real;
    using Lt = Lattice<Q>;
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (j > L.ny || k > L.nz) return;
    const int o1 = L.cell(1, j, k), o0 = L.cell(0, j, k), on = L.cell(L.nx, j, k), oe = L.cell(L.nx + 1, j, k);
    R l0[Q], le[Q], l1[Q], ln[Q], lfg[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        l0[l] = C::template decode<Q>(lam[(size_t)l * L.plane + o0], l);
        le[l] = C::template decode<Q>(lam[(size_t)l * L.plane + oe], l);
        l1[l] = C::template decode<Q>(lam[(size_t)l * L.plane + o1], l);
        ln[l] = C::template decode<Q>(lam[(size_t)l * L.plane + on], l);
        lfg[l] = R(0);
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) lfg[Lt::cx(l) <= 0 ? l : Lt::opposite(l)] += l0[l];
    R s = R(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        // fg_l = f_l - 2 w rho0 (cx cos + cy sin) uu / cs2,  uu = utime * shear(k)
        s += lfg[l] * (R(-6) * weight<Q, R>(l) * rho0 * (R(Lt::cx(l)) * cosd + R(Lt::cy(l)) * sind));
        lam[(size_t)l * L.plane + o1] = C::template encode<Q>(l1[l] + lfg[l], l);
        lam[(size_t)l * L.plane + on] = C::template encode<Q>(ln[l] + le[l], l);
        lam[(size_t)l * L.plane + o0] = C::template encode<Q>(R(0), l);
        lam[(size_t)l * L.plane + oe] = C::template encode<Q>(R(0), l);
    }
    if (gS != nullptr) gS[(size_t)(k - 1) * L.ny + (j - 1)] += (double)(s * utime);
}

// Reverse of the closed-face pass of axis D (fold_closed_face) of step n: lam_y adjoint of X(n),
// lam_x adjoint of X(n-1).  One thread per ghost cell of the two faces (tangential coordinates
// interior), t0/t1 the tangential axes.
template <int Q, class C>
__global__ void __launch_bounds__(64) adjoint_face_kernel(Layout L, typename C::raw* lam_y, typename C::raw* lam_x, int D, int t0, int t1,
                                    int opt_lo, int opt_hi) {
    using R = typename C::real;
    using Lt = Lattice<Q>;
    const int n[3] = {L.nx, L.ny, L.nz};
    const int per[3] = {L.per_i, L.per_j, L.per_k};
    const int a0 = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int a1 = blockIdx.y + 1;
    const int side = blockIdx.z;
    if (a0 > n[t0] || a1 > n[t1]) return;
    int g[3];
    g[t0] = a0;
    g[t1] = a1;
    g[D] = side ? n[D] + 1 : 0;
    const int plane = side ? n[D] : 1;
    const int opt = side ? opt_hi : opt_lo;
    const int og = L.cell(g[0], g[1], g[2]);
    // loads first: the adjoint of the written slots (l, G), (m, G), of X(n-1)'s (l, G), and of
    // the wall-plane sources
    R lm[Q], ll[Q], lxl[Q], lsv[Q];
    size_t os[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const int cn = Lt::c(l, D);
        if (!(side ? cn > 0 : cn < 0)) continue;
        const int m = opt < 0 ? mirror<Q>(l, D, -1) : mirror<Q>(l, D, +1);
        int sc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d == D) { sc[d] = plane; continue; }
            sc[d] = g[d] - Lt::c(l, d);
            if (per[d]) sc[d] = Layout::wrap(sc[d], n[d], 1);
        }
        os[l] = (size_t)l * L.plane + L.cell(sc[0], sc[1], sc[2]);
        lm[l] = C::template decode<Q>(lam_y[(size_t)m * L.plane + og], m);
        ll[l] = C::template decode<Q>(lam_y[(size_t)l * L.plane + og], l);
        lxl[l] = C::template decode<Q>(lam_x[(size_t)l * L.plane + og], l);
        lsv[l] = C::template decode<Q>(lam_y[os[l]], l);
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const int cn = Lt::c(l, D);
        if (!(side ? cn > 0 : cn < 0)) continue;
        const int m = opt < 0 ? mirror<Q>(l, D, -1) : mirror<Q>(l, D, +1);
        // (m, G) was copied from X(n-1)'s (l, G); (l, G) came from the wall-plane cell G - c_l
        // (tangential): an interior cell or a ghost of the inflow axis
        lam_x[(size_t)l * L.plane + og] = C::template encode<Q>(lxl[l] + lm[l], l);
        lam_y[os[l]] = C::template encode<Q>(lsv[l] + ll[l], l);
        lam_y[(size_t)m * L.plane + og] = C::template encode<Q>(R(0), m);
        lam_y[(size_t)l * L.plane + og] = C::template encode<Q>(R(0), l);
    }
}

// Objective: sum over the box of 0.5*rho*|u|^2 of the streamed populations of buffer f (the
// reference's macrovars after a step), added per block b to acc[2b] (FP64); with the tangent
// type its directional derivative is added to acc[2b+1].  The grid size is fixed per box
// (objective_blocks), so the host's ordered sum over the slots is reproducible.
template <int Q, class C>
__global__ void __launch_bounds__(256) objective_kernel(Layout L, const void* f, AdjBox box, double* acc) {
    using R = typename C::real;
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    constexpr bool TAN = is_dual<R>::value;
    const int bx = box.hi[0] - box.lo[0] + 1, by = box.hi[1] - box.lo[1] + 1, bz = box.hi[2] - box.lo[2] + 1;
    const long long nb = (long long)bx * by * bz;
    const raw* __restrict__ p = static_cast<const raw*>(f);
    double sum = 0.0, dsum = 0.0;
    // walk the box in storage order (contiguous axis fastest)
    const int ext[3] = {bx, by, bz};
    const int e0 = ext[L.ax0], e1 = ext[L.ax1];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nb; e += (long long)gridDim.x * blockDim.x) {
        int c[3];
        c[L.ax0] = box.lo[L.ax0] + (int)(e % e0);
        c[L.ax1] = box.lo[L.ax1] + (int)((e / e0) % e1);
        c[L.ax2] = box.lo[L.ax2] + (int)(e / ((long long)e0 * e1));
        const int i = c[0], j = c[1], k = c[2];
        R dr = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const R d = C::template dep<Q>(p[(size_t)l * L.plane + L.cell(Layout::wrap(i - Lt::cx(l), L.nx, L.per_i),
                                                                           Layout::wrap(j - Lt::cy(l), L.ny, L.per_j),
                                                                           Layout::wrap(k - Lt::cz(l), L.nz, L.per_k))], l) /
                          C::kDepScale;
            dr += d;
            jx += R(Lt::cx(l)) * d;
            jy += R(Lt::cy(l)) * d;
            jz += R(Lt::cz(l)) * d;
        }
        if constexpr (TAN) {
            const R ke = R(0.5) * (jx * jx + jy * jy + jz * jz) / (R(1) + dr);
            sum += ke.v;
            dsum += ke.d;
        } else {
            sum += 0.5 * ((double)jx * jx + (double)jy * jy + (double)jz * jz) / (1.0 + (double)dr);
        }
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = sum;
    red[1][threadIdx.x] = dsum;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    // every block owns its pair of slots: the sum is the same in every run
    if (threadIdx.x == 0) {
        acc[2 * blockIdx.x] += red[0][0];
        acc[2 * blockIdx.x + 1] += red[1][0];
    }
}

}  // namespace lbm
