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

// Adjoint of one cell.  Dg(l): departure g_l - w_l of the streamed (pre-collision) population l;
// Lg(l): adjoint of the post-collision population l; kappa: weight of the kinetic energy
// 0.5*rho*|u|^2 of g in the objective (0: not in the objective box); coll = false: objective term
// only.  The adjoint of g_l goes to sink(l, value); lF receives the adjoint of the cell's body
// force.  The per-direction loops are rolled or partly unrolled (the lattice constants of a
// run-time direction are scalar shifts of packed constants) so that the cell's state stays small;
// Dg and Lg may be called more than once per direction.
template <int Q, bool BGK3, bool HRR, bool VRE, typename R, class DG, class LG, class Sink>
__host__ __device__ inline void collide_adjoint_f(const StepArgs<R>& a, DG&& Dg, int k, int oc, bool solid, LG&& Lg,
                                                  R kappa, bool coll, Sink&& sink, R (&lF)[3]) {
    using Lt = Lattice<Q>;
    // The loops that read the adjoint of the outputs run 4 directions per iteration in FP32, so
    // that four loads are in flight together (reverse kernel 344 -> 325 us on 192x96x96 D3Q19);
    // in FP64 that costs occupancy (134 -> 174 VGPRs) and is slower, so they stay rolled.
    constexpr int kLoadUnroll = sizeof(R) == 4 ? 4 : 1;
    // ---- forward recompute (collide_cell's quantities, in plain rather than paired form)
    R dr = 0, jx = 0, jy = 0, jz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const R cx = R(Lt::cx(l)), cy = R(Lt::cy(l)), cz = R(Lt::cz(l));
        const R d = Dg(l);
        dr += d;
        jx += cx * d;
        jy += cy * d;
        jz += cz * d;
        pxx += cx * cx * d;
        pyy += cy * cy * d;
        pzz += cz * cz * d;
        pxy += cx * cy * d;
        pxz += cx * cz * d;
        pyz += cy * cz * d;
    }
    const R rho = R(1) + dr;
    const R inv = R(1) / rho;
    const R ux = jx * inv, uy = jy * inv, uz = jz * inv;
    const R usq = ux * ux + uy * uy + uz * uz;
    lF[0] = lF[1] = lF[2] = R(0);
    if (!coll) {
        // objective only: d(0.5 |j|^2 / rho)/dg_l = c.u - |u|^2/2
#pragma unroll 1
        for (int l = 0; l < Q; ++l) sink(l, kappa * (sdot<R>(Lt::cx(l), Lt::cy(l), Lt::cz(l), ux, uy, uz) - R(0.5) * usq));
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
        t = bb / amag;
        const R eddy = t > a.eps ? a.vconst * std::sqrt(t) : R(0);
        tau = R(3) * (a.kinevisc + eddy + a.ablvisc[k - 1]) + R(0.5);
        fac = R(1) - R(1) / tau;
    }
    const R T = axx + ayy + azz;
    const R gx = axx * ux + axy * uy + axz * uz, gy = axy * ux + ayy * uy + ayz * uz, gz = axz * ux + ayz * uy + azz * uz;

    // ---- reverse
    // adjoint of the population before the solid swap of pairs (l, l+1), l odd
    auto ly = [&](int l) { return Lg(solid && l > 0 ? ((l & 1) ? l + 1 : l - 1) : l); };
    R lrho = 0, lux = 0, luy = 0, luz = 0;
    R lxx = 0, lyy = 0, lzz = 0, lxy = 0, lxz = 0, lyz = 0;   // adjoint of a2
    R lfac = 0;

    // body force: y += feq2(rho, v) - feq2(rho, u), v = u - F/rho (skipped when rho <= 0)
    if (a.Fx != nullptr && rho > R(0)) {
        const R Fx = a.Fx[oc], Fy = a.Fy[oc], Fz = a.Fz[oc];
        const R vx = ux - Fx * inv, vy = uy - Fy * inv, vz = uz - Fz * inv;
        const R vsq = vx * vx + vy * vy + vz * vz;
        R lvx = 0, lvy = 0, lvz = 0;
#pragma unroll kLoadUnroll
        for (int l = 0; l < Q; ++l) {
            const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
            const R w = weight<Q, R>(l);
            const R cv = sdot<R>(cx, cy, cz, vx, vy, vz), cu = sdot<R>(cx, cy, cz, ux, uy, uz);
            const R Pv = R(1) - R(1.5) * vsq + R(3) * cv + R(4.5) * cv * cv;
            const R Pu = R(1) - R(1.5) * usq + R(3) * cu + R(4.5) * cu * cu;
            const R wl = w * ly(l);
            const R wr = wl * rho;
            lrho += wl * (Pv - Pu);
            // grad P2(v) - grad P2(u) = -3 (v - u) + 9 (c.v - c.u) c, and grad P2(v)
            const R dc = R(9) * (cv - cu);
            lux += wr * (R(-3) * (vx - ux) + dc * R(cx));
            luy += wr * (R(-3) * (vy - uy) + dc * R(cy));
            luz += wr * (R(-3) * (vz - uz) + dc * R(cz));
            const R c9 = R(3) + R(9) * cv;
            lvx += wr * (R(-3) * vx + c9 * R(cx));
            lvy += wr * (R(-3) * vy + c9 * R(cy));
            lvz += wr * (R(-3) * vz + c9 * R(cz));
        }
        // dv/drho = F / rho^2, dv/dF = -1/rho (du/du contributions are in lu above)
        lrho += (lvx * Fx + lvy * Fy + lvz * Fz) * inv * inv;
        lF[0] = -lvx * inv;
        lF[1] = -lvy * inv;
        lF[2] = -lvz * inv;
    }

    // relaxation y = feq + fac * fneq
#pragma unroll kLoadUnroll
    for (int l = 0; l < Q; ++l) {
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        const R fx = R(cx), fy = R(cy), fz = R(cz);
        const R w = weight<Q, R>(l);
        const R cu = sdot<R>(cx, cy, cz, ux, uy, uz);
        R P = R(1) - R(1.5) * usq + R(3) * cu + R(4.5) * cu * cu;
        // grad P = -3u + 3c + 9 cu c [+ 4.5 ((3 cu^2 - u.u) c - 2 cu u)]
        R gq = R(3) + R(9) * cu, gu = R(-3);
        if (BGK3 && l > 0) {
            P += R(4.5) * cu * (cu * cu - usq);
            gq += R(4.5) * (R(3) * cu * cu - usq);
            gu -= R(9) * cu;
        }
        const R le = ly(l);
        if (HRR) {
            R cac = R(0);
            if (cx != 0) cac += axx;
            if (cy != 0) cac += ayy;
            if (cz != 0) cac += azz;
            cac += R(2) * sdot<R>(cx * cy, cx * cz, cy * cz, axy, axz, ayz);
            const R cg = sdot<R>(cx, cy, cz, gx, gy, gz);
            const R w45 = R(4.5) * w;
            const R fneq = w45 * ((cac - T * third) + (cu * (R(3) * cac - T) - R(2) * cg));
            const R lw = le * w;
            const R lwr = lw * rho;
            lrho += lw * P;
            lux += lwr * (gu * ux + gq * fx);
            luy += lwr * (gu * uy + gq * fy);
            luz += lwr * (gu * uz + gq * fz);
            lfac += le * fneq;
            const R phi = fac * le * w45;
            // d fneq / d a2 (six symmetric entries)
            const R c2 = R(2) * cu;
            lxx += phi * ((fx * fx - third) + cu * (R(3) * fx * fx - R(1)) - R(2) * fx * ux);
            lyy += phi * ((fy * fy - third) + cu * (R(3) * fy * fy - R(1)) - R(2) * fy * uy);
            lzz += phi * ((fz * fz - third) + cu * (R(3) * fz * fz - R(1)) - R(2) * fz * uz);
            lxy += phi * (R(2) * fx * fy + R(3) * c2 * fx * fy - R(2) * (fx * uy + fy * ux));
            lxz += phi * (R(2) * fx * fz + R(3) * c2 * fx * fz - R(2) * (fx * uz + fz * ux));
            lyz += phi * (R(2) * fy * fz + R(3) * c2 * fy * fz - R(2) * (fy * uz + fz * uy));
            // d fneq / d u = c (3 cAc - T) - 2 a2 c
            const R s3 = R(3) * cac - T;
            lux += phi * (fx * s3 - R(2) * (axx * fx + axy * fy + axz * fz));
            luy += phi * (fy * s3 - R(2) * (axy * fx + ayy * fy + ayz * fz));
            luz += phi * (fz * s3 - R(2) * (axz * fx + ayz * fy + azz * fz));
        } else {
            // y = feq + fac (g - feq)
            const R feq = w * rho * P;
            const R lfe = (R(1) - fac) * le;
            const R lw = lfe * w;
            const R lwr = lw * rho;
            lrho += lw * P;
            lux += lwr * (gu * ux + gq * fx);
            luy += lwr * (gu * uy + gq * fy);
            luz += lwr * (gu * uz + gq * fz);
            lfac += le * ((w + Dg(l)) - feq);
        }
    }

    if (VRE) {
        // fac = 1 - 1/tau, tau = 3 (nu + nu_t + abl) + 1/2, nu_t = c sqrt(t) (t > eps)
        const R ltau = lfac / (tau * tau);
        const R lnut = R(3) * ltau;
        const R lt = t > a.eps ? lnut * a.vconst / (R(2) * std::sqrt(t)) : R(0);
        const R lbb = lt / amag;
        const R lam = amag_raw > a.tiny ? -lt * t / amag : R(0);
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
    const R ljx = lux * inv, ljy = luy * inv, ljz = luz * inv;
    lrho -= (lux * ux + luy * uy + luz * uz) * inv;
    const R kx = kappa * ux, ky = kappa * uy, kz = kappa * uz, k0 = lrho - R(0.5) * kappa * usq;
#pragma unroll 1
    for (int l = 0; l < Q; ++l) {
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        R v = k0 + sdot<R>(cx, cy, cz, ljx + kx, ljy + ky, ljz + kz);
        if (cx != 0) v += lxx;
        if (cy != 0) v += lyy;
        if (cz != 0) v += lzz;
        v += sdot<R>(cx * cy, cx * cz, cy * cz, lxy, lxz, lyz);
        if (!HRR) v += fac * ly(l);
        sink(l, v);
    }
}

// Array form (host checks).
template <int Q, bool BGK3, bool HRR, bool VRE, typename R>
__host__ __device__ inline void collide_adjoint(const StepArgs<R>& a, const R (&D)[Q], int k, int oc, bool solid,
                                                const R (&ly_in)[Q], R kappa, bool coll, R (&lg)[Q], R (&lF)[3]) {
    collide_adjoint_f<Q, BGK3, HRR, VRE, R>(
        a, [&](int l) { return D[l]; }, k, oc, solid, [&](int l) { return ly_in[l]; }, kappa, coll,
        [&](int l, R v) { lg[l] = v; }, lF);
}

// ------------------------------------------------------------------------------------------
// Reverse of the stream + collide kernel for step n.  src: forward buffer X(n-1) (the step's
// input); lam_y: adjoint buffer of X(n) (read, interior); lam_x: adjoint buffer of X(n-1)
// (interior slots assigned, ghost slots accumulated).  box: objective cells (inclusive, 1-based);
// kappa: objective weight of the moments of X(n-1) streamed, i.e. this step's gathered input.
// gF: FP64 accumulators of the body-force adjoint (3 cell planes) or nullptr.
// ------------------------------------------------------------------------------------------
template <int Q, class C, bool BGK3, bool HRR, bool VRE>
__global__ void __launch_bounds__(256) adjoint_step_kernel(const StepArgs<typename C::real> a, const typename C::raw* lam_y,
                                                           typename C::raw* lam_x, AdjBox box, typename C::real kappa,
                                                           int coll, double* gF) {
    using R = typename C::real;
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    const Layout& L = a.lay;
    const int ca = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int cb = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int cc = blockIdx.z + 1;
    if (ca > L.n(L.ax0) || cb > L.n(L.ax1) || cc > L.n(L.ax2)) return;
    int i, j, k;
    L.ijk(ca, cb, cc, i, j, k);
    const int x[3] = {i, j, k};
    const int n[3] = {L.nx, L.ny, L.nz};
    const int per[3] = {L.per_i, L.per_j, L.per_k};
    const raw* __restrict__ src = static_cast<const raw*>(a.src);
    // source cell of direction l (periodic axes folded)
    auto source = [&](int l) {
        return L.cell(Layout::wrap(i - Lt::cx(l), L.nx, L.per_i), Layout::wrap(j - Lt::cy(l), L.ny, L.per_j),
                      Layout::wrap(k - Lt::cz(l), L.nz, L.per_k));
    };
    auto Dg = [&](int l) { return C::template dep<Q>(src[(size_t)l * L.plane + source(l)], l); };
    const int oc = L.cell(i, j, k);
    auto Lg = [&](int l) { return C::template decode<Q>(lam_y[(size_t)l * L.plane + oc], l); };
    bool solid = false;
    if (a.solid != nullptr) {
        const uint32_t word = a.solid[((cc - 1) * L.n(L.ax1) + (cb - 1)) * a.solid_wpr + ((ca - 1) >> 5)];
        solid = (word >> ((ca - 1) & 31)) & 1u;
    }
    const bool inbox = i >= box.lo[0] && i <= box.hi[0] && j >= box.lo[1] && j <= box.hi[1] && k >= box.lo[2] && k <= box.hi[2];
    // a cell next to a non-periodic face: some sources are ghosts, some own slots are unhit
    bool edge = false;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (!per[d] && (x[d] == 1 || x[d] == n[d])) edge = true;
    raw* __restrict__ lx = lam_x;
    R lF[3];
    auto sink = [&](int l, R v) {
        raw* p = lx + (size_t)l * L.plane + source(l);
        if (!edge) {
            *p = C::template encode<Q>(v, l);
            return;
        }
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
        if (ghost) *p = C::template encode<Q>(C::template decode<Q>(*p, l) + v, l);
        else *p = C::template encode<Q>(v, l);
        // this cell's own slot of direction l is pulled by no interior cell: clear it
        if (unhit) lx[(size_t)l * L.plane + oc] = C::template encode<Q>(R(0), l);
    };
    const int ox = L.cidx(i, j, k);   // per-cell arrays (force, its gradient)
    collide_adjoint_f<Q, BGK3, HRR, VRE, R>(a, Dg, k, ox, solid, Lg, inbox ? kappa : R(0), coll != 0, sink, lF);
    if (gF != nullptr && a.Fx != nullptr) {
        gF[ox] += (double)lF[0];
        gF[(size_t)L.cplane + ox] += (double)lF[1];
        gF[2 * (size_t)L.cplane + ox] += (double)lF[2];
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
