// Device kernels of the lattice update.
//
// One time step (the reference's src/main.F90:290-329) is
//   1. step_kernel      : pull-stream + moments + equilibrium + regularization + Vreman tau +
//                         relaxation + inflow-turbulence df + body-force df + solid swap, one
//                         pass over the interior, reading buffer P(n-1) and writing P(n);
//   2. boundary kernels : the non-periodic ghost passes of boundarycond on the buffer just
//                         written, in the reference's pass order (closed i, j, k faces; edges and
//                         corners; inflow/outflow).
// The buffers hold post-collision populations (the reference's f1 after postcoll/solids) in
// the interior and the reference's persistent ghost layers.  The reference's post-streaming
// state f1 = drift(P(n)) is formed on the fly by the next step, and by state_out_kernel when the
// caller asks for it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>

#include "codec.h"
#include "lattice.h"
#include "layout.h"
#include "impl.h"

namespace lbm {

template <typename R>
struct StepArgs {
    Layout lay;
    const void* src;       // buffer holding P(n-1) (or the uploaded state when stream == 0)
    void* dst;             // buffer receiving P(n)
    int stream;            // 1: pull-stream from src; 0: collide the state in place
    R fac0;                // 1 - 1/tau0 when Vreman is off (tau0 = 3*kinevisc + 0.5)
    R kinevisc, vconst, eps, tiny;
    const R* ablvisc;      // [nz]
    const R* ext_rho;      // moments to use instead of the computed ones (first step after
    const R* ext_u;        //   an equilibrium initialisation, src/main.F90:214,243,299);
    const R* ext_v;        //   cell layout, nullptr otherwise
    const R* ext_w;
    const uint32_t* solid; // one bit per interior cell in rows along the contiguous axis,
                           //   solid_wpr words per row; nullptr: no solids
    int solid_wpr;
    int ip;                // inflow-turbulence slice (reference: 10)
    const R* tu;           // turbulence slice (ny*nz, j fastest) for this step; nullptr: off
    const R* tv;
    const R* tw;
    R ampl;
    R* avg;                // six cell planes: sums of u, v, w, u^2, v^2, w^2 of the streamed
                           //   input (the moments after the previous step); nullptr: off
    const R* Fx;           // external forcing (cell layout); nullptr: off
    const R* Fy;
    const R* Fz;
    int flo[3], fhi[3];    // cells (i, j, k; inclusive) whose force is read (the others have none)
    const R* theta;        // potential temperature (cell layout) of the buoyancy forcing; nullptr: off
    R btheta0, bscale;     // its reference temperature and g * time^2 / length
    R* tau_out;            // cell layout; nullptr: tau not stored
    int fold_dir;          // closed axis whose face pass this kernel performs (-1: none)
    int fold_opt_lo, fold_opt_hi;
    // inflow/outflow planes of i written by the cells at i = 1 and i = nx (0: separate pass)
    int inflow_fold;
    const R* shear;        // uvel_shear(nz)
    R utime, cosd, sind, rho0;
    int heavy_first;       // bit 0 / bit 1: cells 1 and n of ax1 / ax2 are dispatched first
    int halo_dir;          // axis whose ghost planes are halos filled by lbm_halo_unpack (-1: none)
    int xcd_groups;        // 1: the blocks of one row group run on one XCD (step_kernel)
    // split step along the halo axis (step_kernel): 0 every cell; 1 the part_edge cells next to
    // each halo plane (the launch covers 2 * part_edge along part_axis); 2 the others (the launch
    // covers n - 2 * part_edge).  part_axis: 1 the halo axis is ax1, 2 it is ax2.
    int part, part_axis, part_edge;
};

// Signed sum  sx*a + sy*b + sz*c  for compile-time signs in {-1,0,1}: adds only.  With no term the
// result is -0, the additive identity, so that a sum it starts folds away.
template <typename R>
__host__ __device__ __forceinline__ R sdot(int sx, int sy, int sz, R a, R b, R c) {
    R s = -R(0);
    bool first = true;
    if (sx != 0) { s = sx > 0 ? a : -a; first = false; }
    if (sy != 0) { s = first ? (sy > 0 ? b : -b) : (sy > 0 ? s + b : s - b); first = false; }
    if (sz != 0) { s = first ? (sz > 0 ? c : -c) : (sz > 0 ? s + c : s - c); }
    return s;
}

// Reciprocal: the hardware's (1 ulp) in FP32 device code, the IEEE quotient otherwise.  Used for
// 1/rho, 1/tau and the Vreman ratio, where the FP32 quotient's correction sequence costs about
// ten instructions per use in an issue-bound kernel.
template <typename R>
__host__ __device__ __forceinline__ R recip(R x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<R, float>::value) return __builtin_amdgcn_rcpf(x);
#endif
    return R(1) / x;
}

// Square root: the hardware's in FP32 device code (the Vreman ratio is a normal number wherever it
// is used), the IEEE root otherwise.  The adjoint uses it; the forward collision keeps the IEEE
// root (a one-ulp change of the eddy viscosity moves the realization of a turbulent run).
template <typename R>
__host__ __device__ __forceinline__ R sqrt_r(R x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<R, float>::value) return __builtin_amdgcn_sqrtf(x);
#endif
    using std::sqrt;   // the tangent type brings its own sqrt
    return sqrt(x);
}

// Hermite equilibrium of direction l (src/m_postcoll_kernel.F90:136-183 in closed form):
//   feq = w rho (1 - 1.5 u.u + cu (3 + 4.5 cu [+ 4.5 (cu^2 - u.u)]))   with cu = c.u
// The bracket is the reference's full H3:A3 contraction, 4.5 rho cu (cu^2 - u.u), applied for
// l > 0 when the third order is on.
template <int Q, bool ORDER3, typename R>
__device__ __forceinline__ R feq_dir(int l, R rho, R ux, R uy, R uz, R usq) {
    using Lt = Lattice<Q>;
    const R cu = sdot<R>(Lt::cx(l), Lt::cy(l), Lt::cz(l), ux, uy, uz);
    R poly = R(3) + R(4.5) * cu;
    if (ORDER3 && l > 0) poly += R(4.5) * (cu * cu - usq);
    return (weight<Q, R>(l) * rho) * ((R(1) - R(1.5) * usq) + cu * poly);
}

// Reflected direction for a wall normal to axis d: normal component negated, tangential
// components multiplied by opt (-1 no-slip, +1 free-slip).
template <int Q>
__host__ __device__ constexpr int mirror(int l, int d, int opt) {
    using Lt = Lattice<Q>;
    return Lt::find(d == 0 ? -Lt::cx(l) : opt * Lt::cx(l), d == 1 ? -Lt::cy(l) : opt * Lt::cy(l),
                    d == 2 ? -Lt::cz(l) : opt * Lt::cz(l));
}

// The per-cell part of a step: moments, collision, turbulence and body-force increments, solid
// swap.  rv holds the streamed populations of cell (i,j,k) (storage coordinates ca,cb,cc, device
// cell offset oc); dev receives the departures S*(f_l - w_l) of its post-collision populations
// (S = C::kDepScale), tau the Vreman relaxation time.
//
// Directions come in opposite pairs (l, l+1) for odd l (src/mod_D3Q27setup.F90:31,43).  Every
// quantity the collision forms per direction is even or odd in c, so each pair is evaluated
// once: moments accumulate from D_l +- D_l+1 and the outputs are (even) +- (odd), with
//   feq_l - w_l = w (rho*base - 1 + 4.5 rho cu^2) + w rho cu (3 [+ 4.5 (cu^2 - u.u)])
//   (src/m_postcoll_kernel.F90:136-183 in closed form; base = 1 - 1.5 u.u; the bracket is the
//   reference's full H3:A3 contraction, applied for l > 0 when the third order is on)
// and the regularized non-equilibrium (src/m_postcoll_kernel.F90:189-244)
//   fneq_l = 4.5 w [ (cAc - T/3) + (cu (3 cAc - T) - 2 c.g) ],  T = tr a2,  g = a2 u.
// Working on departures needs no decode/encode arithmetic for the shifted 16-bit formats and
// forms rho = 1 + sum D/S, j = sum c D/S without the cancellation of sum f.
// Per-cell inputs of the collision besides the populations: the ABL viscosity of the layer and the
// word of the row's solid mask holding the cell's bit.  For 16-bit storage the step kernel reads
// them ahead of the population loads, so that their latency overlaps those loads' instead of
// stalling the collision where they are used (measured on MI355X, D3Q27 reg+Vreman 384x192x128:
// FP16 214 -> 198 us, BF16 215 -> 196 us per step kernel; FP32 352 -> 358 us, so 4- and 8-byte
// storage read them where they are used).  The external force stays where it is used: read ahead,
// its three registers would be held through every collision, forced or not (they cost the D3Q19
// BGK kernel its fifth wave per SIMD).
template <typename R>
struct CellAux {
    R abl;
    uint32_t sword;
};
template <bool VRE, typename R>
__host__ __device__ __forceinline__ CellAux<R> cell_aux(const StepArgs<R>& a, int k, int ca, int cb, int cc) {
    const Layout& L = a.lay;
    CellAux<R> x{R(0), 0u};
    if (VRE) x.abl = a.ablvisc[k - 1];
    if (a.solid != nullptr) x.sword = a.solid[((cc - 1) * L.n(L.ax1) + (cb - 1)) * a.solid_wpr + ((ca - 1) >> 5)];
    return x;
}

// AUX: the per-cell inputs come from aux (read ahead by the caller); otherwise they are read here
// where they are used.
template <int Q, class C, bool BGK3, bool HRR, bool VRE, bool AUX = false>
__host__ __device__ __forceinline__ void collide_cell(const StepArgs<typename C::real>& a, const typename C::raw (&rv)[Q],
                                                      const CellAux<typename C::real>& aux, int i, int j, int k, int ca,
                                                      int cb, int cc, int oc, typename C::real (&dev)[Q],
                                                      typename C::real& tau_out) {
    using R = typename C::real;
    using Lt = Lattice<Q>;
    const Layout& L = a.lay;
    constexpr int NP = (Q - 1) / 2;
    constexpr R S = C::kDepScale;
    constexpr R invS = R(1) / S;

    R D[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) D[l] = C::template dep<Q>(rv[l], l);
    // accumulators start at -0, the additive identity the compiler may fold (0 + x is not x for
    // x = -0)
    const R z0 = -R(0);
    R jx = z0, jy = z0, jz = z0;
    R pxy = z0, pxz = z0, pyz = z0;
    // pair sums by the axes c spans (a face axis, an edge plane, the corners): the density and
    // the diagonal second moments are sums of these (a corner counts in all three diagonals, an
    // edge plane in two)
    R sfx = z0, sfy = z0, sfz = z0, sexy = z0, sexz = z0, seyz = z0, scn = z0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int l = 2 * p + 1;
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        const int n2 = Lt::norm2(l);
        const R sp = D[l] + D[l + 1];
        const R dp = D[l] - D[l + 1];
        if (n2 == 1) (cx != 0 ? sfx : cy != 0 ? sfy : sfz) += sp;
        else if (n2 == 2) (cz == 0 ? sexy : cy == 0 ? sexz : seyz) += sp;
        else scn += sp;
        if (cx > 0) jx += dp; else if (cx < 0) jx -= dp;
        if (cy > 0) jy += dp; else if (cy < 0) jy -= dp;
        if (cz > 0) jz += dp; else if (cz < 0) jz -= dp;
        if (HRR || VRE) {
            if (cx * cy > 0) pxy += sp; else if (cx * cy < 0) pxy -= sp;
            if (cx * cz > 0) pxz += sp; else if (cx * cz < 0) pxz -= sp;
            if (cy * cz > 0) pyz += sp; else if (cy * cz < 0) pyz -= sp;
        }
    }
    R dr = D[0] + (((sfx + sfy) + sfz) + ((sexy + sexz) + seyz) + scn);
    const R pxx = sfx + sexy + sexz + scn, pyy = sfy + sexy + seyz + scn, pzz = sfz + sexz + seyz + scn;
    dr *= invS;
    R rh = R(1) + dr;
    const R invrho = recip(rh);
    R ux = (jx * invS) * invrho, uy = (jy * invS) * invrho, uz = (jz * invS) * invrho;
    if (a.avg != nullptr) {
        // time averaging of the moments after the previous step (src/m_averaging_full_kernel.F90:33-39)
        // (the six sums are read before any is written: one memory round trip instead of six)
        R* p = a.avg + oc;
        const size_t P = (size_t)L.cplane;
        const R s0 = p[0], s1 = p[P], s2 = p[2 * P], s3 = p[3 * P], s4 = p[4 * P], s5 = p[5 * P];
        p[0] = s0 + ux;
        p[P] = s1 + uy;
        p[2 * P] = s2 + uz;
        p[3 * P] = s3 + ux * ux;
        p[4 * P] = s4 + uy * uy;
        p[5 * P] = s5 + uz * uz;
    }
    if (a.ext_rho) { rh = a.ext_rho[oc]; ux = a.ext_u[oc]; uy = a.ext_v[oc]; uz = a.ext_w[oc]; }
    const R usq = ux * ux + uy * uy + uz * uz;
    const R base = R(1) - R(1.5) * usq;
    const R rb1 = rh * base - R(1);      // rho*base - 1
    const R r45 = R(4.5) * rh;

    // Non-equilibrium second-order tensor a2 = sum_l H2_l (f_l - feq_l)
    //   = Pi' - rho u u - (sum_l (f_l - w_l)) I/3   (Pi' = sum_l c c (f_l - w_l)),
    // exact for D3Q19/D3Q27 because sum_l c c feq_l = rho u u + rho I/3.  For the regularized
    // populations the same tensor is the Vreman alpha (src/m_postcoll_kernel.F90:255-268), since
    // sum_l w_l H2_l H2_l : a2 * 4.5 = a2 on these lattices.
    R axx = 0, ayy = 0, azz = 0, axy = 0, axz = 0, ayz = 0;
    if (HRR || VRE) {
        const R third = R(1) / R(3);
        const R rux = rh * ux, ruy = rh * uy, ruz = rh * uz;
        axx = pxx * invS - rux * ux - dr * third;
        ayy = pyy * invS - ruy * uy - dr * third;
        azz = pzz * invS - ruz * uz - dr * third;
        axy = pxy * invS - rux * uy;
        axz = pxz * invS - rux * uz;
        ayz = pyz * invS - ruy * uz;
    }

    R fac = a.fac0;
    R tau = 0;
    if (VRE) {
        // Vreman (2004) eddy viscosity as src/m_postcoll_kernel.F90:270-304 forms it.
        R amag = axx * axx + ayy * ayy + azz * azz + R(2) * (axy * axy + axz * axz + ayz * ayz);
        amag = amag > a.tiny ? amag : a.tiny;
        const R b11 = a.eps + (axx * axx + axy * axy + axz * axz);
        const R b22 = a.eps + (axy * axy + ayy * ayy + ayz * ayz);
        const R b33 = a.eps + (axz * axz + ayz * ayz + azz * azz);
        const R b12 = a.eps + (axx * axy + axy * ayy + axz * ayz);
        const R b13 = a.eps + (axx * axz + axy * ayz + axz * azz);
        const R b23 = a.eps + (axy * axz + ayy * ayz + ayz * azz);
        const R bb = b11 * b22 - b12 * b12 + b11 * b33 - b13 * b13 + b22 * b33 - b23 * b23;
        const R t = bb * recip(amag);
        using std::sqrt;   // the tangent type brings its own sqrt
        const R eddy = t > a.eps ? a.vconst * sqrt(t) : R(0);
        tau = R(3) * (a.kinevisc + eddy + (AUX ? aux.abl : a.ablvisc[k - 1])) + R(0.5);
        fac = R(1) - recip(tau);
    }

    const R T = axx + ayy + azz;
    const R gx = axx * ux + axy * uy + axz * uz;
    const R gy = axy * ux + ayy * uy + ayz * uz;
    const R gz = axz * ux + ayz * uy + azz * uz;

    {   // rest population
        const R wS = weight<Q, R>(0) * S;
        const R e = wS * rb1;
        const R ne = HRR ? -(R(1.5) * wS) * T : D[0] - e;
        dev[0] = e + fac * ne;
    }
    // Factors shared by every direction of a weight class (|c|^2 = 1, 2, 3; directions 1, 7, 19
    // are the first of each class), hoisted out of the pair loop:
    //   even part  ev = kE0 + kE2 cu^2 + kF cAc
    //   odd part   od = kO cu pol + kF (cu (3 cAc - T) + c.(-2 g))
    //                 = cu (kA + kB cu^2 + kC cAc) + kF c.(-2 g)      (regularized)
    // with kE2 = wS*4.5rho, kO = wS*rho, kF = fac*4.5wS, kE0 = wS(rho base - 1) - kF T/3,
    // pol = 3 - 4.5 u.u + 4.5 cu^2 (third order) or 3, so kA = kO (3 - 4.5 u.u) - kF T or
    // 3 kO - kF T, kB = 4.5 kO or 0, kC = 3 kF.
    constexpr int NC = Q == 27 ? 3 : 2;
    R kE0[NC + 1], kE2[NC + 1], kO[NC + 1], kF[NC + 1], kA[NC + 1], kB[NC + 1], kC[NC + 1];
    const R T3 = T * (R(1) / R(3));
    const R P0 = R(3) - R(4.5) * usq;
#pragma unroll
    for (int n = 1; n <= NC; ++n) {
        const R wS = weight<Q, R>(n == 1 ? 1 : n == 2 ? 7 : 19) * S;
        kE2[n] = wS * r45;
        kO[n] = wS * rh;
        kF[n] = fac * (R(4.5) * wS);
        kE0[n] = HRR ? wS * rb1 - kF[n] * T3 : wS * rb1;
        kA[n] = (BGK3 ? kO[n] * P0 : R(3) * kO[n]) - kF[n] * T;
        kB[n] = R(4.5) * kO[n];
        kC[n] = R(3) * kF[n];
    }
    const R a2xy = R(2) * axy, a2xz = R(2) * axz, a2yz = R(2) * ayz;
    const R gmx = R(-2) * gx, gmy = R(-2) * gy, gmz = R(-2) * gz;
    // c.A.c by the axes c spans: one axis a_dd; two axes (a_dd + a_ee) +- 2 a_de; three axes
    // T + sxy 2a_xy + sxz 2a_xz + syz 2a_yz, whose signs multiply to +1, so with sxy fixed the
    // other two are (s, s) or (s, -s): T +- 2a_xy +- (2a_xz +- 2a_yz)
    const R exy = axx + ayy, exz = axx + azz, eyz = ayy + azz;
    const R tpp = T + a2xy, tmm = T - a2xy, up = a2xz + a2yz, um = a2xz - a2yz;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int l = 2 * p + 1;
        const int cx = Lt::cx(l), cy = Lt::cy(l), cz = Lt::cz(l);
        const int n2 = Lt::norm2(l);
        const R cu = sdot<R>(cx, cy, cz, ux, uy, uz);
        const R cu2 = cu * cu;
        if (HRR) {
            R cac;
            if (n2 == 1) cac = cx != 0 ? axx : cy != 0 ? ayy : azz;
            else if (n2 == 2) cac = cz == 0 ? (cx * cy > 0 ? exy + a2xy : exy - a2xy)
                                  : cy == 0 ? (cx * cz > 0 ? exz + a2xz : exz - a2xz)
                                            : (cy * cz > 0 ? eyz + a2yz : eyz - a2yz);
            else cac = cx * cy > 0 ? (cx * cz > 0 ? tpp + up : tpp - up) : (cx * cz > 0 ? tmm + um : tmm - um);
            const R ev = kE0[n2] + kE2[n2] * cu2 + kF[n2] * cac;
            const R in = BGK3 ? kA[n2] + kB[n2] * cu2 + kC[n2] * cac : kA[n2] + kC[n2] * cac;
            const R od = cu * in + kF[n2] * sdot<R>(cx, cy, cz, gmx, gmy, gmz);
            dev[l] = ev + od;
            dev[l + 1] = ev - od;
        } else {
            const R pol = BGK3 ? P0 + R(4.5) * cu2 : R(3);
            const R cp = cu * pol;
            const R e = kE0[n2] + kE2[n2] * cu2;   // S (E - w)
            const R o = kO[n2] * cp;               // S O
            dev[l] = (e + o) + fac * (D[l] - (e + o));
            dev[l + 1] = (e - o) + fac * (D[l + 1] - (e - o));
        }
    }

    // Equilibrium increments feq(rho, v) - feq(rho, u) for the inflow-turbulence forcing on
    // slice ip (src/m_inflow_turbulence_forcing.F90:57-128, added after the collision as
    // src/m_inflow_turbulence_apply.F90:21-26; third order when ibgk = 3) and for the external
    // body force (Kupershtokh exact difference with the 2nd-order equilibrium and du = -F/rho,
    // src/m_forcings_apply_kernel.F90:142-271; a cell with rho <= 0 is skipped).
    auto add_feq_diff = [&](R vx, R vy, R vz, auto order3) {
        constexpr bool O3 = decltype(order3)::value;
        const R vsq = vx * vx + vy * vy + vz * vz;
        const R vrb1 = rh * (R(1) - R(1.5) * vsq) - R(1);
        dev[0] = dev[0] + weight<Q, R>(0) * S * (vrb1 - rb1);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int l = 2 * p + 1;
            const R wS = weight<Q, R>(l) * S;
            const R cv = sdot<R>(Lt::cx(l), Lt::cy(l), Lt::cz(l), vx, vy, vz);
            const R cu = sdot<R>(Lt::cx(l), Lt::cy(l), Lt::cz(l), ux, uy, uz);
            R pv = R(3), pu = R(3);
            if (O3) { pv += R(4.5) * (cv * cv - vsq); pu += R(4.5) * (cu * cu - usq); }
            const R de = wS * ((vrb1 + r45 * cv * cv) - (rb1 + r45 * cu * cu));
            const R dodd = wS * (rh * (cv * pv) - rh * (cu * pu));
            dev[l] = dev[l] + (de + dodd);
            dev[l + 1] = dev[l + 1] + (de - dodd);
        }
    };
    if (a.tu != nullptr && i == a.ip) {
        const int sl = (k - 1) * L.ny + (j - 1);
        add_feq_diff(ux + a.ampl * a.tu[sl], uy + a.ampl * a.tv[sl], uz + a.ampl * a.tw[sl],
                     std::integral_constant<bool, BGK3>());
    }
    // force box: every interior cell for the caller's field, the deposit box for turbines (which the
    // tangent storage never has; the test is compiled out there, where it trips the register allocator)
    bool in_fbox = true;
    if constexpr (sizeof(R) <= 8)
        in_fbox = i >= a.flo[0] && i <= a.fhi[0] && j >= a.flo[1] && j <= a.fhi[1] && k >= a.flo[2] && k <= a.fhi[2];
    // One increment for both producers (a second inlined copy of add_feq_diff costs the unforced
    // kernels registers: D3Q19 BGK FP16 71 -> 101 VGPRs).
    if ((a.Fx != nullptr || a.theta != nullptr) && rh > R(0)) {
        R Fx = 0, Fy = 0, Fz = 0;
        bool forced = false;
        if (a.Fx != nullptr && in_fbox) {
            Fx = a.Fx[oc];
            Fy = a.Fy[oc];
            Fz = a.Fz[oc];
            forced = true;
        }
        if (a.theta != nullptr) {
            // buoyancy added to the forcing of every interior cell (src/m_buoyancy_forcing_kernel.F90:37,
            // the reference's operation order): F_z - scaling * (theta - theta0) / theta0
            Fz = Fz - (a.bscale * (a.theta[oc] - a.btheta0)) / a.btheta0;
            forced = true;
        }
        if (forced) {
            const R ir = recip(rh);
            add_feq_diff(ux - Fx * ir, uy - Fy * ir, uz - Fz * ir, std::integral_constant<bool, false>());
        }
    }

    // Solid cells: swap the opposite pairs (2,3),(4,5),... (src/m_solids.F90:19-32).
    if (a.solid != nullptr) {
        // rows of bits along the contiguous axis
        const uint32_t word = AUX ? aux.sword : a.solid[((cc - 1) * L.n(L.ax1) + (cb - 1)) * a.solid_wpr + ((ca - 1) >> 5)];
        if ((word >> ((ca - 1) & 31)) & 1u) {
#pragma unroll
            for (int l = 1; l + 1 < Q; l += 2) {
                const R t = dev[l];
                dev[l] = dev[l + 1];
                dev[l + 1] = t;
            }
        }
    }
    tau_out = tau;
}
template <int Q, class C, bool BGK3, bool HRR, bool VRE>
__host__ __device__ __forceinline__ void collide_cell(const StepArgs<typename C::real>& a, const typename C::raw (&rv)[Q], int i,
                                                      int j, int k, int ca, int cb, int cc, int oc,
                                                      typename C::real (&dev)[Q], typename C::real& tau_out) {
    collide_cell<Q, C, BGK3, HRR, VRE, false>(a, rv, CellAux<typename C::real>{}, i, j, k, ca, cb, cc, oc, dev, tau_out);
}

// The closed-face pass of axis D (src/m_boundary_{i,j,k}_closed_kernel.F90) performed by the
// step kernel for the cells next to the walls, when D is the only closed axis (the other axes
// periodic, or i carrying the inflow/outflow planes, whose kernel runs after this one).  For
// ghost cell G of the face and outgoing l the reference sets
//   f1(m, G) = f2(l, G)                 (m = l reflected: normal negated, tangential * opt)
//   f1(l, G) = f1(l, G - c_l tangentially, on the wall plane).
// The cell x above G does the first; the second is pushed by the cell G - c_l that computed
// the value (periodic tangential axes wrapped), or, when G - c_l is a stored ghost of the
// inflow axis, copied by x from that ghost (still holding its value from before this step).
template <int Q, class C>
__device__ __forceinline__ void fold_closed_face(const StepArgs<typename C::real>& a, int i, int j, int k,
                                                 const typename C::real (&dev)[Q]) {
    using Lt = Lattice<Q>;
    using raw = typename C::raw;
    const Layout& L = a.lay;
    const int D = a.fold_dir;
    const int x[3] = {i, j, k};
    const int n[3] = {L.nx, L.ny, L.nz};
    const int per[3] = {L.per_i, L.per_j, L.per_k};
    const raw* __restrict__ f2 = static_cast<const raw*>(a.src);
    raw* __restrict__ f1 = static_cast<raw*>(a.dst);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int plane = side ? n[D] : 1;
        if (x[D] != plane) continue;
        const int ghost = side ? n[D] + 1 : 0;
        const int opt = side ? a.fold_opt_hi : a.fold_opt_lo;
        int g[3] = {i, j, k};
        g[D] = ghost;
        const int og = L.cell(g[0], g[1], g[2]);
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const int cn = Lt::c(l, D);
            if (!(side ? cn > 0 : cn < 0)) continue;
            const int m = opt < 0 ? mirror<Q>(l, D, -1) : mirror<Q>(l, D, +1);
            f1[(size_t)m * L.plane + og] = f2[(size_t)l * L.plane + og];
            // push: the ghost this population streams to
            int t[3], s[3];
            bool push = true, from_inflow = false, from_halo = false;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (d == D) { t[d] = ghost; s[d] = plane; continue; }
                const int cd = Lt::c(l, d);
                t[d] = x[d] + cd;
                s[d] = x[d] - cd;
                if (per[d]) {
                    t[d] = Layout::wrap(t[d], n[d], 1);
                    s[d] = Layout::wrap(s[d], n[d], 1);
                } else {
                    if (t[d] < 1 || t[d] > n[d]) push = false;
                    if (s[d] < 1 || s[d] > n[d]) {
                        if (d == a.halo_dir) from_halo = true;
                        else from_inflow = true;
                    }
                }
            }
            if (push) f1[(size_t)l * L.plane + L.cell(t[0], t[1], t[2])] = C::template enc_dep<Q>(dev[l], l);
            // G - c_l outside the face's range.  On a halo plane only: the neighbour's value of this
            // step, pushed by that halo's unpack.  On the inflow axis only: the ghost's value from
            // before this step, copied by the cell that rewrites the ghost when the inflow planes are
            // folded too, else here.  On both (the inflow ghost of a halo plane): the value that
            // ghost holds now, the neighbour's inflow ghost from before this step.
            if (from_inflow && (from_halo || !a.inflow_fold))
                f1[(size_t)l * L.plane + og] = f1[(size_t)l * L.plane + L.cell(s[0], s[1], s[2])];
        }
    }
}

// Inflow ghost values (src/m_boundary_i_inflow_kernel.F90:61-74) in place over the decoded
// populations of cell (1, j, k): fg_l - 2 w_l rho0 (c_x u cos(udir) + c_y u sin(udir)) / cs^2.
// The ghost takes fg_l for c_x <= 0 and fg of the opposite direction otherwise (:89-104).
template <int Q, typename R>
__host__ __device__ __forceinline__ void inflow_ghost_values(R (&fg)[Q], R uu, R cosd, R sind, R rho0) {
    using Lt = Lattice<Q>;
    const R cs2 = R(1) / R(3);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const R wl = weight<Q, R>(l);
        const R cxl = R(Lt::cx(l)), cyl = R(Lt::cy(l));
        fg[l] = fg[l] - R(2) * wl * rho0 * (cxl * uu * cosd + cyl * uu * sind) / cs2;
    }
}

// The inflow/outflow pass for the ghost row (j, k) of i, performed by the cell (i, j, k) with
// i = 1 (inflow ghost i = 0) or i = nx (outflow ghost i = nx+1, a copy of the cell) from the
// values it has just stored.  When the closed face of fold_dir is folded as well, the face pass
// reads this ghost's value from before the step (f1(l, G - c_l) on the wall plane, G - c_l on
// the inflow axis): those copies are made here, before the ghost is rewritten.
template <int Q, class C>
__device__ __forceinline__ void fold_inflow(const StepArgs<typename C::real>& a, int i, int j, int k,
                                            const typename C::real (&dev)[Q]) {
    using R = typename C::real;
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    const Layout& L = a.lay;
    raw* __restrict__ f1 = static_cast<raw*>(a.dst);
    const int gi = i == 1 ? 0 : L.nx + 1;
    const int og = L.cell(gi, j, k);
    const int D = a.fold_dir;
    if (D > 0) {
        const int x[3] = {i, j, k};
        const int n[3] = {L.nx, L.ny, L.nz};
        const int O = 3 - D;   // the other tangential axis: periodic or a halo axis
        const int perO = O == 1 ? L.per_j : L.per_k;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int plane = side ? n[D] : 1;
            if (x[D] != plane) continue;
#pragma unroll
            for (int l = 0; l < Q; ++l) {
                const int cn = Lt::c(l, D);
                if (!(side ? cn > 0 : cn < 0)) continue;
                // the target ghost G = (gi + c_x, x_O + c_O, ghost of D) lies in the face's range
                if (Lt::cx(l) != (i == 1 ? 1 : -1)) continue;
                int t[3];
                t[0] = i;
                t[D] = side ? n[D] + 1 : 0;
                t[O] = x[O] + Lt::c(l, O);
                if (perO) t[O] = Layout::wrap(t[O], n[O], 1);
                else if (t[O] < 1 || t[O] > n[O]) continue;   // a ghost of the halo axis
                f1[(size_t)l * L.plane + L.cell(t[0], t[1], t[2])] = f1[(size_t)l * L.plane + og];
            }
        }
    }
    if (i == 1) {
        R fg[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) fg[l] = C::template decode<Q>(C::template enc_dep<Q>(dev[l], l), l);
        inflow_ghost_values<Q, R>(fg, a.utime * a.shear[k - 1], a.cosd, a.sind, a.rho0);
#pragma unroll
        for (int l = 0; l < Q; ++l)
            f1[(size_t)l * L.plane + og] = C::template encode<Q>(Lt::cx(l) <= 0 ? fg[l] : fg[Lt::opposite(l)], l);
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) f1[(size_t)l * L.plane + og] = C::template enc_dep<Q>(dev[l], l);
    }
}

#if defined(__HIP_DEVICE_COMPILE__)
// Raw buffer access of 2- or 4-byte values: base (scalar), byte offset voff (vector), byte offset
// soff (scalar), bounds nbytes.  The descriptor word 3 is the gfx9-family raw-buffer format.
template <typename raw>
__device__ __forceinline__ raw buffer_load(const void* base, int nbytes, uint32_t voff, uint32_t soff) {
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, nbytes, 0x00020000);
    if constexpr (sizeof(raw) == 2) return __builtin_bit_cast(raw, __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0));
    else return __builtin_bit_cast(raw, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
// NT: the non-temporal cache policy (aux bit 1), as the pointer path's non-temporal stores.
template <typename raw, bool NT>
__device__ __forceinline__ void buffer_store(void* base, int nbytes, uint32_t voff, uint32_t soff, raw v) {
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(base, 0, nbytes, 0x00020000);
    if constexpr (sizeof(raw) == 2)
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)voff, (int)soff, NT ? 2 : 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff, NT ? 2 : 0);
}
#endif

// One cell of the step at storage coordinates (ca, cb, cc) (1-based along ax0, ax1, ax2):
// gather the streamed populations, collide, store, and the folded face pass.
// VAR bit 0: non-temporal stores, bit 1: non-temporal loads (measured slower; kept for tuning),
// bit 3: the buffer is below 4 GiB (32-bit offsets from one base), bit 4: row-blocked buffers of
// any size addressed per slab of the slowest axis (one scalar base per source slab, 32-bit offsets
// within it; needs AXS >= 0 and a slab below 4 GiB).
// AXS >= 0: the contiguous storage axis is AXS (known at compile time, the other two in
// make_layout's order), so i, j, k map to (ca, cb, cc) without run-time selects, and the caller's
// wave-uniform cb, cc keep the row's offsets, wraps and checks in scalar registers.
template <int Q, class C, bool BGK3, bool HRR, bool VRE, int VAR, int AXS = -1>
__device__ __forceinline__ void step_cell(const StepArgs<typename C::real>& a, int ca, int cb, int cc) {
    using R = typename C::real;
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    const Layout& L = a.lay;
    if (ca > L.n(L.ax0) || cb > L.n(L.ax1) || cc > L.n(L.ax2)) return;
    int i, j, k;
    constexpr int A1 = AXS == 0 ? 1 : 0, A2 = AXS == 2 ? 1 : 2;
    if constexpr (AXS < 0) {
        L.ijk(ca, cb, cc, i, j, k);
    } else {
        i = AXS == 0 ? ca : A1 == 0 ? cb : cc;
        j = AXS == 1 ? ca : A1 == 1 ? cb : cc;
        k = AXS == 2 ? ca : A1 == 2 ? cb : cc;
    }

    // Source offsets: population l comes from (i - cx, j - cy, k - cz), periodic directions
    // folded onto the interior.  Index d = c + 1.
    // The cell's own offset is formed once; the neighbours' by adding the uniform stride, and
    // for a periodic axis at its first/last cell by the uniform wrap distance (no per-offset
    // multiplies).
    const int s = a.stream;
    // the contiguous axis has unit stride in both buffer arrangements (make_layout)
    const int oc0 = AXS < 0 ? L.cell(i, j, k)
                            : L.base + ca + cb * (A1 == 0 ? L.si : A1 == 1 ? L.sj : L.sk) +
                                  cc * (A2 == 0 ? L.si : A2 == 1 ? L.sj : L.sk);
    int ox[3], oy[3], oz[3];
    auto offsets = [&](int x, int n, int per, int st, int (&o)[3]) {
        // o[d]: offset change of the source x - s*(d-1) along this axis
        const int wrapd = per ? n * st : 0;
        const int lo = (per && x == n) ? -wrapd : 0;   // x + 1 wraps to 1
        const int hi = (per && x == 1) ? wrapd : 0;    // x - 1 wraps to n
        o[0] = s * (st + lo);                          // source x + 1 (c = -1)
        o[1] = 0;
        o[2] = -s * (st - hi);                         // source x - 1 (c = +1)
    };
    offsets(i, L.nx, L.per_i, L.si, ox);
    offsets(j, L.ny, L.per_j, L.sj, oy);
    offsets(k, L.nz, L.per_k, L.sk, oz);
    // slab form (VAR bit 4): offsets stay relative, the cell's own offset is split into its slab's
    // start and its offset within the slab
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr bool SLAB = (VAR & 16) != 0 && AXS >= 0 && (sizeof(raw) == 2 || sizeof(raw) == 4);
#else
    constexpr bool SLAB = false;
#endif
    if constexpr (!SLAB) {
        ox[0] += oc0;
        ox[1] += oc0;
        ox[2] += oc0;
    }
    // per-cell arrays, and (16-bit storage) the collision's per-cell inputs read ahead of the
    // populations
    const int occ = AXS < 0 ? L.cidx(i, j, k)
                            : L.base + ca + cb * (A1 == 0 ? L.ci : A1 == 1 ? L.cj : L.ck) +
                                  cc * (A2 == 0 ? L.ci : A2 == 1 ? L.cj : L.ck);
    constexpr bool AUX = sizeof(raw) == 2;
    CellAux<R> aux{};
    if constexpr (AUX) aux = cell_aux<VRE>(a, k, ca, cb, cc);
    // Plane bases are uniform (scalar registers); the per-lane part is a 32-bit byte offset
    // within a plane (planes are below 4 GiB), so loads and stores take the scalar-base form
    // without 64-bit address arithmetic per population.
    // With VAR bit 3 the whole buffer is below 4 GiB: one scalar base and 32-bit offsets that
    // include the plane (two scalar registers instead of one pair per plane).
    const char* __restrict__ srcb = static_cast<const char*>(a.src);
    const size_t pbytes = (size_t)L.plane * sizeof(raw);
    const uint32_t pb32 = (uint32_t)pbytes;
    raw rv[Q];
    // 2- and 4-byte storage in a buffer below 4 GiB: raw buffer loads/stores, the plane offset in
    // the scalar offset operand, so an access costs no vector address arithmetic at all.
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr bool BUF = (VAR & 8) != 0 && (sizeof(raw) == 2 || sizeof(raw) == 4);
#else
    constexpr bool BUF = false;
#endif
    // descriptor byte count: an unsigned 32-bit field (the variant runs only below 4 GiB)
    const size_t tbytes = (size_t)L.total * sizeof(raw);
    const int nbytes = (int)(uint32_t)(tbytes < 0xffffffffull ? tbytes : 0xffffffffull);
    // slab form: the element offset of the cell within its slab of the slowest axis and the slab's
    // start; a source's slab start moves by the ax2 offset (wrap included), its offset within the
    // slab by the ax0 and ax1 offsets
    const int s2 = A2 == 0 ? L.si : A2 == 1 ? L.sj : L.sk;
    const int (&d0)[3] = AXS == 0 ? ox : AXS == 1 ? oy : oz;
    const int (&d1)[3] = A1 == 0 ? ox : A1 == 1 ? oy : oz;
    const int (&d2)[3] = A2 == 0 ? ox : A2 == 1 ? oy : oz;
    const int w0 = SLAB ? oc0 - cc * s2 : 0;      // cell offset within its slab
    const size_t slab0 = SLAB ? (size_t)cc * (size_t)s2 : 0;
    const int sbytes = (int)((uint32_t)s2 * (uint32_t)sizeof(raw));
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if constexpr (SLAB) {
#if defined(__HIP_DEVICE_COMPILE__)
            const int c0 = Lt::c(l, AXS), c1 = Lt::c(l, A1), c2 = Lt::c(l, A2);
            const char* sb = srcb + (size_t)((long long)slab0 + d2[c2 + 1]) * sizeof(raw);
            rv[l] = buffer_load<raw>(sb, sbytes, (uint32_t)(w0 + d0[c0 + 1] + d1[c1 + 1]) * (uint32_t)sizeof(raw),
                                     (uint32_t)l * pb32);
#endif
            continue;
        }
        const uint32_t offe = (uint32_t)(oz[Lt::cz(l) + 1] + oy[Lt::cy(l) + 1] + ox[Lt::cx(l) + 1]);
        const uint32_t off = offe * (uint32_t)sizeof(raw);
        if constexpr (BUF) {
#if defined(__HIP_DEVICE_COMPILE__)
            rv[l] = buffer_load<raw>(srcb, nbytes, off, (uint32_t)l * pb32);
#endif
            continue;
        }
        const raw* p;
        if constexpr ((VAR & 8) != 0) p = reinterpret_cast<const raw*>(srcb + (off + (uint32_t)l * pb32));
        else p = reinterpret_cast<const raw*>(srcb + (size_t)l * pbytes + (size_t)offe * sizeof(raw));
        if constexpr ((VAR & 2) != 0) rv[l] = __builtin_nontemporal_load(p);
        else rv[l] = *p;
    }

    R out[Q];
    R tau = 0;
    const int oc = oc0;
    collide_cell<Q, C, BGK3, HRR, VRE, AUX>(a, rv, aux, i, j, k, ca, cb, cc, occ, out, tau);

    char* __restrict__ dstb = static_cast<char*>(a.dst);
    const uint32_t ob = (uint32_t)oc * (uint32_t)sizeof(raw);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if constexpr (SLAB) {
#if defined(__HIP_DEVICE_COMPILE__)
            buffer_store<raw, (VAR & 1) != 0>(dstb + slab0 * sizeof(raw), sbytes, (uint32_t)w0 * (uint32_t)sizeof(raw),
                                              (uint32_t)l * pb32, C::template enc_dep<Q>(out[l], l));
#endif
            continue;
        }
        if constexpr (BUF) {
#if defined(__HIP_DEVICE_COMPILE__)
            buffer_store<raw, (VAR & 1) != 0>(dstb, nbytes, ob, (uint32_t)l * pb32, C::template enc_dep<Q>(out[l], l));
#endif
            continue;
        }
        raw* p;
        if constexpr ((VAR & 8) != 0) p = reinterpret_cast<raw*>(dstb + (ob + (uint32_t)l * pb32));
        else p = reinterpret_cast<raw*>(dstb + (size_t)l * pbytes + (size_t)(uint32_t)oc * sizeof(raw));
        if constexpr ((VAR & 1) != 0) __builtin_nontemporal_store(C::template enc_dep<Q>(out[l], l), p);
        else *p = C::template enc_dep<Q>(out[l], l);
    }
    if (VRE && a.tau_out != nullptr) a.tau_out[occ] = tau;
    if (a.fold_dir >= 0) fold_closed_face<Q, C>(a, i, j, k, out);
    if (a.inflow_fold && (i == 1 || i == L.nx)) fold_inflow<Q, C>(a, i, j, k, out);
}

// Threads run along the storage axes: x along the contiguous axis ax0, y along ax1, z along ax2.
// heavy_first bit 0 (ax1) / bit 1 (ax2): that axis's cells are taken in the order 1, n, 2, 3, ..,
// n-1, so that the rows and planes of a folded face or of the folded inflow/outflow planes, which
// carry extra stores, are dispatched first (blocks start in x, y, z order) instead of forming the
// launch's tail.
__device__ __forceinline__ int heavy_order(int x, int n) { return x == 0 ? 1 : x == 1 ? n : x; }

// With xcd_groups (several blocks along ax0 per row group), the blocks of one row group run on one
// XCD: workgroups are dispatched to the 8 XCDs in turn, so hardware block b (x fastest) works on
// row group ((b / 8) / gx) * 8 + b % 8, segment (b / 8) % gx.  The segments of a row group share the
// cache lines at their edges (the +-1 shift along ax0), which another XCD's L2 would fetch again;
// consecutive row groups still go to different XCDs.  Blocks past the last full set of 8 row
// groups keep their own position.
// AXS >= 0: step_cell's compile-time contiguous axis; the row coordinate is then taken as wave-uniform
// (block_x is a multiple of 64, so each wave lies in one row).
template <int Q, class C, bool BGK3, bool HRR, bool VRE, int VAR, int AXS = -1>
__global__ void __launch_bounds__(sizeof(typename C::raw) >= 8 ? 256 : 1024, sizeof(typename C::raw) >= 8 ? 3 : 1) step_kernel(const StepArgs<typename C::real> a) {
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    // (compiled out for the 16-byte tangent storage, which keeps the plain order: the order code
    // there makes the compiler's register allocator fail)
    if (sizeof(typename C::raw) <= 8 && a.xcd_groups) {
        const int gx = gridDim.x, gy = gridDim.y;
        const int b = bx + gx * (by + gy * bz);
        const int rows = gy * gridDim.z;
        if (b < 8 * gx * (rows / 8)) {
            const int slot = b >> 3;
            const int r = (slot / gx) * 8 + (b & 7);
            bx = slot % gx;
            by = r % gy;
            bz = r / gy;
        }
    }
    int y = by * blockDim.y + threadIdx.y;
    if constexpr (AXS >= 0) y = __builtin_amdgcn_readfirstlane(y);
    if (a.part) {
        // launch coordinate along the halo axis -> the part's cell (part_row)
        const int n = a.lay.n(a.part_axis == 1 ? a.lay.ax1 : a.lay.ax2);
        const int m = part_row(a.part_axis == 1 ? y : bz, a.part, n, a.part_edge);
        if (a.part_axis == 1) y = m;
        else bz = m;
    }
    const int n1 = a.lay.n(a.lay.ax1);
    const int cb = (a.heavy_first & 1) && y < n1 ? heavy_order(y, n1) : y + 1;
    const int cc = (a.heavy_first & 2) ? heavy_order(bz, a.lay.n(a.lay.ax2)) : bz + 1;
    step_cell<Q, C, BGK3, HRR, VRE, VAR, AXS>(a, bx * blockDim.x + threadIdx.x + 1, cb, cc);
}

// ------------------------------------------------------------------------------------------
// Boundary kernels (non-periodic ghost layers).  f1 is the buffer the step just wrote, f2 the
// other one; reads of ghost cells in periodic directions wrap onto the interior.
// ------------------------------------------------------------------------------------------
template <int Q, class C>
__device__ __forceinline__ typename C::real ld(const Layout& L, const void* f, int l, int off) {
    return C::template decode<Q>(static_cast<const typename C::raw*>(f)[(size_t)l * (size_t)L.plane + (size_t)off], l);
}
template <int Q, class C>
__device__ __forceinline__ void st(const Layout& L, void* f, int l, int off, typename C::real v) {
    static_cast<typename C::raw*>(f)[(size_t)l * (size_t)L.plane + (size_t)off] = C::template encode<Q>(v, l);
}
// Raw copy (no decode/encode round trip) between two cells of possibly different buffers.
template <int Q, class C>
__device__ __forceinline__ void cp(const Layout& L, void* fdst, int ldst, int odst, const void* fsrc, int lsrc, int osrc) {
    using raw = typename C::raw;
    (void)ldst;
    static_cast<raw*>(fdst)[(size_t)ldst * (size_t)L.plane + (size_t)odst] =
        static_cast<const raw*>(fsrc)[(size_t)lsrc * (size_t)L.plane + (size_t)osrc];
}


// Closed faces normal to axis DIR (src/m_boundary_i_closed_kernel.F90:44-58, _j_:103-114,
// _k_:157-168).  One thread per tangential position handles both faces: it first loads every
// source (the other buffer's ghost cell and the wall-plane cells), then writes the ghosts.  The
// sources are never written by this kernel (plane n >= 2), so loading first keeps the
// reference's result and lets the loads overlap.  Threads run along the tangential axis TA
// (the one stored first), then TB.
template <int Q, class C, int DIR>
__global__ void bc_closed_kernel(Layout L, void* f1, const void* f2, int opt_lo, int opt_hi, int TA, int TB,
                                 const void* snapv, const SnapList sl) {
    using Lt = Lattice<Q>;
    using raw = typename C::raw;
    const int ta = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int tb = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (ta > L.n(TA) || tb > L.n(TB)) return;
    const int n = L.n(DIR);
    raw* __restrict__ p1 = static_cast<raw*>(f1);
    const raw* __restrict__ p2 = static_cast<const raw*>(f2);
    int og[2];
    raw v2[Q], v1[Q];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        int g[3];
        g[DIR] = side ? n + 1 : 0;
        g[TA] = ta;
        g[TB] = tb;
        og[side] = L.cell(g[0], g[1], g[2]);
        const int plane = side ? n : 1;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const int cn = Lt::c(l, DIR);
            if (!(side == 0 ? cn < 0 : cn > 0)) continue;
            v2[l] = p2[(size_t)l * L.plane + og[side]];
            const int si = DIR == 0 ? plane : g[0] - Lt::cx(l);
            const int sj = DIR == 1 ? plane : g[1] - Lt::cy(l);
            const int sk = DIR == 2 ? plane : g[2] - Lt::cz(l);
            v1[l] = p1[(size_t)l * L.plane + L.cell_w(si, sj, sk)];
            if (sl.n > 0) {
                for (int e = 0; e < sl.n; ++e)
                    if (sl.g[e][0] == si && sl.g[e][1] == sj && sl.g[e][2] == sk)
                        v1[l] = static_cast<const raw*>(snapv)[e * Q + l];
            }
        }
    }
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int opt = side ? opt_hi : opt_lo;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const int cn = Lt::c(l, DIR);
            if (!(side == 0 ? cn < 0 : cn > 0)) continue;
            const int m = opt < 0 ? mirror<Q>(l, DIR, -1) : mirror<Q>(l, DIR, +1);
            p1[(size_t)m * L.plane + og[side]] = v2[l];
            p1[(size_t)l * L.plane + og[side]] = v1[l];
        }
    }
}

// Ghost shell of a buffer (the ghost cells of the non-periodic axes, all populations; periodic
// axes' ghosts are never read) copied from src to dst: a buffer about to receive a step's output
// then holds the ghosts it would hold had it received the step before it (the adjoint's recompute
// writes each step into a snapshot slot of its own instead of ping-ponging two buffers).
// Work items in storage coordinates (a, b, c) along (ax0, ax1, ax2): blocks [0, nr1) copy the rows
// b = 0, n1+1 (every c), blocks [nr1, nr1 + nr2) the rows c = 0, n2+1 (interior b), the remaining
// blocks the two end cells a = 0, n0+1 of the interior rows; nr1, nr2, the end cells are zero for a
// periodic axis.
template <int Q, class C>
__global__ void __launch_bounds__(256) ghost_copy_kernel(Layout L, const void* src, void* dst, int nr1, int nr2) {
    using raw = typename C::raw;
    const int n0 = L.n(L.ax0), n1 = L.n(L.ax1), n2 = L.n(L.ax2);
    const raw* __restrict__ s = static_cast<const raw*>(src);
    raw* __restrict__ d = static_cast<raw*>(dst);
    auto copy = [&](int a, int b, int c) {
        int i, j, k;
        L.ijk(a, b, c, i, j, k);
        const int o = L.cell(i, j, k);
#pragma unroll
        for (int l = 0; l < Q; ++l) d[(size_t)l * L.plane + o] = s[(size_t)l * L.plane + o];
    };
    const int r = blockIdx.x;
    if (r < nr1 + nr2) {
        int b, c;
        if (r < nr1) {
            b = (r & 1) ? n1 + 1 : 0;
            c = r >> 1;
        } else {
            const int r2 = r - nr1;
            c = (r2 & 1) ? n2 + 1 : 0;
            b = (r2 >> 1) + 1;
        }
        for (int a = threadIdx.x; a <= n0 + 1; a += blockDim.x) copy(a, b, c);
        return;
    }
    const int t = (r - nr1 - nr2) * blockDim.x + threadIdx.x;
    if (t >= 2 * n1 * n2) return;
    const int rest = t >> 1;
    copy((t & 1) ? n0 + 1 : 0, rest % n1 + 1, rest / n1 + 1);
}

// Copy of the cells listed in sl (periodic axes wrapped) taken before the closed passes.
template <int Q, class C>
__global__ void bc_snapshot_kernel(Layout L, const void* f1, void* snapv, const SnapList sl) {
    using raw = typename C::raw;
    const int t = threadIdx.x;
    if (t >= sl.n * Q) return;
    const int e = t / Q, l = t % Q;
    static_cast<raw*>(snapv)[t] =
        static_cast<const raw*>(f1)[(size_t)l * L.plane + L.cell_w(sl.g[e][0], sl.g[e][1], sl.g[e][2])];
}

// Ghost edges and corners: both passes of src/m_boundary_{j_k,i_j,i_k}_edges.F90 and
// src/m_boundary_closed_corners.F90.  blockIdx.y selects the entry of the list.  Per cell:
//   pass 1: f1(m, g) = f2(l, g) for each outgoing l, m = mirrored l;
//   pass 2: f1(l, g) = f1(l, g - c_l along the closed axes).
// The two passes write disjoint directions and read only the other buffer and interior cells,
// so one thread performs both, and different entries never touch the same cell.
template <int Q, class C>
__global__ void bc_edges_kernel(Layout L, void* f1, const void* f2, const EdgeList e) {
    using Lt = Lattice<Q>;
    const int s = blockIdx.y;
    const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (a > e.len[s]) return;
    int g[3] = {e.g[s][0], e.g[s][1], e.g[s][2]};
    if (e.kind[s] == 0) g[e.t[s]] = a;
    const int og = L.cell(g[0], g[1], g[2]);
    using raw = typename C::raw;
    raw* __restrict__ p1 = static_cast<raw*>(f1);
    const raw* __restrict__ p2 = static_cast<const raw*>(f2);
    raw v2[Q], v1[Q];
    bool outg[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const int c[3] = {Lt::cx(l), Lt::cy(l), Lt::cz(l)};
        bool outgoing = true;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (e.out[s][d] != 0 && c[d] * e.out[s][d] <= 0) outgoing = false;
        outg[l] = outgoing;
        if (!outgoing) continue;
        v2[l] = p2[(size_t)l * L.plane + og];
        int src[3] = {g[0], g[1], g[2]};
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (e.out[s][d] != 0) src[d] = g[d] - c[d];
        v1[l] = p1[(size_t)l * L.plane + L.cell(src[0], src[1], src[2])];
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if (!outg[l]) continue;
        const int m = Lt::find(e.mir[s][0] * Lt::cx(l), e.mir[s][1] * Lt::cy(l), e.mir[s][2] * Lt::cz(l));
        if (m >= 0) p1[(size_t)m * L.plane + og] = v2[l];
        p1[(size_t)l * L.plane + og] = v1[l];
    }
}

// Inflow at i = 0 and zero-gradient outflow at i = nx+1 (src/m_boundary_i_inflow_kernel.F90:61-111)
// for the ghost row (j, k); all loads precede the stores.
template <int Q, class C>
__device__ __forceinline__ void inflow_cell(const Layout& L, void* f1, const typename C::real* shear,
                                            typename C::real utime, typename C::real cosd, typename C::real sind,
                                            typename C::real rho0, int j, int k) {
    using R = typename C::real;
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    raw* __restrict__ p = static_cast<raw*>(f1);
    const R uu = utime * shear[k - 1];        // uvel(k) = uvel_time(it)*uvel_shear(k), taper = 1
    const int o1 = L.cell(1, j, k), o0 = L.cell(0, j, k), on = L.cell(L.nx, j, k), oe = L.cell(L.nx + 1, j, k);
    R fg[Q];
    raw fo[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        fg[l] = C::template decode<Q>(p[(size_t)l * L.plane + o1], l);
        fo[l] = p[(size_t)l * L.plane + on];
    }
    inflow_ghost_values<Q, R>(fg, uu, cosd, sind, rho0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        p[(size_t)l * L.plane + o0] = C::template encode<Q>(Lt::cx(l) <= 0 ? fg[l] : fg[Lt::opposite(l)], l);
        p[(size_t)l * L.plane + oe] = fo[l];
    }
}

// Threads run along j (jfast = 1) or k, whichever is stored first.
template <int Q, class C>
__global__ void bc_inflow_kernel(Layout L, void* f1, const typename C::real* shear, typename C::real utime,
                                 typename C::real cosd, typename C::real sind, typename C::real rho0, int jfast) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = jfast ? x : y;
    const int k = jfast ? y : x;
    if (j > L.ny || k > L.nz) return;
    inflow_cell<Q, C>(L, f1, shear, utime, cosd, sind, rho0, j, k);
}

// ------------------------------------------------------------------------------------------
// Halo planes of j for a domain tiled along j (src/m_mpi_pack_jplane.F90,
// src/m_mpi_unpack_jplane.F90, exchanged in src/m_mpi_halo_exchange_j.F90:55-120).  A halo holds
// only the populations that cross the plane: pack side 0 sends plane j = 1's populations with
// c_y = -1 (to the neighbour below), side 1 plane j = ny's with c_y = +1; unpack side 0 fills
// ghost j = 0 with c_y = +1 populations, side 1 ghost j = ny+1 with c_y = -1.  Buffer order
// [q][k = 0..nz+1][i = 0..nx+1], q counting those populations in lattice order (both sides: side
// 0's halo, then side 1's); values in the storage encoding.
// ------------------------------------------------------------------------------------------
template <int Q>
__host__ __device__ constexpr int halo_count() { return Q == 27 ? 9 : 5; }

// the q-th population with c_y = cy
template <int Q>
__device__ __forceinline__ int halo_pop(int q, int cy) {
    int n = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l)
        if (Lattice<Q>::cy(l) == cy) {
            if (n == q) return l;
            ++n;
        }
    return -1;
}

// One thread per buffer element e = ((h * count + q) * (nz+2) + k) * (nx+2) + i over the nside
// sides side0 .. side0 + nside - 1 (h = side - side0); with both sides the buffer holds side 0's
// halo and then side 1's.
struct HaloIdx {
    int i, k, q, h;
};
template <int Q>
__device__ __forceinline__ bool halo_index(const Layout& L, int nside, HaloIdx& x) {
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned ni = L.nx + 2, nk = L.nz + 2;
    if (e >= (unsigned)nside * halo_count<Q>() * nk * ni) return false;
    x.i = e % ni;
    const unsigned r = e / ni;
    x.k = r % nk;
    const unsigned qh = r / nk;
    x.q = qh % halo_count<Q>();
    x.h = qh / halo_count<Q>();
    return true;
}

template <int Q, class C>
__global__ void __launch_bounds__(256) halo_pack_kernel(Layout L, const void* f, int side0, int nside, void* buf) {
    using raw = typename C::raw;
    HaloIdx x;
    if (!halo_index<Q>(L, nside, x)) return;
    const int side = side0 + x.h;
    const int l = halo_pop<Q>(x.q, side ? 1 : -1);
    const int j = side ? L.ny : 1;
    static_cast<raw*>(buf)[blockIdx.x * blockDim.x + threadIdx.x] =
        static_cast<const raw*>(f)[(size_t)l * L.plane + L.cell(x.i, j, x.k)];
}

// Unpack, and the closed-face pass's copies whose source lies on this halo plane: for ghost G of
// the face of fold_dir and population l pointing out of the fluid, f(l, G) = f(l, G - c_l) on the
// wall plane (src/m_boundary_k_closed_kernel.F90:39-41); when G - c_l is on the halo plane it is
// the neighbour's value of this step, which only the halo brings.  Sources on a ghost of another
// non-periodic axis (the inflow ghost of the halo plane) are left to the step kernel, which reads
// them before this unpack replaces them.
template <int Q, class C>
__global__ void __launch_bounds__(256) halo_unpack_kernel(Layout L, void* f, int side0, int nside, const void* buf,
                                                          int fold_dir) {
    using raw = typename C::raw;
    using Lt = Lattice<Q>;
    HaloIdx hx;
    if (!halo_index<Q>(L, nside, hx)) return;
    const int i = hx.i, k = hx.k;
    const int side = side0 + hx.h;
    const int cy = side ? -1 : 1;
    const int l = halo_pop<Q>(hx.q, cy);
    const int j = side ? L.ny + 1 : 0;
    raw* __restrict__ p = static_cast<raw*>(f);
    const raw v = static_cast<const raw*>(buf)[blockIdx.x * blockDim.x + threadIdx.x];
    p[(size_t)l * L.plane + L.cell(i, j, k)] = v;
    const int D = fold_dir;
    if (D != 0 && D != 2) return;
    const int x[3] = {i, j, k};
    const int n[3] = {L.nx, L.ny, L.nz};
    const int O = 2 - D;   // the face's tangential axis other than j
    if (x[O] < 1 || x[O] > n[O]) return;
    const int cD = Lt::c(l, D);
    const int side_d = x[D] == 1 && cD < 0 ? 0 : x[D] == n[D] && cD > 0 ? 1 : -1;
    if (side_d < 0) return;
    int t[3];
    t[1] = j + cy;
    t[D] = side_d ? n[D] + 1 : 0;
    t[O] = x[O] + Lt::c(l, O);
    const int perO = O == 0 ? L.per_i : L.per_k;
    if (perO) t[O] = Layout::wrap(t[O], n[O], 1);
    else if (t[O] < 1 || t[O] > n[O]) return;
    p[(size_t)l * L.plane + L.cell(t[0], t[1], t[2])] = v;
}

// ------------------------------------------------------------------------------------------
// Ensembles: many domains of one lattice, storage and collision model advanced by one launch of
// a persistent kernel per step.  Every domain's cells are cut into tiles of blockDim.x consecutive
// cells of one (ax0, ax1) plane; the tiles of all members form one list that the blocks of the
// grid walk with stride gridDim.x.  Per-step values are read from a device clock (clock[0] = it of
// the step being computed, clock[1] = steps done since the clock was set, clock[2] = the value of
// clock[1] at the first step of the uploaded table of inflow values) so that a captured graph of
// steps replays without host updates.  Each member's closed face and inflow/outflow planes are
// folded into the step (at most one closed axis per member).
// ------------------------------------------------------------------------------------------
template <typename R>
struct MemberTab {
    StepArgs<R> a;          // src, dst, tu, tv, tw are set per step
    void* buf[2];
    int cur0;               // buffer holding the current state when the clock was set
    int nrturb, slice;      // turbulence slices of ny*nz reals; nrturb = 0: off
    const R* turb[3];       // slice 0 of uu, vv, ww
    int n0, tiles_plane;    // cells along ax0; tiles per (ax0, ax1) plane
};

// Member owning tile t: the last m with start[m] <= t (start[0] = 0, increasing).
__device__ __forceinline__ int owner(const int* __restrict__ start, int n, int t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// vals: (uvel_time(it), cos(udir), sin(udir)) per step of the table and member.
template <int Q, class C, bool BGK3, bool HRR, bool VRE>
__global__ void __launch_bounds__(1024) step_kernel_batch(const MemberTab<typename C::real>* __restrict__ tab,
                                                          const int* __restrict__ start, int nmem, int total,
                                                          const typename C::real* __restrict__ vals,
                                                          const int* __restrict__ clock) {
    using R = typename C::real;
    const int it = clock[0];
    const int s = clock[1];
    const int sv = s - clock[2];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int m = owner(start, nmem, t);
        const MemberTab<R>& M = tab[m];
        StepArgs<R> a = M.a;
        const int p = (M.cur0 + s) & 1;
        a.src = M.buf[p];
        a.dst = M.buf[p ^ 1];
        if (a.inflow_fold) {
            const R* v = vals + ((size_t)sv * nmem + m) * 3;
            a.utime = v[0];
            a.cosd = v[1];
            a.sind = v[2];
        }
        if (M.nrturb > 0) {
            int lit = it % M.nrturb;             // src/m_inflow_turbulence_forcing.F90:45-46
            if (lit == 0) lit = M.nrturb;
            const size_t off = (size_t)lit * M.slice;
            a.tu = M.turb[0] + off;
            a.tv = M.turb[1] + off;
            a.tw = M.turb[2] + off;
        }
        const int local = t - start[m];
        const int cc = local / M.tiles_plane;
        const int e = (local - cc * M.tiles_plane) * (int)blockDim.x + (int)threadIdx.x;
        const int cb = e / M.n0;
        step_cell<Q, C, BGK3, HRR, VRE, 1>(a, e - cb * M.n0 + 1, cb + 1, cc + 1);
    }
}

template <int Q, class C>
__global__ void clock_tick_kernel(int* clock) {
    if (threadIdx.x == 0) {
        clock[0] += 1;
        clock[1] += 1;
    }
}

// ------------------------------------------------------------------------------------------
// State transfer between the reference's AoS layout f(nl,0:nx+1,0:ny+1,0:nz+1) (H = float or
// double, as the caller's REAL kind) and the device buffers.
// ------------------------------------------------------------------------------------------
template <int Q, class C, typename H>
__global__ void aos_to_soa_kernel(Layout L, const H* __restrict__ aos, void* dstA, void* dstB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i > L.nx + 1) return;
    const size_t base = ((size_t)k * (L.ny + 2) * (L.nx + 2) + (size_t)j * (L.nx + 2) + i) * Q;
    const int oc = L.cell(i, j, k);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const typename C::real v = (typename C::real)aos[base + l];
        st<Q, C>(L, dstA, l, oc, v);
        if (dstB) st<Q, C>(L, dstB, l, oc, v);
    }
}

// The reference's f1 (post-streaming state including ghosts) and/or its moments
// (src/m_macrovars_kernel.F90:41-49, all cells including ghosts).
//   interior: stream_interior ? cur(l, x - c) : cur(l, x)
//   ghosts  : ghost_raw ? gsrc(l, x) as stored : gsrc(l, x with periodic axes wrapped)
template <int Q, class C, typename H>
__global__ void state_out_kernel(Layout L, const void* cur, const void* gsrc, int stream_interior, int ghost_raw,
                                 H* __restrict__ aos, H* rho, H* u, H* v, H* w) {
    using R = typename C::real;
    using Lt = Lattice<Q>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i > L.nx + 1) return;
    const bool interior = i >= 1 && i <= L.nx && j >= 1 && j <= L.ny && k >= 1 && k <= L.nz;
    const size_t cidx = (size_t)k * (L.ny + 2) * (L.nx + 2) + (size_t)j * (L.nx + 2) + i;
    R tr = 0, tu = 0, tv = 0, tw = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        R f;
        if (interior) {
            const int s = stream_interior;
            f = ld<Q, C>(L, cur, l, L.cell_w(i - s * Lt::cx(l), j - s * Lt::cy(l), k - s * Lt::cz(l)));
        } else {
            f = ld<Q, C>(L, gsrc, l, ghost_raw ? L.cell(i, j, k) : L.cell_w(i, j, k));
        }
        if (aos) aos[cidx * Q + l] = (H)f;
        tr += f;
        if (Lt::cx(l) > 0) tu += f; else if (Lt::cx(l) < 0) tu -= f;
        if (Lt::cy(l) > 0) tv += f; else if (Lt::cy(l) < 0) tv -= f;
        if (Lt::cz(l) > 0) tw += f; else if (Lt::cz(l) < 0) tw -= f;
    }
    if (rho) {
        const R inv = R(1) / tr;
        rho[cidx] = (H)tr;
        u[cidx] = (H)(tu * inv);
        v[cidx] = (H)(tv * inv);
        w[cidx] = (H)(tw * inv);
    }
}

// Equilibrium initialisation (src/m_fequil3_kernel.F90 then src/m_fillghosts.F90): every cell,
// ghosts included, gets feq of its clamped interior cell.  The moments (H, reference cell
// layout) are also kept in the device cell layout for the first collision.
template <int Q, class C, bool BGK3, typename H>
__global__ void feq_init_kernel(Layout L, const H* rho, const H* u, const H* v, const H* w, void* dst,
                                typename C::real* mr, typename C::real* mu, typename C::real* mv,
                                typename C::real* mw) {
    using R = typename C::real;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i > L.nx + 1) return;
    const int i1 = i < 1 ? 1 : (i > L.nx ? L.nx : i);
    const int j1 = j < 1 ? 1 : (j > L.ny ? L.ny : j);
    const int k1 = k < 1 ? 1 : (k > L.nz ? L.nz : k);
    const size_t c = (size_t)k1 * (L.ny + 2) * (L.nx + 2) + (size_t)j1 * (L.nx + 2) + i1;
    const R r = (R)rho[c], ux = (R)u[c], uy = (R)v[c], uz = (R)w[c];
    const R usq = ux * ux + uy * uy + uz * uz;
    const int oc = L.cell(i, j, k);
#pragma unroll
    for (int l = 0; l < Q; ++l)
        st<Q, C>(L, dst, l, oc, feq_dir<Q, BGK3, R>(l, r, ux, uy, uz, usq));
    if (i == i1 && j == j1 && k == k1) {
        const int ox = L.cidx(i, j, k);
        mr[ox] = r; mu[ox] = ux; mv[ox] = uy; mw[ox] = uz;
    }
}

// Global mass and momentum of the state (interior cells) in FP64: per-block partial sums,
// finished on the host.  partial[4*block + {0,1,2,3}] = (mass, mx, my, mz).
// Adds the moments of the current state (streamed when stream_interior) to the averaging sums.
template <int Q, class C>
__global__ void average_kernel(Layout L, const void* cur, int stream_interior, typename C::real* avg) {
    using R = typename C::real;
    using Lt = Lattice<Q>;
    const int ca = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int cb = blockIdx.y + 1;
    const int cc = blockIdx.z + 1;
    if (ca > L.n(L.ax0)) return;
    int i, j, k;
    L.ijk(ca, cb, cc, i, j, k);
    const int s = stream_interior;
    R dr = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const R d = C::template dep<Q>(static_cast<const typename C::raw*>(cur)[(size_t)l * L.plane +
                                           L.cell_w(i - s * Lt::cx(l), j - s * Lt::cy(l), k - s * Lt::cz(l))], l) /
                    C::kDepScale;
        dr += d;
        jx += R(Lt::cx(l)) * d;
        jy += R(Lt::cy(l)) * d;
        jz += R(Lt::cz(l)) * d;
    }
    const R inv = R(1) / (R(1) + dr);
    const R ux = jx * inv, uy = jy * inv, uz = jz * inv;
    R* p = avg + L.cidx(i, j, k);
    const size_t P = (size_t)L.cplane;
    p[0] += ux;
    p[P] += uy;
    p[2 * P] += uz;
    p[3 * P] += ux * ux;
    p[4 * P] += uy * uy;
    p[5 * P] += uz * uz;
}

// Global mass and momentum of the state (interior cells) in FP64, with the smallest cell density
// and the number of non-finite populations (the instability test of src/m_diag.F90:169-173):
// per-block partial results partial[kTotalsSlots*block + {0..5}] = (mass, mx, my, mz, min rho,
// non-finite count), finished on the host.  Every block writes all of its slots.
template <int Q, class C>
__global__ void __launch_bounds__(256) totals_kernel(Layout L, const void* cur, int stream_interior, double* partial) {
    using Lt = Lattice<Q>;
    __shared__ double sh[kTotalsSlots][256];
    const long long ncell = (long long)L.nx * L.ny * L.nz;
    double m = 0, mx = 0, my = 0, mz = 0, rmin = INFINITY, bad = 0;
    const int s = stream_interior;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % L.nx) + 1;
        const int j = (int)((c / L.nx) % L.ny) + 1;
        const int k = (int)(c / ((long long)L.nx * L.ny)) + 1;
        double r = 0;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const double f = (double)ld<Q, C>(L, cur, l, L.cell_w(i - s * Lt::cx(l), j - s * Lt::cy(l), k - s * Lt::cz(l)));
            if (!isfinite(f)) bad += 1;
            r += f;
            mx += Lt::cx(l) * f;
            my += Lt::cy(l) * f;
            mz += Lt::cz(l) * f;
        }
        m += r;
        rmin = r < rmin ? r : rmin;
    }
    const int t = threadIdx.x;
    sh[0][t] = m; sh[1][t] = mx; sh[2][t] = my; sh[3][t] = mz; sh[4][t] = rmin; sh[5][t] = bad;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < 4; ++q) sh[q][t] += sh[q][t + w];
            sh[4][t] = sh[4][t + w] < sh[4][t] ? sh[4][t + w] : sh[4][t];
            sh[5][t] += sh[5][t + w];
        }
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < kTotalsSlots; ++q) partial[kTotalsSlots * blockIdx.x + q] = sh[q][0];
}

}  // namespace lbm
