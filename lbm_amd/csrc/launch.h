// Host launch wrappers for one (lattice, codec) pair; included once per codec translation unit.
#pragma once

#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstring>

#include "impl.h"
#include "kernels.h"
#include "adjoint.h"
#include "turbines.h"

namespace lbm {

template <typename R>
struct Limits;
template <>
struct Limits<float> {
    static float tiny() { return FLT_MIN; }
    static float eps() { return sqrtf(FLT_MIN); }   // sqrt(tiny(1.0)), src/m_postcoll.F90:47
};
template <>
struct Limits<double> {
    static double tiny() { return DBL_MIN; }
    static double eps() { return std::sqrt(DBL_MIN); }
};
template <>
struct Limits<Dual> {
    static Dual tiny() { return Dual(DBL_MIN); }
    static Dual eps() { return Dual(std::sqrt(DBL_MIN)); }
};

inline dim3 grid_cells(const Layout& L, int bx) { return dim3((L.nx + 2 + bx - 1) / bx, L.ny + 2, L.nz + 2); }

template <int Q, class C>
struct Launch {
    using R = typename C::real;
    // non-temporal loads and stores take scalar types only (not the tangent pair)
    static constexpr bool kNT = sizeof(typename C::raw) <= 8;

    template <bool B3, bool H, bool V>
    static hipError_t step_t(const Layout& L, const StepArgs<R>& a, dim3 grid, dim3 block, int var, int lds, hipStream_t s) {
        const bool nt = var & 1;
        // one 32-bit offset range covers every plane of the buffer (var bit 1 forces the large form)
        const bool small = (double)L.total * sizeof(typename C::raw) < 4294967296.0 && !(var & 2);
        if constexpr (kNT) {
            // the fast form is compiled per contiguous axis (the others in make_layout's order); its
            // wave-uniform rows need whole waves along that axis
            const int ax = L.ax0;
            const bool axes_std = L.ax1 == (ax == 0 ? 1 : 0) && L.ax2 == (ax == 2 ? 1 : 2) && block.x % 64 == 0;
            // the slab form (bit 4, one base per slab of the slowest axis): row-blocked 4-byte buffers of
            // every size (measured on MI355X, D3Q27 384x192x128: 0.717 of the HBM peak against 0.714 for
            // the one-base form; 768x256x192, 5.2 GB: 0.74 against 0.66 for the plane arrangement), and
            // 2-byte buffers of 4 GiB and more (below, the one-base form measured 1.5 % faster)
            const bool slab = L.rowblock && axes_std && sizeof(typename C::raw) <= 4 && (sizeof(typename C::raw) == 4 || !small);
            bool done = false;
            // variant bit 2: the one-base form with the axes mapped at run time (no per-axis compilation)
            if (var & 4) {
                if (nt && small) {
                    hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 9>), grid, block, lds, s, a);
                    done = true;
                }
            } else if constexpr (sizeof(typename C::raw) != 4) {
                done = true;
                if (nt && small && axes_std && ax == 0) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 9, 0>), grid, block, lds, s, a);
                else if (nt && small && axes_std && ax == 1) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 9, 1>), grid, block, lds, s, a);
                else if (nt && small && axes_std) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 9, 2>), grid, block, lds, s, a);
                else done = false;
            }
            if (done) {
            } else if (nt && slab && ax == 0) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 17, 0>), grid, block, lds, s, a);
            else if (nt && slab && ax == 1) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 17, 1>), grid, block, lds, s, a);
            else if (nt && slab) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 17, 2>), grid, block, lds, s, a);
            else if (nt) hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 1>), grid, block, lds, s, a);
            else hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 0>), grid, block, lds, s, a);
        } else {
            hipLaunchKernelGGL((step_kernel<Q, C, B3, H, V, 0>), grid, block, lds, s, a);
        }
        return hipGetLastError();
    }

    static StepArgs<R> args_of(const Layout& L, const StepHost& h) {
        StepArgs<R> a;
        a.lay = L;
        a.src = h.src;
        a.dst = h.dst;
        a.stream = h.stream;
        a.fac0 = (R)h.fac0;
        a.kinevisc = (R)h.kinevisc;
        a.vconst = (R)h.vconst;
        a.eps = Limits<R>::eps();
        a.tiny = Limits<R>::tiny();
        a.ablvisc = static_cast<const R*>(h.ablvisc);
        a.ext_rho = static_cast<const R*>(h.ext[0]);
        a.ext_u = static_cast<const R*>(h.ext[1]);
        a.ext_v = static_cast<const R*>(h.ext[2]);
        a.ext_w = static_cast<const R*>(h.ext[3]);
        a.solid = h.solid;
        a.solid_wpr = h.solid_wpr;
        a.ip = h.ip;
        a.tu = static_cast<const R*>(h.turb[0]);
        a.tv = static_cast<const R*>(h.turb[1]);
        a.tw = static_cast<const R*>(h.turb[2]);
        a.ampl = (R)h.ampl;
        a.avg = static_cast<R*>(h.avg);
        a.Fx = static_cast<const R*>(h.F[0]);
        a.Fy = static_cast<const R*>(h.F[1]);
        a.Fz = static_cast<const R*>(h.F[2]);
        for (int d = 0; d < 3; ++d) {
            a.flo[d] = h.flo[d];
            a.fhi[d] = h.fhi[d];
        }
        a.theta = static_cast<const R*>(h.theta);
        a.btheta0 = (R)h.theta0;
        a.bscale = (R)h.bscale;
        a.tau_out = static_cast<R*>(h.tau_out);
        a.fold_dir = h.fold_dir;
        a.fold_opt_lo = h.fold_opt_lo;
        a.fold_opt_hi = h.fold_opt_hi;
        a.inflow_fold = h.inflow_fold;
        a.shear = static_cast<const R*>(h.shear);
        a.utime = (R)h.utime;
        a.cosd = (R)h.cosd;
        a.sind = (R)h.sind;
        a.rho0 = (R)h.rho0;
        a.heavy_first = h.heavy_first;
        a.halo_dir = h.halo_dir;
        a.part = h.part;
        a.part_axis = h.part_axis;
        a.part_edge = h.part_edge;
        return a;
    }

    static hipError_t step(const Layout& L, const StepHost& h, hipStream_t s) {
        StepArgs<R> a = args_of(L, h);
        const dim3 block(h.block_x, h.block_y, 1);
        // a split step covers 2e (part 1) or n - 2e (part 2) along the halo axis (n > 2e)
        int n1 = L.n(L.ax1), n2 = L.n(L.ax2);
        if (h.part) {
            int& n = h.part_axis == 1 ? n1 : n2;
            n = h.part == 1 ? 2 * h.part_edge : n - 2 * h.part_edge;
            // the kernel's part mapping replaces the dispatch order along the halo axis
            a.heavy_first &= h.part_axis == 1 ? ~1 : ~2;
        }
        const dim3 grid((L.n(L.ax0) + h.block_x - 1) / h.block_x, (n1 + h.block_y - 1) / h.block_y, n2);
        static const bool xcd = !(getenv("LBM_XCD_GROUPS") && atoi(getenv("LBM_XCD_GROUPS")) == 0);
        a.xcd_groups = xcd && grid.x > 1;
        const int key = (h.bgk3 ? 4 : 0) | (h.hrr ? 2 : 0) | (h.vre ? 1 : 0);
        switch (key) {
            case 0: return step_t<false, false, false>(L, a, grid, block, h.variant, h.lds_bytes, s);
            case 1: return step_t<false, false, true>(L, a, grid, block, h.variant, h.lds_bytes, s);
            case 2: return step_t<false, true, false>(L, a, grid, block, h.variant, h.lds_bytes, s);
            case 3: return step_t<false, true, true>(L, a, grid, block, h.variant, h.lds_bytes, s);
            case 4: return step_t<true, false, false>(L, a, grid, block, h.variant, h.lds_bytes, s);
            case 5: return step_t<true, false, true>(L, a, grid, block, h.variant, h.lds_bytes, s);
            case 6: return step_t<true, true, false>(L, a, grid, block, h.variant, h.lds_bytes, s);
            default: return step_t<true, true, true>(L, a, grid, block, h.variant, h.lds_bytes, s);
        }
    }

    static hipError_t snapshot(const Layout& L, const void* f1, void* snap, const SnapList& sl, hipStream_t s) {
        if (sl.n == 0) return hipSuccess;
        hipLaunchKernelGGL((bc_snapshot_kernel<Q, C>), dim3(1), dim3(256), 0, s, L, f1, snap, sl);
        return hipGetLastError();
    }

    // side 0 or 1, or 2: both (side 0's halo first in buf)
    static hipError_t ghost_copy(const Layout& L, const void* src, void* dst, hipStream_t s) {
        const int per[3] = {L.per_i, L.per_j, L.per_k};
        const int n1 = L.n(L.ax1), n2 = L.n(L.ax2);
        const int nr1 = per[L.ax1] ? 0 : 2 * (n2 + 2);
        const int nr2 = per[L.ax2] ? 0 : 2 * n1;
        const long long ne = per[L.ax0] ? 0 : 2LL * n1 * n2;
        const long long blocks = nr1 + nr2 + (ne + 255) / 256;
        if (blocks == 0) return hipSuccess;
        hipLaunchKernelGGL((ghost_copy_kernel<Q, C>), dim3((unsigned)blocks), dim3(256), 0, s, L, src, dst, nr1, nr2);
        return hipGetLastError();
    }

    static hipError_t turb_step(const Layout& L, const TurbHost& h, const void* src, int stream,
                                const void* const ext[4], hipStream_t s) {
        if constexpr (sizeof(R) > 8) {
            return hipErrorNotSupported;   // the tangent storage has no turbines
        } else {
            TurbArgs<R> a;
            a.lay = L;
            a.src = src;
            a.stream = stream;
            for (int q = 0; q < 4; ++q) a.ext[q] = static_cast<const R*>(ext[q]);
            a.nturb = h.nturb;
            a.nblades = h.nblades;
            a.nchords = h.nchords;
            a.nang_max = h.nang_max;
            a.krad = h.krad;
            a.sigma2 = (R)h.sigma2;
            a.rot = static_cast<const R*>(h.rot);
            a.theta = static_cast<R*>(h.theta);
            a.blade = static_cast<const R*>(h.blade);
            a.nfoil = h.nfoil;
            a.nang = h.nang;
            a.foil = static_cast<const R*>(h.foil);
            a.pts = static_cast<R*>(h.pts);
            a.tbox = h.tbox;
            for (int d = 0; d < 3; ++d) {
                a.lo[d] = h.lo[d];
                a.hi[d] = h.hi[d];
            }
            for (int q = 0; q < 3; ++q) a.F[q] = static_cast<R*>(h.F[q]);
            hipLaunchKernelGGL((turbine_point_kernel<Q, C>), dim3(h.npts), dim3(256), 0, s, a);
            const long long cells = (long long)(h.hi[0] - h.lo[0] + 1) * (h.hi[1] - h.lo[1] + 1) * (h.hi[2] - h.lo[2] + 1);
            // at least one block: block 0 also stores the advanced azimuths
            const long long blocks = cells > 0 ? (cells + 255) / 256 : 1;
            hipLaunchKernelGGL((turbine_deposit_kernel<R>), dim3((unsigned)blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        }
    }

    static dim3 halo_grid(const Layout& L, int nside) {
        const long long n = (long long)nside * halo_count<Q>() * (L.nz + 2) * (L.nx + 2);
        return dim3((unsigned)((n + 255) / 256));
    }
    static hipError_t halo_pack(const Layout& L, const void* f, int side, void* buf, hipStream_t s) {
        const int nside = side == 2 ? 2 : 1;
        hipLaunchKernelGGL((halo_pack_kernel<Q, C>), halo_grid(L, nside), dim3(256), 0, s, L, f, side == 2 ? 0 : side,
                           nside, buf);
        return hipGetLastError();
    }

    static hipError_t halo_unpack(const Layout& L, void* f, int side, const void* buf, int fold_dir, hipStream_t s) {
        const int nside = side == 2 ? 2 : 1;
        hipLaunchKernelGGL((halo_unpack_kernel<Q, C>), halo_grid(L, nside), dim3(256), 0, s, L, f,
                           side == 2 ? 0 : side, nside, buf, fold_dir);
        return hipGetLastError();
    }

    static hipError_t closed(const Layout& L, void* f1, const void* f2, int dir, int opt_lo, int opt_hi,
                             const void* snap, const SnapList& sl, hipStream_t s) {
        // tangential axes in storage order: threads run along the first
        const int order[3] = {L.ax0, L.ax1, L.ax2};
        int t[2], q = 0;
        for (int a = 0; a < 3; ++a)
            if (order[a] != dir) t[q++] = order[a];
        const dim3 block(64, 4, 1);
        const dim3 grid((L.n(t[0]) + 63) / 64, (L.n(t[1]) + 3) / 4, 1);
        if (dir == 0) hipLaunchKernelGGL((bc_closed_kernel<Q, C, 0>), grid, block, 0, s, L, f1, f2, opt_lo, opt_hi, t[0], t[1], snap, sl);
        else if (dir == 1) hipLaunchKernelGGL((bc_closed_kernel<Q, C, 1>), grid, block, 0, s, L, f1, f2, opt_lo, opt_hi, t[0], t[1], snap, sl);
        else hipLaunchKernelGGL((bc_closed_kernel<Q, C, 2>), grid, block, 0, s, L, f1, f2, opt_lo, opt_hi, t[0], t[1], snap, sl);
        return hipGetLastError();
    }

    static hipError_t edges(const Layout& L, void* f1, const void* f2, const EdgeList& e, hipStream_t s) {
        if (e.n == 0) return hipSuccess;
        int mx = 1;
        for (int q = 0; q < e.n; ++q) mx = e.len[q] > mx ? e.len[q] : mx;
        const dim3 block(64, 1, 1);
        const dim3 grid((mx + 63) / 64, e.n, 1);
        hipLaunchKernelGGL((bc_edges_kernel<Q, C>), grid, block, 0, s, L, f1, f2, e);
        return hipGetLastError();
    }

    static hipError_t inflow(const Layout& L, void* f1, const void* shear, double utime, double cosd, double sind,
                             double rho0, hipStream_t s) {
        // threads along j or k, whichever is stored first
        const int jfast = (L.ax0 == 1) || (L.ax0 == 0 && L.ax1 == 1);
        const int nx_ = jfast ? L.ny : L.nz, ny_ = jfast ? L.nz : L.ny;
        const dim3 block(64, 4, 1);
        const dim3 grid((nx_ + 63) / 64, (ny_ + 3) / 4, 1);
        hipLaunchKernelGGL((bc_inflow_kernel<Q, C>), grid, block, 0, s, L, f1, static_cast<const R*>(shear), (R)utime,
                           (R)cosd, (R)sind, (R)rho0, jfast);
        return hipGetLastError();
    }

    static hipError_t aos_in(const Layout& L, const void* aos, int realkind, void* dA, void* dB, hipStream_t s) {
        const dim3 block(128, 1, 1);
        const dim3 grid = grid_cells(L, 128);
        if (realkind == 8)
            hipLaunchKernelGGL((aos_to_soa_kernel<Q, C, double>), grid, block, 0, s, L, static_cast<const double*>(aos), dA, dB);
        else
            hipLaunchKernelGGL((aos_to_soa_kernel<Q, C, float>), grid, block, 0, s, L, static_cast<const float*>(aos), dA, dB);
        return hipGetLastError();
    }

    static hipError_t state_out(const Layout& L, const void* cur, const void* gsrc, int si, int raw, void* aos,
                                void* rho, void* u, void* v, void* w, int realkind, hipStream_t s) {
        const dim3 block(128, 1, 1);
        const dim3 grid = grid_cells(L, 128);
        if (realkind == 8)
            hipLaunchKernelGGL((state_out_kernel<Q, C, double>), grid, block, 0, s, L, cur, gsrc, si, raw,
                               static_cast<double*>(aos), static_cast<double*>(rho), static_cast<double*>(u),
                               static_cast<double*>(v), static_cast<double*>(w));
        else
            hipLaunchKernelGGL((state_out_kernel<Q, C, float>), grid, block, 0, s, L, cur, gsrc, si, raw,
                               static_cast<float*>(aos), static_cast<float*>(rho), static_cast<float*>(u),
                               static_cast<float*>(v), static_cast<float*>(w));
        return hipGetLastError();
    }

    template <typename H>
    static hipError_t feq_init_h(const Layout& L, int bgk3, const void* r, const void* u, const void* v, const void* w,
                                 void* dst, void* mr, void* mu, void* mv, void* mw, hipStream_t s) {
        const dim3 block(128, 1, 1);
        const dim3 grid = grid_cells(L, 128);
        const H *hr = static_cast<const H*>(r), *hu = static_cast<const H*>(u), *hv = static_cast<const H*>(v),
                *hw = static_cast<const H*>(w);
        R *pr = static_cast<R*>(mr), *pu = static_cast<R*>(mu), *pv = static_cast<R*>(mv), *pw = static_cast<R*>(mw);
        if (bgk3)
            hipLaunchKernelGGL((feq_init_kernel<Q, C, true, H>), grid, block, 0, s, L, hr, hu, hv, hw, dst, pr, pu, pv, pw);
        else
            hipLaunchKernelGGL((feq_init_kernel<Q, C, false, H>), grid, block, 0, s, L, hr, hu, hv, hw, dst, pr, pu, pv, pw);
        return hipGetLastError();
    }

    static hipError_t feq_init(const Layout& L, int bgk3, const void* r, const void* u, const void* v, const void* w,
                               int realkind, void* dst, void* mr, void* mu, void* mv, void* mw, hipStream_t s) {
        if (realkind == 8) return feq_init_h<double>(L, bgk3, r, u, v, w, dst, mr, mu, mv, mw, s);
        return feq_init_h<float>(L, bgk3, r, u, v, w, dst, mr, mu, mv, mw, s);
    }

    static hipError_t totals(const Layout& L, const void* cur, int si, double* partial, int nblocks, hipStream_t s) {
        hipLaunchKernelGGL((totals_kernel<Q, C>), dim3(nblocks), dim3(256), 0, s, L, cur, si, partial);
        return hipGetLastError();
    }

    static hipError_t average(const Layout& L, const void* cur, int si, void* avg, hipStream_t s) {
        const dim3 grid((L.n(L.ax0) + 127) / 128, L.n(L.ax1), L.n(L.ax2));
        hipLaunchKernelGGL((average_kernel<Q, C>), grid, dim3(128), 0, s, L, cur, si, static_cast<R*>(avg));
        return hipGetLastError();
    }

    // ---- ensembles
    static void ens_pack(const MemberHost& m, int block, void* out, int* tiles) {
        const Layout& L = m.lay;
        MemberTab<R> t{};
        t.a = args_of(L, m.h);
        t.a.stream = 1;
        t.a.tu = t.a.tv = t.a.tw = nullptr;
        t.buf[0] = m.buf[0];
        t.buf[1] = m.buf[1];
        t.cur0 = m.cur0;
        t.nrturb = m.turb[0] ? m.nrturb : 0;
        t.slice = L.ny * L.nz;
        for (int q = 0; q < 3; ++q) t.turb[q] = static_cast<const R*>(m.turb[q]);
        t.n0 = L.n(L.ax0);
        const long long plane = (long long)L.n(L.ax0) * L.n(L.ax1);
        t.tiles_plane = (int)((plane + block - 1) / block);
        t.a.inflow_fold = m.inflow;
        t.a.shear = static_cast<const R*>(m.shear);
        t.a.rho0 = (R)m.rho0;
        std::memcpy(out, &t, sizeof(t));
        *tiles = t.tiles_plane * L.n(L.ax2);
    }

    template <bool B3, bool H, bool V>
    static void ens_step_t(const EnsLaunch& e, hipStream_t s) {
        if constexpr (kNT)
            hipLaunchKernelGGL((step_kernel_batch<Q, C, B3, H, V>), dim3(e.grid), dim3(e.block), 0, s,
                               static_cast<const MemberTab<R>*>(e.tab), e.start, e.nmem, e.total,
                               static_cast<const R*>(e.vals), e.clock);
    }

    static hipError_t ens_step(const EnsLaunch& e, hipStream_t s) {
        const int key = (e.bgk3 ? 4 : 0) | (e.hrr ? 2 : 0) | (e.vre ? 1 : 0);
        switch (key) {
            case 0: ens_step_t<false, false, false>(e, s); break;
            case 1: ens_step_t<false, false, true>(e, s); break;
            case 2: ens_step_t<false, true, false>(e, s); break;
            case 3: ens_step_t<false, true, true>(e, s); break;
            case 4: ens_step_t<true, false, false>(e, s); break;
            case 5: ens_step_t<true, false, true>(e, s); break;
            case 6: ens_step_t<true, true, false>(e, s); break;
            default: ens_step_t<true, true, true>(e, s); break;
        }
        hipLaunchKernelGGL((clock_tick_kernel<Q, C>), dim3(1), dim3(64), 0, s, e.clock);
        return hipGetLastError();
    }

    // ---- adjoint
    static constexpr bool kAdj = sizeof(typename C::raw) == 4 || sizeof(typename C::raw) == 8;

    template <bool B3, bool H, bool V>
    static void adj_step_t(const Layout& L, const StepArgs<R>& a, const void* ly, void* lx, const AdjBox& box, R kappa,
                           int coll, double* gF, hipStream_t s) {
        if constexpr (kAdj) {
            const dim3 block(256, 1, 1);
            const dim3 grid((L.n(L.ax0) * L.n(L.ax1) + 255) / 256, L.n(L.ax2), 1);
            const auto* lyp = static_cast<const typename C::raw*>(ly);
            auto* lxp = static_cast<typename C::raw*>(lx);
            // 32-bit byte offsets from one base when a buffer is below 4 GiB
            if ((double)L.total * sizeof(typename C::raw) < 4294967296.0)
                hipLaunchKernelGGL((adjoint_step_kernel<Q, C, B3, H, V, true>), grid, block, 0, s, a, lyp, lxp, box, kappa,
                                   coll, gF);
            else
                hipLaunchKernelGGL((adjoint_step_kernel<Q, C, B3, H, V, false>), grid, block, 0, s, a, lyp, lxp, box, kappa,
                                   coll, gF);
        }
    }

    static hipError_t adj_step(const Layout& L, const StepHost& h, const void* ly, void* lx, const AdjBox& box,
                               double kappa, int coll, double* gF, hipStream_t s) {
        const StepArgs<R> a = args_of(L, h);
        const int key = (h.bgk3 ? 4 : 0) | (h.hrr ? 2 : 0) | (h.vre ? 1 : 0);
        const R k = (R)kappa;
        switch (key) {
            case 0: adj_step_t<false, false, false>(L, a, ly, lx, box, k, coll, gF, s); break;
            case 1: adj_step_t<false, false, true>(L, a, ly, lx, box, k, coll, gF, s); break;
            case 2: adj_step_t<false, true, false>(L, a, ly, lx, box, k, coll, gF, s); break;
            case 3: adj_step_t<false, true, true>(L, a, ly, lx, box, k, coll, gF, s); break;
            case 4: adj_step_t<true, false, false>(L, a, ly, lx, box, k, coll, gF, s); break;
            case 5: adj_step_t<true, false, true>(L, a, ly, lx, box, k, coll, gF, s); break;
            case 6: adj_step_t<true, true, false>(L, a, ly, lx, box, k, coll, gF, s); break;
            default: adj_step_t<true, true, true>(L, a, ly, lx, box, k, coll, gF, s); break;
        }
        return hipGetLastError();
    }

    static hipError_t adj_inflow(const Layout& L, void* lam, double utime, double cosd, double sind, double rho0,
                                 double* gS, hipStream_t s) {
        if constexpr (kAdj) {
            const dim3 grid((L.ny + 63) / 64, L.nz, 1);
            hipLaunchKernelGGL((adjoint_inflow_kernel<Q, C>), grid, dim3(64), 0, s, L, static_cast<typename C::raw*>(lam),
                               (R)utime, (R)cosd, (R)sind, (R)rho0, gS);
        }
        return hipGetLastError();
    }

    static hipError_t adj_face(const Layout& L, void* ly, void* lx, int dir, int opt_lo, int opt_hi, hipStream_t s) {
        if constexpr (kAdj) {
            // tangential axes in storage order: threads run along the first
            const int order[3] = {L.ax0, L.ax1, L.ax2};
            int tt[2], q = 0;
            for (int a = 0; a < 3; ++a)
                if (order[a] != dir) tt[q++] = order[a];
            const int t0 = tt[0], t1 = tt[1];
            const dim3 grid((L.n(t0) + 63) / 64, L.n(t1), 2);
            hipLaunchKernelGGL((adjoint_face_kernel<Q, C>), grid, dim3(64), 0, s, L, static_cast<typename C::raw*>(ly),
                               static_cast<typename C::raw*>(lx), dir, t0, t1, opt_lo, opt_hi);
        }
        return hipGetLastError();
    }

    static hipError_t objective(const Layout& L, const void* f, const AdjBox& box, double* acc, hipStream_t s) {
        hipLaunchKernelGGL((objective_kernel<Q, C>), dim3(objective_blocks(box)), dim3(256), 0, s, L, f, box, acc);
        return hipGetLastError();
    }

    static const Ops* ops() {
        static const Ops o = {&step, &closed, &snapshot, &edges, &inflow, &aos_in, &state_out, &feq_init, &totals, &average,
                              sizeof(MemberTab<R>), &ens_pack, &ens_step,
                              kAdj ? &adj_step : nullptr, kAdj ? &adj_inflow : nullptr, kAdj ? &adj_face : nullptr,
                              &objective, &halo_pack, &halo_unpack, &ghost_copy, &turb_step, (int)sizeof(R),
                              (int)sizeof(typename C::raw)};
        return &o;
    }
};

}  // namespace lbm
