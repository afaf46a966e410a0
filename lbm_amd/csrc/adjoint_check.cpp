// CPU check of the per-cell adjoint (collide_adjoint, adjoint.h) against central differences of
// the collision the step kernel runs (collide_cell, kernels.h), both executed on the host.
//
//   adjoint_check            prints one JSON line per case and exits 0
//
// For random pre-collision populations g (departures from equilibrium of random rho, u plus a
// random non-equilibrium part), body force F, adjoint weights ly and objective weight kappa, the
// functional  J(g, F) = sum_l ly_l y_l(g, F) + kappa * 0.5 rho|u|^2(g)  is differentiated in a
// random direction (dg, dF): the adjoint gives  lg.dg + lF.dF,  the double-precision central
// difference (J(+h) - J(-h)) / 2h with h = 1e-6 gives the reference value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <type_traits>
#include <vector>

#include "adjoint.h"

using namespace lbm;

namespace {

template <typename R>
struct Cell {
    Layout lay;
    R abl[4] = {R(0), R(0), R(0), R(0)};
    std::vector<R> F[3];       // cell planes of the 3x3x3 layout
    uint32_t solid[4] = {0, 0, 0, 0};
    StepArgs<R> a{};
    int oc = 0;

    Cell(bool force, bool solid_cell, double nu, double vconst) {
        lay = make_layout(3, 3, 3, sizeof(R), 0, 0, 0);
        oc = lay.cidx(1, 1, 1);
        for (int c = 0; c < 3; ++c) F[c].assign((size_t)lay.cplane, R(0));
        a.lay = lay;
        a.stream = 1;
        a.kinevisc = (R)nu;
        a.fac0 = R(1) - R(1) / (R(3) * (R)nu + R(0.5));
        a.vconst = (R)vconst;
        a.tiny = sizeof(R) == 4 ? (R)1.17549435e-38f : (R)2.2250738585072014e-308;
        a.eps = std::sqrt(a.tiny);
        a.ablvisc = abl;
        a.Fx = force ? F[0].data() : nullptr;
        a.Fy = force ? F[1].data() : nullptr;
        a.Fz = force ? F[2].data() : nullptr;
        for (int d = 0; d < 3; ++d) {   // the force box: every interior cell
            a.flo[d] = 1;
            a.fhi[d] = 3;
        }
        solid[0] = solid_cell ? 1u : 0u;
        a.solid = solid;
        a.solid_wpr = 1;
        a.ip = -5;
    }
};

template <int Q, bool B3, bool H, bool V, typename R>
double functional(Cell<R>& c, const double* g, const double* F, const double* ly, double kappa) {
    using C = typename std::conditional<sizeof(R) == 4, CodecF32, CodecF64>::type;
    typename C::raw rv[Q];
    for (int l = 0; l < Q; ++l) rv[l] = (typename C::raw)g[l];
    for (int q = 0; q < 3; ++q) c.F[q][c.oc] = (R)F[q];
    R dev[Q];
    R tau = 0;
    collide_cell<Q, C, B3, H, V>(c.a, rv, 1, 1, 1, 1, 1, 1, c.oc, dev, tau);
    double J = 0.0, rho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    for (int l = 0; l < Q; ++l) {
        J += ly[l] * ((double)weight<Q, R>(l) + (double)dev[l]);
        rho += g[l];
        jx += Lattice<Q>::cx(l) * g[l];
        jy += Lattice<Q>::cy(l) * g[l];
        jz += Lattice<Q>::cz(l) * g[l];
    }
    return J + kappa * 0.5 * (jx * jx + jy * jy + jz * jz) / rho;
}

template <int Q, bool B3, bool H, bool V>
void run_case(std::mt19937_64& rng, bool force, bool solid, double kappa, int trial) {
    using Lt = Lattice<Q>;
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double nu = 1e-3, vconst = 2.5 * 0.16733 * 0.16733;
    // state: feq(rho, u) + non-equilibrium
    const double rho = 1.0 + 0.01 * U(rng), ux = 0.05 + 0.02 * U(rng), uy = 0.02 * U(rng), uz = 0.02 * U(rng);
    double g[Q], dg[Q], ly[Q], F[3], dF[3];
    for (int l = 0; l < Q; ++l) {
        const double cu = Lt::cx(l) * ux + Lt::cy(l) * uy + Lt::cz(l) * uz, usq = ux * ux + uy * uy + uz * uz;
        const double feq = Lt::w(l) * rho * (1.0 - 1.5 * usq + 3.0 * cu + 4.5 * cu * cu);
        g[l] = feq + 2e-3 * Lt::w(l) * U(rng);
        dg[l] = Lt::w(l) * U(rng);
        ly[l] = U(rng);
    }
    for (int q = 0; q < 3; ++q) {
        F[q] = force ? 1e-4 * U(rng) : 0.0;
        dF[q] = force ? U(rng) : 0.0;
    }
    const double h = 1e-6;
    double gp[Q], gm[Q], Fp[3], Fm[3];
    for (int l = 0; l < Q; ++l) {
        gp[l] = g[l] + h * dg[l];
        gm[l] = g[l] - h * dg[l];
    }
    for (int q = 0; q < 3; ++q) {
        Fp[q] = F[q] + h * dF[q];
        Fm[q] = F[q] - h * dF[q];
    }
    Cell<double> cd(force, solid, nu, vconst);
    const double fd = (functional<Q, B3, H, V, double>(cd, gp, Fp, ly, kappa) - functional<Q, B3, H, V, double>(cd, gm, Fm, ly, kappa)) / (2 * h);
    double res[2];
    for (int prec = 0; prec < 2; ++prec) {
        double dot = 0.0;
        if (prec == 0) {
            Cell<double> c(force, solid, nu, vconst);
            double D[Q], lyy[Q], lg[Q], lF[3];
            for (int l = 0; l < Q; ++l) {
                D[l] = g[l] - Lt::w(l);
                lyy[l] = ly[l];
            }
            for (int q = 0; q < 3; ++q) c.F[q][c.oc] = F[q];
            collide_adjoint<Q, B3, H, V, double>(c.a, D, 1, c.oc, solid, lyy, kappa, true, lg, lF);
            for (int l = 0; l < Q; ++l) dot += lg[l] * dg[l];
            for (int q = 0; q < 3; ++q) dot += lF[q] * dF[q];
        } else {
            Cell<float> c(force, solid, nu, vconst);
            float D[Q], lyy[Q], lg[Q], lF[3];
            for (int l = 0; l < Q; ++l) {
                D[l] = (float)g[l] - Lattice<Q>::wf(l);
                lyy[l] = (float)ly[l];
            }
            for (int q = 0; q < 3; ++q) c.F[q][c.oc] = (float)F[q];
            collide_adjoint<Q, B3, H, V, float>(c.a, D, 1, c.oc, solid, lyy, (float)kappa, true, lg, lF);
            for (int l = 0; l < Q; ++l) dot += (double)lg[l] * dg[l];
            for (int q = 0; q < 3; ++q) dot += (double)lF[q] * dF[q];
        }
        res[prec] = dot;
    }
    const double scale = std::fabs(fd) > 1e-300 ? std::fabs(fd) : 1.0;
    printf("{\"q\": %d, \"bgk3\": %d, \"hrr\": %d, \"vre\": %d, \"force\": %d, \"solid\": %d, \"kappa\": %g, \"trial\": %d, "
           "\"fd\": %.15e, \"adj64\": %.15e, \"adj32\": %.9e, \"rel64\": %.3e, \"rel32\": %.3e}\n",
           Q, (int)B3, (int)H, (int)V, (int)force, (int)solid, kappa, trial, fd, res[0], res[1],
           std::fabs(res[0] - fd) / scale, std::fabs(res[1] - fd) / scale);
}

template <int Q, bool B3, bool H, bool V>
void run_all(std::mt19937_64& rng) {
    for (int trial = 0; trial < 3; ++trial)
        for (int f = 0; f < 2; ++f)
            for (int s = 0; s < 2; ++s) run_case<Q, B3, H, V>(rng, f, s, trial == 2 ? 0.7 : 0.0, trial);
}

}  // namespace

int main() {
    std::mt19937_64 rng(12345);
    run_all<19, false, false, false>(rng);
    run_all<19, false, true, true>(rng);
    run_all<19, true, true, true>(rng);
    run_all<19, false, false, true>(rng);
    run_all<19, true, true, false>(rng);
    run_all<27, true, true, true>(rng);
    run_all<27, false, false, false>(rng);
    run_all<27, true, false, true>(rng);
    return 0;
}
