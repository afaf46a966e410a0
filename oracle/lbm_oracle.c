/*
 * lbm_oracle.c -- CPU restatement of geirev/LBM's per-timestep lattice update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in lbm_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never does.
 *
 * It restates, loop for loop and in the reference's floating-point evaluation order, the
 * routines the reference's time loop calls (src/main.F90:283-329):
 *   postcoll_kernel            src/m_postcoll_kernel.F90:7-255
 *   macrovars_kernel           src/m_macrovars_kernel.F90:6-53
 *   drift_kernel               src/m_drift_kernel.F90:7-47
 *   boundarycond (pass order)  src/m_boundarycond.F90:14-468 and the kernels it calls
 *   solids                     src/m_solids.F90:3-42
 *   forcings_apply_kernel      src/m_forcings_apply_kernel.F90:7-216
 *   inflow_turbulence_forcing  src/m_inflow_turbulence_forcing.F90:5-135
 *   fequil3_block_kernel       src/m_fequil3_block_kernel.F90:7-117
 *   inflow_turbulence_apply    src/m_inflow_turbulence_apply.F90:4-31
 *   turbine_forcing (CPU path) src/m_turbine_forcing.F90:8-140 and the routines it calls
 *   buoyancy_forcing_kernel    src/m_buoyancy_forcing_kernel.F90:6-49
 *   fequil3_kernel             src/m_fequil3_kernel.F90:7-124 (initialisation)
 *   fillghosts                 src/m_fillghosts.F90:4-28 (initialisation)
 * Array layout is the reference's: f(nl,0:nx+1,0:ny+1,0:nz+1), column major, so element
 * (l,i,j,k) (l from 0 here) sits at l + nl*(i + (nx+2)*(j + (ny+2)*k)).
 *
 * Built twice: -DREAL=float (the reference's default build) and -DREAL=double (its
 * -fdefault-real-8 build).  Compile with -O2 -ffp-contract=off and no fast-math so that
 * every operation rounds as the reference's x86-64 build does.
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif

#define OR_EXPORT __attribute__((visibility("default")))

/* Lattice tables: src/mod_D3Q27setup.F90:28-35 (D3Q19) and :40-48 (D3Q27).  The order of
 * the directions is part of the restart-file format, so it is kept as is. */
static const int cx19[19] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1,-1, 1, 0, 0,-1, 1, 0, 0};
static const int cy19[19] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1, 0, 0,-1, 1};
static const int cz19[19] = {0, 0, 0, 0, 0,-1, 1, 0, 0, 0, 0,-1, 1, 1,-1, 1,-1, 1,-1};
static const int cx27[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1,-1, 1, 0, 0,-1, 1, 0, 0,-1, 1,-1, 1, 1,-1,-1, 1};
static const int cy27[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1, 0, 0,-1, 1, 1,-1,-1, 1, 1,-1, 1,-1};
static const int cz27[27] = {0, 0, 0, 0, 0,-1, 1, 0, 0, 0, 0,-1, 1, 1,-1, 1,-1, 1,-1, 1,-1,-1, 1,-1, 1,-1, 1};

typedef struct {
    int nx, ny, nz, nl;
    int ibgk, ihrr, ivreman;
    int ibnd, jbnd, kbnd;
    int iturb_pos;           /* slice of the inflow-turbulence forcing (reference: 10) */
    REAL kinevisc;           /* non-dimensional kinematic viscosity */
    REAL smagorinsky;        /* Vreman constant c = 2.5*smagorinsky**2 (src/m_postcoll.F90:46) */
    REAL rho0;               /* reference density, 1.0 (src/m_readinfile.F90:211) */
} or_params;

typedef struct {
    int nl;
    int cx[27], cy[27], cz[27];
    REAL w[27];
    REAL H2[27][3][3];       /* H2[l][q][p] <-> H2(p,q,l) */
    REAL H3[27][3][3][3];    /* H3[l][r][q][p] <-> H3(p,q,r,l) */
    REAL cs2, cs4, cs6, inv1cs2, inv2cs4, inv6cs6;
} lattice;

static void lattice_init(lattice *L, int nl)
{
    const int *cx = nl == 19 ? cx19 : cx27, *cy = nl == 19 ? cy19 : cy27, *cz = nl == 19 ? cz19 : cz27;
    L->nl = nl;
    L->cs2 = (REAL)1.0 / (REAL)3.0;
    L->cs4 = (REAL)1.0 / (REAL)9.0;
    L->cs6 = (REAL)1.0 / (REAL)27.0;
    L->inv1cs2 = (REAL)1.0 / L->cs2;
    L->inv2cs4 = (REAL)1.0 / ((REAL)2.0 * L->cs4);
    L->inv6cs6 = (REAL)1.0 / ((REAL)6.0 * L->cs6);
    for (int l = 0; l < nl; l++) {
        L->cx[l] = cx[l]; L->cy[l] = cy[l]; L->cz[l] = cz[l];
        int n2 = cx[l] * cx[l] + cy[l] * cy[l] + cz[l] * cz[l];
        if (nl == 19)
            L->w[l] = n2 == 0 ? (REAL)1.0 / (REAL)3.0 : n2 == 1 ? (REAL)1.0 / (REAL)18.0 : (REAL)1.0 / (REAL)36.0;
        else
            L->w[l] = n2 == 0 ? (REAL)8.0 / (REAL)27.0 : n2 == 1 ? (REAL)2.0 / (REAL)27.0
                    : n2 == 2 ? (REAL)1.0 / (REAL)54.0 : (REAL)1.0 / (REAL)216.0;
        REAL c[3] = {(REAL)cx[l], (REAL)cy[l], (REAL)cz[l]};
        /* src/mod_D3Q27setup.F90:130 and :143-146 */
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++)
                L->H2[l][q][p] = c[p] * c[q] - L->cs2 * (REAL)(p == q);
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++)
                for (int r = 0; r < 3; r++)
                    L->H3[l][r][q][p] = c[p] * c[q] * c[r]
                        - L->cs2 * (c[p] * (REAL)(q == r) + c[q] * (REAL)(p == r) + c[r] * (REAL)(p == q));
    }
}

#define IDX(P, l, i, j, k) ((size_t)(l) + (size_t)(P)->nl * ((size_t)(i) + (size_t)((P)->nx + 2) * ((size_t)(j) + (size_t)((P)->ny + 2) * (size_t)(k))))
#define CIDX(P, i, j, k) ((size_t)(i) + (size_t)((P)->nx + 2) * ((size_t)(j) + (size_t)((P)->ny + 2) * (size_t)(k)))

/* ---------------------------------------------------------------------------------- */
/* feq of one cell, as in postcoll_kernel lines 136-183 (src/m_postcoll_kernel.F90). */
static void feq_postcoll(const lattice *L, int ibgk, REAL dens, const REAL vel[3], REAL *tmpeq)
{
    REAL A2[3][3], A3[3][3][3];
    REAL ratio = L->inv6cs6 / L->inv2cs4;
    for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++)
            A2[q][p] = dens * vel[p] * vel[q] * L->inv2cs4;
    if (ibgk == 3)
        for (int r = 0; r < 3; r++) {
            REAL vratio = vel[r] * ratio;
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++)
                    A3[r][q][p] = A2[q][p] * vratio;
        }
    for (int l = 0; l < L->nl; l++) {
        REAL cu = (REAL)L->cx[l] * vel[0] + (REAL)L->cy[l] * vel[1] + (REAL)L->cz[l] * vel[2];
        REAL t = dens * ((REAL)1.0 + cu * L->inv1cs2);
        for (int q = 0; q < 3; q++)
            for (int p = 0; p < 3; p++)
                t = t + L->H2[l][q][p] * A2[q][p];
        if (ibgk == 3 && l > 0)
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        t = t + L->H3[l][r][q][p] * A3[r][q][p];
        tmpeq[l] = L->w[l] * t;
    }
}

/* feq as fequil3_block_kernel / fequil3_kernel compute it (tmp1 + tmp2 + tmp3 summed
 * separately; src/m_fequil3_block_kernel.F90:62-105).  order3 selects the H3 term. */
static void feq_block(const lattice *L, int order3, REAL rr, const REAL vel[3], REAL *feq)
{
    REAL A2[3][3], A3[3][3][3];
    REAL ratio = L->inv6cs6 / L->inv2cs4;
    for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++)
            A2[q][p] = rr * vel[p] * vel[q] * L->inv2cs4;
    if (order3)
        for (int r = 0; r < 3; r++) {
            REAL vratio = vel[r] * ratio;
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++)
                    A3[r][q][p] = A2[q][p] * vratio;
        }
    for (int l = 0; l < L->nl; l++) {
        REAL cu = (REAL)L->cx[l] * vel[0] + (REAL)L->cy[l] * vel[1] + (REAL)L->cz[l] * vel[2];
        REAL tmp1 = rr * ((REAL)1.0 + cu * L->inv1cs2);
        REAL tmp2 = 0;
        for (int q = 0; q < 3; q++)
            for (int p = 0; p < 3; p++)
                tmp2 = tmp2 + L->H2[l][q][p] * A2[q][p];
        REAL tmp3 = 0;
        if (order3)
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        tmp3 = tmp3 + L->H3[l][r][q][p] * A3[r][q][p];
        feq[l] = L->w[l] * (tmp1 + tmp2 + tmp3);
    }
}

/* ---------------------------------------------------------------------------------- */
/* postcoll: src/m_postcoll.F90:4-65 and src/m_postcoll_kernel.F90:7-255. */
OR_EXPORT void or_postcoll(const or_params *P, REAL *f, REAL *tau, const REAL *rho, const REAL *u,
                           const REAL *v, const REAL *w, const REAL *ablvisc)
{
    lattice L; lattice_init(&L, P->nl);
    const int nl = P->nl;
    REAL cnst = (REAL)2.5 * (P->smagorinsky * P->smagorinsky);   /* 2.5*smagorinsky**2 */
    REAL eps = (sizeof(REAL) == 4) ? (REAL)sqrtf(FLT_MIN) : (REAL)sqrt(DBL_MIN);
    REAL tiny = (sizeof(REAL) == 4) ? (REAL)FLT_MIN : (REAL)DBL_MIN;
    REAL tmpeq[27], tmpneq[27];
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++)
    for (int i = 1; i <= P->nx; i++) {
        size_t c = CIDX(P, i, j, k);
        REAL vel[3] = {u[c], v[c], w[c]};
        REAL dens = rho[c];
        REAL *fc = f + IDX(P, 0, i, j, k);
        feq_postcoll(&L, P->ibgk, dens, vel, tmpeq);
        if (P->ihrr == 1) {
            REAL A2[3][3] = {{0}}, A3[3][3][3];
            for (int l = 0; l < nl; l++) {
                tmpneq[l] = fc[l] - tmpeq[l];
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        A2[q][p] = A2[q][p] + L.H2[l][q][p] * tmpneq[l];
            }
            /* Malaspinas recursion (lines 201-208) */
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        A3[r][q][p] = vel[p] * A2[r][q] + vel[q] * A2[p][r] + vel[r] * A2[q][p];
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++)
                    A2[q][p] = A2[q][p] * L.inv2cs4;
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        A3[r][q][p] = A3[r][q][p] * L.inv6cs6;
            for (int l = 0; l < nl; l++) {
                REAL t = 0;
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        t = t + L.H2[l][q][p] * A2[q][p];
                for (int r = 0; r < 3; r++)
                    for (int q = 0; q < 3; q++)
                        for (int p = 0; p < 3; p++)
                            t = t + L.H3[l][r][q][p] * A3[r][q][p];
                tmpneq[l] = L.w[l] * t;
            }
        } else {
            for (int l = 0; l < nl; l++)
                tmpneq[l] = fc[l] - tmpeq[l];
        }
        if (P->ivreman == 1) {
            /* lines 255-304 */
            REAL alpha[3][3] = {{0}}, beta[3][3];
            for (int l = 0; l < nl; l++)
                for (int q = 0; q < 3; q++)
                    for (int p = 0; p < 3; p++)
                        alpha[q][p] = alpha[q][p] + L.H2[l][q][p] * tmpneq[l];
            REAL alphamag = 0;
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++)
                    alphamag = alphamag + alpha[q][p] * alpha[q][p];
            alphamag = alphamag > tiny ? alphamag : tiny;
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++) {
                    REAL b = eps;
                    for (int m = 0; m < 3; m++)
                        b = b + alpha[p][m] * alpha[q][m];   /* alpha(m,p)*alpha(m,q) */
                    beta[q][p] = b;
                }
            /* beta[q][p] <-> beta(p,q) */
            REAL b11 = beta[0][0], b22 = beta[1][1], b33 = beta[2][2];
            REAL b12 = beta[1][0], b13 = beta[2][0], b23 = beta[2][1];
            REAL Bbeta = b11 * b22 - b12 * b12 + b11 * b33 - b13 * b13 + b22 * b33 - b23 * b23;
            REAL tautmp = Bbeta / alphamag;
            REAL eddy = (tautmp > eps) ? cnst * ((sizeof(REAL) == 4) ? (REAL)sqrtf((float)tautmp) : (REAL)sqrt((double)tautmp)) : (REAL)0;
            tau[c] = (REAL)3.0 * (P->kinevisc + eddy + ablvisc[k - 1]) + (REAL)0.5;
        }
        REAL fac = (REAL)1.0 - (REAL)1.0 / tau[c];
        for (int l = 0; l < nl; l++)
            fc[l] = tmpeq[l] + fac * tmpneq[l];
    }
}

/* macrovars: src/m_macrovars_kernel.F90:6-53, all cells including ghosts. */
OR_EXPORT void or_macrovars(const or_params *P, const REAL *f, REAL *rho, REAL *u, REAL *v, REAL *w)
{
    lattice L; lattice_init(&L, P->nl);
    for (int k = 0; k <= P->nz + 1; k++)
    for (int j = 0; j <= P->ny + 1; j++)
    for (int i = 0; i <= P->nx + 1; i++) {
        const REAL *fc = f + IDX(P, 0, i, j, k);
        REAL r = 0, tu = 0, tv = 0, tw = 0;
        for (int l = 0; l < P->nl; l++) {
            REAL t = fc[l];
            r = r + t;
            tu = tu + (REAL)L.cx[l] * t;
            tv = tv + (REAL)L.cy[l] * t;
            tw = tw + (REAL)L.cz[l] * t;
        }
        REAL inv = (REAL)1.0 / r;
        size_t c = CIDX(P, i, j, k);
        rho[c] = r; u[c] = tu * inv; v[c] = tv * inv; w[c] = tw * inv;
    }
}

/* drift: f(l,i,j,k) = feq(l, i-cx, j-cy, k-cz) over the interior
 * (src/m_drift_kernel.F90:37-42).  Ghosts of the target are not touched. */
OR_EXPORT void or_drift(const or_params *P, REAL *fout, const REAL *fin)
{
    lattice L; lattice_init(&L, P->nl);
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++)
    for (int i = 1; i <= P->nx; i++)
        for (int l = 0; l < P->nl; l++)
            fout[IDX(P, l, i, j, k)] = fin[IDX(P, l, i - L.cx[l], j - L.cy[l], k - L.cz[l])];
}

/* solids: swap f(l)<->f(l+1), l=2,4,..,nl-1 (1-based) in blanked cells (src/m_solids.F90:19-32). */
OR_EXPORT void or_solids(const or_params *P, REAL *f, const uint8_t *blank)
{
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++)
    for (int i = 1; i <= P->nx; i++)
        if (blank[CIDX(P, i, j, k)]) {
            REAL *fc = f + IDX(P, 0, i, j, k);
            for (int l = 1; l < P->nl - 1; l += 2) {
                REAL t = fc[l]; fc[l] = fc[l + 1]; fc[l + 1] = t;
            }
        }
}

/* forcings_apply_kernel: Kupershtokh exact difference with 2nd-order feq
 * (src/m_forcings_apply_kernel.F90:142-271).  F is external_forcing(3,0:nx+1,0:ny+1,0:nz+1).
 * The GPU kernel skips a cell with rho<=0 (line 157); this follows the GPU semantics. */
OR_EXPORT void or_forcings_apply(const or_params *P, REAL *f, const REAL *F, const REAL *rho,
                                 const REAL *u, const REAL *v, const REAL *w)
{
    lattice L; lattice_init(&L, P->nl);
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++)
    for (int i = 1; i <= P->nx; i++) {
        size_t c = CIDX(P, i, j, k);
        REAL Fx = F[3 * c], Fy = F[3 * c + 1], Fz = F[3 * c + 2];
        REAL rr = rho[c];
        if (rr <= (REAL)0) continue;
        REAL vel0[3] = {u[c], v[c], w[c]};
        REAL A0[3][3], A1[3][3];
        for (int q = 0; q < 3; q++)
            for (int p = 0; p < 3; p++)
                A0[q][p] = rr * vel0[p] * vel0[q] * L.inv2cs4;
        REAL du = -Fx / rr, dv = -Fy / rr, dw = -Fz / rr;
        REAL vel1[3] = {vel0[0] + du, vel0[1] + dv, vel0[2] + dw};
        for (int q = 0; q < 3; q++)
            for (int p = 0; p < 3; p++)
                A1[q][p] = rr * vel1[p] * vel1[q] * L.inv2cs4;
        REAL *fc = f + IDX(P, 0, i, j, k);
        for (int l = 0; l < P->nl; l++) {
            REAL cu = (REAL)L.cx[l] * vel0[0] + (REAL)L.cy[l] * vel0[1] + (REAL)L.cz[l] * vel0[2];
            REAL tmp1 = rr * ((REAL)1.0 + cu * L.inv1cs2);
            REAL tmp2 = 0, tmp3 = 0;
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++)
                    tmp2 = tmp2 + L.H2[l][q][p] * A0[q][p];
            REAL feq0 = L.w[l] * (tmp1 + tmp2 + tmp3);
            cu = (REAL)L.cx[l] * vel1[0] + (REAL)L.cy[l] * vel1[1] + (REAL)L.cz[l] * vel1[2];
            tmp1 = rr * ((REAL)1.0 + cu * L.inv1cs2);
            tmp2 = 0;
            for (int q = 0; q < 3; q++)
                for (int p = 0; p < 3; p++)
                    tmp2 = tmp2 + L.H2[l][q][p] * A1[q][p];
            REAL feq1 = L.w[l] * (tmp1 + tmp2 + tmp3);
            fc[l] = fc[l] + (feq1 - feq0);
        }
    }
}

/* buoyancy_forcing_kernel: external_forcing(3,idx) - scaling*(pottemp(idx) - theta0)/theta0 for the
 * interior cells, evaluated left to right as the Fortran expression reads
 * (src/m_buoyancy_forcing_kernel.F90:20-37).  pottemp is (0:nx+1,0:ny+1,0:nz+1). */
OR_EXPORT void or_buoyancy_forcing(const or_params *P, REAL *F, const REAL *pottemp, REAL scaling, REAL theta0)
{
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++)
    for (int i = 1; i <= P->nx; i++) {
        size_t c = CIDX(P, i, j, k);
        REAL d = pottemp[c] - theta0;
        REAL s = scaling * d;
        F[3 * c + 2] = F[3 * c + 2] - s / theta0;
    }
}

/* inflow_turbulence_forcing: df(l,j,k) = feq(rho, u + ampl*(uu,vv,ww)(j,k)) - feq(rho,u) on
 * slice i=ip from the pre-collision moments (src/m_inflow_turbulence_forcing.F90:57-128).
 * uu,vv,ww here are the (ny,nz) slices for lit = mod(it,nrturb) (0 -> nrturb), column major.
 * df is turbulence_df(nl,ny,nz). */
/* ---------------------------------- turbines ---------------------------------- */
/* The actuator-line turbines of src/m_turbine_forcing.F90:8-140 on the CPU path the reference's
 * serial/OpenMP build takes (no MPI, no _CUDA): azimuth update, point distribution
 * (src/m_turbine_distribute_points.F90:6-64, basis src/m_turbine_rotor_basis.F90:13-23), point
 * forces (src/m_turbine_point_forces_kernel.F90:5-68 with src/m_turbine_interpolation.F90:25-114,
 * src/m_nrelliftdrag.F90:7-35, src/m_turbine_compute_bladeforce.F90:20-65) and the Gaussian deposit
 * (src/m_turbine_deposit.F90:7-91), every operation in the reference's order. */
typedef struct {
    int nturb, nblades, nchords, nfoils, nang_max;
    const REAL *xhub, *yhub, *zhub, *yaw, *tilt, *omegand, *pitch;   /* [nturb] */
    REAL *theta;                                                      /* [nturb], advanced per step */
    const REAL *relm, *dc, *chord, *twist;                            /* [nchords] */
    const int *nfoil;                                                 /* [nchords], 1-based */
    const int *nang;                                                  /* [nfoils] */
    const REAL *foil;                                                 /* [nfoils][nang_max][3]: deg, cl, cd */
} or_turbines;

/* math in the REAL kind (cosf for float, as the FP32 build calls it; cos for double) */
#define RCOS(x) _Generic((x), float: cosf, default: cos)(x)
#define RSIN(x) _Generic((x), float: sinf, default: sin)(x)
#define REXP(x) _Generic((x), float: expf, default: exp)(x)
#define RFLOOR(x) _Generic((x), float: floorf, default: floor)(x)
#define RCEIL(x) _Generic((x), float: ceilf, default: ceil)(x)
#define RSQRT(x) _Generic((x), float: sqrtf, default: sqrt)(x)
#define RATAN2(y, x) _Generic((y), float: atan2f, default: atan2)(y, x)

static const REAL OR_PI = (REAL)3.14159265358979323846;   /* src/mod_turbines.F90:10-11 */

static void rotor_basis(REAL yaw, REAL tilt, REAL ea[3], REAL e1[3], REAL e2[3])
{
    ea[0] = RCOS(yaw) * RCOS(tilt); ea[1] = RSIN(yaw) * RCOS(tilt); ea[2] = -RSIN(tilt);
    e1[0] = -RSIN(yaw);            e1[1] = RCOS(yaw);             e1[2] = (REAL)0.0;
    e2[0] = -RCOS(yaw) * RSIN(tilt); e2[1] = -RSIN(yaw) * RSIN(tilt); e2[2] = -RCOS(tilt);
}

static REAL trilinear(REAL c000, REAL c100, REAL c010, REAL c110, REAL c001, REAL c101, REAL c011, REAL c111,
                      REAL fx, REAL fy, REAL fz)
{
    REAL one = (REAL)1;
    return c000 * (one - fx) * (one - fy) * (one - fz) + c100 * fx * (one - fy) * (one - fz) +
           c010 * (one - fx) * fy * (one - fz) + c110 * fx * fy * (one - fz) +
           c001 * (one - fx) * (one - fy) * fz + c101 * fx * (one - fy) * fz +
           c011 * (one - fx) * fy * fz + c111 * fx * fy * fz;
}

static REAL interp(const or_params *P, const REAL *a, int i, int j, int k, REAL fx, REAL fy, REAL fz)
{
    return trilinear(a[CIDX(P, i, j, k)], a[CIDX(P, i + 1, j, k)], a[CIDX(P, i, j + 1, k)], a[CIDX(P, i + 1, j + 1, k)],
                     a[CIDX(P, i, j, k + 1)], a[CIDX(P, i + 1, j, k + 1)], a[CIDX(P, i, j + 1, k + 1)],
                     a[CIDX(P, i + 1, j + 1, k + 1)], fx, fy, fz);
}

static int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* F (3,0:nx+1,0:ny+1,0:nz+1) receives the turbine forcing of this step (the deposit; zero elsewhere) */
OR_EXPORT void or_turbine_forcing(const or_params *P, const or_turbines *T, const REAL *rho, const REAL *u,
                                  const REAL *v, const REAL *w, REAL *F)
{
    const REAL pi2 = (REAL)2.0 * OR_PI;
    for (int t = 0; t < T->nturb; t++) T->theta[t] = T->theta[t] + T->omegand[t];
    int np = T->nturb * T->nblades * T->nchords;
    REAL *pts = (REAL *)malloc(sizeof(REAL) * 6 * (size_t)(np > 0 ? np : 1));   /* xg, yg, zg, F */
    int p = 0;
    for (int t = 0; t < T->nturb; t++) {
        REAL ea[3], e1[3], e2[3];
        rotor_basis(T->yaw[t], T->tilt[t], ea, e1, e2);
        for (int ib = 1; ib <= T->nblades; ib++) {
            REAL theta = T->theta[t] + (REAL)(ib - 1) * pi2 / (REAL)T->nblades;
            for (int ic = 0; ic < T->nchords; ic++, p++) {
                REAL er[3];
                for (int d = 0; d < 3; d++) er[d] = RCOS(theta) * e1[d] + RSIN(theta) * e2[d];
                REAL xg = T->xhub[t] + T->relm[ic] * er[0];
                REAL yg = T->yhub[t] + T->relm[ic] * er[1];
                REAL zg = T->zhub[t] + T->relm[ic] * er[2];
                /* turbine_interpolate_velocity, j_start = 1 */
                int ig = (int)RFLOOR(xg), jg = (int)RFLOOR(yg), kg = (int)RFLOOR(zg);
                REAL fx = xg - (REAL)ig, fy = yg - (REAL)jg, fz = zg - (REAL)kg;
                int i = clampi(ig, 1, P->nx - 1), j = clampi(jg, 1, P->ny - 1), k = clampi(kg, 1, P->nz - 1);
                REAL ux = interp(P, u, i, j, k, fx, fy, fz), uy = interp(P, v, i, j, k, fx, fy, fz);
                REAL uz = interp(P, w, i, j, k, fx, fy, fz), dens = interp(P, rho, i, j, k, fx, fy, fz);
                REAL costheta = RCOS(theta), sintheta = RSIN(theta);
                REAL utheta = T->omegand[t] * T->relm[ic] - uz * costheta - uy * sintheta;
                REAL phi = RATAN2(ux, utheta);
                REAL ang = phi * (REAL)180.0 / OR_PI - T->twist[ic] - T->pitch[t];
                /* nrelliftdrag */
                REAL cl = 0, cd = 0;
                int kf = T->nfoil[ic] - 1;
                const REAL *tab = T->foil + (size_t)kf * T->nang_max * 3;
                for (int a = 0; a < T->nang[kf] - 1; a++) {
                    if (tab[3 * a] < ang && ang <= tab[3 * (a + 1)]) {
                        REAL slope = (tab[3 * (a + 1) + 1] - tab[3 * a + 1]) / (tab[3 * (a + 1)] - tab[3 * a]);
                        cl = tab[3 * a + 1] + slope * (ang - tab[3 * a]);
                        slope = (tab[3 * (a + 1) + 2] - tab[3 * a + 2]) / (tab[3 * (a + 1)] - tab[3 * a]);
                        cd = tab[3 * a + 2] + slope * (ang - tab[3 * a]);
                        break;
                    }
                }
                /* turbine_compute_bladeforce */
                REAL et[3];
                for (int d = 0; d < 3; d++) { er[d] = RCOS(theta) * e1[d] + RSIN(theta) * e2[d]; et[d] = -RSIN(theta) * e1[d] + RCOS(theta) * e2[d]; }
                REAL speed2 = ux * ux + utheta * utheta;
                REAL area = T->chord[ic] * T->dc[ic];
                REAL tiploss = (REAL)1.0;
                REAL Lf = (REAL)0.5 * dens * speed2 * cl * area * tiploss;
                REAL Df = (REAL)0.5 * dens * speed2 * cd * area * tiploss;
                REAL ph = RATAN2(ux, utheta), sinphi = RSIN(ph), cosphi = RCOS(ph);
                pts[6 * p] = xg; pts[6 * p + 1] = yg; pts[6 * p + 2] = zg;
                for (int d = 0; d < 3; d++)
                    pts[6 * p + 3 + d] = Lf * (cosphi * ea[d] + sinphi * et[d]) - Df * (sinphi * ea[d] - cosphi * et[d]);
            }
        }
    }
    /* turbine_deposit into F_turb = 0, then external_forcing = F_turb */
    size_t ncell = (size_t)(P->nx + 2) * (P->ny + 2) * (P->nz + 2);
    memset(F, 0, sizeof(REAL) * 3 * ncell);
    REAL epsilon = (REAL)2.0;
    REAL sigma = epsilon / (REAL)RSQRT((REAL)2.0);
    int krad = (int)RCEIL((REAL)3.0 * sigma);
    if (krad > 5) krad = 5;
    for (p = 0; p < np; p++) {
        REAL xg = pts[6 * p], yg = pts[6 * p + 1], zg = pts[6 * p + 2];
        REAL Fp[3] = {pts[6 * p + 3], pts[6 * p + 4], pts[6 * p + 5]};
        if (Fp[0] == (REAL)0 && Fp[1] == (REAL)0 && Fp[2] == (REAL)0) continue;
        int i0 = (int)RFLOOR(xg), j0 = (int)RFLOOR(yg), k0 = (int)RFLOOR(zg);
        int ii1 = i0 - krad > 1 ? i0 - krad : 1, ii2 = i0 + krad < P->nx ? i0 + krad : P->nx;
        int jj1 = j0 - krad > 1 ? j0 - krad : 1, jj2 = j0 + krad < P->ny ? j0 + krad : P->ny;
        int kk1 = k0 - krad > 1 ? k0 - krad : 1, kk2 = k0 + krad < P->nz ? k0 + krad : P->nz;
        REAL sumW = 0;
        for (int kk = kk1; kk <= kk2; kk++) {
            REAL dz = (REAL)kk - zg;
            for (int jj = jj1; jj <= jj2; jj++) {
                REAL dy = (REAL)jj - yg;
                for (int ii = ii1; ii <= ii2; ii++) {
                    REAL dx = (REAL)ii - xg;
                    sumW = sumW + REXP(-(dx * dx + dy * dy + dz * dz) / ((REAL)2.0 * sigma * sigma));
                }
            }
        }
        if (sumW < (REAL)1.0e-6) continue;
        for (int kk = kk1; kk <= kk2; kk++) {
            REAL dz = (REAL)kk - zg;
            for (int jj = jj1; jj <= jj2; jj++) {
                REAL dy = (REAL)jj - yg;
                for (int ii = ii1; ii <= ii2; ii++) {
                    REAL dx = (REAL)ii - xg;
                    REAL weight = REXP(-(dx * dx + dy * dy + dz * dz) / ((REAL)2.0 * sigma * sigma)) / sumW;
                    size_t c = CIDX(P, ii, jj, kk);
                    F[3 * c] = F[3 * c] + weight * Fp[0];
                    F[3 * c + 1] = F[3 * c + 1] + weight * Fp[1];
                    F[3 * c + 2] = F[3 * c + 2] + weight * Fp[2];
                }
            }
        }
    }
    free(pts);
}

OR_EXPORT void or_turbulence_forcing(const or_params *P, const REAL *rho, const REAL *u, const REAL *v,
                                     const REAL *w, const REAL *uu, const REAL *vv, const REAL *ww,
                                     REAL ampl, REAL *df)
{
    lattice L; lattice_init(&L, P->nl);
    int ip = P->iturb_pos;
    REAL f1[27], f2[27];
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++) {
        size_t c = CIDX(P, ip, j, k);
        size_t s = (size_t)(j - 1) + (size_t)P->ny * (size_t)(k - 1);
        REAL rr = rho[c];
        REAL vel[3] = {u[c], v[c], w[c]};
        feq_block(&L, P->ibgk == 3, rr, vel, f1);
        REAL vel2[3] = {vel[0] + ampl * uu[s], vel[1] + ampl * vv[s], vel[2] + ampl * ww[s]};
        feq_block(&L, P->ibgk == 3, rr, vel2, f2);
        for (int l = 0; l < P->nl; l++)
            df[(size_t)l + (size_t)P->nl * s] = f2[l] - f1[l];
    }
}

/* inflow_turbulence_apply: f(:,ip,j,k) += df (src/m_inflow_turbulence_apply.F90:21-26). */
OR_EXPORT void or_turbulence_apply(const or_params *P, REAL *f, const REAL *df)
{
    int ip = P->iturb_pos;
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++) {
        size_t s = (size_t)(j - 1) + (size_t)P->ny * (size_t)(k - 1);
        for (int l = 0; l < P->nl; l++)
            f[IDX(P, l, ip, j, k)] = f[IDX(P, l, ip, j, k)] + df[(size_t)l + (size_t)P->nl * s];
    }
}

/* fequil3 (init): feq over the interior (src/m_fequil3_kernel.F90:60-124 uses the
 * postcoll-style accumulation, tmp = tmp + H2*A2 ...). */
OR_EXPORT void or_fequil3(const or_params *P, REAL *f, const REAL *rho, const REAL *u,
                          const REAL *v, const REAL *w)
{
    lattice L; lattice_init(&L, P->nl);
    REAL feq[27];
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++)
    for (int i = 1; i <= P->nx; i++) {
        size_t c = CIDX(P, i, j, k);
        REAL vel[3] = {u[c], v[c], w[c]};
        feq_postcoll(&L, P->ibgk, rho[c], vel, feq);
        for (int l = 0; l < P->nl; l++)
            f[IDX(P, l, i, j, k)] = feq[l];
    }
}

/* fillghosts: ghost = clamped interior cell (src/m_fillghosts.F90:191-197). */
OR_EXPORT void or_fillghosts(const or_params *P, REAL *f)
{
    for (int k = 0; k <= P->nz + 1; k++)
    for (int j = 0; j <= P->ny + 1; j++)
    for (int i = 0; i <= P->nx + 1; i++) {
        int i1 = i < 1 ? 1 : (i > P->nx ? P->nx : i);
        int j1 = j < 1 ? 1 : (j > P->ny ? P->ny : j);
        int k1 = k < 1 ? 1 : (k > P->nz ? P->nz : k);
        for (int l = 0; l < P->nl; l++)
            f[IDX(P, l, i, j, k)] = f[IDX(P, l, i1, j1, k1)];
    }
}

/* ----------------------------------- boundaries ----------------------------------- */
static void bc_periodic_i(const or_params *P, REAL *f)
{   /* src/m_boundary_i_periodic_kernel.F90:26-33 */
    for (int k = 0; k <= P->nz + 1; k++)
    for (int j = 0; j <= P->ny + 1; j++)
        for (int l = 0; l < P->nl; l++) {
            f[IDX(P, l, 0, j, k)] = f[IDX(P, l, P->nx, j, k)];
            f[IDX(P, l, P->nx + 1, j, k)] = f[IDX(P, l, 1, j, k)];
        }
}

static void bc_periodic_j(const or_params *P, REAL *f)
{   /* src/m_boundary_j_periodic_kernel.F90:69-76 */
    for (int k = 0; k <= P->nz + 1; k++)
    for (int i = 0; i <= P->nx + 1; i++)
        for (int l = 0; l < P->nl; l++) {
            f[IDX(P, l, i, 0, k)] = f[IDX(P, l, i, P->ny, k)];
            f[IDX(P, l, i, P->ny + 1, k)] = f[IDX(P, l, i, 1, k)];
        }
}

static void bc_periodic_k(const or_params *P, REAL *f)
{   /* src/m_boundary_k_periodic_kernel.F90:26-27 */
    size_t plane = (size_t)P->nl * (P->nx + 2) * (P->ny + 2);
    memcpy(f, f + plane * P->nz, plane * sizeof(REAL));
    memcpy(f + plane * (P->nz + 1), f + plane, plane * sizeof(REAL));
}

static int find_dir(const lattice *L, int mx, int my, int mz)
{
    for (int m = 0; m < L->nl; m++)
        if (L->cx[m] == mx && L->cy[m] == my && L->cz[m] == mz)
            return m;
    return -1;
}

/* Closed walls: src/m_boundary_i_closed_kernel.F90:44-58, ..._j_...:103-114, ..._k_...:157-168.
 * dir: 0=i,1=j,2=k; plane = 1 or n (n = nx, ny or nz). */
static void bc_closed(const or_params *P, const lattice *L, REAL *f1, const REAL *f2, int dir, int plane, int opt)
{
    int n = dir == 0 ? P->nx : dir == 1 ? P->ny : P->nz;
    int ghost = plane == 1 ? 0 : n + 1;
    int na = dir == 0 ? P->ny : P->nx;   /* first tangential range */
    int nb = dir == 2 ? P->ny : P->nz;   /* second tangential range */
    for (int b = 1; b <= nb; b++)
    for (int a = 1; a <= na; a++) {
        int i, j, k;
        if (dir == 0) { i = ghost; j = a; k = b; }
        else if (dir == 1) { i = a; j = ghost; k = b; }
        else { i = a; j = b; k = ghost; }
        for (int l = 0; l < L->nl; l++) {
            int cn = dir == 0 ? L->cx[l] : dir == 1 ? L->cy[l] : L->cz[l];
            if (!((plane == 1 && cn < 0) || (plane == n && cn > 0)))
                continue;
            int mx, my, mz;
            if (dir == 0) { mx = -L->cx[l]; my = opt * L->cy[l]; mz = opt * L->cz[l]; }
            else if (dir == 1) { mx = opt * L->cx[l]; my = -L->cy[l]; mz = opt * L->cz[l]; }
            else { mx = opt * L->cx[l]; my = opt * L->cy[l]; mz = -L->cz[l]; }
            int m = find_dir(L, mx, my, mz);
            if (m < 0) continue;
            f1[IDX(P, m, i, j, k)] = f2[IDX(P, l, i, j, k)];
            int si = dir == 0 ? plane : i - L->cx[l];
            int sj = dir == 1 ? plane : j - L->cy[l];
            int sk = dir == 2 ? plane : k - L->cz[l];
            f1[IDX(P, l, i, j, k)] = f1[IDX(P, l, si, sj, sk)];
        }
    }
}

/* Edges: src/m_boundary_j_k_edges.F90, _i_j_edges.F90, _i_k_edges.F90.  d1,d2 are the two
 * closed directions, s1,s2 in {0 (low), 1 (high)} pick the edge.  mirror[3] gives the
 * multiplier applied to each velocity component for the incoming direction. */
static void bc_edge(const or_params *P, const lattice *L, REAL *f1, const REAL *f2,
                    int d1, int s1, int d2, int s2, const int mirror[3])
{
    int n[3] = {P->nx, P->ny, P->nz};
    int t = 3 - d1 - d2;              /* tangential direction */
    int g[3], in[3];
    g[d1] = s1 ? n[d1] + 1 : 0; in[d1] = s1 ? n[d1] : 1;
    g[d2] = s2 ? n[d2] + 1 : 0; in[d2] = s2 ? n[d2] : 1;
    /* pass 1: incoming populations from the other buffer */
    for (int a = 1; a <= n[t]; a++) {
        g[t] = a;
        for (int l = 0; l < L->nl; l++) {
            int c[3] = {L->cx[l], L->cy[l], L->cz[l]};
            int out1 = s1 ? c[d1] > 0 : c[d1] < 0;
            int out2 = s2 ? c[d2] > 0 : c[d2] < 0;
            if (!(out1 && out2)) continue;
            int m = find_dir(L, mirror[0] * c[0], mirror[1] * c[1], mirror[2] * c[2]);
            if (m < 0) continue;
            f1[IDX(P, m, g[0], g[1], g[2])] = f2[IDX(P, l, g[0], g[1], g[2])];
        }
    }
    /* pass 2: outgoing populations from the interior edge cell (normal shift only) */
    for (int a = 1; a <= n[t]; a++) {
        g[t] = a; in[t] = a;
        for (int l = 0; l < L->nl; l++) {
            int c[3] = {L->cx[l], L->cy[l], L->cz[l]};
            int out1 = s1 ? c[d1] > 0 : c[d1] < 0;
            int out2 = s2 ? c[d2] > 0 : c[d2] < 0;
            if (!(out1 && out2)) continue;
            f1[IDX(P, l, g[0], g[1], g[2])] = f1[IDX(P, l, in[0], in[1], in[2])];
        }
    }
}

static void bc_edges_jk(const or_params *P, const lattice *L, REAL *f1, const REAL *f2, int opt)
{   /* src/m_boundary_j_k_edges.F90: tangential x flipped by opt, except the (ny+1,nz+1)
     * edge whose mirror keeps cx (line 129). */
    int mir[3] = {opt, -1, -1}, mir_hh[3] = {1, -1, -1};
    bc_edge(P, L, f1, f2, 1, 0, 2, 0, mir);
    bc_edge(P, L, f1, f2, 1, 1, 2, 0, mir);
    bc_edge(P, L, f1, f2, 1, 0, 2, 1, mir);
    bc_edge(P, L, f1, f2, 1, 1, 2, 1, mir_hh);
}

static void bc_edges_ij(const or_params *P, const lattice *L, REAL *f1, const REAL *f2, int opt)
{   /* src/m_boundary_i_j_edges.F90 */
    int mir[3] = {-1, -1, opt};
    bc_edge(P, L, f1, f2, 0, 0, 1, 0, mir);
    bc_edge(P, L, f1, f2, 0, 1, 1, 0, mir);
    bc_edge(P, L, f1, f2, 0, 0, 1, 1, mir);
    bc_edge(P, L, f1, f2, 0, 1, 1, 1, mir);
}

static void bc_edges_ik(const or_params *P, const lattice *L, REAL *f1, const REAL *f2, int opt)
{   /* src/m_boundary_i_k_edges.F90 */
    int mir[3] = {-1, opt, -1};
    bc_edge(P, L, f1, f2, 0, 0, 2, 0, mir);
    bc_edge(P, L, f1, f2, 0, 1, 2, 0, mir);
    bc_edge(P, L, f1, f2, 0, 0, 2, 1, mir);
    bc_edge(P, L, f1, f2, 0, 1, 2, 1, mir);
}

/* The reference's corner table is reshape([24 values], [8,3]) (lines 177-186).  reshape
 * fills column major, so corner m is (list[m], list[8+m], list[16+m]) of the 24 listed
 * values, not the m-th listed triple.  Only entries whose three indices are ghost indices of
 * their own axis pass the outgoing test below; the others do nothing.  The table is kept
 * as the reference builds it. */
OR_EXPORT void or_corner_table(int nx, int ny, int nz, int cs[8][3])
{
    int list[24] = {0, 0, 0,   0, 0, nz + 1,   0, ny + 1, 0,   0, ny + 1, nz + 1,
                    nx + 1, 0, 0,   nx + 1, 0, nz + 1,   nx + 1, ny + 1, 0,   nx + 1, ny + 1, nz + 1};
    for (int m = 0; m < 8; m++)
        for (int d = 0; d < 3; d++)
            cs[m][d] = list[8 * d + m];
}

static void bc_corners(const or_params *P, const lattice *L, REAL *f1, const REAL *f2, int opt)
{   /* src/m_boundary_closed_corners.F90:191-252 */
    int nx = P->nx, ny = P->ny, nz = P->nz;
    int cs[8][3];
    or_corner_table(nx, ny, nz, cs);
    for (int m = 0; m < 8; m++) {
        int ic = cs[m][0], jc = cs[m][1], kc = cs[m][2];
        for (int pass = 0; pass < 2; pass++)
            for (int l = 0; l < L->nl; l++) {
                int ox = (ic == 0 && L->cx[l] < 0) || (ic == nx + 1 && L->cx[l] > 0);
                int oy = (jc == 0 && L->cy[l] < 0) || (jc == ny + 1 && L->cy[l] > 0);
                int oz = (kc == 0 && L->cz[l] < 0) || (kc == nz + 1 && L->cz[l] > 0);
                if (!(ox && oy && oz)) continue;
                if (pass == 0) {
                    int mm = find_dir(L, opt * L->cx[l], opt * L->cy[l], opt * L->cz[l]);
                    if (mm >= 0) f1[IDX(P, mm, ic, jc, kc)] = f2[IDX(P, l, ic, jc, kc)];
                } else {
                    f1[IDX(P, l, ic, jc, kc)] = f1[IDX(P, l, ic - L->cx[l], jc - L->cy[l], kc - L->cz[l])];
                }
            }
    }
}

/* Inflow at i=0 / zero-gradient outflow at i=nx+1 (src/m_boundary_i_inflow_kernel.F90:61-111),
 * taper = 1 (src/m_boundarycond.F90:334-335). */
static void bc_inflow(const or_params *P, const lattice *L, REAL *f, const REAL *uvel, REAL udir)
{
    const REAL pi = (REAL)3.1415927410125732;
    REAL fghost[27];
    for (int k = 1; k <= P->nz; k++)
    for (int j = 1; j <= P->ny; j++) {
        int ka = k < 1 ? 1 : (k > P->nz ? P->nz : k);
        REAL uu = uvel[ka - 1] * (REAL)1.0 * (REAL)1.0;
        REAL ang = udir * pi / (REAL)180.0;
        REAL cs = (sizeof(REAL) == 4) ? (REAL)cosf((float)ang) : (REAL)cos((double)ang);
        REAL sn = (sizeof(REAL) == 4) ? (REAL)sinf((float)ang) : (REAL)sin((double)ang);
        for (int l = 0; l < L->nl; l++) {
            REAL wl = L->w[l], cxl = (REAL)L->cx[l], cyl = (REAL)L->cy[l];
            fghost[l] = f[IDX(P, l, 1, j, k)]
                - (REAL)2.0 * wl * P->rho0 * (cxl * uu * cs + cyl * uu * sn) / L->cs2;
        }
        for (int l = 0; l < L->nl; l++) {
            if (L->cx[l] <= 0)
                f[IDX(P, l, 0, j, k)] = fghost[l];
            else {
                int m = find_dir(L, -L->cx[l], -L->cy[l], -L->cz[l]);
                f[IDX(P, l, 0, j, k)] = fghost[m];
            }
        }
        for (int l = 0; l < L->nl; l++)
            f[IDX(P, l, P->nx + 1, j, k)] = f[IDX(P, l, P->nx, j, k)];
    }
}

static void opts(int code, int *o1, int *oN)
{   /* src/m_boundarycond.F90:166-172 */
    switch (code) {
    case 11: *o1 = -1; *oN = -1; break;
    case 12: *o1 = -1; *oN = +1; break;
    case 21: *o1 = +1; *oN = -1; break;
    case 22: *o1 = +1; *oN = +1; break;
    default: *o1 = 0; *oN = 0;
    }
}

/* boundarycond: the fixed pass order of src/m_boundarycond.F90:69-463. */
OR_EXPORT void or_boundarycond(const or_params *P, REAL *f1, REAL *f2, const REAL *uvel, REAL udir)
{
    lattice L; lattice_init(&L, P->nl);
    if (P->ibnd == 0) bc_periodic_i(P, f1);
    if (P->jbnd == 0) bc_periodic_j(P, f1);
    if (P->kbnd == 0) bc_periodic_k(P, f1);
    int o1, oN;
    opts(P->ibnd, &o1, &oN);
    if (o1 != 0) { bc_closed(P, &L, f1, f2, 0, 1, o1); bc_closed(P, &L, f1, f2, 0, P->nx, oN); }
    opts(P->jbnd, &o1, &oN);
    if (o1 != 0) { bc_closed(P, &L, f1, f2, 1, 1, o1); bc_closed(P, &L, f1, f2, 1, P->ny, oN); }
    opts(P->kbnd, &o1, &oN);
    if (o1 != 0) { bc_closed(P, &L, f1, f2, 2, 1, o1); bc_closed(P, &L, f1, f2, 2, P->nz, oN); }
    if (P->jbnd > 10 && P->kbnd > 10)
        bc_edges_jk(P, &L, f1, f2, (P->jbnd == 22 && P->kbnd == 22) ? 1 : -1);
    if (P->ibnd > 10 && P->jbnd > 10)
        bc_edges_ij(P, &L, f1, f2, (P->ibnd == 22 && P->jbnd == 22) ? 1 : -1);
    if (P->ibnd > 10 && P->kbnd > 10)
        bc_edges_ik(P, &L, f1, f2, (P->ibnd == 22 && P->kbnd == 22) ? 1 : -1);
    if (P->ibnd > 10 && P->jbnd > 10 && P->kbnd > 10)
        bc_corners(P, &L, f1, f2, -1);
    if (P->ibnd == 1) bc_inflow(P, &L, f1, uvel, udir);
    if (P->ibnd == 0) bc_periodic_i(P, f1);
    if (P->jbnd == 0) bc_periodic_j(P, f1);
    if (P->kbnd == 0) bc_periodic_k(P, f1);
}

/* ---------------------------------------------------------------------------------- */
/* One pass of the time loop of src/main.F90:283-329 over steps it0+1..it1.
 * State on entry: fA (f1) and fB (f2) as after readrestart + macrovars + fB=fA
 * (src/main.F90:255-258), or as the previous call left them.  rho,u,v,w hold the moments
 * of f1.  On exit the roles have been swapped (it1-it0) times; the function returns 1 if
 * the final f1 is fB, else 0.
 *   uvel_shear[nz], uvel_time[it], udir_time[it] indexed by absolute step it (arrays of
 *   length it1+1); turb (may be NULL): uu,vv,ww as (ny,nz,0:nrturb); F (may be NULL):
 *   external_forcing applied when apply_forcing != 0; blank (may be NULL) enables solids.
 *   pottemp (may be NULL): it1 - it0 potential-temperature fields (0:nx+1,0:ny+1,0:nz+1), the one
 *   of step it0 + 1 + m at offset m; each step's external_forcing is then the turbines' (or zero)
 *   plus the buoyancy of its field (src/main.F90:290-293), applied by forcings_apply (:306). */
OR_EXPORT int or_run(const or_params *P, REAL *fA, REAL *fB, REAL *tau, REAL *rho, REAL *u, REAL *v, REAL *w,
                     const REAL *ablvisc, const uint8_t *blank,
                     const REAL *uvel_shear, const REAL *uvel_time, const REAL *udir_time,
                     const REAL *uu, const REAL *vv, const REAL *ww, int nrturb, REAL ampl,
                     const REAL *F, int apply_forcing, int it0, int it1, const or_turbines *T,
                     const REAL *pottemp, REAL bscale, REAL btheta0)
{
    REAL *f1 = fA, *f2 = fB;
    REAL *Ft = NULL;   /* external_forcing of the producers (src/main.F90:290-293) */
    const size_t ncell = (size_t)(P->nx + 2) * (P->ny + 2) * (P->nz + 2);
    if (T || pottemp) {
        Ft = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)(P->nx + 2) * (P->ny + 2) * (P->nz + 2));
        F = Ft;
        apply_forcing = 1;
    }
    REAL *uvel = (REAL *)malloc(sizeof(REAL) * (size_t)P->nz);
    REAL *df = NULL;
    size_t sl = (size_t)P->ny * (size_t)P->nz;
    if (uu) df = (REAL *)malloc(sizeof(REAL) * sl * (size_t)P->nl);
    int swapped = 0;
    for (int it = it0 + 1; it <= it1; it++) {
        if (T) or_turbine_forcing(P, T, rho, u, v, w, Ft);
        else if (pottemp) memset(Ft, 0, sizeof(REAL) * 3 * ncell);
        if (pottemp) or_buoyancy_forcing(P, Ft, pottemp + ncell * (size_t)(it - it0 - 1), bscale, btheta0);
        if (uu) {
            int lit = it % nrturb;
            if (lit == 0) lit = nrturb;
            or_turbulence_forcing(P, rho, u, v, w, uu + sl * lit, vv + sl * lit, ww + sl * lit, ampl, df);
        }
        or_postcoll(P, f1, tau, rho, u, v, w, ablvisc);
        if (uu) or_turbulence_apply(P, f1, df);
        if (F && apply_forcing) or_forcings_apply(P, f1, F, rho, u, v, w);
        if (blank) or_solids(P, f1, blank);
        for (int k = 0; k < P->nz; k++) uvel[k] = uvel_time[it] * uvel_shear[k];
        or_boundarycond(P, f1, f2, uvel, udir_time[it]);
        or_drift(P, f2, f1);
        REAL *t = f1; f1 = f2; f2 = t;
        swapped ^= 1;
        or_macrovars(P, f1, rho, u, v, w);
    }
    free(uvel);
    free(df);
    free(Ft);
    return swapped;
}
