// ndd.hip -- device-resident implicit NDD time step.
//
// One step = ions, Poisson, electrons, electron energy, each a
// predictor-corrector (PC) loop of up to maxITR_HES_* iterations with one
// 5-point solve per iteration (FluidModule::solve_rhoi_constTe,
// solve_phi_couple_wdTe_wSEE_PC, solve_rhoe_wdTe_wSEE_PC,
// solve_Te_wdTe_wSEE_PC; include/fluidModule.hpp:1405-4957 and the order of
// src/main.cpp:196-231).  The domain is one rectangular block with the
// reference's block-4 boundary set: left/right z walls (phi = 0 / V_bias,
// thermal + secondary-emission wall flux), the axis below (zero face
// coefficient) and an open upper face (grad phi = 0, electron flux = ion
// flux).  Ions use the electron Scharfetter-Gummel scheme at charge +1 with the
// kinetic wall flux of fluidModule.hpp:1502-1526 evaluated at the drift
// velocity mu_i E of the wall face.
//
// Layout: every field is (ni+2) x (nj+2) doubles, row-major [i][j] (j = r
// contiguous), exactly the host layout, so state transfers are flat copies.
// The 5-point systems are written straight into the multigrid's level-0
// arrays (row pitch ld0).
//
// The assembly, prediction and correction kernels evaluate every expression
// in the source order of the restated reference loops and are compiled
// without multiply-add contraction, so they reproduce a sequential FP64
// evaluation up to the last-ulp differences of exp/expm1/erf.
#include <algorithm>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <utility>
#include <memory>
#include <vector>

#include "geometry.hpp"
#include "learned.hpp"
#include "mg.hpp"
#include "ndd_core.hpp"

namespace cathode {


// isotropic mobility evaluated through the tensor formulas of
// fluidModule.hpp:2314-2328 (B = 0)
static Mob mobility(double mu)
{
    const double mu_para = (mu + mu) / 2.0, mu_perp = (mu + mu) / 2.0;
    const double mu_rr = (mu + mu) / 2.0, mu_xr = (0.0 + 0.0) / 2.0;
    Mob m;
    m.mu_star = mu_perp * 1.0;
    m.inv_par = mu_rr / (mu_para * mu_perp);
    m.inv_cross = -mu_xr / (mu_para * mu_perp);
    return m;
}

// Scharfetter-Gummel weights: bp = vd/(1 - exp(-z)) (upstream density),
// bm = vd/(exp(z) - 1) (downstream density), with the exponent clamped at
// lim_exp as in fluidModule.hpp:3239-3243; differences taken with expm1,
// z == 0 returns the limit D/h.
__device__ inline double bern_bp(double vd, double z, double Doh)
{
    if (z == 0.0) return Doh;
    return vd / (-expm1(fmin(-z, K_LIM_EXP)));
}
__device__ inline double bern_bm(double vd, double z, double Doh)
{
    if (z == 0.0) return Doh;
    return vd / expm1(fmin(z, K_LIM_EXP));
}

struct Face {
    double vd, D, z;
};

// face drift / diffusion / Peclet number between an upstream cell (Ta) and a
// downstream cell (Tc) at spacing h with face field E (fluidModule.hpp:3595-3606)
__device__ inline Face sg_face(const NddView& V, int sp, double mu_star, double E, double Ta, double Tc,
                               double h)
{
#pragma clang fp contract(off)
    Face F;
    if (sp == SP_ION) {
        F.vd = mu_star * (E) + K_EPS_VD;
        F.D = K_BOLTZ * V.Ti / K_E0 * mu_star;
    } else {
        const double T_f = (Tc + Ta) / 2.0;
        const double dT = (Tc - Ta) / h;
        F.vd = mu_star * (-E - K_BOLTZ / K_E0 * dT) + K_EPS_VD;
        F.D = K_BOLTZ * T_f / K_E0 * mu_star;
    }
    F.z = F.vd * h / F.D;
    return F;
}

// the face's Scharfetter-Gummel weights bp (upstream) and bm (downstream)
__device__ inline void sg_weights(const NddView& V, int sp, double mu_star, double E, double Ta, double Tc, double h,
                                  double& bp, double& bm)
{
#pragma clang fp contract(off)
    const Face F = sg_face(V, sp, mu_star, E, Ta, Tc, h);
    bp = bern_bp(F.vd, F.z, F.D / h);
    bm = bern_bm(F.vd, F.z, F.D / h);
}

// func_Te, miniFuncs.hpp:35-46
__device__ inline double func_Te(const double* v, int n, double Te, double dTe)
{
#pragma clang fp contract(off)
    const int iL = (int)(Te / dTe);
    const int iR = (int)(Te / dTe) + 1;
    const double TeL = iL * dTe;
    if (iR < n - 1) return (v[iR] - v[iL]) / (dTe + 1e-100) * (Te - TeL) + v[iL];
    return v[n - 1];
}

__device__ inline double eisee_sum(const NddView& V, double Te)
{
#pragma clang fp contract(off)
    if (V.n_see <= 1) return 0.0;
    const double eb = func_Te(V.see_eb, V.n_see, Te, V.dTe_SEE);
    const double rd = func_Te(V.see_rd, V.n_see, Te, V.dTe_SEE);
    const double ts = func_Te(V.see_ts, V.n_see, Te, V.dTe_SEE);
    return eb + rd + ts;
}

// electron thermal wall flux per density (fluidModule.hpp:2350-2352)
__device__ inline double wall_gpn(const NddView& V, double Te, double dphi, double sign)
{
#pragma clang fp contract(off)
    const double vth = sqrt(8.0 * K_BOLTZ * Te / (M_PI * V.masse));
    return sign * 0.25 * vth * exp(fmin(-K_E0 * dphi / (K_BOLTZ * Te + 1e-100), K_LIM_EXP));
}

// ion kinetic wall flux per density (fluidModule.hpp:1502-1526)
__device__ inline double ion_wall_gpn(const NddView& V, double U, double sign)
{
#pragma clang fp contract(off)
    const double vth = sqrt(8.0 * K_BOLTZ * V.Ti / (M_PI * V.massi));
    const double um = sqrt(2.0 * K_BOLTZ * V.Ti / V.massi);
    const double u = U / (um + 1e-100);
    return sign * 0.25 * vth * (exp(-u * u) + u * sqrt(M_PI) * (erf(u) + sign));
}

// ================================================================== Poisson
struct PoissonArgs {
    const double *phi, *rhoe, *rhoi, *Te, *rhoe_old, *Ex, *Er, *Ex_old, *Er_old, *rhoUix, *rhoUir, *S;
    const double *nUex_t, *nUer_t;
    double *nUex, *nUer;
    Mob m;
};

// predicted x flux on face (i, j) (fluidModule.hpp:2266-2462): inner faces
// the drift-diffusion flux plus the mobility-tensor correction, wall faces
// the thermal + emission wall flux of the plasma cell beside the wall
template <class G>
__device__ __forceinline__ double poisson_flux_x(const G& g, const NddView& V, const PoissonArgs& A, int i, int j)
{
#pragma clang fp contract(off)
    const double dx = V.dx;
    const double* r = V.r;
    const int t = g.xf(V, i, j) & GF_TYPE;
    if (t == GF_INNER) {
        const double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
        const double n_f = (A.rhoe[IX(i, j)] + A.rhoe[IX(i - 1, j)]) / 2.0;
        const double dphidx = (A.phi[IX(i, j)] - A.phi[IX(i - 1, j)]) / dx;
        const double dpedx =
            K_BOLTZ * (A.rhoe[IX(i, j)] * A.Te[IX(i, j)] - A.rhoe[IX(i - 1, j)] * A.Te[IX(i - 1, j)]) / dx;
        const double bLr = tmask_r(g.rf(V, i, j)), bLrD = tmask_r(g.rf(V, i - 1, j));
        const double bRr = tmask_r(g.rf(V, i, j + 1)), bRrD = tmask_r(g.rf(V, i - 1, j + 1));
        const double tr = (rR * A.nUer_t[IX(i, j + 1)] * bRr + rL * A.nUer_t[IX(i, j)] * bLr +
                           rR * A.nUer_t[IX(i - 1, j + 1)] * bRrD + rL * A.nUer_t[IX(i - 1, j)] * bLrD) /
                          (4.0 * r[j]);
        return A.nUex_t[IX(i, j)] - (A.m.inv_par * A.nUex_t[IX(i, j)] + A.m.inv_cross * tr) * A.m.mu_star +
                           A.m.mu_star * (n_f * dphidx - dpedx / K_E0);
    } else if (t == GF_WLO) {
        const double Te = A.Te[IX(i, j)];
        const double dphi = fmax(-A.Ex_old[IX(i, j)] * dx / 2.0, 0.0);
        const double gpn = wall_gpn(V, Te, dphi, -1.0);
        const double see = -V.coefIISEE_ts * A.rhoUix[IX(i, j)] - eisee_sum(V, Te) * A.rhoe_old[IX(i, j)] * gpn;
        return A.rhoe[IX(i, j)] * gpn + see;
    } else {
        const int c = i - 1;
        const double Te = A.Te[IX(c, j)];
        const double dphi = fmax(A.Ex_old[IX(i, j)] * dx / 2.0, 0.0);
        const double gpn = wall_gpn(V, Te, dphi, 1.0);
        const double see = -V.coefIISEE_ts * A.rhoUix[IX(i, j)] - eisee_sum(V, Te) * A.rhoe_old[IX(c, j)] * gpn;
        return A.rhoe[IX(c, j)] * gpn + see;
    }
}

// predicted r flux on face (i, j) (fluidModule.hpp:2466-2621): inner faces
// as for x, wall faces the wall flux, the open face the ion flux
template <class G>
__device__ __forceinline__ double poisson_flux_r(const G& g, const NddView& V, const PoissonArgs& A, int i, int j)
{
#pragma clang fp contract(off)
    const double dr = V.dr;
    const double* r = V.r;
    const int t = g.rf(V, i, j) & GF_TYPE;
    if (t == GF_INNER) {
        const double r_t = (r[j] + r[j - 1]) / 2.0;
        const double n_f = (A.rhoe[IX(i, j)] + A.rhoe[IX(i, j - 1)]) / 2.0;
        const double dphidr = (A.phi[IX(i, j)] - A.phi[IX(i, j - 1)]) / dr;
        const double dpedr =
            K_BOLTZ * (A.rhoe[IX(i, j)] * A.Te[IX(i, j)] - A.rhoe[IX(i, j - 1)] * A.Te[IX(i, j - 1)]) / dr;
        const double bLx = tmask_x(g.xf(V, i, j)), bLxD = tmask_x(g.xf(V, i, j - 1));
        const double bRx = tmask_x(g.xf(V, i + 1, j)), bRxD = tmask_x(g.xf(V, i + 1, j - 1));
        const double tx = (r[j] * A.nUex_t[IX(i + 1, j)] * bRx + r[j] * A.nUex_t[IX(i, j)] * bLx +
                           r[j - 1] * A.nUex_t[IX(i + 1, j - 1)] * bRxD + r[j - 1] * A.nUex_t[IX(i, j - 1)] * bLxD) /
                          (4.0 * r_t);
        return A.nUer_t[IX(i, j)] - (A.m.inv_cross * tx + A.m.inv_par * A.nUer_t[IX(i, j)]) * A.m.mu_star +
                           A.m.mu_star * (n_f * dphidr - dpedr / K_E0);
    } else if (t == GF_OPEN) {
        const double uir = V.mu_i * A.Er[IX(i, j)] + K_EPS_VD;
        return A.rhoi[IX(i, j - 1)] * uir;
    } else if (t == GF_WLO) {
        const double Te = A.Te[IX(i, j)];
        const double dphi = fmax(-A.Er_old[IX(i, j)] * dr / 2.0, 0.0);
        const double gpn = wall_gpn(V, Te, dphi, -1.0);
        const double see = -V.coefIISEE_ts * A.rhoUir[IX(i, j)] - eisee_sum(V, Te) * A.rhoe_old[IX(i, j)] * gpn;
        return A.rhoe[IX(i, j)] * gpn + see;
    } else {
        const int c = j - 1;
        const double Te = A.Te[IX(i, c)];
        const double dphi = fmax(A.Er_old[IX(i, j)] * dr / 2.0, 0.0);
        const double gpn = wall_gpn(V, Te, dphi, 1.0);
        const double see = -V.coefIISEE_ts * A.rhoUir[IX(i, j)] - eisee_sum(V, Te) * A.rhoe_old[IX(i, c)] * gpn;
        return A.rhoe[IX(i, c)] * gpn + see;
    }
}


// 5-point Poisson correction system and its BC folds (fluidModule.hpp:2627-2824)
// (the predicted face fluxes are stored after every load of the cell, so the
// loads need not wait for the stores).  The plasma conduction term of a face
// counts only between two plasma cells; Dirichlet walls fold aP += a_nb, the
// open face aP -= a_nb, in the reference's order.
template <class G>
__device__ __forceinline__ Row5 poisson_assemble_cell(const G& g, const NddView& V, const PoissonArgs& A, int i,
                                                      int j)
{
#pragma clang fp contract(off)
    const double dx = V.dx, dr = V.dr, dt = V.dt;
    const double* r = V.r;
    const double qR = (r[j] + r[j + 1]) / 2.0 / r[j];
    const double qL = (r[j] + r[j - 1]) / 2.0 / r[j];
    const double mu = A.m.mu_star;
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, tN = g.rf(V, i, j + 1) & GF_TYPE;
    const double nLx = (A.rhoe[IX(i, j)] + A.rhoe[IX(i - 1, j)]) / 2.0;
    const double nRx = (A.rhoe[IX(i, j)] + A.rhoe[IX(i + 1, j)]) / 2.0;
    const double nLr = (A.rhoe[IX(i, j)] + A.rhoe[IX(i, j - 1)]) / 2.0;
    const double nRr = (A.rhoe[IX(i, j)] + A.rhoe[IX(i, j + 1)]) / 2.0;
    const double bLx = (double)(tL == GF_INNER), bRx = (double)(tR == GF_INNER);
    const double bLr = (double)(tS == GF_INNER), bRr = (double)(tN == GF_INNER);
    const double eRx = K_EPS0 + K_E0 * dt * nRx * mu * bRx;
    const double eLx = K_EPS0 + K_E0 * dt * nLx * mu * bLx;
    const double eRr = K_EPS0 + K_E0 * dt * nRr * mu * bRr;
    const double eLr = K_EPS0 + K_E0 * dt * nLr * mu * bLr;
    double vP = (eRx + eLx) * dr * dr + (qR * eRr + qL * eLr) * dx * dx;
    double vE = eRx * dr * dr;
    double vW = eLx * dr * dr;
    double vN = qR * eRr * dx * dx;
    double vS = qL * eLr * dx * dx;
    // predicted fluxes of the four faces, stored as in the SG assembly; the
    // axis face is not predicted
    const double fxL = poisson_flux_x(g, V, A, i, j), fxR = poisson_flux_x(g, V, A, i + 1, j);
    const double frL = tS != GF_AXIS ? poisson_flux_r(g, V, A, i, j) : A.nUer[IX(i, j)];
    const double frR = poisson_flux_r(g, V, A, i, j + 1);
    const double div = (fxR - fxL) / dx + (qR * frR - qL * frL) / dr;
    const double lap = -(A.Ex[IX(i + 1, j)] - A.Ex[IX(i, j)]) / dx - (qR * A.Er[IX(i, j + 1)] - qL * A.Er[IX(i, j)]) / dr;
    const double S = A.S ? A.S[IX(i, j)] : 0.0;
    const double vb = (K_E0 * A.rhoi[IX(i, j)] - K_E0 * A.rhoe[IX(i, j)] - K_E0 * dt * S + K_EPS0 * lap + K_E0 * dt * div) *
                      dx * dx * dr * dr;
    if (tL == GF_WLO) {
        vP = vP + vW;
        vW = 0.0;
    }
    if (tR == GF_WHI) {
        vP = vP + vE;
        vE = 0.0;
    }
    if (tS == GF_WLO) {
        vP = vP + vS;
        vS = 0.0;
    }
    if (tN == GF_WHI) {
        vP = vP + vN;
        vN = 0.0;
    }
    if (tN == GF_OPEN) {
        vP = vP - vN;
        vN = 0.0;
    }
    A.nUex[IX(i, j)] = fxL;
    if (tR == GF_WHI) A.nUex[IX(i + 1, j)] = fxR;
    if (tS != GF_AXIS) A.nUer[IX(i, j)] = frL;
    if (tN == GF_WHI || tN == GF_OPEN) A.nUer[IX(i, j + 1)] = frR;
    return Row5{vP, vE, vW, vN, vS, vb};
}

// with s_out set, also the multigrid's level-0 preparation (prep_row)
template <class G>
__global__ void k_poisson_assemble(G g, NddView V, PoissonArgs A, double* aP, double* aE, double* aW, double* aN,
                                   double* aS, double* b, double* s_out, double* part)
{
    THREAD_IJ(1, 1);
    double amax = 0.0, bb = 0.0;
    if (i <= V.ni && j <= V.nj) {
        if (g.act(V, i, j)) {
            Row5 R = poisson_assemble_cell(g, V, A, i, j);
            if (s_out) prep_row(g, V, i, j, R, s_out, amax, bb);
            store_row(V, i, j, R, aP, aE, aW, aN, aS, b);
        } else {
            store_idle_row(V, i, j, aP, aE, aW, aN, aS, b, s_out);
        }
    }
    if (s_out) prep_partials(amax, bb, part);
}

// phi ghosts before the PC loop (fluidModule.hpp:2152-2249): every cell of
// the extended grid with a ghost rule takes 2 V - phi (V = 0 or V_bias) or
// phi of its source cell (the same values also into cp, the first pass's
// copy of phi)
template <class G>
__global__ void k_poisson_ghost(G g, NddView V, double* __restrict__ phi, double* __restrict__ cp)
{
#pragma clang fp contract(off)
    THREAD_IJ(0, 0);
    if (i > V.ni + 1 || j > V.nj + 1) return;
    const int q = g.ghost(V, i, j);
    if (q < 0) return;
    const double ps = phi[q >> 2];
    const double v = (q & 2) ? ((q & 1) ? 2.0 * V.V_bias - ps : -ps) : ps;
    phi[IX(i, j)] = v;
    cp[IX(i, j)] = v;
}

// the correction at (i, j) of the extended grid: the solution in plasma
// cells, its odd mirror on Dirichlet ghosts and its copy on open ghosts
template <class G>
__device__ inline double dphi_at(const G& g, const NddView& V, const double* __restrict__ sol, int i, int j)
{
    if (g.act_nb(V, i, j)) return sol[IX0(i, j)];
    const int q = g.ghost(V, i, j);
    if (q < 0) return 0.0;
    const int src = q >> 2;
    const int si = src / V.ld, sj = src - si * V.ld;
    const double d = sol[IX0(si, sj)];
    return (q & 2) ? -d : d;
}

// phi += dphi on the extended grid with the ghost reflections of dphi, the
// wall-face field update (fluidModule.hpp:2850-2958), then the inner face
// fields E = -grad phi and the flux correction mu* n_f grad dphi
// (fluidModule.hpp:2959-3033).  phi0 is phi at the start of the PC pass, so
// every thread can form the updated phi of its neighbours.  The new phi and
// every face flux the assembly predicted also go to the next pass's
// start-of-pass copies nc, nx, nr (their other entries were copied at the
// routine's start and do not change).  Every load is issued before the first
// store.
template <class G>
__global__ void k_poisson_update(G g, NddView V, const double* __restrict__ sol, const double* __restrict__ phi0,
                                 double* __restrict__ dphi, double* __restrict__ phi, double* __restrict__ Ex,
                                 double* __restrict__ Er, const double* __restrict__ rhoe, double* __restrict__ nUex,
                                 double* __restrict__ nUer, Mob m, double* __restrict__ nc, double* __restrict__ nx,
                                 double* __restrict__ nr)
{
#pragma clang fp contract(off)
    THREAD_IJ(0, 0);
    if (i > V.ni + 1 || j > V.nj + 1) return;
    const bool cell = g.act_nb(V, i, j);
    const double d = dphi_at(g, V, sol, i, j);
    const double p = phi0[IX(i, j)] + d;
    int tL = GF_NONE, tR = GF_NONE, tS = GF_NONE, tN = GF_NONE;
    double dl = 0.0, pl = 0.0, ds = 0.0, ps = 0.0, rc = 0.0, rl = 0.0, rs = 0.0;
    double ux = 0.0, ur = 0.0, uxE = 0.0, urN = 0.0, ex1 = 0.0, exN = 0.0, er1 = 0.0, erN = 0.0;
    if (cell) {
        tL = g.xf(V, i, j) & GF_TYPE;
        tR = g.xf(V, i + 1, j) & GF_TYPE;
        tS = g.rf(V, i, j) & GF_TYPE;
        tN = g.rf(V, i, j + 1) & GF_TYPE;
        ux = nUex[IX(i, j)];
        ur = nUer[IX(i, j)];
        rc = rhoe[IX(i, j)];
        if (tL == GF_INNER) {
            dl = dphi_at(g, V, sol, i - 1, j);
            pl = phi0[IX(i - 1, j)] + dl;
            rl = rhoe[IX(i - 1, j)];
        }
        if (tS == GF_INNER) {
            ds = dphi_at(g, V, sol, i, j - 1);
            ps = phi0[IX(i, j - 1)] + ds;
            rs = rhoe[IX(i, j - 1)];
        }
        if (tL == GF_WLO) ex1 = Ex[IX(i, j)];
        if (tR == GF_WHI) {
            exN = Ex[IX(i + 1, j)];
            uxE = nUex[IX(i + 1, j)];
        }
        if (tS == GF_WLO) er1 = Er[IX(i, j)];
        if (tN == GF_WHI) erN = Er[IX(i, j + 1)];
        if (tN == GF_WHI || tN == GF_OPEN) urN = nUer[IX(i, j + 1)];
    }
    dphi[IX(i, j)] = d;
    phi[IX(i, j)] = p;
    nc[IX(i, j)] = p;
    if (!cell) return;
    // wall faces: the ghost across them holds -d
    if (tL == GF_WLO) Ex[IX(i, j)] = ex1 - (d - (-d)) / V.dx;
    if (tR == GF_WHI) Ex[IX(i + 1, j)] = exN - ((-d) - d) / V.dx;
    if (tS == GF_WLO) Er[IX(i, j)] = er1 - (d - (-d)) / V.dr;
    if (tN == GF_WHI) Er[IX(i, j + 1)] = erN - ((-d) - d) / V.dr;
    if (tN == GF_OPEN) Er[IX(i, j + 1)] = 0.0;
    if (tL == GF_INNER) {
        Ex[IX(i, j)] = -(p - pl) / V.dx;
        const double n_f = (rc + rl) / 2.0;
        const double u = ux + m.mu_star * n_f * ((d - dl) / V.dx);
        nUex[IX(i, j)] = u;
        nx[IX(i, j)] = u;
    } else {
        nx[IX(i, j)] = ux;
    }
    if (tR == GF_WHI) nx[IX(i + 1, j)] = uxE;
    if (tS == GF_INNER) {
        Er[IX(i, j)] = -(p - ps) / V.dr;
        const double n_f = (rc + rs) / 2.0;
        const double u = ur + m.mu_star * n_f * ((d - ds) / V.dr);
        nUer[IX(i, j)] = u;
        nr[IX(i, j)] = u;
    } else {
        nr[IX(i, j)] = ur;
    }
    if (tN == GF_WHI || tN == GF_OPEN) nr[IX(i, j + 1)] = urN;
}

// ================================================================== SG equations
struct SgArgs {
    int sp;
    Mob m;
    double tfac;
    const double *n_old, *T;  // start-of-routine density, temperature of the SG faces
    const double *Ex, *Er, *Te, *rhoi, *rhoe, *rhoe_old, *rhoUix, *rhoUir, *rhoUex, *rhoUer;
    const double *S, *Pabs, *eloss;
    const double *Fx_t, *Fr_t;
    double *n, *Fx, *Fr;
    // the SG weights bp, bm of each cell's x face (i, j) and r face (i, j),
    // planes of wn doubles in the multigrid's row pitch: written by the
    // assembly, read by the correction of the same PC pass
    double* wgt;
    size_t wn;
    int reuse;  // the weights are already in wgt (a later PC pass of the routine)
};

// wall flux of species sp through a wall face whose plasma cell is c = (ci,
// cj), face field E and ion flux Gi at the face, sgn -1 for a wall on the
// cell's low side, +1 on its high side (fluidModule.hpp:3250-3367 and
// 3453-3532, energy :4161-4316 and :4403-4506; ions: the kinetic flux of
// :1502-1526 at the drift mu_i E); h is the spacing across the face.  The
// electrons' sheath barrier is the cell potential above the wall's,
// phi_cell - phi_wall = sgn E h / 2 (fluidModule.hpp:3255, 3335).
__device__ __forceinline__ double sg_wall_flux(const NddView& V, const SgArgs& A, int ci, int cj, double E, double Gi,
                                               double sgn, double h)
{
#pragma clang fp contract(off)
    const double* T = A.T;
    if (A.sp == SP_ION) {
        const double U = V.mu_i * E;
        return A.n[IX(ci, cj)] * ion_wall_gpn(V, U, sgn);
    } else if (A.sp == SP_ELEC) {
        const double Te = A.Te[IX(ci, cj)];
        const double dphi = fmax(sgn * E * h / 2.0, 0.0);
        const double gpn = wall_gpn(V, Te, dphi, sgn);
        const double see = -V.coefIISEE_ts * Gi - eisee_sum(V, Te) * A.rhoe_old[IX(ci, cj)] * gpn;
        return A.n[IX(ci, cj)] * gpn + see;
    } else {
        const double Teo = T[IX(ci, cj)];
        const double dphi = fmax(sgn * E * h / 2.0, 0.0);
        const double vth = sqrt(8.0 * K_BOLTZ * Teo / (M_PI * V.masse));
        const double ex = exp(fmin(-K_E0 * dphi / (K_BOLTZ * Teo + 1e-100), K_LIM_EXP));
        const double gpne = sgn * vth * ex / 3.0 * 1.0;
        const double eps_emit = 2.0 * K_E0 * V.Te_emitSEE_eV;
        const double gpn = sgn * 0.25 * 1.0 * vth * ex;
        const double Te = A.Te[IX(ci, cj)];
        double eb = 0.0, rd = 0.0, ts = 0.0;
        if (V.n_see > 1) {
            eb = func_Te(V.see_eb, V.n_see, Te, V.dTe_SEE);
            rd = func_Te(V.see_rd, V.n_see, Te, V.dTe_SEE);
            ts = func_Te(V.see_ts, V.n_see, Te, V.dTe_SEE);
        }
        const double see = -eps_emit * V.coefIISEE_ts * Gi - (eb + V.ratioEngy_EISEE_rd * rd) * A.n_old[IX(ci, cj)] * gpne -
                           eps_emit * ts * A.rhoe_old[IX(ci, cj)] * gpn;
        return A.n[IX(ci, cj)] * gpne + see;
    }
}

// predicted SG flux on x face (i, j): the Scharfetter-Gummel flux plus the
// mobility-tensor correction of the previous fluxes on inner faces, the wall
// flux on wall faces (fluidModule.hpp:3239-3367, energy :4161-4316).  bp, bm:
// the face's weights bern_bp / bern_bm (inner faces).
template <class G>
__device__ __forceinline__ double sg_flux_x(const G& g, const NddView& V, const SgArgs& A, int i, int j, double bp,
                                            double bm)
{
#pragma clang fp contract(off)
    const double* r = V.r;
    const int t = g.xf(V, i, j) & GF_TYPE;
    if (t == GF_INNER) {
        const double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
        const double bLr = tmask_r(g.rf(V, i, j)), bLrD = tmask_r(g.rf(V, i - 1, j));
        const double bRr = tmask_r(g.rf(V, i, j + 1)), bRrD = tmask_r(g.rf(V, i - 1, j + 1));
        const double tr = (rR * A.Fr_t[IX(i, j + 1)] * bRr + rL * A.Fr_t[IX(i, j)] * bLr +
                           rR * A.Fr_t[IX(i - 1, j + 1)] * bRrD + rL * A.Fr_t[IX(i - 1, j)] * bLrD) /
                          (4.0 * r[j]);
        double sg = bp * A.n[IX(i - 1, j)] - bm * A.n[IX(i, j)];
        if (A.sp == SP_ENERGY) sg = sg * 5.0 / 3.0;
        return sg + A.Fx_t[IX(i, j)] - A.m.mu_star * (A.m.inv_par * A.Fx_t[IX(i, j)] + A.m.inv_cross * tr);
    }
    const bool high = t == GF_WHI;  // the plasma cell is i - 1
    return sg_wall_flux(V, A, high ? i - 1 : i, j, A.Ex[IX(i, j)], A.rhoUix[IX(i, j)], high ? 1.0 : -1.0, V.dx);
}

// predicted SG flux on r face (i, j): inner faces as for x, wall faces the
// wall flux, the open face the ion flux (electrons, ions) or the convected
// energy 5/2 k T max(Gamma_e, 0) (fluidModule.hpp:3443-3542, energy
// :4403-4506).  bp, bm as for sg_flux_x.
template <class G>
__device__ __forceinline__ double sg_flux_r(const G& g, const NddView& V, const SgArgs& A, int i, int j, double bp,
                                            double bm)
{
#pragma clang fp contract(off)
    const double* r = V.r;
    const double* T = A.T;
    const int t = g.rf(V, i, j) & GF_TYPE;
    if (t == GF_INNER) {
        const double r_t = (r[j] + r[j - 1]) / 2.0;
        const double bLx = tmask_x(g.xf(V, i, j)), bLxD = tmask_x(g.xf(V, i, j - 1));
        const double bRx = tmask_x(g.xf(V, i + 1, j)), bRxD = tmask_x(g.xf(V, i + 1, j - 1));
        const double tx = (r[j] * A.Fx_t[IX(i + 1, j)] * bRx + r[j] * A.Fx_t[IX(i, j)] * bLx +
                           r[j - 1] * A.Fx_t[IX(i + 1, j - 1)] * bRxD + r[j - 1] * A.Fx_t[IX(i, j - 1)] * bLxD) /
                          (4.0 * r_t);
        double sg = bp * A.n[IX(i, j - 1)] - bm * A.n[IX(i, j)];
        if (A.sp == SP_ENERGY) sg = sg * 5.0 / 3.0;
        return sg + A.Fr_t[IX(i, j)] - A.m.mu_star * (A.m.inv_cross * tx + A.m.inv_par * A.Fr_t[IX(i, j)]);
    } else if (t == GF_OPEN) {
        if (A.sp == SP_ENERGY) return 5.0 / 2.0 * K_BOLTZ * T[IX(i, j - 1)] * fmax(A.rhoUer[IX(i, j)], 0.0);
        const double uir = V.mu_i * A.Er[IX(i, j)] + K_EPS_VD;
        return A.rhoi[IX(i, j - 1)] * uir;
    }
    const bool high = t == GF_WHI;  // the plasma cell is j - 1
    return sg_wall_flux(V, A, i, high ? j - 1 : j, A.Er[IX(i, j)], A.rhoUir[IX(i, j)], high ? 1.0 : -1.0, V.dr);
}


// the weights of each face are formed once, for the coefficients and the
// predicted flux; the fluxes are stored after every load of the cell
template <class G>
__device__ __forceinline__ Row5 sg_assemble_cell(const G& g, const NddView& V, const SgArgs& A, int i, int j)
{
#pragma clang fp contract(off)
    const double dx = V.dx, dr = V.dr, dt = V.dt;
    const double* r = V.r;
    const double* T = A.T;
    const double rR = (r[j] + r[j + 1]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
    const double qR = (r[j] + r[j + 1]) / 2.0 / r[j];
    const double qL = (r[j] + r[j - 1]) / 2.0 / r[j];
    const double mu = A.m.mu_star;
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, tN = g.rf(V, i, j + 1) & GF_TYPE;
    const double bLx = (double)(tL == GF_INNER), bRx = (double)(tR == GF_INNER);
    const double bLr = (double)(tS == GF_INNER), bRr = (double)(tN == GF_INNER);
    // the weights of the four faces: formed here on the routine's first PC
    // pass, read back on later passes (the fields and temperatures they
    // depend on do not change within the routine); the east and north faces
    // of a cell whose neighbour there is not plasma belong to no cell's store
    const size_t k = IX0(i, j), wn = A.wn;
    double bpLx, bmLx, bpRx, bmRx, bpLr, bmLr, bpRr, bmRr;
    if (A.reuse) {
        bpLx = A.wgt[k];
        bmLx = A.wgt[wn + k];
        bpLr = A.wgt[2 * wn + k];
        bmLr = A.wgt[3 * wn + k];
    } else {
        sg_weights(V, A.sp, mu, A.Ex[IX(i, j)], T[IX(i - 1, j)], T[IX(i, j)], dx, bpLx, bmLx);
        sg_weights(V, A.sp, mu, A.Er[IX(i, j)], T[IX(i, j - 1)], T[IX(i, j)], dr, bpLr, bmLr);
    }
    if (A.reuse && g.act_nb(V, i + 1, j)) {
        bpRx = A.wgt[k + V.ld0];
        bmRx = A.wgt[wn + k + V.ld0];
    } else {
        sg_weights(V, A.sp, mu, A.Ex[IX(i + 1, j)], T[IX(i, j)], T[IX(i + 1, j)], dx, bpRx, bmRx);
    }
    if (A.reuse && g.act_nb(V, i, j + 1)) {
        bpRr = A.wgt[2 * wn + k + 1];
        bmRr = A.wgt[3 * wn + k + 1];
    } else {
        sg_weights(V, A.sp, mu, A.Er[IX(i, j + 1)], T[IX(i, j)], T[IX(i, j + 1)], dr, bpRr, bmRr);
    }
    const double vP = bpRx * bRx * dr + bmLx * bLx * dr + qR * bpRr * bRr * dx + qL * bmLr * bLr * dx;
    Row5 R;
    R.P = vP + A.tfac * dx * dr / dt;
    R.E = bmRx * bRx * dr;
    R.W = bpLx * bLx * dr;
    R.N = qR * bmRr * bRr * dx;
    R.S = qL * bpLr * bLr * dx;
    // the predicted fluxes of the cell's four faces; the cell stores its west
    // and south faces (east and north too where they are boundaries), the
    // axis face is not predicted
    const double fxL = sg_flux_x(g, V, A, i, j, bpLx, bmLx), fxR = sg_flux_x(g, V, A, i + 1, j, bpRx, bmRx);
    const double frL = tS != GF_AXIS ? sg_flux_r(g, V, A, i, j, bpLr, bmLr) : A.Fr[IX(i, j)];
    const double frR = sg_flux_r(g, V, A, i, j + 1, bpRr, bmRr);
    const double div = (fxR - fxL) / dx + (qR * frR - qL * frL) / dr;
    if (A.sp == SP_ENERGY) {
        // J.E with the wall faces left out (fluidModule.hpp:4624-4625)
        const double bLxw = (double)!is_wall(tL), bRxw = (double)!is_wall(tR);
        const double bLrw = (double)!is_wall(tS), bRrw = (double)!is_wall(tN);
        const double JEx = -K_E0 * (A.rhoUex[IX(i + 1, j)] * A.Ex[IX(i + 1, j)] * bRxw + A.rhoUex[IX(i, j)] * A.Ex[IX(i, j)] * bLxw) / 2.0;
        const double JEr = -K_E0 * (rR * A.rhoUer[IX(i, j + 1)] * A.Er[IX(i, j + 1)] * bRrw + rL * A.rhoUer[IX(i, j)] * A.Er[IX(i, j)] * bLrw) /
                           (2.0 * r[j]);
        const double R1 = (-A.n[IX(i, j)] + A.n_old[IX(i, j)]) / dt;
        const double R2 = JEx + JEr;
        const double Pabs = A.Pabs ? A.Pabs[IX(i, j)] : 0.0;
        const double eloss = A.eloss ? A.eloss[IX(i, j)] : 0.0;
        const double R3 = Pabs - A.rhoe[IX(i, j)] * eloss;
        const double R4 = -div;
        R.b = 3.0 / 5.0 * (R1 + R2 + R3 + R4) * dx * dr;
    } else {
        const double S = A.S ? A.S[IX(i, j)] : 0.0;
        R.b = (-div + S - (A.n[IX(i, j)] - A.n_old[IX(i, j)]) / dt) * dx * dr;
    }
    A.Fx[IX(i, j)] = fxL;
    if (tR == GF_WHI) A.Fx[IX(i + 1, j)] = fxR;
    if (tS != GF_AXIS) A.Fr[IX(i, j)] = frL;
    if (tN == GF_WHI || tN == GF_OPEN) A.Fr[IX(i, j + 1)] = frR;
    if (!A.reuse) {
        A.wgt[k] = bpLx;
        A.wgt[wn + k] = bmLx;
        A.wgt[2 * wn + k] = bpLr;
        A.wgt[3 * wn + k] = bmLr;
    }
    return R;
}

// with s_out set, also the multigrid's level-0 preparation (prep_row)
template <class G>
__global__ void k_sg_assemble(G g, NddView V, SgArgs A, double* aP, double* aE, double* aW, double* aN, double* aS,
                              double* b, double* s_out, double* part)
{
    THREAD_IJ(1, 1);
    double amax = 0.0, bb = 0.0;
    if (i <= V.ni && j <= V.nj) {
        if (g.act(V, i, j)) {
            Row5 R = sg_assemble_cell(g, V, A, i, j);
            if (s_out) prep_row(g, V, i, j, R, s_out, amax, bb);
            store_row(V, i, j, R, aP, aE, aW, aN, aS, b);
        } else {
            store_idle_row(V, i, j, aP, aE, aW, aN, aS, b, s_out);
        }
    }
    if (s_out) prep_partials(amax, bb, part);
}

// flux correction on inner faces and the positivity-guarded density
// update (fluidModule.hpp:3682-3757); thread (i, j) owns x-face (i, j),
// r-face (i, j) and cell (i, j).  The faces' SG weights are the ones the
// assembly of this PC pass formed (same fields, same temperatures).  The
// corrected density and every face flux the assembly predicted also go to
// the next pass's start-of-pass copies nc, nx, nr (their other entries were
// copied at the routine's first pass and do not change).
template <class G>
__global__ void k_sg_correct(G g, NddView V, SgArgs A, const double* __restrict__ dn, double* __restrict__ nc,
                             double* __restrict__ nx, double* __restrict__ nr)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj) return;
    if (!g.act(V, i, j)) return;
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, tN = g.rf(V, i, j + 1) & GF_TYPE;
    const size_t k = IX0(i, j), wn = A.wn;
    // every load before the first store
    const double fx = A.Fx[IX(i, j)], fr = A.Fr[IX(i, j)], n0 = A.n[IX(i, j)];
    const double fxE = tR == GF_WHI ? A.Fx[IX(i + 1, j)] : 0.0;
    const double frN = (tN == GF_WHI || tN == GF_OPEN) ? A.Fr[IX(i, j + 1)] : 0.0;
    const double dc = dn[k], dW = dn[IX0(i - 1, j)], dS = dn[IX0(i, j - 1)];
    const double bpx = A.wgt[k], bmx = A.wgt[wn + k], bpr = A.wgt[2 * wn + k], bmr = A.wgt[3 * wn + k];
    double fx1 = fx, fr1 = fr;
    if (tL == GF_INNER) {
        double d = bpx * dW - bmx * dc;
        if (A.sp == SP_ENERGY) d = d * 5.0 / 3.0;
        fx1 = fx + d;
        A.Fx[IX(i, j)] = fx1;
    }
    if (tS == GF_INNER) {
        double d = bpr * dS - bmr * dc;
        if (A.sp == SP_ENERGY) d = d * 5.0 / 3.0;
        fr1 = fr + d;
        A.Fr[IX(i, j)] = fr1;
    }
    const double nn = n0 + dc;
    if (nn > 0.0) A.n[IX(i, j)] = nn;
    nx[IX(i, j)] = fx1;
    nr[IX(i, j)] = fr1;
    nc[IX(i, j)] = nn > 0.0 ? nn : n0;
    if (tR == GF_WHI) nx[IX(i + 1, j)] = fxE;
    if (tN == GF_WHI || tN == GF_OPEN) nr[IX(i, j + 1)] = frN;
}

// ================================================================== reductions
// calcRes (miniFuncs.hpp:292-375) of three (new, old) pairs over the plasma
// cells, the inner x faces and the inner r faces (the reference's cell and
// face boxes); six partial values per block.

template <class G>
__global__ void k_calcres(G g, NddView V, ResPairs P, int mode, double eps, double* __restrict__ part, int nb)
{
    __shared__ double lds[6 * NB_X * NB_Y / 64];
    THREAD_IJ(1, 1);
    double a[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
    if (i <= V.ni && j <= V.nj) {
        // all six loads first (the face boxes start one row or column later;
        // the first face's value is loaded and not used)
        const double vnq[3] = {P.n0[IX(i, j)], P.n1[IX(i, j)], P.n2[IX(i, j)]};
        const double voq[3] = {P.o0[IX(i, j)], P.o1[IX(i, j)], P.o2[IX(i, j)]};
        const bool in[3] = {g.act(V, i, j), (g.xf(V, i, j) & GF_TYPE) == GF_INNER,
                            (g.rf(V, i, j) & GF_TYPE) == GF_INNER};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (!in[q]) continue;
            const double vn = vnq[q], vo = voq[q];
            const double d = vn - vo;
            if (mode == 0) {
                a[q] = fabs(d);
            } else if (mode == 1) {
                a[q] = (d * d) / (vo * vo + eps);
                c[q] = 1.0;
            } else if (mode == 2) {
                a[q] = fabs(d);
                c[q] = fabs(vo);
            } else {
                a[q] = d * d;
                c[q] = vo * vo;
            }
        }
    }
    const bool mx_a = mode == 0 || mode == 2, mx_c = mode == 2;
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    double v[6] = {a[0], c[0], a[1], c[1], a[2], c[2]};
    const bool mx[6] = {mx_a, mx_c, mx_a, mx_c, mx_a, mx_c};
    block_reduce_vec<NB_X * NB_Y, 6>(v, mx, lds);
    if (lin_tid() == 0)
        for (int k = 0; k < 6; k++) part[k * nb + bid] = v[k];
}

// errors e[0..2] and their max e[3] (also into side when set)
__global__ void k_calcres_finish(const double* __restrict__ part, int nb, int mode, double* __restrict__ e,
                                 double* __restrict__ side)
{
    __shared__ double lds[6 * 256 / 64];
    const bool mx_a = mode == 0 || mode == 2, mx_c = mode == 2;
    double v[6];
    const bool mx[6] = {mx_a, mx_c, mx_a, mx_c, mx_a, mx_c};
#pragma unroll
    for (int m = 0; m < 6; m++) v[m] = 0.0;
    // each thread's partials in the order k = t, t + 256, ...; the loads of
    // two rounds are issued before they are combined
    for (int k = threadIdx.x; k < nb; k += 2 * 256) {
        const bool two = k + 256 < nb;
        double p0[6], p1[6];
#pragma unroll
        for (int m = 0; m < 6; m++) {
            p0[m] = part[m * nb + k];
            p1[m] = two ? part[m * nb + k + 256] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < 6; m++) {
            v[m] = mx[m] ? fmax(v[m], p0[m]) : v[m] + p0[m];
            if (two) v[m] = mx[m] ? fmax(v[m], p1[m]) : v[m] + p1[m];
        }
    }
    block_reduce_vec<256, 6>(v, mx, lds);
    double err[3];
    for (int q = 0; q < 3; q++) {
        const double sa = v[2 * q], sc = v[2 * q + 1];
        if (mode == 0)
            err[q] = sa;
        else if (mode == 1)
            err[q] = sqrt(sa / sc);
        else if (mode == 2)
            err[q] = sa / sc;
        else
            err[q] = sqrt(sa / sc);
    }
    if (threadIdx.x == 0) {
        e[0] = err[0];
        e[1] = err[1];
        e[2] = err[2];
        e[3] = fmax(fmax(err[0], err[1]), err[2]);
        if (side) side[0] = e[3];
    }
}

// energy post-processing (fluidModule.hpp:4868-4905)
template <class G>
__global__ void k_energy_guard_max(G g, NddView V, double* __restrict__ rhoeps, const double* __restrict__ rhoeps_old,
                                   const double* __restrict__ rhoe, double* __restrict__ part)
{
    __shared__ double lds[8];
    THREAD_IJ(1, 1);
    double v = 0.0;
    if (i <= V.ni && j <= V.nj && g.act(V, i, j)) {
        if (rhoeps[IX(i, j)] < 0.0) rhoeps[IX(i, j)] = fmax(rhoeps_old[IX(i, j)], 0.0);
        v = rhoe[IX(i, j)];
    }
    const double m = block_max<NB_X * NB_Y>(v, lds);
    if (lin_tid() == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = m;
}

template <class G>
__global__ void k_te_update(G g, NddView V, const double* __restrict__ rhoeps, const double* __restrict__ rhoe,
                            const double* __restrict__ part, int nb, double* __restrict__ Te)
{
#pragma clang fp contract(off)
    __shared__ double lds[NB_X * NB_Y / 64];
    THREAD_IJ(1, 1);
    // max n_e from the guard kernel's block partials (every workgroup gets the
    // same value: a maximum does not depend on the order)
    double mx = 0.0;
    for (int k = lin_tid(); k < nb; k += NB_X * NB_Y) mx = fmax(mx, part[k]);
    const double nmax = block_max_bcast<NB_X * NB_Y>(mx, lds);
    if (i > V.ni || j > V.nj || !g.act(V, i, j)) return;
    const double Te_max = 1000.0 * K_E0 / K_BOLTZ, ratio = 0.001;
    const double t = fmin(rhoeps[IX(i, j)] / (3.0 / 2.0 * (rhoe[IX(i, j)] + nmax * ratio) * K_BOLTZ + 1e-100), Te_max);
    Te[IX(i, j)] = t;
    // the ghosts of the geometry's Te list copy their cell (the rectangle:
    // the whole ghost ring, no corners)
    const int w = g.te_bits(V, i, j);
    if (w & GW_XLO) Te[IX(i - 1, j)] = (w & (GW_XLO << 4)) ? t : 0.0;
    if (w & GW_XHI) Te[IX(i + 1, j)] = (w & (GW_XHI << 4)) ? t : 0.0;
    if (w & GW_RLO) Te[IX(i, j - 1)] = (w & (GW_RLO << 4)) ? t : 0.0;
    if (w & GW_RHI) Te[IX(i, j + 1)] = (w & (GW_RHI << 4)) ? t : 0.0;
}

// the electron-density ghosts of a block geometry after the electron routine
// (fluidModule.hpp:3873-3913): 0 behind walls, a copy across open sides
template <class G>
__global__ void k_ne_ghost(G g, NddView V, double* __restrict__ rhoe)
{
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj || !g.act(V, i, j)) return;
    const int w = g.ne_bits(V, i, j);
    if (!w) return;
    const double t = rhoe[IX(i, j)];
    if (w & GW_XLO) rhoe[IX(i - 1, j)] = (w & (GW_XLO << 4)) ? t : 0.0;
    if (w & GW_XHI) rhoe[IX(i + 1, j)] = (w & (GW_XHI << 4)) ? t : 0.0;
    if (w & GW_RLO) rhoe[IX(i, j - 1)] = (w & (GW_RLO << 4)) ? t : 0.0;
    if (w & GW_RHI) rhoe[IX(i, j + 1)] = (w & (GW_RHI << 4)) ? t : 0.0;
}

// the ion density across the open faces of a block geometry (the rectangle
// uses k_top_copy)
template <class G>
__global__ void k_open_copy(G g, NddView V, double* __restrict__ a)
{
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj || !g.act(V, i, j)) return;
    if ((g.rf(V, i, j + 1) & GF_TYPE) == GF_OPEN) a[IX(i, j + 1)] = a[IX(i, j)];
}

// the multigrid's level-0 activity from the plasma table (pitches ld, ld0)
__global__ void k_act_from_table(NddView V, const uint8_t* __restrict__ ac, unsigned char* __restrict__ act0, int n0)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n0) return;
    const int i = k / V.ld0, j = k - i * V.ld0;
    act0[k] = (i <= V.ni + 1 && j <= V.nj + 1) ? ac[IX(i, j)] : 0;
}

__global__ void k_top_copy(NddView V, double* __restrict__ a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= V.ni) a[IX(i, V.nj + 1)] = a[IX(i, V.nj)];
}


// right-electrode current (fluidModule.hpp:3942-3965): out = {Ii, Ie}
__global__ void k_anode(NddView V, AnodeSegs E, const double* __restrict__ rhoUix, const double* __restrict__ rhoUex,
                        double* __restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double lds[8];
    double si = 0.0, se = 0.0;
    for (int q = 0; q < E.n; q++) {
        const int fi = E.fi[q];
        for (int j = E.j0[q] + threadIdx.x; j <= E.j1[q]; j += blockDim.x) {
            si += rhoUix[IX(fi, j)] * 2.0 * M_PI * V.r[j] * V.dr * (K_E0);
            se += rhoUex[IX(fi, j)] * 2.0 * M_PI * V.r[j] * V.dr * (-K_E0);
        }
    }
    const double ti = block_sum<256>(si, lds);
    const double te = block_sum<256>(se, lds);
    if (threadIdx.x == 0) {
        out[0] = ti;
        out[1] = te;
    }
}

// check_CFL (fluidModule.hpp:57-115), first stage: per block of the
// extended grid the maxima of n_e and Te over every cell (ghosts included)
// and of the electron speed |Gamma_e| / n_e over the interior cells
// (x flux averaged over the cell's two faces, r flux r-weighted)
__global__ void k_cfl_partials(NddView V, const double* __restrict__ rhoe, const double* __restrict__ Te,
                               const double* __restrict__ Fx, const double* __restrict__ Fr, double* __restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double lds[NB_X * NB_Y / 64];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    double n = 0.0, t = 0.0, u = 0.0;
    if (i <= V.ni + 1 && j <= V.nj + 1) {
        n = rhoe[IX(i, j)];
        t = Te[IX(i, j)];
        if (i >= 1 && i <= V.ni && j >= 1 && j <= V.nj) {
            const double rL = (V.r[j] + V.r[j - 1]) / 2.0, rR = (V.r[j] + V.r[j + 1]) / 2.0;
            const double ux = (Fx[IX(i + 1, j)] + Fx[IX(i, j)]) / 2.0 / (n + 1e-100);
            const double ur = (rR * Fr[IX(i, j + 1)] + rL * Fr[IX(i, j)]) / (2.0 * V.r[j]) / (n + 1e-100);
            u = sqrt(ux * ux + ur * ur);
        }
    }
    // the maxima start at 0, as the reference's loops do
    n = block_max<NB_X * NB_Y>(fmax(n, 0.0), lds);
    t = block_max<NB_X * NB_Y>(fmax(t, 0.0), lds);
    u = block_max<NB_X * NB_Y>(u, lds);
    if (lin_tid() == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        part[3 * b] = n;
        part[3 * b + 1] = t;
        part[3 * b + 2] = u;
    }
}

// second stage: out = {dt, CFL, dt_min, max n_e, max Te, max u_e + vth(max Te)};
// adaptive: dt = CFL dt_min, else CFL = dt / dt_min (dt_min = eps0 / (e n mu))
__global__ void k_cfl_finish(const double* __restrict__ part, int nb, double mu_max, double masse, double dt,
                             double CFL, int adaptive, double* __restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double lds[4];
    double n = 0.0, t = 0.0, u = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        n = fmax(n, part[3 * b]);
        t = fmax(t, part[3 * b + 1]);
        u = fmax(u, part[3 * b + 2]);
    }
    n = block_max<256>(n, lds);
    t = block_max<256>(t, lds);
    u = block_max<256>(u, lds);
    if (threadIdx.x == 0) {
        const double vth = sqrt(8.0 * K_BOLTZ * t / (M_PI * masse));
        const double dt_min = K_EPS0 / (K_E0 * n * mu_max);
        out[0] = adaptive ? dt_min * CFL : dt;
        out[1] = adaptive ? CFL : dt / dt_min;
        out[2] = dt_min;
        out[3] = n;
        out[4] = t;
        out[5] = u + vth;
    }
}

__global__ void k_fill_one(int n, unsigned char* a, int ni, int nj, int ld)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = k / ld, j = k % ld;
    a[k] = (i >= 1 && i <= ni && j >= 1 && j <= nj) ? 1 : 0;
}

// ================================================================== host side
// ------------------------------------------------------------------ learned initial guess
// Sum of squares over interior cells of two planes (pitches la, lb; b may be
// NULL): partials in part[bid], part[nb + bid]
__global__ void k_learn_sumsq(NddView V, const double* __restrict__ a, int la, const double* __restrict__ b, int lb,
                              double* __restrict__ part, int nb)
{
    __shared__ double lds[8];
    THREAD_IJ(1, 1);
    double sa = 0.0, sb = 0.0;
    if (i <= V.ni && j <= V.nj) {
        const double va = a[(size_t)i * la + j];
        sa = va * va;
        if (b) {
            const double vb = b[(size_t)i * lb + j];
            sb = vb * vb;
        }
    }
    const double ta = block_sum<NB_X * NB_Y>(sa, lds);
    const double tb = block_sum<NB_X * NB_Y>(sb, lds);
    if (lin_tid() == 0) {
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = ta;
        part[nb + bid] = tb;
    }
}

// out[0], out[1] = root mean square of the two planes
__global__ void k_learn_rms(const double* __restrict__ part, int nb, double ncell, double* __restrict__ out)
{
    __shared__ double lds[8];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        a += part[k];
        b += part[nb + k];
    }
    a = block_sum<256>(a, lds);
    b = block_sum<256>(b, lds);
    if (threadIdx.x == 0) {
        out[0] = sqrt(a / ncell);
        out[1] = sqrt(b / ncell);
    }
}

// feature planes (pitch V.ld): b / rms(b), r / r_top, x_prev / rms(x_prev)
__global__ void k_learn_features(NddView V, const double* __restrict__ b, const double* __restrict__ xprev, int lx,
                                 const double* __restrict__ rms, double r_top, double* __restrict__ feat)
{
    THREAD_IJ(0, 0);
    if (i > V.ni + 1 || j > V.nj + 1) return;
    const size_t plane = (size_t)(V.ni + 2) * V.ld, p = IX(i, j);
    const bool in = i >= 1 && i <= V.ni && j >= 1 && j <= V.nj;
    const double sb = rms[0] > 0.0 ? 1.0 / rms[0] : 0.0, sx = rms[1] > 0.0 ? 1.0 / rms[1] : 0.0;
    feat[p] = in ? b[IX0(i, j)] * sb : 0.0;
    feat[plane + p] = in ? V.r[j] / r_top : 0.0;
    feat[2 * plane + p] = (in && xprev) ? xprev[(size_t)i * lx + j] * sx : 0.0;
}

// previous-step solution (pitch V.ld) and, for training capture, the
// target x / rms(x)
__global__ void k_learn_store(NddView V, const double* __restrict__ x, const double* __restrict__ rms,
                              double* __restrict__ xprev, double* __restrict__ target)
{
    THREAD_IJ(0, 0);
    if (i > V.ni + 1 || j > V.nj + 1) return;
    const bool in = i >= 1 && i <= V.ni && j >= 1 && j <= V.nj;
    const double v = in ? x[IX0(i, j)] : 0.0;
    xprev[IX(i, j)] = v;
    if (target) target[IX(i, j)] = rms[0] > 0.0 ? v / rms[0] : 0.0;
}

static double** field_slot(cathode_ndd_fields* f, int k)
{
    double** a[19] = {&f->phi, &f->rhoe, &f->rhoi, &f->Te, &f->rhoeps, &f->rhoe_old, &f->Ex, &f->rhoUex, &f->nUex,
                      &f->Gx, &f->rhoUix, &f->Er, &f->rhoUer, &f->nUer, &f->Gr, &f->rhoUir, &f->rate_ionize,
                      &f->Pabs, &f->rate_eloss};
    return a[k];
}


// up to 8 whole-plane copies in one launch (blockIdx.y picks the pair)
constexpr int COPY_MAX = 12;
struct CopyList {
    double* dst[COPY_MAX];
    const double* src[COPY_MAX];
};

__global__ void k_copy_planes(CopyList L, size_t n)
{
    double* __restrict__ d = L.dst[blockIdx.y];
    const double* __restrict__ a = L.src[blockIdx.y];
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        d[k] = a[k];
}

// dst_q = src_q for each listed pair, in one launch
void d2d(NddHandle& H, std::initializer_list<std::pair<double*, const double*>> pairs)
{
    CopyList L{};
    int q = 0;
    for (const auto& pr : pairs) {
        if (q == COPY_MAX) throw std::logic_error("d2d: more than COPY_MAX planes");
        L.dst[q] = pr.first;
        L.src[q] = pr.second;
        q++;
    }
    if (q == 0) return;
    const int gx = std::min(cdiv((long)H.n, 256), 512);
    hipLaunchKernelGGL(k_copy_planes, dim3(gx, q), dim3(256), 0, H.s, L, H.n);
    CATHODE_LAUNCH_CHECK();
}

double read_dev(NddHandle& H, const double* d, int count)
{
    fetch_host(H.h_scal, d, count, H.s);
    CATHODE_HIP(hipStreamSynchronize(H.s));
    return H.h_scal[0];
}

static void account(NddHandle& H, const cathode_status& st)
{
    H.rep.solves++;
    H.rep.total_iterations += st.iterations;
    if (st.iterations > H.rep.max_iterations) H.rep.max_iterations = st.iterations;
    if (st.rel_residual > H.rep.max_rel_residual || std::isnan(st.rel_residual)) H.rep.max_rel_residual = st.rel_residual;
    if (!st.converged) H.rep.nonconverged++;
}

// gate: the PC stopping threshold the previous iteration's calcRes is
// tested against (-inf for the first iteration).  Returns false when the
// loop should have stopped before this iteration.
// slot: the solve's warm-start slot; in learned mode its iteration count of
// the previous step is the solver's expected count (fewer host round trips)
static bool solve(NddHandle& H, double gate, int maxITR, double tol, int slot)
{
    cathode_status st{};
    H.mg.side_cnv = gate;
    if (H.p.learned) H.mg.expected_its = H.last_its[slot];
    if (H.cut_on) H.mg.cut_hint = H.cut[slot];
    H.rep.cut_by_solve[slot] = H.cut_on ? H.cut[slot] : -1;
    H.mg.solve(maxITR, tol, H.p.mode, H.s, &st);
    H.mg.side_cnv = -INFINITY;
    if (H.mg.aborted) return false;
    if (H.cut_on) H.cut[slot] = H.mg.cut_next();
    H.last_its[slot] = st.iterations;
    H.rep.iterations_by_solve[slot] += st.iterations;
    account(H, st);
    return true;
}

// calcRes of three pairs without a host round trip: the max error lands in
// the multigrid's side slot and is read with the next solve's first sync
static void calcres3_async(NddHandle& H, const ResPairs& P)
{
    const NddView& V = H.V;
    NDD_GEO_LAUNCH(H, k_calcres, ggrid(V.ni, V.nj), gblk(), V, P, H.p.flag_error, 1e-100, H.part.p, H.nb);
    hipLaunchKernelGGL(k_calcres_finish, dim3(1), dim3(256), 0, H.s, H.part.p, H.nb, H.p.flag_error, H.scal.p,
                       H.mg.side_slot());
    CATHODE_LAUNCH_CHECK();
}

// calcRes of three pairs; returns the max error (host sync); errors also in scal[0..3]
double calcres3(NddHandle& H, const ResPairs& P, double* err_out, int mode)
{
    const NddView& V = H.V;
    if (mode < 0) mode = H.p.flag_error;
    NDD_GEO_LAUNCH(H, k_calcres, ggrid(V.ni, V.nj), gblk(), V, P, mode, 1e-100, H.part.p, H.nb);
    hipLaunchKernelGGL(k_calcres_finish, dim3(1), dim3(256), 0, H.s, H.part.p, H.nb, mode, H.scal.p,
                       (double*)nullptr);
    CATHODE_LAUNCH_CHECK();
    fetch_host(H.h_scal, H.scal.p, 4, H.s);
    CATHODE_HIP(hipStreamSynchronize(H.s));
    if (err_out)
        for (int q = 0; q < 3; q++) err_out[q] = H.h_scal[q];
    return H.h_scal[3];
}

// Features of the assembled first Poisson correction system, the network's
// guess g, and the solver's initial guess alpha g with alpha minimising
// |b - alpha A g| (never worse than x = 0; 0 when the guess does not help).
// With capture set, the features are copied out for training.
enum { EQ_ELEC = 0, EQ_ION = 1, EQ_ENERGY = 2, EQ_POISSON = 3 };
static int warm_slot(int eq, int ncount) { return eq * 2 + (ncount >= 2 ? 1 : 0); }

constexpr double NET_USEFUL = 0.9;  // residual ratio below which the network's term counts as useful
constexpr int NET_PROBE = 16;       // while it is not, it is tried once every NET_PROBE steps

static void warm_guess(NddHandle& H, int slot, const double* g = nullptr);

static void learned_guess(NddHandle& H, bool net)
{
    const NddView& V = H.V;
    const size_t plane = (size_t)(V.ni + 2) * V.ld;
    H.lfeat.alloc(3 * plane);
    H.lguess.alloc(plane);
    H.xprev.alloc(plane);
    H.lpart.alloc(6 * H.nb + 64);
    H.lscal.alloc(8);
    const dim3 g1 = ggrid(V.ni, V.nj), g2 = ggrid(V.ni + 2, V.nj + 2);
    // the previous step's solution of this solve: the learned mode's warm
    // slot (pitch ld0), else the copy learned_store keeps (pitch ld)
    const int slot = warm_slot(EQ_POISSON, 1);
    const bool from_slot = H.p.learned && H.nhist[slot] > 0;
    const double* xp = from_slot ? (const double*)H.wbuf[slot][0].p
                                 : (H.have_xprev ? (const double*)H.xprev.p : (const double*)nullptr);
    const int lxp = from_slot ? V.ld0 : V.ld;
    hipLaunchKernelGGL(k_learn_sumsq, g1, gblk(), 0, H.s, V, (const double*)H.mg.rhs(), V.ld0, xp, lxp, H.lpart.p,
                       H.nb);
    hipLaunchKernelGGL(k_learn_rms, dim3(1), dim3(256), 0, H.s, (const double*)H.lpart.p, H.nb,
                       (double)V.ni * V.nj, H.lscal.p);
    const double r_top = H.p.rmin + V.nj * V.dr;
    hipLaunchKernelGGL(k_learn_features, g2, gblk(), 0, H.s, V, (const double*)H.mg.rhs(), xp, lxp,
                       (const double*)H.lscal.p, r_top, H.lfeat.p);
    CATHODE_LAUNCH_CHECK();
    if (H.cap_feat)
        CATHODE_HIP(hipMemcpyAsync(H.cap_feat, H.lfeat.p, 3 * plane * sizeof(double), hipMemcpyDeviceToDevice, H.s));
    if (!(H.p.learned && H.cnn && net)) {
        warm_guess(H, slot);  // learned mode without the network this step: the warm start alone
        return;
    }
    if (!H.cnn_sc) H.cnn_sc.reset(cnn_scratch_new());
    cnn_infer(H.cnn, V.ni, V.nj, H.lfeat.p, H.lguess.p, H.cnn_sc.get(), H.s);
    // the network's output joins the solve's history in the least-squares
    // combination; the device reports what it adds (h_scal[8])
    warm_guess(H, slot, H.lguess.p);
    H.guesses++;
    H.gate_pending = true;
}

// after the first Poisson correction solve: keep its solution as the next
// step's feature, and write the training target when capturing
static void learned_store(NddHandle& H)
{
    const NddView& V = H.V;
    const dim3 g1 = ggrid(V.ni, V.nj), g2 = ggrid(V.ni + 2, V.nj + 2);
    hipLaunchKernelGGL(k_learn_sumsq, g1, gblk(), 0, H.s, V, (const double*)H.mg.sol(), V.ld0,
                       (const double*)nullptr, V.ld, H.lpart.p, H.nb);
    hipLaunchKernelGGL(k_learn_rms, dim3(1), dim3(256), 0, H.s, (const double*)H.lpart.p, H.nb,
                       (double)V.ni * V.nj, H.lscal.p + 4);
    hipLaunchKernelGGL(k_learn_store, g2, gblk(), 0, H.s, V, (const double*)H.mg.sol(), (const double*)H.lscal.p + 4,
                       H.xprev.p, H.cap_target);
    CATHODE_LAUNCH_CHECK();
    H.have_xprev = true;
}


// learned mode: x0 = the least-squares combination of this solve's last
// solutions (and of the network's guess g, pitch V.ld, when given; then its
// contribution is reported to h_scal[8])
static void warm_guess(NddHandle& H, int slot, const double* g)
{
    if (!H.p.learned) return;
    WarmSet ws;
    for (int q = 0; q < H.nhist[slot]; q++) {
        ws.y[ws.k] = H.wbuf[slot][q].p;
        ws.ld[ws.k] = H.V.ld0;
        ws.k++;
    }
    if (g) {
        ws.y[ws.k] = g;
        ws.ld[ws.k] = H.V.ld;
        ws.k++;
        ws.report = H.h_scal + 8;
    }
    if (ws.k > 0) H.mg.set_warm_start(ws);
}

// after the kernels that read the solution have been launched: the slot
// takes the solution buffer as its newest entry, the solver its oldest
// buffer (no copy)
static void warm_store(NddHandle& H, int slot)
{
    if (!H.p.learned) return;
    auto& h = H.wbuf[slot];
    double* oldest = h[NddHandle::WARM_HIST - 1].p;
    for (int q = NddHandle::WARM_HIST - 1; q > 0; q--) h[q].p = h[q - 1].p;
    h[0].p = H.mg.exchange_solution(oldest);
    H.nhist[slot] = std::min(H.nhist[slot] + 1, NddHandle::WARM_HIST);
}

static void poisson_routine(NddHandle& H, bool report)
{
    const NddView& V = H.V;
    double** F = H.F;
    // start-of-pass copies of phi, nUex, nUer: pass p reads set (p - 1) & 1
    // and its update writes set p & 1 for the next pass
    double* const snap[2][3] = {{H.tmp_c, H.tmp_x, H.tmp_r}, {H.tmp2_c, H.tmp2_x, H.tmp2_r}};
    // the step's old values and the first pass's copies (both flux sets get
    // the entries no update rewrites), one launch; the ghost kernel then sets
    // the ghosts of phi and of its copy
    d2d(H, {{H.phi_old, F[F_PHI]}, {H.Ex_old, F[F_EX]}, {H.Er_old, F[F_ER]}, {H.sx_old, F[F_NUEX]},
            {H.sr_old, F[F_NUER]}, {H.tmp_c, F[F_PHI]}, {H.tmp_x, F[F_NUEX]}, {H.tmp_r, F[F_NUER]},
            {H.tmp2_x, F[F_NUEX]}, {H.tmp2_r, F[F_NUER]}});
    NDD_GEO_LAUNCH(H, k_poisson_ghost, ggrid(V.ni + 2, V.nj + 2), gblk(), V, F[F_PHI], H.tmp_c);
    PoissonArgs A{};
    A.phi = F[F_PHI]; A.rhoe = F[F_RHOE]; A.rhoi = F[F_RHOI]; A.Te = F[F_TE]; A.rhoe_old = F[F_RHOE_OLD];
    A.Ex = F[F_EX]; A.Er = F[F_ER]; A.Ex_old = H.Ex_old; A.Er_old = H.Er_old; A.rhoUix = F[F_RHOUIX];
    A.rhoUir = F[F_RHOUIR];
    A.S = H.has_src[0] ? F[F_S] : nullptr;
    A.nUex = F[F_NUEX]; A.nUer = F[F_NUER];
    A.m = H.me;
    int ncount = 0;
    for (;;) {
        ncount++;
        double* const* S = snap[(ncount - 1) & 1];
        double* const* Sn = snap[ncount & 1];
        A.nUex_t = S[1];
        A.nUer_t = S[2];
        NDD_GEO_LAUNCH(H, k_poisson_assemble, ggrid(V.ni, V.nj), gblk(), V, A, H.mg.aP(), H.mg.aE(), H.mg.aW(),
                       H.mg.aN(), H.mg.aS(), H.mg.rhs(), H.mg.prep_s(), H.mg.prep_part());
        H.mg.prepared = true;
        CATHODE_LAUNCH_CHECK();
        bool use_net = false;
        if (ncount == 1 && H.p.learned && H.cnn && !H.masked) {
            // the previous use's residual ratio (written by the device, read
            // here after the syncs of the steps since)
            if (H.gate_pending) {
                H.gate_pending = false;
                H.net_useful = H.h_scal[8] < NET_USEFUL;
                if (!H.net_useful) H.net_wait = NET_PROBE - 1;
            }
            use_net = H.net_useful || H.net_wait == 0;
            if (!use_net) H.net_wait--;
        }
        const bool learn = ncount == 1 && (use_net || H.cap_feat);
        if (learn) learned_guess(H, use_net);
        else warm_guess(H, warm_slot(EQ_POISSON, ncount));
        // the later passes' operator depends only on n_e, mu and dt, which the
        // routine does not change: the first pass's hierarchy is reused
        H.mg.same_operator = ncount > 1;
        if (!solve(H, ncount > 1 ? H.p.error_cnv_HES_phi : -INFINITY, H.p.maxITR_SOR_phi, H.p.error_cnv_SOR_phi,
                   warm_slot(EQ_POISSON, ncount))) {
            // calcRes of the previous iteration met the PC criterion: undo the prediction
            d2d(H, {{F[F_NUEX], S[1]}, {F[F_NUER], S[2]}});
            break;
        }
        if (learn && (H.cap_feat || !H.p.learned)) learned_store(H);  // learned mode reads its warm slot
        NDD_GEO_LAUNCH(H, k_poisson_update, ggrid(V.ni + 2, V.nj + 2), gblk(), V, (const double*)H.mg.sol(),
                       (const double*)S[0], H.dphi, F[F_PHI], F[F_EX], F[F_ER], (const double*)F[F_RHOE],
                       F[F_NUEX], F[F_NUER], H.me, Sn[0], Sn[1], Sn[2]);
        warm_store(H, warm_slot(EQ_POISSON, ncount));
        CATHODE_LAUNCH_CHECK();
        if (ncount >= H.p.maxITR_HES_phi) break;
        ResPairs P{F[F_PHI], S[0], F[F_NUEX], S[1], F[F_NUER], S[2]};
        calcres3_async(H, P);
    }
    if (report) {
        double e[3];
        ResPairs P{F[F_PHI], H.phi_old, F[F_NUEX], H.sx_old, F[F_NUER], H.sr_old};
        calcres3(H, P, e);
        H.rep.error_phi = e[0];
        H.rep.error_nUex = e[1];
        H.rep.error_nUer = e[2];
    }
}

static void sg_routine(NddHandle& H, int sp, bool report, double* err3)
{
    const NddView& V = H.V;
    double** F = H.F;
    SgArgs A{};
    A.sp = sp;
    A.Ex = F[F_EX]; A.Er = F[F_ER]; A.Te = F[F_TE]; A.rhoi = F[F_RHOI]; A.rhoe = F[F_RHOE];
    A.rhoe_old = F[F_RHOE_OLD]; A.rhoUix = F[F_RHOUIX]; A.rhoUir = F[F_RHOUIR]; A.rhoUex = F[F_RHOUEX];
    A.rhoUer = F[F_RHOUER];
    A.S = H.has_src[0] ? F[F_S] : nullptr;
    A.Pabs = H.has_src[1] ? F[F_PABS] : nullptr;
    A.eloss = H.has_src[2] ? F[F_ELOSS] : nullptr;
    A.wgt = H.sgw.p;
    A.wn = (size_t)(V.ni + 2) * V.ld0;
    int maxit, lin_maxit;
    double cnv, lin_tol;
    double* n_old;
    if (sp == SP_ELEC) {
        A.n = F[F_RHOE]; A.Fx = F[F_RHOUEX]; A.Fr = F[F_RHOUER]; A.T = F[F_TE]; A.m = H.me; A.tfac = 1.0;
        maxit = H.p.maxITR_HES_rhoe; cnv = H.p.error_cnv_HES_rhoe; n_old = F[F_RHOE_OLD];
        lin_maxit = H.p.maxITR_SOR_rhoe; lin_tol = H.p.error_cnv_SOR_rhoe;
    } else if (sp == SP_ION) {
        A.n = F[F_RHOI]; A.Fx = F[F_RHOUIX]; A.Fr = F[F_RHOUIR]; A.T = F[F_TE]; A.m = H.mi; A.tfac = 1.0;
        maxit = H.p.maxITR_HES_rhoi; cnv = H.p.error_cnv_HES_rhoi; n_old = H.rhoi_old;
        lin_maxit = H.p.maxITR_SOR_rhoi; lin_tol = H.p.error_cnv_SOR_rhoi;
    } else {
        A.n = F[F_RHOEPS]; A.Fx = F[F_GX]; A.Fr = F[F_GR]; A.T = H.Te_old; A.m = H.me; A.tfac = 3.0 / 5.0;
        maxit = H.p.maxITR_HES_rhoeps; cnv = H.p.error_cnv_HES_rhoeps; n_old = H.s_old;
        lin_maxit = H.p.maxITR_SOR_rhoeps; lin_tol = H.p.error_cnv_SOR_rhoeps;
    }
    A.n_old = n_old;
    // start-of-pass copies of n, Fx, Fr: pass p reads set (p - 1) & 1 and its
    // correction writes set p & 1 for the next pass
    double* const snap[2][3] = {{H.tmp_c, H.tmp_x, H.tmp_r}, {H.tmp2_c, H.tmp2_x, H.tmp2_r}};
    int ncount = 0;
    for (;;) {
        ncount++;
        double* const* S = snap[(ncount - 1) & 1];
        double* const* Sn = snap[ncount & 1];
        if (ncount == 1) {
            // the step's old values and this pass's, one launch; both flux
            // sets get the entries no correction rewrites
            if (sp == SP_ENERGY)
                d2d(H, {{n_old, A.n}, {H.sx_old, A.Fx}, {H.sr_old, A.Fr}, {S[0], A.n}, {S[1], A.Fx}, {S[2], A.Fr},
                        {Sn[1], A.Fx}, {Sn[2], A.Fr}, {H.Te_old, F[F_TE]}});
            else
                d2d(H, {{n_old, A.n}, {H.sx_old, A.Fx}, {H.sr_old, A.Fr}, {S[0], A.n}, {S[1], A.Fx}, {S[2], A.Fr},
                        {Sn[1], A.Fx}, {Sn[2], A.Fr}});
        }
        A.Fx_t = S[1];
        A.Fr_t = S[2];
        A.reuse = ncount > 1;
        NDD_GEO_LAUNCH(H, k_sg_assemble, ggrid(V.ni, V.nj), gblk(), V, A, H.mg.aP(), H.mg.aE(), H.mg.aW(),
                       H.mg.aN(), H.mg.aS(), H.mg.rhs(), H.mg.prep_s(), H.mg.prep_part());
        H.mg.prepared = true;
        CATHODE_LAUNCH_CHECK();
        const int eq = sp == SP_ELEC ? EQ_ELEC : (sp == SP_ION ? EQ_ION : EQ_ENERGY);
        warm_guess(H, warm_slot(eq, ncount));
        // the later passes read back the first pass's face weights (the
        // fields and temperatures they depend on do not change in the
        // routine), so their operator is bitwise the first pass's
        H.mg.same_operator = ncount > 1;
        if (!solve(H, ncount > 1 ? cnv : -INFINITY, lin_maxit, lin_tol, warm_slot(eq, ncount))) {
            d2d(H, {{A.Fx, S[1]}, {A.Fr, S[2]}});
            break;
        }
        NDD_GEO_LAUNCH(H, k_sg_correct, ggrid(V.ni, V.nj), gblk(), V, A, (const double*)H.mg.sol(), Sn[0], Sn[1],
                       Sn[2]);
        CATHODE_LAUNCH_CHECK();
        warm_store(H, warm_slot(eq, ncount));
        if (ncount >= maxit) break;
        ResPairs P{A.n, S[0], A.Fx, S[1], A.Fr, S[2]};
        calcres3_async(H, P);
    }
    if (report) {
        ResPairs P{A.n, n_old, A.Fx, H.sx_old, A.Fr, H.sr_old};
        calcres3(H, P, err3);
    }
}

// check_CFL before a step (main.cpp:174) into H.scal[16..21]; with
// flag_adp_dt = 1 the step's dt is read back and set (one host round trip)
static void check_cfl(NddHandle& H)
{
    NddView& V = H.V;
    const dim3 g = ggrid(V.ni + 2, V.nj + 2);
    const int nb = (int)(g.x * g.y);
    hipLaunchKernelGGL(k_cfl_partials, g, gblk(), 0, H.s, V, (const double*)H.F[F_RHOE], (const double*)H.F[F_TE],
                       (const double*)H.F[F_RHOUEX], (const double*)H.F[F_RHOUER], H.cfl_part.p);
    hipLaunchKernelGGL(k_cfl_finish, dim3(1), dim3(256), 0, H.s, (const double*)H.cfl_part.p, nb,
                       std::max(0.0, H.p.mu_e), H.p.masse, V.dt, H.p.CFL, H.p.flag_adp_dt == 1 ? 1 : 0,
                       H.scal.p + 16);
    CATHODE_LAUNCH_CHECK();
    if (H.p.flag_adp_dt == 1) {
        V.dt = read_dev(H, H.scal.p + 16);
        if (!(V.dt > 0.0) || !std::isfinite(V.dt)) throw std::domain_error("check_CFL: non-finite adaptive dt");
    }
}

static void ndd_one_step(NddHandle& H, bool report)
{
    // time step: every step when dt adapts, else for the report of the last step
    if (H.p.flag_adp_dt == 1 || report) check_cfl(H);
    const NddView& V = H.V;
    double** F = H.F;
    double e3[3] = {0, 0, 0};
    // 1. ions
    sg_routine(H, SP_ION, report, e3);
    if (report) H.rep.error_rhoi = e3[0];
    if (H.masked)
        hipLaunchKernelGGL(k_open_copy<GeoMask>, ggrid(V.ni, V.nj), gblk(), 0, H.s, H.gm, V, F[F_RHOI]);
    else
        hipLaunchKernelGGL(k_top_copy, dim3(cdiv(V.ni, 256)), dim3(256), 0, H.s, V, F[F_RHOI]);
    // 2. Poisson
    poisson_routine(H, report);
    // 3. electrons
    sg_routine(H, SP_ELEC, report, e3);
    if (H.masked) {
        hipLaunchKernelGGL(k_ne_ghost<GeoMask>, ggrid(V.ni, V.nj), gblk(), 0, H.s, H.gm, V, F[F_RHOE]);
        CATHODE_LAUNCH_CHECK();
    }
    if (report) {
        H.rep.error_rhoe = e3[0];
        H.rep.error_rhoUex = e3[1];
        H.rep.error_rhoUer = e3[2];
    }
    // 4. electron energy
    if (H.p.energy_on) {
        // Te_old = Te is taken with the routine's first copies
        sg_routine(H, SP_ENERGY, report, e3);
        if (report) {
            H.rep.error_rhoeps = e3[0];
            H.rep.error_Gx = e3[1];
            H.rep.error_Gr = e3[2];
        }
        NDD_GEO_LAUNCH(H, k_energy_guard_max, ggrid(V.ni, V.nj), gblk(), V, F[F_RHOEPS], (const double*)H.s_old,
                       (const double*)F[F_RHOE], H.part.p);
        // max n_e: every workgroup of the update reads the guard's partials
        // (one pre-reduced value on large grids, slot 24)
        const double* gp = H.part.p;
        int gnb = H.nb;
        if (gnb > presum_min()) {
            presum_launch(H.part.p, gnb, 1, 1u, H.scal.p + 24, H.s);
            gp = H.scal.p + 24;
            gnb = 1;
        }
        NDD_GEO_LAUNCH(H, k_te_update, ggrid(V.ni, V.nj), gblk(), V, (const double*)F[F_RHOEPS],
                       (const double*)F[F_RHOE], gp, gnb, F[F_TE]);
        CATHODE_LAUNCH_CHECK();
    }
    H.rep.steps++;
}

}  // namespace cathode

using namespace cathode;

extern "C" {

int cathode_ndd_create(const cathode_ndd_params* p, void** handle)
{
    try {
        if (!require_init()) return CATHODE_ERR_NOINIT;
        if (!p || !handle || p->ni < 2 || p->nj < 2 || !(p->dx > 0) || !(p->dr > 0) || !(p->dt > 0) ||
            p->n_see < 0 || p->n_see > CATHODE_SEE_TABLE_MAX || (p->mode != CATHODE_MODE_FP64 && p->mode != CATHODE_MODE_MIXED) ||
            p->maxITR_HES_phi < 1 || p->maxITR_HES_rhoe < 1 || p->maxITR_HES_rhoi < 1 ||
            (p->flag_adp_dt == 1 && !(p->CFL > 0.0)) ||
            (p->energy_on && p->maxITR_HES_rhoeps < 1) || p->flag_error < 0 || p->flag_error > 3) {
            set_error("ndd_create: invalid parameters");
            return CATHODE_ERR_ARG;
        }
        std::unique_ptr<NddHandle> H(new NddHandle());
        H->p = *p;
        {
            const char* e = std::getenv("CATHODE_CUT");
            H->cut_on = !(e && e[0] == '0');
        }
        // each instance has its own stream so independent instances can run
        // side by side (cathode_ndd_batch_step)
        CATHODE_HIP(hipStreamCreateWithFlags(&H->s, hipStreamNonBlocking));
        const int ni = p->ni, nj = p->nj;
        H->n = (size_t)(ni + 2) * (nj + 2);
        H->fields.alloc(19 * H->n);
        CATHODE_HIP(hipMemset(H->fields.p, 0, 19 * H->n * sizeof(double)));
        for (int k = 0; k < 19; k++) H->F[k] = H->fields.p + k * H->n;
        constexpr int NW = 15;
        H->work.alloc(NW * H->n);
        CATHODE_HIP(hipMemset(H->work.p, 0, NW * H->n * sizeof(double)));
        CATHODE_HIP(hipDeviceSynchronize());  // null-stream memsets before the handle's stream runs
        double** w[NW] = {&H->phi_old, &H->Ex_old, &H->Er_old, &H->tmp_c,  &H->tmp_x,  &H->tmp_r,  &H->dphi,    &H->s_old,
                          &H->sx_old,  &H->sr_old, &H->Te_old, &H->rhoi_old, &H->tmp2_c, &H->tmp2_x, &H->tmp2_r};
        for (int k = 0; k < NW; k++) *w[k] = H->work.p + k * H->n;
        std::vector<double> r(nj + 2);
        for (int j = 0; j < nj + 2; j++) r[j] = p->rmin + (j - 0.5) * p->dr;  // initialModule.hpp:86-88
        H->radii.alloc(nj + 2);
        CATHODE_HIP(hipMemcpy(H->radii.p, r.data(), (nj + 2) * sizeof(double), hipMemcpyHostToDevice));
        H->see.alloc(3 * CATHODE_SEE_TABLE_MAX);
        std::vector<double> tab(3 * CATHODE_SEE_TABLE_MAX);
        std::memcpy(tab.data(), p->see_eb, sizeof(p->see_eb));
        std::memcpy(tab.data() + CATHODE_SEE_TABLE_MAX, p->see_rd, sizeof(p->see_rd));
        std::memcpy(tab.data() + 2 * CATHODE_SEE_TABLE_MAX, p->see_ts, sizeof(p->see_ts));
        CATHODE_HIP(hipMemcpy(H->see.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        NddView& V = H->V;
        V.ni = ni; V.nj = nj; V.ld = nj + 2;
        V.dx = p->dx; V.dr = p->dr; V.dt = p->dt;
        V.V_bias = p->V_bias; V.mu_i = p->mu_i; V.Ti = p->Ti; V.masse = p->masse; V.massi = p->massi;
        V.coefIISEE_ts = p->coefIISEE_ts; V.Te_emitSEE_eV = p->Te_emitSEE_eV;
        V.ratioEngy_EISEE_rd = p->ratioEngy_EISEE_rd; V.dTe_SEE = p->dTe_SEE; V.n_see = p->n_see;
        V.r = H->radii.p;
        V.see_eb = H->see.p; V.see_rd = H->see.p + CATHODE_SEE_TABLE_MAX; V.see_ts = H->see.p + 2 * CATHODE_SEE_TABLE_MAX;
        H->me = mobility(p->mu_e);
        H->mi = mobility(p->mu_i);
        H->geo_src = rect_geometry(ni, nj);
        H->mg.configure(ni, nj, H->s);
        V.ld0 = H->mg.ld0();
        const int n0 = (ni + 2) * V.ld0;
        hipLaunchKernelGGL(k_fill_one, dim3(cdiv(n0, 256)), dim3(256), 0, H->s, n0, H->mg.act(), ni, nj, V.ld0);
        CATHODE_LAUNCH_CHECK();
        H->anode.n = 1;
        H->anode.fi[0] = ni + 1;
        H->anode.j0[0] = 1;
        H->anode.j1[0] = nj;
        H->mg.set_weights(H->radii.p);  // conservative rows: w_j = r_j
        H->nb = cdiv(nj, NB_X) * cdiv(ni, NB_Y);
        H->part.alloc(6 * H->nb + 64);
        H->sgw.alloc(4 * (size_t)n0);
        H->scal.alloc(32);
        {
            const dim3 gc = ggrid(ni + 2, nj + 2);
            H->cfl_part.alloc(3 * (size_t)gc.x * gc.y + 64);
        }
        if (p->learned) {
            for (auto& slot : H->wbuf)
                for (auto& wb : slot) {
                    wb.alloc((size_t)n0);
                    CATHODE_HIP(hipMemset(wb.p, 0, (size_t)n0 * sizeof(double)));  // zero ghosts and padding
                }
            CATHODE_HIP(hipDeviceSynchronize());  // the memsets run on the null stream
        }
        CATHODE_HIP(hipHostMalloc(&H->h_scal, 32 * sizeof(double), hipHostMallocDefault));
        CATHODE_HIP(hipStreamSynchronize(H->s));
        *handle = H.release();
        return CATHODE_OK;
    CATHODE_CATCH_END
}

int cathode_ndd_destroy(void* handle)
{
    NddHandle* H = (NddHandle*)handle;
    if (!H) return CATHODE_OK;
    if (H->s) (void)hipStreamSynchronize(H->s);
    if (H->h_scal) (void)hipHostFree(H->h_scal);
    hipStream_t s = H->s;
    delete H;
    if (s) (void)hipStreamDestroy(s);
    return CATHODE_OK;
}

int cathode_ndd_set_geometry(void* handle, const cathode_ndd_geometry* g)
{
    CATHODE_GUARD_BEGIN
    NddHandle* H = (NddHandle*)handle;
    if (!H) {
        set_error("ndd_set_geometry: null handle");
        return CATHODE_ERR_ARG;
    }
    NddView& V = H->V;
    const int n0 = (V.ni + 2) * V.ld0;
    H->geo_version++;
    if (!g) {
        H->masked = false;
        H->geo_src = rect_geometry(V.ni, V.nj);
        hipLaunchKernelGGL(k_fill_one, dim3(cdiv(n0, 256)), dim3(256), 0, H->s, n0, H->mg.act(), V.ni, V.nj, V.ld0);
        CATHODE_LAUNCH_CHECK();
        H->anode.n = 1;
        H->anode.fi[0] = V.ni + 1;
        H->anode.j0[0] = 1;
        H->anode.j1[0] = V.nj;
        CATHODE_HIP(hipStreamSynchronize(H->s));
        return CATHODE_OK;
    }
    GeoTables T;
    std::string err;
    if (!build_geo_tables(V.ni, V.nj, *g, T, err)) {
        set_error(err.c_str());
        return CATHODE_ERR_ARG;
    }
    if ((int)T.anode.size() / 3 > ANODE_MAX) {
        set_error("ndd_set_geometry: more than 4 biased high-x wall segments");
        return CATHODE_ERR_ARG;
    }
    const size_t n = T.xc.size();
    H->geo_u8.alloc(6 * n);
    H->geo_gs.alloc(n);
    std::vector<uint8_t> u8(6 * n);
    std::memcpy(u8.data(), T.xc.data(), n);
    std::memcpy(u8.data() + n, T.rc.data(), n);
    std::memcpy(u8.data() + 2 * n, T.act.data(), n);
    std::memcpy(u8.data() + 3 * n, T.te.data(), n);
    std::memcpy(u8.data() + 4 * n, T.ne.data(), n);
    std::memcpy(u8.data() + 5 * n, T.blk.data(), n);
    CATHODE_HIP(hipMemcpyAsync(H->geo_u8.p, u8.data(), 6 * n, hipMemcpyHostToDevice, H->s));
    CATHODE_HIP(hipMemcpyAsync(H->geo_gs.p, T.gs.data(), n * sizeof(int), hipMemcpyHostToDevice, H->s));
    H->gm.xc = H->geo_u8.p;
    H->gm.rc = H->geo_u8.p + n;
    H->gm.ac = H->geo_u8.p + 2 * n;
    H->gm.teb = H->geo_u8.p + 3 * n;
    H->gm.neb = H->geo_u8.p + 4 * n;
    H->gm.bk = H->geo_u8.p + 5 * n;
    H->gm.gs = H->geo_gs.p;
    hipLaunchKernelGGL(k_act_from_table, dim3(cdiv(n0, 256)), dim3(256), 0, H->s, V, (const uint8_t*)H->gm.ac,
                       H->mg.act(), n0);
    CATHODE_LAUNCH_CHECK();
    H->anode.n = (int)T.anode.size() / 3;
    for (int q = 0; q < H->anode.n; q++) {
        H->anode.fi[q] = T.anode[3 * q];
        H->anode.j0[q] = T.anode[3 * q + 1];
        H->anode.j1[q] = T.anode[3 * q + 2];
    }
    H->masked = true;
    H->geo_src = *g;
    CATHODE_HIP(hipStreamSynchronize(H->s));  // the host tables go out of scope
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_ndd_set_learned(void* handle, void* net)
{
    NddHandle* H = (NddHandle*)handle;
    if (!H) {
        set_error("ndd_set_learned: null handle");
        return CATHODE_ERR_ARG;
    }
    H->cnn = net;
    return CATHODE_OK;
}

int cathode_ndd_capture(void* handle, double* feat_dev, double* target_dev)
{
    NddHandle* H = (NddHandle*)handle;
    if (!H || (!feat_dev) != (!target_dev)) {
        set_error("ndd_capture: null handle or only one of the two buffers");
        return CATHODE_ERR_ARG;
    }
    H->cap_feat = feat_dev;
    H->cap_target = target_dev;
    return CATHODE_OK;
}

int cathode_ndd_set_state(void* handle, const cathode_ndd_fields* host)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || !host) {
            set_error("ndd_set_state: null argument");
            return CATHODE_ERR_ARG;
        }
        cathode_ndd_fields f = *host;
        H->have_xprev = false;  // a new state: no previous Poisson correction
        for (int& v : H->nhist) v = 0;
        for (int& v : H->last_its) v = 0;
        for (int& v : H->cut) v = -1;
        H->net_useful = true;
        H->gate_pending = false;
        H->net_wait = 0;
        for (int k = 0; k < 19; k++) {
            const double* src = *field_slot(&f, k);
            if (!src) {
                if (k < F_S) {
                    set_error("ndd_set_state: state array missing");
                    return CATHODE_ERR_ARG;
                }
                CATHODE_HIP(hipMemsetAsync(H->F[k], 0, H->n * sizeof(double), H->s));
                H->has_src[k - F_S] = false;
                continue;
            }
            CATHODE_HIP(hipMemcpyAsync(H->F[k], src, H->n * sizeof(double), hipMemcpyHostToDevice, H->s));
            if (k >= F_S) H->has_src[k - F_S] = true;
        }
        CATHODE_HIP(hipStreamSynchronize(H->s));
        return CATHODE_OK;
    CATHODE_CATCH_END
}

int cathode_ndd_get_state(void* handle, cathode_ndd_fields* host)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || !host) {
            set_error("ndd_get_state: null argument");
            return CATHODE_ERR_ARG;
        }
        for (int k = 0; k < 19; k++) {
            double* dst = *field_slot(host, k);
            if (!dst) continue;
            CATHODE_HIP(hipMemcpyAsync(dst, H->F[k], H->n * sizeof(double), hipMemcpyDeviceToHost, H->s));
        }
        CATHODE_HIP(hipStreamSynchronize(H->s));
        return CATHODE_OK;
    CATHODE_CATCH_END
}

// nsteps implicit steps of one instance on its own stream; fills rep (may be NULL)
static int ndd_run(NddHandle* H, int nsteps, cathode_ndd_report* rep)
{
    {
        auto t0 = std::chrono::steady_clock::now();
        std::memset(&H->rep, 0, sizeof(H->rep));
        for (int s = 0; s < nsteps; s++) ndd_one_step(*H, rep != nullptr && s == nsteps - 1);
        if (rep && nsteps > 0) {
            hipLaunchKernelGGL(k_anode, dim3(1), dim3(256), 0, H->s, H->V, H->anode, (const double*)H->F[F_RHOUIX],
                               (const double*)H->F[F_RHOUEX], H->scal.p + 12);
            CATHODE_LAUNCH_CHECK();
            fetch_host(H->h_scal, H->scal.p + 12, 2, H->s);
            fetch_host(H->h_scal + 16, H->scal.p + 16, 6, H->s);
        }
        CATHODE_HIP(hipStreamSynchronize(H->s));
        if (rep) {
            *rep = H->rep;
            if (nsteps > 0) {
                rep->Ii_anode = H->h_scal[0];
                rep->Ie_anode = H->h_scal[1];
                rep->I_anode = H->h_scal[0] + H->h_scal[1];
                rep->dt = H->h_scal[16];
                rep->CFL = H->h_scal[17];
                rep->dt_min = H->h_scal[18];
                rep->rhoe_max = H->h_scal[19];
                rep->Te_max = H->h_scal[20];
                rep->ue_max = H->h_scal[21];
            }
            rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        if (!(H->rep.max_rel_residual == H->rep.max_rel_residual)) {
            set_error("ndd_step: non-finite residual in a linear solve");
            return CATHODE_ERR_NUMERIC;
        }
        return CATHODE_OK;
    }
}

// ndd_run with every exception turned into an error code (worker threads)
static int ndd_run_guarded(NddHandle* H, int nsteps, cathode_ndd_report* rep)
{
    try {
        return ndd_run(H, nsteps, rep);
    CATHODE_CATCH_END
}

int cathode_ndd_step(void* handle, int nsteps, cathode_ndd_report* rep)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || nsteps < 0) {
            set_error("ndd_step: invalid argument");
            return CATHODE_ERR_ARG;
        }
        return ndd_run(H, nsteps, rep);
    CATHODE_CATCH_END
}

int cathode_ndd_batch_step(void** handles, int n, int nsteps, int workers, cathode_ndd_report* reps)
{
    if (!handles || n < 0 || nsteps < 0) {
        set_error("ndd_batch_step: invalid argument");
        return CATHODE_ERR_ARG;
    }
    for (int k = 0; k < n; k++)
        if (!handles[k]) {
            set_error("ndd_batch_step: null handle");
            return CATHODE_ERR_ARG;
        }
    if (n == 0 || nsteps == 0) return CATHODE_OK;
    workers = std::max(1, std::min(workers, n));
    // largest instances first, dealt to the least loaded worker (LPT)
    std::vector<int> order(n);
    for (int k = 0; k < n; k++) order[k] = k;
    auto cells = [&](int k) { return (long)((NddHandle*)handles[k])->p.ni * ((NddHandle*)handles[k])->p.nj; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    std::vector<std::vector<int> > queue(workers);
    std::vector<long> load(workers, 0);
    for (int k : order) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        queue[w].push_back(k);
        load[w] += cells(k);
    }
    // one instance per grid shape and precision runs its first step alone, so
    // the tile autotuner times that shape on an otherwise idle GPU; the
    // others reuse its choice
    std::vector<int> done(n, 0);
    std::vector<cathode_ndd_report> first(n);
    std::vector<int> rc(n, CATHODE_OK);
    {
        std::vector<long> seen;
        for (int k : order) {
            NddHandle* H = (NddHandle*)handles[k];
            const long key = ((long)H->p.ni << 32) ^ ((long)H->p.nj << 2) ^ H->p.mode;
            if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
            seen.push_back(key);
            rc[k] = ndd_run_guarded(H, 1, &first[k]);
            done[k] = 1;
            if (rc[k] != CATHODE_OK) return rc[k];
        }
    }
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; w++)
        pool.emplace_back([&, w] {
            for (int k : queue[w]) {
                NddHandle* H = (NddHandle*)handles[k];
                {
                    cathode_ndd_report r{};
                    if (nsteps - done[k] == 0) {
                        if (reps) reps[k] = first[k];
                        continue;
                    }
                    rc[k] = ndd_run_guarded(H, nsteps - done[k], &r);
                    if (reps) {
                        if (done[k]) {  // merge the counters of the first step
                            r.steps += first[k].steps;
                            r.solves += first[k].solves;
                            r.total_iterations += first[k].total_iterations;
                            for (int q = 0; q < 8; q++) r.iterations_by_solve[q] += first[k].iterations_by_solve[q];
                            r.max_iterations = std::max(r.max_iterations, first[k].max_iterations);
                            r.nonconverged += first[k].nonconverged;
                            r.max_rel_residual = std::max(r.max_rel_residual, first[k].max_rel_residual);
                            r.seconds += first[k].seconds;
                        }
                        reps[k] = r;
                    }
                }
            }
        });
    for (auto& t : pool) t.join();
    for (int k = 0; k < n; k++)
        if (rc[k] != CATHODE_OK) return rc[k];
    return CATHODE_OK;
}

int cathode_ndd_profile(void* handle, int enable)
{
    NddHandle* H = (NddHandle*)handle;
    if (!H) return CATHODE_ERR_ARG;
    H->mg.prof.enabled = enable != 0;
    H->mg.prof.launches = 0;
    H->mg.prof.total_ms = 0.0;
    return CATHODE_OK;
}

int cathode_ndd_tile(void* handle, int mixed, int level, int* ti, int* tj, int* nt)
{
    NddHandle* H = (NddHandle*)handle;
    if (!H || level < 0) return CATHODE_ERR_ARG;
    H->mg.tile_shape(mixed, level, ti, tj, nt);
    return CATHODE_OK;
}

int cathode_ndd_profile_read(void* handle, int* launches, double* total_ms, double* bytes_per_launch)
{
    NddHandle* H = (NddHandle*)handle;
    if (!H) return CATHODE_ERR_ARG;
    if (launches) *launches = H->mg.prof.launches;
    if (total_ms) *total_ms = H->mg.prof.total_ms;
    if (bytes_per_launch) *bytes_per_launch = H->mg.prof.bytes_per_launch;
    return CATHODE_OK;
}

}  // extern "C"
