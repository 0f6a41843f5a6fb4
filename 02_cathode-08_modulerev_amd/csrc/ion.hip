// ion.hip -- the ion block of the reference's time step (src/main.cpp:194-213)
// on the NDD handle's grid and plasma geometry: the temporary ion velocities
// (FluidModule::solve_Uix_tmp, solve_Uir_tmp, solve_Uip_tmp,
// include/fluidModule.hpp:123-1398), the ion density (solve_rhoi_constTe,
// :1405-1755) and the velocity / flux correction (correct_Ui_constTe,
// :1763-2072).
//
// Four implicit 5-point systems: Uix on the inner x faces, Uir on the inner
// r faces, Uip and n_i on the plasma cells.  Each is assembled in full by one
// kernel into the multigrid's level-0 arrays and solved with that system's
// own activity mask (the cell mask is put back afterwards).
//
// The reference writes its boundary operations as loops over fixed block
// ranges; here they come from the boundary segments, turned on the host
// into per-face operation bits (coefficient folds and zeros) and ordered
// ghost-write lists, kind by kind in the reference's loop order (low x walls,
// high x walls, low r walls, the axis, high r walls, the open face):
//  - a wall's normal velocity acts on the inner face next to the wall and
//    is reflected onto the wall face;
//  - the tangential rules act on the inner faces that touch the segment's
//    cells and write 0 to the face across the boundary row where that face
//    has no plasma on either side (a face with plasma on one side is a
//    wall face and keeps its normal value);
//  - CATHODE_SEG_UIX_BELOW moves the Uix tangential write of a high-r wall
//    to the row below its cells, unfiltered, as the reference's x1 loop does.
// Ghost lists keep the last write per target.  All expressions follow the
// source order of the restated loops, without multiply-add contraction.
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ndd_core.hpp"

namespace cathode {

enum { IP_RHOI = 0, IP_UIX, IP_UIR, IP_UIP, IP_FX, IP_FR, IP_FXW, IP_FRW, IP_PHI, IP_EX, IP_ER, IP_S, IP_NU, IP_BX,
       IP_BR, IP_RHOI_OLD, IP_UIX_OLD, IP_UIR_OLD, IP_UIP_OLD, IP_ZERO, IP_COUNT };

// operation bits of an inner x face (Uix)
enum : uint8_t { XO_WN = 1, XO_EN = 2, XO_SZ = 4, XO_NZ = 8, XO_WB = 16, XO_EB = 32 };
// of an inner r face (Uir)
enum : uint8_t { RO_WZ = 1, RO_EZ = 2, RO_SN = 4, RO_SA = 8, RO_NN = 16, RO_NO = 32 };
// of a plasma cell (Uip)
enum : uint8_t { CO_W = 1, CO_E = 2, CO_S = 4, CO_N = 8 };
// ghost entries {target, source, type, source row}; type | GT_BIAS for a biased wall
enum { GT_ZERO = 0, GT_WALL_NEG = 1, GT_WALL_POS = 2, GT_OPEN = 3, GT_COPY = 4, GT_SCALE = 5, GT_BIAS = 16 };

struct IonArgs {
    double *rhoi, *Uix, *Uir, *Uip, *fx, *fr, *fxw, *frw;
    const double *phi, *Ex, *Er, *S, *nu, *Bx, *Br;
    const double *rhoi_old, *Uix_old, *Uir_old, *Uip_old;
    const uint8_t *xop, *rop, *cop;
    const double* sol;  // the current solve's solution, pitch ld0
    int* neg;
    double dt, Ti, massi, V_bias, Te_rep, nu_star;
};

template <class G>
__device__ __forceinline__ bool inner_x(const G& g, const NddView& V, int i, int j)
{
    return (g.xf(V, i, j) & GF_TYPE) == GF_INNER;
}
template <class G>
__device__ __forceinline__ bool inner_r(const G& g, const NddView& V, int i, int j)
{
    return (g.rf(V, i, j) & GF_TYPE) == GF_INNER;
}

// sign sqrt(U_old^2 + max(2 e dphi / m_i, 0)) (fluidModule.hpp:231-233); the
// square is the correctly rounded product
__device__ __forceinline__ double ui_wall(double sign, double U_old, double dphi, double massi)
{
#pragma clang fp contract(off)
    const double v = sqrt(U_old * U_old + fmax(2.0 * K_E0 * dphi / massi, 0.0));
    return sign < 0.0 ? -v : v;
}

// kinetic wall flux per density of a drifting Maxwellian (:1502-1526)
__device__ __forceinline__ double ion_gpn(double Ti, double massi, double U, double sign)
{
#pragma clang fp contract(off)
    const double vth = sqrt(8.0 * K_BOLTZ * Ti / (M_PI * massi));
    const double um = sqrt(2.0 * K_BOLTZ * Ti / massi);
    const double u = U / (um + 1e-100);
    if (sign < 0.0) return -0.25 * vth * (exp(-u * u) + u * sqrt(M_PI) * (erf(u) - 1.0));
    return 0.25 * vth * (exp(-u * u) + u * sqrt(M_PI) * (erf(u) + 1.0));
}

__device__ __forceinline__ void put_row(const NddView& V, int i, int j, double* aP, double* aE, double* aW,
                                        double* aN, double* aS, double* b, double P, double E, double W, double N,
                                        double S, double rhs)
{
    const size_t k = IX0(i, j);
    aP[k] = P;
    aE[k] = E;
    aW[k] = W;
    aN[k] = N;
    aS[k] = S;
    b[k] = rhs;
}

// ------------------------------------------------------------------ Uix (:149-367)
template <class G>
__global__ void k_uix_assemble(G g, NddView V, IonArgs A, double* aP, double* aE, double* aW, double* aN, double* aS,
                               double* bb)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj) return;
    if (!inner_x(g, V, i, j)) {
        put_row(V, i, j, aP, aE, aW, aN, aS, bb, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double dx = V.dx, dr = V.dr, dt = A.dt;
    const double* r = V.r;
    const size_t k = IX(i, j), kw = IX(i - 1, j);
    const double rL = (r[j] + r[j - 1]) / 2.0;
    const double rR = (r[j] + r[j + 1]) / 2.0;
    const double qL = rL / r[j];
    const double qR = rR / r[j];
    const double Uix_Lx = (A.Uix_old[k] + A.Uix_old[kw]) / 2.0;
    const double Uix_Rx = (A.Uix_old[k] + A.Uix_old[IX(i + 1, j)]) / 2.0;
    const double Uir_Lr = (A.Uir_old[k] + A.Uir_old[kw]) / 2.0;
    const double Uir_Rr = (A.Uir_old[IX(i, j + 1)] + A.Uir_old[IX(i - 1, j + 1)]) / 2.0;
    const double rhoi_tmp = (A.rhoi[k] + A.rhoi[kw]) / 2.0;
    const double S_tmp = (A.S[k] + A.S[kw]) / 2.0;
    const double nu_tmp = (A.nu[k] + A.nu[kw]) / 2.0;
    const double dndx = (A.rhoi[k] - A.rhoi[kw]) / dx;
    const double Uip_tmp = (A.Uip_old[k] + A.Uip_old[kw]) / 2.0;
    const double Br_tmp = (A.Br[k] + A.Br[kw]) / 2.0;
    const double lorentz = -K_E0 * Uip_tmp * Br_tmp;
    double E = -(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx);
    double W = (Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx);
    double N = -qR * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr);
    double S = qL * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr);
    const double P = 1.0 / dt + S_tmp / (rhoi_tmp + 1e-100) + nu_tmp + E + W + N + S;
    double b = A.Uix_old[k] / dt + (K_E0 * A.Ex[k] + lorentz) / A.massi -
               K_BOLTZ * A.Ti / (A.massi * rhoi_tmp + 1e-100) * dndx;
    const int op = A.xop[k];
    if (op & (XO_WN | XO_EN)) {
        double dphiW = (A.phi[k] + A.phi[kw]) / 2.0;
        double dphiE = dphiW;
        if (op & XO_WB) dphiW = (A.phi[k] + A.phi[kw]) / 2.0 - A.V_bias;
        if (op & XO_EB) dphiE = (A.phi[k] + A.phi[kw]) / 2.0 - A.V_bias;
        if (op & XO_WN) {
            b = b + W * ui_wall(-1.0, A.Uix_old[k], dphiW, A.massi);
            W = 0.0;
        }
        if (op & XO_EN) {
            b = b + E * ui_wall(1.0, A.Uix_old[k], dphiE, A.massi);
            E = 0.0;
        }
    }
    if (op & XO_SZ) S = 0.0;
    if (op & XO_NZ) N = 0.0;
    put_row(V, i, j, aP, aE, aW, aN, aS, bb, P, E, W, N, S, b);
}

// ------------------------------------------------------------------ Uir (:614-808)
template <class G>
__global__ void k_uir_assemble(G g, NddView V, IonArgs A, double* aP, double* aE, double* aW, double* aN, double* aS,
                               double* bb)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj) return;
    if (!inner_r(g, V, i, j)) {
        put_row(V, i, j, aP, aE, aW, aN, aS, bb, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double dx = V.dx, dr = V.dr, dt = A.dt;
    const double* r = V.r;
    const size_t k = IX(i, j), ks = IX(i, j - 1);
    const double r_tmp = (r[j] + r[j - 1]) / 2.0;
    const double rL = (r[j - 1] + r[j - 2]) / 2.0;
    const double rR = (r[j + 1] + r[j]) / 2.0;
    const double Uix_Lx = (r[j] * A.Uix_old[k] + r[j - 1] * A.Uix_old[ks]) / (2.0 * r_tmp);
    const double Uix_Rx = (r[j] * A.Uix_old[IX(i + 1, j)] + r[j - 1] * A.Uix_old[IX(i + 1, j - 1)]) / (2.0 * r_tmp);
    const double Uir_Lr = (r_tmp * A.Uir_old[k] + rL * A.Uir_old[ks]) / (2.0 * r[j - 1]);
    const double Uir_Rr = (r_tmp * A.Uir_old[k] + rR * A.Uir_old[IX(i, j + 1)]) / (2.0 * r[j]);
    const double rhoi_tmp = (A.rhoi[k] + A.rhoi[ks]) / 2.0;
    const double S_tmp = (A.S[k] + A.S[ks]) / 2.0;
    const double nu_tmp = (A.nu[k] + A.nu[ks]) / 2.0;
    const double dndr = (A.rhoi[k] - A.rhoi[ks]) / dr;
    const double Bx_tmp = (r[j] * A.Bx[k] + r[j - 1] * A.Bx[ks]) / (2.0 * r_tmp);
    const double Uip_tmp = (r[j] * A.Uip_old[k] + r[j - 1] * A.Uip_old[ks]) / (2.0 * r_tmp);
    const double lorentz = K_E0 * Uip_tmp * Bx_tmp;
    const double centrifugal = -Uip_tmp * Uip_tmp / r_tmp;
    double E = -(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx);
    double W = (Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx);
    double N = -r[j] / r_tmp * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr);
    double S = r[j - 1] / r_tmp * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr);
    const double P = 1.0 / dt + S_tmp / (rhoi_tmp + 1e-100) + nu_tmp + E + W + N + S;
    double b = A.Uir_old[k] / dt + (K_E0 * A.Er[k] + lorentz) / A.massi -
               K_BOLTZ * A.Ti / (A.massi * rhoi_tmp + 1e-100) * dndr + centrifugal;
    const int op = A.rop[k];
    if (op & RO_WZ) W = 0.0;
    if (op & RO_EZ) E = 0.0;
    const double dphi = (A.phi[k] + A.phi[ks]) / 2.0;
    if (op & RO_SN) {
        b = b + S * ui_wall(-1.0, A.Uir_old[k], dphi, A.massi);
        S = 0.0;
    }
    if (op & RO_SA) S = 0.0;
    if (op & RO_NN) {
        b = b + N * ui_wall(1.0, A.Uir_old[k], dphi, A.massi);
        N = 0.0;
    }
    if (op & RO_NO) {
        b = b + N * r_tmp / rR * fmax(A.Uir_old[k], 0.0);
        N = 0.0;
    }
    put_row(V, i, j, aP, aE, aW, aN, aS, bb, P, E, W, N, S, b);
}

// ------------------------------------------------------------------ Uip (:1066-1205)
template <class G>
__global__ void k_uip_assemble(G g, NddView V, IonArgs A, double* aP, double* aE, double* aW, double* aN, double* aS,
                               double* bb)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj) return;
    if (!g.act(V, i, j)) {
        put_row(V, i, j, aP, aE, aW, aN, aS, bb, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double dx = V.dx, dr = V.dr, dt = A.dt;
    const double* r = V.r;
    const size_t k = IX(i, j);
    const double rL = (r[j] + r[j - 1]) / 2.0;
    const double rR = (r[j] + r[j + 1]) / 2.0;
    const double qL = rL / r[j];
    const double qR = rR / r[j];
    const double Uix_Lx = A.Uix_old[k];
    const double Uix_Rx = A.Uix_old[IX(i + 1, j)];
    const double Uir_Lr = A.Uir_old[k];
    const double Uir_Rr = A.Uir_old[IX(i, j + 1)];
    const double Uix_tmp = (A.Uix_old[IX(i + 1, j)] + A.Uix_old[k]) / 2.0;
    const double Uir_tmp = (rL * A.Uir_old[k] + rR * A.Uir_old[IX(i, j + 1)]) / (2.0 * r[j]);
    const double lorentz = K_E0 * (Uix_tmp * A.Br[k] - Uir_tmp * A.Bx[k]);
    const double centrifugal = A.Uip_old[k] * Uir_tmp / r[j];
    double E = -(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx);
    double W = (Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx);
    double N = -qR * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr);
    double S = qL * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr);
    double P = 1.0 / dt + A.S[k] / (A.rhoi[k] + 1e-100) + A.nu[k] + E + W + N + S;
    const double b = A.Uip_old[k] / dt + lorentz / A.massi + centrifugal;
    const int op = A.cop[k];
    if (op & CO_W) {
        P = P - W;
        W = 0.0;
    }
    if (op & CO_E) {
        P = P - E;
        E = 0.0;
    }
    if (op & CO_S) {
        P = P - S * r[j] / r[j - 1];
        S = 0.0;
    }
    if (op & CO_N) {
        P = P - N * r[j] / r[j + 1];
        N = 0.0;
    }
    put_row(V, i, j, aP, aE, aW, aN, aS, bb, P, E, W, N, S, b);
}

// ------------------------------------------------------------------ n_i (:1418-1579)
template <class G>
__global__ void k_rhoi_assemble(G g, NddView V, IonArgs A, double* aP, double* aE, double* aW, double* aN, double* aS,
                                double* bb)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj) return;
    if (!g.act(V, i, j)) {
        put_row(V, i, j, aP, aE, aW, aN, aS, bb, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double dx = V.dx, dr = V.dr, dt = A.dt;
    const double* r = V.r;
    const size_t k = IX(i, j), kE = IX(i + 1, j), kW = IX(i - 1, j), kN = IX(i, j + 1), kS = IX(i, j - 1);
    const double rL = (r[j] + r[j - 1]) / 2.0;
    const double rR = (r[j] + r[j + 1]) / 2.0;
    const double qL = rL / r[j];
    const double qR = rR / r[j];
    const double Uix_Lx = A.Uix[k], Uix_Rx = A.Uix[kE];
    const double Uir_Lr = A.Uir[k], Uir_Rr = A.Uir[kN];
    const double nLx = (A.nu[k] + A.nu[kW]) / 2.0;
    const double nRx = (A.nu[k] + A.nu[kE]) / 2.0;
    const double nLr = (A.nu[k] + A.nu[kS]) / 2.0;
    const double nRr = (A.nu[k] + A.nu[kN]) / 2.0;
    const double kT = K_BOLTZ * (A.Ti + A.Te_rep);
    const double D_Lx = kT / (A.massi * (nLx + A.nu_star) + 1e-100);
    const double D_Rx = kT / (A.massi * (nRx + A.nu_star) + 1e-100);
    const double D_Lr = kT / (A.massi * (nLr + A.nu_star) + 1e-100);
    const double D_Rr = kT / (A.massi * (nRr + A.nu_star) + 1e-100);
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, tN = g.rf(V, i, j + 1) & GF_TYPE;
    const double bLx = (double)(tL != GF_WLO), bRx = (double)(tR != GF_WHI);
    const double bLr = (double)(tS != GF_WLO) * (double)(tS != GF_AXIS);
    const double bRr = (double)(tN != GF_WHI);  // the open face is inner here, then folded
    const double gLx = tL == GF_WLO ? ion_gpn(A.Ti, A.massi, A.Uix[k], -1.0) : 0.0;
    const double gRx = tR == GF_WHI ? ion_gpn(A.Ti, A.massi, A.Uix[kE], 1.0) : 0.0;
    const double gLr = tS == GF_WLO ? ion_gpn(A.Ti, A.massi, A.Uir[k], -1.0) : 0.0;
    const double gRr = tN == GF_WHI ? ion_gpn(A.Ti, A.massi, A.Uir[kN], 1.0) : 0.0;
    double P = (Uix_Rx + fabs(Uix_Rx)) / (2.0 * dx) * bRx - (Uix_Lx - fabs(Uix_Lx)) / (2.0 * dx) * bLx +
               qR * (Uir_Rr + fabs(Uir_Rr)) / (2.0 * dr) * bRr - qL * (Uir_Lr - fabs(Uir_Lr)) / (2.0 * dr) * bLr +
               ((D_Rx * bRx + D_Lx * bLx) / (dx * dx) + (qR * D_Rr * bRr + qL * D_Lr * bLr) / (dr * dr)) - gLx / dx +
               gRx / dx - qL * gLr / dr + qR * gRr / dr + 1.0 / dt;
    const double E = (-(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx) + D_Rx / (dx * dx)) * bRx;
    const double W = ((Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx) + D_Lx / (dx * dx)) * bLx;
    double N = (-qR * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr) + qR * D_Rr / (dr * dr)) * bRr;
    const double S = (qL * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr) + qL * D_Lr / (dr * dr)) * bLr;
    const double* no = A.rhoi_old;
    const double divD = (D_Rx * no[kE] * bRx - (D_Rx * bRx + D_Lx * bLx) * no[k] + D_Lx * no[kW] * bLx) / (dx * dx) +
                        (qR * D_Rr * no[kN] * bRr - (qR * D_Rr * bRr + qL * D_Lr * bLr) * no[k] +
                         qL * D_Lr * no[kS] * bLr) / (dr * dr);
    const double b = no[k] / dt + A.S[k] - divD;
    if (tN == GF_OPEN) {  // zero gradient (:1571-1579)
        P = P - N;
        N = 0.0;
    }
    put_row(V, i, j, aP, aE, aW, aN, aS, bb, P, E, W, N, S, b);
}

// ------------------------------------------------------------------ stores, ghosts, post-processing
// which: 0 inner x faces into Uix, 1 inner r faces into Uir, 2 plasma cells into Uip
template <class G>
__global__ void k_ion_store(G g, NddView V, IonArgs A, int which)
{
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj) return;
    const double v = A.sol[IX0(i, j)];
    if (which == 0) {
        if (inner_x(g, V, i, j)) A.Uix[IX(i, j)] = v;
    } else if (which == 1) {
        if (inner_r(g, V, i, j)) A.Uir[IX(i, j)] = v;
    } else {
        if (g.act(V, i, j)) A.Uip[IX(i, j)] = v;
    }
}

// one ghost list (independent entries, one per target)
__global__ void k_ion_ghosts(NddView V, IonArgs A, const int* __restrict__ e, int n, int which)
{
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const size_t tgt = (size_t)e[4 * q], src = (size_t)e[4 * q + 1];
    const int type = e[4 * q + 2] & 15, bias = e[4 * q + 2] & GT_BIAS, jr = e[4 * q + 3];
    const double* r = V.r;
    if (which == 0) {
        double v = 0.0;
        if (type != GT_ZERO) {
            double dphi = (A.phi[src] + A.phi[src - V.ld]) / 2.0;
            if (bias) dphi = (A.phi[src] + A.phi[src - V.ld]) / 2.0 - A.V_bias;
            v = ui_wall(type == GT_WALL_NEG ? -1.0 : 1.0, A.Uix_old[src], dphi, A.massi);
        }
        A.Uix[tgt] = v;
    } else if (which == 1) {
        double v = 0.0;
        if (type == GT_WALL_NEG || type == GT_WALL_POS) {
            const double dphi = (A.phi[src] + A.phi[src - 1]) / 2.0;
            v = ui_wall(type == GT_WALL_NEG ? -1.0 : 1.0, A.Uir_old[src], dphi, A.massi);
        } else if (type == GT_OPEN) {
            const double r_tmp = (r[jr] + r[jr - 1]) / 2.0;
            const double rR = (r[jr + 1] + r[jr]) / 2.0;
            v = r_tmp / rR * fmax(A.Uir_old[src], 0.0);
        }
        A.Uir[tgt] = v;
    } else {
        const int tj = (int)(tgt % (size_t)V.ld);
        A.Uip[tgt] = type == GT_COPY ? A.Uip[src] : r[jr] / r[tj] * A.Uip[src];
    }
}

// the density from the solution: wall fluxes of the unreset value (:1590-1715),
// the open face's flux, the reset of negative values and the open ghost
// (:1727-1753)
template <class G>
__global__ void k_rhoi_post(G g, NddView V, IonArgs A)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj || !g.act(V, i, j)) return;
    const size_t k = IX(i, j), kE = IX(i + 1, j), kN = IX(i, j + 1);
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, tN = g.rf(V, i, j + 1) & GF_TYPE;
    const double s = A.sol[IX0(i, j)];
    if (tL == GF_WLO) A.fxw[k] = s * ion_gpn(A.Ti, A.massi, A.Uix[k], -1.0);
    if (tR == GF_WHI) A.fxw[kE] = s * ion_gpn(A.Ti, A.massi, A.Uix[kE], 1.0);
    if (tS == GF_WLO) A.frw[k] = s * ion_gpn(A.Ti, A.massi, A.Uir[k], -1.0);
    if (tN == GF_WHI) A.frw[kN] = s * ion_gpn(A.Ti, A.massi, A.Uir[kN], 1.0);
    if (tN == GF_OPEN) {
        A.fr[kN] = s * A.Uir[kN];
        A.frw[kN] = 0.0;
    }
    double out = s;
    if (s < 0.0) {
        out = fmax(A.rhoi_old[k], 0.0);
        atomicAdd(A.neg, 1);
    }
    A.rhoi[k] = out;
    if (tN == GF_OPEN) A.rhoi[kN] = out;
}

// correct_Ui_constTe (:1773-1822): x faces (axis 0) over the inner faces, r
// faces (axis 1) over the inner and the open faces
template <class G>
__global__ void k_ion_correct(G g, NddView V, IonArgs A, int axis)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj + 1) return;
    size_t k, kb;
    double h;
    double *U, *F;
    if (axis == 0) {
        if (j > V.nj || !inner_x(g, V, i, j)) return;
        k = IX(i, j);
        kb = IX(i - 1, j);
        h = V.dx;
        U = A.Uix;
        F = A.fx;
    } else {
        const int t = g.rf(V, i, j) & GF_TYPE;
        if (t != GF_INNER && t != GF_OPEN) return;
        k = IX(i, j);
        kb = IX(i, j - 1);
        h = V.dr;
        U = A.Uir;
        F = A.fr;
    }
    const double nu_tmp = (A.nu[k] + A.nu[kb]) / 2.0;
    const double D = K_BOLTZ * (A.Ti + A.Te_rep) / (A.massi * (nu_tmp + A.nu_star) + 1e-100);
    const double rhoi_tmp = (A.rhoi[k] + A.rhoi[kb]) / 2.0;
    const double dn_old = (A.rhoi_old[k] - A.rhoi_old[kb]) / h;
    const double dn = (A.rhoi[k] - A.rhoi[kb]) / h;
    const double u = U[k];
    const double flux = A.rhoi[kb] * (u + fabs(u)) / 2.0 + A.rhoi[k] * (u - fabs(u)) / 2.0;
    F[k] = flux - D * (dn - dn_old);
    U[k] = u - D / (rhoi_tmp + 1e-100) * (dn - dn_old);
}

// the multigrid's activity: 0 inner x faces, 1 inner r faces, 2 plasma cells
template <class G>
__global__ void k_ion_act(G g, NddView V, unsigned char* __restrict__ act0, int n0, int which)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n0) return;
    const int i = k / V.ld0, j = k - i * V.ld0;
    bool a = i >= 1 && i <= V.ni && j >= 1 && j <= V.nj;
    if (a) {
        if (which == 0) a = inner_x(g, V, i, j);
        else if (which == 1) a = inner_r(g, V, i, j);
        else a = g.act(V, i, j);
    }
    act0[k] = a ? 1 : 0;
}

// ------------------------------------------------------------------ host: the operation plan
struct IonPlan {
    std::vector<uint8_t> xop, rop, cop;
    std::vector<int> gx, gr, gc;  // 4 ints per entry
};

namespace {

using Entry = std::array<int, 4>;

// the entries in order with only the last write per target kept
std::vector<int> last_per_target(const std::vector<Entry>& in)
{
    std::map<int, size_t> last;
    for (size_t q = 0; q < in.size(); q++) last[in[q][0]] = q;
    std::vector<int> out;
    for (size_t q = 0; q < in.size(); q++) {
        if (last[in[q][0]] != q) continue;
        out.insert(out.end(), in[q].begin(), in[q].end());
    }
    return out;
}

bool build_ion_plan(int ni, int nj, const cathode_ndd_geometry& g, IonPlan& P, std::string& err)
{
    GeoTables T;
    if (!build_geo_tables(ni, nj, g, T, err)) return false;
    const int ld = nj + 2;
    const size_t n = (size_t)(ni + 2) * ld;
    auto at = [&](int i, int j) { return (int)((size_t)i * ld + j); };
    auto xin = [&](int i, int j) {
        return i >= 1 && i <= ni + 1 && j >= 1 && j <= nj && (T.xc[at(i, j)] & GF_TYPE) == GF_INNER;
    };
    auto rin = [&](int i, int j) {
        return i >= 1 && i <= ni && j >= 1 && j <= nj + 1 && (T.rc[at(i, j)] & GF_TYPE) == GF_INNER;
    };
    auto pl = [&](int i, int j) { return i >= 0 && i <= ni + 1 && j >= 0 && j <= nj + 1 && T.act[at(i, j)] != 0; };
    auto xghost = [&](int i, int j) { return !pl(i - 1, j) && !pl(i, j); };
    auto rghost = [&](int i, int j) { return !pl(i, j - 1) && !pl(i, j); };
    P.xop.assign(n, 0);
    P.rop.assign(n, 0);
    P.cop.assign(n, 0);
    std::vector<Entry> gxn, gxt, grt, grn, gc;
    auto fail = [&](int s, const char* what) {
        err = "ndd_ion_update: segment " + std::to_string(s) + ": " + what;
        return false;
    };
    const int order[6] = {CATHODE_SEG_X_LOW, CATHODE_SEG_X_HIGH, CATHODE_SEG_R_LOW, CATHODE_SEG_R_AXIS,
                          CATHODE_SEG_R_HIGH, CATHODE_SEG_R_OPEN};
    for (int kind : order)
        for (int s = 0; s < g.n_seg; s++) {
            const cathode_ndd_segment& q = g.seg[s];
            if (q.kind != kind) continue;
            const int c = q.cell;
            const int bias = (q.flag & CATHODE_SEG_BIAS) ? GT_BIAS : 0;
            if (kind == CATHODE_SEG_X_LOW || kind == CATHODE_SEG_X_HIGH) {
                const bool lo = kind == CATHODE_SEG_X_LOW;
                const int fin = lo ? c + 1 : c, fw = lo ? c : c + 1, gcol = lo ? c - 1 : c + 1;
                for (int j = q.lo; j <= q.hi; j++) {
                    if (!xin(fin, j)) return fail(s, "an x wall needs an inner x face next to it");
                    P.xop[at(fin, j)] |= lo ? (XO_WN | (bias ? XO_WB : 0)) : (XO_EN | (bias ? XO_EB : 0));
                    gxn.push_back({at(fw, j), at(fin, j), (lo ? GT_WALL_NEG : GT_WALL_POS) | bias, j});
                    P.cop[at(c, j)] |= lo ? CO_W : CO_E;
                    gc.push_back({at(gcol, j), at(c, j), GT_COPY, j});
                }
                for (int fj = q.lo; fj <= q.hi + 1; fj++) {
                    if (!rin(c, fj)) continue;
                    P.rop[at(c, fj)] |= lo ? RO_WZ : RO_EZ;
                    if (rghost(gcol, fj)) grt.push_back({at(gcol, fj), at(c, fj), GT_ZERO, fj});
                }
            } else if (kind == CATHODE_SEG_R_LOW || kind == CATHODE_SEG_R_AXIS) {
                for (int i = q.lo; i <= q.hi; i++) {
                    if (!rin(i, c + 1)) return fail(s, "a low r boundary needs an inner r face above it");
                    if (kind == CATHODE_SEG_R_AXIS) {
                        P.rop[at(i, c + 1)] |= RO_SA;
                        grn.push_back({at(i, c), at(i, c + 1), GT_ZERO, c + 1});
                        continue;
                    }
                    P.rop[at(i, c + 1)] |= RO_SN;
                    grn.push_back({at(i, c), at(i, c + 1), GT_WALL_NEG, c + 1});
                    P.cop[at(i, c)] |= CO_S;
                    gc.push_back({at(i, c - 1), at(i, c), GT_SCALE, c});
                }
                if (kind == CATHODE_SEG_R_AXIS) continue;
                for (int fi = q.lo; fi <= q.hi + 1; fi++) {
                    if (!xin(fi, c)) continue;
                    P.xop[at(fi, c)] |= XO_SZ;
                    if (xghost(fi, c - 1)) gxt.push_back({at(fi, c - 1), at(fi, c), GT_ZERO, c});
                }
            } else {  // R_HIGH, R_OPEN
                const bool open = kind == CATHODE_SEG_R_OPEN;
                for (int i = q.lo; i <= q.hi; i++) {
                    if (!rin(i, c)) return fail(s, "a high r boundary needs an inner r face below it");
                    P.rop[at(i, c)] |= open ? RO_NO : RO_NN;
                    grn.push_back({at(i, c + 1), at(i, c), open ? GT_OPEN : GT_WALL_POS, c});
                    P.cop[at(i, c)] |= CO_N;
                    gc.push_back({at(i, c + 1), at(i, c), GT_SCALE, c});
                }
                const bool below = !open && (q.flag & CATHODE_SEG_UIX_BELOW);
                for (int fi = q.lo; fi <= q.hi + 1; fi++) {
                    if (!xin(fi, c)) continue;
                    P.xop[at(fi, c)] |= XO_NZ;
                    if (below) gxt.push_back({at(fi, c - 1), at(fi, c), GT_ZERO, c});
                    else if (xghost(fi, c + 1)) gxt.push_back({at(fi, c + 1), at(fi, c), GT_ZERO, c});
                }
            }
        }
    std::vector<Entry> gx = gxn, gr = grt;
    gx.insert(gx.end(), gxt.begin(), gxt.end());
    gr.insert(gr.end(), grn.begin(), grn.end());
    P.gx = last_per_target(gx);
    P.gr = last_per_target(gr);
    P.gc = last_per_target(gc);
    return true;
}

}  // namespace

}  // namespace cathode

using namespace cathode;

// the block on host arrays (dev = false: staged through the handle's planes)
// or on the caller's device arrays (dev = true)
static int ion_entry(void* handle, const cathode_ion_params* p, cathode_ion_fields* f, cathode_ion_report* rep,
                     bool dev)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || !p || !f) {
            set_error("ndd_ion_update: null argument");
            return CATHODE_ERR_ARG;
        }
        if (!(p->dt_i > 0.0) || !(p->Ti > 0.0) || !(p->massi > 0.0) || !(p->rmax > 0.0) || !(p->Te_rep_eV >= 0.0) ||
            !std::isfinite(p->V_bias) || p->maxITR_Ui < 1 || p->maxITR_rhoi < 1 || !(p->tol_Ui > 0.0) ||
            !(p->tol_rhoi > 0.0) || (p->mode != CATHODE_MODE_FP64 && p->mode != CATHODE_MODE_MIXED)) {
            set_error("ndd_ion_update: invalid parameters");
            return CATHODE_ERR_ARG;
        }
        if (!f->rhoi || !f->Uix || !f->Uir || !f->Uip || !f->rhoUix || !f->rhoUir || !f->rhoUix_wall ||
            !f->rhoUir_wall || !f->phi) {
            set_error("ndd_ion_update: rhoi, the velocities, the fluxes and phi are required");
            return CATHODE_ERR_ARG;
        }
        auto t0 = std::chrono::steady_clock::now();
        const NddView& V = H->V;
        const size_t n = H->n;
        const int n0 = (V.ni + 2) * V.ld0;
        // the operation plan of the current geometry
        if (H->ion_plan_version != H->geo_version) {
            IonPlan P;
            std::string err;
            if (!build_ion_plan(V.ni, V.nj, H->geo_src, P, err)) {
                set_error(err);
                return CATHODE_ERR_ARG;
            }
            H->ion_ops.alloc(3 * n);
            std::vector<uint8_t> ops(3 * n);
            std::memcpy(ops.data(), P.xop.data(), n);
            std::memcpy(ops.data() + n, P.rop.data(), n);
            std::memcpy(ops.data() + 2 * n, P.cop.data(), n);
            const size_t ng = P.gx.size() + P.gr.size() + P.gc.size();
            H->ion_ghosts.alloc(ng + 4);
            std::vector<int> gl;
            gl.insert(gl.end(), P.gx.begin(), P.gx.end());
            gl.insert(gl.end(), P.gr.begin(), P.gr.end());
            gl.insert(gl.end(), P.gc.begin(), P.gc.end());
            CATHODE_HIP(hipMemcpyAsync(H->ion_ops.p, ops.data(), 3 * n, hipMemcpyHostToDevice, H->s));
            if (!gl.empty())
                CATHODE_HIP(hipMemcpyAsync(H->ion_ghosts.p, gl.data(), gl.size() * sizeof(int), hipMemcpyHostToDevice,
                                           H->s));
            CATHODE_HIP(hipStreamSynchronize(H->s));
            H->ion_ng[0] = (int)P.gx.size() / 4;
            H->ion_ng[1] = (int)P.gr.size() / 4;
            H->ion_ng[2] = (int)P.gc.size() / 4;
            H->ion_plan_version = H->geo_version;
        }
        H->ionp.alloc(IP_COUNT * n);
        H->ion_cnt.alloc(1);
        double* Q[IP_COUNT];
        for (int q = 0; q < IP_COUNT; q++) Q[q] = H->ionp.p + q * n;
        CATHODE_HIP(hipMemsetAsync(Q[IP_ZERO], 0, n * sizeof(double), H->s));
        CATHODE_HIP(hipMemsetAsync(H->ion_cnt.p, 0, sizeof(int), H->s));
        auto up = [&](int q, const double* src) -> const double* {
            if (!src) return Q[IP_ZERO];
            if (dev) {
                Q[q] = const_cast<double*>(src);
                return src;
            }
            CATHODE_HIP(hipMemcpyAsync(Q[q], src, n * sizeof(double), hipMemcpyHostToDevice, H->s));
            return Q[q];
        };
        up(IP_RHOI, f->rhoi);
        up(IP_UIX, f->Uix);
        up(IP_UIR, f->Uir);
        up(IP_UIP, f->Uip);
        up(IP_FX, f->rhoUix);
        up(IP_FR, f->rhoUir);
        up(IP_FXW, f->rhoUix_wall);
        up(IP_FRW, f->rhoUir_wall);
        IonArgs A{};
        A.rhoi = Q[IP_RHOI]; A.Uix = Q[IP_UIX]; A.Uir = Q[IP_UIR]; A.Uip = Q[IP_UIP];
        A.fx = Q[IP_FX]; A.fr = Q[IP_FR]; A.fxw = Q[IP_FXW]; A.frw = Q[IP_FRW];
        A.phi = up(IP_PHI, f->phi);
        A.Ex = up(IP_EX, f->Ex);
        A.Er = up(IP_ER, f->Er);
        A.S = up(IP_S, f->rate_ionize);
        A.nu = up(IP_NU, f->nui_m);
        A.Bx = up(IP_BX, f->Bx);
        A.Br = up(IP_BR, f->Br);
        A.rhoi_old = Q[IP_RHOI_OLD]; A.Uix_old = Q[IP_UIX_OLD]; A.Uir_old = Q[IP_UIR_OLD]; A.Uip_old = Q[IP_UIP_OLD];
        A.xop = H->ion_ops.p;
        A.rop = H->ion_ops.p + n;
        A.cop = H->ion_ops.p + 2 * n;
        A.neg = H->ion_cnt.p;
        A.dt = p->dt_i;
        A.Ti = p->Ti;
        A.massi = p->massi;
        A.V_bias = p->V_bias;
        A.Te_rep = p->Te_rep_eV * K_E0 / K_BOLTZ;                        // :1438
        A.nu_star = 1.0 / p->rmax * std::sqrt(K_BOLTZ * A.Te_rep / p->massi);  // :1439
        // the old values (:139-146)
        d2d(*H, {{Q[IP_RHOI_OLD], Q[IP_RHOI]}, {Q[IP_UIX_OLD], Q[IP_UIX]}, {Q[IP_UIR_OLD], Q[IP_UIR]},
                 {Q[IP_UIP_OLD], Q[IP_UIP]}});
        const dim3 gg = ggrid(V.ni, V.nj), gb = gblk();
        cathode_status st[4] = {};
        auto solve = [&](int which, int maxITR, double tol, cathode_status* out) {
            NDD_GEO_LAUNCH(*H, k_ion_act, dim3(cdiv(n0, 256)), dim3(256), V, H->mg.act(), n0, which);
            CATHODE_LAUNCH_CHECK();
            H->mg.prepared = false;
            H->mg.same_operator = false;
            H->mg.keep_guess = false;
            H->mg.expected_its = 0;
            H->mg.side_cnv = -INFINITY;
            H->mg.solve(maxITR, tol, p->mode, H->s, out);
            A.sol = H->mg.sol();
        };
        const int* gl = H->ion_ghosts.p;
        auto ghosts = [&](int which, const int* e, int cnt) {
            if (cnt <= 0) return;
            hipLaunchKernelGGL(k_ion_ghosts, dim3(cdiv(cnt, 256)), dim3(256), 0, H->s, V, A, e, cnt, which);
            CATHODE_LAUNCH_CHECK();
        };
        double *aP = H->mg.aP(), *aE = H->mg.aE(), *aW = H->mg.aW(), *aN = H->mg.aN(), *aS = H->mg.aS(),
               *rb = H->mg.rhs();
        // Uix
        NDD_GEO_LAUNCH(*H, k_uix_assemble, gg, gb, V, A, aP, aE, aW, aN, aS, rb);
        CATHODE_LAUNCH_CHECK();
        solve(0, p->maxITR_Ui, p->tol_Ui, &st[0]);
        NDD_GEO_LAUNCH(*H, k_ion_store, gg, gb, V, A, 0);
        ghosts(0, gl, H->ion_ng[0]);
        // Uir
        NDD_GEO_LAUNCH(*H, k_uir_assemble, gg, gb, V, A, aP, aE, aW, aN, aS, rb);
        CATHODE_LAUNCH_CHECK();
        solve(1, p->maxITR_Ui, p->tol_Ui, &st[1]);
        NDD_GEO_LAUNCH(*H, k_ion_store, gg, gb, V, A, 1);
        ghosts(1, gl + 4 * H->ion_ng[0], H->ion_ng[1]);
        // Uip
        NDD_GEO_LAUNCH(*H, k_uip_assemble, gg, gb, V, A, aP, aE, aW, aN, aS, rb);
        CATHODE_LAUNCH_CHECK();
        solve(2, p->maxITR_Ui, p->tol_Ui, &st[2]);
        NDD_GEO_LAUNCH(*H, k_ion_store, gg, gb, V, A, 2);
        ghosts(2, gl + 4 * (H->ion_ng[0] + H->ion_ng[1]), H->ion_ng[2]);
        // n_i
        NDD_GEO_LAUNCH(*H, k_rhoi_assemble, gg, gb, V, A, aP, aE, aW, aN, aS, rb);
        CATHODE_LAUNCH_CHECK();
        solve(2, p->maxITR_rhoi, p->tol_rhoi, &st[3]);
        NDD_GEO_LAUNCH(*H, k_rhoi_post, gg, gb, V, A);
        CATHODE_LAUNCH_CHECK();
        // correction
        NDD_GEO_LAUNCH(*H, k_ion_correct, gg, gb, V, A, 0);
        NDD_GEO_LAUNCH(*H, k_ion_correct, ggrid(V.ni, V.nj + 1), gb, V, A, 1);
        CATHODE_LAUNCH_CHECK();
        // the cell activity of the NDD step's solves is the plasma mask again
        NDD_GEO_LAUNCH(*H, k_ion_act, dim3(cdiv(n0, 256)), dim3(256), V, H->mg.act(), n0, 2);
        CATHODE_LAUNCH_CHECK();
        auto down = [&](double* dst, int q) {
            if (!dev) CATHODE_HIP(hipMemcpyAsync(dst, Q[q], n * sizeof(double), hipMemcpyDeviceToHost, H->s));
        };
        down(f->rhoi, IP_RHOI);
        down(f->Uix, IP_UIX);
        down(f->Uir, IP_UIR);
        down(f->Uip, IP_UIP);
        down(f->rhoUix, IP_FX);
        down(f->rhoUir, IP_FR);
        down(f->rhoUix_wall, IP_FXW);
        down(f->rhoUir_wall, IP_FRW);
        int neg = 0;
        CATHODE_HIP(hipMemcpyAsync(&neg, H->ion_cnt.p, sizeof(int), hipMemcpyDeviceToHost, H->s));
        CATHODE_HIP(hipStreamSynchronize(H->s));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rep) {
            std::memset(rep, 0, sizeof(*rep));
            rep->solve_Uix = st[0];
            rep->solve_Uir = st[1];
            rep->solve_Uip = st[2];
            rep->solve_rhoi = st[3];
            for (cathode_status* s : {&rep->solve_Uix, &rep->solve_Uir, &rep->solve_Uip, &rep->solve_rhoi})
                s->seconds = secs;
            rep->neg_rhoi = neg;
        }
        return CATHODE_OK;
    CATHODE_CATCH_END
}

extern "C" int cathode_ndd_ion_update(void* handle, const cathode_ion_params* p, cathode_ion_fields* f,
                                      cathode_ion_report* rep)
{
    return ion_entry(handle, p, f, rep, false);
}

extern "C" int cathode_ndd_ion_update_dev(void* handle, const cathode_ion_params* p, cathode_ion_fields* f,
                                          cathode_ion_report* rep)
{
    return ion_entry(handle, p, f, rep, true);
}
