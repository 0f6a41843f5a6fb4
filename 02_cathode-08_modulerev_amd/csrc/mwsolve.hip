// mwsolve.hip -- the coupled (E_x, E_r, E_phi) microwave solve of
// EmfieldModule::solve_Microwave (include/emfieldModule.hpp:2249-3321) and the
// E_r-only impedance test (solve_Microwave_impedanceTest, :5528-6141) on the
// device (gfx950).
//
// Rows (one per unknown face or cell, act_* = 1), multiplied by dx^2 dr^2:
//   aP E - sum_slots a_slot E_target = b,
// assembled on the union of the blocks' boxes from the cold magnetised-plasma
// conductivity (:2302-2454), then folded by the twelve boundary lists in the
// reference's order (:2457-2773).  A slot whose target is not an unknown is
// dropped (:2834-2991).  Slot targets:
//   E_x (i, j):  E W N S -> E_x(i+-1, j), E_x(i, j+-1); NE SE NW SW -> E_r(i, j+1),
//                E_r(i, j), E_r(i-1, j+1), E_r(i-1, j); EE WW -> E_phi(i, j), E_phi(i-1, j)
//   E_r (i, j):  E W N S; NE SE NW SW -> E_x(i+1, j), E_x(i+1, j-1), E_x(i, j), E_x(i, j-1);
//                NN SS -> E_phi(i, j), E_phi(i, j-1)
//   E_phi (i, j): E W N S; EE WW -> E_x(i+1, j), E_x(i, j); NN SS -> E_r(i, j+1), E_r(i, j)
// Every coupling reaches at most one x column to either side, so with the
// unknowns of x column i gathered into one block (slots: component position
// x ld + j; inactive slots are identity rows) the system is block
// tridiagonal: A_i x_{i-1} + B_i x_i + C_i x_{i+1} = f_i over the ni + 2
// columns.  It is solved directly by block cyclic reduction: at each level
// the odd rows of the active list are eliminated (B^-1 by Gauss-Jordan with
// partial pivoting, one workgroup per block; the Schur updates as batched
// complex GEMMs), the even rows recurse, and the eliminated rows are
// recovered on the way back.  The factors are kept, so refinement steps
// (sparse residual in FP64, solve with the same factors) cost GEMVs only.
// After the solve the ghost rules of the lists run (:3026-3264), and the
// impedance test normalises the forward power at the reference plane
// (:5917-6139).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <set>
#include <utility>
#include <vector>

#include "common.hpp"
#include "cplx.hpp"

namespace cathode {

namespace {

constexpr double MWS_EPS0 = 8.8541878128e-12;  // include/constants.hpp:7
constexpr double MWS_MU0 = 1.2566370614e-6;    // :8
constexpr double MWS_E0 = 1.60217662e-19;      // :6
constexpr int NSLOT = 14;
enum { SP = 0, SE, SW, SN, SS, SNE, SNW, SSE, SSW, SEE, SWW, SNN, SSS, SB };

// slot targets {component, di, dj}; component -1: no slot
__constant__ int c_target[3][NSLOT][3] = {
    {{0, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {1, 0, 1}, {1, -1, 1}, {1, 0, 0}, {1, -1, 0},
     {2, 0, 0}, {2, -1, 0}, {-1, 0, 0}, {-1, 0, 0}, {-1, 0, 0}},
    {{1, 0, 0}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 0}, {0, 0, 0}, {0, 1, -1}, {0, 0, -1},
     {-1, 0, 0}, {-1, 0, 0}, {2, 0, 0}, {2, 0, -1}, {-1, 0, 0}},
    {{2, 0, 0}, {2, 1, 0}, {2, -1, 0}, {2, 0, 1}, {2, 0, -1}, {-1, 0, 0}, {-1, 0, 0}, {-1, 0, 0}, {-1, 0, 0},
     {0, 1, 0}, {0, 0, 0}, {1, 0, 1}, {1, 0, 0}, {-1, 0, 0}},
};

struct MwsView {
    int ni, nj, ld, which;
    int pos[3];   // slot block of each component in a column (-1: not solved)
    int m;        // slots per column
    int n_bl;
    int j_flc0;
    double dx, dr, omega, masse, J1r_exc, r1, r2;
    int box[3][4 * CATHODE_MW_MAX_BLOCKS];   // 0 x faces, 1 r faces, 2 cells
    const double* r;
};

struct Fields {
    double *rhoe, *nu;
    const double *Bx, *Br, *epsr, *Er_dc;
    const int *act[3];
    cxd *J[3];   // excitation currents (never NULL here: a zero array stands in)
    cxd *E[3];
};

__device__ __host__ inline size_t qidx(int ld, int i, int j) { return (size_t)i * ld + j; }

__device__ inline bool in_union(const MwsView& V, int kind, int i, int j)
{
    for (int b = 0; b < V.n_bl; b++) {
        const int* x = V.box[kind] + 4 * b;
        if (i >= x[0] && i <= x[1] && j >= x[2] && j <= x[3]) return true;
    }
    return false;
}

__device__ inline bool active(const MwsView& V, const Fields& F, int c, int i, int j)
{
    if (c < 0 || V.pos[c] < 0 || i < 0 || i > V.ni + 1 || j < 0 || j > V.nj + 1) return false;
    return F.act[c][qidx(V.ld, i, j)] == 1;
}

// cold-plasma conductivity factors: h_x, h_r and coef = n e^2 / (m nu*) / (1 + h^2)
__device__ inline void cond(double n, double nu, double Bx, double Br, double w, double masse, cxd& hx, cxd& hr,
                            cxd& coef)
{
    const cxd EPS(1e-100, 1e-100);
    const cxd den = cxd(masse * nu, masse * w) + EPS;
    hx = cxd(MWS_E0 * Bx) / den;
    hr = cxd(MWS_E0 * Br) / den;
    const cxd h2 = hx * hx + hr * hr;
    coef = cxd(n * MWS_E0 * MWS_E0) / den / (cxd(1.0) + h2 + EPS);
}

__device__ inline cxd sc(cxd a, double s) { return cxd(a.re * s, a.im * s); }

#pragma clang fp contract(off)

// ---------------------------------------------------------------- assembly (:2302-2454)
__global__ __launch_bounds__(256) void k_mws_assemble(MwsView V, Fields F, cxd* __restrict__ co)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j > V.nj + 1 || i > V.ni + 1) return;
    const int ld = V.ld;
    const size_t n2 = (size_t)(V.ni + 2) * ld, q = qidx(ld, i, j);
    const double dx = V.dx, dr = V.dr, w = V.omega;
    const double* r = V.r;
    const cxd iwm(0.0, w * MWS_MU0);
    const cxd bsrc(0.0, -w * MWS_MU0 * dx * dx * dr * dr);
    auto S = [&](int c, int k) -> cxd& { return co[((size_t)c * NSLOT + k) * n2 + q]; };
    cxd hx, hr, coef;
    if (V.pos[0] >= 0 && in_union(V, 0, i, j)) {
        const size_t qW = q - ld;
        const double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100), qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
        cond((F.rhoe[q] + F.rhoe[qW]) / 2.0, (F.nu[q] + F.nu[qW]) / 2.0, (F.Bx[q] + F.Bx[qW]) / 2.0,
             (F.Br[q] + F.Br[qW]) / 2.0, w, V.masse, hx, hr, coef);
        const double ep = (F.epsr[q] + F.epsr[qW]) / 2.0;
        const cxd sxr = coef * (hx * hr), sxp = coef * hr, sxx = coef * (cxd(1.0) + hx * hx);
        const double h4 = dx * dx * dr * dr;
        S(0, SP) = cxd(2.0 * dx * dx + 2.0 * dr * dr - MWS_EPS0 * ep * MWS_MU0 * w * w * dx * dx * dr * dr) +
                   sc(iwm * sxx, h4);
        S(0, SE) = cxd(dr * dr);
        S(0, SW) = cxd(dr * dr);
        S(0, SN) = cxd(qR * dx * dx);
        S(0, SS) = cxd(qL * dx * dx);
        S(0, SB) = bsrc * F.J[0][q];
        S(0, SEE) = sc(-(iwm * sxp), 0.5 * h4);
        S(0, SWW) = sc(-(iwm * sxp), 0.5 * h4);
        S(0, SNE) = sc(-(iwm * sxr), 0.25 * qR * h4);
        S(0, SNW) = sc(-(iwm * sxr), 0.25 * qR * h4);
        S(0, SSE) = sc(-(iwm * sxr), 0.25 * qL * h4);
        S(0, SSW) = sc(-(iwm * sxr), 0.25 * qL * h4);
    }
    if (V.pos[1] >= 0 && in_union(V, 1, i, j)) {
        const size_t qS = q - 1;
        const double rt = (r[j - 1] + r[j]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0, rL = (r[j - 2] + r[j - 1]) / 2.0;
        cond((F.rhoe[q] + F.rhoe[qS]) / 2.0, (F.nu[q] + F.nu[qS]) / 2.0,
             (r[j] * F.Bx[q] + r[j - 1] * F.Bx[qS]) / (2.0 * rt), (r[j] * F.Br[q] + r[j - 1] * F.Br[qS]) / (2.0 * rt),
             w, V.masse, hx, hr, coef);
        const double ep = (F.epsr[q] + F.epsr[qS]) / 2.0;
        const cxd srr = coef * (cxd(1.0) + hr * hr), srp = coef * (-hx), srx = coef * (hx * hr);
        const double h4 = dx * dx * dr * dr;
        S(1, SP) = cxd(rt * (1.0 / r[j] + 1.0 / r[j - 1]) * dx * dx + 2.0 * dr * dr -
                       MWS_EPS0 * ep * MWS_MU0 * w * w * dx * dx * dr * dr) +
                   sc(iwm * srr, h4);
        S(1, SE) = cxd(dr * dr);
        S(1, SW) = cxd(dr * dr);
        S(1, SN) = cxd(rR / r[j] * dx * dx);
        S(1, SS) = cxd(rL / r[j - 1] * dx * dx);
        S(1, SB) = bsrc * F.J[1][q];
        S(1, SNN) = sc(-(iwm * srp), 0.5 * r[j] / rt * h4);
        S(1, SSS) = sc(-(iwm * srp), 0.5 * r[j - 1] / rt * h4);
        S(1, SNE) = sc(-(iwm * srx), 0.25 * r[j] / rt * h4);
        S(1, SNW) = sc(-(iwm * srx), 0.25 * r[j] / rt * h4);
        S(1, SSE) = sc(-(iwm * srx), 0.25 * r[j - 1] / rt * h4);
        S(1, SSW) = sc(-(iwm * srx), 0.25 * r[j - 1] / rt * h4);
    }
    if (V.pos[2] >= 0 && in_union(V, 2, i, j)) {
        const double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100), qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
        cond(F.rhoe[q], F.nu[q], F.Bx[q], F.Br[q], w, V.masse, hx, hr, coef);
        const cxd spr = coef * hx, spp = coef, spx = coef * (-hr);
        const double h4 = dx * dx * dr * dr;
        S(2, SP) = cxd(2.0 * dx * dx + 2.0 * dr * dr + dx * dx * dr * dr / (r[j] * r[j]) -
                       MWS_EPS0 * F.epsr[q] * MWS_MU0 * w * w * dx * dx * dr * dr) +
                   sc(iwm * spp, h4);
        S(2, SE) = cxd(dr * dr);
        S(2, SW) = cxd(dr * dr);
        S(2, SN) = cxd(qR * dx * dx);
        S(2, SS) = cxd(qL * dx * dx);
        S(2, SB) = bsrc * F.J[2][q];
        S(2, SEE) = sc(-(iwm * spx), 0.5 * h4);
        S(2, SWW) = sc(-(iwm * spx), 0.5 * h4);
        S(2, SNN) = sc(-(iwm * spr), 0.5 * qR * h4);
        S(2, SSS) = sc(-(iwm * spr), 0.5 * qL * h4);
    }
}

// ---------------------------------------------------------------- boundary lists
struct ListSet {
    const int* e;    // entries (i, j, s) of all twelve lists, x W E S N, r W E S N, p W E S N
    int off[13];
};

__device__ inline double c0_of(double epsr) { return sqrt(1.0 / (MWS_EPS0 * epsr * MWS_MU0)); }

// first-order absorbing reflection factor on an x face (:2584-2586) and its denominator
__device__ inline cxd refl_x(double c0, double w, double dx, double* deno_out = nullptr)
{
    const double a = 2.0 * c0, b = w * dx;
    const double deno = a * a + b * b;
    if (deno_out) *deno_out = deno;
    return cxd(a * a - b * b, -4.0 * c0 * w * dx) / cxd(deno + 1e-100);
}

// the same at the high-r wall (:2540-2550)
__device__ inline cxd refl_r(double c0, double w, double dr, double rj, double rj1)
{
    const double rt = (rj1 + rj) / 2.0;
    const double qL = rj / (rt + 1e-100), qR = rj1 / (rt + 1e-100);
    const double t = 2.0 * c0 * sqrt(qR), u = w * dr * qR;
    const double deno = t * t + u * u;
    const cxd R((2.0 * c0) * (2.0 * c0) * sqrt(qL * qR) - (w * dr) * (w * dr) * qL * qR,
                -2.0 * c0 * w * dr * (qL * sqrt(qR) + qR * sqrt(qL)));
    return R / cxd(deno + 1e-100);
}

// TEM excitation of E_r (:2606-2620): ghost = Rd E + Sd
__device__ inline void excite(const MwsView& V, double epsr, double rj, double rjm, cxd& Rd, cxd& Sd)
{
    const double c0 = c0_of(epsr);
    const cxd deno = cxd(2.0 * c0, V.omega * V.dx) + cxd(1e-100);
    const cxd R(2.0 * c0, -V.omega * V.dx);
    const double rt = (rj + rjm) / 2.0;
    const double Efw = V.J1r_exc / rt / log(V.r2 / V.r1) * 1e3;
    const cxd S = cxd(0.0, 4.0 * V.omega * V.dx) * cxd(Efw);
    Rd = R / deno;
    Sd = S / deno;
}

// The folds (:2457-2773).  One workgroup; the lists run in order, the entries
// of one list side by side (the host checks that no two share a face or cell).
__global__ __launch_bounds__(256) void k_mws_fold(MwsView V, Fields F, ListSet L, cxd* __restrict__ co)
{
    const int ld = V.ld;
    const size_t n2 = (size_t)(V.ni + 2) * ld;
    const double dx = V.dx, dr = V.dr, w = V.omega, *r = V.r;
    for (int list = 0; list < 12; list++) {
        const int c = list / 4, side = list % 4;
        if (V.pos[c] < 0) continue;   // uniform over the workgroup
        for (int k = L.off[list] + threadIdx.x; k < L.off[list + 1]; k += blockDim.x) {
            const int i = L.e[3 * k], j = L.e[3 * k + 1], s = L.e[3 * k + 2];
            const size_t q = qidx(ld, i, j);
            auto A = [&](int slot) -> cxd& { return co[((size_t)c * NSLOT + slot) * n2 + q]; };
            cxd &P = A(SP), &E = A(SE), &W = A(SW), &N = A(SN), &Sd = A(SS), &NE = A(SNE), &NW = A(SNW);
            cxd &SEs = A(SSE), &SWs = A(SSW), &Bv = A(SB);
            if (c == 0) {
                const double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                const double qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                const double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                if (side == 0) {
                    if (s == 0) {
                        W = cxd(0.0);
                    } else if (s == 2) {
                        P = P - W;
                        NW = NW + sc(sc(W, dx / dr), qR);
                        SWs = SWs - sc(sc(W, dx / dr), qL);
                        W = cxd(0.0);
                    } else if (s == 3) {
                        const double deno = 1.0 + dx / dr * qL;
                        P = P - W / cxd(deno);
                        NW = NW + sc(W, dx / dr * qR) / cxd(deno);
                        W = cxd(0.0);
                    } else if (s == 4) {
                        const double deno = 1.0 + dx / dr * qR;
                        P = P - W / cxd(deno);
                        SWs = SWs - sc(W, dx / dr * qL) / cxd(deno);
                        W = cxd(0.0);
                    }
                } else if (side == 1) {
                    if (s == 0) {
                        E = cxd(0.0);
                    } else if (s == 2) {
                        P = P - E;
                        NE = NE - sc(sc(E, dx / dr), qR);
                        SEs = SEs + sc(sc(E, dx / dr), qL);
                        E = cxd(0.0);
                    } else if (s == 4) {
                        const double deno = 1.0 + dx / dr * qR;
                        P = P - E / cxd(deno);
                        SEs = SEs + sc(E, dx / dr * qL) / cxd(deno);
                        E = cxd(0.0);
                    }
                } else if (side == 2) {
                    if (s == 0) {
                        P = P + sc(Sd, r[j] / r[j - 1]);
                        Sd = cxd(0.0);
                    } else if (s == 6) {
                        P = P - sc(Sd, dr / dx * r[j] / rL);
                        W = W + sc(Sd, -dr / dx * r[j] / rL);
                        NW = NW + sc(Sd, rR / rL);
                        Sd = cxd(0.0);
                    }
                } else {
                    if (s == 0) {
                        P = P + sc(N, r[j] / r[j + 1]);
                        N = cxd(0.0);
                    } else if (s == 1) {
                        const cxd R = refl_r(c0_of((F.epsr[q] + F.epsr[q - ld]) / 2.0), w, dr, r[j], r[j + 1]);
                        P = P - N * R;
                        N = cxd(0.0);
                    } else if (s == 5) {
                        P = P - sc(N, dr / dx * r[j] / rR);
                        E = E + sc(N, -dr / dx * r[j] / rR);
                        SEs = SEs + sc(N, rL / rR);
                        N = cxd(0.0);
                    } else if (s == 6) {
                        P = P - sc(N, dr / dx * r[j] / rR);
                        W = W + sc(N, -dr / dx * r[j] / rR);
                        SWs = SWs + sc(N, -rL / rR);
                        N = cxd(0.0);
                    }
                }
            } else if (c == 1) {
                const double ep = (F.epsr[q] + F.epsr[q - 1]) / 2.0;
                if (side == 0) {
                    if (s == 0) {
                        P = P + W;
                        W = cxd(0.0);
                    } else if (s == 1) {
                        P = P - W * refl_x(c0_of(ep), w, dx);
                        W = cxd(0.0);
                    } else if (s == 5) {
                        const double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                        P = P - sc(W, dx / dr * rL / r[j]);
                        N = N + sc(W, -dx / dr * rR / r[j]);
                        NE = NE + sc(W, -1.0);
                        W = cxd(0.0);
                    } else if (s == 6) {
                        const double rR = (r[j] + r[j - 1]) / 2.0, rL = (r[j - 1] + r[j - 2]) / 2.0;
                        P = P - sc(W, dx / dr * rR / r[j - 1]);
                        Sd = Sd + sc(W, -dx / dr * rL / r[j - 1]);
                        SEs = SEs + sc(W, 1.0);
                        W = cxd(0.0);
                    } else if (s == 7) {
                        cxd Rd, Sx;
                        excite(V, ep, r[j], r[j - 1], Rd, Sx);
                        P = P - W * Rd;
                        Bv = Bv + W * Sx;
                        W = cxd(0.0);
                    }
                } else if (side == 1) {
                    if (s == 0) {
                        P = P + E;
                        E = cxd(0.0);
                    } else if (s == 1) {
                        P = P - E * refl_x(c0_of(ep), w, dx);
                        E = cxd(0.0);
                    } else if (s == 5) {
                        const double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                        P = P - sc(E, dx / dr * rL / r[j]);
                        N = N + sc(E, -dx / dr * rR / r[j]);
                        NW = NW + sc(E, 1.0);
                        E = cxd(0.0);
                    }
                } else if (side == 2) {
                    const double rR = (r[j] + r[j - 1]) / 2.0, rL = (r[j - 1] + r[j - 2]) / 2.0;
                    if (s == 2) {
                        P = P - sc(Sd, rR / rL);
                        SEs = SEs + sc(Sd, dr / dx * r[j - 1] / rL);
                        SWs = SWs - sc(Sd, dr / dx * r[j - 1] / rL);
                        Sd = cxd(0.0);
                    } else if (s == 3) {
                        double deno = dr / dx * r[j - 1] / rL;
                        if (r[j] < (double)V.j_flc0) deno = 1.0;
                        P = P - sc(Sd, rR / rL / deno);
                        NE = NE + sc(Sd, dr / dx * r[j - 1] / rL / deno);
                        Sd = cxd(0.0);
                    }
                } else {
                    const double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                    const double deno = dr / dx * r[j] / rR;
                    if (s == 2) {
                        P = P - sc(N, rL / rR);
                        NE = NE - sc(N, dr / dx * r[j] / rR);
                        NW = NW + sc(N, dr / dx * r[j] / rR);
                        N = cxd(0.0);
                    } else if (s == 3) {
                        P = P - sc(N, rL / rR / deno);
                        SEs = SEs - sc(N, dr / dx * r[j] / rR / deno);
                        N = cxd(0.0);
                    } else if (s == 4) {
                        P = P - sc(N, rL / rR / deno);
                        SWs = SWs + sc(N, dr / dx * r[j] / rR / deno);
                        N = cxd(0.0);
                    }
                }
            } else {
                if (side == 0) {
                    if (s == 0) {
                        P = P + W;
                        W = cxd(0.0);
                    } else if (s == 1) {
                        P = P - W * refl_x(c0_of(F.epsr[q]), w, dx);
                        W = cxd(0.0);
                    }
                } else if (side == 1) {
                    if (s == 0) {
                        P = P + E;
                        E = cxd(0.0);
                    }
                } else if (side == 2) {
                    if (s == 0) {
                        P = P + sc(Sd, r[j] / r[j - 1]);
                        Sd = cxd(0.0);
                    }
                } else {
                    if (s == 0) {
                        P = P + sc(N, r[j] / r[j + 1]);
                        N = cxd(0.0);
                    } else if (s == 1) {
                        P = P - N * refl_r(c0_of(F.epsr[q]), w, dr, r[j], r[j + 1]);
                        N = cxd(0.0);
                    }
                }
            }
        }
        __syncthreads();
    }
}

// The ghost rules (:3026-3264).  One workgroup, lists in order.
__global__ __launch_bounds__(256) void k_mws_post(MwsView V, Fields F, ListSet L)
{
    const int ld = V.ld;
    const double dx = V.dx, dr = V.dr, w = V.omega, *r = V.r;
    cxd *Ex = F.E[0], *Er = F.E[1], *Ep = F.E[2];
    for (int list = 0; list < 12; list++) {
        const int c = list / 4, side = list % 4;
        if (V.pos[c] < 0) continue;
        for (int k = L.off[list] + threadIdx.x; k < L.off[list + 1]; k += blockDim.x) {
            const int i = L.e[3 * k], j = L.e[3 * k + 1], s = L.e[3 * k + 2];
            const size_t q = qidx(ld, i, j);
            if (c == 0) {
                const double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                const double qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                if (side == 0) {
                    if (s == 0)
                        Ex[q - ld] = cxd(0.0);
                    else if (s == 2)
                        Ex[q - ld] = Ex[q] + sc(sc(Er[q - ld + 1], qR) - sc(Er[q - ld], qL), dx / dr);
                    else if (s == 3) {
                        const double deno = 1.0 + dx / dr * qL;
                        Ex[q - ld] = Ex[q] / cxd(deno) + sc(sc(Er[q - ld + 1], qR), dx / dr) / cxd(deno);
                    } else if (s == 4) {
                        const double deno = 1.0 + dx / dr * qR;
                        Ex[q - ld] = Ex[q] / cxd(deno) - sc(sc(Er[q - ld], qL), dx / dr) / cxd(deno);
                    }
                } else if (side == 1) {
                    if (s == 0)
                        Ex[q + ld] = cxd(0.0);
                    else if (s == 2)
                        Ex[q + ld] = Ex[q] - sc(sc(Er[q + 1], qR) - sc(Er[q], qL), dx / dr);
                    else if (s == 4) {
                        const double deno = 1.0 + dx / dr * qR;
                        Ex[q + ld] = Ex[q] / cxd(deno) + sc(Er[q], dx / dr * qL) / cxd(deno);
                    }
                } else if (side == 2) {
                    if (s == 0) Ex[q - 1] = sc(Ex[q], -r[j] / r[j - 1]);
                } else {
                    if (s == 0)
                        Ex[q + 1] = sc(Ex[q], -r[j] / r[j + 1]);
                    else if (s == 1)
                        Ex[q + 1] = refl_r(c0_of((F.epsr[q] + F.epsr[q - ld]) / 2.0), w, dr, r[j], r[j + 1]) * Ex[q];
                }
            } else if (c == 1) {
                const double ep = (F.epsr[q] + F.epsr[q - 1]) / 2.0;
                if (side == 0) {
                    if (s == 0)
                        Er[q - ld] = -Er[q];
                    else if (s == 1)
                        Er[q - ld] = refl_x(c0_of(ep), w, dx) * Er[q];
                    else if (s == 7) {
                        cxd Rd, Sx;
                        excite(V, ep, r[j], r[j - 1], Rd, Sx);
                        Er[q - ld] = Rd * Er[q] + Sx;
                    }
                } else if (side == 1) {
                    if (s == 0)
                        Er[q + ld] = -Er[q];
                    else if (s == 1) {
                        // the reference scales the static field gr.Er by R / deno (:3162)
                        double deno;
                        const cxd R = refl_x(c0_of(ep), w, dx, &deno);
                        const double edc = F.Er_dc ? F.Er_dc[q] : 0.0;
                        Er[q + ld] = R / cxd(deno + 1e-100) * cxd(edc);
                    }
                } else if (side == 2) {
                    const double rR = (r[j] + r[j - 1]) / 2.0, rL = (r[j - 1] + r[j - 2]) / 2.0;
                    if (s == 2)
                        Er[q - 1] = sc(Er[q], rR / rL) + sc(Ex[q + ld - 1] - Ex[q - 1], dr / dx * r[j - 1] / rL);
                    else if (s == 3) {
                        double deno = dr / dx * r[j - 1] / rL;
                        if (r[j] < (double)V.j_flc0) deno = 1.0;
                        Er[q - 1] = sc(Er[q], rR / rL) / cxd(deno) + sc(Ex[q + ld - 1], dr / dx * r[j - 1] / rL) / cxd(deno);
                    }
                } else {
                    const double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                    const double deno = dr / dx * r[j] / rR;
                    if (s == 2)
                        Er[q + 1] = sc(Er[q], rL / rR) - sc(Ex[q + ld] - Ex[q], dr / dx * r[j] / rR);
                    else if (s == 3)
                        Er[q + 1] = sc(Er[q], rL / rR) / cxd(deno) - sc(Ex[q + ld], dr / dx * r[j] / rR) / cxd(deno);
                    else if (s == 4)
                        Er[q + 1] = sc(Er[q], rL / rR) / cxd(deno) + sc(Ex[q], dr / dx * r[j] / rR) / cxd(deno);
                }
            } else {
                if (side == 0) {
                    if (s == 0)
                        Ep[q - ld] = -Ep[q];
                    else if (s == 1)
                        Ep[q - ld] = refl_x(c0_of(F.epsr[q]), w, dx) * Ep[q];
                } else if (side == 1) {
                    Ep[q + ld] = -Ep[q];
                } else if (side == 2) {
                    Ep[q - 1] = sc(Ep[q], -r[j] / r[j - 1]);
                } else {
                    if (s == 0)
                        Ep[q + 1] = sc(Ep[q], -r[j] / r[j + 1]);
                    else if (s == 1)
                        Ep[q + 1] = refl_r(c0_of(F.epsr[q]), w, dr, r[j], r[j + 1]) * Ep[q];
                }
            }
        }
        __syncthreads();
    }
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------- block system
// Column i's blocks: B_i (diagonal), A_i (to column i-1), C_i (to column i+1),
// m x m row-major, and the right-hand side f_i.  Inactive slots: identity rows.
// A_i and C_i of level 0 sit at AC + 2 i mm and AC + (2 i + 1) mm.
__global__ __launch_bounds__(256) void k_mws_blocks(MwsView V, Fields F, const cxd* __restrict__ co,
                                                   cxd* __restrict__ Bb, cxd* __restrict__ AC, cxd* __restrict__ f)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    const int m = V.m, ld = V.ld;
    if (s >= m) return;
    const size_t mm = (size_t)m * m, n2 = (size_t)(V.ni + 2) * ld;
    const int blk = s / ld, j = s % ld;
    int c = -1;
    for (int k = 0; k < 3; k++)
        if (V.pos[k] == blk) c = k;
    cxd* Bi = Bb + (size_t)i * mm + (size_t)s * m;
    if (c < 0 || !active(V, F, c, i, j)) {
        Bi[s] = cxd(1.0);
        f[(size_t)i * m + s] = cxd(0.0);
        return;
    }
    const size_t q = qidx(ld, i, j);
    Bi[s] = co[((size_t)c * NSLOT + SP) * n2 + q];
    f[(size_t)i * m + s] = co[((size_t)c * NSLOT + SB) * n2 + q];
    for (int k = 1; k < SB; k++) {
        const int tc = c_target[c][k][0];
        if (tc < 0) continue;
        if (V.which != CATHODE_MW_COUPLED && k > SS) continue;   // the impedance test keeps the five-point part
        const int ti = i + c_target[c][k][1], tj = j + c_target[c][k][2];
        if (!active(V, F, tc, ti, tj)) continue;
        const int col = V.pos[tc] * ld + tj;
        const cxd a = -co[((size_t)c * NSLOT + k) * n2 + q];
        if (ti == i)
            Bi[col] = a;
        else if (ti == i - 1)
            AC[(size_t)(2 * i) * mm + (size_t)s * m + col] = a;
        else
            AC[(size_t)(2 * i + 1) * mm + (size_t)s * m + col] = a;
    }
}

// ---------------------------------------------------------------- dense block inverses
// In-place inverses of a batch of m x m blocks: blocked Gauss-Jordan with
// partial (row) pivoting.  For each panel of nb columns K = [k0, k0 + nb):
//  k_gj_panel   one workgroup per block holds the m x nb panel in LDS and runs
//               the Gauss-Jordan steps of its columns (pivot = largest |a_ik|,
//               i >= k; row interchange; elimination above and below).  The
//               panel then holds the columns K of the step transform T, which
//               differs from the identity only there; the panel's row
//               interchanges are composed into a gather map of rows k0..m-1.
//  k_gj_rows    the other columns: rows gathered through the map, rows K
//               copied out (tmp) and set to 0;
//  k_gj_update  those columns += T[:, K] tmp (off K a row keeps its own value,
//               on K it is T[K, K] tmp).
// At the end the row interchanges come back as column interchanges
// (k_gj_cols), as in the unblocked in-place inverse.  A zero pivot sets *bad.
__global__ __launch_bounds__(256) void k_gj_panel(cxd* const* __restrict__ mats, int m, int k0, int nbk, int nb,
                                                 int* __restrict__ gmap, int* __restrict__ perm, int* __restrict__ bad)
{
    extern __shared__ unsigned char lds_raw[];
    cxd* P = reinterpret_cast<cxd*>(lds_raw);         // m x nb panel
    cxd* fcol = P + (size_t)m * nb;                   // column kk before the step
    cxd* srow = fcol + m;                             // scaled pivot row (nb)
    int* map = reinterpret_cast<int*>(srow + nb);     // gather map of rows k0..m-1
    __shared__ double best_v[4];
    __shared__ int best_i[4];
    __shared__ int piv_row;
    cxd* a = mats[blockIdx.x];
    int* gm = gmap + (size_t)blockIdx.x * m;
    int* pm = perm + (size_t)blockIdx.x * m;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    for (int i = wid; i < m; i += 4)
        for (int c = lane; c < nbk; c += 64) P[(size_t)i * nb + c] = a[(size_t)i * m + k0 + c];
    for (int r = k0 + t; r < m; r += 256) map[r - k0] = r;
    __syncthreads();
    for (int kk = 0; kk < nbk; kk++) {
        const int k = k0 + kk;
        double bv = -1.0;
        int bi = m;
        for (int i = k + t; i < m; i += 256) {
            const double v = abs2v(P[(size_t)i * nb + kk]);
            if (v > bv) {
                bv = v;
                bi = i;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            best_v[wid] = bv;
            best_i[wid] = bi;
        }
        __syncthreads();
        if (t == 0) {
            double v = best_v[0];
            int p = best_i[0];
            for (int w = 1; w < 4; w++)
                if (best_v[w] > v || (best_v[w] == v && best_i[w] < p)) {
                    v = best_v[w];
                    p = best_i[w];
                }
            if (!(v > 0.0)) {
                *bad = 1;
                p = k;
            }
            piv_row = p;
            pm[k] = p;
            const int tmp = map[k - k0];
            map[k - k0] = map[p - k0];
            map[p - k0] = tmp;
        }
        __syncthreads();
        const int p = piv_row;
        if (p != k && t < nbk) {
            const cxd x = P[(size_t)k * nb + t];
            P[(size_t)k * nb + t] = P[(size_t)p * nb + t];
            P[(size_t)p * nb + t] = x;
        }
        __syncthreads();
        const cxd akk = P[(size_t)k * nb + kk];
        const cxd piv = abs2v(akk) > 0.0 ? cxd(1.0) / akk : cxd(0.0);
        if (t < nbk) srow[t] = t == kk ? piv : P[(size_t)k * nb + t] * piv;
        for (int i = t; i < m; i += 256) fcol[i] = P[(size_t)i * nb + kk];
        __syncthreads();
        for (int e = t; e < m * nbk; e += 256) {
            const int i = e / nbk, c = e - i * nbk;
            cxd& x = P[(size_t)i * nb + c];
            if (i == k)
                x = srow[c];
            else if (c == kk)
                x = -(fcol[i] * piv);
            else
                x = x - fcol[i] * srow[c];
        }
        __syncthreads();
    }
    for (int i = wid; i < m; i += 4)
        for (int c = lane; c < nbk; c += 64) a[(size_t)i * m + k0 + c] = P[(size_t)i * nb + c];
    for (int r = k0 + t; r < m; r += 256) gm[r] = map[r - k0];
}

// The same panel with the values in registers: 16 columns, thread t holds
// column t / 16 of the rows t % 16 + 16 q (q < QMAX, so m <= 16 QMAX).  The
// 16 threads of a column are 16 adjacent lanes of one wave, so the pivot
// search reduces inside that lane group; LDS carries only the pivot row, the
// pivot column and the chosen pivot.
template <int QMAX>
__global__ __launch_bounds__(256) void k_gj_panel_reg(cxd* const* __restrict__ mats, int m, int k0, int nbk,
                                                     int* __restrict__ gmap, int* __restrict__ perm,
                                                     int* __restrict__ bad)
{
    constexpr int NB = 16, G = 16;
    extern __shared__ unsigned char lds_raw[];
    cxd* fcol = reinterpret_cast<cxd*>(lds_raw);   // m
    int* map = reinterpret_cast<int*>(fcol + m);    // m - k0
    __shared__ cxd rk[NB], rp[NB];
    __shared__ int piv_row;
    cxd* a = mats[blockIdx.x];
    int* gm = gmap + (size_t)blockIdx.x * m;
    int* pm = perm + (size_t)blockIdx.x * m;
    const int t = threadIdx.x;
    const int c = t / G, rg = t % G;
    const bool col_ok = c < nbk;
    cxd v[QMAX];
#pragma unroll
    for (int q = 0; q < QMAX; q++) {
        const int i = rg + G * q;
        v[q] = (i < m && col_ok) ? a[(size_t)i * m + k0 + c] : cxd(0.0);
    }
    for (int r = k0 + t; r < m; r += 256) map[r - k0] = r;
    __syncthreads();
    for (int kk = 0; kk < nbk; kk++) {
        const int k = k0 + kk;
        // pivot: largest |a_ik|^2, i >= k, first index on ties (column kk's lane group)
        double bv = -1.0;
        int bi = m;
        if (c == kk) {
#pragma unroll
            for (int q = 0; q < QMAX; q++) {
                const int i = rg + G * q;
                if (i >= k && i < m) {
                    const double x = abs2v(v[q]);
                    if (x > bv) {
                        bv = x;
                        bi = i;
                    }
                }
            }
        }
        for (int off = 8; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (c == kk && rg == 0) {
            int p = bi;
            if (!(bv > 0.0)) {
                *bad = 1;
                p = k;
            }
            piv_row = p;
            pm[k] = p;
            const int tmp = map[k - k0];
            map[k - k0] = map[p - k0];
            map[p - k0] = tmp;
        }
        __syncthreads();
        const int p = piv_row;
        // rows k and p through LDS
#pragma unroll
        for (int q = 0; q < QMAX; q++) {
            const int i = rg + G * q;
            if (i == k && c < NB) rk[c] = v[q];
            if (i == p && c < NB) rp[c] = v[q];
        }
        __syncthreads();
        const cxd nk_kk = p != k ? rp[kk] : rk[kk];   // new pivot row, column kk
        const cxd nk_c = p != k ? rp[c] : rk[c];
        const cxd piv = abs2v(nk_kk) > 0.0 ? cxd(1.0) / nk_kk : cxd(0.0);
        const cxd srow_c = c == kk ? piv : nk_c * piv;
#pragma unroll
        for (int q = 0; q < QMAX; q++) {
            const int i = rg + G * q;
            if (p != k) {
                if (i == k) v[q] = rp[c];
                if (i == p) v[q] = rk[c];
            }
            if (c == kk && i < m) fcol[i] = v[q];
        }
        __syncthreads();
        if (col_ok) {
#pragma unroll
            for (int q = 0; q < QMAX; q++) {
                const int i = rg + G * q;
                if (i >= m) continue;
                if (i == k)
                    v[q] = srow_c;
                else if (c == kk)
                    v[q] = -(fcol[i] * piv);
                else
                    v[q] = v[q] - fcol[i] * srow_c;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < QMAX; q++) {
        const int i = rg + G * q;
        if (i < m && col_ok) a[(size_t)i * m + k0 + c] = v[q];
    }
    __syncthreads();
    for (int r = k0 + t; r < m; r += 256) gm[r] = map[r - k0];
}

// the panel's row interchanges on the other columns; rows K into tmp (nb x m) and zeroed
__global__ __launch_bounds__(256) void k_gj_rows(cxd* const* __restrict__ mats, int m, int k0, int nbk, int cw,
                                                const int* __restrict__ gmap, cxd* __restrict__ tmp, int nb)
{
    extern __shared__ unsigned char lds_raw[];
    cxd* G = reinterpret_cast<cxd*>(lds_raw);   // (m - k0) x cw gathered
    cxd* a = mats[blockIdx.y];
    const int* gm = gmap + (size_t)blockIdx.y * m;
    cxd* tb = tmp + (size_t)blockIdx.y * nb * m;
    const int c0 = blockIdx.x * cw;
    const int rows = m - k0;
    for (int e = threadIdx.x; e < rows * cw; e += 256) {
        const int r = e / cw, c = c0 + (e - r * cw);
        if (c < m && (c < k0 || c >= k0 + nbk)) G[e] = a[(size_t)gm[k0 + r] * m + c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < rows * cw; e += 256) {
        const int r = e / cw, c = c0 + (e - r * cw);
        if (c >= m) continue;
        if (c >= k0 && c < k0 + nbk) {
            if (r < nbk) tb[(size_t)r * m + c] = cxd(0.0);
            continue;
        }
        if (r < nbk) {
            tb[(size_t)r * m + c] = G[e];
            a[(size_t)(k0 + r) * m + c] = cxd(0.0);
        } else {
            a[(size_t)(k0 + r) * m + c] = G[e];
        }
    }
}

// a[:, c] += a[:, K] tmp[:, c] for the columns c outside K; 32 x 32 outputs per workgroup
__global__ __launch_bounds__(256) void k_gj_update(cxd* const* __restrict__ mats, int m, int k0, int nbk,
                                                  const cxd* __restrict__ tmp, int nb)
{
    __shared__ cxd As[32][17];
    __shared__ cxd Bs[16][33];
    cxd* a = mats[blockIdx.z];
    const cxd* tb = tmp + (size_t)blockIdx.z * nb * m;
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int row0 = blockIdx.y * 32, col0 = blockIdx.x * 32;
    if (col0 >= k0 && col0 + 32 <= k0 + nbk) return;   // the whole tile lies in the panel
    cxd acc[2][2] = {{cxd(0.0), cxd(0.0)}, {cxd(0.0), cxd(0.0)}};
    for (int q0 = 0; q0 < nbk; q0 += 16) {
        for (int e = t; e < 512; e += 256) {
            const int rr = e >> 4, kk = e & 15;
            const int gr = row0 + rr, gk = q0 + kk;
            As[rr][kk] = (gr < m && gk < nbk) ? a[(size_t)gr * m + k0 + gk] : cxd(0.0);
            const int kb = e >> 5, cc = e & 31;
            const int gkb = q0 + kb, gc = col0 + cc;
            Bs[kb][cc] = (gkb < nbk && gc < m) ? tb[(size_t)gkb * m + gc] : cxd(0.0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            const cxd a0 = As[ty * 2][kk], a1 = As[ty * 2 + 1][kk];
            const cxd b0 = Bs[kk][tx * 2], b1 = Bs[kk][tx * 2 + 1];
            acc[0][0] = acc[0][0] + a0 * b0;
            acc[0][1] = acc[0][1] + a0 * b1;
            acc[1][0] = acc[1][0] + a1 * b0;
            acc[1][1] = acc[1][1] + a1 * b1;
        }
        __syncthreads();
    }
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) {
            const int gr = row0 + ty * 2 + x, gc = col0 + tx * 2 + y;
            if (gr >= m || gc >= m || (gc >= k0 && gc < k0 + nbk)) continue;
            cxd* c = a + (size_t)gr * m + gc;
            *c = *c + acc[x][y];
        }
}

// the row interchanges of all panels undone as column interchanges, last first
__global__ __launch_bounds__(256) void k_gj_cols(cxd* const* __restrict__ mats, int m, const int* __restrict__ perm)
{
    extern __shared__ unsigned char lds_raw[];
    int* idx = reinterpret_cast<int*>(lds_raw);                        // m
    cxd* rowbuf = reinterpret_cast<cxd*>(lds_raw + ((m * sizeof(int) + 15) / 16) * 16);   // 4 x m
    cxd* a = mats[blockIdx.x];
    const int* pm = perm + (size_t)blockIdx.x * m;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    for (int c = t; c < m; c += 256) idx[c] = c;
    __syncthreads();
    if (t == 0)
        for (int k = m - 1; k >= 0; k--) {
            const int p = pm[k], x = idx[k];
            idx[k] = idx[p];
            idx[p] = x;
        }
    __syncthreads();
    cxd* buf = rowbuf + (size_t)wid * m;
    for (int i = wid; i < m; i += 4) {
        cxd* ai = a + (size_t)i * m;
        // each lane writes back only the entries it gathered
        for (int c = lane; c < m; c += 64) buf[c] = ai[idx[c]];
        for (int c = lane; c < m; c += 64) ai[c] = buf[c];
    }
}

// C = beta C + alpha A B over a batch of m x m blocks; 32 x 32 outputs per
// workgroup, 2 x 2 per thread, k in steps of 16 through LDS.
struct GemmOp {
    cxd* C;
    const cxd* A;
    const cxd* B;
    double alpha, beta;
};

__global__ __launch_bounds__(256) void k_zgemm_batched(const GemmOp* __restrict__ ops, int m)
{
    __shared__ cxd As[32][17];
    __shared__ cxd Bs[16][33];
    const GemmOp op = ops[blockIdx.z];
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int row0 = blockIdx.y * 32, col0 = blockIdx.x * 32;
    cxd acc[2][2] = {{cxd(0.0), cxd(0.0)}, {cxd(0.0), cxd(0.0)}};
    for (int k0 = 0; k0 < m; k0 += 16) {
        for (int e = t; e < 512; e += 256) {
            const int rr = e >> 4, kk = e & 15;
            const int gr = row0 + rr, gk = k0 + kk;
            As[rr][kk] = (gr < m && gk < m) ? op.A[(size_t)gr * m + gk] : cxd(0.0);
            const int kb = e >> 5, cc = e & 31;
            const int gkb = k0 + kb, gc = col0 + cc;
            Bs[kb][cc] = (gkb < m && gc < m) ? op.B[(size_t)gkb * m + gc] : cxd(0.0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            const cxd a0 = As[ty * 2][kk], a1 = As[ty * 2 + 1][kk];
            const cxd b0 = Bs[kk][tx * 2], b1 = Bs[kk][tx * 2 + 1];
            acc[0][0] = acc[0][0] + a0 * b0;
            acc[0][1] = acc[0][1] + a0 * b1;
            acc[1][0] = acc[1][0] + a1 * b0;
            acc[1][1] = acc[1][1] + a1 * b1;
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const int gr = row0 + ty * 2 + a, gc = col0 + tx * 2 + b;
            if (gr >= m || gc >= m) continue;
            cxd* c = op.C + (size_t)gr * m + gc;
            const cxd v = sc(acc[a][b], op.alpha);
            *c = op.beta == 0.0 ? v : sc(*c, op.beta) + v;
        }
}

// y = s0 v0 + s1 M1 x1 + s2 M2 x2 over a batch (M2 / v0 may be NULL); one
// workgroup per 16 rows of an op.
struct GemvOp {
    cxd* y;
    const cxd* v0;
    const cxd* M1;
    const cxd* x1;
    const cxd* M2;
    const cxd* x2;
    double s0, s1, s2;
};

__global__ __launch_bounds__(256) void k_zgemv_batched(const GemvOp* __restrict__ ops, int m)
{
    // 16 rows per workgroup, 4 per wave; lanes across the row
    const GemvOp op = ops[blockIdx.y];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int r0 = blockIdx.x * 16 + wid * 4;
    double re1[4] = {0, 0, 0, 0}, im1[4] = {0, 0, 0, 0}, re2[4] = {0, 0, 0, 0}, im2[4] = {0, 0, 0, 0};
    for (int c = lane; c < m; c += 64) {
        const cxd xa = op.x1[c];
        const cxd xb = op.M2 ? op.x2[c] : cxd(0.0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rw = r0 + q;
            if (rw >= m) continue;
            const cxd a = op.M1[(size_t)rw * m + c] * xa;
            re1[q] += a.re;
            im1[q] += a.im;
            if (op.M2) {
                const cxd b = op.M2[(size_t)rw * m + c] * xb;
                re2[q] += b.re;
                im2[q] += b.im;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double a_re = wave_sum(re1[q]), a_im = wave_sum(im1[q]);
        const double b_re = wave_sum(re2[q]), b_im = wave_sum(im2[q]);
        const int rw = r0 + q;
        if (lane == 0 && rw < m) {
            cxd v = sc(cxd(a_re, a_im), op.s1);
            if (op.M2) v = v + sc(cxd(b_re, b_im), op.s2);
            if (op.v0) v = v + sc(op.v0[rw], op.s0);
            op.y[rw] = v;
        }
    }
}

// E (unknowns) = x, or E += x
__global__ __launch_bounds__(256) void k_mws_scatter(MwsView V, Fields F, const cxd* __restrict__ x, int add)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (s >= V.m) return;
    const int blk = s / V.ld, j = s % V.ld;
    for (int c = 0; c < 3; c++)
        if (V.pos[c] == blk && active(V, F, c, i, j)) {
            const size_t q = qidx(V.ld, i, j);
            const cxd v = x[(size_t)i * V.m + s];
            F.E[c][q] = add ? F.E[c][q] + v : v;
        }
}

// r = b - A E on the unknowns, into the block vector; per-workgroup partial
// sums of |r|^2 and |b|^2
__global__ __launch_bounds__(256) void k_mws_residual(MwsView V, Fields F, const cxd* __restrict__ co,
                                                     cxd* __restrict__ rv, double* __restrict__ part)
{
    __shared__ double lds[8];
    const int s = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    const int ld = V.ld;
    const size_t n2 = (size_t)(V.ni + 2) * ld;
    double rr = 0.0, bb = 0.0;
    if (s < V.m) {
        const int blk = s / ld, j = s % ld;
        int c = -1;
        for (int k = 0; k < 3; k++)
            if (V.pos[k] == blk) c = k;
        cxd out(0.0);
        if (c >= 0 && active(V, F, c, i, j)) {
            const size_t q = qidx(ld, i, j);
            cxd y = co[((size_t)c * NSLOT + SP) * n2 + q] * F.E[c][q];
            for (int k = 1; k < SB; k++) {
                const int tc = c_target[c][k][0];
                if (tc < 0) continue;
                if (V.which != CATHODE_MW_COUPLED && k > SS) continue;
                const int ti = i + c_target[c][k][1], tj = j + c_target[c][k][2];
                if (!active(V, F, tc, ti, tj)) continue;
                y = y - co[((size_t)c * NSLOT + k) * n2 + q] * F.E[tc][qidx(ld, ti, tj)];
            }
            const cxd b = co[((size_t)c * NSLOT + SB) * n2 + q];
            out = b - y;
            rr = abs2v(out);
            bb = abs2v(b);
        }
        rv[(size_t)i * V.m + s] = out;
    }
    double v2[2] = {rr, bb};
    block_sum_vec_bcast<256, 2>(v2, lds);
    if (threadIdx.x == 0) {
        const size_t w = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * w] = v2[0];
        part[2 * w + 1] = v2[1];
    }
}

// impedance test: the reference-plane sums (:5926-6003) and the Poynting
// flux through the east face of block 0 (mWLoss_at_boundary_xR, :8537-8579)
__global__ __launch_bounds__(64) void k_mws_refplane(MwsView V, Fields F, int i_ref, int jlo, int jhi, int io,
                                                    int jo_lo, int jo_hi, double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    const int ld = V.ld;
    const double* r = V.r;
    const cxd* Er = F.E[1];
    const cxd* Ex = F.E[0];
    const cxd* Ep = F.E[2];
    cxd Is(0.0), cs(0.0);
    int n = 0;
    for (int j = jlo; j <= jhi; j++) {
        const double rt = (r[j] + r[j - 1]) / 2.0;
        const size_t q = qidx(ld, i_ref, j);
        const cxd H = sc(Er[q] - Er[q - ld], 1.0 / V.dx / (V.omega * MWS_MU0)) * cxd(0.0, 1.0);
        Is = Is + sc(H, 2.0 * M_PI * rt);
        cs = cs + sc(sc(Er[q] + Er[q - ld], 0.5), rt);
        n++;
    }
    double P = 0.0;
    const cxd iwm(0.0, 1.0 / (V.omega * MWS_MU0));
    for (int j = jo_lo; j <= jo_hi; j++) {
        const double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
        const size_t q = qidx(ld, io, j);
        const cxd HL = iwm * (sc(Er[q + ld] - Er[q], 1.0 / V.dx) - sc(Ex[q + ld] - Ex[q + ld - 1], 1.0 / V.dr));
        const cxd HR = iwm * (sc(Er[q + ld + 1] - Er[q + 1], 1.0 / V.dx) - sc(Ex[q + ld + 1] - Ex[q + ld], 1.0 / V.dr));
        const cxd Hp = sc(sc(HR, rR) + sc(HL, rL), 1.0 / (2.0 * r[j]));
        const cxd Ept = sc(Ep[q + ld] + Ep[q], 1.0 / (2.0 * rR));
        const cxd ErL = sc(sc(Er[q + ld + 1], rR) + sc(Er[q + ld], rL), 1.0 / (2.0 * r[j]));
        const cxd ErR = sc(sc(Er[q + 1], rR) + sc(Er[q], rL), 1.0 / (2.0 * r[j]));
        const cxd Ert = sc(ErR + ErL, 0.5);
        const cxd ExH = Ert * conj(Hp) - Ept * conj(Hp);
        P += 2.0 * M_PI * r[j] * (0.5 * ExH.re) * V.dr;
    }
    out[0] = Is.re;
    out[1] = Is.im;
    out[2] = cs.re;
    out[3] = cs.im;
    out[4] = (double)n;
    out[5] = P;
}

__global__ __launch_bounds__(256) void k_mws_scale(cxd* a, size_t n, double s)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = sc(a[k], s);
}

__global__ __launch_bounds__(256) void k_mws_fill(double* a, size_t n, double v)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

// ---------------------------------------------------------------- host side
struct MwsBuffers {
    DevBuf<unsigned char> work;      // coefficients, blocks, vectors, pointer tables
    DevBuf<double> stage;            // host-array staging (real)
    DevBuf<int> istage;              // list entries
    DevBuf<int> acts;                // host-array staging (masks)
    DevBuf<cxd> cstage;              // host-array staging (complex) / zero stand-ins
    DevBuf<double> radii;
};
MwsBuffers g_mws;

struct Plan {
    std::vector<std::vector<int> > R;   // active rows per level
};

Plan make_plan(int ncol)
{
    Plan P;
    std::vector<int> r(ncol);
    for (int i = 0; i < ncol; i++) r[i] = i;
    P.R.push_back(r);
    while (P.R.back().size() > 1) {
        const std::vector<int>& cur = P.R.back();
        std::vector<int> nxt;
        for (size_t t = 0; t < cur.size(); t += 2) nxt.push_back(cur[t]);
        P.R.push_back(nxt);
    }
    return P;
}

bool check_lists(const cathode_mw_params* p, const cathode_mw_bc* bc, std::string& why)
{
    const cathode_mw_bc_list* all[12];
    for (int s = 0; s < 4; s++) {
        all[s] = &bc->x[s];
        all[4 + s] = &bc->r[s];
        all[8 + s] = &bc->p[s];
    }
    static const int gdi[4] = {-1, 1, 0, 0}, gdj[4] = {0, 0, -1, 1};
    for (int l = 0; l < 12; l++) {
        const cathode_mw_bc_list* L = all[l];
        if (L->n < 0 || (L->n > 0 && (!L->i || !L->j || !L->s))) {
            why = "list " + std::to_string(l) + " has no entries array";
            return false;
        }
        std::set<std::pair<int, int> > seen;
        for (int k = 0; k < L->n; k++) {
            const int i = L->i[k], j = L->j[k];
            if (i < 1 || i > p->ni || j < 1 || j > p->nj || (l >= 4 && l < 8 && j < 2)) {
                why = "list " + std::to_string(l) + " entry " + std::to_string(k) +
                      " outside 1..ni x 1..nj (x 2..nj for E_r)";
                return false;
            }
            if (!seen.insert({i, j}).second) {
                why = "list " + std::to_string(l) + " names (" + std::to_string(i) + ", " + std::to_string(j) +
                      ") twice";
                return false;
            }
        }
        // a ghost written by one entry must not be another entry of the same list
        for (int k = 0; k < L->n; k++) {
            const int gi = L->i[k] + gdi[l % 4], gj = L->j[k] + gdj[l % 4];
            if (seen.count({gi, gj})) {
                why = "list " + std::to_string(l) + ": the ghost of entry " + std::to_string(k) +
                      " is another entry of the list";
                return false;
            }
        }
    }
    return true;
}

int mws_run(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o, Fields F,
            cathode_mw_solve_report* rep, hipStream_t st, const double* radii)
{
    const int ni = p->ni, nj = p->nj, ld = nj + 2, ncol = ni + 2;
    const size_t n2 = (size_t)ncol * ld;
    MwsView V{};
    V.ni = ni;
    V.nj = nj;
    V.ld = ld;
    V.which = o->which;
    if (o->which == CATHODE_MW_COUPLED) {
        V.pos[0] = 0;
        V.pos[1] = 1;
        V.pos[2] = 2;
        V.m = 3 * ld;
    } else {
        V.pos[0] = -1;
        V.pos[1] = 0;
        V.pos[2] = -1;
        V.m = ld;
    }
    V.n_bl = p->n_bl;
    V.j_flc0 = o->j_flc0;
    V.dx = p->dx;
    V.dr = p->dr;
    V.omega = p->omega;
    V.masse = p->masse;
    V.J1r_exc = p->J1r_exc;
    V.r1 = p->r1;
    V.r2 = p->r2;
    V.r = radii;
    std::memcpy(V.box[0], p->xface_box, sizeof(p->xface_box));
    std::memcpy(V.box[1], p->rface_box, sizeof(p->rface_box));
    std::memcpy(V.box[2], p->cell_box, sizeof(p->cell_box));
    const int m = V.m;
    const size_t mm = (size_t)m * m;
    const Plan plan = make_plan(ncol);
    const int L = (int)plan.R.size() - 1;   // levels with an elimination

    // pointer-table entries and workspace layout
    size_t nlevel_rows = 0;
    for (int l = 0; l <= L; l++) nlevel_rows += plan.R[l].size();
    std::vector<std::vector<int> > elim(L);
    for (int l = 0; l < L; l++)
        for (size_t t = 1; t < plan.R[l].size(); t += 2) elim[l].push_back(plan.R[l][t]);
    Arena ar;
    size_t inv_max = 1;
    for (int l = 0; l < L; l++) inv_max = std::max(inv_max, elim[l].size());
    // Gauss-Jordan panel width and gather width: the panel (plus pivot buffers) and
    // the gathered column strip each within 64 KiB of LDS
    int nb = m <= 16 * 32 ? 16 : 32, cw = 32;
    while (nb > 1 && ((size_t)m * nb + m + nb) * sizeof(cxd) + (size_t)m * sizeof(int) > 65536) nb /= 2;
    while (cw > 1 && (size_t)m * cw * sizeof(cxd) > 65536) cw /= 2;
    const size_t panel_lds = ((size_t)m * nb + m + nb) * sizeof(cxd) + (size_t)m * sizeof(int);
    const size_t cols_lds = ((m * sizeof(int) + 15) / 16) * 16 + 4 * (size_t)m * sizeof(cxd);
    if (panel_lds > 65536 || (size_t)m * cw * sizeof(cxd) > 65536 || cols_lds > 65536) {
        set_error("mw_solve: column block of " + std::to_string(m) + " slots exceeds the inverse's LDS buffers");
        return CATHODE_ERR_ARG;
    }
    cxd* tmpb;
    int *gmapb, *permb;
    auto layout = [&](Arena& a, cxd** co, cxd** Bm, cxd** GA, cxd** GC, cxd** AC, cxd** fv, cxd** gv, cxd** xv,
                      double** part, int** bad, GemmOp** gops, GemvOp** vops, cxd*** mptr, size_t ngemm, size_t ngemv,
                      size_t nparts) {
        tmpb = a.take<cxd>(inv_max * nb * m);
        gmapb = a.take<int>(inv_max * m);
        permb = a.take<int>(inv_max * m);
        *co = a.take<cxd>(3 * NSLOT * n2);
        *Bm = a.take<cxd>(ncol * mm);
        *GA = a.take<cxd>(ncol * mm);
        *GC = a.take<cxd>(ncol * mm);
        *AC = a.take<cxd>(2 * nlevel_rows * mm);
        *fv = a.take<cxd>((size_t)ncol * m);
        *gv = a.take<cxd>((size_t)ncol * m);
        *xv = a.take<cxd>((size_t)ncol * m);
        *part = a.take<double>(2 * nparts + 16);
        *bad = a.take<int>(16);
        *gops = a.take<GemmOp>(ngemm);
        *vops = a.take<GemvOp>(ngemv);
        *mptr = a.take<cxd*>((size_t)ncol + 16);
    };
    // the operation lists (host), built against the workspace base below
    const dim3 gcol(cdiv(m, 256), ncol);
    const size_t nparts = (size_t)gcol.x * gcol.y;
    size_t ngemm = 0, ngemv = 0;
    for (int l = 0; l < L; l++) {
        const size_t K = plan.R[l].size();
        ngemm += 6 * K + 8;
        ngemv += 3 * K + 8;
    }
    ngemv += 8;
    cxd *co, *Bm, *GA, *GC, *AC, *fv, *gv, *xv;
    double* part;
    int* bad;
    GemmOp* gops;
    GemvOp* vops;
    cxd** mptr;
    ar.base = nullptr;
    layout(ar, &co, &Bm, &GA, &GC, &AC, &fv, &gv, &xv, &part, &bad, &gops, &vops, &mptr, ngemm, ngemv, nparts);
    const size_t bytes = ar.used;
    if (bytes > ((size_t)96 << 30)) {
        set_error("mw_solve: the dense column blocks need more than 96 GiB (m = " + std::to_string(m) +
                  ", columns = " + std::to_string(ncol) + ")");
        return CATHODE_ERR_ARG;
    }
    g_mws.work.alloc(bytes);
    Arena wa;
    wa.base = g_mws.work.p;
    layout(wa, &co, &Bm, &GA, &GC, &AC, &fv, &gv, &xv, &part, &bad, &gops, &vops, &mptr, ngemm, ngemv, nparts);
    // level l's A and C of its t-th active row
    std::vector<size_t> lvl_base(L + 1);
    {
        size_t acc = 0;
        for (int l = 0; l <= L; l++) {
            lvl_base[l] = acc;
            acc += plan.R[l].size();
        }
    }
    auto Al = [&](int l, size_t t) { return AC + (2 * (lvl_base[l] + t)) * mm; };
    auto Cl = [&](int l, size_t t) { return AC + (2 * (lvl_base[l] + t) + 1) * mm; };
    auto Bp = [&](int row) { return Bm + (size_t)row * mm; };
    auto GAp = [&](int row) { return GA + (size_t)row * mm; };
    auto GCp = [&](int row) { return GC + (size_t)row * mm; };
    auto fp = [&](int row) { return fv + (size_t)row * m; };
    auto gp = [&](int row) { return gv + (size_t)row * m; };
    auto xp = [&](int row) { return xv + (size_t)row * m; };

    // ---- rows: assemble, fold, blocks
    CATHODE_HIP(hipMemsetAsync(co, 0, 3 * NSLOT * n2 * sizeof(cxd), st));
    {
        dim3 g(cdiv(ld, 256), ncol);
        hipLaunchKernelGGL(k_mws_assemble, g, dim3(256), 0, st, V, F, co);
    }
    std::vector<int> ent;
    ListSet LS{};
    {
        const cathode_mw_bc_list* all[12];
        for (int s = 0; s < 4; s++) {
            all[s] = &bc->x[s];
            all[4 + s] = &bc->r[s];
            all[8 + s] = &bc->p[s];
        }
        for (int l = 0; l < 12; l++) {
            LS.off[l] = (int)(ent.size() / 3);
            for (int k = 0; k < all[l]->n; k++) {
                ent.push_back(all[l]->i[k]);
                ent.push_back(all[l]->j[k]);
                ent.push_back(all[l]->s[k]);
            }
        }
        LS.off[12] = (int)(ent.size() / 3);
    }
    g_mws.istage.alloc(ent.size() + 16);
    if (!ent.empty())
        CATHODE_HIP(hipMemcpyAsync(g_mws.istage.p, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice, st));
    LS.e = g_mws.istage.p;
    hipLaunchKernelGGL(k_mws_fold, dim3(1), dim3(256), 0, st, V, F, LS, co);
    CATHODE_HIP(hipMemsetAsync(Bm, 0, ncol * mm * sizeof(cxd), st));
    CATHODE_HIP(hipMemsetAsync(AC, 0, 2 * nlevel_rows * mm * sizeof(cxd), st));
    hipLaunchKernelGGL(k_mws_blocks, gcol, dim3(256), 0, st, V, F, (const cxd*)co, Bm, Al(0, 0), fv);
    CATHODE_LAUNCH_CHECK();

    // ---- factor: block cyclic reduction
    std::vector<GemmOp> hg;
    std::vector<GemvOp> hv;
    std::vector<cxd*> hm;
    struct Launch {
        int kind;       // 0 inverse, 1 gemm, 2 gemv
        size_t first, count;
    };
    std::vector<Launch> factor, solve;
    auto inv_batch = [&](std::vector<Launch>& seq, const std::vector<int>& rows) {
        if (rows.empty()) return;
        const size_t f0 = hm.size();
        for (int rw : rows) hm.push_back(Bp(rw));
        seq.push_back({0, f0, rows.size()});
    };
    auto gemm_batch = [&](std::vector<Launch>& seq, const std::vector<GemmOp>& ops) {
        if (ops.empty()) return;
        seq.push_back({1, hg.size(), ops.size()});
        hg.insert(hg.end(), ops.begin(), ops.end());
    };
    auto gemv_batch = [&](std::vector<Launch>& seq, const std::vector<GemvOp>& ops) {
        if (ops.empty()) return;
        seq.push_back({2, hv.size(), ops.size()});
        hv.insert(hv.end(), ops.begin(), ops.end());
    };
    std::vector<char> has_right(ncol, 0);   // GC of an eliminated row exists
    for (int l = 0; l < L; l++) {
        const std::vector<int>& R = plan.R[l];
        const size_t K = R.size();
        inv_batch(factor, elim[l]);
        std::vector<GemmOp> g1, g2, g3, g4;
        for (size_t t = 1; t < K; t += 2) {
            const int e = R[t];
            g1.push_back({GAp(e), Bp(e), Al(l, t), 1.0, 0.0});
            if (t + 1 < K) {
                g1.push_back({GCp(e), Bp(e), Cl(l, t), 1.0, 0.0});
                has_right[e] = 1;
            }
        }
        for (size_t t = 0; t < K; t += 2) {
            const size_t tn = t / 2;
            if (t >= 1) {   // left eliminated neighbour R[t-1]
                const int eL = R[t - 1];
                g2.push_back({Bp(R[t]), Al(l, t), GCp(eL), -1.0, 1.0});
                g4.push_back({Al(l + 1, tn), Al(l, t), GAp(eL), -1.0, 0.0});
            }
            if (t + 1 < K) {   // right eliminated neighbour R[t+1]
                const int eR = R[t + 1];
                g3.push_back({Bp(R[t]), Cl(l, t), GAp(eR), -1.0, 1.0});
                if (has_right[eR]) g4.push_back({Cl(l + 1, tn), Cl(l, t), GCp(eR), -1.0, 0.0});
            }
        }
        gemm_batch(factor, g1);
        gemm_batch(factor, g2);
        gemm_batch(factor, g3);
        gemm_batch(factor, g4);
    }
    inv_batch(factor, plan.R[L]);

    // ---- solve (right-hand side in fv; solution in xv)
    for (int l = 0; l < L; l++) {
        const std::vector<int>& R = plan.R[l];
        const size_t K = R.size();
        std::vector<GemvOp> v1, v2;
        for (size_t t = 1; t < K; t += 2) v1.push_back({gp(R[t]), nullptr, Bp(R[t]), fp(R[t]), nullptr, nullptr, 0.0, 1.0, 0.0});
        for (size_t t = 0; t < K; t += 2) {
            const bool lft = t >= 1, rgt = t + 1 < K;
            if (!lft && !rgt) continue;
            GemvOp op{fp(R[t]), fp(R[t]), nullptr, nullptr, nullptr, nullptr, 1.0, -1.0, -1.0};
            if (lft) {
                op.M1 = Al(l, t);
                op.x1 = gp(R[t - 1]);
                if (rgt) {
                    op.M2 = Cl(l, t);
                    op.x2 = gp(R[t + 1]);
                }
            } else {
                op.M1 = Cl(l, t);
                op.x1 = gp(R[t + 1]);
            }
            v2.push_back(op);
        }
        gemv_batch(solve, v1);
        gemv_batch(solve, v2);
    }
    {
        const int last = plan.R[L][0];
        gemv_batch(solve, {{xp(last), nullptr, Bp(last), fp(last), nullptr, nullptr, 0.0, 1.0, 0.0}});
    }
    for (int l = L - 1; l >= 0; l--) {
        const std::vector<int>& R = plan.R[l];
        const size_t K = R.size();
        std::vector<GemvOp> v;
        for (size_t t = 1; t < K; t += 2) {
            const int e = R[t];
            GemvOp op{xp(e), gp(e), GAp(e), xp(R[t - 1]), nullptr, nullptr, 1.0, -1.0, -1.0};
            if (has_right[e]) {
                op.M2 = GCp(e);
                op.x2 = xp(R[t + 1]);
            }
            v.push_back(op);
        }
        gemv_batch(solve, v);
    }
    if (hg.size() > ngemm || hv.size() > ngemv || hm.size() > (size_t)ncol + 16) {
        set_error("mw_solve: internal plan size");
        return CATHODE_ERR_HIP;
    }
    if (!hg.empty()) CATHODE_HIP(hipMemcpyAsync(gops, hg.data(), hg.size() * sizeof(GemmOp), hipMemcpyHostToDevice, st));
    if (!hv.empty()) CATHODE_HIP(hipMemcpyAsync(vops, hv.data(), hv.size() * sizeof(GemvOp), hipMemcpyHostToDevice, st));
    if (!hm.empty()) CATHODE_HIP(hipMemcpyAsync(mptr, hm.data(), hm.size() * sizeof(cxd*), hipMemcpyHostToDevice, st));
    CATHODE_HIP(hipMemsetAsync(bad, 0, 16 * sizeof(int), st));
    auto run = [&](const std::vector<Launch>& seq) {
        for (const Launch& Lc : seq) {
            if (Lc.kind == 0) {
                cxd* const* mats = mptr + Lc.first;
                const unsigned cnt = (unsigned)Lc.count;
                const int tiles = cdiv(m, 32);
                for (int k0 = 0; k0 < m; k0 += nb) {
                    const int nbk = std::min(nb, m - k0);
                    const size_t reg_lds = (size_t)m * sizeof(cxd) + (size_t)m * sizeof(int);
                    if (nb == 16 && m <= 16 * 16)
                        hipLaunchKernelGGL(k_gj_panel_reg<16>, dim3(cnt), dim3(256), reg_lds, st, mats, m, k0, nbk,
                                           gmapb, permb, bad);
                    else if (nb == 16 && m <= 16 * 32)
                        hipLaunchKernelGGL(k_gj_panel_reg<32>, dim3(cnt), dim3(256), reg_lds, st, mats, m, k0, nbk,
                                           gmapb, permb, bad);
                    else
                        hipLaunchKernelGGL(k_gj_panel, dim3(cnt), dim3(256), panel_lds, st, mats, m, k0, nbk, nb,
                                           gmapb, permb, bad);
                    hipLaunchKernelGGL(k_gj_rows, dim3(cdiv(m, cw), cnt), dim3(256), (size_t)m * cw * sizeof(cxd), st,
                                       mats, m, k0, nbk, cw, (const int*)gmapb, tmpb, nb);
                    hipLaunchKernelGGL(k_gj_update, dim3(tiles, tiles, cnt), dim3(256), 0, st, mats, m, k0, nbk,
                                       (const cxd*)tmpb, nb);
                }
                hipLaunchKernelGGL(k_gj_cols, dim3(cnt), dim3(256), cols_lds, st, mats, m, (const int*)permb);
            } else if (Lc.kind == 1) {
                const int tiles = cdiv(m, 32);
                hipLaunchKernelGGL(k_zgemm_batched, dim3(tiles, tiles, (unsigned)Lc.count), dim3(256), 0, st,
                                   (const GemmOp*)(gops + Lc.first), m);
            } else {
                hipLaunchKernelGGL(k_zgemv_batched, dim3(cdiv(m, 16), (unsigned)Lc.count), dim3(256), 0, st,
                                   (const GemvOp*)(vops + Lc.first), m);
            }
        }
        CATHODE_LAUNCH_CHECK();
    };
    run(factor);
    int hbad = 0;
    CATHODE_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    CATHODE_HIP(hipStreamSynchronize(st));
    if (hbad) {
        set_error("mw_solve: singular column block (zero pivot)");
        return CATHODE_ERR_NUMERIC;
    }
    run(solve);
    hipLaunchKernelGGL(k_mws_scatter, gcol, dim3(256), 0, st, V, F, (const cxd*)xv, 0);
    // refinement with the same factors
    std::vector<double> hp(2 * nparts);
    double relres = 0.0;
    int steps = 0;
    for (;;) {
        hipLaunchKernelGGL(k_mws_residual, gcol, dim3(256), 0, st, V, F, (const cxd*)co, fv, part);
        CATHODE_LAUNCH_CHECK();
        CATHODE_HIP(hipMemcpyAsync(hp.data(), part, 2 * nparts * sizeof(double), hipMemcpyDeviceToHost, st));
        CATHODE_HIP(hipStreamSynchronize(st));
        double rr = 0.0, bb = 0.0;
        for (size_t k = 0; k < nparts; k++) {
            rr += hp[2 * k];
            bb += hp[2 * k + 1];
        }
        relres = std::sqrt(rr) / (bb > 0.0 ? std::sqrt(bb) : 1.0);
        if (!(relres > o->tol) || steps >= o->max_refine) break;
        run(solve);
        hipLaunchKernelGGL(k_mws_scatter, gcol, dim3(256), 0, st, V, F, (const cxd*)xv, 1);
        steps++;
    }
    if (!std::isfinite(relres)) {
        set_error("mw_solve: the solution is not finite");
        return CATHODE_ERR_NUMERIC;
    }
    hipLaunchKernelGGL(k_mws_post, dim3(1), dim3(256), 0, st, V, F, LS);
    CATHODE_LAUNCH_CHECK();
    std::memset(rep, 0, sizeof(*rep));
    {
        // unknowns per component
        std::vector<int> h(n2);
        int cnt[3] = {0, 0, 0};
        for (int c = 0; c < 3; c++) {
            if (V.pos[c] < 0) continue;
            CATHODE_HIP(hipMemcpyAsync(h.data(), F.act[c], n2 * sizeof(int), hipMemcpyDeviceToHost, st));
            CATHODE_HIP(hipStreamSynchronize(st));
            for (size_t q = 0; q < n2; q++) cnt[c] += h[q] == 1;
        }
        rep->n_x = cnt[0];
        rep->n_r = cnt[1];
        rep->n_p = cnt[2];
    }
    rep->block = m;
    rep->refine = steps;
    rep->rel_residual = relres;
    if (o->which == CATHODE_MW_IMPEDANCE) {
        hipLaunchKernelGGL(k_mws_refplane, dim3(1), dim3(64), 0, st, V, F, p->i_ref, p->rface_box[2], p->rface_box[3],
                           p->cell_box[1], p->cell_box[2], p->cell_box[3], part);
        CATHODE_LAUNCH_CHECK();
        double h[6];
        CATHODE_HIP(hipMemcpyAsync(h, part, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
        CATHODE_HIP(hipStreamSynchronize(st));
        typedef std::complex<double> C128;
        const double cnt = h[4];
        const C128 Iref = C128(h[0], h[1]) / cnt;
        const C128 Vref = C128(h[2], h[3]) / cnt * std::log(p->r2 / p->r1);
        const C128 Zref = Vref / (Iref + 1e-100);
        const C128 a2 = (Vref - p->Z0 * Iref) / (2.0 * std::sqrt(2.0 * p->Z0));
        const C128 b2 = (Vref + p->Z0 * Iref) / (2.0 * std::sqrt(2.0 * p->Z0));
        const double Pf = std::norm(b2), Pr = std::norm(a2), Pout = h[5];
        const double Perr = Pf - Pr - Pout;
        const C128 S11(p->S11[0], p->S11[1]), S12(p->S12[0], p->S12[1]), S21(p->S21[0], p->S21[1]),
            S22(p->S22[0], p->S22[1]);
        const C128 a1 = (b2 - S22 * a2) / S21;
        const C128 b1 = S11 * a1 + S12 * a2;
        const double ratio = p->Pmw / Pf, sr = std::sqrt(ratio);
        cxd* sc6[6] = {F.E[0], F.E[1], F.E[2], F.J[0], F.J[1], F.J[2]};
        for (int k = 0; k < 6; k++)
            hipLaunchKernelGGL(k_mws_scale, dim3(cdiv((long)n2, 256)), dim3(256), 0, st, sc6[k], n2, sr);
        CATHODE_LAUNCH_CHECK();
        rep->V_ref[0] = Vref.real();
        rep->V_ref[1] = Vref.imag();
        rep->I_ref[0] = Iref.real();
        rep->I_ref[1] = Iref.imag();
        rep->Z_ref[0] = Zref.real();
        rep->Z_ref[1] = Zref.imag();
        rep->ratio = ratio;
        rep->J1r_exc = p->J1r_exc * sr;
        rep->P_fwd = Pf * ratio;
        rep->P_ref = Pr * ratio;
        rep->P_fwd_all = std::norm(a1) * ratio;
        rep->P_ref_all = std::norm(b1) * ratio;
        rep->P_err = Perr * ratio;
        rep->P_out = Pout * ratio;
    }
    return CATHODE_OK;
}

int mws_check(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
              const cathode_mw_solve_fields* f, const cathode_mw_solve_report* rep)
{
    if (!p || !bc || !o || !f || !rep || p->ni < 1 || p->nj < 2 || p->n_bl < 1 ||
        p->n_bl > CATHODE_MW_MAX_BLOCKS || !(p->dx > 0) || !(p->dr > 0) || !(p->omega > 0) || !(p->masse > 0) ||
        (o->which != CATHODE_MW_COUPLED && o->which != CATHODE_MW_IMPEDANCE) || o->max_refine < 0 ||
        o->max_refine > 8 || !f->rhoe || !f->nu_m || !f->Bx || !f->Br || !f->epsr || !f->act_x || !f->act_r ||
        !f->act_p || !f->E1x || !f->E1r || !f->E1p) {
        set_error("mw_solve: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    for (int b = 0; b < p->n_bl; b++) {
        const int *x = p->xface_box + 4 * b, *r = p->rface_box + 4 * b, *c = p->cell_box + 4 * b;
        if (x[0] < 1 || x[1] > p->ni + 1 || x[2] < 1 || x[3] > p->nj || r[0] < 0 || r[1] > p->ni + 1 || r[2] < 2 ||
            r[3] > p->nj || c[0] < 0 || c[1] > p->ni + 1 || c[2] < 1 || c[3] > p->nj) {
            set_error("mw_solve: a block box leaves x faces 1..ni+1 x 1..nj, r faces 0..ni+1 x 2..nj or cells "
                      "0..ni+1 x 1..nj");
            return CATHODE_ERR_ARG;
        }
    }
    std::string why;
    if (!check_lists(p, bc, why)) {
        set_error("mw_solve: " + why);
        return CATHODE_ERR_ARG;
    }
    bool needs_feed = false;
    for (int k = 0; k < bc->r[0].n; k++) needs_feed |= bc->r[0].s[k] == CATHODE_MWBC_EXCITE;
    if ((needs_feed || o->which == CATHODE_MW_IMPEDANCE) && !(p->r2 > p->r1 && p->r1 > 0)) {
        set_error("mw_solve: the excitation needs feed radii 0 < r1 < r2");
        return CATHODE_ERR_ARG;
    }
    if (o->which == CATHODE_MW_IMPEDANCE) {
        if (p->i_ref < 1 || p->i_ref > p->ni + 1 || p->rface_box[2] > p->rface_box[3] || p->cell_box[1] > p->ni ||
            p->cell_box[1] < 0 || !(p->Z0 > 0) || !(p->Pmw > 0) || (p->S21[0] == 0.0 && p->S21[1] == 0.0)) {
            set_error("mw_solve: impedance test needs i_ref in 1..ni+1, block 0's boxes, Z0 > 0, Pmw > 0, S21 != 0");
            return CATHODE_ERR_ARG;
        }
    }
    return CATHODE_OK;
}

}  // namespace

void mws_release()
{
    g_mws.work.release();
    g_mws.stage.release();
    g_mws.istage.release();
    g_mws.acts.release();
    g_mws.cstage.release();
    g_mws.radii.release();
}

}  // namespace cathode

using namespace cathode;

static const double* upload_radii(const cathode_mw_params* p, hipStream_t s)
{
    std::vector<double> r(p->nj + 2);
    for (int j = 0; j < p->nj + 2; j++) r[j] = p->rmin + (j - 0.5) * p->dr;   // initialModule.hpp:86-88
    g_mws.radii.alloc(r.size());
    CATHODE_HIP(hipMemcpyAsync(g_mws.radii.p, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice, s));
    return g_mws.radii.p;
}

extern "C" int cathode_mw_solve(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                                cathode_mw_solve_fields* f, cathode_mw_solve_report* rep)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    int rc = mws_check(p, bc, o, f, rep);
    if (rc != CATHODE_OK) return rc;
    hipStream_t s = ctx().stream;
    const size_t n2 = (size_t)(p->ni + 2) * (p->nj + 2);
    g_mws.stage.alloc(6 * n2);
    g_mws.acts.alloc(3 * n2);
    g_mws.cstage.alloc(6 * n2);
    double* R = g_mws.stage.p;
    const double* rin[6] = {f->rhoe, f->nu_m, f->Bx, f->Br, f->epsr, f->Er_dc};
    for (int k = 0; k < 6; k++) {
        if (rin[k])
            CATHODE_HIP(hipMemcpyAsync(R + k * n2, rin[k], n2 * sizeof(double), hipMemcpyHostToDevice, s));
        else
            CATHODE_HIP(hipMemsetAsync(R + k * n2, 0, n2 * sizeof(double), s));
    }
    const int* ain[3] = {f->act_x, f->act_r, f->act_p};
    for (int k = 0; k < 3; k++)
        CATHODE_HIP(hipMemcpyAsync(g_mws.acts.p + k * n2, ain[k], n2 * sizeof(int), hipMemcpyHostToDevice, s));
    cxd* C = g_mws.cstage.p;
    double* cin[6] = {f->J1x_exc, f->J1r_exc, f->J1p_exc, f->E1x, f->E1r, f->E1p};
    for (int k = 0; k < 6; k++) {
        if (cin[k])
            CATHODE_HIP(hipMemcpyAsync(C + k * n2, cin[k], n2 * sizeof(cxd), hipMemcpyHostToDevice, s));
        else
            CATHODE_HIP(hipMemsetAsync(C + k * n2, 0, n2 * sizeof(cxd), s));
    }
    const double* radii = upload_radii(p, s);
    if (o->which == CATHODE_MW_IMPEDANCE) {   // the reference zeroes n_e and sets nu = nu_eff (:5531-5536)
        CATHODE_HIP(hipMemsetAsync(R, 0, n2 * sizeof(double), s));
        hipLaunchKernelGGL(k_mws_fill, dim3(cdiv((long)n2, 256)), dim3(256), 0, s, R + n2, n2, o->nu_eff);
    }
    Fields F{};
    F.rhoe = R;
    F.nu = R + n2;
    F.Bx = R + 2 * n2;
    F.Br = R + 3 * n2;
    F.epsr = R + 4 * n2;
    F.Er_dc = f->Er_dc ? R + 5 * n2 : nullptr;
    for (int k = 0; k < 3; k++) {
        F.act[k] = g_mws.acts.p + k * n2;
        F.J[k] = C + k * n2;
        F.E[k] = C + (3 + k) * n2;
    }
    rc = mws_run(p, bc, o, F, rep, s, radii);
    if (rc != CATHODE_OK) return rc;
    for (int k = 0; k < 3; k++)
        CATHODE_HIP(hipMemcpyAsync(cin[3 + k], C + (3 + k) * n2, n2 * sizeof(cxd), hipMemcpyDeviceToHost, s));
    if (o->which == CATHODE_MW_IMPEDANCE) {
        for (int k = 0; k < 3; k++)
            if (cin[k]) CATHODE_HIP(hipMemcpyAsync(cin[k], C + k * n2, n2 * sizeof(cxd), hipMemcpyDeviceToHost, s));
        CATHODE_HIP(hipMemcpyAsync(f->rhoe, R, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
        CATHODE_HIP(hipMemcpyAsync(f->nu_m, R + n2, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    CATHODE_HIP(hipStreamSynchronize(s));
    return CATHODE_OK;
    CATHODE_GUARD_END
}

extern "C" int cathode_mw_solve_dev(const cathode_mw_params* p, const cathode_mw_bc* bc,
                                    const cathode_mw_solve_opts* o, cathode_mw_solve_fields* f,
                                    cathode_mw_solve_report* rep)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    int rc = mws_check(p, bc, o, f, rep);
    if (rc != CATHODE_OK) return rc;
    hipStream_t s = ctx().stream;
    const size_t n2 = (size_t)(p->ni + 2) * (p->nj + 2);
    // zero stand-ins for excitation currents the caller leaves out
    g_mws.cstage.alloc(n2);
    CATHODE_HIP(hipMemsetAsync(g_mws.cstage.p, 0, n2 * sizeof(cxd), s));
    const double* radii = upload_radii(p, s);
    if (o->which == CATHODE_MW_IMPEDANCE) {
        CATHODE_HIP(hipMemsetAsync(f->rhoe, 0, n2 * sizeof(double), s));
        hipLaunchKernelGGL(k_mws_fill, dim3(cdiv((long)n2, 256)), dim3(256), 0, s, f->nu_m, n2, o->nu_eff);
    }
    Fields F{};
    F.rhoe = f->rhoe;
    F.nu = f->nu_m;
    F.Bx = f->Bx;
    F.Br = f->Br;
    F.epsr = f->epsr;
    F.Er_dc = f->Er_dc;
    F.act[0] = f->act_x;
    F.act[1] = f->act_r;
    F.act[2] = f->act_p;
    double* jin[3] = {f->J1x_exc, f->J1r_exc, f->J1p_exc};
    double* ein[3] = {f->E1x, f->E1r, f->E1p};
    bool zero_used = false;
    for (int k = 0; k < 3; k++) {
        F.J[k] = jin[k] ? reinterpret_cast<cxd*>(jin[k]) : g_mws.cstage.p;
        zero_used |= !jin[k];
        F.E[k] = reinterpret_cast<cxd*>(ein[k]);
    }
    if (zero_used && o->which == CATHODE_MW_IMPEDANCE) {
        // the normalisation scales the excitation currents: never the shared zero array
        for (int k = 0; k < 3; k++)
            if (!jin[k]) {
                set_error("mw_solve_dev: the impedance test rescales J1x_exc, J1r_exc, J1p_exc; pass all three");
                return CATHODE_ERR_ARG;
            }
    }
    rc = mws_run(p, bc, o, F, rep, s, radii);
    if (rc != CATHODE_OK) return rc;
    CATHODE_HIP(hipStreamSynchronize(s));
    return CATHODE_OK;
    CATHODE_GUARD_END
}
