/*
 * orc_mwsolve.c -- CPU FP64 restatement of the coupled microwave solve,
 * EmfieldModule::solve_Microwave (include/emfieldModule.hpp:2249-3321), and of
 * EmfieldModule::solve_Microwave_impedanceTest (:5528-6141).  TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Unknowns: E_x on the x faces with act_x = 1, E_r on the r faces with
 * act_r = 1, E_phi on the cells with act_p = 1.  Each row is the vector
 * Helmholtz equation of its component multiplied by dx^2 dr^2:
 *   aP E - sum_nb a_nb E_nb = b,  b = -i omega mu0 dx^2 dr^2 J_exc,
 * with the magnetised cold-plasma conductivity entering the diagonal
 * (i omega mu0 sigma_cc dx^2 dr^2) and the coupling slots (-i omega mu0
 * sigma_cc' times the averaging weights of the other two components):
 *   E_x row (:2302-2350): E, W, N, S -> E_x(i+-1, j), E_x(i, j+-1);
 *     NE, SE, NW, SW -> E_r(i, j+1), E_r(i, j), E_r(i-1, j+1), E_r(i-1, j);
 *     EE, WW -> E_phi(i, j), E_phi(i-1, j);
 *   E_r row (:2353-2412): E, W, N, S -> E_r neighbours;
 *     NE, SE, NW, SW -> E_x(i+1, j), E_x(i+1, j-1), E_x(i, j), E_x(i, j-1);
 *     NN, SS -> E_phi(i, j), E_phi(i, j-1);
 *   E_phi row (:2415-2454): E, W, N, S; EE, WW -> E_x(i+1, j), E_x(i, j);
 *     NN, SS -> E_r(i, j+1), E_r(i, j).
 * A slot whose target is not an unknown is dropped (:2834-2991).  The twelve
 * boundary lists fold the wall couplings (:2457-2773) and, after the solve,
 * write the ghost values (:3026-3264), in the reference's list order.
 * CATHODE_MW_IMPEDANCE keeps the E_r rows alone with their five-point part
 * (:5538-5809), overwrites rhoe = 0 and nu = nu_eff (:5531-5536), and
 * normalises the forward power at the reference plane (:5917-6139).
 *
 * The system is solved by banded LU with partial pivoting (row interchanges
 * within the band, the algorithm of LAPACK's zgbtrf) over the unknowns
 * numbered by x column (i outer; E_x, E_r, E_phi; j inner).  pow(x, 2) of
 * the reference is x * x (the same correctly rounded value).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

#define NSL ORC_MWS_NSLOT

static cplx cg(const double* a, size_t k) { return a[2 * k] + I * a[2 * k + 1]; }
static void cs(double* a, size_t k, cplx v)
{
    a[2 * k] = creal(v);
    a[2 * k + 1] = cimag(v);
}

typedef struct {
    int ni, nj, ld;
    double dx, dr, w;
    double* r;
    cplx* co;   /* [3][NSL][(ni+2)(nj+2)] */
} sys_t;

static cplx* slot(sys_t* S, int c, int k) { return S->co + ((size_t)c * NSL + k) * (size_t)(S->ni + 2) * S->ld; }

/* cold-plasma conductivity factors (the reference's Hallcmpx, Hallcmpr, coef) */
static void cond(double n, double nu, double Bx, double Br, double w, double masse, cplx* hx, cplx* hr, cplx* coef)
{
    const cplx EPS = 1e-100 + 1e-100 * I;
    cplx nus = nu + I * w;
    *hx = ORC_E0 * Bx / (masse * nus + EPS);
    *hr = ORC_E0 * Br / (masse * nus + EPS);
    cplx h2 = *hx * *hx + *hr * *hr;
    *coef = n * ORC_E0 * ORC_E0 / (masse * nus + EPS) / (1.0 + h2 + EPS);
}

/* target (component, i, j) of slot k of a row of component c at (i, j); -1: none */
static int target(int c, int k, int i, int j, int* ti, int* tj)
{
    static const int tab[3][NSL][3] = {
        /* E_x */ {{0, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {1, 0, 1}, {1, -1, 1}, {1, 0, 0},
                   {1, -1, 0}, {2, 0, 0}, {2, -1, 0}, {-1, 0, 0}, {-1, 0, 0}, {-1, 0, 0}},
        /* E_r */ {{1, 0, 0}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 0}, {0, 0, 0}, {0, 1, -1},
                   {0, 0, -1}, {-1, 0, 0}, {-1, 0, 0}, {2, 0, 0}, {2, 0, -1}, {-1, 0, 0}},
        /* E_phi */ {{2, 0, 0}, {2, 1, 0}, {2, -1, 0}, {2, 0, 1}, {2, 0, -1}, {-1, 0, 0}, {-1, 0, 0}, {-1, 0, 0},
                     {-1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {1, 0, 1}, {1, 0, 0}, {-1, 0, 0}},
    };
    *ti = i + tab[c][k][1];
    *tj = j + tab[c][k][2];
    return tab[c][k][0];
}

static int in_box(const int* b, int i, int j) { return i >= b[0] && i <= b[1] && j >= b[2] && j <= b[3]; }

static int in_union(const int* boxes, int n_bl, int i, int j)
{
    for (int b = 0; b < n_bl; b++)
        if (in_box(boxes + 4 * b, i, j)) return 1;
    return 0;
}

/* assembly of the three components (:2302-2454) on the boxes' union */
static void assemble(sys_t* S, const cathode_mw_params* p, const cathode_mw_solve_fields* f, int which)
{
    const int ni = p->ni, nj = p->nj, ld = S->ld;
    const double dx = p->dx, dr = p->dr, w = p->omega, *r = S->r;
    const double h4 = dx * dx * dr * dr;
    const cplx iwm = I * (w * ORC_MU0);
    const cplx bsrc = -I * (w * ORC_MU0 * dx * dx * dr * dr);
    const double *ne = f->rhoe, *nu = f->nu_m, *Bx = f->Bx, *Br = f->Br, *er = f->epsr;
    for (int i = 0; i <= ni + 1; i++)
        for (int j = 0; j <= nj + 1; j++) {
            const size_t q = (size_t)i * ld + j, qW = q - ld, qS = q - 1;
            cplx hx, hr, coef;
            if (which == CATHODE_MW_COUPLED && in_union(p->xface_box, p->n_bl, i, j)) {
                double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100), qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                double nut = (nu[q] + nu[qW]) / 2.0, rho = (ne[q] + ne[qW]) / 2.0;
                double bx = (Bx[q] + Bx[qW]) / 2.0, br = (Br[q] + Br[qW]) / 2.0, ep = (er[q] + er[qW]) / 2.0;
                cond(rho, nut, bx, br, w, p->masse, &hx, &hr, &coef);
                cplx sxr = coef * (hx * hr), sxp = coef * hr, sxx = coef * (1.0 + hx * hx);
                slot(S, 0, ORC_MWS_P)[q] = 2.0 * dx * dx + 2.0 * dr * dr
                    - ORC_EPS0 * ep * ORC_MU0 * w * w * dx * dx * dr * dr + iwm * sxx * dx * dx * dr * dr;
                slot(S, 0, ORC_MWS_E)[q] = dr * dr;
                slot(S, 0, ORC_MWS_W)[q] = dr * dr;
                slot(S, 0, ORC_MWS_N)[q] = qR * dx * dx;
                slot(S, 0, ORC_MWS_S)[q] = qL * dx * dx;
                slot(S, 0, ORC_MWS_B)[q] = bsrc * cg(f->J1x_exc, q);
                slot(S, 0, ORC_MWS_EE)[q] = -iwm * sxp * 0.5 * dx * dx * dr * dr;
                slot(S, 0, ORC_MWS_WW)[q] = -iwm * sxp * 0.5 * dx * dx * dr * dr;
                slot(S, 0, ORC_MWS_NE)[q] = -iwm * sxr * 0.25 * qR * dx * dx * dr * dr;
                slot(S, 0, ORC_MWS_NW)[q] = -iwm * sxr * 0.25 * qR * dx * dx * dr * dr;
                slot(S, 0, ORC_MWS_SE)[q] = -iwm * sxr * 0.25 * qL * dx * dx * dr * dr;
                slot(S, 0, ORC_MWS_SW)[q] = -iwm * sxr * 0.25 * qL * dx * dx * dr * dr;
            }
            if (in_union(p->rface_box, p->n_bl, i, j)) {
                double rt = (r[j - 1] + r[j]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0, rL = (r[j - 2] + r[j - 1]) / 2.0;
                double nut = (nu[q] + nu[qS]) / 2.0, rho = (ne[q] + ne[qS]) / 2.0;
                double bx = (r[j] * Bx[q] + r[j - 1] * Bx[qS]) / (2.0 * rt);
                double br = (r[j] * Br[q] + r[j - 1] * Br[qS]) / (2.0 * rt);
                double ep = (er[q] + er[qS]) / 2.0;
                cond(rho, nut, bx, br, w, p->masse, &hx, &hr, &coef);
                cplx srr = coef * (1.0 + hr * hr), srp = coef * (-hx), srx = coef * (hx * hr);
                double wx = orc_planted() == ORC_PLANT_MWS_RX_WEIGHT ? 0.5 : 0.25;
                slot(S, 1, ORC_MWS_P)[q] = rt * (1.0 / r[j] + 1.0 / r[j - 1]) * dx * dx + 2.0 * dr * dr
                    - ORC_EPS0 * ep * ORC_MU0 * w * w * dx * dx * dr * dr + iwm * srr * dx * dx * dr * dr;
                slot(S, 1, ORC_MWS_E)[q] = dr * dr;
                slot(S, 1, ORC_MWS_W)[q] = dr * dr;
                slot(S, 1, ORC_MWS_N)[q] = rR / r[j] * dx * dx;
                slot(S, 1, ORC_MWS_S)[q] = rL / r[j - 1] * dx * dx;
                slot(S, 1, ORC_MWS_B)[q] = bsrc * cg(f->J1r_exc, q);
                slot(S, 1, ORC_MWS_NN)[q] = -iwm * srp * 0.5 * r[j] / rt * dx * dx * dr * dr;
                slot(S, 1, ORC_MWS_SS)[q] = -iwm * srp * 0.5 * r[j - 1] / rt * dx * dx * dr * dr;
                slot(S, 1, ORC_MWS_NE)[q] = -iwm * srx * wx * r[j] / rt * dx * dx * dr * dr;
                slot(S, 1, ORC_MWS_NW)[q] = -iwm * srx * wx * r[j] / rt * dx * dx * dr * dr;
                slot(S, 1, ORC_MWS_SE)[q] = -iwm * srx * wx * r[j - 1] / rt * dx * dx * dr * dr;
                slot(S, 1, ORC_MWS_SW)[q] = -iwm * srx * wx * r[j - 1] / rt * dx * dx * dr * dr;
            }
            if (which == CATHODE_MW_COUPLED && in_union(p->cell_box, p->n_bl, i, j)) {
                double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100), qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                cond(ne[q], nu[q], Bx[q], Br[q], w, p->masse, &hx, &hr, &coef);
                cplx spr = coef * hx, spp = coef * 1.0, spx = coef * (-hr);
                slot(S, 2, ORC_MWS_P)[q] = 2.0 * dx * dx + 2.0 * dr * dr + dx * dx * dr * dr / (r[j] * r[j])
                    - ORC_EPS0 * er[q] * ORC_MU0 * w * w * dx * dx * dr * dr + iwm * spp * dx * dx * dr * dr;
                slot(S, 2, ORC_MWS_E)[q] = dr * dr;
                slot(S, 2, ORC_MWS_W)[q] = dr * dr;
                slot(S, 2, ORC_MWS_N)[q] = qR * dx * dx;
                slot(S, 2, ORC_MWS_S)[q] = qL * dx * dx;
                slot(S, 2, ORC_MWS_B)[q] = bsrc * cg(f->J1p_exc, q);
                slot(S, 2, ORC_MWS_EE)[q] = -iwm * spx * 0.5 * dx * dx * dr * dr;
                slot(S, 2, ORC_MWS_WW)[q] = -iwm * spx * 0.5 * dx * dx * dr * dr;
                slot(S, 2, ORC_MWS_NN)[q] = -iwm * spr * 0.5 * qR * dx * dx * dr * dr;
                slot(S, 2, ORC_MWS_SS)[q] = -iwm * spr * 0.5 * qL * dx * dx * dr * dr;
            }
            (void)h4;
        }
}

/* first-order absorbing reflection factors: x faces (:2584-2586) and the
 * high-r wall (:2540-2550) */
static cplx refl_x(double c0, double w, double dx)
{
    double deno = (2.0 * c0) * (2.0 * c0) + (w * dx) * (w * dx);
    cplx R = ((2.0 * c0) * (2.0 * c0) - (w * dx) * (w * dx)) - I * (4.0 * c0 * w * dx);
    return R / (deno + 1e-100);
}
static double refl_x_deno(double c0, double w, double dx) { return (2.0 * c0) * (2.0 * c0) + (w * dx) * (w * dx); }

static cplx refl_r(double c0, double w, double dr, double rj, double rj1)
{
    double rt = (rj1 + rj) / 2.0;
    double qL = rj / (rt + 1e-100), qR = rj1 / (rt + 1e-100);
    double t = 2.0 * c0 * sqrt(qR), u = w * dr * qR;
    double deno = t * t + u * u;
    cplx R = ((2.0 * c0) * (2.0 * c0) * sqrt(qL * qR) - (w * dr) * (w * dr) * qL * qR)
        - I * (2.0 * c0 * w * dr * (qL * sqrt(qR) + qR * sqrt(qL)));
    return R / (deno + 1e-100);
}

static double c0_of(double epsr) { return sqrt(1.0 / (ORC_EPS0 * epsr * ORC_MU0)); }

/* the E_r excitation (:2606-2620): ghost = Rx E + Sx */
static void excite(const cathode_mw_params* p, double epsr, double rj, double rjm, cplx* Rd, cplx* Sd)
{
    double c0 = c0_of(epsr);
    cplx deno = 2.0 * c0 + I * (p->omega * p->dx);
    cplx R = 2.0 * c0 - I * (p->omega * p->dx);
    double rt = (rj + rjm) / 2.0;
    double Efw = p->J1r_exc / rt / log(p->r2 / p->r1) * 1e3;
    double sg = orc_planted() == ORC_PLANT_MWS_EXCITE_SIGN ? -1.0 : 1.0;
    cplx S = (I * (sg * 4.0 * p->omega * p->dx)) * Efw;
    *Rd = R / (deno + 1e-100);
    *Sd = S / (deno + 1e-100);
}

static void fold(sys_t* S, const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_fields* f,
                 const cathode_mw_solve_opts* o)
{
    const int ld = S->ld;
    const double dx = p->dx, dr = p->dr, w = p->omega, *r = S->r, *er = f->epsr;
    if (o->which == CATHODE_MW_COUPLED) {
        for (int side = 0; side < 4; side++) {
            const cathode_mw_bc_list* L = &bc->x[side];
            for (int k = 0; k < L->n; k++) {
                const int i = L->i[k], j = L->j[k], s = L->s[k];
                const size_t q = (size_t)i * ld + j;
                cplx *P = slot(S, 0, ORC_MWS_P) + q, *E = slot(S, 0, ORC_MWS_E) + q, *W = slot(S, 0, ORC_MWS_W) + q;
                cplx *N = slot(S, 0, ORC_MWS_N) + q, *Ss = slot(S, 0, ORC_MWS_S) + q;
                cplx *NE = slot(S, 0, ORC_MWS_NE) + q, *NW = slot(S, 0, ORC_MWS_NW) + q;
                cplx *SE = slot(S, 0, ORC_MWS_SE) + q, *SW = slot(S, 0, ORC_MWS_SW) + q;
                double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100), qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                if (side == 0) {
                    if (s == 0) {
                        *W = 0.0;
                    } else if (s == 2) {
                        *P -= *W;
                        *NW += *W * dx / dr * qR;
                        *SW -= *W * dx / dr * qL;
                        *W = 0.0;
                    } else if (s == 3) {
                        double deno = 1.0 + dx / dr * (orc_planted() == ORC_PLANT_MWS_CONCAVE_DENO ? qR : qL);
                        *P -= *W / deno;
                        *NW += *W * (dx / dr * qR) / deno;
                        *W = 0.0;
                    } else if (s == 4) {
                        double deno = 1.0 + dx / dr * qR;
                        *P -= *W / deno;
                        *SW -= *W * (dx / dr * qL) / deno;
                        *W = 0.0;
                    }
                } else if (side == 1) {
                    if (s == 0) {
                        *E = 0.0;
                    } else if (s == 2) {
                        *P -= *E;
                        *NE -= *E * dx / dr * qR;
                        *SE += *E * dx / dr * qL;
                        *E = 0.0;
                    } else if (s == 4) {
                        double deno = 1.0 + dx / dr * qR;
                        *P -= *E / deno;
                        *SE += *E * (dx / dr * qL) / deno;
                        *E = 0.0;
                    }
                } else if (side == 2) {
                    if (s == 0) {
                        *P += r[j] / r[j - 1] * *Ss;
                        *Ss = 0.0;
                    } else if (s == 6) {
                        *P -= *Ss * (dr / dx * r[j] / rL);
                        *W += *Ss * (-dr / dx * r[j] / rL);
                        *NW += *Ss * (rR / rL);
                        *Ss = 0.0;
                    }
                } else {
                    if (s == 0) {
                        *P += r[j] / r[j + 1] * *N;
                        *N = 0.0;
                    } else if (s == 1) {
                        cplx R = refl_r(c0_of((er[q] + er[q - ld]) / 2.0), w, dr, r[j], r[j + 1]);
                        *P -= *N * R;
                        *N = 0.0;
                    } else if (s == 5) {
                        *P -= *N * (dr / dx * r[j] / rR);
                        *E += *N * (-dr / dx * r[j] / rR);
                        *SE += *N * (rL / rR);
                        *N = 0.0;
                    } else if (s == 6) {
                        *P -= *N * (dr / dx * r[j] / rR);
                        *W += *N * (-dr / dx * r[j] / rR);
                        *SW += *N * (-rL / rR);
                        *N = 0.0;
                    }
                }
            }
        }
    }
    for (int side = 0; side < 4; side++) {
        const cathode_mw_bc_list* L = &bc->r[side];
        for (int k = 0; k < L->n; k++) {
            const int i = L->i[k], j = L->j[k], s = L->s[k];
            const size_t q = (size_t)i * ld + j;
            cplx *P = slot(S, 1, ORC_MWS_P) + q, *E = slot(S, 1, ORC_MWS_E) + q, *W = slot(S, 1, ORC_MWS_W) + q;
            cplx *N = slot(S, 1, ORC_MWS_N) + q, *Ss = slot(S, 1, ORC_MWS_S) + q, *B = slot(S, 1, ORC_MWS_B) + q;
            cplx *NE = slot(S, 1, ORC_MWS_NE) + q, *NW = slot(S, 1, ORC_MWS_NW) + q;
            cplx *SE = slot(S, 1, ORC_MWS_SE) + q, *SW = slot(S, 1, ORC_MWS_SW) + q;
            const double ep = (er[q] + er[q - 1]) / 2.0;
            if (side == 0) {
                if (s == 0) {
                    *P += *W;
                    *W = 0.0;
                } else if (s == 1) {
                    *P -= *W * refl_x(c0_of(ep), w, dx);
                    *W = 0.0;
                } else if (s == 5) {
                    double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                    *P -= *W * (dx / dr * rL / r[j]);
                    *N += *W * (-dx / dr * rR / r[j]);
                    *NE += *W * (-1.0);
                    *W = 0.0;
                } else if (s == 6) {
                    double rR = (r[j] + r[j - 1]) / 2.0, rL = (r[j - 1] + r[j - 2]) / 2.0;
                    *P -= *W * (dx / dr * rR / r[j - 1]);
                    *Ss += *W * (-dx / dr * rL / r[j - 1]);
                    *SE += *W * (1.0);
                    *W = 0.0;
                } else if (s == 7) {
                    cplx Rd, Sd;
                    excite(p, ep, r[j], r[j - 1], &Rd, &Sd);
                    *P -= *W * Rd;
                    *B += *W * Sd;
                    *W = 0.0;
                }
            } else if (side == 1) {
                if (s == 0) {
                    *P += *E;
                    *E = 0.0;
                } else if (s == 1) {
                    *P -= *E * refl_x(c0_of(ep), w, dx);
                    *E = 0.0;
                } else if (s == 5) {
                    double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                    *P -= *E * (dx / dr * rL / r[j]);
                    *N += *E * (-dx / dr * rR / r[j]);
                    *NW += *E * (1.0);
                    *E = 0.0;
                }
            } else if (side == 2) {
                double rR = (r[j] + r[j - 1]) / 2.0, rL = (r[j - 1] + r[j - 2]) / 2.0;
                if (s == 2) {
                    if (orc_planted() == ORC_PLANT_MWS_GAUSS_RADII)
                        *P -= *Ss * rL / rR;
                    else
                        *P -= *Ss * rR / rL;
                    *SE += *Ss * dr / dx * r[j - 1] / rL;
                    *SW -= *Ss * dr / dx * r[j - 1] / rL;
                    *Ss = 0.0;
                } else if (s == 3) {
                    double deno = dr / dx * r[j - 1] / rL;
                    if (r[j] < (double)o->j_flc0) deno = 1.0;
                    *P -= *Ss * rR / rL / deno;
                    *NE += *Ss * dr / dx * r[j - 1] / rL / deno;
                    *Ss = 0.0;
                }
            } else {
                double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                double deno = dr / dx * r[j] / rR;
                if (s == 2) {
                    *P -= *N * rL / rR;
                    *NE -= *N * dr / dx * r[j] / rR;
                    *NW += *N * dr / dx * r[j] / rR;
                    *N = 0.0;
                } else if (s == 3) {
                    *P -= *N * rL / rR / deno;
                    *SE -= *N * dr / dx * r[j] / rR / deno;
                    *N = 0.0;
                } else if (s == 4) {
                    *P -= *N * rL / rR / deno;
                    *SW += *N * dr / dx * r[j] / rR / deno;
                    *N = 0.0;
                }
            }
        }
    }
    if (o->which != CATHODE_MW_COUPLED) return;
    for (int side = 0; side < 4; side++) {
        const cathode_mw_bc_list* L = &bc->p[side];
        for (int k = 0; k < L->n; k++) {
            const int i = L->i[k], j = L->j[k], s = L->s[k];
            const size_t q = (size_t)i * ld + j;
            cplx *P = slot(S, 2, ORC_MWS_P) + q, *E = slot(S, 2, ORC_MWS_E) + q, *W = slot(S, 2, ORC_MWS_W) + q;
            cplx *N = slot(S, 2, ORC_MWS_N) + q, *Ss = slot(S, 2, ORC_MWS_S) + q;
            if (side == 0) {
                if (s == 0) {
                    *P += *W;
                    *W = 0.0;
                } else if (s == 1) {
                    *P -= *W * refl_x(c0_of(er[q]), w, dx);
                    *W = 0.0;
                }
            } else if (side == 1) {
                if (s == 0) {
                    *P += *E;
                    *E = 0.0;
                }
            } else if (side == 2) {
                if (s == 0) {
                    *P += r[j] / r[j - 1] * *Ss;
                    *Ss = 0.0;
                }
            } else {
                if (s == 0) {
                    *P += r[j] / r[j + 1] * *N;
                    *N = 0.0;
                } else if (s == 1) {
                    *P -= *N * refl_r(c0_of(er[q]), w, dr, r[j], r[j + 1]);
                    *N = 0.0;
                }
            }
        }
    }
}

/* ghost values after the solve (:3026-3264; E_r only for the impedance test, :5811-5914) */
static void post(const sys_t* S, const cathode_mw_params* p, const cathode_mw_bc* bc,
                 const cathode_mw_solve_fields* f, const cathode_mw_solve_opts* o)
{
    const int ld = S->ld;
    const double dx = p->dx, dr = p->dr, w = p->omega, *r = S->r, *er = f->epsr;
    double *Ex = f->E1x, *Er = f->E1r, *Ep = f->E1p;
    if (o->which == CATHODE_MW_COUPLED) {
        for (int side = 0; side < 4; side++) {
            const cathode_mw_bc_list* L = &bc->x[side];
            for (int k = 0; k < L->n; k++) {
                const int i = L->i[k], j = L->j[k], s = L->s[k];
                const size_t q = (size_t)i * ld + j;
                double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100), qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                if (side == 0) {
                    if (s == 0)
                        cs(Ex, q - ld, 0.0);
                    else if (s == 2)
                        cs(Ex, q - ld, cg(Ex, q) + dx / dr * (qR * cg(Er, q - ld + 1) - qL * cg(Er, q - ld)));
                    else if (s == 3) {
                        double deno = 1.0 + dx / dr * qL;
                        cs(Ex, q - ld, cg(Ex, q) / deno + dx / dr * (qR * cg(Er, q - ld + 1)) / deno);
                    } else if (s == 4) {
                        double deno = 1.0 + dx / dr * qR;
                        cs(Ex, q - ld, cg(Ex, q) / deno - dx / dr * (qL * cg(Er, q - ld)) / deno);
                    }
                } else if (side == 1) {
                    if (s == 0)
                        cs(Ex, q + ld, 0.0);
                    else if (s == 2)
                        cs(Ex, q + ld, cg(Ex, q) - dx / dr * (qR * cg(Er, q + 1) - qL * cg(Er, q)));
                    else if (s == 4) {
                        double deno = 1.0 + dx / dr * qR;
                        cs(Ex, q + ld, cg(Ex, q) / deno + dx / dr * qL * cg(Er, q) / deno);
                    }
                } else if (side == 2) {
                    if (s == 0) cs(Ex, q - 1, -r[j] / r[j - 1] * cg(Ex, q));
                } else {
                    if (s == 0)
                        cs(Ex, q + 1, -r[j] / r[j + 1] * cg(Ex, q));
                    else if (s == 1)
                        cs(Ex, q + 1, refl_r(c0_of((er[q] + er[q - ld]) / 2.0), w, dr, r[j], r[j + 1]) * cg(Ex, q));
                }
            }
        }
    }
    for (int side = 0; side < 4; side++) {
        const cathode_mw_bc_list* L = &bc->r[side];
        for (int k = 0; k < L->n; k++) {
            const int i = L->i[k], j = L->j[k], s = L->s[k];
            const size_t q = (size_t)i * ld + j;
            const double ep = (er[q] + er[q - 1]) / 2.0;
            if (side == 0) {
                if (s == 0)
                    cs(Er, q - ld, -cg(Er, q));
                else if (s == 1)
                    cs(Er, q - ld, refl_x(c0_of(ep), w, dx) * cg(Er, q));
                else if (s == 7) {
                    cplx Rd, Sd;
                    excite(p, ep, r[j], r[j - 1], &Rd, &Sd);
                    cs(Er, q - ld, Rd * cg(Er, q) + Sd);
                }
            } else if (side == 1) {
                if (s == 0)
                    cs(Er, q + ld, -cg(Er, q));
                else if (s == 1) {
                    double c0 = c0_of(ep);
                    cplx R = refl_x(c0, w, dx);
                    double edc = f->Er_dc ? f->Er_dc[q] : 0.0;
                    cs(Er, q + ld, R / (refl_x_deno(c0, w, dx) + 1e-100) * edc);
                }
            } else if (side == 2) {
                double rR = (r[j] + r[j - 1]) / 2.0, rL = (r[j - 1] + r[j - 2]) / 2.0;
                if (s == 2)
                    cs(Er, q - 1,
                       rR / rL * cg(Er, q) + dr / dx * r[j - 1] / rL * (cg(Ex, q + ld - 1) - cg(Ex, q - 1)));
                else if (s == 3) {
                    double deno = dr / dx * r[j - 1] / rL;
                    if (r[j] < (double)o->j_flc0) deno = 1.0;
                    cs(Er, q - 1, rR / rL * cg(Er, q) / deno + dr / dx * r[j - 1] / rL * (cg(Ex, q + ld - 1)) / deno);
                }
            } else {
                double rR = (r[j + 1] + r[j]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                double deno = dr / dx * r[j] / rR;
                if (s == 2)
                    cs(Er, q + 1, rL / rR * cg(Er, q) - dr / dx * r[j] / rR * (cg(Ex, q + ld) - cg(Ex, q)));
                else if (s == 3)
                    cs(Er, q + 1, rL / rR * cg(Er, q) / deno - dr / dx * r[j] / rR * (cg(Ex, q + ld)) / deno);
                else if (s == 4)
                    cs(Er, q + 1, rL / rR * cg(Er, q) / deno + dr / dx * r[j] / rR * (cg(Ex, q)) / deno);
            }
        }
    }
    if (o->which != CATHODE_MW_COUPLED) return;
    for (int side = 0; side < 4; side++) {
        const cathode_mw_bc_list* L = &bc->p[side];
        for (int k = 0; k < L->n; k++) {
            const int i = L->i[k], j = L->j[k], s = L->s[k];
            const size_t q = (size_t)i * ld + j;
            if (side == 0) {
                if (s == 0)
                    cs(Ep, q - ld, -cg(Ep, q));
                else if (s == 1)
                    cs(Ep, q - ld, refl_x(c0_of(er[q]), w, dx) * cg(Ep, q));
            } else if (side == 1) {
                cs(Ep, q + ld, -cg(Ep, q));
            } else if (side == 2) {
                cs(Ep, q - 1, -r[j] / r[j - 1] * cg(Ep, q));
            } else {
                if (s == 0)
                    cs(Ep, q + 1, -r[j] / r[j + 1] * cg(Ep, q));
                else if (s == 1)
                    cs(Ep, q + 1, refl_r(c0_of(er[q]), w, dr, r[j], r[j + 1]) * cg(Ep, q));
            }
        }
    }
}

static const int* act_of(const cathode_mw_solve_fields* f, int c) { return c == 0 ? f->act_x : c == 1 ? f->act_r : f->act_p; }

static int comp_on(int which, int c) { return which == CATHODE_MW_COUPLED || c == 1; }

/* slot k of a component-c row is part of the system (the impedance test keeps the five-point part) */
static int slot_on(int which, int k) { return which == CATHODE_MW_COUPLED || k <= ORC_MWS_S; }

static int check_args(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                      const cathode_mw_solve_fields* f)
{
    if (!p || !bc || !o || !f || p->ni < 1 || p->nj < 2 || p->n_bl < 1 || p->n_bl > CATHODE_MW_MAX_BLOCKS) return -1;
    if (o->which != CATHODE_MW_COUPLED && o->which != CATHODE_MW_IMPEDANCE) return -1;
    for (int b = 0; b < p->n_bl; b++) {
        const int *x = p->xface_box + 4 * b, *rr = p->rface_box + 4 * b, *c = p->cell_box + 4 * b;
        if (x[0] < 1 || x[1] > p->ni + 1 || x[2] < 1 || x[3] > p->nj) return -1;
        if (rr[0] < 0 || rr[1] > p->ni + 1 || rr[2] < 2 || rr[3] > p->nj) return -1;
        if (c[0] < 0 || c[1] > p->ni + 1 || c[2] < 1 || c[3] > p->nj) return -1;
    }
    const cathode_mw_bc_list* all[12];
    for (int s = 0; s < 4; s++) {
        all[s] = &bc->x[s];
        all[4 + s] = &bc->r[s];
        all[8 + s] = &bc->p[s];
    }
    for (int l = 0; l < 12; l++)
        for (int k = 0; k < all[l]->n; k++) {
            int i = all[l]->i[k], j = all[l]->j[k];
            if (i < 1 || i > p->ni || j < 1 || j > p->nj) return -1;
            if (l >= 4 && l < 8 && j < 2) return -1;
        }
    return 0;
}

int orc_mw_rows(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                const cathode_mw_solve_fields* f, double* coef)
{
    if (check_args(p, bc, o, f)) return -1;
    sys_t S;
    S.ni = p->ni;
    S.nj = p->nj;
    S.ld = p->nj + 2;
    S.dx = p->dx;
    S.dr = p->dr;
    S.w = p->omega;
    S.r = malloc(sizeof(double) * (p->nj + 2));
    for (int j = 0; j < p->nj + 2; j++) S.r[j] = p->rmin + (j - 0.5) * p->dr;
    const size_t n2 = (size_t)(p->ni + 2) * S.ld;
    S.co = calloc(3 * NSL * n2, sizeof(cplx));
    assemble(&S, p, f, o->which);
    fold(&S, p, bc, f, o);
    for (size_t k = 0; k < 3 * NSL * n2; k++) cs(coef, k, S.co[k]);
    free(S.co);
    free(S.r);
    return 0;
}

static int active(const cathode_mw_params* p, const cathode_mw_solve_fields* f, int which, int c, int i, int j)
{
    if (!comp_on(which, c) || c < 0) return 0;
    if (i < 0 || i > p->ni + 1 || j < 0 || j > p->nj + 1) return 0;
    return act_of(f, c)[(size_t)i * (p->nj + 2) + j] == 1;
}

int orc_mw_apply(const cathode_mw_params* p, const cathode_mw_solve_fields* f, int which, const double* coef,
                 const double* Ex, const double* Er, const double* Ep, double* yx, double* yr, double* yp)
{
    const int ld = p->nj + 2;
    const size_t n2 = (size_t)(p->ni + 2) * ld;
    const double* E[3] = {Ex, Er, Ep};
    double* Y[3] = {yx, yr, yp};
    for (int c = 0; c < 3; c++) {
        memset(Y[c], 0, 2 * n2 * sizeof(double));
        for (int i = 0; i <= p->ni + 1; i++)
            for (int j = 0; j <= p->nj + 1; j++) {
                if (!active(p, f, which, c, i, j)) continue;
                const size_t q = (size_t)i * ld + j;
                cplx y = cg(coef, ((size_t)c * NSL + ORC_MWS_P) * n2 + q) * cg(E[c], q);
                for (int k = 1; k < ORC_MWS_B; k++) {
                    int ti, tj, tc = target(c, k, i, j, &ti, &tj);
                    if (tc < 0 || !slot_on(which, k) || !active(p, f, which, tc, ti, tj)) continue;
                    y -= cg(coef, ((size_t)c * NSL + k) * n2 + q) * cg(E[tc], (size_t)ti * ld + tj);
                }
                cs(Y[c], q, y);
            }
    }
    return 0;
}

/* mWLoss_at_boundary_xR (:8537-8579): Poynting flux through the x face i + 1/2 */
static double loss_xR(const cathode_mw_params* p, const double* r, const double* Ex, const double* Er,
                      const double* Ep, int i, int jlo, int jhi)
{
    const int ld = p->nj + 2;
    const cplx iwm = I * (1.0 / (p->omega * ORC_MU0));
    double P = 0.0;
    for (int j = jlo; j <= jhi; j++) {
        double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
        size_t q = (size_t)i * ld + j;
        cplx dErdx_L = (cg(Er, q + ld) - cg(Er, q)) / p->dx;
        cplx dExdr_L = (cg(Ex, q + ld) - cg(Ex, q + ld - 1)) / p->dr;
        cplx dErdx_R = (cg(Er, q + ld + 1) - cg(Er, q + 1)) / p->dx;
        cplx dExdr_R = (cg(Ex, q + ld + 1) - cg(Ex, q + ld)) / p->dr;
        cplx HL = iwm * (dErdx_L - dExdr_L), HR = iwm * (dErdx_R - dExdr_R);
        cplx Hp = (rR * HR + rL * HL) / (2.0 * r[j]);
        cplx Ept = (cg(Ep, q + ld) + cg(Ep, q)) / (2.0 * rR);
        cplx ErL = (rR * cg(Er, q + ld + 1) + rL * cg(Er, q + ld)) / (2.0 * r[j]);
        cplx ErR = (rR * cg(Er, q + 1) + rL * cg(Er, q)) / (2.0 * r[j]);
        cplx Ert = (ErR + ErL) / 2.0;
        cplx ExH = Ert * conj(Hp) - Ept * conj(Hp);
        P += 2.0 * M_PI * r[j] * (0.5 * creal(ExH)) * p->dr;
    }
    return P;
}

int orc_mw_solve(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                 cathode_mw_solve_fields* f, cathode_mw_solve_report* rep)
{
    if (check_args(p, bc, o, f)) return -1;
    const int ni = p->ni, nj = p->nj, ld = nj + 2;
    const size_t n2 = (size_t)(ni + 2) * ld;
    if (o->which == CATHODE_MW_IMPEDANCE)
        for (size_t q = 0; q < n2; q++) {
            f->rhoe[q] = 0.0;
            f->nu_m[q] = o->nu_eff;
        }
    double* coef = malloc(sizeof(double) * 2 * 3 * NSL * n2);
    orc_mw_rows(p, bc, o, f, coef);
    double* r = malloc(sizeof(double) * ld);
    for (int j = 0; j < ld; j++) r[j] = p->rmin + (j - 0.5) * p->dr;

    /* number the unknowns column by column */
    int* id = malloc(sizeof(int) * 3 * n2);
    int N = 0, cnt[3] = {0, 0, 0};
    for (int i = 0; i <= ni + 1; i++)
        for (int c = 0; c < 3; c++)
            for (int j = 0; j <= nj + 1; j++) {
                size_t q = (size_t)i * ld + j;
                if (active(p, f, o->which, c, i, j)) {
                    id[c * n2 + q] = N++;
                    cnt[c]++;
                } else
                    id[c * n2 + q] = -1;
            }
    /* bandwidths */
    int kl = 0, ku = 0;
    for (int c = 0; c < 3; c++)
        for (int i = 0; i <= ni + 1; i++)
            for (int j = 0; j <= nj + 1; j++) {
                int row = id[c * n2 + (size_t)i * ld + j];
                if (row < 0) continue;
                for (int k = 1; k < ORC_MWS_B; k++) {
                    int ti, tj, tc = target(c, k, i, j, &ti, &tj);
                    if (tc < 0 || !slot_on(o->which, k) || !active(p, f, o->which, tc, ti, tj)) continue;
                    int col = id[tc * n2 + (size_t)ti * ld + tj];
                    if (col < row && row - col > kl) kl = row - col;
                    if (col > row && col - row > ku) ku = col - row;
                }
            }
    const int Wd = 2 * kl + ku + 1;
    cplx* AB = calloc((size_t)N * Wd, sizeof(cplx));
    cplx* x = calloc((size_t)N, sizeof(cplx));
    cplx* bvec = calloc((size_t)N, sizeof(cplx));
#define AT(rw, cl) AB[(size_t)(rw) * Wd + ((cl) - (rw) + kl)]
    for (int c = 0; c < 3; c++)
        for (int i = 0; i <= ni + 1; i++)
            for (int j = 0; j <= nj + 1; j++) {
                size_t q = (size_t)i * ld + j;
                int row = id[c * n2 + q];
                if (row < 0) continue;
                AT(row, row) = cg(coef, ((size_t)c * NSL + ORC_MWS_P) * n2 + q);
                bvec[row] = cg(coef, ((size_t)c * NSL + ORC_MWS_B) * n2 + q);
                for (int k = 1; k < ORC_MWS_B; k++) {
                    int ti, tj, tc = target(c, k, i, j, &ti, &tj);
                    if (tc < 0 || !slot_on(o->which, k) || !active(p, f, o->which, tc, ti, tj)) continue;
                    int col = id[tc * n2 + (size_t)ti * ld + tj];
                    AT(row, col) = -cg(coef, ((size_t)c * NSL + k) * n2 + q);
                }
            }
    /* banded LU with partial pivoting */
    int* piv = malloc(sizeof(int) * (N > 0 ? N : 1));
    int rc = 0;
    for (int k = 0; k < N; k++) {
        int last = k + kl < N - 1 ? k + kl : N - 1;
        int pr = k;
        double best = cabs(AT(k, k));
        for (int rw = k + 1; rw <= last; rw++)
            if (cabs(AT(rw, k)) > best) {
                best = cabs(AT(rw, k));
                pr = rw;
            }
        piv[k] = pr;
        if (best == 0.0) {
            rc = -2;
            break;
        }
        int cend = k + ku + kl < N - 1 ? k + ku + kl : N - 1;
        if (pr != k)
            for (int cl = k; cl <= cend; cl++) {
                cplx t = AT(k, cl);
                AT(k, cl) = AT(pr, cl);
                AT(pr, cl) = t;
            }
        for (int rw = k + 1; rw <= last; rw++) {
            cplx l = AT(rw, k) / AT(k, k);
            AT(rw, k) = l;
            if (l != 0.0)
                for (int cl = k + 1; cl <= cend; cl++) AT(rw, cl) -= l * AT(k, cl);
        }
    }
    if (rc == 0) {
        for (int k = 0; k < N; k++) x[k] = bvec[k];
        for (int k = 0; k < N; k++) {
            if (piv[k] != k) {
                cplx t = x[k];
                x[k] = x[piv[k]];
                x[piv[k]] = t;
            }
            int last = k + kl < N - 1 ? k + kl : N - 1;
            for (int rw = k + 1; rw <= last; rw++) x[rw] -= AT(rw, k) * x[k];
        }
        for (int k = N - 1; k >= 0; k--) {
            int cend = k + ku + kl < N - 1 ? k + ku + kl : N - 1;
            cplx s = x[k];
            for (int cl = k + 1; cl <= cend; cl++) s -= AT(k, cl) * x[cl];
            x[k] = s / AT(k, k);
        }
        double* E[3] = {f->E1x, f->E1r, f->E1p};
        for (int c = 0; c < 3; c++)
            for (size_t q = 0; q < n2; q++)
                if (id[c * n2 + q] >= 0) cs(E[c], q, x[id[c * n2 + q]]);
    }
#undef AT
    if (rep) {
        memset(rep, 0, sizeof(*rep));
        rep->n_x = cnt[0];
        rep->n_r = cnt[1];
        rep->n_p = cnt[2];
        rep->block = kl;
    }
    if (rc == 0) {
        /* residual of the solved field with the sparse rows */
        double* y[3];
        for (int c = 0; c < 3; c++) y[c] = calloc(2 * n2, sizeof(double));
        orc_mw_apply(p, f, o->which, coef, f->E1x, f->E1r, f->E1p, y[0], y[1], y[2]);
        double rr = 0.0, bb = 0.0;
        for (int c = 0; c < 3; c++)
            for (size_t q = 0; q < n2; q++) {
                if (id[c * n2 + q] < 0) continue;
                cplx bq = cg(coef, ((size_t)c * NSL + ORC_MWS_B) * n2 + q);
                cplx d = bq - cg(y[c], q);
                rr += creal(d) * creal(d) + cimag(d) * cimag(d);
                bb += creal(bq) * creal(bq) + cimag(bq) * cimag(bq);
            }
        for (int c = 0; c < 3; c++) free(y[c]);
        if (rep) rep->rel_residual = sqrt(rr) / (sqrt(bb) > 0.0 ? sqrt(bb) : 1.0);
        post(&(sys_t){ni, nj, ld, p->dx, p->dr, p->omega, r, NULL}, p, bc, f, o);
        if (o->which == CATHODE_MW_IMPEDANCE) {
            /* reference plane, port waves and the power normalisation (:5917-6139) */
            const double* Er = f->E1r;
            const int i = p->i_ref;
            const int jlo = p->rface_box[2], jhi = p->rface_box[3];
            cplx Isum = 0.0, csum = 0.0;
            int n = 0;
            for (int j = jlo; j <= jhi; j++) {
                double rt = (r[j] + r[j - 1]) / 2.0;
                size_t q = (size_t)i * ld + j;
                cplx H = (cg(Er, q) - cg(Er, q - ld)) / p->dx / (p->omega * ORC_MU0) * I;
                Isum += 2.0 * M_PI * rt * H;
                csum += (cg(Er, q) + cg(Er, q - ld)) / 2.0 * rt;
                n++;
            }
            cplx Iref = Isum / (double)n;
            cplx V = csum / (double)n * log(p->r2 / p->r1);
            cplx Z = V / (Iref + 1e-100);
            cplx a2 = (V - p->Z0 * Iref) / (2.0 * sqrt(2.0 * p->Z0));
            cplx b2 = (V + p->Z0 * Iref) / (2.0 * sqrt(2.0 * p->Z0));
            double Pf = creal(b2) * creal(b2) + cimag(b2) * cimag(b2);
            double Pr = creal(a2) * creal(a2) + cimag(a2) * cimag(a2);
            double Pout = loss_xR(p, r, f->E1x, f->E1r, f->E1p, p->cell_box[1], p->cell_box[2], p->cell_box[3]);
            double Perr = Pf - Pr - Pout;
            cplx S11 = p->S11[0] + I * p->S11[1], S12 = p->S12[0] + I * p->S12[1];
            cplx S21 = p->S21[0] + I * p->S21[1], S22 = p->S22[0] + I * p->S22[1];
            cplx a1 = (b2 - S22 * a2) / S21;
            cplx b1 = S11 * a1 + S12 * a2;
            double Pfa = creal(a1) * creal(a1) + cimag(a1) * cimag(a1);
            double Pra = creal(b1) * creal(b1) + cimag(b1) * cimag(b1);
            double ratio = p->Pmw / Pf, sr = sqrt(ratio);
            double* sc[6] = {f->E1x, f->E1r, f->E1p, f->J1x_exc, f->J1r_exc, f->J1p_exc};
            for (int a = 0; a < 6; a++)
                if (sc[a])
                    for (size_t q = 0; q < 2 * n2; q++) sc[a][q] = sc[a][q] * sr;
            if (rep) {
                rep->V_ref[0] = creal(V);
                rep->V_ref[1] = cimag(V);
                rep->I_ref[0] = creal(Iref);
                rep->I_ref[1] = cimag(Iref);
                rep->Z_ref[0] = creal(Z);
                rep->Z_ref[1] = cimag(Z);
                rep->ratio = ratio;
                rep->J1r_exc = p->J1r_exc * sr;
                rep->P_fwd = Pf * ratio;
                rep->P_ref = Pr * ratio;
                rep->P_fwd_all = Pfa * ratio;
                rep->P_ref_all = Pra * ratio;
                rep->P_err = Perr * ratio;
                rep->P_out = Pout * ratio;
            }
        }
    }
    free(piv);
    free(AB);
    free(x);
    free(bvec);
    free(id);
    free(r);
    free(coef);
    return rc;
}
