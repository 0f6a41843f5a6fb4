/*
 * orc_transport.c -- CPU FP64 restatement of FluidModule::
 * update_transport_coef (include/fluidModule.hpp:5799-5976) and of the table
 * lookups func_Te, func_dTe, func_Ei (include/miniFuncs.hpp:35-91), over the
 * geometry of orc_ndd_set_geometry.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of cathode_ndd_transport_update.
 *
 * The plasma cells and the wall flags (the reference's bLx_wall, bRx_wall,
 * bLr_wall, bRr_wall of :5826-5846) come from the cell boxes and from the
 * cells each boundary segment lists; the loops run block by block as the
 * reference's.  A table index below 0 (Te <= -dTe, where the reference reads
 * before its vector) reads entry 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IXT(i, j) ((size_t)(i) * (size_t)ld + (size_t)(j))

/* func_Te: linear between entries, the last entry from iR = n - 1 on (the
 * reference compares iR with size() - 1 as unsigned values) */
double orc_func_Te_tab(const double* v, int n, double Te, double dTe)
{
    int iL = (int)(Te / dTe);
    int iR = (int)(Te / dTe) + 1;
    double TeL = iL * dTe;
    if ((size_t)(long)iR < (size_t)n - 1) {
        if (iL < 0) iL = 0;
        return (v[iR] - v[iL]) / (dTe + 1e-100) * (Te - TeL) + v[iL];
    }
    return v[n - 1];
}

/* func_dTe: the slope of the interval, the last interval from iR = n - 1 on */
double orc_func_dTe_tab(const double* v, int n, double Te, double dTe)
{
    int iL = (int)(Te / dTe);
    int iR = (int)(Te / dTe) + 1;
    if ((size_t)(long)iR >= (size_t)n - 1) {
        iL = n - 2;
        iR = n - 1;
    }
    double TeL = iL * dTe;
    double TeR = iR * dTe;
    return (v[iR] - v[iL < 0 ? 0 : iL]) / (TeR - TeL + 1e-100);
}

/* func_Ei: linear, and beyond the table the last entry's cross-section
 * (rate / speed) times the speed sqrt(Ei / m_i) */
double orc_func_Ei_tab(const double* v, int n, double Ei, double dEi, double massi)
{
    int iL = (int)(Ei / dEi);
    int iR = (int)(Ei / dEi) + 1;
    double EiL = iL * dEi;
    if ((size_t)(long)iR < (size_t)n - 1) {
        if (iL < 0) iL = 0;
        return (v[iR] - v[iL]) / (dEi + 1e-100) * (Ei - EiL) + v[iL];
    }
    double imax = n - 1;
    double v_max = sqrt(imax * dEi / massi);
    double v_tmp = sqrt(Ei / massi);
    double sigma_max = v[n - 1] / v_max;
    return sigma_max * v_tmp;
}

static const double* zo(const double* a, const double* z) { return a ? a : z; }

int orc_transport_update(const cathode_ndd_params* grid, const cathode_transport_params* p,
                         const cathode_rate_tables* T, cathode_transport_fields* f)
{
    const int ni = grid->ni, nj = grid->nj, ld = nj + 2;
    const size_t n = (size_t)(ni + 2) * (size_t)ld;
    cathode_ndd_geometry G;
    orc_ndd_geometry(ni, nj, &G);
    double* r = (double*)calloc(nj + 2, sizeof(double));
    for (int j = 0; j < nj + 2; j++) r[j] = grid->rmin + (j - 0.5) * grid->dr; /* initialModule.hpp:86-88 */
    double* zero = (double*)calloc(n, sizeof(double));
    /* wall flags per cell (1 = no wall on that side) */
    unsigned char *wL = malloc(n), *wR = malloc(n), *wS = malloc(n), *wN = malloc(n);
    memset(wL, 1, n);
    memset(wR, 1, n);
    memset(wS, 1, n);
    memset(wN, 1, n);
    for (int s = 0; s < G.n_seg; s++) {
        const cathode_ndd_segment* q = &G.seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            if (q->kind == CATHODE_SEG_X_LOW) wL[IXT(q->cell, t)] = 0;
            else if (q->kind == CATHODE_SEG_X_HIGH) wR[IXT(q->cell, t)] = 0;
            else if (q->kind == CATHODE_SEG_R_LOW) wS[IXT(t, q->cell)] = 0;
            else if (q->kind == CATHODE_SEG_R_HIGH) wN[IXT(t, q->cell)] = 0;
        }
    }
    const double *rhoe = zo(f->rhoe, zero), *rhoi = zo(f->rhoi, zero), *rhon = zo(f->rhon, zero);
    const double *rhom = zo(f->rhom, zero), *Te = f->Te, *Bx = zo(f->Bx, zero), *Br = zo(f->Br, zero);
    const double *nu_ano = zo(f->nu_ano, zero), *Uix = zo(f->Uix, zero), *Uir = zo(f->Uir, zero);
    const double *Uip = zo(f->Uip, zero), *rhoUex = zo(f->rhoUex, zero), *rhoUer = zo(f->rhoUer, zero);
    double* iat = f->nu_ano_IAT;
    const double masse = p->masse, massi = p->massi, Ti = p->Ti;
    /* :5807-5818 */
    double rhoe_max = 0.0;
    for (int b = 0; b < G.n_bl; b++) {
        const int* B = G.cell_box + 4 * b;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++)
                if (rhoe_max < rhoe[IXT(i, j)]) rhoe_max = rhoe[IXT(i, j)];
    }
    /* :5821-5899 */
    for (int b = 0; b < G.n_bl; b++) {
        const int* B = G.cell_box + 4 * b;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                const size_t k = IXT(i, j);
                const double bLx = wL[k], bRx = wR[k], bLr = wS[k], bRr = wN[k];
                const double ut = sqrt(2.0 * ORC_BOLTZ * Te[k] / masse);
                const double rL = (r[j] + r[j - 1]) / 2.0;
                const double rR = (r[j] + r[j + 1]) / 2.0;
                const double fx = (rhoUex[IXT(i + 1, j)] * bRx + rhoUex[k] * bLx) / 2.0;
                const double fr = (rR * rhoUer[IXT(i, j + 1)] * bRr + rL * rhoUer[k] * bLr) / (2.0 * r[j]);
                const double Uex = fx / (rhoe[k] + rhoe_max * 0.001);
                const double Uer = fr / (rhoe[k] + rhoe_max * 0.001);
                const double Ux = (Uix[k] + Uix[IXT(i + 1, j)]) / 2.0;
                const double Ur = (rL * Uir[k] + rR * Uir[IXT(i, j + 1)]) / (2.0 * r[j]);
                if (f->Mache) f->Mache[k] = sqrt(pow(Uex, 2) + pow(Uer, 2)) / (ut + 100.0);
                const double cs = sqrt(2.0 * ORC_BOLTZ * (Te[k] + Ti) / massi);
                const double n_pi = orc_planted() == ORC_PLANT_TR_IAT_DENSITY ? rhoe[k] : rhoi[k];
                const double omega_pi = sqrt(n_pi * ORC_E0 * ORC_E0 / (massi * ORC_EPS0));
                const double ud = sqrt(pow(Uex - Ux, 2) + pow(Uer - Ur, 2));
                iat[k] = 0.01 * omega_pi * ud / cs * Te[k] / Ti;
            }
    }
    /* :5901-5974 */
    const int ne = T->n_e;
    const double dTe = T->dTe;
    for (int i = 0; i < ni + 2; i++)
        for (int j = 0; j < nj + 2; j++) {
            const size_t k = IXT(i, j);
            const double t = Te[k];
            const double v_elas = rhon[k] * orc_func_Te_tab(T->nu_elas, ne, t, dTe);
            const double v_exc = rhon[k] * orc_func_Te_tab(T->nu_exc, ne, t, dTe);
            const double v_reso = rhon[k] * orc_func_Te_tab(T->nu_excReso, ne, t, dTe);
            const double v_meta = rhon[k] * orc_func_Te_tab(T->nu_excMeta, ne, t, dTe);
            const double v_ionz = rhon[k] * orc_func_Te_tab(T->nu_ionz, ne, t, dTe);
            const double v_ionzStep = rhom[k] * orc_func_Te_tab(T->nu_ionzStep, ne, t, dTe);
            const double v_excStep = rhom[k] * orc_func_Te_tab(T->nu_excStep, ne, t, dTe);
            const double v_super = rhom[k] * orc_func_Te_tab(T->nu_super, ne, t, dTe);
            double* outs[8] = {f->nu_elas, f->nu_exc, f->nu_excReso, f->nu_excMeta,
                               f->nu_ionz, f->nu_ionzStep, f->nu_excStep, f->nu_super};
            const double vals[8] = {v_elas, v_exc, v_reso, v_meta, v_ionz, v_ionzStep, v_excStep, v_super};
            for (int q = 0; q < 8; q++)
                if (outs[q]) outs[q][k] = vals[q];
            const double nu_m_real = rhon[k] * orc_func_Te_tab(T->nu_elas, ne, t, dTe);
            const double nu_m = nu_m_real + (double)p->flag_Bohm * nu_ano[k] + (double)p->flag_Sagdeev * iat[k];
            if (f->nu_m) f->nu_m[k] = nu_m;
            if (f->nu_m1) f->nu_m1[k] = nu_m_real + p->nu_eff;
            if (f->rate_ionize) f->rate_ionize[k] = rhoe[k] * (v_ionz + v_ionzStep);
            double vv = 0.0;
            if (i != 0 && i != ni + 1 && j != 0 && j != nj + 1) {
                const double rL = (r[j] + r[j - 1]) / 2.0;
                const double rR = (r[j] + r[j + 1]) / 2.0;
                const double Ux = (Uix[k] + Uix[IXT(i + 1, j)]) / 2.0;
                const double Ur = (rL * Uir[k] + rR * Uir[IXT(i, j + 1)]) / (2.0 * r[j]);
                const double Up = Uip[k];
                vv = Ux * Ux + Ur * Ur + Up * Up;
            }
            const double Ei = massi * vv;
            if (f->nui_m) f->nui_m[k] = rhon[k] * orc_func_Ei_tab(T->nui_m, T->n_i, Ei, T->dEi, massi);
            if (f->rate_eloss)
                f->rate_eloss[k] = rhon[k] * orc_func_Te_tab(T->eloss_n, ne, t, dTe) +
                                   rhom[k] * orc_func_Te_tab(T->eloss_m, ne, t, dTe);
            if (f->delossdeps)
                f->delossdeps[k] = 2.0 / (3.0 * ORC_BOLTZ) * rhon[k] * orc_func_dTe_tab(T->eloss_n, ne, t, dTe) +
                                   2.0 / (3.0 * ORC_BOLTZ) * rhom[k] * orc_func_dTe_tab(T->eloss_m, ne, t, dTe);
            const double mu_para = ORC_E0 / (masse * nu_m + 1e-100);
            const double bx = Bx[k], br = Br[k];
            const double Bmag = sqrt(bx * bx + br * br);
            const double Om = ORC_E0 * Bmag / (masse * nu_m + 1e-100);
            const double Ox = ORC_E0 * bx / (masse * nu_m + 1e-100);
            const double Or = ORC_E0 * br / (masse * nu_m + 1e-100);
            const double deno = (orc_planted() == ORC_PLANT_TR_DENO ? 2.0 : 1.0) + Om * Om;
            if (f->mu_para) f->mu_para[k] = mu_para;
            if (f->mu_xx) f->mu_xx[k] = mu_para * (1.0 + Ox * Ox) / deno;
            if (f->mu_rr) f->mu_rr[k] = mu_para * (1.0 + Or * Or) / deno;
            if (f->mu_xr) f->mu_xr[k] = mu_para * (Ox * Or) / deno;
            if (f->mu_perp) f->mu_perp[k] = mu_para / deno;
            if (f->mu_min) f->mu_min[k] = mu_para * 2.0 / deno;
        }
    free(r);
    free(zero);
    free(wL);
    free(wR);
    free(wS);
    free(wN);
    return 0;
}
