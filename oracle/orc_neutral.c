/*
 * orc_neutral.c -- CPU FP64 restatement of FluidModule::update_rhon
 * (include/fluidModule.hpp:4963-5411) and FluidModule::update_rhom
 * (:5416-5792) over the block geometry of orc_ndd_set_geometry.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of cathode_ndd_neutral_update.
 *
 * The restatement keeps the reference's structure: per-cell boundary flags
 * (the bLx_wall, bRx_wall, bLr_wall, bLr_cen, bRr_wall, bRr_in, bRr_open
 * products of :5024-5046 and :5480-5501), here derived from the boundary
 * segments each cell lies on instead of the reference's hard-coded block
 * indices; the row of each plasma cell block by block; one 5-point solve
 * over the plasma cell boxes from a zero guess; the inner-face flux loops
 * over the face boxes, then the boundary loops segment by segment in list
 * order (the reference's z0, z1, z2, z4, z3, z5, x0, x2, x1, x4, x5 loops).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IXN(i, j) ((size_t)(i) * (size_t)ld + (size_t)(j))

static const double* or_zero(const double* a, const double* z) { return a ? a : z; }

enum { SP_N = 0, SP_M = 1 };

typedef struct {
    /* 1 = no boundary of that kind on the cell's face (the reference's flags) */
    unsigned char bLx_wall, bRx_wall, bLr_wall, bLr_cen, bRr_wall, bRr_in, bRr_open;
    int block; /* plasma block of the cell, -1 outside */
} cellflags_t;

/* the flags of every cell for one species: walls, the axis, the open face;
 * the inlet segment's cells at x <= x_in are inlet cells for the neutrals
 * and walls for the metastables (:5042-5043, :5496) */
static void cell_flags(const cathode_ndd_geometry* G, int ni, int nj, const cathode_neutral_params* p, double dx,
                       int sp, cellflags_t* F)
{
    const int ld = nj + 2;
    const size_t n = (size_t)(ni + 2) * (size_t)ld;
    for (size_t k = 0; k < n; k++) {
        cellflags_t c = {1, 1, 1, 1, 1, 1, 1, -1};
        F[k] = c;
    }
    for (int b = 0; b < G->n_bl; b++) {
        const int* B = G->cell_box + 4 * b;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) F[IXN(i, j)].block = b;
    }
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* q = &G->seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            switch (q->kind) {
            case CATHODE_SEG_X_LOW: F[IXN(q->cell, t)].bLx_wall = 0; break;
            case CATHODE_SEG_X_HIGH: F[IXN(q->cell, t)].bRx_wall = 0; break;
            case CATHODE_SEG_R_LOW: F[IXN(t, q->cell)].bLr_wall = 0; break;
            case CATHODE_SEG_R_AXIS: F[IXN(t, q->cell)].bLr_cen = 0; break;
            case CATHODE_SEG_R_OPEN: F[IXN(t, q->cell)].bRr_open = 0; break;
            default: {
                const double x = p->xL + (t - 0.5) * dx; /* gc.x, initialModule.hpp:84 */
                if (sp == SP_N && (q->flag & CATHODE_SEG_INLET) && x <= p->x_in)
                    F[IXN(t, q->cell)].bRr_in = 0;
                else
                    F[IXN(t, q->cell)].bRr_wall = 0;
                break;
            }
            }
        }
    }
}

int orc_neutral_update(const cathode_ndd_params* grid, const cathode_neutral_params* p, cathode_neutral_fields* f,
                       cathode_neutral_report* rep, int solver)
{
    const int ni = grid->ni, nj = grid->nj, ld = nj + 2;
    const size_t n = (size_t)(ni + 2) * (size_t)ld;
    const double dx = grid->dx, dr = grid->dr;
    cathode_ndd_geometry G;
    orc_ndd_geometry(ni, nj, &G);
    double* r = (double*)calloc(nj + 2, sizeof(double));
    for (int j = 0; j < nj + 2; j++) r[j] = grid->rmin + (j - 0.5) * dr; /* initialModule.hpp:86-88 */
    double* zero = (double*)calloc(n, sizeof(double));
    double *aP = calloc(n, sizeof(double)), *aE = calloc(n, sizeof(double)), *aW = calloc(n, sizeof(double));
    double *aN = calloc(n, sizeof(double)), *aS = calloc(n, sizeof(double)), *b = calloc(n, sizeof(double));
    double *x = calloc(n, sizeof(double)), *n_old = calloc(n, sizeof(double)), *fr_old = calloc(n, sizeof(double));
    cellflags_t* F = (cellflags_t*)calloc(n, sizeof(cellflags_t));
    const double* rhoe = or_zero(f->rhoe, zero);
    const double* excStep = or_zero(f->nu_excStep, zero);
    const double* super = or_zero(f->nu_super, zero);
    const double* excMeta = or_zero(f->nu_excMeta, zero);
    const double* exc = or_zero(f->nu_exc, zero);
    const double* ionz = or_zero(f->nu_ionz, zero);
    const double* ionzStep = or_zero(f->nu_ionzStep, zero);
    const double* Gix = or_zero(f->rhoUix_wall, zero);
    const double* Gir = or_zero(f->rhoUir_wall, zero);
    if (rep) memset(rep, 0, sizeof(*rep));
    const double vth = sqrt(8.0 * ORC_BOLTZ * p->Tn / (M_PI * p->massi)); /* :4977, :5433 */

    for (int sp = SP_N; sp <= SP_M; sp++) {
        if (!(p->which & (1 << sp))) continue;
        double* dens = sp == SP_N ? f->rhon : f->rhom;
        double* fx = sp == SP_N ? f->rhoUnx : f->rhoUmx;
        double* fr = sp == SP_N ? f->rhoUnr : f->rhoUmr;
        double* fxw = sp == SP_N ? f->rhoUnx_wall : f->rhoUmx_wall;
        double* frw = sp == SP_N ? f->rhoUnr_wall : f->rhoUmr_wall;
        const double* Gmx = or_zero(f->rhoUmx_wall, zero);
        const double* Gmr = or_zero(f->rhoUmr_wall, zero);
        const double dt = sp == SP_N ? p->dt_n : p->dt_m;
        double DN;
        if (sp == SP_N) {
            double sigma = M_PI * (2.0 * p->ri) * (2.0 * p->ri); /* :4979 */
            if (orc_planted() == ORC_PLANT_NEUT_SIGMA) sigma = M_PI * p->ri * p->ri;
            DN = 2.0 / (3.0 * sigma) * sqrt(ORC_BOLTZ * p->Tn / (M_PI * p->massi)); /* :4983 */
        } else {
            DN = p->DmN;
        }
        /* neutrals: D from the old neutral density; metastables: from the current one */
        memcpy(n_old, dens, n * sizeof(double));
        memcpy(fr_old, fr, n * sizeof(double));
        const double* nD = sp == SP_N ? n_old : f->rhon;
        cell_flags(&G, ni, nj, p, dx, sp, F);
        memset(aP, 0, n * sizeof(double));
        memset(aE, 0, n * sizeof(double));
        memset(aW, 0, n * sizeof(double));
        memset(aN, 0, n * sizeof(double));
        memset(aS, 0, n * sizeof(double));
        memset(b, 0, n * sizeof(double));
        for (int bl = 0; bl < G.n_bl; bl++) {
            const int* B = G.cell_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) {
                    const size_t k = IXN(i, j);
                    const cellflags_t c = F[k];
                    const double rL = (r[j] + r[j - 1]) / 2.0;
                    const double rR = (r[j] + r[j + 1]) / 2.0;
                    const double qL = (r[j] + r[j - 1]) / 2.0 / r[j];
                    const double qR = (r[j] + r[j + 1]) / 2.0 / r[j];
                    const double D_Lx = DN / ((nD[k] + nD[IXN(i - 1, j)]) / 2.0 + 1e-100);
                    const double D_Rx = DN / ((nD[k] + nD[IXN(i + 1, j)]) / 2.0 + 1e-100);
                    const double D_Lr = DN / ((nD[k] + nD[IXN(i, j - 1)]) / 2.0 + 1e-100);
                    const double D_Rr = DN / ((nD[k] + nD[IXN(i, j + 1)]) / 2.0 + 1e-100);
                    const double bLx = (double)c.bLx_wall, bRx = (double)c.bRx_wall;
                    const double bLr = (double)(c.bLr_wall * c.bLr_cen);
                    const double bRr = (double)(c.bRr_wall * c.bRr_in * c.bRr_open);
                    aE[k] = D_Rx * dr * dr * bRx;
                    aW[k] = D_Lx * dr * dr * bLx;
                    aN[k] = qR * D_Rr * dx * dx * bRr;
                    aS[k] = qL * D_Lr * dx * dx * bLr;
                    if (sp == SP_N) {
                        /* :5069-5091 */
                        double GwLx = 0.0, GwRx = 0.0, GwLr = 0.0, GwRr = 0.0, Gin = 0.0, Gop = 0.0;
                        if (!c.bLx_wall) GwLx = -Gix[k] - Gmx[k];
                        if (!c.bRx_wall) GwRx = -Gix[IXN(i + 1, j)] - Gmx[IXN(i + 1, j)];
                        if (!c.bLr_wall) GwLr = -Gir[k] - Gmr[k];
                        if (!c.bRr_wall) GwRr = -Gir[IXN(i, j + 1)] - Gmr[IXN(i, j + 1)];
                        if (!c.bRr_in) {
                            const double feed = orc_planted() == ORC_PLANT_NEUT_INLET_SIGN ? -p->fn_In : p->fn_In;
                            Gin = -Gir[IXN(i, j + 1)] - Gmr[IXN(i, j + 1)] - feed;
                        }
                        if (!c.bRr_open) Gop = fmax(rL / rR * fr_old[k], 0.0);
                        aP[k] = (D_Rx * bRx + D_Lx * bLx) * dr * dr + (qR * D_Rr * bRr + qL * D_Lr * bLr) * dx * dx +
                                dx * dx * dr * dr / dt;
                        /* :5104-5110 */
                        const double Gs = (excStep[k] + super[k]) * rhoe[k];
                        const double Ls = (excMeta[k] + 0.5 * exc[k] + ionz[k]) * rhoe[k];
                        b[k] = n_old[k] / dt * dx * dx * dr * dr + (Gs - Ls) * dx * dx * dr * dr +
                               (GwLx) * dr * dx * dr - (GwRx) * dr * dx * dr + qL * (GwLr) * dx * dx * dr -
                               qR * (GwRr + Gop + Gin) * dx * dx * dr;
                    } else {
                        /* :5508-5545: wall flux per density, implicit */
                        const double mean = orc_planted() == ORC_PLANT_META_VTH
                                                ? sqrt(2.0 * ORC_BOLTZ * p->Tn / p->massi)
                                                : vth;
                        const double gLx = !c.bLx_wall ? -0.25 * mean : 0.0;
                        const double gRx = !c.bRx_wall ? 0.25 * mean : 0.0;
                        const double gLr = !c.bLr_wall ? -0.25 * mean : 0.0;
                        const double gRr = !c.bRr_wall ? 0.25 * mean : 0.0;
                        const double Gop = !c.bRr_open ? rL / rR * fr_old[k] : 0.0;
                        const double qRw = orc_planted() == ORC_PLANT_META_RWALL_Q ? 1.0 : qR;
                        aP[k] = (D_Rx * bRx + D_Lx * bLx) * dr * dr + (qR * D_Rr * bRr + qL * D_Lr * bLr) * dx * dx -
                                gLx * dr * dx * dr + gRx * dr * dx * dr - qL * gLr * dx * dx * dr +
                                qRw * gRr * dx * dx * dr + dx * dx * dr * dr / dt;
                        /* :5555-5572; the explicit wall terms there are 0.0 */
                        const double Gs = rhoe[k] * (excMeta[k] + 0.5 * exc[k]);
                        const double Ls = rhoe[k] * (ionzStep[k] + excStep[k] + super[k]);
                        b[k] = n_old[k] / dt * dx * dx * dr * dr + (Gs - Ls) * dx * dx * dr * dr -
                               qR * (Gop) * dx * dx * dr;
                    }
                }
        }
        /* solver_SMG over the plasma cell boxes, zero initial guess (:5114-5118, :5577-5581) */
        memset(x, 0, n * sizeof(double));
        double rel = 0.0;
        int its = 1;
        if (solver == 0) {
            orc_struct5_direct(aP, aE, aW, aN, aS, b, G.cell_box, G.n_bl, ni + 2, nj + 2, ld, x);
            rel = orc_rel_residual(aP, aE, aW, aN, aS, b, G.cell_box, G.n_bl, ni + 2, nj + 2, ld, x);
        } else {
            its = orc_mg_solve(aP, aE, aW, aN, aS, b, G.cell_box, G.n_bl, ni + 2, nj + 2, ld, p->maxITR, p->tol, x,
                               &rel);
        }
        for (int bl = 0; bl < G.n_bl; bl++) {
            const int* B = G.cell_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) dens[IXN(i, j)] = x[IXN(i, j)];
        }
        int nneg = 0;
        /* metastables: negative values in the first guard_blocks blocks go back
         * to max(old, 0) before the fluxes (:5582-5595) */
        if (sp == SP_M) {
            for (int bl = 0; bl < G.n_bl && bl < p->guard_blocks; bl++) {
                const int* B = G.cell_box + 4 * bl;
                for (int i = B[0]; i <= B[1]; i++)
                    for (int j = B[2]; j <= B[3]; j++)
                        if (dens[IXN(i, j)] < 0.0) {
                            nneg++;
                            dens[IXN(i, j)] = fmax(n_old[IXN(i, j)], 0.0);
                        }
            }
        }
        /* inner faces (:5121-5144, :5597-5618) */
        for (int bl = 0; bl < G.n_bl; bl++) {
            const int* X = G.xface_box + 4 * bl;
            for (int i = X[0]; i <= X[1]; i++)
                for (int j = X[2]; j <= X[3]; j++) {
                    const double D = DN / ((nD[IXN(i, j)] + nD[IXN(i - 1, j)]) / 2.0 + 1e-100);
                    fx[IXN(i, j)] = -D * (dens[IXN(i, j)] - dens[IXN(i - 1, j)]) / dx;
                }
            const int* R = G.rface_box + 4 * bl;
            for (int i = R[0]; i <= R[1]; i++)
                for (int j = R[2]; j <= R[3]; j++) {
                    const double D = DN / ((nD[IXN(i, j)] + nD[IXN(i, j - 1)]) / 2.0 + 1e-100);
                    fr[IXN(i, j)] = -D * (dens[IXN(i, j)] - dens[IXN(i, j - 1)]) / dr;
                }
        }
        /* boundary loops, segment by segment (:5147-5262, :5620-5693) */
        for (int s = 0; s < G.n_seg; s++) {
            const cathode_ndd_segment* q = &G.seg[s];
            for (int t = q->lo; t <= q->hi; t++) {
                if (q->kind == CATHODE_SEG_X_LOW || q->kind == CATHODE_SEG_X_HIGH) {
                    const int i = q->cell, j = t;
                    const int fi = q->kind == CATHODE_SEG_X_LOW ? i : i + 1;
                    const double sgn = q->kind == CATHODE_SEG_X_LOW ? -1.0 : 1.0;
                    if (sp == SP_N) fx[IXN(fi, j)] = -Gix[IXN(fi, j)] - Gmx[IXN(fi, j)];
                    else fx[IXN(fi, j)] = sgn * 0.25 * dens[IXN(i, j)] * vth;
                    fxw[IXN(fi, j)] = fx[IXN(fi, j)];
                } else if (q->kind == CATHODE_SEG_R_LOW) {
                    const int i = t, j = q->cell;
                    if (sp == SP_N) fr[IXN(i, j)] = -Gir[IXN(i, j)] - Gmr[IXN(i, j)];
                    else fr[IXN(i, j)] = -0.25 * dens[IXN(i, j)] * vth;
                    frw[IXN(i, j)] = fr[IXN(i, j)];
                } else if (q->kind == CATHODE_SEG_R_HIGH) {
                    const int i = t, j = q->cell;
                    const double xi = p->xL + (i - 0.5) * dx;
                    if (sp == SP_N && (q->flag & CATHODE_SEG_INLET) && xi < p->x_in) {
                        /* the inlet part of x1 (:5232-5235): no wall copy */
                        fr[IXN(i, j + 1)] = -Gir[IXN(i, j + 1)] - Gmr[IXN(i, j + 1)] - p->fn_In;
                    } else {
                        if (sp == SP_N) fr[IXN(i, j + 1)] = -Gir[IXN(i, j + 1)] - Gmr[IXN(i, j + 1)];
                        else fr[IXN(i, j + 1)] = 0.25 * dens[IXN(i, j)] * vth;
                        frw[IXN(i, j + 1)] = fr[IXN(i, j + 1)];
                    }
                } else if (q->kind == CATHODE_SEG_R_OPEN) {
                    const int i = t, j = q->cell;
                    const double rL = (r[j] + r[j - 1]) / 2.0;
                    const double rR = (r[j] + r[j + 1]) / 2.0;
                    if (sp == SP_N) fr[IXN(i, j + 1)] = fmax(rL / rR * fr_old[IXN(i, j)], 0.0);
                    else fr[IXN(i, j + 1)] = fr_old[IXN(i, j)] * rL / rR;
                    frw[IXN(i, j + 1)] = 0.0;
                }
            }
        }
        /* neutrals: every negative value goes back to max(old, 0) after the
         * fluxes (:5383-5399) */
        if (sp == SP_N) {
            for (int bl = 0; bl < G.n_bl; bl++) {
                const int* B = G.cell_box + 4 * bl;
                for (int i = B[0]; i <= B[1]; i++)
                    for (int j = B[2]; j <= B[3]; j++)
                        if (dens[IXN(i, j)] < 0.0) {
                            nneg++;
                            dens[IXN(i, j)] = fmax(n_old[IXN(i, j)], 0.0);
                        }
            }
        }
        const double err = orc_calc_res_boxes(dens, n_old, ld, G.cell_box, G.n_bl, 1e-100, p->flag_error);
        if (rep) {
            cathode_status* st = sp == SP_N ? &rep->solve_n : &rep->solve_m;
            st->iterations = its;
            st->converged = solver == 0 || rel <= p->tol;
            st->rel_residual = rel;
            if (sp == SP_N) {
                rep->error_rhon = err;
                rep->neg_rhon = nneg;
            } else {
                rep->error_rhom = err;
                rep->neg_rhom = nneg;
            }
        }
    }
    free(r);
    free(zero);
    free(aP);
    free(aE);
    free(aW);
    free(aN);
    free(aS);
    free(b);
    free(x);
    free(n_old);
    free(fr_old);
    free(F);
    return 0;
}
