/*
 * orc_ion.c -- CPU FP64 restatement of the reference's ion block
 * (src/main.cpp:194-213): FluidModule::solve_Uix_tmp, solve_Uir_tmp,
 * solve_Uip_tmp (include/fluidModule.hpp:123-1398), solve_rhoi_constTe
 * (:1405-1755) and correct_Ui_constTe (:1763-2072, without its residual
 * diagnostics).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of cathode_ndd_ion_update.
 *
 * This one follows the reference literally: its boundary loops are written
 * against the five plasma blocks' boxes (gc.i_flc_bl / gx.i_flx_bl /
 * gr.i_flr_bl of blocks 0..4, taken from the geometry's cell, x-face and
 * r-face boxes in that order), with the reference's own ranges, including
 * the x1 loop of solve_Uix_tmp that writes the row below the wall row
 * (:520-525).  It therefore needs the five-block geometry of
 * cathode_ndd_reference_geometry; the device derives the same operations
 * from the boundary segments.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IXI(i, j) ((size_t)(i) * (size_t)ld + (size_t)(j))
/* the block boxes by the reference's names: [b][0] i0, [b][1] i1 */
#define IC(b, k) (G.cell_box[4 * (b) + (k)])
#define JC(b, k) (G.cell_box[4 * (b) + 2 + (k)])
#define IX_(b, k) (G.xface_box[4 * (b) + (k)])
#define JX(b, k) (G.xface_box[4 * (b) + 2 + (k)])
#define IR(b, k) (G.rface_box[4 * (b) + (k)])
#define JR(b, k) (G.rface_box[4 * (b) + 2 + (k)])

typedef struct {
    int ni, nj, ld;
    size_t n;
    double dx, dr;
    double *aP, *aE, *aW, *aN, *aS, *b, *x;
} sys_t;

static void sys_clear(sys_t* S)
{
    double* a[] = {S->aP, S->aE, S->aW, S->aN, S->aS, S->b, S->x};
    for (int k = 0; k < 7; k++) memset(a[k], 0, S->n * sizeof(double));
}

/* solver_SMG over the given boxes from a zero guess; the solution lands in
 * dst inside the boxes only (solverFuncs.hpp:139-146) */
static void sys_solve(sys_t* S, const int* boxes, int nbox, int maxITR, double tol, int solver, double* dst,
                      cathode_status* st)
{
    memset(S->x, 0, S->n * sizeof(double));
    double rel = 0.0;
    int its = 1;
    if (solver == 0) {
        orc_struct5_direct(S->aP, S->aE, S->aW, S->aN, S->aS, S->b, boxes, nbox, S->ni + 2, S->nj + 2, S->ld, S->x);
        rel = orc_rel_residual(S->aP, S->aE, S->aW, S->aN, S->aS, S->b, boxes, nbox, S->ni + 2, S->nj + 2, S->ld,
                               S->x);
    } else {
        its = orc_mg_solve(S->aP, S->aE, S->aW, S->aN, S->aS, S->b, boxes, nbox, S->ni + 2, S->nj + 2, S->ld, maxITR,
                           tol, S->x, &rel);
    }
    const int ld = S->ld;
    for (int bl = 0; bl < nbox; bl++) {
        const int* B = boxes + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) dst[IXI(i, j)] = S->x[IXI(i, j)];
    }
    if (st) {
        st->iterations = its;
        st->rel_residual = rel;
        st->converged = solver == 0 || rel <= tol;
    }
}

/* Ui at a wall by energy conservation from the adjacent unknown
 * (fluidModule.hpp:231-233): sign sqrt(U_old^2 + max(2 e dphi / m_i, 0)) */
static double ui_wall(double sign, double U_old, double dphi, double massi)
{
    const double v = sqrt(pow(U_old, 2) + fmax(2.0 * ORC_E0 * dphi / massi, 0.0));
    return sign < 0.0 ? -v : v;
}

/* kinetic wall flux per density of a drifting Maxwellian (:1502-1526) */
static double ion_gpn(double Ti, double massi, double U, double sign)
{
    const double vth = sqrt(8.0 * ORC_BOLTZ * Ti / (M_PI * massi));
    const double um = sqrt(2.0 * ORC_BOLTZ * Ti / massi);
    const double u = U / (um + 1e-100);
    if (sign < 0.0) return -0.25 * vth * (exp(-u * u) + u * sqrt(M_PI) * (erf(u) - 1.0));
    return 0.25 * vth * (exp(-u * u) + u * sqrt(M_PI) * (erf(u) + 1.0));
}

int orc_ion_update(const cathode_ndd_params* grid, const cathode_ion_params* p, cathode_ion_fields* f,
                   cathode_ion_report* rep, int solver)
{
    const int ni = grid->ni, nj = grid->nj, ld = nj + 2;
    const size_t n = (size_t)(ni + 2) * (size_t)ld;
    const double dx = grid->dx, dr = grid->dr;
    cathode_ndd_geometry G;
    orc_ndd_geometry(ni, nj, &G);
    if (G.n_bl != 5) return -1;
    const int nbl = 5;
    double* r = (double*)calloc(nj + 2, sizeof(double));
    for (int j = 0; j < nj + 2; j++) r[j] = grid->rmin + (j - 0.5) * dr; /* initialModule.hpp:86-88 */
    sys_t S = {ni, nj, ld, n, dx, dr};
    double** bufs[] = {&S.aP, &S.aE, &S.aW, &S.aN, &S.aS, &S.b, &S.x};
    for (int k = 0; k < 7; k++) *bufs[k] = (double*)calloc(n, sizeof(double));
    double* zero = (double*)calloc(n, sizeof(double));
    double* Uix_old = (double*)malloc(n * sizeof(double));
    double* Uir_old = (double*)malloc(n * sizeof(double));
    double* Uip_old = (double*)malloc(n * sizeof(double));
    double* rhoi_old = (double*)malloc(n * sizeof(double));
    double *Uix = f->Uix, *Uir = f->Uir, *Uip = f->Uip, *rhoi = f->rhoi;
    const double* phi = f->phi;
    const double* Ex = f->Ex ? f->Ex : zero;
    const double* Er = f->Er ? f->Er : zero;
    const double* S_ion = f->rate_ionize ? f->rate_ionize : zero;
    const double* nui_m = f->nui_m ? f->nui_m : zero;
    const double* Bx = f->Bx ? f->Bx : zero;
    const double* Br = f->Br ? f->Br : zero;
    const double dt = p->dt_i;
    const double massi = p->massi, Ti = p->Ti;
    if (rep) memset(rep, 0, sizeof(*rep));

    /* ---------------------------------------------------------- solve_Uix_tmp */
    memcpy(Uix_old, Uix, n * sizeof(double)); /* :139-146 */
    memcpy(Uir_old, Uir, n * sizeof(double));
    memcpy(Uip_old, Uip, n * sizeof(double));
    memcpy(rhoi_old, rhoi, n * sizeof(double));
    sys_clear(&S);
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IX_(bl, 0); i <= IX_(bl, 1); i++)
            for (int j = JX(bl, 0); j <= JX(bl, 1); j++) {
                const size_t k = IXI(i, j);
                const double rL = (r[j] + r[j - 1]) / 2.0;
                const double rR = (r[j] + r[j + 1]) / 2.0;
                const double qL = rL / r[j];
                const double qR = rR / r[j];
                const double Uix_Lx = (Uix_old[k] + Uix_old[IXI(i - 1, j)]) / 2.0;
                const double Uix_Rx = (Uix_old[k] + Uix_old[IXI(i + 1, j)]) / 2.0;
                const double Uir_Lr = (Uir_old[k] + Uir_old[IXI(i - 1, j)]) / 2.0;
                const double Uir_Rr = (Uir_old[IXI(i, j + 1)] + Uir_old[IXI(i - 1, j + 1)]) / 2.0;
                const double rhoi_tmp = (rhoi[k] + rhoi[IXI(i - 1, j)]) / 2.0;
                const double S_tmp = (S_ion[k] + S_ion[IXI(i - 1, j)]) / 2.0;
                const double nu_tmp = (nui_m[k] + nui_m[IXI(i - 1, j)]) / 2.0;
                const double dndx = (rhoi[k] - rhoi[IXI(i - 1, j)]) / dx;
                const double Uip_tmp = (Uip_old[k] + Uip_old[IXI(i - 1, j)]) / 2.0;
                const double Br_tmp = (Br[k] + Br[IXI(i - 1, j)]) / 2.0;
                const double lorentz = -ORC_E0 * Uip_tmp * Br_tmp;
                S.aE[k] = -(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx);
                S.aW[k] = (Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx);
                S.aN[k] = -qR * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr);
                S.aS[k] = qL * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr);
                S.aP[k] = 1.0 / dt + S_tmp / (rhoi_tmp + 1e-100) + nu_tmp + S.aE[k] + S.aW[k] + S.aN[k] + S.aS[k];
                S.b[k] = Uix_old[k] / dt + (ORC_E0 * Ex[k] + lorentz) / massi -
                         ORC_BOLTZ * Ti / (massi * rhoi_tmp + 1e-100) * dndx;
            }
    /* normal walls (:221-313): z0, z1, z2, z4 fold aW, z3, z5 fold aE */
    {
        struct { int i, j0, j1; double sign; int bias; } zl[6] = {
            {IX_(0, 0), JX(0, 0), JX(0, 1), -1.0, 0}, {IX_(2, 0), JX(2, 0), JX(0, 0) - 1, -1.0, 0},
            {IX_(3, 0), JX(3, 0), JX(2, 0) - 1, -1.0, 0}, {IX_(4, 0), JX(1, 0), JX(4, 1), -1.0, 0},
            {IX_(1, 1), JX(1, 0), JX(1, 1), 1.0, 0}, {IX_(4, 1), JX(4, 0), JX(4, 1), 1.0, 1}};
        for (int q = 0; q < 6; q++)
            for (int j = zl[q].j0; j <= zl[q].j1; j++) {
                const int i = zl[q].i;
                const size_t k = IXI(i, j);
                double dphi = (phi[k] + phi[IXI(i - 1, j)]) / 2.0;
                if (zl[q].bias) dphi = (phi[k] + phi[IXI(i - 1, j)]) / 2.0 - p->V_bias;
                const double Uw = ui_wall(zl[q].sign, Uix_old[k], dphi, massi);
                if (zl[q].sign < 0.0) {
                    S.b[k] = S.b[k] + S.aW[k] * Uw;
                    S.aW[k] = 0.0;
                } else {
                    S.b[k] = S.b[k] + S.aE[k] * Uw;
                    S.aE[k] = 0.0;
                }
            }
    }
    /* tangential walls and the open face (:330-367) */
    for (int i = IX_(0, 0); i <= IX_(0, 1); i++) S.aS[IXI(i, JX(0, 0))] = 0.0;  /* x0 */
    for (int i = IX_(2, 0); i <= IX_(2, 1); i++) S.aS[IXI(i, JX(2, 0))] = 0.0;  /* x2 */
    for (int i = IX_(0, 0); i <= IX_(1, 1); i++) S.aN[IXI(i, JX(0, 1))] = 0.0;  /* x1 */
    for (int i = IX_(1, 1) + 1; i <= IX_(3, 1); i++) S.aN[IXI(i, JX(3, 1))] = 0.0;  /* x4 */
    for (int i = IX_(4, 0); i <= IX_(4, 1); i++) S.aN[IXI(i, JX(4, 1))] = 0.0;  /* x5 */
    sys_solve(&S, G.xface_box, nbl, p->maxITR_Ui, p->tol_Ui, solver, Uix, rep ? &rep->solve_Uix : NULL);
    /* ghosts (:381-491): the wall faces take Ui_wall, the tangential rows 0 */
    {
        struct { int i, j0, j1; double sign; int bias; } zl[6] = {
            {IX_(0, 0), JX(0, 0), JX(0, 1), -1.0, 0}, {IX_(2, 0), JX(2, 0), JX(0, 0) - 1, -1.0, 0},
            {IX_(3, 0), JX(3, 0), JX(2, 0) - 1, -1.0, 0}, {IX_(4, 0), JX(1, 0), JX(4, 1), -1.0, 0},
            {IX_(1, 1), JX(1, 0), JX(1, 1), 1.0, 0}, {IX_(4, 1), JX(4, 0), JX(4, 1), 1.0, 1}};
        for (int q = 0; q < 6; q++)
            for (int j = zl[q].j0; j <= zl[q].j1; j++) {
                const int i = zl[q].i;
                const size_t k = IXI(i, j);
                double dphi = (phi[k] + phi[IXI(i - 1, j)]) / 2.0;
                if (zl[q].bias) dphi = (phi[k] + phi[IXI(i - 1, j)]) / 2.0 - p->V_bias;
                const double Uw = ui_wall(zl[q].sign, Uix_old[k], dphi, massi);
                Uix[IXI(zl[q].sign < 0.0 ? i - 1 : i + 1, j)] = Uw;
            }
    }
    for (int i = IX_(0, 0); i <= IX_(0, 1) - 1; i++) Uix[IXI(i, JX(0, 0) - 1)] = 0.0;      /* x0 */
    for (int i = IX_(2, 0); i <= IX_(2, 1) - 1; i++) Uix[IXI(i, JX(2, 0) - 1)] = 0.0;      /* x2 */
    for (int i = IX_(0, 0); i <= IX_(1, 1); i++) Uix[IXI(i, JX(0, 1) - 1)] = 0.0;          /* x1, the row below */
    for (int i = IX_(1, 1) + 2; i <= IX_(3, 1) - 1; i++) Uix[IXI(i, JX(3, 1) + 1)] = 0.0;  /* x4 */
    for (int i = IX_(4, 0); i <= IX_(4, 1); i++) Uix[IXI(i, JX(4, 1) + 1)] = 0.0;          /* x5 */

    /* ---------------------------------------------------------- solve_Uir_tmp */
    sys_clear(&S);
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IR(bl, 0); i <= IR(bl, 1); i++)
            for (int j = JR(bl, 0); j <= JR(bl, 1); j++) {
                const size_t k = IXI(i, j);
                const double r_tmp = (r[j] + r[j - 1]) / 2.0;
                const double rL = (r[j - 1] + r[j - 2]) / 2.0;
                const double rR = (r[j + 1] + r[j]) / 2.0;
                const double Uix_Lx = (r[j] * Uix_old[k] + r[j - 1] * Uix_old[IXI(i, j - 1)]) / (2.0 * r_tmp);
                const double Uix_Rx =
                    (r[j] * Uix_old[IXI(i + 1, j)] + r[j - 1] * Uix_old[IXI(i + 1, j - 1)]) / (2.0 * r_tmp);
                const double Uir_Lr = (r_tmp * Uir_old[k] + rL * Uir_old[IXI(i, j - 1)]) / (2.0 * r[j - 1]);
                const double Uir_Rr = (r_tmp * Uir_old[k] + rR * Uir_old[IXI(i, j + 1)]) / (2.0 * r[j]);
                const double rhoi_tmp = (rhoi[k] + rhoi[IXI(i, j - 1)]) / 2.0;
                const double S_tmp = (S_ion[k] + S_ion[IXI(i, j - 1)]) / 2.0;
                const double nu_tmp = (nui_m[k] + nui_m[IXI(i, j - 1)]) / 2.0;
                const double dndr = (rhoi[k] - rhoi[IXI(i, j - 1)]) / dr;
                const double Bx_tmp = (r[j] * Bx[k] + r[j - 1] * Bx[IXI(i, j - 1)]) / (2.0 * r_tmp);
                const double Uip_tmp = (r[j] * Uip_old[k] + r[j - 1] * Uip_old[IXI(i, j - 1)]) / (2.0 * r_tmp);
                const double lorentz = ORC_E0 * Uip_tmp * Bx_tmp;
                const double centrifugal = -Uip_tmp * Uip_tmp / r_tmp;
                S.aE[k] = -(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx);
                S.aW[k] = (Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx);
                S.aN[k] = -r[j] / r_tmp * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr);
                S.aS[k] = r[j - 1] / r_tmp * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr);
                S.aP[k] = 1.0 / dt + S_tmp / (rhoi_tmp + 1e-100) + nu_tmp + S.aE[k] + S.aW[k] + S.aN[k] + S.aS[k];
                S.b[k] = Uir_old[k] / dt + (ORC_E0 * Er[k] + lorentz) / massi -
                         ORC_BOLTZ * Ti / (massi * rhoi_tmp + 1e-100) * dndr + centrifugal;
            }
    /* tangential z walls (:676-730) */
    for (int j = JR(0, 0); j <= JR(0, 1); j++) S.aW[IXI(IR(0, 0), j)] = 0.0;          /* z0 */
    for (int j = JR(2, 0); j <= JR(0, 0) - 1; j++) S.aW[IXI(IR(2, 0), j)] = 0.0;      /* z1 */
    for (int j = JR(3, 0); j <= JR(2, 0) - 1; j++) S.aW[IXI(IR(3, 0), j)] = 0.0;      /* z2 */
    for (int j = JR(1, 0); j <= JR(4, 1); j++) S.aW[IXI(IR(4, 0), j)] = 0.0;          /* z4 */
    for (int j = JR(1, 0); j <= JR(1, 1); j++) S.aE[IXI(IR(1, 1), j)] = 0.0;          /* z3 */
    for (int j = JR(4, 0); j <= JR(4, 1); j++) S.aE[IXI(IR(4, 1), j)] = 0.0;          /* z5 */
    /* normal r walls, the axis and the open face (:734-808) */
    {
        struct { int i0, i1, j; double sign; } xl[2] = {{IR(0, 0), IR(0, 1), JR(0, 0), -1.0},
                                                        {IR(2, 0), IR(2, 1), JR(2, 0), -1.0}};
        for (int q = 0; q < 2; q++)
            for (int i = xl[q].i0; i <= xl[q].i1; i++) {
                const int j = xl[q].j;
                const size_t k = IXI(i, j);
                const double dphi = (phi[k] + phi[IXI(i, j - 1)]) / 2.0;
                const double Uw = ui_wall(-1.0, Uir_old[k], dphi, massi);
                S.b[k] = S.b[k] + S.aS[k] * Uw;
                S.aS[k] = 0.0;
            }
    }
    for (int i = IR(3, 0); i <= IR(4, 1); i++) S.aS[IXI(i, JR(3, 0))] = 0.0; /* x6 */
    {
        struct { int i0, i1, j; } xu[2] = {{IR(0, 0), IR(1, 1), JR(0, 1)}, {IR(1, 1) + 1, IR(3, 1), JR(3, 1)}};
        for (int q = 0; q < 2; q++)
            for (int i = xu[q].i0; i <= xu[q].i1; i++) {
                const int j = xu[q].j;
                const size_t k = IXI(i, j);
                const double dphi = (phi[k] + phi[IXI(i, j - 1)]) / 2.0;
                const double Uw = ui_wall(1.0, Uir_old[k], dphi, massi);
                S.b[k] = S.b[k] + S.aN[k] * Uw;
                S.aN[k] = 0.0;
            }
    }
    for (int i = IR(4, 0); i <= IR(4, 1); i++) { /* x5, backflow blocked */
        const int j = JR(4, 1);
        const size_t k = IXI(i, j);
        const double r_tmp = (r[j] + r[j - 1]) / 2.0;
        const double rR = (r[j + 1] + r[j]) / 2.0;
        S.b[k] = S.b[k] + S.aN[k] * r_tmp / rR * fmax(Uir_old[k], 0);
        S.aN[k] = 0.0;
    }
    sys_solve(&S, G.rface_box, nbl, p->maxITR_Ui, p->tol_Ui, solver, Uir, rep ? &rep->solve_Uir : NULL);
    /* ghosts (:827-941) */
    for (int j = JR(0, 0); j <= JR(0, 1); j++) Uir[IXI(IR(0, 0) - 1, j)] = 0.0;       /* z0 */
    for (int j = JR(2, 0); j <= JR(0, 0) - 2; j++) Uir[IXI(IR(2, 0) - 1, j)] = 0.0;   /* z1 */
    for (int j = JR(3, 0); j <= JR(2, 0) - 2; j++) Uir[IXI(IR(3, 0) - 1, j)] = 0.0;   /* z2 */
    for (int j = JR(1, 0) + 1; j <= JR(4, 1); j++) Uir[IXI(IR(4, 0) - 1, j)] = 0.0;   /* z4 */
    for (int j = JR(1, 0) + 1; j <= JR(1, 1); j++) Uir[IXI(IR(1, 1) + 1, j)] = 0.0;   /* z3 */
    for (int j = JR(4, 0); j <= JR(4, 1); j++) Uir[IXI(IR(4, 1) + 1, j)] = 0.0;       /* z5 */
    {
        struct { int i0, i1, j; } xl[2] = {{IR(0, 0), IR(0, 1), JR(0, 0)}, {IR(2, 0), IR(2, 1), JR(2, 0)}};
        for (int q = 0; q < 2; q++)
            for (int i = xl[q].i0; i <= xl[q].i1; i++) {
                const int j = xl[q].j;
                const double dphi = (phi[IXI(i, j)] + phi[IXI(i, j - 1)]) / 2.0;
                Uir[IXI(i, j - 1)] = ui_wall(-1.0, Uir_old[IXI(i, j)], dphi, massi);
            }
    }
    for (int i = IR(3, 0); i <= IR(4, 1); i++) Uir[IXI(i, JR(3, 0) - 1)] = 0.0; /* x6 */
    {
        struct { int i0, i1, j; } xu[2] = {{IR(0, 0), IR(1, 1), JR(0, 1)}, {IR(1, 1) + 1, IR(3, 1), JR(3, 1)}};
        for (int q = 0; q < 2; q++)
            for (int i = xu[q].i0; i <= xu[q].i1; i++) {
                const int j = xu[q].j;
                const double dphi = (phi[IXI(i, j)] + phi[IXI(i, j - 1)]) / 2.0;
                Uir[IXI(i, j + 1)] = ui_wall(1.0, Uir_old[IXI(i, j)], dphi, massi);
            }
    }
    for (int i = IR(4, 0); i <= IR(4, 1); i++) {
        const int j = JR(4, 1);
        const double r_tmp = (r[j] + r[j - 1]) / 2.0;
        const double rR = (r[j + 1] + r[j]) / 2.0;
        Uir[IXI(i, j + 1)] = r_tmp / rR * fmax(Uir_old[IXI(i, j)], 0);
    }

    /* ---------------------------------------------------------- solve_Uip_tmp */
    sys_clear(&S);
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IC(bl, 0); i <= IC(bl, 1); i++)
            for (int j = JC(bl, 0); j <= JC(bl, 1); j++) {
                const size_t k = IXI(i, j);
                const double rL = (r[j] + r[j - 1]) / 2.0;
                const double rR = (r[j] + r[j + 1]) / 2.0;
                const double qL = rL / r[j];
                const double qR = rR / r[j];
                const double Uix_Lx = Uix_old[k];
                const double Uix_Rx = Uix_old[IXI(i + 1, j)];
                const double Uir_Lr = Uir_old[k];
                const double Uir_Rr = Uir_old[IXI(i, j + 1)];
                const double Uix_tmp = (Uix_old[IXI(i + 1, j)] + Uix_old[k]) / 2.0;
                const double Uir_tmp = (rL * Uir_old[k] + rR * Uir_old[IXI(i, j + 1)]) / (2.0 * r[j]);
                double lorentz = ORC_E0 * (Uix_tmp * Br[k] - Uir_tmp * Bx[k]);
                if (orc_planted() == ORC_PLANT_ION_LORENTZ_SIGN) lorentz = -lorentz;
                const double centrifugal = Uip[k] * Uir_tmp / r[j];
                S.aE[k] = -(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx);
                S.aW[k] = (Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx);
                S.aN[k] = -qR * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr);
                S.aS[k] = qL * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr);
                S.aP[k] = 1.0 / dt + S_ion[k] / (rhoi[k] + 1e-100) + nui_m[k] + S.aE[k] + S.aW[k] + S.aN[k] + S.aS[k];
                S.b[k] = Uip_old[k] / dt + lorentz / massi + centrifugal;
            }
    /* continuity at every boundary (:1117-1205) */
    {
        struct { int i, j0, j1, east; } zl[6] = {
            {IC(0, 0), JC(0, 0), JC(0, 1), 0}, {IC(2, 0), JC(2, 0), JC(0, 0) - 1, 0},
            {IC(3, 0), JC(3, 0), JC(2, 0) - 1, 0}, {IC(4, 0), JC(1, 0), JC(4, 1), 0},
            {IC(1, 1), JC(1, 0), JC(1, 1), 1}, {IC(4, 1), JC(4, 0), JC(4, 1), 1}};
        for (int q = 0; q < 6; q++)
            for (int j = zl[q].j0; j <= zl[q].j1; j++) {
                const size_t k = IXI(zl[q].i, j);
                if (zl[q].east) {
                    S.aP[k] = S.aP[k] - S.aE[k];
                    S.aE[k] = 0.0;
                } else {
                    S.aP[k] = S.aP[k] - S.aW[k];
                    S.aW[k] = 0.0;
                }
            }
        struct { int i0, i1, j, up; } xl[5] = {
            {IC(0, 0), IC(0, 1), JC(0, 0), 0}, {IC(2, 0), IC(2, 1), JC(2, 0), 0}, {IC(0, 0), IC(1, 1), JC(0, 1), 1},
            {IC(1, 1) + 1, IC(3, 1), JC(3, 1), 1}, {IC(4, 0), IC(4, 1), JC(4, 1), 1}};
        for (int q = 0; q < 5; q++)
            for (int i = xl[q].i0; i <= xl[q].i1; i++) {
                const int j = xl[q].j;
                const size_t k = IXI(i, j);
                if (xl[q].up) {
                    S.aP[k] = S.aP[k] - S.aN[k] * r[j] / r[j + 1];
                    S.aN[k] = 0.0;
                } else {
                    S.aP[k] = S.aP[k] - S.aS[k] * r[j] / r[j - 1];
                    S.aS[k] = 0.0;
                }
            }
        sys_solve(&S, G.cell_box, nbl, p->maxITR_Ui, p->tol_Ui, solver, Uip, rep ? &rep->solve_Uip : NULL);
        /* ghosts (:1223-1307) */
        for (int q = 0; q < 6; q++)
            for (int j = zl[q].j0; j <= zl[q].j1; j++) {
                const int i = zl[q].i;
                Uip[IXI(zl[q].east ? i + 1 : i - 1, j)] = Uip[IXI(i, j)];
            }
        for (int q = 0; q < 5; q++)
            for (int i = xl[q].i0; i <= xl[q].i1; i++) {
                const int j = xl[q].j;
                if (xl[q].up) Uip[IXI(i, j + 1)] = r[j] / r[j + 1] * Uip[IXI(i, j)];
                else Uip[IXI(i, j - 1)] = r[j] / r[j - 1] * Uip[IXI(i, j)];
            }
    }

    /* ---------------------------------------------------------- solve_rhoi_constTe */
    sys_clear(&S);
    const double Te_rep = p->Te_rep_eV * ORC_E0 / ORC_BOLTZ;
    const double nu_star = 1.0 / p->rmax * sqrt(ORC_BOLTZ * Te_rep / massi);
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IC(bl, 0); i <= IC(bl, 1); i++)
            for (int j = JC(bl, 0); j <= JC(bl, 1); j++) {
                const size_t k = IXI(i, j);
                const double rL = (r[j] + r[j - 1]) / 2.0;
                const double rR = (r[j] + r[j + 1]) / 2.0;
                const double qL = rL / r[j];
                const double qR = rR / r[j];
                const double Uix_Lx = Uix[k], Uix_Rx = Uix[IXI(i + 1, j)];
                const double Uir_Lr = Uir[k], Uir_Rr = Uir[IXI(i, j + 1)];
                const double nLx = (nui_m[k] + nui_m[IXI(i - 1, j)]) / 2.0;
                const double nRx = (nui_m[k] + nui_m[IXI(i + 1, j)]) / 2.0;
                const double nLr = (nui_m[k] + nui_m[IXI(i, j - 1)]) / 2.0;
                const double nRr = (nui_m[k] + nui_m[IXI(i, j + 1)]) / 2.0;
                const double D_Lx = ORC_BOLTZ * (Ti + Te_rep) / (massi * (nLx + nu_star) + 1e-100);
                const double D_Rx = ORC_BOLTZ * (Ti + Te_rep) / (massi * (nRx + nu_star) + 1e-100);
                const double D_Lr = ORC_BOLTZ * (Ti + Te_rep) / (massi * (nLr + nu_star) + 1e-100);
                const double D_Rr = ORC_BOLTZ * (Ti + Te_rep) / (massi * (nRr + nu_star) + 1e-100);
                /* the reference's adjacency flags (:1447-1473) */
                double bLx_wall = (double)(i != IC(0, 0) || bl != 0);
                bLx_wall = bLx_wall * (double)(i != IC(2, 0) || j > JC(0, 0) - 1);
                bLx_wall = bLx_wall * (double)(i != IC(3, 0) || j > JC(2, 0) - 1);
                bLx_wall = bLx_wall * (double)(i != IC(4, 0) || j < JC(1, 0));
                const double bLx = bLx_wall;
                double bRx_wall = (double)(i != IC(1, 1) || j < JC(1, 0));
                bRx_wall = bRx_wall * (double)(i != IC(4, 1));
                const double bRx = bRx_wall;
                double bLr_wall = (double)(j != JC(0, 0) || i > IC(0, 1));
                bLr_wall = bLr_wall * (double)(j != JC(2, 0) || i > IC(2, 1));
                const double bLr_cen = (double)(j != JC(3, 0));
                const double bLr = bLr_wall * bLr_cen;
                double bRr_wall = (double)(j != JC(0, 1) || bl > 2);
                bRr_wall = bRr_wall * (double)(j != JC(3, 1) || i < IC(1, 1) + 1 || i > IC(3, 1));
                const double bRr = bRr_wall;
                double gLx = 0.0, gRx = 0.0, gLr = 0.0, gRr = 0.0;
                if (bLx_wall == 0) gLx = ion_gpn(Ti, massi, Uix[k], -1.0);
                if (bRx_wall == 0) gRx = ion_gpn(Ti, massi, Uix[IXI(i + 1, j)], 1.0);
                if (bLr_wall == 0) gLr = ion_gpn(Ti, massi, Uir[k], -1.0);
                if (bRr_wall == 0) gRr = ion_gpn(Ti, massi, Uir[IXI(i, j + 1)], 1.0);
                S.aP[k] = (Uix_Rx + fabs(Uix_Rx)) / (2.0 * dx) * bRx - (Uix_Lx - fabs(Uix_Lx)) / (2.0 * dx) * bLx +
                          qR * (Uir_Rr + fabs(Uir_Rr)) / (2.0 * dr) * bRr - qL * (Uir_Lr - fabs(Uir_Lr)) / (2.0 * dr) * bLr +
                          ((D_Rx * bRx + D_Lx * bLx) / (dx * dx) + (qR * D_Rr * bRr + qL * D_Lr * bLr) / (dr * dr)) -
                          gLx / dx + gRx / dx - qL * gLr / dr + qR * gRr / dr + 1.0 / dt;
                S.aE[k] = (-(Uix_Rx - fabs(Uix_Rx)) / (2.0 * dx) + D_Rx / (dx * dx)) * bRx;
                S.aW[k] = ((Uix_Lx + fabs(Uix_Lx)) / (2.0 * dx) + D_Lx / (dx * dx)) * bLx;
                S.aN[k] = (-qR * (Uir_Rr - fabs(Uir_Rr)) / (2.0 * dr) + qR * D_Rr / (dr * dr)) * bRr;
                S.aS[k] = (qL * (Uir_Lr + fabs(Uir_Lr)) / (2.0 * dr) + qL * D_Lr / (dr * dr)) * bLr;
                const double divD =
                    (D_Rx * rhoi_old[IXI(i + 1, j)] * bRx - (D_Rx * bRx + D_Lx * bLx) * rhoi_old[k] +
                     D_Lx * rhoi_old[IXI(i - 1, j)] * bLx) / (dx * dx) +
                    (qR * D_Rr * rhoi_old[IXI(i, j + 1)] * bRr - (qR * D_Rr * bRr + qL * D_Lr * bLr) * rhoi_old[k] +
                     qL * D_Lr * rhoi_old[IXI(i, j - 1)] * bLr) / (dr * dr);
                S.b[k] = rhoi_old[k] / dt + S_ion[k] - divD;
            }
    for (int i = IC(4, 0); i <= IC(4, 1); i++) { /* x5 zero gradient (:1571-1579) */
        const size_t k = IXI(i, JC(4, 1));
        S.aP[k] = S.aP[k] - S.aN[k];
        S.aN[k] = 0.0;
    }
    sys_solve(&S, G.cell_box, nbl, p->maxITR_rhoi, p->tol_rhoi, solver, rhoi, rep ? &rep->solve_rhoi : NULL);
    /* wall fluxes from the unreset density (:1590-1715) */
    {
        struct { int i, j0, j1; double sign; } zl[6] = {
            {IC(0, 0), JC(0, 0), JC(0, 1), -1.0}, {IC(2, 0), JC(2, 0), JC(0, 0) - 1, -1.0},
            {IC(3, 0), JC(3, 0), JC(2, 0) - 1, -1.0}, {IC(4, 0), JC(1, 0), JC(4, 1), -1.0},
            {IC(1, 1), JC(1, 0), JC(1, 1), 1.0}, {IC(4, 1), JC(4, 0), JC(4, 1), 1.0}};
        for (int q = 0; q < 6; q++)
            for (int j = zl[q].j0; j <= zl[q].j1; j++) {
                const int i = zl[q].i;
                const int fi = zl[q].sign < 0.0 ? i : i + 1;
                f->rhoUix_wall[IXI(fi, j)] = rhoi[IXI(i, j)] * ion_gpn(Ti, massi, Uix[IXI(fi, j)], zl[q].sign);
            }
        struct { int i0, i1, j; double sign; } xl[4] = {
            {IC(0, 0), IC(0, 1), JC(0, 0), -1.0}, {IC(2, 0), IC(2, 1), JC(2, 0), -1.0},
            {IC(0, 0), IC(1, 1), JC(0, 1), 1.0}, {IC(1, 1) + 1, IC(3, 1), JC(3, 1), 1.0}};
        for (int q = 0; q < 4; q++)
            for (int i = xl[q].i0; i <= xl[q].i1; i++) {
                const int j = xl[q].j;
                const int fj = xl[q].sign < 0.0 ? j : j + 1;
                f->rhoUir_wall[IXI(i, fj)] = rhoi[IXI(i, j)] * ion_gpn(Ti, massi, Uir[IXI(i, fj)], xl[q].sign);
            }
        for (int i = IC(4, 0); i <= IC(4, 1); i++) {
            const int j = JC(4, 1);
            f->rhoUir[IXI(i, j + 1)] = rhoi[IXI(i, j)] * Uir[IXI(i, j + 1)];
            f->rhoUir_wall[IXI(i, j + 1)] = 0.0;
        }
    }
    int nneg = 0; /* :1727-1741 */
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IC(bl, 0); i <= IC(bl, 1); i++)
            for (int j = JC(bl, 0); j <= JC(bl, 1); j++)
                if (rhoi[IXI(i, j)] < 0.0) {
                    nneg++;
                    rhoi[IXI(i, j)] = fmax(rhoi_old[IXI(i, j)], 0.0);
                }
    for (int i = IC(4, 0); i <= IC(4, 1); i++) rhoi[IXI(i, JC(4, 1) + 1)] = rhoi[IXI(i, JC(4, 1))];
    if (rep) rep->neg_rhoi = nneg;

    /* ---------------------------------------------------------- correct_Ui_constTe */
    const double Te_c = orc_planted() == ORC_PLANT_ION_CORR_TE ? 0.0 : Te_rep;
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IX_(bl, 0); i <= IX_(bl, 1); i++)
            for (int j = JX(bl, 0); j <= JX(bl, 1); j++) {
                const size_t k = IXI(i, j), kw = IXI(i - 1, j);
                const double nu_tmp = (nui_m[k] + nui_m[kw]) / 2.0;
                const double D = ORC_BOLTZ * (Ti + Te_c) / (massi * (nu_tmp + nu_star) + 1e-100);
                const double rhoi_tmp = (rhoi[k] + rhoi[kw]) / 2.0;
                const double dn_old = (rhoi_old[k] - rhoi_old[kw]) / dx;
                const double dn = (rhoi[k] - rhoi[kw]) / dx;
                const double flux = rhoi[kw] * (Uix[k] + fabs(Uix[k])) / 2.0 + rhoi[k] * (Uix[k] - fabs(Uix[k])) / 2.0;
                f->rhoUix[k] = flux - D * (dn - dn_old);
                Uix[k] = Uix[k] - D / (rhoi_tmp + 1e-100) * (dn - dn_old);
            }
    for (int bl = 0; bl < nbl; bl++)
        for (int i = IR(bl, 0); i <= IR(bl, 1); i++)
            for (int j = JR(bl, 0); j <= JR(bl, 1) + (int)(bl == 4); j++) {
                const size_t k = IXI(i, j), ks = IXI(i, j - 1);
                const double nu_tmp = (nui_m[k] + nui_m[ks]) / 2.0;
                const double D = ORC_BOLTZ * (Ti + Te_c) / (massi * (nu_tmp + nu_star) + 1e-100);
                const double rhoi_tmp = (rhoi[k] + rhoi[ks]) / 2.0;
                const double dn_old = (rhoi_old[k] - rhoi_old[ks]) / dr;
                const double dn = (rhoi[k] - rhoi[ks]) / dr;
                const double flux = rhoi[ks] * (Uir[k] + fabs(Uir[k])) / 2.0 + rhoi[k] * (Uir[k] - fabs(Uir[k])) / 2.0;
                f->rhoUir[k] = flux - D * (dn - dn_old);
                Uir[k] = Uir[k] - D / (rhoi_tmp + 1e-100) * (dn - dn_old);
            }

    free(r);
    for (int k = 0; k < 7; k++) free(*bufs[k]);
    free(zero);
    free(Uix_old);
    free(Uir_old);
    free(Uip_old);
    free(rhoi_old);
    return 0;
}
