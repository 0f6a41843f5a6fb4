/*
 * orc_ndd.c -- CPU FP64 restatement of one implicit NDD time step.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Domain: one rectangular block of cells i = 1..ni (z), j = 1..nj (r) with the
 * boundary set of the reference's block 4 (include/initialModule.hpp:239-267,
 * fluidModule.hpp:2178-2240):
 *   left  x-face 1     wall "z4": phi = 0, thermal (+SEE) wall flux
 *   right x-face ni+1  wall "z5": phi = V_bias, thermal (+SEE) wall flux
 *   lower r-face 1     axis "x6": zero face coefficient (q_L = 0 there)
 *   upper r-face nj+1  open "x5": grad phi = 0, electron flux = ion flux
 * Interior x-faces are i = 2..ni, interior r-faces j = 2..nj (the flx/flr
 * block extents of initialModule.hpp:270-330 reduced to one block).
 *
 * Mobility is isotropic and uniform (B = 0): mu_para = mu_perp = mu_xx =
 * mu_rr = mu, mu_xr = 0, evaluated through the reference's tensor formulas so
 * that the predictor keeps its rounding (fluidModule.hpp:2314-2339).
 *
 * Ions use the electron Scharfetter-Gummel routine with charge +1
 * (vd = mu_i E + EPS, D_i = k Ti mu_i / e, no temperature-gradient drift),
 * the reference's kinetic wall flux (fluidModule.hpp:1502-1526) evaluated at
 * the drift velocity mu_i E of the wall face, and an open upper face carrying
 * n_i (mu_i E_r + EPS) (the zero-gradient limit of the SG flux,
 * fluidModule.hpp:1563-1573, 1738-1748).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static int g_literal = 0;
void orc_set_literal_bernoulli(int on) { g_literal = on; }

/* Planted mistakes for the tests that show each independent check catches
 * the error it guards against (ORC_PLANT_*, oracle.h); 0 = the restatement. */
static int g_plant = 0;
void orc_plant(int which) { g_plant = which; }
int orc_planted(void) { return g_plant; }

double orc_bern_bp(double vd, double z, double D_over_h, int literal)
{
    if (literal) return vd / (1.0 - exp(fmin(-z, ORC_LIM_EXP)));
    if (z == 0.0) return D_over_h;
    return vd / (-expm1(fmin(-z, ORC_LIM_EXP)));
}

double orc_bern_bm(double vd, double z, double D_over_h, int literal)
{
    if (literal) return vd / (exp(fmin(z, ORC_LIM_EXP)) - 1.0);
    if (z == 0.0) return D_over_h;
    return vd / expm1(fmin(z, ORC_LIM_EXP));
}

/* func_Te, miniFuncs.hpp:35-46 */
double orc_func_Te(const double* v, int n, double Te, double dTe)
{
    int iL = (int)(Te / dTe);
    int iR = (int)(Te / dTe) + 1;
    double TeL = iL * dTe;
    if (iR < n - 1) return (v[iR] - v[iL]) / (dTe + 1e-100) * (Te - TeL) + v[iL];
    return v[n - 1];
}

/* calcRes, miniFuncs.hpp:292-375, over the union of nbox boxes {i0, i1, j0, j1}
 * (block by block, as the reference loops) */
double orc_calc_res_boxes(const double* vn, const double* vo, int ld, const int* boxes, int nbox, double eps,
                          int mode)
{
    double err = 0.0, nrm = 0.0;
    int n = 0;
    for (int b = 0; b < nbox; b++) {
        const int* B = boxes + 4 * b;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double d = vn[i * ld + j] - vo[i * ld + j];
                if (mode == 0) {
                    double e = fabs(d);
                    if (e > err) err = e;
                } else if (mode == 1) {
                    err += (d * d) / (vo[i * ld + j] * vo[i * ld + j] + eps);
                    n++;
                } else if (mode == 2) {
                    double e = fabs(d);
                    if (e > err) err = e;
                    double a = fabs(vo[i * ld + j]);
                    if (a > nrm) nrm = a;
                } else {
                    err += d * d;
                    nrm += vo[i * ld + j] * vo[i * ld + j];
                }
            }
    }
    if (mode == 1) err = sqrt(err / (double)n);
    else if (mode == 2) err = err / nrm;
    else if (mode == 3) err = sqrt(err / nrm);
    return err;
}

/* calcRes over one box */
double orc_calc_res(const double* vn, const double* vo, int ld, int i0, int i1,
                    int j0, int j1, double eps, int mode)
{
    int box[4] = {i0, i1, j0, j1};
    return orc_calc_res_boxes(vn, vo, ld, box, 1, eps, mode);
}

/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* Plasma geometry: the reference's plasma blocks (cell boxes, inner face  */
/* boxes) and its boundary segments (cathode_hip.h).  NULL: the rectangle */
/* with the block-4 boundary set.  Face categories of the masks below.    */
/* ------------------------------------------------------------------ */
enum { FC_NONE = 0, FC_INNER = 1, FC_WALL = 2, FC_OPEN = 3, FC_AXIS = 4 };
static int g_have_geo = 0;
static cathode_ndd_geometry g_geo;
void orc_ndd_set_geometry(const cathode_ndd_geometry* g)
{
    g_have_geo = g != NULL;
    if (g) g_geo = *g;
}

static void rect_geo(int ni, int nj, cathode_ndd_geometry* g)
{
    memset(g, 0, sizeof(*g));
    g->n_bl = 1;
    int c[4] = {1, ni, 1, nj}, x[4] = {2, ni, 1, nj}, r[4] = {1, ni, 2, nj};
    for (int q = 0; q < 4; q++) {
        g->cell_box[q] = c[q];
        g->xface_box[q] = x[q];
        g->rface_box[q] = r[q];
    }
    cathode_ndd_segment seg[4] = {{CATHODE_SEG_X_LOW, 1, 1, nj, 0}, {CATHODE_SEG_X_HIGH, ni, 1, nj, 1},
                                  {CATHODE_SEG_R_OPEN, nj, 1, ni, 0}, {CATHODE_SEG_R_AXIS, 1, 1, ni, 0}};
    cathode_ndd_segment te[4] = {{CATHODE_SEG_X_LOW, 1, 1, nj, 1}, {CATHODE_SEG_X_HIGH, ni, 1, nj, 1},
                                 {CATHODE_SEG_R_LOW, 1, 1, ni, 1}, {CATHODE_SEG_R_HIGH, nj, 1, ni, 1}};
    g->n_seg = 4;
    g->n_te = 4;
    for (int q = 0; q < 4; q++) {
        g->seg[q] = seg[q];
        g->te[q] = te[q];
    }
}

/* the geometry of the following calls on an ni x nj grid (the rectangle when none is set) */
void orc_ndd_geometry(int ni, int nj, cathode_ndd_geometry* out)
{
    if (g_have_geo) *out = g_geo;
    else rect_geo(ni, nj, out);
}

typedef struct {
    const cathode_ndd_params* p;
    int ni, nj, ld;
    size_t n;
    const cathode_ndd_geometry* G;
    cathode_ndd_geometry Gbuf;
    unsigned char *xc, *rc; /* face categories */
    double* r; /* cell-centre radii r[0..nj+1] */
    double *aP, *aE, *aW, *aN, *aS, *b;
    double *x, *tmp_c, *tmp_x, *tmp_r, *Ex_old, *Er_old;
    double *zero;
    int solver;
    cathode_ndd_report* rep;
} ctx_t;

#define IX(i, j) ((size_t)(i) * (size_t)C->ld + (size_t)(j))

static double* src_or_zero(ctx_t* C, double* a) { return a ? a : C->zero; }

static void ctx_init(ctx_t* C, const cathode_ndd_params* p, int solver, cathode_ndd_report* rep)
{
    C->p = p;
    C->ni = p->ni;
    C->nj = p->nj;
    C->ld = p->nj + 2;
    C->n = (size_t)(p->ni + 2) * (size_t)(p->nj + 2);
    C->r = (double*)calloc(p->nj + 2, sizeof(double));
    for (int j = 0; j < p->nj + 2; j++) C->r[j] = p->rmin + (j - 0.5) * p->dr; /* initialModule.hpp:86-88 */
    double** bufs[] = {&C->aP, &C->aE, &C->aW, &C->aN, &C->aS, &C->b, &C->x,
                       &C->tmp_c, &C->tmp_x, &C->tmp_r, &C->Ex_old, &C->Er_old, &C->zero};
    for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); k++) *bufs[k] = (double*)calloc(C->n, sizeof(double));
    C->solver = solver;
    C->rep = rep;
    if (g_have_geo) C->Gbuf = g_geo;
    else rect_geo(p->ni, p->nj, &C->Gbuf);
    C->G = &C->Gbuf;
    const cathode_ndd_geometry* G = C->G;
    C->xc = (unsigned char*)calloc(C->n, 1);
    C->rc = (unsigned char*)calloc(C->n, 1);
    for (int b = 0; b < G->n_bl; b++) {
        const int *X = G->xface_box + 4 * b, *R = G->rface_box + 4 * b;
        for (int i = X[0]; i <= X[1]; i++)
            for (int j = X[2]; j <= X[3]; j++) C->xc[IX(i, j)] = FC_INNER;
        for (int i = R[0]; i <= R[1]; i++)
            for (int j = R[2]; j <= R[3]; j++) C->rc[IX(i, j)] = FC_INNER;
    }
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* q = &G->seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            switch (q->kind) {
            case CATHODE_SEG_X_LOW: C->xc[IX(q->cell, t)] = FC_WALL; break;
            case CATHODE_SEG_X_HIGH: C->xc[IX(q->cell + 1, t)] = FC_WALL; break;
            case CATHODE_SEG_R_LOW: C->rc[IX(t, q->cell)] = FC_WALL; break;
            case CATHODE_SEG_R_HIGH: C->rc[IX(t, q->cell + 1)] = FC_WALL; break;
            case CATHODE_SEG_R_OPEN: C->rc[IX(t, q->cell + 1)] = FC_OPEN; break;
            default: C->rc[IX(t, q->cell)] = FC_AXIS; break;
            }
        }
    }
}

static void ctx_free(ctx_t* C)
{
    free(C->r);
    free(C->xc);
    free(C->rc);
    double* bufs[] = {C->aP, C->aE, C->aW, C->aN, C->aS, C->b, C->x,
                      C->tmp_c, C->tmp_x, C->tmp_r, C->Ex_old, C->Er_old, C->zero};
    for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); k++) free(bufs[k]);
}

/* the linear solve handed to solver_SMG in the reference */
static void solve5(ctx_t* C, double* x, int maxITR, double tol)
{
    /* solver_SMG over the plasma blocks' cell boxes */
    const int* box = C->G->cell_box;
    const int nbox = C->G->n_bl;
    double rel = 0.0;
    int its = 0;
    if (C->solver == 0) {
        orc_struct5_direct(C->aP, C->aE, C->aW, C->aN, C->aS, C->b, box, nbox, C->ni + 2, C->nj + 2, C->ld, x);
        rel = orc_rel_residual(C->aP, C->aE, C->aW, C->aN, C->aS, C->b, box, nbox, C->ni + 2, C->nj + 2, C->ld, x);
        its = 1;
    } else {
        its = orc_mg_solve(C->aP, C->aE, C->aW, C->aN, C->aS, C->b, box, nbox, C->ni + 2, C->nj + 2, C->ld,
                           maxITR, tol, x, &rel);
    }
    if (C->rep) {
        C->rep->solves++;
        C->rep->total_iterations += its;
        if (its > C->rep->max_iterations) C->rep->max_iterations = its;
        if (rel > C->rep->max_rel_residual) C->rep->max_rel_residual = rel;
        if (C->solver != 0 && !(rel <= tol)) C->rep->nonconverged++;
    }
}

/* isotropic mobility through the reference's tensor algebra */
typedef struct {
    double mu_star, inv_par, inv_cross;
} mob_t;
static mob_t mobility(double mu)
{
    double mu_para = (mu + mu) / 2.0, mu_perp = (mu + mu) / 2.0;
    double mu_rr = (mu + mu) / 2.0, mu_xr = (0.0 + 0.0) / 2.0;
    mob_t m;
    m.mu_star = mu_perp * 1.0; /* c_mu = 1, fluidModule.hpp:2106 */
    m.inv_par = mu_rr / (mu_para * mu_perp);
    m.inv_cross = -mu_xr / (mu_para * mu_perp);
    return m;
}

/* electron thermal wall flux per density: +-0.25 vth exp(-e dphi / k Te) */
static double wall_gpn(const cathode_ndd_params* p, double Te, double dphi, double sign)
{
    double vth = sqrt(8.0 * ORC_BOLTZ * Te / (M_PI * p->masse));
    return sign * 0.25 * vth * exp(fmin(-ORC_E0 * dphi / (ORC_BOLTZ * Te + 1e-100), ORC_LIM_EXP));
}

static double eisee_sum(const cathode_ndd_params* p, double Te)
{
    if (p->n_see <= 1) return 0.0;
    double eb = orc_func_Te(p->see_eb, p->n_see, Te, p->dTe_SEE);
    double rd = orc_func_Te(p->see_rd, p->n_see, Te, p->dTe_SEE);
    double ts = orc_func_Te(p->see_ts, p->n_see, Te, p->dTe_SEE);
    return eb + rd + ts;
}

/* ion kinetic wall flux per density, fluidModule.hpp:1502-1526; sign -1 left, +1 right */
static double ion_wall_gpn(const cathode_ndd_params* p, double U, double sign)
{
    double vth = sqrt(8.0 * ORC_BOLTZ * p->Ti / (M_PI * p->massi));
    double um = sqrt(2.0 * ORC_BOLTZ * p->Ti / p->massi);
    double u = U / (um + 1e-100);
    return sign * 0.25 * vth * (exp(-u * u) + u * sqrt(M_PI) * (erf(u) + sign));
}

/* ================================================================== */
/* Poisson, fluidModule.hpp:2080-3075                                  */
/* ================================================================== */
/* masks of the transverse-flux interpolation (the reference's bLr, bRr, bLx,
 * bRx flags): 0 across a wall, and across the axis for r faces */
static double tmask_r(ctx_t* C, int i, int j) { int c = C->rc[IX(i, j)]; return (double)(c != FC_WALL && c != FC_AXIS); }
static double tmask_x(ctx_t* C, int i, int j) { return (double)(C->xc[IX(i, j)] != FC_WALL); }

/* electron thermal + emission wall flux through a wall face whose plasma
 * cell is (ci, cj): E_face the field at the face (the routine's start value
 * in the Poisson routine), Gi the ion flux at the face, h the spacing,
 * sgn -1 for a wall on the cell's low side, +1 on its high side
 * (fluidModule.hpp:3250-3367, 3453-3532) */
static double elec_wall_flux(ctx_t* C, const cathode_ndd_fields* f, const double* n, int ci, int cj, double E_face,
                             double Gi, double h, double sgn)
{
    const cathode_ndd_params* p = C->p;
    double Te = f->Te[IX(ci, cj)];
    /* the cell potential above the wall's: phi_cell - phi_wall = sgn E h / 2 */
    double dphi = fmax(sgn * E_face * h / 2.0, 0.0);
    double gpn = wall_gpn(p, Te, dphi, sgn);
    const double* nold = g_plant == ORC_PLANT_SEE_DENSITY ? n : f->rhoe_old;
    const double gi = g_plant == ORC_PLANT_SEE_ION_SIGN ? -p->coefIISEE_ts : p->coefIISEE_ts;
    double see = -gi * Gi - eisee_sum(p, Te) * nold[IX(ci, cj)] * gpn;
    return n[IX(ci, cj)] * gpn + see;
}

static void poisson_predict_assemble(ctx_t* C, cathode_ndd_fields* f, const double* nUex_t,
                                     const double* nUer_t)
{
    const cathode_ndd_params* p = C->p;
    const cathode_ndd_geometry* G = C->G;
    const double dx = p->dx, dr = p->dr, dt = p->dt;
    const double* r = C->r;
    mob_t m = mobility(p->mu_e);
    double* S = src_or_zero(C, f->rate_ionize);

    /* predicted x-flux on the blocks' inner x faces (:2266-2342) */
    for (int bl = 0; bl < G->n_bl; bl++) {
        const int* B = G->xface_box + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
                double n_f = (f->rhoe[IX(i, j)] + f->rhoe[IX(i - 1, j)]) / 2.0;
                double dphidx = (f->phi[IX(i, j)] - f->phi[IX(i - 1, j)]) / dx;
                double dpedx = ORC_BOLTZ * (f->rhoe[IX(i, j)] * f->Te[IX(i, j)] - f->rhoe[IX(i - 1, j)] * f->Te[IX(i - 1, j)]) / dx;
                double bLr = tmask_r(C, i, j), bLrD = tmask_r(C, i - 1, j);
                double bRr = tmask_r(C, i, j + 1), bRrD = tmask_r(C, i - 1, j + 1);
                double tr = (rR * nUer_t[IX(i, j + 1)] * bRr + rL * nUer_t[IX(i, j)] * bLr +
                             rR * nUer_t[IX(i - 1, j + 1)] * bRrD + rL * nUer_t[IX(i - 1, j)] * bLrD) /
                            (4.0 * r[j]);
                f->nUex[IX(i, j)] = nUex_t[IX(i, j)] - (m.inv_par * nUex_t[IX(i, j)] + m.inv_cross * tr) * m.mu_star +
                                    m.mu_star * (n_f * dphidx - dpedx / ORC_E0);
            }
    }
    /* wall x faces (:2346-2462), segment by segment, with the routine's
     * starting field */
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* q = &G->seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            if (q->kind == CATHODE_SEG_X_LOW)
                f->nUex[IX(q->cell, t)] = elec_wall_flux(C, f, f->rhoe, q->cell, t, C->Ex_old[IX(q->cell, t)],
                                                         f->rhoUix[IX(q->cell, t)], dx, -1.0);
            else if (q->kind == CATHODE_SEG_X_HIGH)
                f->nUex[IX(q->cell + 1, t)] = elec_wall_flux(C, f, f->rhoe, q->cell, t, C->Ex_old[IX(q->cell + 1, t)],
                                                             f->rhoUix[IX(q->cell + 1, t)], dx, 1.0);
        }
    }
    /* predicted r-flux on the blocks' inner r faces (:2466-2545) */
    for (int bl = 0; bl < G->n_bl; bl++) {
        const int* B = G->rface_box + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double r_t = (r[j] + r[j - 1]) / 2.0;
                double n_f = (f->rhoe[IX(i, j)] + f->rhoe[IX(i, j - 1)]) / 2.0;
                double dphidr = (f->phi[IX(i, j)] - f->phi[IX(i, j - 1)]) / dr;
                double dpedr = ORC_BOLTZ * (f->rhoe[IX(i, j)] * f->Te[IX(i, j)] - f->rhoe[IX(i, j - 1)] * f->Te[IX(i, j - 1)]) / dr;
                double bLx = tmask_x(C, i, j), bLxD = tmask_x(C, i, j - 1);
                double bRx = tmask_x(C, i + 1, j), bRxD = tmask_x(C, i + 1, j - 1);
                double tx = (r[j] * nUex_t[IX(i + 1, j)] * bRx + r[j] * nUex_t[IX(i, j)] * bLx +
                             r[j - 1] * nUex_t[IX(i + 1, j - 1)] * bRxD + r[j - 1] * nUex_t[IX(i, j - 1)] * bLxD) /
                            (4.0 * r_t);
                f->nUer[IX(i, j)] = nUer_t[IX(i, j)] - (m.inv_cross * tx + m.inv_par * nUer_t[IX(i, j)]) * m.mu_star +
                                    m.mu_star * (n_f * dphidr - dpedr / ORC_E0);
            }
    }
    /* wall r faces (:2549-2613) and the open face (:2614-2621): electron
     * flux = ion flux there */
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* q = &G->seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            if (q->kind == CATHODE_SEG_R_LOW) {
                f->nUer[IX(t, q->cell)] = elec_wall_flux(C, f, f->rhoe, t, q->cell, C->Er_old[IX(t, q->cell)],
                                                         f->rhoUir[IX(t, q->cell)], dr, -1.0);
            } else if (q->kind == CATHODE_SEG_R_HIGH) {
                f->nUer[IX(t, q->cell + 1)] = elec_wall_flux(C, f, f->rhoe, t, q->cell, C->Er_old[IX(t, q->cell + 1)],
                                                             f->rhoUir[IX(t, q->cell + 1)], dr, 1.0);
            } else if (q->kind == CATHODE_SEG_R_OPEN) {
                double uir = p->mu_i * f->Er[IX(t, q->cell + 1)] + ORC_EPS_VD;
                f->nUer[IX(t, q->cell + 1)] = f->rhoi[IX(t, q->cell)] * uir;
            }
        }
    }
    /* coefficients on the blocks' cells (:2627-2719); the plasma term of a
     * face counts between two plasma cells only */
    for (int bl = 0; bl < G->n_bl; bl++) {
        const int* B = G->cell_box + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double qR = (r[j] + r[j + 1]) / 2.0 / r[j];
                double qL = (r[j] + r[j - 1]) / 2.0 / r[j];
                double mu = m.mu_star; /* (mu_perp[i][j] + mu_perp[nb]) / 2 * c_mu */
                const double* ne = f->rhoe;
                double dte = dt;
                if (g_plant == ORC_PLANT_POISSON_DT) dte = 2.0 * dt;
                if (g_plant == ORC_PLANT_POISSON_MU) mu = mobility(p->mu_i).mu_star;
                if (g_plant == ORC_PLANT_POISSON_DENSITY) ne = f->rhoi;
                double nLx = (ne[IX(i, j)] + ne[IX(i - 1, j)]) / 2.0;
                double nRx = (ne[IX(i, j)] + ne[IX(i + 1, j)]) / 2.0;
                double nLr = (ne[IX(i, j)] + ne[IX(i, j - 1)]) / 2.0;
                double nRr = (ne[IX(i, j)] + ne[IX(i, j + 1)]) / 2.0;
                double bLx = (double)(C->xc[IX(i, j)] == FC_INNER), bRx = (double)(C->xc[IX(i + 1, j)] == FC_INNER);
                double bLr = (double)(C->rc[IX(i, j)] == FC_INNER), bRr = (double)(C->rc[IX(i, j + 1)] == FC_INNER);
                double eRx = ORC_EPS0 + ORC_E0 * dte * nRx * mu * bRx;
                double eLx = ORC_EPS0 + ORC_E0 * dte * nLx * mu * bLx;
                double eRr = ORC_EPS0 + ORC_E0 * dte * nRr * mu * bRr;
                double eLr = ORC_EPS0 + ORC_E0 * dte * nLr * mu * bLr;
                C->aP[IX(i, j)] = (eRx + eLx) * dr * dr + (qR * eRr + qL * eLr) * dx * dx;
                C->aE[IX(i, j)] = eRx * dr * dr;
                C->aW[IX(i, j)] = eLx * dr * dr;
                C->aN[IX(i, j)] = qR * eRr * dx * dx;
                C->aS[IX(i, j)] = qL * eLr * dx * dx;
                double div = (f->nUex[IX(i + 1, j)] - f->nUex[IX(i, j)]) / dx +
                             (qR * f->nUer[IX(i, j + 1)] - qL * f->nUer[IX(i, j)]) / dr;
                double lap = -(f->Ex[IX(i + 1, j)] - f->Ex[IX(i, j)]) / dx - (qR * f->Er[IX(i, j + 1)] - qL * f->Er[IX(i, j)]) / dr;
                const double s_on = g_plant == ORC_PLANT_POISSON_SOURCE ? 0.0 : 1.0;
                C->b[IX(i, j)] = (ORC_E0 * f->rhoi[IX(i, j)] - ORC_E0 * f->rhoe[IX(i, j)] - s_on * ORC_E0 * dt * S[IX(i, j)] +
                                  ORC_EPS0 * lap + ORC_E0 * dt * div) *
                                 dx * dx * dr * dr;
            }
    }
    /* BC folds (:2721-2824), segment by segment: Dirichlet walls by ghost
     * mirror aP += a_nb, the open face grad phi = 0: aP -= a_nb */
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* q = &G->seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            if (q->kind == CATHODE_SEG_X_LOW) {
                size_t k = IX(q->cell, t);
                C->aP[k] = C->aP[k] + C->aW[k];
                C->aW[k] = 0.0;
            } else if (q->kind == CATHODE_SEG_X_HIGH) {
                size_t k = IX(q->cell, t);
                C->aP[k] = C->aP[k] + C->aE[k];
                C->aE[k] = 0.0;
            } else if (q->kind == CATHODE_SEG_R_LOW) {
                size_t k = IX(t, q->cell);
                C->aP[k] = C->aP[k] + C->aS[k];
                C->aS[k] = 0.0;
            } else if (q->kind == CATHODE_SEG_R_HIGH) {
                if (g_plant == ORC_PLANT_GEO_FOLD) continue;
                size_t k = IX(t, q->cell);
                C->aP[k] = C->aP[k] + C->aN[k];
                C->aN[k] = 0.0;
            } else if (q->kind == CATHODE_SEG_R_OPEN) {
                size_t k = IX(t, q->cell);
                C->aP[k] = C->aP[k] - C->aN[k];
                C->aN[k] = 0.0;
            }
        }
    }
}

/* the calcRes of the three planes of a routine over the blocks' boxes */
static void calc_res3(ctx_t* C, const double* n, const double* n0, const double* fx, const double* fx0,
                      const double* fr, const double* fr0, double* e)
{
    const cathode_ndd_geometry* G = C->G;
    int mode = C->p->flag_error;
    e[0] = orc_calc_res_boxes(n, n0, C->ld, G->cell_box, G->n_bl, 1e-100, mode);
    e[1] = orc_calc_res_boxes(fx, fx0, C->ld, G->xface_box, G->n_bl, 1e-100, mode);
    e[2] = orc_calc_res_boxes(fr, fr0, C->ld, G->rface_box, G->n_bl, 1e-100, mode);
}

static void poisson_routine(ctx_t* C, cathode_ndd_fields* f, double* phi_old, double* nUex_old,
                            double* nUer_old)
{
    const cathode_ndd_params* p = C->p;
    const cathode_ndd_geometry* G = C->G;
    const double dx = p->dx, dr = p->dr;
    mob_t m = mobility(p->mu_e);
    double* dphi = C->x;
    memcpy(phi_old, f->phi, C->n * sizeof(double));
    memcpy(C->Ex_old, f->Ex, C->n * sizeof(double));
    memcpy(C->Er_old, f->Er, C->n * sizeof(double));
    memcpy(nUex_old, f->nUex, C->n * sizeof(double));
    memcpy(nUer_old, f->nUer, C->n * sizeof(double));
    for (size_t k = 0; k < C->n; k++) dphi[k] = f->phi[k] - phi_old[k];
    /* ghost potential (:2152-2249), segment by segment: phi = 0 or V_bias
     * walls, grad phi = 0 on the open face */
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* q = &G->seg[s];
        for (int t = q->lo; t <= q->hi; t++) {
            if (q->kind == CATHODE_SEG_X_LOW)
                f->phi[IX(q->cell - 1, t)] = (q->flag & CATHODE_SEG_BIAS) ? 2.0 * p->V_bias - f->phi[IX(q->cell, t)] : -f->phi[IX(q->cell, t)];
            else if (q->kind == CATHODE_SEG_X_HIGH)
                f->phi[IX(q->cell + 1, t)] = (q->flag & CATHODE_SEG_BIAS) ? 2.0 * p->V_bias - f->phi[IX(q->cell, t)] : -f->phi[IX(q->cell, t)];
            else if (q->kind == CATHODE_SEG_R_LOW)
                f->phi[IX(t, q->cell - 1)] = -f->phi[IX(t, q->cell)];
            else if (q->kind == CATHODE_SEG_R_HIGH)
                f->phi[IX(t, q->cell + 1)] = -f->phi[IX(t, q->cell)];
            else if (q->kind == CATHODE_SEG_R_OPEN)
                f->phi[IX(t, q->cell + 1)] = f->phi[IX(t, q->cell)];
        }
    }

    double* phi_t = C->tmp_c;
    double* nUex_t = C->tmp_x;
    double* nUer_t = C->tmp_r;
    int ncount = 0;
    for (;;) {
        ncount++;
        memcpy(phi_t, f->phi, C->n * sizeof(double));
        memcpy(nUex_t, f->nUex, C->n * sizeof(double));
        memcpy(nUer_t, f->nUer, C->n * sizeof(double));
        poisson_predict_assemble(C, f, nUex_t, nUer_t);
        solve5(C, dphi, p->maxITR_SOR_phi, p->error_cnv_SOR_phi);
        /* ghost correction and wall-face field (:2840-2943), segment by segment */
        for (int s = 0; s < G->n_seg; s++) {
            const cathode_ndd_segment* q = &G->seg[s];
            for (int t = q->lo; t <= q->hi; t++) {
                if (q->kind == CATHODE_SEG_X_LOW) {
                    int i = q->cell, j = t;
                    dphi[IX(i - 1, j)] = -dphi[IX(i, j)];
                    f->Ex[IX(i, j)] = f->Ex[IX(i, j)] - (dphi[IX(i, j)] - dphi[IX(i - 1, j)]) / dx;
                } else if (q->kind == CATHODE_SEG_X_HIGH) {
                    int i = q->cell, j = t;
                    dphi[IX(i + 1, j)] = -dphi[IX(i, j)];
                    f->Ex[IX(i + 1, j)] = f->Ex[IX(i + 1, j)] - (dphi[IX(i + 1, j)] - dphi[IX(i, j)]) / dx;
                } else if (q->kind == CATHODE_SEG_R_LOW) {
                    int i = t, j = q->cell;
                    dphi[IX(i, j - 1)] = -dphi[IX(i, j)];
                    f->Er[IX(i, j)] = f->Er[IX(i, j)] - (dphi[IX(i, j)] - dphi[IX(i, j - 1)]) / dr;
                } else if (q->kind == CATHODE_SEG_R_HIGH) {
                    int i = t, j = q->cell;
                    dphi[IX(i, j + 1)] = -dphi[IX(i, j)];
                    f->Er[IX(i, j + 1)] = f->Er[IX(i, j + 1)] - (dphi[IX(i, j + 1)] - dphi[IX(i, j)]) / dr;
                } else if (q->kind == CATHODE_SEG_R_OPEN) {
                    int i = t, j = q->cell;
                    dphi[IX(i, j + 1)] = dphi[IX(i, j)];
                    f->Er[IX(i, j + 1)] = 0.0;
                }
            }
        }
        for (size_t k = 0; k < C->n; k++) f->phi[k] = f->phi[k] + dphi[k];
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->xface_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) f->Ex[IX(i, j)] = -(f->phi[IX(i, j)] - f->phi[IX(i - 1, j)]) / dx;
        }
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->rface_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) f->Er[IX(i, j)] = -(f->phi[IX(i, j)] - f->phi[IX(i, j - 1)]) / dr;
        }
        /* flux correction on the inner faces (:2980-3033) */
        if (g_plant == ORC_PLANT_POISSON_FLUX_MU) m = mobility(p->mu_i);
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->xface_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) {
                    double n_f = (f->rhoe[IX(i, j)] + f->rhoe[IX(i - 1, j)]) / 2.0;
                    f->nUex[IX(i, j)] = f->nUex[IX(i, j)] + m.mu_star * n_f * ((dphi[IX(i, j)] - dphi[IX(i - 1, j)]) / dx);
                }
        }
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->rface_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) {
                    double n_f = (f->rhoe[IX(i, j)] + f->rhoe[IX(i, j - 1)]) / 2.0;
                    f->nUer[IX(i, j)] = f->nUer[IX(i, j)] + m.mu_star * n_f * ((dphi[IX(i, j)] - dphi[IX(i, j - 1)]) / dr);
                }
        }
        double e[3];
        calc_res3(C, f->phi, phi_t, f->nUex, nUex_t, f->nUer, nUer_t, e);
        double eg = fmax(fmax(e[0], e[1]), e[2]);
        if (eg < p->error_cnv_HES_phi || ncount >= p->maxITR_HES_phi) break;
    }
    if (C->rep) {
        double e[3];
        calc_res3(C, f->phi, phi_old, f->nUex, nUex_old, f->nUer, nUer_old, e);
        C->rep->error_phi = e[0];
        C->rep->error_nUex = e[1];
        C->rep->error_nUer = e[2];
    }
}

/* ================================================================== */
/* Scharfetter-Gummel continuity: electrons (fluidModule.hpp:3082-4000) */
/* and ions (charge +1), energy (:4006-4957)                          */
/* ================================================================== */
enum { SP_ELEC = 0, SP_ION = 1, SP_ENERGY = 2 };

typedef struct {
    double vd, D, z;
} face_t;

/* face drift/diffusion of species sp between cells (a) upstream and (c) at
 * spacing h with face field E; Tup/Tdn: temperatures [K] of the two cells */
static face_t sg_face(const cathode_ndd_params* p, int sp, double mu_star, double E, double Ta,
                      double Tc, double h)
{
    face_t F;
    if (sp == SP_ION) {
        F.vd = mu_star * (E) + ORC_EPS_VD;
        F.D = ORC_BOLTZ * p->Ti / ORC_E0 * mu_star;
    } else {
        double T_f = (Tc + Ta) / 2.0;
        double dT = (Tc - Ta) / h;
        F.vd = mu_star * (-E - ORC_BOLTZ / ORC_E0 * dT) + ORC_EPS_VD;
        F.D = ORC_BOLTZ * T_f / ORC_E0 * mu_star;
    }
    F.z = F.vd * h / F.D;
    return F;
}

typedef struct {
    double* n;      /* transported density (rhoe, rhoi, rhoeps) */
    double* Fx;     /* x flux */
    double* Fr;     /* r flux */
    const double* T; /* temperature used by the SG faces (Te or Te_old) */
    double mu;
    double cfac;    /* flux factor: 1 or 5/3 */
    double tfac;    /* time-term factor: 1 or 3/5 */
    int maxit, lin_maxit;   /* PC iterations; linear-solve cap (maxITR_SOR_*) */
    double cnv, lin_tol;    /* PC threshold; linear-solve tolerance (error_cnv_SOR_*) */
} sg_eq_t;

/* wall flux of species sp through a wall face whose plasma cell is (ci, cj):
 * E the face field, Gi the ion flux at the face, h the spacing, sgn -1 for a
 * wall on the cell's low side, +1 on its high side (electrons
 * fluidModule.hpp:3250-3367, 3453-3532; energy :4161-4316, :4403-4506; ions
 * the kinetic flux of :1502-1526 at the drift mu_i E) */
static double sg_wall_flux(ctx_t* C, const cathode_ndd_fields* f, int sp, const sg_eq_t* q, const double* n_old,
                           int ci, int cj, double E, double Gi, double h, double sgn)
{
    const cathode_ndd_params* p = C->p;
    const double* n = q->n;
    if (sp == SP_ION) {
        double U = p->mu_i * E;
        return n[IX(ci, cj)] * ion_wall_gpn(p, U, sgn);
    } else if (sp == SP_ELEC) {
        return elec_wall_flux(C, f, n, ci, cj, E, Gi, h, sgn);
    }
    const double* T = q->T;
    double Teo = T[IX(ci, cj)];
    double dphi = fmax(sgn * E * h / 2.0, 0.0);
    double vth = sqrt(8.0 * ORC_BOLTZ * Teo / (M_PI * p->masse));
    double ex = exp(fmin(-ORC_E0 * dphi / (ORC_BOLTZ * Teo + 1e-100), ORC_LIM_EXP));
    double gpne = sgn * vth * ex / 3.0 * 1.0;
    double eps_emit = (g_plant == ORC_PLANT_SEE_EMIT_ENERGY ? 1.0 : 2.0) * ORC_E0 * p->Te_emitSEE_eV;
    double gpn = sgn * 0.25 * 1.0 * vth * ex;
    double Te = f->Te[IX(ci, cj)];
    double eb = 0.0, rd = 0.0, ts = 0.0;
    if (p->n_see > 1) {
        eb = orc_func_Te(p->see_eb, p->n_see, Te, p->dTe_SEE);
        rd = orc_func_Te(p->see_rd, p->n_see, Te, p->dTe_SEE);
        ts = orc_func_Te(p->see_ts, p->n_see, Te, p->dTe_SEE);
    }
    double see = -eps_emit * p->coefIISEE_ts * Gi -
                 (eb + (g_plant == ORC_PLANT_SEE_RD_RATIO ? 1.0 : p->ratioEngy_EISEE_rd) * rd) * n_old[IX(ci, cj)] * gpne -
                 eps_emit * ts * f->rhoe_old[IX(ci, cj)] * gpn;
    return n[IX(ci, cj)] * gpne + see;
}

static void sg_predict_assemble(ctx_t* C, cathode_ndd_fields* f, int sp, const sg_eq_t* q,
                                const double* Fx_t, const double* Fr_t, const double* n_old)
{
    const cathode_ndd_params* p = C->p;
    const cathode_ndd_geometry* G = C->G;
    const double dx = p->dx, dr = p->dr, dt = p->dt;
    const double* r = C->r;
    const int lit = g_literal;
    mob_t m = mobility(q->mu);
    const double* T = q->T;
    double* n = q->n;

    /* the blocks' inner x faces (:3160-3244) */
    for (int bl = 0; bl < G->n_bl; bl++) {
        const int* B = G->xface_box + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
                face_t F = sg_face(p, sp, m.mu_star, f->Ex[IX(i, j)], T[IX(i - 1, j)], T[IX(i, j)], dx);
                double bLr = tmask_r(C, i, j), bLrD = tmask_r(C, i - 1, j);
                double bRr = tmask_r(C, i, j + 1), bRrD = tmask_r(C, i - 1, j + 1);
                double tr = (rR * Fr_t[IX(i, j + 1)] * bRr + rL * Fr_t[IX(i, j)] * bLr +
                             rR * Fr_t[IX(i - 1, j + 1)] * bRrD + rL * Fr_t[IX(i - 1, j)] * bLrD) /
                            (4.0 * r[j]);
                double sg = orc_bern_bp(F.vd, F.z, F.D / dx, lit) * n[IX(i - 1, j)] - orc_bern_bm(F.vd, F.z, F.D / dx, lit) * n[IX(i, j)];
                if (sp == SP_ENERGY) sg = sg * 5.0 / 3.0;
                q->Fx[IX(i, j)] = sg + Fx_t[IX(i, j)] - m.mu_star * (m.inv_par * Fx_t[IX(i, j)] + m.inv_cross * tr);
            }
    }
    /* wall x faces (:3250-3367), segment by segment */
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* g = &G->seg[s];
        for (int t = g->lo; t <= g->hi; t++) {
            if (g->kind == CATHODE_SEG_X_LOW)
                q->Fx[IX(g->cell, t)] = sg_wall_flux(C, f, sp, q, n_old, g->cell, t, f->Ex[IX(g->cell, t)],
                                                     f->rhoUix[IX(g->cell, t)], dx, -1.0);
            else if (g->kind == CATHODE_SEG_X_HIGH)
                q->Fx[IX(g->cell + 1, t)] = sg_wall_flux(C, f, sp, q, n_old, g->cell, t, f->Ex[IX(g->cell + 1, t)],
                                                         f->rhoUix[IX(g->cell + 1, t)], dx, 1.0);
        }
    }
    /* the blocks' inner r faces (:3370-3450) */
    for (int bl = 0; bl < G->n_bl; bl++) {
        const int* B = G->rface_box + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double r_t = (r[j] + r[j - 1]) / 2.0;
                face_t F = sg_face(p, sp, m.mu_star, f->Er[IX(i, j)], T[IX(i, j - 1)], T[IX(i, j)], dr);
                double bLx = tmask_x(C, i, j), bLxD = tmask_x(C, i, j - 1);
                double bRx = tmask_x(C, i + 1, j), bRxD = tmask_x(C, i + 1, j - 1);
                double tx = (r[j] * Fx_t[IX(i + 1, j)] * bRx + r[j] * Fx_t[IX(i, j)] * bLx +
                             r[j - 1] * Fx_t[IX(i + 1, j - 1)] * bRxD + r[j - 1] * Fx_t[IX(i, j - 1)] * bLxD) /
                            (4.0 * r_t);
                double sg = orc_bern_bp(F.vd, F.z, F.D / dr, lit) * n[IX(i, j - 1)] - orc_bern_bm(F.vd, F.z, F.D / dr, lit) * n[IX(i, j)];
                if (sp == SP_ENERGY) sg = sg * 5.0 / 3.0;
                q->Fr[IX(i, j)] = sg + Fr_t[IX(i, j)] - m.mu_star * (m.inv_cross * tx + m.inv_par * Fr_t[IX(i, j)]);
            }
    }
    /* wall r faces (:3453-3532) and the open face (:3533-3542): the ion
     * flux, or for the energy 5/2 k T max(Gamma_e, 0) (:4492-4499) */
    for (int s = 0; s < G->n_seg; s++) {
        const cathode_ndd_segment* g = &G->seg[s];
        for (int t = g->lo; t <= g->hi; t++) {
            if (g->kind == CATHODE_SEG_R_LOW) {
                q->Fr[IX(t, g->cell)] = sg_wall_flux(C, f, sp, q, n_old, t, g->cell, f->Er[IX(t, g->cell)],
                                                     f->rhoUir[IX(t, g->cell)], dr, -1.0);
            } else if (g->kind == CATHODE_SEG_R_HIGH) {
                q->Fr[IX(t, g->cell + 1)] = sg_wall_flux(C, f, sp, q, n_old, t, g->cell, f->Er[IX(t, g->cell + 1)],
                                                         f->rhoUir[IX(t, g->cell + 1)], dr, 1.0);
            } else if (g->kind == CATHODE_SEG_R_OPEN) {
                int i = t, j = g->cell;
                if (sp == SP_ENERGY) {
                    q->Fr[IX(i, j + 1)] = 5.0 / 2.0 * ORC_BOLTZ * T[IX(i, j)] * fmax(f->rhoUer[IX(i, j + 1)], 0.0);
                } else {
                    double uir = p->mu_i * f->Er[IX(i, j + 1)] + ORC_EPS_VD;
                    q->Fr[IX(i, j + 1)] = f->rhoi[IX(i, j)] * uir;
                }
            }
        }
    }
    /* coefficients on the blocks' cells (:3544-3664, :4508-4642) */
    double* S = src_or_zero(C, f->rate_ionize);
    double* Pabs = src_or_zero(C, f->Pabs);
    double* eloss = src_or_zero(C, f->rate_eloss);
    for (int bl = 0; bl < G->n_bl; bl++) {
        const int* B = G->cell_box + 4 * bl;
        for (int i = B[0]; i <= B[1]; i++)
            for (int j = B[2]; j <= B[3]; j++) {
                double rR = (r[j] + r[j + 1]) / 2.0, rL = (r[j] + r[j - 1]) / 2.0;
                double qR = (r[j] + r[j + 1]) / 2.0 / r[j];
                double qL = (r[j] + r[j - 1]) / 2.0 / r[j];
                face_t Lx = sg_face(p, sp, m.mu_star, f->Ex[IX(i, j)], T[IX(i - 1, j)], T[IX(i, j)], dx);
                face_t Rx = sg_face(p, sp, m.mu_star, f->Ex[IX(i + 1, j)], T[IX(i, j)], T[IX(i + 1, j)], dx);
                face_t Lr = sg_face(p, sp, m.mu_star, f->Er[IX(i, j)], T[IX(i, j - 1)], T[IX(i, j)], dr);
                face_t Rr = sg_face(p, sp, m.mu_star, f->Er[IX(i, j + 1)], T[IX(i, j)], T[IX(i, j + 1)], dr);
                /* the face coefficients count between two plasma cells only
                 * (walls, open face and axis are explicit fluxes) */
                double bLx = (double)(C->xc[IX(i, j)] == FC_INNER), bRx = (double)(C->xc[IX(i + 1, j)] == FC_INNER);
                double bLr = (double)(C->rc[IX(i, j)] == FC_INNER), bRr = (double)(C->rc[IX(i, j + 1)] == FC_INNER);
                if (g_plant == ORC_PLANT_GEO_WALL_COEF) {
                    bLx = (double)(C->xc[IX(i, j)] != FC_NONE);
                    bRx = (double)(C->xc[IX(i + 1, j)] != FC_NONE);
                }
                double aP = orc_bern_bp(Rx.vd, Rx.z, Rx.D / dx, lit) * bRx * dr + orc_bern_bm(Lx.vd, Lx.z, Lx.D / dx, lit) * bLx * dr +
                            qR * orc_bern_bp(Rr.vd, Rr.z, Rr.D / dr, lit) * bRr * dx +
                            qL * orc_bern_bm(Lr.vd, Lr.z, Lr.D / dr, lit) * bLr * dx;
                C->aP[IX(i, j)] = aP + q->tfac * dx * dr / dt;
                C->aE[IX(i, j)] = orc_bern_bm(Rx.vd, Rx.z, Rx.D / dx, lit) * bRx * dr;
                C->aW[IX(i, j)] = orc_bern_bp(Lx.vd, Lx.z, Lx.D / dx, lit) * bLx * dr;
                C->aN[IX(i, j)] = qR * orc_bern_bm(Rr.vd, Rr.z, Rr.D / dr, lit) * bRr * dx;
                C->aS[IX(i, j)] = qL * orc_bern_bp(Lr.vd, Lr.z, Lr.D / dr, lit) * bLr * dx;
                double div = (q->Fx[IX(i + 1, j)] - q->Fx[IX(i, j)]) / dx + (qR * q->Fr[IX(i, j + 1)] - qL * q->Fr[IX(i, j)]) / dr;
                if (sp == SP_ENERGY) {
                    /* J.E without the wall faces (:4624-4625) */
                    double bLxw = (double)(C->xc[IX(i, j)] != FC_WALL), bRxw = (double)(C->xc[IX(i + 1, j)] != FC_WALL);
                    double bLrw = (double)(C->rc[IX(i, j)] != FC_WALL), bRrw = (double)(C->rc[IX(i, j + 1)] != FC_WALL);
                    double JEx = -ORC_E0 * (f->rhoUex[IX(i + 1, j)] * f->Ex[IX(i + 1, j)] * bRxw + f->rhoUex[IX(i, j)] * f->Ex[IX(i, j)] * bLxw) / 2.0;
                    double JEr = -ORC_E0 * (rR * f->rhoUer[IX(i, j + 1)] * f->Er[IX(i, j + 1)] * bRrw + rL * f->rhoUer[IX(i, j)] * f->Er[IX(i, j)] * bLrw) /
                                 (2.0 * r[j]);
                    double R1 = (-n[IX(i, j)] + n_old[IX(i, j)]) / dt;
                    double R2 = JEx + JEr;
                    double R3 = Pabs[IX(i, j)] - f->rhoe[IX(i, j)] * eloss[IX(i, j)];
                    double R4 = -div;
                    C->b[IX(i, j)] = 3.0 / 5.0 * (R1 + R2 + R3 + R4) * dx * dr;
                } else {
                    C->b[IX(i, j)] = (-div + S[IX(i, j)] - (n[IX(i, j)] - n_old[IX(i, j)]) / dt) * dx * dr;
                }
            }
    }
}

static void sg_routine(ctx_t* C, cathode_ndd_fields* f, int sp, const sg_eq_t* q, double* n_old,
                       double* Fx_old, double* Fr_old, double* err3)
{
    const cathode_ndd_params* p = C->p;
    const cathode_ndd_geometry* G = C->G;
    const double dx = p->dx, dr = p->dr;
    const int lit = g_literal;
    mob_t m = mobility(q->mu);
    double* dn = C->x;
    memcpy(n_old, q->n, C->n * sizeof(double));
    memcpy(Fx_old, q->Fx, C->n * sizeof(double));
    memcpy(Fr_old, q->Fr, C->n * sizeof(double));
    for (size_t k = 0; k < C->n; k++) dn[k] = q->n[k] - n_old[k];
    double* n_t = C->tmp_c;
    double* Fx_t = C->tmp_x;
    double* Fr_t = C->tmp_r;
    int ncount = 0;
    for (;;) {
        ncount++;
        memcpy(n_t, q->n, C->n * sizeof(double));
        memcpy(Fx_t, q->Fx, C->n * sizeof(double));
        memcpy(Fr_t, q->Fr, C->n * sizeof(double));
        sg_predict_assemble(C, f, sp, q, Fx_t, Fr_t, n_old);
        solve5(C, dn, q->lin_maxit, q->lin_tol);
        /* flux correction on the inner faces (:3682-3743) */
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->xface_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) {
                    face_t F = sg_face(p, sp, m.mu_star, f->Ex[IX(i, j)], q->T[IX(i - 1, j)], q->T[IX(i, j)], dx);
                    double d = orc_bern_bp(F.vd, F.z, F.D / dx, lit) * dn[IX(i - 1, j)] - orc_bern_bm(F.vd, F.z, F.D / dx, lit) * dn[IX(i, j)];
                    if (sp == SP_ENERGY) d = d * 5.0 / 3.0;
                    q->Fx[IX(i, j)] = q->Fx[IX(i, j)] + d;
                }
        }
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->rface_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) {
                    face_t F = sg_face(p, sp, m.mu_star, f->Er[IX(i, j)], q->T[IX(i, j - 1)], q->T[IX(i, j)], dr);
                    double d = orc_bern_bp(F.vd, F.z, F.D / dr, lit) * dn[IX(i, j - 1)] - orc_bern_bm(F.vd, F.z, F.D / dr, lit) * dn[IX(i, j)];
                    if (sp == SP_ENERGY) d = d * 5.0 / 3.0;
                    q->Fr[IX(i, j)] = q->Fr[IX(i, j)] + d;
                }
        }
        /* positivity-guarded update on the blocks' cells (:3748-3757) */
        for (int bl = 0; bl < G->n_bl; bl++) {
            const int* B = G->cell_box + 4 * bl;
            for (int i = B[0]; i <= B[1]; i++)
                for (int j = B[2]; j <= B[3]; j++) {
                    double nn = q->n[IX(i, j)] + dn[IX(i, j)];
                    if (nn > 0.0) q->n[IX(i, j)] = nn;
                }
        }
        double e[3];
        calc_res3(C, q->n, n_t, q->Fx, Fx_t, q->Fr, Fr_t, e);
        double eg = fmax(fmax(e[0], e[1]), e[2]);
        if (eg < q->cnv || ncount >= q->maxit) break;
    }
    calc_res3(C, q->n, n_old, q->Fx, Fx_old, q->Fr, Fr_old, err3);
}

/* ghost writes after a routine: the neighbour across each listed face
 * takes the cell's value (flag 1) or 0 (flag 0), in list order */
static void ghost_list(ctx_t* C, double* a, const cathode_ndd_segment* list, int cnt)
{
    for (int s = 0; s < cnt; s++) {
        const cathode_ndd_segment* g = &list[s];
        for (int t = g->lo; t <= g->hi; t++) {
            int ci, cj, gi, gj;
            if (g->kind == CATHODE_SEG_X_LOW) { ci = g->cell; cj = t; gi = ci - 1; gj = cj; }
            else if (g->kind == CATHODE_SEG_X_HIGH) { ci = g->cell; cj = t; gi = ci + 1; gj = cj; }
            else if (g->kind == CATHODE_SEG_R_LOW) { ci = t; cj = g->cell; gi = ci; gj = cj - 1; }
            else { ci = t; cj = g->cell; gi = ci; gj = cj + 1; }
            a[IX(gi, gj)] = g->flag ? a[IX(ci, cj)] : 0.0;
        }
    }
}

static void set_eq(const cathode_ndd_params* p, cathode_ndd_fields* f, int sp, const double* Te_old, sg_eq_t* q)
{
    if (sp == SP_ELEC) {
        q->n = f->rhoe; q->Fx = f->rhoUex; q->Fr = f->rhoUer; q->T = f->Te; q->mu = p->mu_e;
        q->cfac = 1.0; q->tfac = 1.0; q->maxit = p->maxITR_HES_rhoe; q->cnv = p->error_cnv_HES_rhoe;
        q->lin_maxit = p->maxITR_SOR_rhoe; q->lin_tol = p->error_cnv_SOR_rhoe;
    } else if (sp == SP_ION) {
        q->n = f->rhoi; q->Fx = f->rhoUix; q->Fr = f->rhoUir; q->T = f->Te; q->mu = p->mu_i;
        q->cfac = 1.0; q->tfac = 1.0; q->maxit = p->maxITR_HES_rhoi; q->cnv = p->error_cnv_HES_rhoi;
        q->lin_maxit = p->maxITR_SOR_rhoi; q->lin_tol = p->error_cnv_SOR_rhoi;
    } else {
        q->n = f->rhoeps; q->Fx = f->Gx; q->Fr = f->Gr; q->T = Te_old; q->mu = p->mu_e;
        q->cfac = 5.0 / 3.0; q->tfac = 3.0 / 5.0; q->maxit = p->maxITR_HES_rhoeps; q->cnv = p->error_cnv_HES_rhoeps;
        q->lin_maxit = p->maxITR_SOR_rhoeps; q->lin_tol = p->error_cnv_SOR_rhoeps;
    }
}

/* ------------------------------------------------------------------ */
/* check_CFL, fluidModule.hpp:57-115: maxima of n_e and Te over every cell
 * (ghosts included), of mu_para (uniform mu_e here, B = 0) and of the
 * electron speed |Gamma_e| / n_e over the interior (the x flux averaged over
 * the cell's two faces, the r flux r-weighted), the thermal speed at max Te
 * added; dt_min = eps0 / (e max n_e max mu).  out = {dt_min, rhoe_max,
 * mu_max, Te_max, ue_max}. */
void orc_check_cfl(const cathode_ndd_params* p, const cathode_ndd_fields* f, double* out)
{
    const int ni = p->ni, nj = p->nj, ld = nj + 2;
    double rhoe_max = 0.0, Te_max = 0.0, ue_max = 0.0;
    const double mu_max = fmax(0.0, p->mu_e);
    for (int i = 0; i < ni + 2; i++)
        for (int j = 0; j < nj + 2; j++) {
            if (rhoe_max < f->rhoe[i * ld + j]) rhoe_max = f->rhoe[i * ld + j];
            if (Te_max < f->Te[i * ld + j]) Te_max = f->Te[i * ld + j];
        }
    for (int i = 1; i < ni + 1; i++)
        for (int j = 1; j < nj + 1; j++) {
            double rj = p->rmin + (j - 0.5) * p->dr, rjm = p->rmin + (j - 1.5) * p->dr,
                   rjp = p->rmin + (j + 0.5) * p->dr;
            double rL = (rj + rjm) / 2.0, rR = (rj + rjp) / 2.0;
            double n = f->rhoe[i * ld + j] + 1e-100;
            double ux = (f->rhoUex[(i + 1) * ld + j] + f->rhoUex[i * ld + j]) / 2.0 / n;
            double ur = (rR * f->rhoUer[i * ld + j + 1] + rL * f->rhoUer[i * ld + j]) / (2.0 * rj) / n;
            double u = sqrt(ux * ux + ur * ur);
            if (ue_max < u) ue_max = u;
        }
    double vth = sqrt(8.0 * ORC_BOLTZ * Te_max / (M_PI * p->masse));
    out[0] = ORC_EPS0 / (ORC_E0 * rhoe_max * mu_max);
    out[1] = rhoe_max;
    out[2] = mu_max;
    out[3] = Te_max;
    out[4] = ue_max + vth;
}

int orc_ndd_step(const cathode_ndd_params* pin, cathode_ndd_fields* f, int nsteps, int solver,
                 cathode_ndd_report* rep)
{
    ctx_t Cx, *C = &Cx;
    /* a copy whose dt check_CFL may set (flag_adp_dt = 1) */
    cathode_ndd_params pl = *pin;
    const cathode_ndd_params* p = &pl;
    if (rep) memset(rep, 0, sizeof(*rep));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    ctx_init(C, p, solver, rep);
    const cathode_ndd_geometry* G = C->G;
    double* s_old = (double*)calloc(C->n, sizeof(double));
    double* sx_old = (double*)calloc(C->n, sizeof(double));
    double* sr_old = (double*)calloc(C->n, sizeof(double));
    double* Te_old = (double*)calloc(C->n, sizeof(double));
    double* rhoi_old = (double*)calloc(C->n, sizeof(double));
    double err3[3];
    sg_eq_t q;
    for (int step = 0; step < nsteps; step++) {
        /* 0. time step (main.cpp:174): adaptive dt, or the CFL number of the fixed one */
        double cfl[5];
        orc_check_cfl(p, f, cfl);
        if (pl.flag_adp_dt == 1) pl.dt = cfl[0] * pl.CFL;
        else pl.CFL = pl.dt / cfl[0];
        if (rep) {
            rep->dt = pl.dt;
            rep->CFL = pl.CFL;
            rep->dt_min = cfl[0];
            rep->rhoe_max = cfl[1];
            rep->Te_max = cfl[3];
            rep->ue_max = cfl[4];
        }
        /* 1. ions */
        set_eq(p, f, SP_ION, NULL, &q);
        sg_routine(C, f, SP_ION, &q, rhoi_old, sx_old, sr_old, err3);
        /* the ion density across the open faces */
        for (int s = 0; s < G->n_seg; s++)
            if (G->seg[s].kind == CATHODE_SEG_R_OPEN)
                for (int i = G->seg[s].lo; i <= G->seg[s].hi; i++)
                    f->rhoi[IX(i, G->seg[s].cell + 1)] = f->rhoi[IX(i, G->seg[s].cell)];
        if (rep) rep->error_rhoi = err3[0];
        /* 2. Poisson */
        poisson_routine(C, f, s_old, sx_old, sr_old);
        /* 3. electrons */
        set_eq(p, f, SP_ELEC, NULL, &q);
        sg_routine(C, f, SP_ELEC, &q, f->rhoe_old, sx_old, sr_old, err3);
        /* electron-density ghosts (:3873-3913): 0 behind walls, copies across open sides */
        ghost_list(C, f->rhoe, G->ne, G->n_ne);
        if (rep) {
            rep->error_rhoe = err3[0];
            rep->error_rhoUex = err3[1];
            rep->error_rhoUer = err3[2];
            /* the electrode current over the faces at V_bias (:3942-3965) */
            double Ii = 0.0, Ie = 0.0;
            for (int s = 0; s < G->n_seg; s++) {
                const cathode_ndd_segment* g = &G->seg[s];
                if (g->kind != CATHODE_SEG_X_HIGH || !(g->flag & CATHODE_SEG_BIAS)) continue;
                for (int j = g->lo; j <= g->hi; j++) {
                    double di = f->rhoUix[IX(g->cell + 1, j)] * 2.0 * M_PI * C->r[j] * p->dr * (ORC_E0);
                    double de = f->rhoUex[IX(g->cell + 1, j)] * 2.0 * M_PI * C->r[j] * p->dr * (-ORC_E0);
                    Ii += di;
                    Ie += de;
                }
            }
            rep->Ii_anode = Ii;
            rep->Ie_anode = Ie;
            rep->I_anode = Ii + Ie;
        }
        /* 4. electron energy */
        if (p->energy_on) {
            memcpy(Te_old, f->Te, C->n * sizeof(double));
            set_eq(p, f, SP_ENERGY, Te_old, &q);
            sg_routine(C, f, SP_ENERGY, &q, s_old, sx_old, sr_old, err3);
            if (rep) {
                rep->error_rhoeps = err3[0];
                rep->error_Gx = err3[1];
                rep->error_Gr = err3[2];
            }
            /* positivity guard, max n_e and Te on the blocks' cells (:4866-4901) */
            for (int bl = 0; bl < G->n_bl; bl++) {
                const int* B = G->cell_box + 4 * bl;
                for (int i = B[0]; i <= B[1]; i++)
                    for (int j = B[2]; j <= B[3]; j++)
                        if (f->rhoeps[IX(i, j)] < 0.0) f->rhoeps[IX(i, j)] = fmax(s_old[IX(i, j)], 0.0);
            }
            double nmax = 0.0;
            for (int bl = 0; bl < G->n_bl; bl++) {
                const int* B = G->cell_box + 4 * bl;
                for (int i = B[0]; i <= B[1]; i++)
                    for (int j = B[2]; j <= B[3]; j++)
                        if (nmax < f->rhoe[IX(i, j)]) nmax = f->rhoe[IX(i, j)];
            }
            const double Te_max = 1000.0 * ORC_E0 / ORC_BOLTZ, ratio = 0.001;
            for (int bl = 0; bl < G->n_bl; bl++) {
                const int* B = G->cell_box + 4 * bl;
                for (int i = B[0]; i <= B[1]; i++)
                    for (int j = B[2]; j <= B[3]; j++)
                        f->Te[IX(i, j)] = fmin(f->rhoeps[IX(i, j)] / (3.0 / 2.0 * (f->rhoe[IX(i, j)] + nmax * ratio) * ORC_BOLTZ + 1e-100), Te_max);
            }
            /* Te ghosts (:4906-4946) */
            ghost_list(C, f->Te, G->te, G->n_te);
        }
        if (rep) rep->steps++;
    }
    free(s_old);
    free(sx_old);
    free(sr_old);
    free(Te_old);
    free(rhoi_old);
    ctx_free(C);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (rep) rep->seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    return 0;
}

int orc_ndd_assemble(const cathode_ndd_params* p, const cathode_ndd_fields* fin, int which,
                     double* aP, double* aE, double* aW, double* aN, double* aS, double* b,
                     double* fx, double* fr)
{
    ctx_t Cx, *C = &Cx;
    ctx_init(C, p, 0, NULL);
    /* work on a shallow copy whose flux arrays are the caller's outputs */
    cathode_ndd_fields f = *fin;
    size_t nb = C->n * sizeof(double);
    double* n_old = (double*)calloc(C->n, sizeof(double));
    int rc = 0;
    if (which == 0) {
        memcpy(fx, fin->nUex, nb);
        memcpy(fr, fin->nUer, nb);
        f.nUex = fx;
        f.nUer = fr;
        memcpy(C->Ex_old, fin->Ex, nb);
        memcpy(C->Er_old, fin->Er, nb);
        memcpy(C->tmp_x, fin->nUex, nb);
        memcpy(C->tmp_r, fin->nUer, nb);
        poisson_predict_assemble(C, &f, C->tmp_x, C->tmp_r);
    } else if (which >= 1 && which <= 3) {
        int sp = which == 1 ? SP_ELEC : which == 2 ? SP_ION : SP_ENERGY;
        sg_eq_t q;
        set_eq(p, &f, sp, fin->Te, &q);
        memcpy(fx, q.Fx, nb);
        memcpy(fr, q.Fr, nb);
        memcpy(C->tmp_x, q.Fx, nb);
        memcpy(C->tmp_r, q.Fr, nb);
        memcpy(n_old, q.n, nb);
        q.Fx = fx;
        q.Fr = fr;
        sg_predict_assemble(C, &f, sp, &q, C->tmp_x, C->tmp_r, n_old);
    } else {
        rc = -1;
    }
    if (!rc) {
        memcpy(aP, C->aP, nb);
        memcpy(aE, C->aE, nb);
        memcpy(aW, C->aW, nb);
        memcpy(aN, C->aN, nb);
        memcpy(aS, C->aS, nb);
        memcpy(b, C->b, nb);
    }
    free(n_old);
    ctx_free(C);
    return rc;
}
