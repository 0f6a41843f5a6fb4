/*
 * orc_mg.c -- host copy of the GPU multigrid algorithm (CPU baseline).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Algorithm (identical to the plain-FP64 GPU solve):
 *   weights    the fine rows are scaled by a radial weight w_j before
 *              aggregation (the conservative r-weighted form: the reference
 *              divides every row by r_j, fluidModule.hpp:2634-2641).  w_j is
 *              inferred from the matrix: w_{j+1}/w_j = sum_i aN(i,j) /
 *              sum_i aS(i,j+1) (= r_{j+1}/r_j for the r-weighted operators),
 *              each ratio clipped to [1e-2, 1e2];
 *   hierarchy  2x2 cell aggregation; the coarse operator is the Galerkin
 *              product with piecewise-constant transfers in which the
 *              couplings along each coarsened direction are halved; each
 *              fine row excess s = aP - sum(a_nb) >= 0 is split into a mass
 *              part m (the minimum excess over the cell and its neighbours)
 *              that is summed and a boundary part s - m that is halved like
 *              a coupling -- exact re-discretization for constant-coefficient
 *              diffusion + mass with Dirichlet walls, always 5-point;
 *   smoother   red-black point Gauss-Seidel, nu1 = nu2 = 2 sweeps;
 *   transfers  restriction = w-weighted sum of the children's residuals,
 *              prolongation = piecewise constant;
 *   coarsest   the first level of <= 32 cells, solved by dense LU with
 *              partial pivoting (the device inverts it by Gauss-Jordan
 *              elimination without pivoting: the same solution to rounding);
 *   cut        a mass-dominated operator stops its V-cycles early: with the
 *              row excess fraction delta_l = min over active cells of
 *              (aP - sum a_nb) / aP, level 0 when delta_0 >= 0.9, else the
 *              device's second candidate level (the last level of its first
 *              fine launch unit below level 0) when its delta >= 0.35.  The
 *              cut level is smoothed (2 + 2 sweeps from x = 0) with no
 *              coarse-grid correction.  The device decides from the same
 *              solve one time step earlier (its launches are enqueued before
 *              the setup's fractions reach the host); this copy from the
 *              solve's own operator;
 *   outer      BiCGStab right-preconditioned by one V-cycle per application,
 *              from x = 0 (solverFuncs.hpp:106-109), stopping when the true
 *              residual ||b-Ax||_2/||b||_2 <= tol or after maxITR iterations.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define MG_MAX_LEVELS 24
static int MG_COARSE_MAX = 32;

typedef struct {
    int ni, nj, ld;
    int cx, cy; /* coarsen to the next level in i / j */
    double *aP, *aE, *aW, *aN, *aS, *x, *b, *r;
    double* m;  /* mass part of the (weighted) row excess */
    double* w;  /* radial row weight per column j (level 0), 1 on coarse levels */
    unsigned char* act;
} level_t;

static int g_threads = 1;
void orc_mg_set_threads(int n) { g_threads = n > 0 ? n : 1; }
static int g_cut_on = 1;
static int g_last_cut = -1;
void orc_mg_set_cut(int on) { g_cut_on = on != 0; }
int orc_mg_last_cut(void) { return g_last_cut; }
void orc_mg_set_coarse_max(int n) { MG_COARSE_MAX = n > 0 ? n : 1; }

#define L_AT(L, a, i, j) ((L)->a[(size_t)(i) * (L)->ld + (j)])

static void level_alloc(level_t* L, int ni, int nj)
{
    L->ni = ni;
    L->nj = nj;
    L->ld = nj + 2;
    size_t n = (size_t)(ni + 2) * (nj + 2);
    double** b[] = {&L->aP, &L->aE, &L->aW, &L->aN, &L->aS, &L->x, &L->b, &L->r, &L->m};
    for (int k = 0; k < 9; k++) *b[k] = (double*)calloc(n, sizeof(double));
    L->w = (double*)malloc((nj + 2) * sizeof(double));
    for (int j = 0; j < nj + 2; j++) L->w[j] = 1.0;
    L->act = (unsigned char*)calloc(n, 1);
}

static void level_free(level_t* L)
{
    double* b[] = {L->aP, L->aE, L->aW, L->aN, L->aS, L->x, L->b, L->r, L->m};
    for (int k = 0; k < 9; k++) free(b[k]);
    free(L->w);
    free(L->act);
}

static void coarsen(const level_t* F, level_t* Cl)
{
    int fi = F->cx ? 2 : 1, fj = F->cy ? 2 : 1;
    int ni = F->cx ? (F->ni + 1) / 2 : F->ni;
    int nj = F->cy ? (F->nj + 1) / 2 : F->nj;
    level_alloc(Cl, ni, nj);
    double hx = F->cx ? 0.5 : 1.0, hy = F->cy ? 0.5 : 1.0;
    double hb = (F->cx && F->cy) ? 0.5 : ((F->cx || F->cy) ? 0.75 : 1.0);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int I = 1; I <= ni; I++)
        for (int J = 1; J <= nj; J++) {
            int i0 = fi * (I - 1) + 1, i1 = fi * I < F->ni ? fi * I : F->ni;
            int j0 = fj * (J - 1) + 1, j1 = fj * J < F->nj ? fj * J : F->nj;
            double S = 0.0, M = 0.0, E = 0.0, W = 0.0, N = 0.0, So = 0.0;
            int any = 0;
            for (int i = i0; i <= i1; i++)
                for (int j = j0; j <= j1; j++) {
                    if (!L_AT(F, act, i, j)) continue;
                    any = 1;
                    double w = F->w[j];
                    double aE = L_AT(F, aE, i, j), aW = L_AT(F, aW, i, j);
                    double aN = L_AT(F, aN, i, j), aS = L_AT(F, aS, i, j);
                    double s = L_AT(F, aP, i, j) - (aE + aW + aN + aS);
                    S += w * (s > 0.0 ? s : 0.0);
                    M += L_AT(F, m, i, j);
                    if (i == i1) E += w * aE;
                    if (i == i0) W += w * aW;
                    if (j == j1) N += w * aN;
                    if (j == j0) So += w * aS;
                }
            L_AT(Cl, act, I, J) = (unsigned char)any;
            if (!any) continue;
            E *= hx;
            W *= hx;
            N *= hy;
            So *= hy;
            double ex = M + hb * (S - M);
            L_AT(Cl, aE, I, J) = E;
            L_AT(Cl, aW, I, J) = W;
            L_AT(Cl, aN, I, J) = N;
            L_AT(Cl, aS, I, J) = So;
            L_AT(Cl, m, I, J) = M;
            L_AT(Cl, aP, I, J) = ex + (E + W + N + So);
        }
}

/* level-0 radial weights and mass estimate */
static void fine_setup(level_t* L)
{
    for (int j = 0; j < L->nj + 2; j++) L->w[j] = 1.0;
    for (int j = 1; j < L->nj; j++) {
        double sN = 0.0, sS = 0.0;
        for (int i = 1; i <= L->ni; i++) {
            sN += L_AT(L, aN, i, j);
            sS += L_AT(L, aS, i, j + 1);
        }
        double rho = (sN > 0.0 && sS > 0.0) ? sN / sS : 1.0;
        rho = rho < 1e-2 ? 1e-2 : (rho > 1e2 ? 1e2 : rho);
        L->w[j + 1] = L->w[j] * rho;
    }
    size_t n = (size_t)(L->ni + 2) * L->ld;
    double* s = (double*)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; k++) s[k] = INFINITY;
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            if (!L_AT(L, act, i, j)) continue;
            double v = L_AT(L, aP, i, j) - (L_AT(L, aE, i, j) + L_AT(L, aW, i, j) + L_AT(L, aN, i, j) + L_AT(L, aS, i, j));
            s[(size_t)i * L->ld + j] = v > 0.0 ? v : 0.0;
        }
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            if (!L_AT(L, act, i, j)) continue;
            size_t k = (size_t)i * L->ld + j;
            double v = s[k];
            if (s[k + L->ld] < v) v = s[k + L->ld];
            if (s[k - L->ld] < v) v = s[k - L->ld];
            if (s[k + 1] < v) v = s[k + 1];
            if (s[k - 1] < v) v = s[k - 1];
            L_AT(L, m, i, j) = L->w[j] * v;
        }
    free(s);
}

static void smooth(level_t* L, int color)
{
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 1; i <= L->ni; i++) {
        int j0 = 1 + ((i + 1 + color) & 1);
        for (int j = j0; j <= L->nj; j += 2) {
            if (!L_AT(L, act, i, j)) continue;
            double s = L_AT(L, b, i, j) + L_AT(L, aE, i, j) * L_AT(L, x, i + 1, j) + L_AT(L, aW, i, j) * L_AT(L, x, i - 1, j) +
                       L_AT(L, aN, i, j) * L_AT(L, x, i, j + 1) + L_AT(L, aS, i, j) * L_AT(L, x, i, j - 1);
            L_AT(L, x, i, j) = s / L_AT(L, aP, i, j);
        }
    }
}

static void residual(level_t* L)
{
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            if (!L_AT(L, act, i, j)) {
                L_AT(L, r, i, j) = 0.0;
                continue;
            }
            L_AT(L, r, i, j) = L_AT(L, b, i, j) - L_AT(L, aP, i, j) * L_AT(L, x, i, j) + L_AT(L, aE, i, j) * L_AT(L, x, i + 1, j) +
                               L_AT(L, aW, i, j) * L_AT(L, x, i - 1, j) + L_AT(L, aN, i, j) * L_AT(L, x, i, j + 1) +
                               L_AT(L, aS, i, j) * L_AT(L, x, i, j - 1);
        }
}

static void restrict_res(const level_t* F, level_t* Cl)
{
    int fi = F->cx ? 2 : 1, fj = F->cy ? 2 : 1;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int I = 1; I <= Cl->ni; I++)
        for (int J = 1; J <= Cl->nj; J++) {
            int i0 = fi * (I - 1) + 1, i1 = fi * I < F->ni ? fi * I : F->ni;
            int j0 = fj * (J - 1) + 1, j1 = fj * J < F->nj ? fj * J : F->nj;
            double s = 0.0;
            for (int i = i0; i <= i1; i++)
                for (int j = j0; j <= j1; j++) s += F->w[j] * L_AT(F, r, i, j);
            L_AT(Cl, b, I, J) = s;
            L_AT(Cl, x, I, J) = 0.0;
        }
}

static void prolong_add(level_t* F, const level_t* Cl)
{
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 1; i <= F->ni; i++)
        for (int j = 1; j <= F->nj; j++) {
            if (!L_AT(F, act, i, j)) continue;
            int I = F->cx ? (i + 1) / 2 : i, J = F->cy ? (j + 1) / 2 : j;
            L_AT(F, x, i, j) += L_AT(Cl, x, I, J);
        }
}

/* dense LU of the coarsest level (partial pivoting) */
typedef struct {
    int n;
    int* idx; /* cell -> row */
    double* lu;
    int* piv;
} dense_t;

static void dense_setup(const level_t* L, dense_t* D)
{
    size_t nc = (size_t)(L->ni + 2) * L->ld;
    D->idx = (int*)malloc(nc * sizeof(int));
    int n = 0;
    for (size_t k = 0; k < nc; k++) D->idx[k] = L->act[k] ? n++ : -1;
    D->n = n;
    D->lu = (double*)calloc((size_t)n * n + 1, sizeof(double));
    D->piv = (int*)calloc(n + 1, sizeof(int));
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            int k = D->idx[i * L->ld + j];
            if (k < 0) continue;
            D->lu[k * n + k] = L_AT(L, aP, i, j);
            int nb[4] = {D->idx[(i + 1) * L->ld + j], D->idx[(i - 1) * L->ld + j], D->idx[i * L->ld + j + 1], D->idx[i * L->ld + j - 1]};
            double a[4] = {L_AT(L, aE, i, j), L_AT(L, aW, i, j), L_AT(L, aN, i, j), L_AT(L, aS, i, j)};
            for (int q = 0; q < 4; q++)
                if (nb[q] >= 0) D->lu[k * n + nb[q]] -= a[q];
        }
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(D->lu[i * n + k]) > fabs(D->lu[p * n + k])) p = i;
        D->piv[k] = p;
        if (p != k)
            for (int c = 0; c < n; c++) {
                double t = D->lu[k * n + c];
                D->lu[k * n + c] = D->lu[p * n + c];
                D->lu[p * n + c] = t;
            }
        double d = D->lu[k * n + k];
        if (d == 0.0) d = D->lu[k * n + k] = 1e-300;
        for (int i = k + 1; i < n; i++) {
            double l = D->lu[i * n + k] / d;
            D->lu[i * n + k] = l;
            for (int c = k + 1; c < n; c++) D->lu[i * n + c] -= l * D->lu[k * n + c];
        }
    }
}

static void dense_solve(level_t* L, const dense_t* D)
{
    int n = D->n;
    double* v = (double*)calloc(n + 1, sizeof(double));
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            int k = D->idx[i * L->ld + j];
            if (k >= 0) v[k] = L_AT(L, b, i, j);
        }
    for (int k = 0; k < n; k++) {
        int p = D->piv[k];
        if (p != k) {
            double t = v[k];
            v[k] = v[p];
            v[p] = t;
        }
        for (int i = k + 1; i < n; i++) v[i] -= D->lu[i * n + k] * v[k];
    }
    for (int k = n - 1; k >= 0; k--) {
        double s = v[k];
        for (int c = k + 1; c < n; c++) s -= D->lu[k * n + c] * v[c];
        v[k] = s / D->lu[k * n + k];
    }
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            int k = D->idx[i * L->ld + j];
            if (k >= 0) L_AT(L, x, i, j) = v[k];
        }
    free(v);
}

static void vcycle(level_t* Ls, int l, int nl, const dense_t* D, int cut)
{
    level_t* L = &Ls[l];
    if (l == nl - 1) {
        dense_solve(L, D);
        return;
    }
    if (l == cut) {
        /* the last level visited: pre- and post-smoothing, no correction */
        for (int s = 0; s < 4; s++) {
            smooth(L, 0);
            smooth(L, 1);
        }
        return;
    }
    for (int s = 0; s < 2; s++) {
        smooth(L, 0);
        smooth(L, 1);
    }
    residual(L);
    restrict_res(L, &Ls[l + 1]);
    vcycle(Ls, l + 1, nl, D, cut);
    prolong_add(L, &Ls[l + 1]);
    for (int s = 0; s < 2; s++) {
        smooth(L, 0);
        smooth(L, 1);
    }
}

static double dotp(const double* a, const double* b, size_t n)
{
    double s = 0.0;
    for (size_t k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}

/* out = A in on the active cells of level 0 (inactive: 0) */
static void apply_A(level_t* L, const double* in, double* out)
{
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < L->ni + 2; i++)
        for (int j = 0; j < L->ld; j++) {
            size_t k = (size_t)i * L->ld + j;
            if (i < 1 || i > L->ni || j < 1 || j > L->nj || !L->act[k]) {
                out[k] = 0.0;
                continue;
            }
            out[k] = L->aP[k] * in[k] - L->aE[k] * in[k + L->ld] - L->aW[k] * in[k - L->ld] -
                     L->aN[k] * in[k + 1] - L->aS[k] * in[k - 1];
        }
}

/* out = V-cycle(in) from a zero initial guess */
static void precond(level_t* Ls, int nl, const dense_t* D, int cut, const double* in, double* out)
{
    level_t* L0 = &Ls[0];
    size_t n0 = (size_t)(L0->ni + 2) * L0->ld;
    for (size_t k = 0; k < n0; k++) L0->b[k] = L0->act[k] ? in[k] : 0.0;
    memset(L0->x, 0, n0 * sizeof(double));
    vcycle(Ls, 0, nl, D, cut);
    memcpy(out, L0->x, n0 * sizeof(double));
}

/* minimum row excess fraction (aP - sum a_nb) / aP over the active cells */
static double excess_fraction(const level_t* L)
{
    double d = INFINITY;
    for (int i = 1; i <= L->ni; i++)
        for (int j = 1; j <= L->nj; j++) {
            if (!L_AT(L, act, i, j)) continue;
            double a = L_AT(L, aP, i, j);
            double e = a - (L_AT(L, aE, i, j) + L_AT(L, aW, i, j) + L_AT(L, aN, i, j) + L_AT(L, aS, i, j));
            double f = a > 0.0 ? (e > 0.0 ? e : 0.0) / a : 0.0;
            if (f < d) d = f;
        }
    return d;
}

/* The V-cycle cut (-1: none) with the device's candidate levels: level 0
 * when the finest level runs in the fine-level kernels (the one-workgroup
 * tail takes the levels from lt on: a first level of <= 2048 cells and at
 * most four further ones of <= 1024), and tcut, the last level of the first
 * launch unit below level 0 (levels 1 and 2 together when level 1 coarsens
 * in both directions and 2 < lt, else level 1), when 1 <= tcut < lt. */
static int mg_cut(const level_t* Ls, int nl)
{
    int lt = nl - 1;
    while (lt > 0 && (long)Ls[lt - 1].ni * Ls[lt - 1].nj <= 2048 && (long)Ls[lt].ni * Ls[lt].nj <= 1024 &&
           nl - (lt - 1) <= 5)
        lt--;
    if (lt <= 0) return -1;
    if (excess_fraction(&Ls[0]) >= 0.9) return 0;
    if (lt <= 1) return -1;
    int tcut = (2 < lt && Ls[1].cx && Ls[1].cy) ? 2 : 1;
    return excess_fraction(&Ls[tcut]) >= 0.35 ? tcut : -1;
}

int orc_mg_solve(const double* aP, const double* aE, const double* aW, const double* aN,
                 const double* aS, const double* b, const int* boxes, int n_bl, int ni2,
                 int nj2, int ld, int maxITR, double tol, double* x, double* relres)
{
    level_t Ls[MG_MAX_LEVELS];
    memset(Ls, 0, sizeof(Ls));
    level_t* L0 = &Ls[0];
    level_alloc(L0, ni2 - 2, nj2 - 2);
    for (int bl = 0; bl < n_bl; bl++) {
        const int* bx = boxes + 4 * bl;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++) L_AT(L0, act, i, j) = 1;
    }
    double bb = 0.0;
    for (int i = 1; i <= L0->ni; i++)
        for (int j = 1; j <= L0->nj; j++) {
            if (!L_AT(L0, act, i, j)) continue;
            size_t k = (size_t)i * ld + j;
            L_AT(L0, aP, i, j) = aP[k];
            L_AT(L0, aE, i, j) = L_AT(L0, act, i + 1, j) ? aE[k] : 0.0;
            L_AT(L0, aW, i, j) = L_AT(L0, act, i - 1, j) ? aW[k] : 0.0;
            L_AT(L0, aN, i, j) = L_AT(L0, act, i, j + 1) ? aN[k] : 0.0;
            L_AT(L0, aS, i, j) = L_AT(L0, act, i, j - 1) ? aS[k] : 0.0;
            L_AT(L0, b, i, j) = b[k];
            bb += b[k] * b[k];
        }
    fine_setup(L0);
    int nl = 1;
    while (nl < MG_MAX_LEVELS) {
        level_t* F = &Ls[nl - 1];
        if ((long)F->ni * F->nj <= MG_COARSE_MAX || (F->ni == 1 && F->nj == 1)) break;
        F->cx = F->ni > 1;
        F->cy = F->nj > 1;
        coarsen(F, &Ls[nl]);
        nl++;
    }
    dense_t D;
    dense_setup(&Ls[nl - 1], &D);
    const int cut = g_cut_on ? mg_cut(Ls, nl) : -1;
    g_last_cut = cut;
    double bnorm = sqrt(bb);
    int it = 0;
    double res = 0.0;
    const size_t n0 = (size_t)(L0->ni + 2) * L0->ld;
    /* BiCGStab vectors; x = 0 on the box union (the reference's zero guess) */
    double* xs = (double*)calloc(n0, sizeof(double));
    double* bs = (double*)malloc(n0 * sizeof(double));
    double* vec = (double*)calloc(8 * n0, sizeof(double));
    double *r = vec, *r0 = vec + n0, *p = vec + 2 * n0, *v = vec + 3 * n0, *sv = vec + 4 * n0,
           *t = vec + 5 * n0, *ph = vec + 6 * n0, *sh = vec + 7 * n0;
    memcpy(bs, L0->b, n0 * sizeof(double));
    double rho = 1.0, rho1 = 0.0, al = 1.0, om = 1.0;
    int restart = 1;
    if (bnorm > 0.0) {
        while (it < maxITR) {
            if (restart) {
                memcpy(L0->x, xs, n0 * sizeof(double));
                memcpy(L0->b, bs, n0 * sizeof(double));
                residual(L0);
                memcpy(r, L0->r, n0 * sizeof(double));
                memcpy(r0, r, n0 * sizeof(double));
                memset(p, 0, n0 * sizeof(double));
                memset(v, 0, n0 * sizeof(double));
                rho1 = dotp(r0, r, n0);
                rho = al = om = 1.0;
            }
            it++;
            if (restart) {
                memcpy(p, r, n0 * sizeof(double));
            } else {
                double beta = (rho1 / rho) * (al / om);
                for (size_t k = 0; k < n0; k++) p[k] = r[k] + beta * (p[k] - om * v[k]);
            }
            restart = 0;
            precond(Ls, nl, &D, cut, p, ph);
            apply_A(L0, ph, v);
            al = rho1 / dotp(r0, v, n0);
            for (size_t k = 0; k < n0; k++) sv[k] = r[k] - al * v[k];
            precond(Ls, nl, &D, cut, sv, sh);
            apply_A(L0, sh, t);
            double tt = dotp(t, t, n0);
            om = tt > 0.0 ? dotp(t, sv, n0) / tt : 0.0;
            for (size_t k = 0; k < n0; k++) {
                xs[k] += al * ph[k] + om * sh[k];
                r[k] = sv[k] - om * t[k];
            }
            rho = rho1;
            rho1 = dotp(r0, r, n0);
            res = sqrt(dotp(r, r, n0)) / bnorm;
            if (!isfinite(res)) break;
            if (res <= tol) {
                memcpy(L0->x, xs, n0 * sizeof(double));
                memcpy(L0->b, bs, n0 * sizeof(double));
                residual(L0);
                res = sqrt(dotp(L0->r, L0->r, n0)) / bnorm;
                if (res <= tol || !isfinite(res)) break;
                restart = 1;
            }
        }
    }
    for (int i = 1; i <= L0->ni; i++)
        for (int j = 1; j <= L0->nj; j++)
            if (L_AT(L0, act, i, j)) x[(size_t)i * ld + j] = xs[(size_t)i * L0->ld + j];
    if (relres) *relres = res;
    free(xs);
    free(bs);
    free(vec);
    free(D.idx);
    free(D.lu);
    free(D.piv);
    for (int l = 0; l < nl; l++) level_free(&Ls[l]);
    return it;
}
