/*
 * orc_solve.c -- exact and reference-semantics linear solves (oracle side).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference hands each 5-point system to HYPRE SMG
 * (include/solverFuncs.hpp:26-162) and the coupled microwave system to Eigen
 * SparseLU (include/emfieldModule.hpp:2994-2998).  Neither library is
 * available, so the oracle solves the same system exactly with a banded LU in
 * natural (i-major, j-minor) order over the bounding rectangle of the box
 * union.  Cells of the rectangle that are outside every box get identity rows
 * (x = 0), which is HYPRE's off-grid convention of dropping couplings to cells
 * that are not part of the grid.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define AT(a, i, j) ((a)[(size_t)(i) * (size_t)ld + (size_t)(j)])

typedef struct {
    int i0, i1, j0, j1; /* bounding rectangle (inclusive) */
    int nr, nc;         /* rows / columns of the rectangle */
    unsigned char* in;  /* in-box flag per rectangle cell */
} box_map;

static int make_map(const int* boxes, int n_bl, int ni2, int nj2, box_map* m)
{
    if (n_bl <= 0) return -1;
    m->i0 = m->j0 = 1 << 30;
    m->i1 = m->j1 = -1;
    for (int b = 0; b < n_bl; b++) {
        const int* bx = boxes + 4 * b;
        if (bx[0] < 1 || bx[1] > ni2 - 2 || bx[2] < 1 || bx[3] > nj2 - 2) return -1;
        if (bx[0] > bx[1] || bx[2] > bx[3]) continue;
        if (bx[0] < m->i0) m->i0 = bx[0];
        if (bx[1] > m->i1) m->i1 = bx[1];
        if (bx[2] < m->j0) m->j0 = bx[2];
        if (bx[3] > m->j1) m->j1 = bx[3];
    }
    if (m->i1 < 0) return -1;
    m->nr = m->i1 - m->i0 + 1;
    m->nc = m->j1 - m->j0 + 1;
    m->in = (unsigned char*)calloc((size_t)m->nr * m->nc, 1);
    for (int b = 0; b < n_bl; b++) {
        const int* bx = boxes + 4 * b;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++)
                m->in[(size_t)(i - m->i0) * m->nc + (j - m->j0)] = 1;
    }
    return 0;
}

static inline int map_in(const box_map* m, int i, int j)
{
    if (i < m->i0 || i > m->i1 || j < m->j0 || j > m->j1) return 0;
    return m->in[(size_t)(i - m->i0) * m->nc + (j - m->j0)];
}

/* ---------------------------------------------------------------- */
/* real banded LU without pivoting                                    */
/* ---------------------------------------------------------------- */
int orc_struct5_direct(const double* aP, const double* aE, const double* aW,
                       const double* aN, const double* aS, const double* b,
                       const int* boxes, int n_bl, int ni2, int nj2, int ld,
                       double* x)
{
    box_map m;
    if (make_map(boxes, n_bl, ni2, nj2, &m)) return -2;
    const long N = (long)m.nr * m.nc;
    const int kl = m.nc, ku = m.nc, W = kl + ku + 1;
    double* A = (double*)calloc((size_t)N * W, sizeof(double));
    double* r = (double*)calloc((size_t)N, sizeof(double));
#define BA(row, col) A[(size_t)(row) * W + ((col) - (row) + kl)]
    for (int i = m.i0; i <= m.i1; i++)
        for (int j = m.j0; j <= m.j1; j++) {
            long k = (long)(i - m.i0) * m.nc + (j - m.j0);
            if (!map_in(&m, i, j)) {
                BA(k, k) = 1.0;
                r[k] = 0.0;
                continue;
            }
            BA(k, k) = AT(aP, i, j);
            r[k] = AT(b, i, j);
            if (map_in(&m, i + 1, j)) BA(k, k + m.nc) = -AT(aE, i, j);
            if (map_in(&m, i - 1, j)) BA(k, k - m.nc) = -AT(aW, i, j);
            if (map_in(&m, i, j + 1)) BA(k, k + 1) = -AT(aN, i, j);
            if (map_in(&m, i, j - 1)) BA(k, k - 1) = -AT(aS, i, j);
        }
    int rc = 0;
    for (long k = 0; k < N && !rc; k++) {
        double piv = BA(k, k);
        if (piv == 0.0) {
            rc = -1;
            break;
        }
        long iend = k + kl < N - 1 ? k + kl : N - 1;
        long jend = k + ku < N - 1 ? k + ku : N - 1;
        for (long i = k + 1; i <= iend; i++) {
            double l = BA(i, k);
            if (l == 0.0) continue;
            l /= piv;
            BA(i, k) = 0.0;
            double* ri = A + (size_t)i * W + (kl - i); /* ri[c] == BA(i, c) */
            const double* rk = A + (size_t)k * W + (kl - k);
            for (long c = k + 1; c <= jend; c++) ri[c] -= l * rk[c];
            r[i] -= l * r[k];
        }
    }
    if (!rc) {
        for (long k = N - 1; k >= 0; k--) {
            double s = r[k];
            long jend = k + ku < N - 1 ? k + ku : N - 1;
            for (long c = k + 1; c <= jend; c++) s -= BA(k, c) * r[c];
            r[k] = s / BA(k, k);
        }
        for (int i = m.i0; i <= m.i1; i++)
            for (int j = m.j0; j <= m.j1; j++)
                if (map_in(&m, i, j)) AT(x, i, j) = r[(long)(i - m.i0) * m.nc + (j - m.j0)];
    }
#undef BA
    free(A);
    free(r);
    free(m.in);
    return rc;
}

/* ---------------------------------------------------------------- */
/* complex banded LU with partial pivoting (general band)             */
/* ---------------------------------------------------------------- */
typedef double complex cplx;

/* Solves the banded system whose rows are filled by `fill`; kl = ku = half
 * bandwidth of the unpivoted matrix.  Storage width 2*kl+ku+1 leaves room
 * for the fill that row exchanges create. */
static int band_lu_pivot(long N, int kl, int ku, cplx* A, cplx* r)
{
    const int W = 2 * kl + ku + 1;
#define CB(row, col) A[(size_t)(row) * W + ((col) - (row) + kl)]
    for (long k = 0; k < N; k++) {
        long iend = k + kl < N - 1 ? k + kl : N - 1;
        long p = k;
        double best = cabs(CB(k, k));
        for (long i = k + 1; i <= iend; i++) {
            double v = cabs(CB(i, k));
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (best == 0.0) return -1;
        long cend = k + kl + ku < N - 1 ? k + kl + ku : N - 1;
        if (p != k) {
            for (long c = k; c <= cend; c++) {
                cplx t = CB(k, c);
                CB(k, c) = CB(p, c);
                CB(p, c) = t;
            }
            cplx t = r[k];
            r[k] = r[p];
            r[p] = t;
        }
        cplx piv = CB(k, k);
        for (long i = k + 1; i <= iend; i++) {
            cplx l = CB(i, k);
            if (l == 0.0) continue;
            l /= piv;
            CB(i, k) = 0.0;
            for (long c = k + 1; c <= cend; c++) CB(i, c) -= l * CB(k, c);
            r[i] -= l * r[k];
        }
    }
    for (long k = N - 1; k >= 0; k--) {
        cplx s = r[k];
        long cend = k + kl + ku < N - 1 ? k + kl + ku : N - 1;
        for (long c = k + 1; c <= cend; c++) s -= CB(k, c) * r[c];
        r[k] = s / CB(k, k);
    }
#undef CB
    return 0;
}

int orc_helm5_direct(const double* aP, const double* aE, const double* aW,
                     const double* aN, const double* aS, const double* b,
                     const int* boxes, int n_bl, int ni2, int nj2, int ld,
                     double* x)
{
    box_map m;
    if (make_map(boxes, n_bl, ni2, nj2, &m)) return -2;
    const long N = (long)m.nr * m.nc;
    const int kl = m.nc, ku = m.nc, W = 2 * kl + ku + 1;
    cplx* A = (cplx*)calloc((size_t)N * W, sizeof(cplx));
    cplx* r = (cplx*)calloc((size_t)N, sizeof(cplx));
#define CB(row, col) A[(size_t)(row) * W + ((col) - (row) + kl)]
#define CV(a, i, j) ((a)[2 * ((size_t)(i) * ld + (j))] + I * (a)[2 * ((size_t)(i) * ld + (j)) + 1])
    for (int i = m.i0; i <= m.i1; i++)
        for (int j = m.j0; j <= m.j1; j++) {
            long k = (long)(i - m.i0) * m.nc + (j - m.j0);
            if (!map_in(&m, i, j)) {
                CB(k, k) = 1.0;
                continue;
            }
            CB(k, k) = CV(aP, i, j);
            r[k] = CV(b, i, j);
            if (map_in(&m, i + 1, j)) CB(k, k + m.nc) = -CV(aE, i, j);
            if (map_in(&m, i - 1, j)) CB(k, k - m.nc) = -CV(aW, i, j);
            if (map_in(&m, i, j + 1)) CB(k, k + 1) = -CV(aN, i, j);
            if (map_in(&m, i, j - 1)) CB(k, k - 1) = -CV(aS, i, j);
        }
    int rc = band_lu_pivot(N, kl, ku, A, r);
    if (!rc)
        for (int i = m.i0; i <= m.i1; i++)
            for (int j = m.j0; j <= m.j1; j++)
                if (map_in(&m, i, j)) {
                    cplx v = r[(long)(i - m.i0) * m.nc + (j - m.j0)];
                    x[2 * ((size_t)i * ld + j)] = creal(v);
                    x[2 * ((size_t)i * ld + j) + 1] = cimag(v);
                }
#undef CV
#undef CB
    free(A);
    free(r);
    free(m.in);
    return rc;
}

int orc_lu9_direct(const double* aP, const double* aE, const double* aW,
                   const double* aN, const double* aS, const double* aEN,
                   const double* aES, const double* aWN, const double* aWS,
                   const double* b, const int* boxes, int n_bl, int ni2,
                   int nj2, int ld, double* x)
{
    box_map m;
    if (make_map(boxes, n_bl, ni2, nj2, &m)) return -2;
    const long N = (long)m.nr * m.nc;
    const int kl = m.nc + 1, ku = m.nc + 1, W = 2 * kl + ku + 1;
    cplx* A = (cplx*)calloc((size_t)N * W, sizeof(cplx));
    cplx* r = (cplx*)calloc((size_t)N, sizeof(cplx));
#define CB(row, col) A[(size_t)(row) * W + ((col) - (row) + kl)]
    for (int i = m.i0; i <= m.i1; i++)
        for (int j = m.j0; j <= m.j1; j++) {
            long k = (long)(i - m.i0) * m.nc + (j - m.j0);
            if (!map_in(&m, i, j)) {
                CB(k, k) = 1.0;
                continue;
            }
            CB(k, k) = AT(aP, i, j);
            r[k] = AT(b, i, j);
            if (map_in(&m, i + 1, j)) CB(k, k + m.nc) = -AT(aE, i, j);
            if (map_in(&m, i - 1, j)) CB(k, k - m.nc) = -AT(aW, i, j);
            if (map_in(&m, i, j + 1)) CB(k, k + 1) = -AT(aN, i, j);
            if (map_in(&m, i, j - 1)) CB(k, k - 1) = -AT(aS, i, j);
            if (map_in(&m, i + 1, j + 1)) CB(k, k + m.nc + 1) = -AT(aEN, i, j);
            if (map_in(&m, i + 1, j - 1)) CB(k, k + m.nc - 1) = -AT(aES, i, j);
            if (map_in(&m, i - 1, j + 1)) CB(k, k - m.nc + 1) = -AT(aWN, i, j);
            if (map_in(&m, i - 1, j - 1)) CB(k, k - m.nc - 1) = -AT(aWS, i, j);
        }
    int rc = band_lu_pivot(N, kl, ku, A, r);
    if (!rc)
        for (int i = m.i0; i <= m.i1; i++)
            for (int j = m.j0; j <= m.j1; j++)
                if (map_in(&m, i, j)) AT(x, i, j) = creal(r[(long)(i - m.i0) * m.nc + (j - m.j0)]);
#undef CB
    free(A);
    free(r);
    free(m.in);
    return rc;
}

/* ---------------------------------------------------------------- */
/* point SOR, include/solverFuncs.hpp:169-216                         */
/* ---------------------------------------------------------------- */
int orc_sor_solve(const double* aP, const double* aE, const double* aW,
                  const double* aN, const double* aS, const double* b,
                  const int* boxes, int n_bl, int ni2, int nj2, int ld,
                  double alpha, int maxITR, double tol, double* x,
                  double* err_out)
{
    (void)ni2;
    (void)nj2;
    int ncount = 0;
    double error = 0.0;
    do {
        ncount = ncount + 1;
        error = 0.0;
        for (int bl = 0; bl < n_bl; bl++) {
            const int* bx = boxes + 4 * bl;
            for (int i = bx[0]; i <= bx[1]; i++)
                for (int j = bx[2]; j <= bx[3]; j++) {
                    double t = (AT(aE, i, j) * AT(x, i + 1, j) + AT(aW, i, j) * AT(x, i - 1, j) +
                                AT(aN, i, j) * AT(x, i, j + 1) + AT(aS, i, j) * AT(x, i, j - 1) +
                                AT(b, i, j)) /
                               (AT(aP, i, j) + 1e-100);
                    double xn = (1.0 - alpha) * AT(x, i, j) + alpha * t;
                    double e = fabs((xn - AT(x, i, j)) / (xn + AT(x, i, j) + 1e-6)) * 2.0;
                    if (e > error) error = e;
                    AT(x, i, j) = xn;
                }
        }
        if (ncount > maxITR) break;
    } while (error > tol);
    if (err_out) *err_out = error;
    return ncount;
}

double orc_rel_residual(const double* aP, const double* aE, const double* aW,
                        const double* aN, const double* aS, const double* b,
                        const int* boxes, int n_bl, int ni2, int nj2, int ld,
                        const double* x)
{
    box_map m;
    if (make_map(boxes, n_bl, ni2, nj2, &m)) return -1.0;
    double rr = 0.0, bb = 0.0;
    for (int i = m.i0; i <= m.i1; i++)
        for (int j = m.j0; j <= m.j1; j++) {
            if (!map_in(&m, i, j)) continue;
            double s = AT(b, i, j) - AT(aP, i, j) * AT(x, i, j);
            if (map_in(&m, i + 1, j)) s += AT(aE, i, j) * AT(x, i + 1, j);
            if (map_in(&m, i - 1, j)) s += AT(aW, i, j) * AT(x, i - 1, j);
            if (map_in(&m, i, j + 1)) s += AT(aN, i, j) * AT(x, i, j + 1);
            if (map_in(&m, i, j - 1)) s += AT(aS, i, j) * AT(x, i, j - 1);
            rr += s * s;
            bb += AT(b, i, j) * AT(b, i, j);
        }
    free(m.in);
    if (bb == 0.0) return rr == 0.0 ? 0.0 : INFINITY;
    return sqrt(rr / bb);
}
