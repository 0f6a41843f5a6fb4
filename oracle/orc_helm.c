/*
 * orc_helm.c -- E_theta microwave system, absorbed power and ABCD algebra.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * E_theta block of EmfieldModule::solve_Microwave with B = 0
 * (include/emfieldModule.hpp:2414-2454): the Hall couplings to E_x / E_r
 * vanish and the E_theta rows decouple into a complex 5-point system.
 */
#include <complex.h>
#include <math.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

#define CI(a, i, j) (2 * ((size_t)(i) * ld + (size_t)(j)))
static inline void cset(double* a, size_t k, cplx v)
{
    a[k] = creal(v);
    a[k + 1] = cimag(v);
}
static inline cplx cget(const double* a, size_t k) { return a[k] + I * a[k + 1]; }

/* sigma_theta-theta, emfieldModule.hpp:2426-2432 (Hall terms zero) */
static cplx sigma_pp(const orc_helm_params* p, double ne)
{
    const cplx EPS = 1e-100 + 1e-100 * I;
    cplx nu_cmp = p->nu + I * p->omega;
    cplx coef = ne * ORC_E0 * ORC_E0 / (p->masse * nu_cmp + EPS) / (1.0 + 0.0 + EPS);
    return coef * 1.0;
}

int orc_helm_assemble(const orc_helm_params* p, const double* ne, const double* J, double* aP,
                      double* aE, double* aW, double* aN, double* aS, double* b)
{
    const int ni = p->ni, nj = p->nj, ld = nj + 2;
    const double dx = p->dx, dr = p->dr, w = p->omega;
    double r[nj + 2];
    for (int j = 0; j < nj + 2; j++) r[j] = p->rmin + (j - 0.5) * dr;
    size_t n2 = 2 * (size_t)(ni + 2) * ld;
    memset(aP, 0, n2 * sizeof(double));
    memset(aE, 0, n2 * sizeof(double));
    memset(aW, 0, n2 * sizeof(double));
    memset(aN, 0, n2 * sizeof(double));
    memset(aS, 0, n2 * sizeof(double));
    memset(b, 0, n2 * sizeof(double));
    const cplx iwmu0 = I * (w * ORC_MU0);
    for (int i = 1; i <= ni; i++)
        for (int j = 1; j <= nj; j++) {
            double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100);
            double qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
            cplx sg = sigma_pp(p, ne[(size_t)i * ld + j]);
            cplx ap = 2.0 * dx * dx + 2.0 * dr * dr + dx * dx * dr * dr / (r[j] * r[j]) -
                      ORC_EPS0 * p->epsr * ORC_MU0 * w * w * dx * dx * dr * dr + iwmu0 * sg * dx * dx * dr * dr;
            cset(aP, CI(a, i, j), ap);
            cset(aE, CI(a, i, j), dr * dr);
            cset(aW, CI(a, i, j), dr * dr);
            cset(aN, CI(a, i, j), qR * dx * dx);
            cset(aS, CI(a, i, j), qL * dx * dx);
            cset(b, CI(a, i, j), (0.0 + I * (-w * ORC_MU0 * dx * dx * dr * dr)) * cget(J, CI(a, i, j)));
        }
    /* left wall (:2713-2727) */
    for (int j = 1; j <= nj; j++) {
        size_t k = CI(a, 1, j);
        if (p->pec_z) {
            cset(aP, k, cget(aP, k) + cget(aW, k));
        } else {
            double c0 = sqrt(1.0 / (ORC_EPS0 * p->epsr * ORC_MU0));
            double deno = pow(2.0 * c0, 2) + pow(w * dx, 2);
            cplx R = (pow(2.0 * c0, 2) - pow(w * dx, 2)) + I * (-4.0 * c0 * w * dx);
            R = R / (deno + 1e-100);
            cset(aP, k, cget(aP, k) - cget(aW, k) * R);
        }
        cset(aW, k, 0.0);
    }
    /* right wall (:2729-2737) */
    for (int j = 1; j <= nj; j++) {
        size_t k = CI(a, ni, j);
        cset(aP, k, cget(aP, k) + cget(aE, k));
        cset(aE, k, 0.0);
    }
    /* lower boundary: the axis (r_{1/2} = 0) needs no fold; a PEC inner wall at
     * rmin > 0 is folded as rE = 0 on the face (:2739-2747) */
    if (p->rmin > 0.0)
        for (int i = 1; i <= ni; i++) {
            size_t k = CI(a, i, 1);
            cset(aP, k, cget(aP, k) + r[1] / r[0] * cget(aS, k));
            cset(aS, k, 0.0);
        }
    /* upper boundary (:2749-2771) */
    for (int i = 1; i <= ni; i++) {
        size_t k = CI(a, i, nj);
        int j = nj;
        if (p->pec_rmax) {
            cset(aP, k, cget(aP, k) + r[j] / r[j + 1] * cget(aN, k));
        } else {
            double c0 = sqrt(1.0 / (ORC_EPS0 * p->epsr * ORC_MU0));
            double r_t = (r[j + 1] + r[j]) / 2.0;
            double qL = r[j] / (r_t + 1e-100);
            double qR = r[j + 1] / (r_t + 1e-100);
            double deno = pow(2.0 * c0 * sqrt(qR), 2) + pow(w * dr * qR, 2);
            cplx R = (pow(2.0 * c0, 2) * sqrt(qL * qR) - pow(w * dr, 2) * qL * qR) +
                     I * (-2.0 * c0 * w * dr * (qL * sqrt(qR) + qR * sqrt(qL)));
            R = R / (deno + 1e-100);
            cset(aP, k, cget(aP, k) - cget(aN, k) * R);
        }
        cset(aN, k, 0.0);
    }
    return 0;
}

double orc_helm_post(const orc_helm_params* p, const double* ne, double* E, double* Pabs)
{
    const int ni = p->ni, nj = p->nj, ld = nj + 2;
    const double dx = p->dx, dr = p->dr;
    double r[nj + 2];
    for (int j = 0; j < nj + 2; j++) r[j] = p->rmin + (j - 0.5) * dr;
    /* ghost cells (:3213-3264) */
    for (int j = 1; j <= nj; j++) {
        if (p->pec_z) {
            cset(E, CI(a, 0, j), -cget(E, CI(a, 1, j)));
        } else {
            double c0 = sqrt(1.0 / (ORC_EPS0 * p->epsr * ORC_MU0));
            double deno = pow(2.0 * c0, 2) + pow(p->omega * dx, 2);
            cplx R = (pow(2.0 * c0, 2) - pow(p->omega * dx, 2)) + I * (-4.0 * c0 * p->omega * dx);
            R = R / (deno + 1e-100);
            cset(E, CI(a, 0, j), R * cget(E, CI(a, 1, j)));
        }
        cset(E, CI(a, ni + 1, j), -cget(E, CI(a, ni, j)));
    }
    if (p->rmin > 0.0)
        for (int i = 1; i <= ni; i++) cset(E, CI(a, i, 0), -r[1] / r[0] * cget(E, CI(a, i, 1)));
    for (int i = 1; i <= ni; i++) {
        int j = nj;
        if (p->pec_rmax) {
            cset(E, CI(a, i, j + 1), -r[j] / r[j + 1] * cget(E, CI(a, i, j)));
        } else {
            double c0 = sqrt(1.0 / (ORC_EPS0 * p->epsr * ORC_MU0));
            double r_t = (r[j + 1] + r[j]) / 2.0;
            double qL = r[j] / (r_t + 1e-100);
            double qR = r[j + 1] / (r_t + 1e-100);
            double deno = pow(2.0 * c0 * sqrt(qR), 2) + pow(p->omega * dr * qR, 2);
            cplx R = (pow(2.0 * c0, 2) * sqrt(qL * qR) - pow(p->omega * dr, 2) * qL * qR) +
                     I * (-2.0 * c0 * p->omega * dr * (qL * sqrt(qR) + qR * sqrt(qL)));
            R = R / (deno + 1e-100);
            cset(E, CI(a, i, j + 1), R * cget(E, CI(a, i, j)));
        }
    }
    /* J = sigma E, Pabs = 0.5 Re(J E*) (emfieldModule.hpp:7733-7743) */
    double P = 0.0;
    for (int i = 1; i <= ni; i++)
        for (int j = 1; j <= nj; j++) {
            cplx e = cget(E, CI(a, i, j));
            cplx Jp = sigma_pp(p, ne[(size_t)i * ld + j]) * e;
            double pa = 0.5 * creal(Jp * conj(e));
            Pabs[(size_t)i * ld + j] = pa;
            P = P + pa * (dx * dr * 2.0 * M_PI * r[j]);
        }
    return P;
}

/* S -> ABCD (initialModule.hpp:1689-1708); S = {S11, S12, S21, S22} as (re, im) */
void orc_s_to_abcd(const double* S, double Z0, double* T)
{
    cplx S11 = S[0] + I * S[1], S12 = S[2] + I * S[3], S21 = S[4] + I * S[5], S22 = S[6] + I * S[7];
    cplx den = 2.0 * S21;
    cplx A = ((1.0 + S11) * (1.0 - S22) + S12 * S21) / den;
    cplx B = Z0 * ((1.0 + S11) * (1.0 + S22) - S12 * S21) / den;
    cplx C = 1.0 / Z0 * ((1.0 - S11) * (1.0 - S22) - S12 * S21) / den;
    cplx D = ((1.0 - S11) * (1.0 + S22) + S12 * S21) / den;
    cplx v[4] = {A, B, C, D};
    for (int k = 0; k < 4; k++) {
        T[2 * k] = creal(v[k]);
        T[2 * k + 1] = cimag(v[k]);
    }
}

/* ABCD -> S (initialModule.hpp:1715-1739) */
void orc_abcd_to_s(const double* T, double Z0, double* S)
{
    cplx A = T[0] + I * T[1], B = T[2] + I * T[3], C = T[4] + I * T[5], D = T[6] + I * T[7];
    cplx den = A + B / Z0 + C * Z0 + D;
    cplx S11 = (A + B / Z0 - C * Z0 - D) / den;
    cplx S21 = 2.0 / den;
    cplx S12 = 2.0 * (A * D - B * C) / den;
    cplx S22 = (-A + B / Z0 - C * Z0 + D) / den;
    cplx v[4] = {S11, S12, S21, S22};
    for (int k = 0; k < 4; k++) {
        S[2 * k] = creal(v[k]);
        S[2 * k + 1] = cimag(v[k]);
    }
}
