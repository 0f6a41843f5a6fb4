/*
 * orc_mw.c -- CPU FP64 restatement of the microwave power profile and port
 * normalisation, EmfieldModule::update_energy_profile
 * (include/emfieldModule.hpp:7630-8105) with the Poynting fluxes of
 * mWLoss_at_boundary_xL / rR (:8488-8668).  TEST INFRASTRUCTURE ONLY (see
 * oracle.h).
 *
 * Conductivity (:7648-7660 and the r / phi loops): with nu* = nu + i omega and
 * h = e B / (m nu*) per component, coef = n e^2 / (m nu*) / (1 + h_x^2 + h_r^2)
 * and
 *   sigma = coef [[1 + h_x^2, h_x h_r,   h_r ],
 *                 [h_x h_r,   1 + h_r^2, -h_x],
 *                 [-h_r,      h_x,       1   ]]   (rows x, r, phi; columns x, r, phi)
 * evaluated on x faces, r faces and cells with the reference's face averages
 * of n, nu, B and of the other two field components.  The tiny complex
 * EPS = (1e-100, 1e-100) guards of the reference are kept.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

#define MW_MU0 1.2566370614e-6 /* constants.hpp:8 */

static cplx cget(const double* a, size_t k) { return a[2 * k] + I * a[2 * k + 1]; }
static void st(double* a, size_t k, cplx v)
{
    a[2 * k] = creal(v);
    a[2 * k + 1] = cimag(v);
}

typedef struct {
    cplx hx, hr, coef;
} cond_t;

static cond_t conductivity(double n, double nu, double Bx, double Br, double omega, double masse)
{
    const cplx EPS = 1e-100 + 1e-100 * I;
    cond_t c;
    cplx nus = nu + I * omega;
    c.hx = ORC_E0 * Bx / (masse * nus + EPS);
    c.hr = ORC_E0 * Br / (masse * nus + EPS);
    cplx h2 = c.hx * c.hx + c.hr * c.hr;
    c.coef = n * ORC_E0 * ORC_E0 / (masse * nus + EPS) / (1.0 + h2 + EPS);
    return c;
}

/* Poynting flux through the x face i over cells jlo..jhi (mWLoss_at_boundary_xL) */
static double loss_xL(const cathode_mw_params* p, const double* r, const double* Ex, const double* Er,
                      const double* Ep, int i, int jlo, int jhi)
{
    const int ld = p->nj + 2;
    const cplx iwm = I * (1.0 / (p->omega * MW_MU0));
    double P = 0.0;
    for (int j = jlo; j <= jhi; j++) {
        double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
#define AT(a, ii, jj) cget(a, (size_t)(ii) * ld + (jj))
        cplx dErdx_L = (AT(Er, i, j) - AT(Er, i - 1, j)) / p->dx;
        cplx dExdr_L = (AT(Ex, i, j) - AT(Ex, i, j - 1)) / p->dr;
        cplx dErdx_R = (AT(Er, i, j + 1) - AT(Er, i - 1, j + 1)) / p->dx;
        cplx dExdr_R = (AT(Ex, i, j + 1) - AT(Ex, i, j)) / p->dr;
        cplx Hp_L = iwm * (dErdx_L - dExdr_L);
        cplx Hp_R = iwm * (dErdx_R - dExdr_R);
        cplx Hp = (rR * Hp_R + rL * Hp_L) / (2.0 * r[j]);
        cplx Epf = (AT(Ep, i, j) + AT(Ep, i - 1, j)) / (2.0 * rR);
        cplx Er_Lx = (rR * AT(Er, i, j + 1) + rL * AT(Er, i, j)) / (2.0 * r[j]);
        cplx Er_Rx = (rR * AT(Er, i - 1, j + 1) + rL * AT(Er, i - 1, j)) / (2.0 * r[j]);
        cplx Erf = (Er_Rx + Er_Lx) / 2.0;
        cplx Hc = orc_planted() == ORC_PLANT_MW_POYNTING_CONJ ? Hp : conj(Hp);
        cplx ExH = Erf * Hc - Epf * conj(Hp);
        P = P + 2.0 * M_PI * r[j] * (0.5 * creal(ExH)) * p->dr;
    }
    return P;
}

/* Poynting flux through the r face above cells (iMin..iMax, j) (mWLoss_at_boundary_rR) */
static double loss_rR(const cathode_mw_params* p, const double* r, const double* Ex, const double* Er,
                      const double* Ep, int j, int imin, int imax)
{
    const int ld = p->nj + 2;
    const cplx iwm = I * (1.0 / (p->omega * MW_MU0));
    double P = 0.0;
    for (int i = imin; i <= imax; i++) {
        double rt = (r[j] + r[j + 1]) / 2.0;
        cplx dErdx_L = (AT(Er, i, j + 1) - AT(Er, i - 1, j + 1)) / p->dx;
        cplx dExdr_L = (AT(Ex, i, j + 1) - AT(Ex, i, j)) / p->dr;
        cplx dErdx_R = (AT(Er, i + 1, j + 1) - AT(Er, i, j + 1)) / p->dx;
        cplx dExdr_R = (AT(Ex, i + 1, j + 1) - AT(Ex, i + 1, j)) / p->dr;
        cplx Hp_L = iwm * (dErdx_L - dExdr_L);
        cplx Hp_R = iwm * (dErdx_R - dExdr_R);
        cplx Hp = (Hp_R + Hp_L) / 2.0;
        cplx dEpdr = (r[j + 1] * AT(Ep, i, j + 1) - r[j] * AT(Ep, i, j)) / (rt * p->dr);
        cplx Hx = iwm * dEpdr;
        cplx Epf = (r[j + 1] * AT(Ep, i, j + 1) + r[j] * AT(Ep, i, j)) / (2.0 * rt);
        cplx Ex_Lx = (r[j + 1] * AT(Ex, i, j + 1) + r[j] * AT(Ex, i, j)) / (2.0 * rt);
        cplx Ex_Rx = (r[j + 1] * AT(Ex, i + 1, j + 1) + r[j] * AT(Ex, i + 1, j)) / (2.0 * rt);
        cplx Exf = (Ex_Rx + Ex_Lx) / 2.0;
        cplx Hc = orc_planted() == ORC_PLANT_MW_POYNTING_CONJ ? Hp : conj(Hp);
        cplx ExH = Epf * conj(Hx) - Exf * Hc;
        P = P + 2.0 * M_PI * rt * (0.5 * creal(ExH)) * p->dx;
    }
    return P;
}

int orc_mw_energy_profile(const cathode_mw_params* p, cathode_mw_fields* f, cathode_mw_report* rep)
{
    const int ni = p->ni, nj = p->nj, ld = nj + 2;
    if (p->n_bl < 1 || p->n_bl > CATHODE_MW_MAX_BLOCKS) return -1;
    double* r = (double*)malloc((size_t)(nj + 2) * sizeof(double));
    for (int j = 0; j < nj + 2; j++) r[j] = p->rmin + (j - 0.5) * p->dr; /* initialModule.hpp:86-88 */
    memset(rep, 0, sizeof(*rep));
    const double vol = p->dx * p->dr * 2.0 * M_PI;
#define RE(a, ii, jj) (a)[(size_t)(ii) * ld + (jj)]
    /* x faces (:7641-7680) */
    double Px = 0.0;
    for (int b = 0; b < p->n_bl; b++) {
        const int* bx = p->xface_box + 4 * b;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++) {
                double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
                cond_t c = conductivity((RE(f->rhoe, i, j) + RE(f->rhoe, i - 1, j)) / 2.0,
                                        (RE(f->nu_m, i, j) + RE(f->nu_m, i - 1, j)) / 2.0,
                                        (RE(f->Bx, i, j) + RE(f->Bx, i - 1, j)) / 2.0,
                                        (RE(f->Br, i, j) + RE(f->Br, i - 1, j)) / 2.0, p->omega, p->masse);
                cplx sxr = c.coef * (c.hx * c.hr), sxp = c.coef * c.hr, sxx = c.coef * (1.0 + c.hx * c.hx);
                if (orc_planted() == ORC_PLANT_MW_HALL_SIGN) sxp = -sxp;
                cplx Ex = AT(f->E1x, i, j);
                cplx Er = (rL * (AT(f->E1r, i, j) + AT(f->E1r, i - 1, j)) +
                           rR * (AT(f->E1r, i, j + 1) + AT(f->E1r, i - 1, j + 1))) /
                          (4.0 * r[j]);
                cplx Ep = (AT(f->E1p, i, j) + AT(f->E1p, i - 1, j)) / 2.0;
                cplx J = sxr * Er + sxp * Ep + sxx * Ex;
                st(f->J1x, (size_t)i * ld + j, J);
                double Pa = 0.5 * creal(J * conj(Ex));
                RE(f->Pabsx, i, j) = Pa;
                Px = Px + Pa * (vol * r[j]);
            }
    }
    /* r faces */
    double Pr = 0.0;
    for (int b = 0; b < p->n_bl; b++) {
        const int* bx = p->rface_box + 4 * b;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++) {
                double rt = (r[j - 1] + r[j]) / 2.0;
                cond_t c = conductivity((RE(f->rhoe, i, j) + RE(f->rhoe, i, j - 1)) / 2.0,
                                        (RE(f->nu_m, i, j) + RE(f->nu_m, i, j - 1)) / 2.0,
                                        (r[j] * RE(f->Bx, i, j) + r[j - 1] * RE(f->Bx, i, j - 1)) / (2.0 * rt),
                                        (r[j] * RE(f->Br, i, j) + r[j - 1] * RE(f->Br, i, j - 1)) / (2.0 * rt),
                                        p->omega, p->masse);
                cplx srr = c.coef * (1.0 + c.hr * c.hr), srp = c.coef * (-c.hx), srx = c.coef * (c.hx * c.hr);
                cplx Ex = (r[j] * (AT(f->E1x, i, j) + AT(f->E1x, i + 1, j)) +
                           r[j - 1] * (AT(f->E1x, i, j - 1) + AT(f->E1x, i + 1, j - 1))) /
                          (4.0 * rt);
                cplx Er = AT(f->E1r, i, j);
                cplx Ep = (r[j] * AT(f->E1p, i, j) + r[j - 1] * AT(f->E1p, i, j - 1)) / (2.0 * rt);
                cplx J = srr * Er + srp * Ep + srx * Ex;
                st(f->J1r, (size_t)i * ld + j, J);
                double Pa = 0.5 * creal(J * conj(Er));
                RE(f->Pabsr, i, j) = Pa;
                Pr = Pr + Pa * (vol * (orc_planted() == ORC_PLANT_MW_RFACE_WEIGHT ? r[j] : rt));
            }
    }
    /* cells (phi) */
    double Pp = 0.0;
    for (int b = 0; b < p->n_bl; b++) {
        const int* bx = p->cell_box + 4 * b;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++) {
                double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                double qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
                cond_t c = conductivity(RE(f->rhoe, i, j), RE(f->nu_m, i, j), RE(f->Bx, i, j), RE(f->Br, i, j),
                                        p->omega, p->masse);
                cplx spr = c.coef * c.hx, spp = c.coef * 1.0, spx = c.coef * (-c.hr);
                cplx Ex = (AT(f->E1x, i, j) + AT(f->E1x, i + 1, j)) / 2.0;
                cplx Er = (qR * AT(f->E1r, i, j + 1) + qL * AT(f->E1r, i, j)) / 2.0;
                cplx Ep = AT(f->E1p, i, j);
                cplx J = spr * Er + spp * Ep + spx * Ex;
                st(f->J1p, (size_t)i * ld + j, J);
                double Pa = 0.5 * creal(J * conj(Ep));
                RE(f->Pabsp, i, j) = Pa;
                Pp = Pp + Pa * (vol * r[j]);
            }
    }
    double power_sum = Px + Pr + Pp;
    rep->Px = Px;
    rep->Pr = Pr;
    rep->Pp = Pp;
    /* cell absorbed power (:7750-7790) */
    for (int b = 0; b < p->n_bl; b++) {
        const int* bx = p->cell_box + 4 * b;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++)
                RE(f->Pabs, i, j) = (RE(f->Pabsx, i, j) + RE(f->Pabsx, i + 1, j)) / 2.0 +
                                    (RE(f->Pabsr, i, j) + RE(f->Pabsr, i, j + 1)) / 2.0 + RE(f->Pabsp, i, j);
    }
    const size_t n = (size_t)(ni + 2) * ld;
    if (p->flag_mwRef == 1) {
        double P_out = loss_rR(p, r, f->E1x, f->E1r, f->E1p, p->j_out, p->i_out_lo, p->i_out_hi);
        double P_in = loss_xL(p, r, f->E1x, f->E1r, f->E1p, p->i_ref, p->j_in_lo, p->j_in_hi);
        /* current and voltage at the reference plane (:7874-7920), averaged over the feed's r faces */
        const int i = p->i_ref;
        const int* fb = p->rface_box + 4 * (p->n_bl - 1);
        int cnt = 0;
        cplx Iref = 0.0, coef = 0.0;
        for (int j = fb[2]; j <= fb[3]; j++) {
            double rt = (r[j] + r[j - 1]) / 2.0;
            cplx H = (AT(f->E1r, i, j) - AT(f->E1r, i - 1, j)) / p->dx / (p->omega * MW_MU0) * I;
            Iref = Iref + 2.0 * M_PI * rt * H;
            coef = coef + (AT(f->E1r, i, j) + AT(f->E1r, i - 1, j)) / 2.0 * rt;
            cnt++;
        }
        Iref = Iref / (double)cnt;
        coef = coef / (double)cnt;
        cplx Vref = coef * log(p->r2 / p->r1);
        cplx Zref = Vref / (Iref + 1e-100);
        cplx a2 = (Vref - p->Z0 * Iref) / (2.0 * sqrt(2.0 * p->Z0));
        cplx b2 = (Vref + p->Z0 * Iref) / (2.0 * sqrt(2.0 * p->Z0));
        if (orc_planted() == ORC_PLANT_MW_PORT_WAVES) {
            cplx t = a2;
            a2 = b2;
            b2 = t;
        }
        double P_fwd = creal(b2 * conj(b2)), P_ref = creal(a2 * conj(a2));
        double P_err = P_fwd - P_ref - power_sum - P_out;
        cplx S11 = p->S11[0] + I * p->S11[1], S12 = p->S12[0] + I * p->S12[1];
        cplx S21 = p->S21[0] + I * p->S21[1], S22 = p->S22[0] + I * p->S22[1];
        cplx a1 = (b2 - S22 * a2) / S21;
        cplx b1 = S11 * a1 + S12 * a2;
        double P_fwd_all = creal(a1 * conj(a1)), P_ref_all = creal(b1 * conj(b1));
        double ratio = p->Pmw / P_fwd_all;
        double ratio_eff = ratio * (power_sum + P_err) / power_sum;
        double sr = sqrt(ratio_eff);
        double* cfield[] = {f->E1x, f->E1r, f->E1p, f->J1x, f->J1r, f->J1p, f->J1x_exc, f->J1r_exc, f->J1p_exc};
        for (size_t k = 0; k < n; k++) {
            f->Pabs[k] = f->Pabs[k] * ratio_eff;
            for (int q = 0; q < 9; q++)
                if (cfield[q]) st(cfield[q], k, cget(cfield[q], k) * sr);
        }
        rep->J1r_exc = p->J1r_exc * sqrt(ratio);
        power_sum = power_sum + P_err;
        rep->P_fwd = P_fwd * ratio;
        rep->P_ref = P_ref * ratio;
        rep->P_fwd_all = P_fwd_all * ratio;
        rep->P_ref_all = P_ref_all * ratio;
        rep->power_sum = power_sum * ratio;
        rep->P_err = P_err * ratio;
        rep->P_out = P_out * ratio;
        rep->P_in = P_in;
        rep->ratio = ratio;
        rep->ratio_eff = ratio_eff;
        cplx Gp = a2 / b2, Zp = p->Z0 * (1.0 + Gp) / (1.0 - Gp);
        cplx Ga = b1 / a1, Za = p->Z0 * (1.0 + Ga) / (1.0 - Ga);
        cplx out[] = {Vref, Iref, Zref, Gp, Zp, Ga, Za};
        double* dst[] = {rep->V_ref, rep->I_ref, rep->Z_ref, rep->Gamma_p, rep->Zp, rep->Gamma_all, rep->Zall};
        for (int q = 0; q < 7; q++) {
            dst[q][0] = creal(out[q]);
            dst[q][1] = cimag(out[q]);
        }
    } else {
        double ratio = p->Pmw / power_sum;
        rep->J1r_exc = p->J1r_exc * sqrt(ratio);
        for (size_t k = 0; k < n; k++) f->Pabs[k] = f->Pabs[k] * ratio;
        rep->ratio = ratio;
        rep->ratio_eff = ratio;
        rep->power_sum = power_sum;
    }
    free(r);
    return 0;
}
