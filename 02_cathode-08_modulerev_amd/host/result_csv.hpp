// result_csv.hpp -- the reference's result file (results/result<n>.csv,
// OutputModule::output, include/outputModule.hpp:185-426), written column for
// column from the state the device step keeps (download it once, then call
// write_result_csv), so the reference's plotting scripts read it unchanged.
//
// Layout: one header line of 102 names (below), then one row per interior
// cell, i = 1..ni (outer), j = 1..nj (inner).  Values go through the C++
// stream's default formatting (six significant digits; "-0", "nan" and "inf"
// as the stream prints them), as the reference writes them.  Face quantities
// are averaged to the cell centre (x faces i, i+1; r faces j, j+1 weighted
// by their radii), the plasma-side columns are multiplied by jdgBnd_flc and
// the microwave amplitudes by jdgBnd_Ep, as there.  The velocity of the
// electrons divides by n_e + 1e-3 max n_e (max over blocks 0..n_bl-2).
//
// Fields are (ni+2)*(nj+2) row-major [i][j] arrays named as the reference's
// members (gc.rhoe -> "rhoe", gx.rhoUex -> "rhoUex", ...; jdgBnd_flc and
// jdgBnd_Ep as doubles); a field that is not given reads as 0.
#pragma once

#include <cmath>
#include <complex>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace cathode_host {

inline const std::vector<std::string>& result_columns()
{
    static const std::vector<std::string> c = {
        "i", "j", "x", "r", "rhoi", "rhoe", "Uix", "Uir", "Uip", "Ui_mag", "rhoUix", "rhoUir", "rhoUip", "Te",
        "Te_eff", "rhoeps", "Gx", "Gr", "phi", "Ex", "Er", "rhoUex", "rhoUer", "rhoUep", "Uex", "Uer", "Uep",
        "Ue_mag", "rate_ionize", "nu_m", "nu_m1", "nu_ionz", "nu_ionzStep", "nu_excStep", "nu_exc", "nu_excMeta",
        "nui_m", "mu_para", "mu_perp", "mu_x", "mu_r", "mu_c", "Ap", "Bx", "Br", "Bmag", "Omega", "Omegax",
        "Omegar", "Pabs", "rate_eloss", "Loss", "P_ES", "Pnet", "nabla_rhoUe", "nabla_rhoUa", "nabla_rhoUi",
        "nabla_rhoUm", "rate_rhom", "rateG_rhom", "rateL_rhom", "nabla_G", "nabla_Ga", "rhom", "rhoUmx", "rhoUmr",
        "Umx", "Umr", "Debye", "phi_analytical", "|Ex_mw|", "|Er_mw|", "|Ep_mw|", "rate_ionize_direct",
        "rate_ionize_stepwise", "rhon", "d|B|/dx", "d|B|/dr", "rate_eloss_n", "rate_eloss_m", "distECR", "nu_eff",
        "Jx_exc", "Jr_exc", "Jp_exc", "scx", "scr", "rhoq_sc", "nUex", "nUer", "nabla_nUe", "rhoUnx", "rhoUnr",
        "nabla_rhoUn", "rate_rhon", "nu_super", "deltaB", "nu_ano", "nu_ano_IAT", "jdgBnd_flc", "Mache", "zero"};
    return c;
}

struct ResultGrid {
    int ni = 0, nj = 0;
    double xL = 0.0, dx = 0.0, rmin = 0.0, dr = 0.0;
    double dt = 0.0;                      // Params::dt
    int ndt_i = 1, ndt_m = 1, ndt_n = 1;  // time-step multipliers of ions, metastables, neutrals
    double masse = 0.0, massi = 0.0;
    double nu_eff = 0.0, deltaECR = 1.0;  // collisionless-heating frequency and ECR layer width
    double dTe = 1.0;                     // Te step of the energy-loss tables [K]
    int n_bl = 0;                         // blocks, the coaxial feed last (gc.i/j_flc_bl)
    std::vector<int> cell_box;            // 4 per block: ilo, ihi, jlo, jhi
};

struct ResultFields {
    std::map<std::string, const double*> real;
    std::map<std::string, const std::complex<double>*> cplx;   // E1x, E1r, E1p, J1x, J1r, J1p
    std::vector<double> eloss_n, eloss_m;                      // bo.rate_eloss_n_N, bo.rate_eloss_m_N
};

namespace detail_result {

// FluidModule's table lookup func_Te (include/miniFuncs.hpp:35-46): linear
// between entries, the last entry from the last interval on
inline double func_Te(const std::vector<double>& v, double Te, double dTe)
{
    if (v.empty()) return 0.0;
    int iL = (int)(Te / dTe);
    const int iR = (int)(Te / dTe) + 1;
    const double TeL = iL * dTe;
    if ((unsigned)iR < (unsigned)(v.size() - 1)) {
        iL = iL < 0 ? 0 : iL;
        return (v[iR] - v[iL]) / (dTe + 1e-100) * (Te - TeL) + v[iL];
    }
    return v[v.size() - 1];
}

}  // namespace detail_result

inline void write_result_csv(const std::string& path, const ResultGrid& g, const ResultFields& F)
{
    using std::complex;
    const double E0 = 1.60217662e-19, BOLTZ = 1.380649e-23, EPS0 = 8.8541878128e-12;  // constants.hpp:5-8
    const int ni = g.ni, nj = g.nj, ld = nj + 2;
    if (ni < 1 || nj < 1 || g.n_bl < 1 || (int)g.cell_box.size() < 4 * g.n_bl)
        throw std::runtime_error("write_result_csv: grid or block boxes missing");
    const size_t n = (size_t)(ni + 2) * ld;
    const std::vector<double> zeros(n, 0.0);
    const std::vector<complex<double> > czeros(n, 0.0);
    auto R = [&](const char* name) -> const double* {
        auto it = F.real.find(name);
        return it != F.real.end() && it->second ? it->second : zeros.data();
    };
    auto C = [&](const char* name) -> const complex<double>* {
        auto it = F.cplx.find(name);
        return it != F.cplx.end() && it->second ? it->second : czeros.data();
    };
    std::vector<double> r(ld), x(ni + 2);
    for (int j = 0; j < ld; j++) r[j] = g.rmin + (j - 0.5) * g.dr;   // initialModule.hpp:83-88
    for (int i = 0; i < ni + 2; i++) x[i] = g.xL + (i - 0.5) * g.dx;
    auto box = [&](int b, int k) { return g.cell_box[4 * b + k]; };
    auto I = [&](int i, int j) { return (size_t)i * ld + j; };

    const double *rhoe = R("rhoe"), *rhoe_old = R("rhoe_old"), *rhoi = R("rhoi"), *rhoi_old = R("rhoi_old");
    const double *Uix = R("Uix"), *Uir = R("Uir"), *Uip = R("Uip"), *rhoUix = R("rhoUix"), *rhoUir = R("rhoUir");
    const double *rhoUix_wall = R("rhoUix_wall"), *rhoUir_wall = R("rhoUir_wall");
    const double *rhoUex = R("rhoUex"), *rhoUer = R("rhoUer"), *Te = R("Te"), *rhoeps = R("rhoeps");
    const double *rhoeps_old = R("rhoeps_old"), *Gx = R("Gx"), *Gr = R("Gr"), *phi = R("phi"), *Ex = R("Ex");
    const double *Er = R("Er"), *rate_ionize = R("rate_ionize"), *nu_m = R("nu_m"), *nu_m1 = R("nu_m1");
    const double *nu_ionz = R("nu_ionz"), *nu_ionzStep = R("nu_ionzStep"), *nu_excStep = R("nu_excStep");
    const double *nu_exc = R("nu_exc"), *nu_excMeta = R("nu_excMeta"), *nui_m = R("nui_m");
    const double *mu_para = R("mu_para"), *mu_perp = R("mu_perp"), *mu_xx = R("mu_xx"), *mu_rr = R("mu_rr");
    const double *mu_xr = R("mu_xr"), *Ap = R("Ap"), *Bx = R("Bx"), *Br = R("Br"), *Pabs = R("Pabs");
    const double *rate_eloss = R("rate_eloss"), *rhom = R("rhom"), *rhom_old = R("rhom_old");
    const double *rhoUmx = R("rhoUmx"), *rhoUmr = R("rhoUmr"), *rhon = R("rhon"), *rhon_old = R("rhon_old");
    const double *rhoUnx = R("rhoUnx"), *rhoUnr = R("rhoUnr"), *distECR = R("distECR"), *scx = R("scx");
    const double *scr = R("scr"), *nUex = R("nUex"), *nUer = R("nUer"), *nu_super = R("nu_super");
    const double *nu_ano = R("nu_ano"), *nu_ano_IAT = R("nu_ano_IAT"), *flc = R("jdgBnd_flc");
    const double* actp = R("jdgBnd_Ep");
    const complex<double>*E1x = C("E1x"), *E1r = C("E1r"), *E1p = C("E1p");
    const complex<double>*J1x = C("J1x"), *J1r = C("J1r"), *J1p = C("J1p");

    // the n_e floor of the velocities: 1e-3 of the largest n_e over blocks 0..n_bl-2 (:187-199)
    double rhoe_max = 0.0;
    const double ratio_rhoe = 0.001;
    for (int b = 0; b < g.n_bl - 1; b++)
        for (int i = box(b, 0); i <= box(b, 1); i++)
            for (int j = box(b, 2); j <= box(b, 3); j++)
                if (rhoe_max < rhoe[I(i, j)]) rhoe_max = rhoe[I(i, j)];
    const double dt_i = g.dt * double(g.ndt_i), dt_m = g.dt * double(g.ndt_m), dt_n = g.dt * double(g.ndt_n);
    const bool five = g.n_bl >= 5;   // the wall flags name blocks 0..4

    std::ofstream out(path);
    if (!out) throw std::runtime_error("write_result_csv: cannot open " + path);
    const std::vector<std::string>& cols = result_columns();
    for (size_t k = 0; k < cols.size(); k++) out << (k ? "," : "") << cols[k];
    out << "\n";
    for (int i = 1; i <= ni; i++)
        for (int j = 1; j <= nj; j++) {
            const size_t q = I(i, j);
            const double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
            const double qR = (r[j] + r[j + 1]) / 2.0 / r[j], qL = (r[j] + r[j - 1]) / 2.0 / r[j];
            // faces next to a wall drop the convective ion flux (the z0..z5, x0..x6 flags, :231-260)
            double bLx_wall = 1.0, bRx_wall = 1.0, bLr_wall = 1.0, bRr_wall = 1.0;
            if (five) {
                bLx_wall = double(i != box(0, 0) || j < box(0, 2));
                bLx_wall *= double(i != box(2, 0) || j > box(0, 2) - 1);
                bLx_wall *= double(i != box(3, 0) || j > box(2, 2) - 1);
                bLx_wall *= double(i != box(4, 0) || j < box(1, 2));
                bRx_wall = double(i != box(1, 1) || j < box(1, 2));
                bRx_wall *= double(i != box(4, 1));
                bLr_wall = double(j != box(0, 2) || i > box(0, 1));
                bLr_wall *= double(j != box(2, 2) || i > box(2, 1));
                bRr_wall = double(j != box(0, 3) || i > box(1, 1));
                bRr_wall *= double(j != box(3, 3) || i < box(1, 1) + 1 || i > box(3, 1));
            }
            const double f = flc[q];
            const double Bmag = std::sqrt(Bx[q] * Bx[q] + Br[q] * Br[q]);
            auto absB = [&](size_t k) { return std::sqrt(Bx[k] * Bx[k] + Br[k] * Br[k]); };
            const double dBdx = (absB(I(i + 1, j)) - absB(I(i - 1, j))) / (2.0 * g.dx);
            const double dBdr = (absB(I(i, j + 1)) - absB(I(i, j - 1))) / (2.0 * g.dr);
            const double Omega = E0 * Bmag / (g.masse * (nu_m[q]) + 1e-100);
            const double Omega_x = E0 * Bx[q] / (g.masse * nu_m[q] + 1e-100);
            const double Omega_r = E0 * Br[q] / (g.masse * nu_m[q] + 1e-100);
            const double gBx = (Bx[I(i + 1, j)] - Bx[I(i - 1, j)]) / (2.0 * g.dx);
            const double gBr = (Br[I(i + 1, j)] - Br[I(i - 1, j)]) / (2.0 * g.dr);
            const double nablaB = std::sqrt(gBx * gBx + gBr * gBr);
            const double deltaB = Bmag / (nablaB + 1e-100);
            const double Uix_c = (Uix[q] + Uix[I(i + 1, j)]) / 2.0;
            const double Uir_c = (rL * Uir[q] + rR * Uir[I(i, j + 1)]) / (2.0 * r[j]);
            const double Uip_c = Uip[q];
            const double rhoUix_c = (rhoUix[q] + rhoUix[I(i + 1, j)]) / 2.0;
            const double rhoUir_c = (rL * rhoUir[q] + rR * rhoUir[I(i, j + 1)]) / (2.0 * r[j]);
            const double rhoUip_c = rhoi[q] * Uip[q];
            const double Ui_mag = std::sqrt(Uix_c * Uix_c + Uir_c * Uir_c);
            const double rhoUex_c = (rhoUex[q] + rhoUex[I(i + 1, j)]) / 2.0;
            const double rhoUer_c = (rL * rhoUer[q] + rR * rhoUer[I(i, j + 1)]) / (2.0 * r[j]);
            const double rhoUep_c = rhoUer_c * Omega_x - rhoUex_c * Omega_r;
            const double nfl = rhoe[q] + ratio_rhoe * rhoe_max;
            const double Uex_c = rhoUex_c / nfl, Uer_c = rhoUer_c / nfl, Uep_c = rhoUep_c / nfl;
            const double Ue_mag = std::sqrt(Uex_c * Uex_c + Uer_c * Uer_c);
            const double Te_eff = Te[q] + 1.0 / (3.0 * BOLTZ) * g.masse * Ue_mag * Ue_mag;
            const double rhoUmx_c = (rhoUmx[q] + rhoUmx[I(i + 1, j)]) / 2.0;
            const double rhoUmr_c = (rL * rhoUmr[q] + rR * rhoUmr[I(i, j + 1)]) / (2.0 * r[j]);
            const double rhoUnx_c = (rhoUnx[q] + rhoUnx[I(i + 1, j)]) / 2.0;
            const double rhoUnr_c = (rL * rhoUnr[q] + rR * rhoUnr[I(i, j + 1)]) / (2.0 * r[j]);
            const double Ex_c = (Ex[q] + Ex[I(i + 1, j)]) / 2.0;
            const double Er_c = (rL * Er[q] + rR * Er[I(i, j + 1)]) / (2.0 * r[j]);
            const double rhoi_Lr = (rhoi[q] + rhoi[I(i, j - 1)]) / (2.0);
            const double nabla_rhoUe = (rhoUex[I(i + 1, j)] - rhoUex[q]) / g.dx +
                                       (qR * rhoUer[I(i, j + 1)] - qL * rhoUer[q]) / g.dr;
            const double nabla_rhoUi = (rhoUix[I(i + 1, j)] * bRx_wall - rhoUix[q] * bLx_wall) / g.dx +
                                       (qR * rhoUir[I(i, j + 1)] * bRr_wall - qL * rhoUir[q] * bLr_wall) / g.dr +
                                       (rhoUix_wall[I(i + 1, j)] - rhoUix_wall[q]) / g.dx +
                                       (qR * rhoUir_wall[I(i, j + 1)] - qL * rhoUir_wall[q]) / g.dr;
            const double nabla_rhoUm = (rhoUmx[I(i + 1, j)] - rhoUmx[q]) / g.dx +
                                       (qR * rhoUmr[I(i, j + 1)] - qL * rhoUmr[q]) / g.dr;
            const double nabla_rhoUn = (rhoUnx[I(i + 1, j)] - rhoUnx[q]) / g.dx +
                                       (qR * rhoUnr[I(i, j + 1)] - qL * rhoUnr[q]) / g.dr;
            const double lapPhi = -(Ex[I(i + 1, j)] - Ex[q]) / g.dx - (qR * Er[I(i, j + 1)] - qL * Er[q]) / g.dr;
            const double nabla_nUe = (nUex[I(i + 1, j)] - nUex[q]) / g.dx + (qR * nUer[I(i, j + 1)] - qL * nUer[q]) / g.dr;
            const double nUex_c = (nUex[q] + nUex[I(i + 1, j)]) / 2.0;
            const double nUer_c = (rL * nUer[q] + rR * nUer[I(i, j + 1)]) / (2.0 * r[j]);
            const double Gm = rhoe[q] * (nu_excMeta[q] + 0.5 * nu_exc[q]);
            const double Lm = rhoe[q] * (nu_ionzStep[q] + nu_excStep[q] + nu_super[q]);
            const double Gn = (nu_excStep[q] + nu_super[q]) * rhoe[q];
            const double Ln = (nu_excMeta[q] + 0.5 * nu_exc[q] + nu_ionz[q]) * rhoe[q];
            const double nabla_G = (Gx[I(i + 1, j)] - Gx[q]) / g.dx + (qR * Gr[I(i, j + 1)] - qL * Gr[q]) / g.dr;
            const double P_ES = -(E0 * (Ex_c * rhoUex_c + Er_c * rhoUer_c));
            const double Ex_mw = std::abs(E1x[q] + E1x[I(i + 1, j)]) / 2.0;
            const double Er_mw = std::abs(rL * E1r[I(i, j + 1)] + rR * E1r[q]) / (2.0 * r[j]);
            const double Ep_mw = std::abs(E1p[q]);
            const double Jx_c = std::abs(J1x[q] + J1x[I(i + 1, j)]) / 2.0;
            const double Jr_c = std::abs(rL * J1r[I(i, j + 1)] + rR * J1r[q]) / (2.0 * r[j]);
            const double Jp_c = std::abs(J1p[q]);
            const double ut = std::sqrt(2.0 * BOLTZ * (Te[q] + 1.0 * E0 / BOLTZ) / g.masse);
            const double Mache = std::sqrt(Uex_c * Uex_c + Uer_c * Uer_c) / (ut + 100.0);
            const double rhoq_sc = (scx[q] * 2.0 * M_PI * r[j] * g.dr + scr[q] * 2.0 * M_PI * rR * g.dx) /
                                   (2.0 * M_PI * r[j] * g.dx * g.dr);
            const double ep = actp[q];
            const double dE = distECR[q] / g.deltaECR;
            out << i << "," << j << "," << x[i] << "," << r[j] << "," << rhoi[q] * f << "," << rhoe[q] * f << ","
                << Uix_c * f << "," << Uir_c * f << "," << Uip_c * f << "," << Ui_mag * f << "," << rhoUix_c * f << ","
                << rhoUir_c * f << "," << rhoUip_c * f << "," << Te[q] * BOLTZ / E0 * f << ","
                << Te_eff * BOLTZ / E0 * f << "," << rhoeps[q] * f << "," << Gx[q] * f << "," << Gr[q] * f << ","
                << phi[q] * f << "," << Ex[q] * f << "," << Er[q] * f << "," << rhoUex_c * f << "," << rhoUer_c * f
                << "," << rhoUep_c * f << "," << Uex_c * f << "," << Uer_c * f << "," << Uep_c * f << ","
                << Ue_mag * f << "," << rate_ionize[q] * f << "," << nu_m[q] * f << "," << nu_m1[q] * f << ","
                << nu_ionz[q] * f << "," << nu_ionzStep[q] * f << "," << nu_excStep[q] * f << "," << nu_exc[q] * f
                << "," << nu_excMeta[q] * f << "," << nui_m[q] * f << "," << mu_para[q] * f << ","
                << mu_perp[q] * f << "," << mu_xx[q] * f << "," << mu_rr[q] * f << "," << mu_xr[q] * f << ","
                << Ap[q] * f << "," << Bx[q] * f << "," << Br[q] * f << "," << Bmag * f << "," << Omega * f << ","
                << Omega_x * f << "," << Omega_r * f << "," << Pabs[q] * f << "," << rate_eloss[q] * f << ","
                << rhoe[q] * rate_eloss[q] * f << "," << P_ES * f << ","
                << (Pabs[q] - rhoe[q] * rate_eloss[q] + P_ES) * f << ","
                << (nabla_rhoUe + (rhoe[q] - rhoe_old[q]) / g.dt) * f << "," << (nabla_rhoUe) * f << ","
                << (nabla_rhoUi + (rhoi[q] - rhoi_old[q]) / dt_i) * f << ","
                << (nabla_rhoUm + (rhom[q] - rhom_old[q]) / dt_m) * f << "," << (Gm - Lm) * f << "," << (Gm) * f
                << "," << (Lm) * f << ","
                << (nabla_G + (rhoeps[q] - rhoeps_old[q]) / g.dt - P_ES + rhoe[q] * rate_eloss[q]) * f << ","
                << (nabla_G + rhoe[q] * rate_eloss[q]) * f << "," << rhom[q] * f << "," << rhoUmx_c * f << ","
                << rhoUmr_c * f << "," << rhoUmx_c / (rhom[q] + 1e-100) * f << ","
                << rhoUmr_c / (rhom[q] + 1e-100) * f << ","
                << std::sqrt(EPS0 * BOLTZ * Te[q] / ((rhoe[q] + 1) * E0 * E0)) * f << ","
                << BOLTZ * Te[q] / E0 *
                       (0.5 * std::log(g.massi / (2.0 * M_PI * g.masse)) +
                        std::log(std::fabs(rhoi_Lr * Uir[q] / (rhoUer[q] + 1e-100)))) *
                       f
                << "," << Ex_mw * ep << "," << Er_mw * ep << "," << Ep_mw * ep << "," << rhoe[q] * nu_ionz[q] * f
                << "," << rhoe[q] * nu_ionzStep[q] * f << "," << rhon[q] * f << "," << dBdx * f << "," << dBdr * f
                << "," << rhon[q] * detail_result::func_Te(F.eloss_n, Te[q], g.dTe) * f << ","
                << rhom[q] * detail_result::func_Te(F.eloss_m, Te[q], g.dTe) * f << "," << distECR[q] * f << ","
                << g.nu_eff / (1.0 + 4.0 * (dE * dE)) * f << "," << Jx_c << "," << Jr_c << "," << Jp_c << ","
                << scx[q] * f << "," << scr[q] * f << "," << rhoq_sc * f << "," << nUex_c * f << "," << nUer_c * f
                << "," << (nabla_nUe + (rhoi[q] - rhoe_old[q]) / g.dt + EPS0 / (E0 * g.dt) * lapPhi) * f << ","
                << rhoUnx_c * f << "," << rhoUnr_c * f << ","
                << (nabla_rhoUn + (rhon[q] - rhon_old[q]) / dt_n) * f << "," << (Gn - Ln) * f << ","
                << nu_super[q] * f << "," << deltaB * f << "," << nu_ano[q] * f << "," << nu_ano_IAT[q] * f << ","
                << f << "," << Mache * f << "," << 0.0 << "\n";
        }
    if (!out) throw std::runtime_error("write_result_csv: write failed for " + path);
}

}  // namespace cathode_host
