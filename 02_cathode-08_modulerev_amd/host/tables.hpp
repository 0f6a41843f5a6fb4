// tables.hpp -- readers for the reference's collision-rate and secondary-
// emission tables: input/rateCoef_e.csv and input/rateCoef_i.csv into the
// lookup tables of cathode_ndd_set_rate_tables (InputModule::
// input_BOLSIG_data, include/inputModule.hpp:1407-1654), and
// input/coefEISEE.csv into the yield tables of cathode_ndd_params
// (InputModule::input_SEE_data, :606-659).
//
// rateCoef_e.csv: row 1 names each column's process (Elastic, Excitation,
// Excitation_reso, Excitation_meta, Ionization, Ionization_step,
// Excitation_step; any other name is super-elastic), row 2 its threshold
// [eV], then one row per mean energy [eV, uniform step] of rate
// coefficients [m^3/s].  The tables are the per-row sums of the columns of
// each kind, in column order, and the energy-loss rates sum rate x threshold
// over the ground-state kinds (elastic .. ionization) and the metastable
// kinds (stepwise ionization, stepwise excitation, super-elastic).  The
// lookup step in Te is 2/3 of the energy step (mean energy 3/2 k Te).
// rateCoef_i.csv: the same layout over drift energy; every column sums into
// the ion-neutral table, step dEi in J.  coefEISEE.csv: energy [eV], then the
// backscattered, rediffused and true-secondary yields; the Te step is half
// the energy step.
//
// Numbers are read with strtod, fields split on ','; errors are thrown as
// std::runtime_error (the reference prints and aborts).
#pragma once

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cathode_hip.h"

namespace cathode_host {

constexpr double TAB_E0 = 1.60217662e-19;    // include/constants.hpp:6
constexpr double TAB_BOLTZ = 1.380649e-23;   // :5

struct RateTables {
    std::vector<double> nu_elas, nu_exc, nu_excReso, nu_excMeta, nu_ionz, nu_ionzStep, nu_excStep, nu_super;
    std::vector<double> eloss_n, eloss_m, nui_m;
    double dTe = 0.0, dEi = 0.0;

    // pointers into the vectors, for cathode_ndd_set_rate_tables
    cathode_rate_tables view() const
    {
        cathode_rate_tables t{};
        t.n_e = (int)nu_elas.size();
        t.nu_elas = nu_elas.data();
        t.nu_exc = nu_exc.data();
        t.nu_excReso = nu_excReso.data();
        t.nu_excMeta = nu_excMeta.data();
        t.nu_ionz = nu_ionz.data();
        t.nu_ionzStep = nu_ionzStep.data();
        t.nu_excStep = nu_excStep.data();
        t.nu_super = nu_super.data();
        t.eloss_n = eloss_n.data();
        t.eloss_m = eloss_m.data();
        t.dTe = dTe;
        t.n_i = (int)nui_m.size();
        t.nui_m = nui_m.data();
        t.dEi = dEi;
        return t;
    }
};

namespace detail {

inline std::vector<std::string> fields(const std::string& line)
{
    std::vector<std::string> out;
    std::istringstream in(line);
    std::string f;
    while (std::getline(in, f, ',')) out.push_back(f);
    return out;
}

inline double number(const std::string& s, const std::string& where)
{
    const char* b = s.c_str();
    char* e = nullptr;
    const double v = std::strtod(b, &e);
    if (e == b) throw std::runtime_error("[" + where + "] not a number: '" + s + "'");
    return v;
}

// a table file: row 1 and row 2 as text fields, the numeric rows, and the
// step of the first column
struct Csv {
    std::vector<std::string> head, second;
    std::vector<std::vector<double> > rows;  // rows[k][c], c = 0 the abscissa
    double step = 0.0;
};

inline Csv read_csv(const std::string& path, int min_cols, bool second_row)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("[" + path + "] Error! File not found.");
    Csv t;
    std::string line;
    if (!std::getline(in, line)) throw std::runtime_error("[" + path + "] empty file");
    t.head = fields(line);
    if ((int)t.head.size() < min_cols) throw std::runtime_error("[" + path + "] too few columns");
    if (second_row) {
        if (!std::getline(in, line)) throw std::runtime_error("[" + path + "] no threshold row");
        t.second = fields(line);
    }
    while (std::getline(in, line)) {
        std::vector<std::string> f = fields(line);
        if (f.size() < t.head.size()) throw std::runtime_error("[" + path + "] short row");
        std::vector<double> v(t.head.size());
        for (size_t c = 0; c < t.head.size(); c++) v[c] = number(f[c], path);
        t.rows.push_back(v);
    }
    if (t.rows.size() < 2) throw std::runtime_error("[" + path + "] fewer than two data rows");
    t.step = t.rows[1][0] - t.rows[0][0];
    return t;
}

}  // namespace detail

// process kind of a rateCoef_e.csv column name
inline int rate_kind(const std::string& name)
{
    static const char* kinds[7] = {"Elastic", "Excitation", "Excitation_reso", "Excitation_meta", "Ionization",
                                   "Ionization_step", "Excitation_step"};
    for (int k = 0; k < 7; k++)
        if (name == kinds[k]) return k;
    return 7;  // super-elastic
}

inline RateTables read_rate_tables(const std::string& path_e, const std::string& path_i)
{
    RateTables T;
    const detail::Csv e = detail::read_csv(path_e, 2, true);
    const size_t nc = e.head.size() - 1, nr = e.rows.size();
    if (e.second.size() < nc + 1) throw std::runtime_error("[" + path_e + "] short threshold row");
    std::vector<int> kind(nc);
    std::vector<double> thr(nc);
    for (size_t c = 0; c < nc; c++) {
        kind[c] = rate_kind(e.head[c + 1]);
        thr[c] = detail::number(e.second[c + 1], path_e) * TAB_E0;
    }
    std::vector<double>* sum[8] = {&T.nu_elas, &T.nu_exc, &T.nu_excReso, &T.nu_excMeta,
                                   &T.nu_ionz, &T.nu_ionzStep, &T.nu_excStep, &T.nu_super};
    for (auto* v : sum) v->assign(nr, 0.0);
    T.eloss_n.assign(nr, 0.0);
    T.eloss_m.assign(nr, 0.0);
    for (size_t k = 0; k < nr; k++)
        for (size_t c = 0; c < nc; c++) {
            const double v = e.rows[k][c + 1];
            (*sum[kind[c]])[k] = (*sum[kind[c]])[k] + v;
            std::vector<double>& loss = kind[c] <= 4 ? T.eloss_n : T.eloss_m;
            loss[k] = loss[k] + v * thr[c];
        }
    const detail::Csv ion = detail::read_csv(path_i, 2, true);
    T.nui_m.assign(ion.rows.size(), 0.0);
    for (size_t k = 0; k < ion.rows.size(); k++)
        for (size_t c = 1; c < ion.head.size(); c++) T.nui_m[k] = T.nui_m[k] + ion.rows[k][c];
    T.dTe = e.step * 2.0 / 3.0 * TAB_E0 / TAB_BOLTZ;
    T.dEi = ion.step * TAB_E0;
    return T;
}

// coefEISEE.csv into the yield tables of p (at most CATHODE_SEE_TABLE_MAX rows)
inline void read_eisee_table(const std::string& path, cathode_ndd_params& p)
{
    const detail::Csv t = detail::read_csv(path, 4, false);
    if ((int)t.rows.size() > CATHODE_SEE_TABLE_MAX) throw std::runtime_error("[" + path + "] more rows than the table holds");
    p.n_see = (int)t.rows.size();
    for (size_t k = 0; k < t.rows.size(); k++) {
        p.see_eb[k] = t.rows[k][1];
        p.see_rd[k] = t.rows[k][2];
        p.see_ts[k] = t.rows[k][3];
    }
    p.dTe_SEE = t.step * 0.5 * TAB_E0 / TAB_BOLTZ;
}

}  // namespace cathode_host
