// setup_yaml.hpp -- reader for the reference's input/setup.yaml (the keys of
// InputModule's READ_DOUBLE / READ_INT lists, include/inputModule.hpp:151-267)
// into SetupParams, and the mapping of the keys the implicit step uses onto
// cathode_ndd_params.
//
// The file is a two-level YAML mapping: "category:" lines at column 0 and
// "  key: value  # comment" lines under them.  That subset is all setup.yaml
// uses, so this reader handles exactly it: comments, blank lines, scalar
// values (numbers, plain words).  A key repeated inside one category keeps
// its first value.  Errors follow the reference: every missing key is
// collected and reported together ("[setup.yaml] Error! Missing YAML
// parameters:" with "category.key" entries), a value that is not a number
// of the requested type is a "Type conversion error: category.key should be
// double|int"; both are thrown as std::runtime_error (the reference prints
// and aborts).
#pragma once

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cathode_hip.h"

namespace cathode_host {

struct SetupYaml {
    std::map<std::string, std::map<std::string, std::string> > cat;

    bool has(const std::string& c, const std::string& k) const
    {
        auto it = cat.find(c);
        return it != cat.end() && it->second.count(k);
    }
};

inline std::string yaml_trim(const std::string& s)
{
    const char* ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

inline SetupYaml parse_setup_yaml(std::istream& in)
{
    SetupYaml y;
    std::string line, current;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        if (yaml_trim(line).empty()) continue;
        const size_t colon = line.find(':');
        if (colon == std::string::npos) throw std::runtime_error("[setup.yaml] cannot parse line: " + line);
        const bool indented = line[0] == ' ' || line[0] == '\t';
        const std::string key = yaml_trim(line.substr(0, colon));
        const std::string val = yaml_trim(line.substr(colon + 1));
        if (!indented) {
            current = key;
            y.cat[current];
        } else {
            if (current.empty()) throw std::runtime_error("[setup.yaml] key outside a category: " + key);
            y.cat[current].emplace(key, val);  // first occurrence kept
        }
    }
    return y;
}

inline SetupYaml read_setup_yaml(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("[setup.yaml] cannot open " + path);
    return parse_setup_yaml(f);
}

// Collects missing keys; converts present ones (strict: the whole value must parse).
class SetupReader {
public:
    explicit SetupReader(const SetupYaml& y) : y_(y) {}

    void read(double& dst, const std::string& c, const std::string& k)
    {
        const std::string* v = find(c, k);
        if (!v) return;
        char* end = nullptr;
        const double d = std::strtod(v->c_str(), &end);
        if (v->empty() || *end != '\0') throw std::runtime_error("Type conversion error: " + c + "." + k + " should be double");
        dst = d;
    }
    void read(int& dst, const std::string& c, const std::string& k)
    {
        const std::string* v = find(c, k);
        if (!v) return;
        char* end = nullptr;
        const long d = std::strtol(v->c_str(), &end, 10);
        if (v->empty() || *end != '\0') throw std::runtime_error("Type conversion error: " + c + "." + k + " should be int");
        dst = (int)d;
    }
    // Throws the reference's missing-parameter report if anything was missing.
    void finish() const
    {
        if (missing_.empty()) return;
        std::string m = "[setup.yaml] Error! Missing YAML parameters:";
        for (const auto& s : missing_) m += "\n  - " + s;
        throw std::runtime_error(m);
    }

private:
    const std::string* find(const std::string& c, const std::string& k)
    {
        auto it = y_.cat.find(c);
        if (it == y_.cat.end()) {
            missing_.push_back(c + "." + k + "  [missing category: " + c + "]");
            return nullptr;
        }
        auto kt = it->second.find(k);
        if (kt == it->second.end()) {
            missing_.push_back(c + "." + k);
            return nullptr;
        }
        return &kt->second;
    }
    const SetupYaml& y_;
    std::vector<std::string> missing_;
};

// Every key InputModule::inputParam reads (include/inputModule.hpp:151-267,
// Params fields of include/params.hpp:8-155), under the reference's names.
struct SetupParams {
    // domain
    int ni = 0, nj = 0;
    double xL = 0, xR = 0, rmin = 0, rmax = 0, width_neutIn = 0;
    // operating_condition
    double V_bias = 0, Q_neutIn_mgs = 0, fmw = 0, Pmw = 0;
    int flag_mwRef = 0;
    // physical_parameter
    double masse = 0, massi = 0, ri = 0, Ti = 0, Tn = 0, DmN = 0, epsr_diele = 0;
    // transmission_line
    double S11_mag = 0, S11_arg_deg = 0, S21_mag = 0, S21_arg_deg = 0, S12_mag = 0, S12_arg_deg = 0, S22_mag = 0,
           S22_arg_deg = 0, Z0 = 0, h0_stab = 0, dh_stab = 0, dl_cable = 0, phase_pre_deg = 0;
    // secondary_electron_emission
    double coefIISEE_ts = 0, coefMISEE_ts = 0, Te_emitSEE_eV = 0, ratioEngy_EISEE_rd = 0;
    // governing_equation
    int flag_inertia = 0;
    double scale_inertia = 0;
    // anomalous_scattering_model
    int flag_Sagdeev = 0, flag_Bohm = 0;
    double alpha_Bohm = 0;
    // collisionless_heating_model
    double nu_eff = 0;
    // initial_condition (Te_ini converted from eV to K on reading, as the reference does)
    double rhon_ini = 0, rhoe_ini = 0, Te_ini = 0, J1r_exc = 0;
    // numerical_scheme
    int flag_PC = 0;
    double Te_rep_eV = 0;
    // solver_setting
    double error_cnv_SOR_Ui = 0;
    int maxITR_SOR_Ui = 0, flag_iter_Ui = 0;
    double error_cnv_SOR_rhoi = 0;
    int maxITR_SOR_rhoi = 0, flag_iter_rhoi = 0;
    double error_cnv_SOR_phi = 0;
    int maxITR_SOR_phi = 0, flag_iter_phi = 0;
    double error_cnv_HES_phi = 0;
    int maxITR_HES_phi = 0;
    double error_cnv_SOR_rhoe = 0;
    int maxITR_SOR_rhoe = 0, flag_iter_rhoe = 0;
    double error_cnv_HES_rhoe = 0;
    int maxITR_HES_rhoe = 0;
    double beta_rhoe = 0, beta_rhoUe = 0;
    double error_cnv_SOR_rhoeps = 0;
    int maxITR_SOR_rhoeps = 0, flag_iter_rhoeps = 0;
    double error_cnv_HES_rhoeps = 0;
    int maxITR_HES_rhoeps = 0;
    // simulation_control
    int flag_adp_dt = 0;
    double dt_ini = 0, CFL = 0;
    int ntime = 0, ndt_i = 0, ndt_m = 0, ndt_n = 0, ndiv_MW = 0, flag_error = 0, flag_autoFinish = 0;
    // output_control
    int flag_chk = 0, flag_gnuRes = 0, ndiv_fout = 0;
    // test_mode
    int flag_impTest = 0;
};

// One key of the table: category, name and where it goes (exactly one of
// the two pointers-to-member is set).
struct SetupKey {
    const char* cat;
    const char* key;
    double SetupParams::*d;
    int SetupParams::*i;
};

// The reference's keys in its reading order; 85 entries.
inline const std::vector<SetupKey>& setup_keys()
{
    using P = SetupParams;
    static const std::vector<SetupKey> keys = [] {
        std::vector<SetupKey> k;
        auto D = [&](const char* c, const char* n, double P::*m) { k.push_back({c, n, m, nullptr}); };
        auto I = [&](const char* c, const char* n, int P::*m) { k.push_back({c, n, nullptr, m}); };
        const char* c = "domain";
        I(c, "ni", &P::ni); I(c, "nj", &P::nj); D(c, "xL", &P::xL); D(c, "xR", &P::xR); D(c, "rmin", &P::rmin);
        D(c, "rmax", &P::rmax); D(c, "width_neutIn", &P::width_neutIn);
        c = "operating_condition";
        D(c, "V_bias", &P::V_bias); D(c, "Q_neutIn_mgs", &P::Q_neutIn_mgs); D(c, "fmw", &P::fmw);
        D(c, "Pmw", &P::Pmw); I(c, "flag_mwRef", &P::flag_mwRef);
        c = "physical_parameter";
        D(c, "masse", &P::masse); D(c, "massi", &P::massi); D(c, "ri", &P::ri); D(c, "Ti", &P::Ti);
        D(c, "Tn", &P::Tn); D(c, "DmN", &P::DmN); D(c, "epsr_diele", &P::epsr_diele);
        c = "transmission_line";
        D(c, "S11_mag", &P::S11_mag); D(c, "S11_arg_deg", &P::S11_arg_deg); D(c, "S21_mag", &P::S21_mag);
        D(c, "S21_arg_deg", &P::S21_arg_deg); D(c, "S12_mag", &P::S12_mag); D(c, "S12_arg_deg", &P::S12_arg_deg);
        D(c, "S22_mag", &P::S22_mag); D(c, "S22_arg_deg", &P::S22_arg_deg); D(c, "Z0", &P::Z0);
        D(c, "h0_stab", &P::h0_stab); D(c, "dh_stab", &P::dh_stab); D(c, "dl_cable", &P::dl_cable);
        D(c, "phase_pre_deg", &P::phase_pre_deg);
        c = "secondary_electron_emission";
        D(c, "coefIISEE_ts", &P::coefIISEE_ts); D(c, "coefMISEE_ts", &P::coefMISEE_ts);
        D(c, "Te_emitSEE_eV", &P::Te_emitSEE_eV); D(c, "ratioEngy_EISEE_rd", &P::ratioEngy_EISEE_rd);
        c = "governing_equation";
        I(c, "flag_inertia", &P::flag_inertia); D(c, "scale_inertia", &P::scale_inertia);
        c = "anomalous_scattering_model";
        I(c, "flag_Sagdeev", &P::flag_Sagdeev); I(c, "flag_Bohm", &P::flag_Bohm); D(c, "alpha_Bohm", &P::alpha_Bohm);
        c = "collisionless_heating_model";
        D(c, "nu_eff", &P::nu_eff);
        c = "initial_condition";
        D(c, "rhon_ini", &P::rhon_ini); D(c, "rhoe_ini", &P::rhoe_ini); D(c, "Te_ini", &P::Te_ini);
        D(c, "J1r_exc", &P::J1r_exc);
        c = "numerical_scheme";
        I(c, "flag_PC", &P::flag_PC); D(c, "Te_rep_eV", &P::Te_rep_eV);
        c = "solver_setting";
        D(c, "error_cnv_SOR_Ui", &P::error_cnv_SOR_Ui); I(c, "maxITR_SOR_Ui", &P::maxITR_SOR_Ui);
        I(c, "flag_iter_Ui", &P::flag_iter_Ui);
        D(c, "error_cnv_SOR_rhoi", &P::error_cnv_SOR_rhoi); I(c, "maxITR_SOR_rhoi", &P::maxITR_SOR_rhoi);
        I(c, "flag_iter_rhoi", &P::flag_iter_rhoi);
        D(c, "error_cnv_SOR_phi", &P::error_cnv_SOR_phi); I(c, "maxITR_SOR_phi", &P::maxITR_SOR_phi);
        I(c, "flag_iter_phi", &P::flag_iter_phi); D(c, "error_cnv_HES_phi", &P::error_cnv_HES_phi);
        I(c, "maxITR_HES_phi", &P::maxITR_HES_phi);
        D(c, "error_cnv_SOR_rhoe", &P::error_cnv_SOR_rhoe); I(c, "maxITR_SOR_rhoe", &P::maxITR_SOR_rhoe);
        I(c, "flag_iter_rhoe", &P::flag_iter_rhoe); D(c, "error_cnv_HES_rhoe", &P::error_cnv_HES_rhoe);
        I(c, "maxITR_HES_rhoe", &P::maxITR_HES_rhoe); D(c, "beta_rhoe", &P::beta_rhoe);
        D(c, "beta_rhoUe", &P::beta_rhoUe);
        D(c, "error_cnv_SOR_rhoeps", &P::error_cnv_SOR_rhoeps); I(c, "maxITR_SOR_rhoeps", &P::maxITR_SOR_rhoeps);
        I(c, "flag_iter_rhoeps", &P::flag_iter_rhoeps); D(c, "error_cnv_HES_rhoeps", &P::error_cnv_HES_rhoeps);
        I(c, "maxITR_HES_rhoeps", &P::maxITR_HES_rhoeps);
        c = "simulation_control";
        I(c, "flag_adp_dt", &P::flag_adp_dt); D(c, "dt_ini", &P::dt_ini); D(c, "CFL", &P::CFL);
        I(c, "ntime", &P::ntime); I(c, "ndt_i", &P::ndt_i); I(c, "ndt_m", &P::ndt_m); I(c, "ndt_n", &P::ndt_n);
        I(c, "ndiv_MW", &P::ndiv_MW); I(c, "flag_error", &P::flag_error);
        I(c, "flag_autoFinish", &P::flag_autoFinish);
        c = "output_control";
        I(c, "flag_chk", &P::flag_chk); I(c, "flag_gnuRes", &P::flag_gnuRes); I(c, "ndiv_fout", &P::ndiv_fout);
        c = "test_mode";
        I(c, "flag_impTest", &P::flag_impTest);
        return k;
    }();
    return keys;
}

// All keys, with the reference's reports: every missing key listed at once,
// a value of the wrong type thrown at once; Te_ini from eV to K with the
// reference's constants (include/constants.hpp:5-6, inputModule.hpp:212-213).
inline SetupParams setup_params_from_yaml(const SetupYaml& y)
{
    SetupParams sp;
    SetupReader r(y);
    for (const SetupKey& k : setup_keys()) {
        if (k.d) r.read(sp.*(k.d), k.cat, k.key);
        else r.read(sp.*(k.i), k.cat, k.key);
    }
    r.finish();
    sp.Te_ini = sp.Te_ini * 1.60217662e-19 / 1.380649e-23;
    return sp;
}

// Grid and solver settings of the implicit step, from setup.yaml:
// domain (ni, nj, xL, xR, rmin, rmax -> dx, dr), operating_condition.V_bias,
// physical_parameter (masse, massi, Ti), secondary_electron_emission,
// solver_setting (maxITR_SOR_*, error_cnv_SOR_*, maxITR_HES_*,
// error_cnv_HES_*) and simulation_control (dt_ini, flag_error, flag_adp_dt, CFL).
// Mobilities, the energy switch, the precision mode and the SEE table are not
// setup.yaml keys (the reference computes mobilities from BOLSIG tables) and
// keep the values already in *p.
inline void ndd_params_from_setup(const SetupYaml& y, cathode_ndd_params* p)
{
    (void)setup_params_from_yaml(y);  // every key of the reference must be there, typed as it reads them
    SetupReader r(y);
    double xL = 0, xR = 0, rmax = 0;
    r.read(p->ni, "domain", "ni");
    r.read(p->nj, "domain", "nj");
    r.read(xL, "domain", "xL");
    r.read(xR, "domain", "xR");
    r.read(p->rmin, "domain", "rmin");
    r.read(rmax, "domain", "rmax");
    r.read(p->V_bias, "operating_condition", "V_bias");
    r.read(p->masse, "physical_parameter", "masse");
    r.read(p->massi, "physical_parameter", "massi");
    r.read(p->Ti, "physical_parameter", "Ti");
    const char* see = "secondary_electron_emission";
    r.read(p->coefIISEE_ts, see, "coefIISEE_ts");
    r.read(p->coefMISEE_ts, see, "coefMISEE_ts");
    r.read(p->Te_emitSEE_eV, see, "Te_emitSEE_eV");
    r.read(p->ratioEngy_EISEE_rd, see, "ratioEngy_EISEE_rd");
    const char* ss = "solver_setting";
    r.read(p->error_cnv_SOR_rhoi, ss, "error_cnv_SOR_rhoi");
    r.read(p->maxITR_SOR_rhoi, ss, "maxITR_SOR_rhoi");
    r.read(p->error_cnv_SOR_phi, ss, "error_cnv_SOR_phi");
    r.read(p->maxITR_SOR_phi, ss, "maxITR_SOR_phi");
    r.read(p->error_cnv_HES_phi, ss, "error_cnv_HES_phi");
    r.read(p->maxITR_HES_phi, ss, "maxITR_HES_phi");
    r.read(p->error_cnv_SOR_rhoe, ss, "error_cnv_SOR_rhoe");
    r.read(p->maxITR_SOR_rhoe, ss, "maxITR_SOR_rhoe");
    r.read(p->error_cnv_HES_rhoe, ss, "error_cnv_HES_rhoe");
    r.read(p->maxITR_HES_rhoe, ss, "maxITR_HES_rhoe");
    r.read(p->error_cnv_SOR_rhoeps, ss, "error_cnv_SOR_rhoeps");
    r.read(p->maxITR_SOR_rhoeps, ss, "maxITR_SOR_rhoeps");
    r.read(p->error_cnv_HES_rhoeps, ss, "error_cnv_HES_rhoeps");
    r.read(p->maxITR_HES_rhoeps, ss, "maxITR_HES_rhoeps");
    r.read(p->dt, "simulation_control", "dt_ini");
    r.read(p->flag_error, "simulation_control", "flag_error");
    r.read(p->flag_adp_dt, "simulation_control", "flag_adp_dt");
    r.read(p->CFL, "simulation_control", "CFL");
    r.finish();
    if (p->ni <= 0 || p->nj <= 0) throw std::runtime_error("[setup.yaml] domain.ni / domain.nj must be positive");
    p->dx = (xR - xL) / p->ni;
    p->dr = (rmax - p->rmin) / p->nj;
    // the reference has no separate ion PC settings (solve_rhoi_constTe is
    // not a PC loop); the ion SG solve uses the electron's
    p->maxITR_HES_rhoi = p->maxITR_HES_rhoe;
    p->error_cnv_HES_rhoi = p->error_cnv_HES_rhoe;
}

}  // namespace cathode_host
