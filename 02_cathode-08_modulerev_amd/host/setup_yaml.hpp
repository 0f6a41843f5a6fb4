// setup_yaml.hpp -- reader for the reference's input/setup.yaml (the keys of
// InputModule's READ_DOUBLE / READ_INT lists, include/inputModule.hpp:151-267)
// and the mapping of the keys this library uses onto cathode_ndd_params.
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

// Grid and solver settings of the implicit step, from setup.yaml:
// domain (ni, nj, xL, xR, rmin, rmax -> dx, dr), operating_condition.V_bias,
// physical_parameter (masse, massi, Ti), secondary_electron_emission,
// solver_setting (maxITR_SOR_*, error_cnv_SOR_*, maxITR_HES_*,
// error_cnv_HES_*) and simulation_control (dt_ini, flag_error).
// Mobilities, the energy switch, the precision mode and the SEE table are not
// setup.yaml keys (the reference computes mobilities from BOLSIG tables) and
// keep the values already in *p.
inline void ndd_params_from_setup(const SetupYaml& y, cathode_ndd_params* p)
{
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
