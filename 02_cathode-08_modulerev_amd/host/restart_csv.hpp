// restart_csv.hpp -- the reference's restart CSV (results/restart<n>.csv,
// written by include/outputModule.hpp:428-468, read back by
// InputModule::input_restart_data, include/inputModule.hpp:667-1150), kept
// column for column so restart files move between the reference and this
// library unchanged, and its mapping onto the NDD step's host fields.
//
// Layout: one header line of 55 names (below), then one row per cell for
// i = 0..ni+1 (outer), j = 0..nj+1 (inner): i and j as integers, every other
// value in scientific notation with max_digits10 significant digits, values
// of magnitude below 1e-100 written as 0 (miniFuncs.hpp:99-107, not applied
// to x and r).  Te columns are in eV (the state keeps Kelvin).
// Reading: header names are trimmed of blanks and CR; i, j, x, r are
// required; other columns are matched by name, unknown ones are kept (and
// listed by unknown_columns()), absent ones read as 0.
#pragma once

#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cathode_hip.h"

namespace cathode_host {

inline const std::vector<std::string>& restart_columns()
{
    static const std::vector<std::string> c = {
        "i", "j", "x", "r", "rhoi", "rhoi_old", "Uix", "Uix_old", "Uir", "Uir_old", "Uip", "Uip_old", "rhoe",
        "rhoe_old", "rhoUex", "rhoUex_old", "rhoUer", "rhoUer_old", "Te", "Te_old", "Gx", "Gx_old", "Gr", "Gr_old",
        "phi", "phi_old", "rhom", "rhom_old", "rhoUmx", "rhoUmr", "rate_ionize", "scx", "scr", "nUex", "nUex_old",
        "nUer", "nUer_old", "rhon", "rhon_old", "rhoUnx", "rhoUnr", "rhoeps", "rhoeps_old", "rhoUix_wall",
        "rhoUir_wall", "rhoUex_wall", "rhoUer_wall", "rhoUmx_wall", "rhoUmr_wall", "rhoUnx_wall", "rhoUnr_wall",
        "Gx_wall", "Gr_wall", "zero"};
    return c;
}

struct RestartTable {
    int ni = 0, nj = 0;
    std::vector<std::string> cols;
    std::map<std::string, std::vector<double> > data;  // column -> (ni+2)(nj+2) values, [i][j] row-major

    size_t cells() const { return (size_t)(ni + 2) * (nj + 2); }
    std::vector<double>& col(const std::string& n)
    {
        auto it = data.find(n);
        if (it == data.end()) {
            cols.push_back(n);
            it = data.emplace(n, std::vector<double>(cells(), 0.0)).first;
        }
        return it->second;
    }
    const std::vector<double>* find(const std::string& n) const
    {
        auto it = data.find(n);
        return it == data.end() ? nullptr : &it->second;
    }
};

// Empty table in the reference layout with the cell coordinates
// x_i = xL + (i - 1/2) dx, r_j = rmin + (j - 1/2) dr (initialModule.hpp:83-88).
inline RestartTable make_restart_table(int ni, int nj, double xL, double dx, double rmin, double dr)
{
    RestartTable T;
    T.ni = ni;
    T.nj = nj;
    for (const auto& n : restart_columns()) T.col(n);
    for (int i = 0; i < ni + 2; i++)
        for (int j = 0; j < nj + 2; j++) {
            const size_t k = (size_t)i * (nj + 2) + j;
            T.data["i"][k] = i;
            T.data["j"][k] = j;
            T.data["x"][k] = xL + (i - 0.5) * dx;
            T.data["r"][k] = rmin + (j - 0.5) * dr;
        }
    return T;
}

inline void write_restart_csv(const std::string& path, const RestartTable& T)
{
    std::ofstream f(path);
    if (!f) throw std::runtime_error("[restart.csv] cannot write " + path);
    const auto& cols = restart_columns();
    for (size_t c = 0; c < cols.size(); c++) f << (c ? "," : "") << cols[c];
    f << "\n" << std::setprecision(std::numeric_limits<double>::max_digits10) << std::scientific;
    std::vector<const std::vector<double>*> src(cols.size());
    for (size_t c = 0; c < cols.size(); c++) src[c] = T.find(cols[c]);
    for (int i = 0; i < T.ni + 2; i++)
        for (int j = 0; j < T.nj + 2; j++) {
            const size_t k = (size_t)i * (T.nj + 2) + j;
            f << i << "," << j;
            for (size_t c = 2; c < cols.size(); c++) {
                double v = src[c] ? (*src[c])[k] : 0.0;
                if (c >= 4 && std::fabs(v) < 1e-100) v = 0.0;
                f << "," << v;
            }
            f << "\n";
        }
}

inline RestartTable read_restart_csv(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("[restart.csv] File not found: " + path);
    std::string line;
    if (!std::getline(f, line)) throw std::runtime_error("[restart.csv] Empty file.");
    auto trim = [](const std::string& s) {
        const char* ws = " \t\r\n";
        const size_t a = s.find_first_not_of(ws);
        return a == std::string::npos ? std::string() : s.substr(a, s.find_last_not_of(ws) - a + 1);
    };
    std::vector<std::string> header;
    {
        std::stringstream ss(line);
        std::string n;
        while (std::getline(ss, n, ',')) header.push_back(trim(n));
    }
    int ci = -1, cj = -1;
    for (size_t c = 0; c < header.size(); c++) {
        if (header[c] == "i") ci = (int)c;
        if (header[c] == "j") cj = (int)c;
    }
    for (const char* req : {"i", "j", "x", "r"}) {
        bool ok = false;
        for (const auto& h : header) ok = ok || h == req;
        if (!ok) throw std::runtime_error(std::string("[restart.csv] Required column is missing: ") + req);
    }
    std::vector<std::vector<double> > rows;
    int lineno = 1, imax = -1, jmax = -1;
    while (std::getline(f, line)) {
        lineno++;
        if (trim(line).empty()) continue;
        std::vector<double> v;
        std::stringstream ss(line);
        std::string t;
        while (std::getline(ss, t, ',')) v.push_back(std::stod(t));
        if (v.size() != header.size())
            throw std::runtime_error("[restart.csv] Malformed row at line " + std::to_string(lineno) + ".");
        imax = std::max(imax, (int)v[ci]);
        jmax = std::max(jmax, (int)v[cj]);
        rows.push_back(std::move(v));
    }
    RestartTable T;
    T.ni = imax - 1;
    T.nj = jmax - 1;
    if (T.ni < 1 || T.nj < 1) throw std::runtime_error("[restart.csv] no interior cells");
    for (const auto& h : header) T.col(h);
    for (const auto& v : rows) {
        const int i = (int)v[ci], j = (int)v[cj];
        if (i < 0 || j < 0) throw std::runtime_error("[restart.csv] negative cell index");
        const size_t k = (size_t)i * (T.nj + 2) + j;
        for (size_t c = 0; c < header.size(); c++) T.data[header[c]][k] = v[c];
    }
    return T;
}

// columns of the file that neither the layout nor the NDD state uses
inline std::vector<std::string> unknown_columns(const RestartTable& T)
{
    std::vector<std::string> out;
    for (const auto& c : T.cols) {
        bool known = false;
        for (const auto& k : restart_columns()) known = known || k == c;
        if (!known) out.push_back(c);
    }
    return out;
}

constexpr double kBoltz = 1.380649e-23, kE0 = 1.60217662e-19;

// (column, field) pairs the NDD state shares with the restart file; Te is
// converted eV <-> K.
struct RestartField {
    const char* column;
    double* cathode_ndd_fields::*member;
};
inline const std::vector<RestartField>& restart_fields()
{
    static const std::vector<RestartField> m = {
        {"rhoi", &cathode_ndd_fields::rhoi},      {"rhoe", &cathode_ndd_fields::rhoe},
        {"rhoe_old", &cathode_ndd_fields::rhoe_old}, {"Te", &cathode_ndd_fields::Te},
        {"phi", &cathode_ndd_fields::phi},        {"rhoeps", &cathode_ndd_fields::rhoeps},
        {"rhoUex", &cathode_ndd_fields::rhoUex},  {"rhoUer", &cathode_ndd_fields::rhoUer},
        {"nUex", &cathode_ndd_fields::nUex},      {"nUer", &cathode_ndd_fields::nUer},
        {"Gx", &cathode_ndd_fields::Gx},          {"Gr", &cathode_ndd_fields::Gr},
        {"rate_ionize", &cathode_ndd_fields::rate_ionize}};
    return m;
}

// restart table -> host fields ((ni+2)(nj+2) arrays; NULL members skipped)
inline void restart_to_fields(const RestartTable& T, cathode_ndd_fields& f)
{
    for (const auto& m : restart_fields()) {
        double* dst = f.*(m.member);
        const std::vector<double>* src = T.find(m.column);
        if (!dst || !src) continue;
        const bool te = std::string(m.column) == "Te";
        for (size_t k = 0; k < T.cells(); k++) dst[k] = te ? (*src)[k] * kE0 / kBoltz : (*src)[k];
    }
}

// host fields -> restart table (other columns untouched)
inline void fields_to_restart(const cathode_ndd_fields& f, RestartTable& T)
{
    for (const auto& m : restart_fields()) {
        const double* src = f.*(m.member);
        if (!src) continue;
        std::vector<double>& dst = T.col(m.column);
        const bool te = std::string(m.column) == "Te";
        for (size_t k = 0; k < T.cells(); k++) dst[k] = te ? src[k] * kBoltz / kE0 : src[k];
    }
}

}  // namespace cathode_host
