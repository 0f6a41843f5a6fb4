// host_check.cpp -- checks of the C++ host headers (02_cathode-08_modulerev_amd/host).
//
//   host_check cpu <tmpdir> <setup.yaml>   CmpMat2 / S<->ABCD, setup.yaml reader, restart CSV
//   host_check tables <rateCoef_e> <rateCoef_i> <coefEISEE> <out>
//                             the rate and yield tables read by host/tables.hpp,
//                             written to <out> as "name n v0 v1 ..." lines (%.17g)
//   host_check result <dir> <out.csv>
//                             the result CSV of host/result_csv.hpp from the fields
//                             in <dir> (<name>.bin float64, <name>.cbin complex128,
//                             grid.txt "key value" lines, eloss_n.bin / eloss_m.bin)
//   host_check gpu            solver_SMG / solver_SOR / solver_LU_9p through the
//                             drop-in header against the oracle (liboracle.so)
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <cstring>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../02_cathode-08_modulerev_amd/host/cmpMat2.hpp"
#include "../../02_cathode-08_modulerev_amd/host/restart_csv.hpp"
#include "../../02_cathode-08_modulerev_amd/host/result_csv.hpp"
#include "../../02_cathode-08_modulerev_amd/host/setup_yaml.hpp"
#include "../../02_cathode-08_modulerev_amd/host/solverFuncs.hpp"
#include "../../02_cathode-08_modulerev_amd/host/tables.hpp"
#include "../../oracle/oracle.h"

using namespace cathode_host;
using Grid = std::vector<std::vector<double> >;

static int fails = 0;
#define CHECK(cond, ...)                                 \
    do {                                                 \
        if (!(cond)) {                                   \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                    \
            std::printf("\n");                           \
            fails++;                                     \
            return;                                      \
        }                                                \
    } while (0)

static void check_abcd()
{
    // round trip against the oracle's restatement and the identity
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> u(-0.6, 0.6);
    for (int t = 0; t < 20; t++) {
        cplx S11(u(g), u(g)), S21(0.5 + u(g) * 0.3, u(g)), S12(0.5 + u(g) * 0.3, u(g)), S22(u(g), u(g));
        CmpMat2 T;
        sParamToABCD(T, S11, S21, S12, S22, 50.0);
        double s8[8] = {S11.real(), S11.imag(), S12.real(), S12.imag(), S21.real(), S21.imag(), S22.real(), S22.imag()};
        double t8[8];
        orc_s_to_abcd(s8, 50.0, t8);
        const cplx ref[4] = {{t8[0], t8[1]}, {t8[2], t8[3]}, {t8[4], t8[5]}, {t8[6], t8[7]}};
        const cplx got[4] = {T.A, T.B, T.C, T.D};
        for (int k = 0; k < 4; k++) CHECK(std::abs(got[k] - ref[k]) <= 1e-12 * (1 + std::abs(ref[k])), "abcd %d", k);
        cplx a, b, c, d;
        ABCDToSParam(a, b, c, d, T, 50.0);
        CHECK(std::abs(a - S11) < 1e-12 && std::abs(b - S21) < 1e-12 && std::abs(c - S12) < 1e-12 &&
                  std::abs(d - S22) < 1e-12,
              "S round trip");
        // reciprocal network (S12 = S21) <=> det(ABCD) = 1
        CmpMat2 R;
        sParamToABCD(R, S11, S21, S21, S22, 50.0);
        CHECK(std::abs(R.det() - 1.0) < 1e-12, "reciprocal det");
        CmpMat2 I = R * R.inv();
        CHECK(std::abs(I.A - 1.0) < 1e-12 && std::abs(I.B) < 1e-9 && std::abs(I.C) < 1e-12 && std::abs(I.D - 1.0) < 1e-12,
              "inverse");
    }
    // a matched lossless line: S11 = S22 = 0, S21 = exp(-j theta)
    const double th = 0.7;
    cplx S11, S21, S12, S22;
    ABCDToSParam(S11, S21, S12, S22, line_abcd(th, 50.0), 50.0);
    CHECK(std::abs(S11) < 1e-14 && std::abs(S22) < 1e-14 && std::abs(S21 - std::exp(cplx(0, -th))) < 1e-14, "line");
    // no stub change and no cable: the S-parameters are unchanged
    SParams s0{{0.2, -0.1}, {0.8, 0.1}, {0.79, 0.12}, {0.1, 0.05}};
    SParams s1 = stub_cable_correction(s0, 50.0, 2 * M_PI * 4.25e9, 2.08, 0.25, 0.0, 0.0, 4605.3);
    CHECK(std::abs(s1.S11 - s0.S11) < 1e-12 && std::abs(s1.S21 - s0.S21) < 1e-12 && std::abs(s1.S22 - s0.S22) < 1e-12,
          "zero correction");
    // a cable only rotates S21, S12 by beta dl and S22 by twice that
    SParams s2 = stub_cable_correction(s0, 50.0, 2 * M_PI * 4.25e9, 2.08, 0.25, 0.0, 8e-3, 0.0);
    const double beta = 2 * M_PI * 4.25e9 * std::sqrt(8.8541878128e-12 * 2.08 * 1.2566370614e-6);
    CHECK(std::abs(s2.S21 - s0.S21 * std::exp(cplx(0, -beta * 8e-3))) < 1e-12 &&
              std::abs(s2.S22 - s0.S22 * std::exp(cplx(0, -2 * beta * 8e-3))) < 1e-12,
          "cable");
    bool threw = false;
    try {
        CmpMat2 T;
        sParamToABCD(T, 0.1, 0.0, 0.1, 0.1, 50.0);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw, "S21 = 0 must throw");
    std::printf("ok abcd\n");
}

// every key of the reference's reader, from a file in the shipped layout
static void check_yaml(const std::string& path)
{
    std::ifstream f(path);
    CHECK((bool)f, "cannot open %s", path.c_str());
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string text = buf.str();
    std::istringstream in(text);
    const SetupYaml y = parse_setup_yaml(in);
    CHECK(setup_keys().size() == 85, "the reference reads 85 keys, table has %zu", setup_keys().size());
    int cats = 0;
    for (const auto& c : y.cat) cats += c.second.empty() ? 0 : 1;
    CHECK(cats == 14, "14 categories, found %d", cats);
    const SetupParams sp = setup_params_from_yaml(y);
    CHECK(sp.ni == 96 && sp.nj == 64 && sp.xR == 0.0096 && sp.rmax == 0.0064 && sp.width_neutIn == 0.002, "domain");
    CHECK(sp.V_bias == 30.0 && sp.Pmw == 10.0 && sp.fmw == 4.25e9 && sp.flag_mwRef == 1, "operating_condition");
    CHECK(sp.S11_mag == 0.25 && sp.S11_arg_deg == -550.0 && sp.Z0 == 50.0 && sp.phase_pre_deg == 4600.0,
          "transmission_line: first of a repeated key kept");
    CHECK(sp.flag_Bohm == 1 && sp.flag_Sagdeev == 0 && sp.alpha_Bohm == 16.0 && sp.nu_eff == 2.0e9, "models");
    CHECK(std::fabs(sp.Te_ini - 5.0 * 1.60217662e-19 / 1.380649e-23) < 1e-9, "Te_ini eV -> K");
    CHECK(sp.flag_PC == 1 && sp.Te_rep_eV == 0.1 && sp.beta_rhoUe == 1.0 && sp.maxITR_HES_rhoeps == 2, "solver");
    CHECK(sp.flag_adp_dt == 0 && sp.dt_ini == 1e-10 && sp.CFL == 1.0 && sp.ntime == 200 && sp.ndiv_MW == 1000 &&
          sp.flag_error == 3 && sp.ndiv_fout == 2000 && sp.flag_impTest == 0, "control");
    cathode_ndd_params p{};
    ndd_params_from_setup(y, &p);
    CHECK(p.ni == 96 && p.nj == 64 && std::fabs(p.dx - 1e-4) < 1e-18 && std::fabs(p.dr - 1e-4) < 1e-18, "grid");
    CHECK(p.V_bias == 30.0 && p.maxITR_SOR_phi == 30 && p.error_cnv_SOR_rhoe == 1e-10 && p.maxITR_HES_rhoeps == 2,
          "solver keys");
    CHECK(p.dt == 1e-10 && p.flag_error == 3 && p.coefIISEE_ts == 0.08, "control keys");
    // every key missing at once is reported, as the reference does
    auto drop = [&](const std::string& t, const std::string& line) {
        std::string b(t);
        const size_t at = b.find(line);
        if (at != std::string::npos) b.erase(at, b.find('\n', at) - at + 1);
        return b;
    };
    bool threw = false;
    try {
        std::istringstream in2(drop(drop(text, "  V_bias:"), "  flag_impTest:"));
        cathode_ndd_params q{};
        ndd_params_from_setup(parse_setup_yaml(in2), &q);
    } catch (const std::runtime_error& e) {
        const std::string m = e.what();
        threw = m.find("Missing YAML parameters") != std::string::npos &&
                m.find("operating_condition.V_bias") != std::string::npos &&
                m.find("test_mode.flag_impTest") != std::string::npos;
    }
    CHECK(threw, "missing keys");
    // a key the step does not use is still required
    threw = false;
    try {
        std::istringstream in3(drop(text, "  nu_eff:"));
        (void)setup_params_from_yaml(parse_setup_yaml(in3));
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find("collisionless_heating_model.nu_eff") != std::string::npos;
    }
    CHECK(threw, "unused key still required");
    std::string bad2(text);
    bad2.replace(bad2.find("  nj: 64"), 8, "  nj: 6.4");
    std::istringstream in4(bad2);
    threw = false;
    try {
        cathode_ndd_params q{};
        ndd_params_from_setup(parse_setup_yaml(in4), &q);
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()) == "Type conversion error: domain.nj should be int";
    }
    CHECK(threw, "type error");
    std::printf("ok yaml\n");
}

static void check_restart(const std::string& dir)
{
    const int ni = 7, nj = 5;
    RestartTable T = make_restart_table(ni, nj, 0.0, 1e-4, 0.0, 2e-4);
    const size_t n = T.cells();
    std::vector<double> a[6];
    std::mt19937_64 g(9);
    std::uniform_real_distribution<double> u(0.5, 2.0);
    for (auto& v : a) {
        v.resize(n);
        for (auto& x : v) x = u(g) * 1e17;
    }
    for (auto& x : a[3]) x = u(g) * 5e4;  // Te [K]
    a[0][3] = 1e-120;                     // written as 0
    cathode_ndd_fields f{};
    f.rhoi = a[0].data();
    f.rhoe = a[1].data();
    f.phi = a[2].data();
    f.Te = a[3].data();
    f.Gx = a[4].data();
    f.rhoeps = a[5].data();
    fields_to_restart(f, T);
    T.col("Uix")[7] = 3.25;  // a column outside the NDD state survives the round trip
    const std::string path = dir + "/restart1.csv";
    write_restart_csv(path, T);
    RestartTable R = read_restart_csv(path);
    CHECK(R.ni == ni && R.nj == nj && unknown_columns(R).empty(), "shape");
    CHECK((*R.find("Uix"))[7] == 3.25 && (*R.find("x"))[(nj + 2) * 2] == (*T.find("x"))[(nj + 2) * 2] &&
              std::fabs((*R.find("x"))[(nj + 2) * 2] - 1.5e-4) < 1e-19,
          "other columns");
    std::vector<double> b[6];
    for (auto& v : b) v.assign(n, -1.0);
    cathode_ndd_fields h{};
    h.rhoi = b[0].data();
    h.rhoe = b[1].data();
    h.phi = b[2].data();
    h.Te = b[3].data();
    h.Gx = b[4].data();
    h.rhoeps = b[5].data();
    restart_to_fields(R, h);
    for (int k = 0; k < 6; k++)
        for (size_t q = 0; q < n; q++) {
            const double want = (k == 0 && q == 3) ? 0.0 : a[k][q];
            // exact except Te (two roundings of the eV <-> K conversion)
            const double tol = k == 3 ? 4e-16 * std::fabs(want) : 0.0;
            CHECK(std::fabs(b[k][q] - want) <= tol, "field %d cell %zu: %.17g vs %.17g", k, q, b[k][q], want);
        }
    // header line is the reference's
    std::FILE* fp = std::fopen(path.c_str(), "r");
    char line[2048];
    CHECK(fp && std::fgets(line, sizeof line, fp), "header read");
    std::fclose(fp);
    CHECK(std::strncmp(line, "i,j,x,r,rhoi,rhoi_old,Uix,Uix_old,", 34) == 0 &&
              std::strstr(line, ",Gx_wall,Gr_wall,zero") != nullptr,
          "header");
    std::printf("ok restart\n");
}

// ---------------------------------------------------------------- GPU part
static Grid grid(int ni2, int nj2, const std::vector<double>& flat)
{
    Grid g(ni2, std::vector<double>(nj2));
    for (int i = 0; i < ni2; i++)
        for (int j = 0; j < nj2; j++) g[i][j] = flat[(size_t)i * nj2 + j];
    return g;
}

static void check_dropin_gpu()
{
    const int ni = 40, nj = 24, ni2 = ni + 2, nj2 = nj + 2;
    const size_t n = (size_t)ni2 * nj2;
    std::mt19937_64 g(17);
    std::uniform_real_distribution<double> u(0.5, 2.0), v(-1.0, 1.0);
    std::vector<double> c[10], b(n);
    for (int k = 0; k < 10; k++) c[k].resize(n);
    for (size_t q = 0; q < n; q++) {
        for (int k = 1; k < 5; k++) c[k][q] = u(g);
        for (int k = 5; k < 9; k++) c[k][q] = 0.2 * v(g);
        c[0][q] = c[1][q] + c[2][q] + c[3][q] + c[4][q] + 0.01;
        c[9][q] = c[0][q] + std::fabs(c[5][q]) + std::fabs(c[6][q]) + std::fabs(c[7][q]) + std::fabs(c[8][q]);
        b[q] = v(g);
    }
    for (int j = 0; j < nj2; j++) {  // Dirichlet fold on the left wall
        c[0][(size_t)nj2 + j] += c[2][(size_t)nj2 + j];
        c[2][(size_t)nj2 + j] = 0;
    }
    // two boxes (cathode-like notch)
    std::vector<std::vector<int> > i_bl = {{1, 20}, {21, ni}}, j_bl = {{8, nj}, {1, nj}};
    const int boxes[8] = {1, 20, 8, nj, 21, ni, 1, nj};
    // SMG drop-in vs the oracle's exact solution
    Grid sol(ni2, std::vector<double>(nj2, 4.0));
    solver_SMG(grid(ni2, nj2, c[0]), grid(ni2, nj2, c[1]), grid(ni2, nj2, c[2]), grid(ni2, nj2, c[3]),
               grid(ni2, nj2, c[4]), grid(ni2, nj2, b), i_bl, j_bl, 2, 30, 1e-10, 1, sol);
    std::vector<double> xd(n, 4.0);
    CHECK(orc_struct5_direct(c[0].data(), c[1].data(), c[2].data(), c[3].data(), c[4].data(), b.data(), boxes, 2, ni2,
                             nj2, nj2, xd.data()) == 0,
          "oracle direct");
    double num = 0, den = 0;
    for (int i = 0; i < ni2; i++)
        for (int j = 0; j < nj2; j++) {
            const double d = sol[i][j] - xd[(size_t)i * nj2 + j];
            num += d * d;
            den += xd[(size_t)i * nj2 + j] * xd[(size_t)i * nj2 + j];
            if (xd[(size_t)i * nj2 + j] == 4.0) CHECK(sol[i][j] == 4.0, "off-box cell written");
        }
    CHECK(std::sqrt(num / den) < 1e-8, "SMG rel L2 %.3e", std::sqrt(num / den));
    cathode_status st;
    cathode_last_status(&st);
    CHECK(st.converged == 1 && st.rel_residual <= 1e-10, "SMG status");
    // SOR drop-in: bit for bit with the oracle's sequential sweep
    Grid s2(ni2, std::vector<double>(nj2, 0.1));
    std::vector<double> xo(n, 0.1);
    solver_SOR(grid(ni2, nj2, c[0]), grid(ni2, nj2, c[1]), grid(ni2, nj2, c[2]), grid(ni2, nj2, c[3]),
               grid(ni2, nj2, c[4]), grid(ni2, nj2, b), i_bl, j_bl, 2, 1.7, 12, 1e-12, 1, s2);
    double err = 0;
    orc_sor_solve(c[0].data(), c[1].data(), c[2].data(), c[3].data(), c[4].data(), b.data(), boxes, 2, ni2, nj2, nj2,
                  1.7, 12, 1e-12, xo.data(), &err);
    for (int i = 0; i < ni2; i++)
        for (int j = 0; j < nj2; j++) CHECK(s2[i][j] == xo[(size_t)i * nj2 + j], "SOR bit-exact at %d %d", i, j);
    // LU_9p drop-in vs the oracle's banded LU
    Grid s3(ni2, std::vector<double>(nj2, 0.0));
    solver_LU_9p(grid(ni2, nj2, c[9]), grid(ni2, nj2, c[1]), grid(ni2, nj2, c[2]), grid(ni2, nj2, c[3]),
                 grid(ni2, nj2, c[4]), grid(ni2, nj2, c[5]), grid(ni2, nj2, c[6]), grid(ni2, nj2, c[7]),
                 grid(ni2, nj2, c[8]), grid(ni2, nj2, b), i_bl, j_bl, 2, 1, s3);
    std::vector<double> x9(n, 0.0);
    CHECK(orc_lu9_direct(c[9].data(), c[1].data(), c[2].data(), c[3].data(), c[4].data(), c[5].data(), c[6].data(),
                         c[7].data(), c[8].data(), b.data(), boxes, 2, ni2, nj2, nj2, x9.data()) == 0,
          "oracle lu9");
    num = den = 0;
    for (int i = 0; i < ni2; i++)
        for (int j = 0; j < nj2; j++) {
            const double d = s3[i][j] - x9[(size_t)i * nj2 + j];
            num += d * d;
            den += x9[(size_t)i * nj2 + j] * x9[(size_t)i * nj2 + j];
        }
    CHECK(std::sqrt(num / den) < 1e-11, "LU_9p rel L2 %.3e", std::sqrt(num / den));
    std::printf("ok dropin\n");
}

static void dump_tables(const char* pe, const char* pi, const char* psee, const char* out)
{
    const RateTables T = read_rate_tables(pe, pi);
    cathode_ndd_params p{};
    read_eisee_table(psee, p);
    FILE* f = std::fopen(out, "w");
    auto put = [&](const char* name, const double* v, size_t n) {
        std::fprintf(f, "%s %zu", name, n);
        for (size_t k = 0; k < n; k++) std::fprintf(f, " %.17g", v[k]);
        std::fprintf(f, "\n");
    };
    const std::vector<double>* tabs[11] = {&T.nu_elas, &T.nu_exc, &T.nu_excReso, &T.nu_excMeta, &T.nu_ionz,
                                           &T.nu_ionzStep, &T.nu_excStep, &T.nu_super, &T.eloss_n, &T.eloss_m,
                                           &T.nui_m};
    const char* names[11] = {"nu_elas", "nu_exc", "nu_excReso", "nu_excMeta", "nu_ionz", "nu_ionzStep", "nu_excStep",
                             "nu_super", "eloss_n", "eloss_m", "nui_m"};
    for (int q = 0; q < 11; q++) put(names[q], tabs[q]->data(), tabs[q]->size());
    put("dTe", &T.dTe, 1);
    put("dEi", &T.dEi, 1);
    put("see_eb", p.see_eb, (size_t)p.n_see);
    put("see_rd", p.see_rd, (size_t)p.n_see);
    put("see_ts", p.see_ts, (size_t)p.n_see);
    put("dTe_SEE", &p.dTe_SEE, 1);
    std::fclose(f);
    const cathode_rate_tables v = T.view();
    CHECK(v.n_e == (int)T.nu_elas.size() && v.n_i == (int)T.nui_m.size() && v.nu_super == T.nu_super.data(),
          "view");
    std::printf("ok tables\n");
}

template <class T>
static std::vector<T> read_bin(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    std::vector<T> v;
    if (!in) return v;
    in.seekg(0, std::ios::end);
    const size_t bytes = (size_t)in.tellg();
    in.seekg(0);
    v.resize(bytes / sizeof(T));
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

static void write_result(const std::string& dir, const std::string& out)
{
    ResultGrid g;
    std::ifstream gin(dir + "/grid.txt");
    std::string key;
    double v;
    while (gin >> key >> v) {
        if (key == "ni") g.ni = (int)v;
        else if (key == "nj") g.nj = (int)v;
        else if (key == "xL") g.xL = v;
        else if (key == "dx") g.dx = v;
        else if (key == "rmin") g.rmin = v;
        else if (key == "dr") g.dr = v;
        else if (key == "dt") g.dt = v;
        else if (key == "ndt_i") g.ndt_i = (int)v;
        else if (key == "ndt_m") g.ndt_m = (int)v;
        else if (key == "ndt_n") g.ndt_n = (int)v;
        else if (key == "masse") g.masse = v;
        else if (key == "massi") g.massi = v;
        else if (key == "nu_eff") g.nu_eff = v;
        else if (key == "deltaECR") g.deltaECR = v;
        else if (key == "dTe") g.dTe = v;
        else if (key == "n_bl") g.n_bl = (int)v;
        else if (key == "box") g.cell_box.push_back((int)v);
    }
    std::vector<std::vector<double> > keepr;
    std::vector<std::vector<std::complex<double> > > keepc;
    ResultFields F;
    const std::vector<std::string> real = {
        "rhoe", "rhoe_old", "rhoi", "rhoi_old", "Uix", "Uir", "Uip", "rhoUix", "rhoUir", "rhoUix_wall",
        "rhoUir_wall", "rhoUex", "rhoUer", "Te", "rhoeps", "rhoeps_old", "Gx", "Gr", "phi", "Ex", "Er",
        "rate_ionize", "nu_m", "nu_m1", "nu_ionz", "nu_ionzStep", "nu_excStep", "nu_exc", "nu_excMeta", "nui_m",
        "mu_para", "mu_perp", "mu_xx", "mu_rr", "mu_xr", "Ap", "Bx", "Br", "Pabs", "rate_eloss", "rhom",
        "rhom_old", "rhoUmx", "rhoUmr", "rhon", "rhon_old", "rhoUnx", "rhoUnr", "distECR", "scx", "scr", "nUex",
        "nUer", "nu_super", "nu_ano", "nu_ano_IAT", "jdgBnd_flc", "jdgBnd_Ep"};
    keepr.reserve(real.size());
    for (const std::string& n : real) {
        keepr.push_back(read_bin<double>(dir + "/" + n + ".bin"));
        if (!keepr.back().empty()) F.real[n] = keepr.back().data();
    }
    const std::vector<std::string> cplx = {"E1x", "E1r", "E1p", "J1x", "J1r", "J1p"};
    keepc.reserve(cplx.size());
    for (const std::string& n : cplx) {
        keepc.push_back(read_bin<std::complex<double> >(dir + "/" + n + ".cbin"));
        if (!keepc.back().empty()) F.cplx[n] = keepc.back().data();
    }
    F.eloss_n = read_bin<double>(dir + "/eloss_n.bin");
    F.eloss_m = read_bin<double>(dir + "/eloss_m.bin");
    write_result_csv(out, g, F);
    std::printf("ok result\n");
}

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "cpu") {
        check_abcd();
        check_yaml(argc > 3 ? argv[3] : "tests/golden/setup_reference_layout.yaml");
        check_restart(argc > 2 ? argv[2] : ".");
    } else if (mode == "result" && argc > 3) {
        write_result(argv[2], argv[3]);
    } else if (mode == "tables" && argc > 5) {
        dump_tables(argv[2], argv[3], argv[4], argv[5]);
    } else if (mode == "gpu") {
        check_dropin_gpu();
    } else if (mode == "noinit") {
        // without a device the drop-in throws instead of returning garbage
        bool threw = false;
        try {
            Grid z(5, std::vector<double>(5, 0.0));
            std::vector<std::vector<int> > ib = {{1, 3}}, jb = {{1, 3}};
            solver_SMG(z, z, z, z, z, z, ib, jb, 1, 10, 1e-10, 0, z);
        } catch (const std::runtime_error& e) {
            threw = std::strstr(e.what(), "cathode_init") != nullptr;
        }
        if (!threw) {
            std::printf("FAIL noinit\n");
            return 1;
        }
        std::printf("ok noinit\n");
    }
    return fails ? 1 : 0;
}
