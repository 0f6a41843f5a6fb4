// solverFuncs.hpp -- drop-in replacement of include/solverFuncs.hpp.
//
// solver_SMG, solver_SOR and solver_LU_9p keep the reference's exact C++
// signatures (include/solverFuncs.hpp:26-36, :169-179, :223-237): arguments by
// value as vector<vector<double>> / vector<vector<int>>, the solution by
// reference, updated on the cells of the n_bl boxes only.  Each call flattens
// the arrays into one contiguous (ni+2) x (nj+2) staging block and hands it to
// libcathode_hip.so (include/cathode_hip.h), which solves on the GPU:
//   solver_SMG   -> cathode_struct5_solve  (multigrid-preconditioned BiCGStab,
//                   stop at ||b - A x|| / ||b|| <= error_cnv or maxITR,
//                   zero initial guess, as HYPRE SMG)
//   solver_SOR   -> cathode_sor_solve      (the same sweeps, bit for bit)
//   solver_LU_9p -> cathode_lu9_solve      (9-point, to relative residual 1e-13)
// flag_msg == 1 prints the lines the reference prints.  The reference returns
// void and reports nothing; here a failure of the GPU library (no device,
// HIP error, non-finite residual) throws std::runtime_error instead of
// returning a wrong answer.  Non-convergence stays silent, as in the
// reference; cathode_last_status() reports it.
//
// Precision: FP64 by default; CATHODE_SOLVER_MODE=mixed in the environment
// selects FP32 V-cycles with FP64 refinement (same stopping test).
// The library context (device 0, or CATHODE_DEVICE) is created on the first
// call and kept until exit; calls are serialised inside the library.
#pragma once

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cathode_hip.h"

namespace cathode_host {

inline void check(int rc, const char* what)
{
    if (rc != CATHODE_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + cathode_last_error());
}

// Lazily initialised library context (one per process).
inline void ensure_init()
{
    static bool ready = false;
    if (ready) return;
    const char* d = std::getenv("CATHODE_DEVICE");
    check(cathode_init(d ? std::atoi(d) : 0), "cathode_init");
    ready = true;
    std::atexit([] { cathode_finalize(); });
}

inline int solver_mode()
{
    const char* m = std::getenv("CATHODE_SOLVER_MODE");
    return (m && std::strcmp(m, "mixed") == 0) ? CATHODE_MODE_MIXED : CATHODE_MODE_FP64;
}

// vector<vector<double>> (ni+2) x (nj+2) -> contiguous rows of length ld = nj+2
inline void flatten(const std::vector<std::vector<double> >& a, int ni2, int nj2, std::vector<double>& out)
{
    if ((int)a.size() != ni2) throw std::invalid_argument("solverFuncs: array row count differs from solution");
    out.resize((size_t)ni2 * nj2);
    for (int i = 0; i < ni2; i++) {
        if ((int)a[i].size() != nj2) throw std::invalid_argument("solverFuncs: array row length differs from solution");
        std::memcpy(out.data() + (size_t)i * nj2, a[i].data(), nj2 * sizeof(double));
    }
}

// i_bl[b] = {ilo, ihi}, j_bl[b] = {jlo, jhi} -> boxes[4 b + 0..3]
inline std::vector<int> boxes_of(const std::vector<std::vector<int> >& i_bl, const std::vector<std::vector<int> >& j_bl,
                                 int n_bl)
{
    std::vector<int> bx(4 * (size_t)n_bl);
    for (int b = 0; b < n_bl; b++) {
        bx[4 * b + 0] = i_bl.at(b).at(0);
        bx[4 * b + 1] = i_bl.at(b).at(1);
        bx[4 * b + 2] = j_bl.at(b).at(0);
        bx[4 * b + 3] = j_bl.at(b).at(1);
    }
    return bx;
}

// in-box cells of the flat solution back into the caller's vector<vector>
inline void scatter(const std::vector<double>& x, const std::vector<int>& bx, int n_bl, int nj2,
                    std::vector<std::vector<double> >& solution)
{
    for (int b = 0; b < n_bl; b++)
        for (int i = bx[4 * b]; i <= bx[4 * b + 1]; i++)
            for (int j = bx[4 * b + 2]; j <= bx[4 * b + 3]; j++) solution[i][j] = x[(size_t)i * nj2 + j];
}

}  // namespace cathode_host

inline void solver_SMG(std::vector<std::vector<double> > aP, std::vector<std::vector<double> > aE,
                       std::vector<std::vector<double> > aW, std::vector<std::vector<double> > aN,
                       std::vector<std::vector<double> > aS, std::vector<std::vector<double> > b,
                       std::vector<std::vector<int> > i_bl, std::vector<std::vector<int> > j_bl, int n_bl, int maxITR,
                       double error_cnv, int flag_msg, std::vector<std::vector<double> >& solution)
{
    using namespace cathode_host;
    ensure_init();
    const int ni2 = (int)solution.size(), nj2 = (int)solution.at(0).size();
    std::vector<double> f[6];
    const std::vector<std::vector<double> >* src[6] = {&aP, &aE, &aW, &aN, &aS, &b};
    for (int k = 0; k < 6; k++) flatten(*src[k], ni2, nj2, f[k]);
    std::vector<int> bx = boxes_of(i_bl, j_bl, n_bl);
    std::vector<double> x((size_t)ni2 * nj2, 0.0);
    cathode_status st{};
    check(cathode_struct5_solve(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), bx.data(),
                                n_bl, ni2, nj2, nj2, maxITR, error_cnv, solver_mode(), x.data(), &st),
          "solver_SMG");
    scatter(x, bx, n_bl, nj2, solution);
    if (flag_msg == 1) {
        std::cout << "solver time = " << st.seconds << " sec" << std::endl;
        std::cout << "Number of Iterations: " << st.iterations << std::endl;
        std::cout << "Final Relative Residual Norm: " << st.rel_residual << std::endl;
    }
}

inline void solver_SOR(std::vector<std::vector<double> > aP, std::vector<std::vector<double> > aE,
                       std::vector<std::vector<double> > aW, std::vector<std::vector<double> > aN,
                       std::vector<std::vector<double> > aS, std::vector<std::vector<double> > b,
                       std::vector<std::vector<int> > i_bl, std::vector<std::vector<int> > j_bl, int n_bl,
                       double alpha_SOR, int maxITR, double error_cnv, int flag_msg,
                       std::vector<std::vector<double> >& solution)
{
    using namespace cathode_host;
    ensure_init();
    const int ni2 = (int)solution.size(), nj2 = (int)solution.at(0).size();
    std::vector<double> f[6], x;
    const std::vector<std::vector<double> >* src[6] = {&aP, &aE, &aW, &aN, &aS, &b};
    for (int k = 0; k < 6; k++) flatten(*src[k], ni2, nj2, f[k]);
    flatten(solution, ni2, nj2, x);  // SOR starts from (and reads ghosts of) the caller's solution
    std::vector<int> bx = boxes_of(i_bl, j_bl, n_bl);
    cathode_status st{};
    check(cathode_sor_solve(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), bx.data(),
                            n_bl, ni2, nj2, nj2, alpha_SOR, maxITR, error_cnv, x.data(), &st),
          "solver_SOR");
    scatter(x, bx, n_bl, nj2, solution);
    if (flag_msg == 1) std::cout << "ncount = " << st.iterations << " error_max = " << st.rel_residual << std::endl;
}

inline void solver_LU_9p(std::vector<std::vector<double> > aP, std::vector<std::vector<double> > aE,
                         std::vector<std::vector<double> > aW, std::vector<std::vector<double> > aN,
                         std::vector<std::vector<double> > aS, std::vector<std::vector<double> > aEN,
                         std::vector<std::vector<double> > aES, std::vector<std::vector<double> > aWN,
                         std::vector<std::vector<double> > aWS, std::vector<std::vector<double> > b,
                         std::vector<std::vector<int> > i_bl, std::vector<std::vector<int> > j_bl, int n_bl,
                         int flag_msg, std::vector<std::vector<double> >& solution)
{
    using namespace cathode_host;
    ensure_init();
    const int ni2 = (int)solution.size(), nj2 = (int)solution.at(0).size();
    std::vector<double> f[10];
    const std::vector<std::vector<double> >* src[10] = {&aP, &aE, &aW, &aN, &aS, &aEN, &aES, &aWN, &aWS, &b};
    for (int k = 0; k < 10; k++) flatten(*src[k], ni2, nj2, f[k]);
    std::vector<int> bx = boxes_of(i_bl, j_bl, n_bl);
    std::vector<double> x((size_t)ni2 * nj2, 0.0);
    cathode_status st{};
    check(cathode_lu9_solve(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), f[6].data(),
                            f[7].data(), f[8].data(), f[9].data(), bx.data(), n_bl, ni2, nj2, nj2, x.data(), &st),
          "solver_LU_9p");
    scatter(x, bx, n_bl, nj2, solution);
    if (flag_msg == 1) std::cout << "Actual error: " << st.rel_residual << std::endl;
}
