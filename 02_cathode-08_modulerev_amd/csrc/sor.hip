// sor.hip -- lexicographic point SOR (include/solverFuncs.hpp:169-216).
//
// The reference sweeps blocks in order, then i, then j, each update using
// the newest values of the neighbours visited earlier in the sweep and the
// previous values of the others.  A cell's wavefront level is one more than
// the largest level of its earlier-visited neighbours; all cells of a level
// are independent, so processing the levels in order inside one workgroup
// (a barrier between levels) performs exactly the sequential arithmetic.
// Multiply-adds are not contracted, so the iterates are bit-identical to a
// sequential FP64 sweep evaluated in source order.
#include <chrono>
#include <cmath>
#include <vector>

#include "sor.hpp"

namespace cathode {

constexpr int SOR_THREADS = 1024;

__global__ void __launch_bounds__(SOR_THREADS)
    k_sor(int ld, const double* __restrict__ aP, const double* __restrict__ aE,
          const double* __restrict__ aW, const double* __restrict__ aN, const double* __restrict__ aS,
          const double* __restrict__ b, double* __restrict__ x, const int* __restrict__ lev_ptr, int nlev,
          const int* __restrict__ cells, double alpha, int maxITR, double tol, int* __restrict__ out_n,
          double* __restrict__ out_err)
{
#pragma clang fp contract(off)
    __shared__ double lds[SOR_THREADS / 64];
    int ncount = 0;
    double error = 0.0;
    do {
        ncount = ncount + 1;
        double e_loc = 0.0;
        for (int l = 0; l < nlev; l++) {
            for (int c = lev_ptr[l] + (int)threadIdx.x; c < lev_ptr[l + 1]; c += SOR_THREADS) {
                const int k = cells[c];
                const double t = (aE[k] * x[k + ld] + aW[k] * x[k - ld] + aN[k] * x[k + 1] + aS[k] * x[k - 1] +
                                  b[k]) /
                                 (aP[k] + 1e-100);
                const double xo = x[k];
                const double xn = (1.0 - alpha) * xo + alpha * t;
                const double e = fabs((xn - xo) / (xn + xo + 1e-6)) * 2.0;
                if (e > e_loc) e_loc = e;
                x[k] = xn;
            }
            __syncthreads();
        }
        error = block_max<SOR_THREADS>(e_loc, lds);
        if (threadIdx.x == 0) lds[0] = error;
        __syncthreads();
        error = lds[0];
        __syncthreads();
        if (ncount > maxITR) break;
    } while (error > tol);
    if (threadIdx.x == 0) {
        *out_n = ncount;
        *out_err = error;
    }
}

struct SorPlan {
    std::vector<int> key;
    DevBuf<int> lev_ptr, cells;
    int nlev = 0;
};
static SorPlan g_plan;
static DevBuf<double> g_arr;
static DevBuf<int> g_out_n;
static DevBuf<double> g_out_err;

void sor_release()
{
    g_plan.lev_ptr.release();
    g_plan.cells.release();
    g_plan.key.clear();
    g_arr.release();
    g_out_n.release();
    g_out_err.release();
}

static int make_plan(const int* boxes, int n_bl, int ni2, int nj2, int ld)
{
    std::vector<int> key(boxes, boxes + 4 * n_bl);
    key.push_back(ni2);
    key.push_back(nj2);
    key.push_back(ld);
    if (key == g_plan.key) return CATHODE_OK;
    std::vector<int> pos((size_t)ni2 * ld, -1), lev((size_t)ni2 * ld, -1);
    std::vector<int> order;
    int p = 0;
    for (int bl = 0; bl < n_bl; bl++) {
        const int* bx = boxes + 4 * bl;
        if (bx[0] < 1 || bx[1] > ni2 - 2 || bx[2] < 1 || bx[3] > nj2 - 2) {
            set_error("sor: boxes must lie inside the interior cells");
            return CATHODE_ERR_ARG;
        }
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++) {
                const size_t k = (size_t)i * ld + j;
                if (pos[k] >= 0) {
                    set_error("sor: overlapping boxes are not supported");
                    return CATHODE_ERR_ARG;
                }
                pos[k] = p++;
                order.push_back((int)k);
            }
    }
    int nlev = 0;
    for (int k : order) {
        int l = 0;
        const int nb[4] = {k + ld, k - ld, k + 1, k - 1};
        for (int q = 0; q < 4; q++)
            if (pos[nb[q]] >= 0 && pos[nb[q]] < pos[k]) l = std::max(l, lev[nb[q]] + 1);
        lev[k] = l;
        nlev = std::max(nlev, l + 1);
    }
    std::vector<int> ptr(nlev + 1, 0), cells(order.size());
    for (int k : order) ptr[lev[k] + 1]++;
    for (int l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int k : order) cells[fill[lev[k]]++] = k;
    g_plan.lev_ptr.alloc(nlev + 1);
    g_plan.cells.alloc(cells.size() + 1);
    CATHODE_HIP(hipMemcpy(g_plan.lev_ptr.p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    CATHODE_HIP(hipMemcpy(g_plan.cells.p, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
    g_plan.nlev = nlev;
    g_plan.key = key;
    return CATHODE_OK;
}

int sor_solve_host(const double* aP, const double* aE, const double* aW, const double* aN,
                   const double* aS, const double* b, const int* boxes, int n_bl, int ni2, int nj2,
                   int ld, double alpha, int maxITR, double tol, double* x, hipStream_t s,
                   cathode_status* st)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!aP || !aE || !aW || !aN || !aS || !b || !x || !boxes || n_bl <= 0 || ni2 < 3 || nj2 < 3 ||
        ld < nj2) {
        set_error("sor: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    int rc = make_plan(boxes, n_bl, ni2, nj2, ld);
    if (rc) return rc;
    const size_t n = (size_t)ni2 * ld;
    g_arr.alloc(7 * n);
    g_out_n.alloc(1);
    g_out_err.alloc(1);
    const double* src[7] = {aP, aE, aW, aN, aS, b, x};
    for (int k = 0; k < 7; k++)
        CATHODE_HIP(hipMemcpyAsync(g_arr.p + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, s));
    double* d = g_arr.p;
    hipLaunchKernelGGL(k_sor, dim3(1), dim3(SOR_THREADS), 0, s, ld, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n,
                       d + 5 * n, d + 6 * n, g_plan.lev_ptr.p, g_plan.nlev, g_plan.cells.p, alpha, maxITR, tol,
                       g_out_n.p, g_out_err.p);
    CATHODE_LAUNCH_CHECK();
    int ncount = 0;
    double err = 0.0;
    CATHODE_HIP(hipMemcpyAsync(x, d + 6 * n, n * sizeof(double), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipMemcpyAsync(&ncount, g_out_n.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipMemcpyAsync(&err, g_out_err.p, sizeof(double), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    cathode_status loc{};
    loc.iterations = ncount;
    loc.rel_residual = err;
    loc.converged = !(err > tol);
    loc.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    record_status(loc);
    if (st) *st = loc;
    return CATHODE_OK;
}

}  // namespace cathode
