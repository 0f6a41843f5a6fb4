// capi.hip -- extern "C" entry points of libcathode_hip (include/cathode_hip.h):
// lifecycle, the drop-in structured solves (solver_SMG / solver_SOR /
// solver_LU_9p of include/solverFuncs.hpp) and the error/status side channel.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "helm.hpp"
#include "mg.hpp"
#include "sor.hpp"

namespace cathode {

static std::string g_err;
static std::mutex g_err_mu;  // batch worker threads report errors concurrently
static cathode_status g_last{};
static std::mutex g_mu;
std::mutex& api_mutex() { return g_mu; }

void set_error(const std::string& m)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = m;
}
const char* get_error()
{
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_err_mu);
    copy = g_err;
    return copy.c_str();
}

Context& ctx()
{
    static Context c;
    return c;
}

bool require_init()
{
    if (!ctx().ready) {
        set_error("cathode_init() has not been called");
        return false;
    }
    return true;
}

void record_status(const cathode_status& s) { g_last = s; }

// generic solver instance (one per library; calls are serialised)
static std::unique_ptr<MGSolver> g_solver;
static DevBuf<double> g_stage;                 // device staging for host inputs
static std::vector<double> g_hx;              // host staging for the solution
static DevBuf<unsigned char> g_actbuf;

// activity mask (interior-only boxes) built on the host
static bool build_mask(const int* boxes, int n_bl, int ni2, int nj2, int ld_dst,
                       std::vector<unsigned char>& m)
{
    m.assign((size_t)ni2 * ld_dst, 0);
    if (n_bl <= 0 || !boxes) return false;
    for (int b = 0; b < n_bl; b++) {
        const int* bx = boxes + 4 * b;
        if (bx[0] < 1 || bx[1] > ni2 - 2 || bx[2] < 1 || bx[3] > nj2 - 2) return false;
        for (int i = bx[0]; i <= bx[1]; i++)
            for (int j = bx[2]; j <= bx[3]; j++) m[(size_t)i * ld_dst + j] = 1;
    }
    return true;
}

__global__ void k_scatter_solution(int ni2, int nj2, int ld_src, const double* __restrict__ xs,
                                   const unsigned char* __restrict__ act, int ld_dst,
                                   double* __restrict__ x)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (i >= ni2 || j >= nj2) return;
    if (act[(size_t)i * ld_src + j]) x[(size_t)i * ld_dst + j] = xs[(size_t)i * ld_src + j];
}

// ------------------------------------------------------------------ no-allocation solve
// cathode_struct5_solve_dev_noalloc runs the multigrid of one solve inside
// `scratch`, a static device array each GPU holds for the life of the
// process (sized for grids like the reference's shipped 194 x 70 and up to
// about 256 x 256 cells; the bytes a shape needs are counted before use).
// In host code `scratch` names a host-side stand-in, not the device copy:
// cathode_init looks up each device's address once (g_scratch_dev) and only
// that address is used.  Every call shares the array and the host block, so
// calls that overlap on one GPU would race; the library's API mutex and the
// synchronisation at the end of each call keep one process's calls apart.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
constexpr size_t NOALLOC_HOST_DOUBLES = 16384;
constexpr int NOALLOC_MAX_BOXES = 64;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static unsigned char* g_scratch_dev[MAX_DEVICES];   // device address of scratch, per device
static double g_noalloc_host[NOALLOC_HOST_DOUBLES];  // what the solve's kernels leave for the host
static MGSolver g_noalloc_solver;                    // its buffers are pieces of scratch

struct BoxList {
    int n;
    int b[4 * NOALLOC_MAX_BOXES];  // ilo, ihi, jlo, jhi per box
};

// act[i][j] = 1 on the union of the boxes, 0 elsewhere (ghosts and padding too)
__global__ void k_mask_from_boxes(BoxList L, int ni2, int ld0, unsigned char* __restrict__ act)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (i >= ni2 || j >= ld0) return;
    unsigned char on = 0;
    for (int q = 0; q < L.n; q++) {
        const int* bx = L.b + 4 * q;
        if (i >= bx[0] && i <= bx[1] && j >= bx[2] && j <= bx[3]) on = 1;
    }
    act[(size_t)i * ld0 + j] = on;
}

// common body of the host and device struct5 entry points
static int struct5_impl(const double* aP, const double* aE, const double* aW, const double* aN,
                        const double* aS, const double* b, const int* boxes, int n_bl, int ni2,
                        int nj2, int ld, int maxITR, double tol, int mode, double* x, bool device,
                        hipStream_t s, cathode_status* st)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!aP || !aE || !aW || !aN || !aS || !b || !x || ni2 < 3 || nj2 < 3 || ld < nj2 ||
        (mode != CATHODE_MODE_FP64 && mode != CATHODE_MODE_MIXED)) {
        set_error("struct5: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    if (!g_solver) g_solver.reset(new MGSolver());
    MGSolver& S = *g_solver;
    S.configure(ni2 - 2, nj2 - 2, s);
    const int ld0 = S.ld0();
    std::vector<unsigned char> mask;
    if (!build_mask(boxes, n_bl, ni2, nj2, ld0, mask)) {
        set_error("struct5: boxes must lie inside the interior cells 1..ni, 1..nj");
        return CATHODE_ERR_ARG;
    }
    CATHODE_HIP(hipMemcpyAsync(S.act(), mask.data(), mask.size(), hipMemcpyHostToDevice, s));
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const double* src[6] = {aP, aE, aW, aN, aS, b};
    double* dst[6] = {S.aP(), S.aE(), S.aW(), S.aN(), S.aS(), S.rhs()};
    for (int k = 0; k < 6; k++)
        CATHODE_HIP(hipMemcpy2DAsync(dst[k], ld0 * sizeof(double), src[k], ld * sizeof(double),
                                     nj2 * sizeof(double), ni2, kind, s));
    S.set_weights(nullptr);
    S.enable_nine(false);
    cathode_status local{};
    S.solve(maxITR, tol, mode, s, &local);
    if (device) {
        hipLaunchKernelGGL(k_scatter_solution, dim3(cdiv(nj2, 64), ni2), dim3(64), 0, s, ni2, nj2, ld0,
                           S.sol(), S.act(), ld, x);
        CATHODE_LAUNCH_CHECK();
        CATHODE_HIP(hipStreamSynchronize(s));
    } else {
        g_hx.resize((size_t)ni2 * ld0);
        CATHODE_HIP(hipMemcpyAsync(g_hx.data(), S.sol(), g_hx.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        CATHODE_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < ni2; i++)
            for (int j = 0; j < nj2; j++)
                if (mask[(size_t)i * ld0 + j]) x[(size_t)i * ld + j] = g_hx[(size_t)i * ld0 + j];
    }
    local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    record_status(local);
    if (st) *st = local;
    if (!std::isfinite(local.rel_residual)) {
        set_error("struct5: non-finite residual (NaN/Inf in the system)");
        return CATHODE_ERR_NUMERIC;
    }
    return CATHODE_OK;
}

}  // namespace cathode

using namespace cathode;

extern "C" {

int cathode_abi_version(void) { return CATHODE_ABI_VERSION; }

int cathode_set_presum_min(int n)
{
    set_presum_min(n);
    return CATHODE_OK;
}

const char* cathode_last_error(void) { return get_error(); }

void cathode_last_status(cathode_status* st)
{
    if (st) *st = g_last;
}

int cathode_init(int device)
{
    CATHODE_GUARD_BEGIN
    Context& c = ctx();
    if (c.ready && c.device == device) return CATHODE_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("no HIP device visible");
        return CATHODE_ERR_HIP;
    }
    if (device < 0 || device >= n) {
        set_error("device index out of range");
        return CATHODE_ERR_ARG;
    }
    CATHODE_HIP(hipSetDevice(device));
    if (device < MAX_DEVICES && !g_scratch_dev[device]) {
        // this device's copy of the no-allocation scratch (the name itself is
        // a host-side stand-in that kernels must never receive)
        void* d = nullptr;
        CATHODE_HIP(hipGetSymbolAddress(&d, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = (unsigned char*)d;
    }
    if (!c.stream) CATHODE_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.device = device;
    c.ready = true;
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_finalize(void)
{
    CATHODE_GUARD_BEGIN
    Context& c = ctx();
    if (!c.ready) return CATHODE_OK;
    g_solver.reset();
    g_stage.release();
    g_actbuf.release();
    sor_release();
    helm_release();
    mw_release();
    mws_release();
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    c.ready = false;
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_struct5_solve(const double* aP, const double* aE, const double* aW, const double* aN,
                          const double* aS, const double* b, const int* boxes, int n_bl, int ni2,
                          int nj2, int ld, int maxITR, double tol, int mode, double* x,
                          cathode_status* st)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    return struct5_impl(aP, aE, aW, aN, aS, b, boxes, n_bl, ni2, nj2, ld, maxITR, tol, mode, x, false,
                        ctx().stream, st);
    CATHODE_GUARD_END
}

int cathode_struct5_solve_dev(const double* aP, const double* aE, const double* aW, const double* aN,
                              const double* aS, const double* b, const int* boxes, int n_bl, int ni2,
                              int nj2, int ld, int maxITR, double tol, int mode, double* x,
                              void* stream, cathode_status* st)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    hipStream_t s = stream ? (hipStream_t)stream : ctx().stream;
    return struct5_impl(aP, aE, aW, aN, aS, b, boxes, n_bl, ni2, nj2, ld, maxITR, tol, mode, x, true, s, st);
    CATHODE_GUARD_END
}

int cathode_struct5_solve_dev_noalloc(const double* aP, const double* aE, const double* aW, const double* aN,
                                      const double* aS, const double* b, const int* boxes, int n_bl, int ni2,
                                      int nj2, int ld, int maxITR, double tol, int mode, double* x, void* stream,
                                      cathode_status* st)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = stream ? (hipStream_t)stream : ctx().stream;
    if (!aP || !aE || !aW || !aN || !aS || !b || !x || !boxes || ni2 < 3 || nj2 < 3 || ld < nj2 ||
        (mode != CATHODE_MODE_FP64 && mode != CATHODE_MODE_MIXED)) {
        set_error("struct5_noalloc: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    if (n_bl < 1 || n_bl > NOALLOC_MAX_BOXES) {
        set_error("struct5_noalloc: n_bl must be 1.." + std::to_string(NOALLOC_MAX_BOXES));
        return CATHODE_ERR_ARG;
    }
    BoxList bl;
    bl.n = n_bl;
    for (int q = 0; q < n_bl; q++) {
        const int* bx = boxes + 4 * q;
        if (bx[0] < 1 || bx[1] > ni2 - 2 || bx[2] < 1 || bx[3] > nj2 - 2) {
            set_error("struct5_noalloc: boxes must lie inside the interior cells 1..ni, 1..nj");
            return CATHODE_ERR_ARG;
        }
        for (int k = 0; k < 4; k++) bl.b[4 * q + k] = bx[k];
    }
    const int dev = ctx().device;
    unsigned char* base = dev < MAX_DEVICES ? g_scratch_dev[dev] : nullptr;
    if (!base) {
        set_error("struct5_noalloc: no scratch address for this device");
        return CATHODE_ERR_HIP;
    }
    MGSolver& S = g_noalloc_solver;  // another device's address differs, so its layout is rebuilt
    const size_t need = S.configure_in(ni2 - 2, nj2 - 2, base, NOALLOC_SCRATCH_BYTES, g_noalloc_host,
                                       NOALLOC_HOST_DOUBLES, s);
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error("struct5_noalloc: a " + std::to_string(ni2 - 2) + " x " + std::to_string(nj2 - 2) +
                  " grid needs " + (need == SIZE_MAX ? std::string("a larger host block") :
                                    std::to_string(need) + " bytes of scratch") +
                  "; the static scratch holds " + std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes");
        return CATHODE_ERR_ARG;
    }
    const int ld0 = S.ld0();
    hipLaunchKernelGGL(k_mask_from_boxes, dim3(cdiv(ld0, 64), ni2), dim3(64), 0, s, bl, ni2, ld0, S.act());
    CATHODE_LAUNCH_CHECK();
    const double* src[6] = {aP, aE, aW, aN, aS, b};
    double* dst[6] = {S.aP(), S.aE(), S.aW(), S.aN(), S.aS(), S.rhs()};
    for (int k = 0; k < 6; k++)
        CATHODE_HIP(hipMemcpy2DAsync(dst[k], ld0 * sizeof(double), src[k], ld * sizeof(double),
                                     nj2 * sizeof(double), ni2, hipMemcpyDeviceToDevice, s));
    S.set_weights(nullptr);
    S.enable_nine(false);
    cathode_status local{};
    S.solve(maxITR, tol, mode, s, &local);
    hipLaunchKernelGGL(k_scatter_solution, dim3(cdiv(nj2, 64), ni2), dim3(64), 0, s, ni2, nj2, ld0, S.sol(), S.act(),
                       ld, x);
    CATHODE_LAUNCH_CHECK();
    CATHODE_HIP(hipStreamSynchronize(s));
    local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    record_status(local);
    if (st) *st = local;
    if (!std::isfinite(local.rel_residual)) {
        set_error("struct5_noalloc: non-finite residual (NaN/Inf in the system)");
        return CATHODE_ERR_NUMERIC;
    }
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_sor_solve(const double* aP, const double* aE, const double* aW, const double* aN,
                      const double* aS, const double* b, const int* boxes, int n_bl, int ni2, int nj2,
                      int ld, double alpha, int maxITR, double tol, double* x, cathode_status* st)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    return sor_solve_host(aP, aE, aW, aN, aS, b, boxes, n_bl, ni2, nj2, ld, alpha, maxITR, tol, x,
                          ctx().stream, st);
    CATHODE_GUARD_END
}

int cathode_dev_alloc(size_t bytes, void** ptr)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if (!ptr || bytes == 0) {
        set_error("dev_alloc: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error("dev_alloc: out of device memory");
        return CATHODE_ERR_NOMEM;
    }
    CATHODE_HIP(hipMemset(p, 0, bytes));
    CATHODE_HIP(hipDeviceSynchronize());  // null-stream memset before any stream use
    *ptr = p;
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_dev_free(void* ptr)
{
    CATHODE_GUARD_BEGIN
    if (ptr) CATHODE_HIP(hipFree(ptr));
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_dev_copy(void* dst, const void* src, size_t bytes, int kind)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if ((!dst || !src) && bytes) {
        set_error("dev_copy: null pointer");
        return CATHODE_ERR_ARG;
    }
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) {
        set_error("dev_copy: kind must be 0 (to device), 1 (to host) or 2 (device to device)");
        return CATHODE_ERR_ARG;
    }
    CATHODE_HIP(hipMemcpyAsync(dst, src, bytes, k, ctx().stream));
    CATHODE_HIP(hipStreamSynchronize(ctx().stream));
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_lu9_solve(const double* aP, const double* aE, const double* aW, const double* aN, const double* aS,
                      const double* aEN, const double* aES, const double* aWN, const double* aWS, const double* b,
                      const int* boxes, int n_bl, int ni2, int nj2, int ld, double* x, cathode_status* st)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    auto t0 = std::chrono::steady_clock::now();
    if (!aP || !aE || !aW || !aN || !aS || !aEN || !aES || !aWN || !aWS || !b || !x || ni2 < 3 || nj2 < 3 ||
        ld < nj2) {
        set_error("lu9: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    hipStream_t s = ctx().stream;
    if (!g_solver) g_solver.reset(new MGSolver());
    MGSolver& S = *g_solver;
    S.configure(ni2 - 2, nj2 - 2, s);
    const int ld0 = S.ld0();
    std::vector<unsigned char> mask;
    if (!build_mask(boxes, n_bl, ni2, nj2, ld0, mask)) {
        set_error("lu9: boxes must lie inside the interior cells 1..ni, 1..nj");
        return CATHODE_ERR_ARG;
    }
    CATHODE_HIP(hipMemcpyAsync(S.act(), mask.data(), mask.size(), hipMemcpyHostToDevice, s));
    S.enable_nine(true);
    const size_t n0 = S.n0();
    const double* src[10] = {aP, b, aE, aW, aN, aS, aEN, aES, aWN, aWS};
    double* dst[10] = {S.aP(), S.rhs()};
    for (int q = 0; q < 8; q++) dst[2 + q] = S.op9() + q * n0;
    for (int k = 0; k < 10; k++)
        CATHODE_HIP(hipMemcpy2DAsync(dst[k], ld0 * sizeof(double), src[k], ld * sizeof(double), nj2 * sizeof(double),
                                     ni2, hipMemcpyHostToDevice, s));
    S.lump_nine(s);
    S.set_weights(nullptr);
    cathode_status local{};
    try {
        // solver_LU_9p factorizes exactly: iterate to round-off level
        S.solve(200, 1e-13, CATHODE_MODE_FP64, s, &local);
    } catch (...) {
        S.enable_nine(false);
        throw;
    }
    S.enable_nine(false);
    g_hx.resize((size_t)ni2 * ld0);
    CATHODE_HIP(hipMemcpyAsync(g_hx.data(), S.sol(), g_hx.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < ni2; i++)
        for (int j = 0; j < nj2; j++)
            if (mask[(size_t)i * ld0 + j]) x[(size_t)i * ld + j] = g_hx[(size_t)i * ld0 + j];
    local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    record_status(local);
    if (st) *st = local;
    if (!std::isfinite(local.rel_residual)) {
        set_error("lu9: non-finite residual (NaN/Inf in the system)");
        return CATHODE_ERR_NUMERIC;
    }
    return CATHODE_OK;
    CATHODE_GUARD_END
}

}  // extern "C"
