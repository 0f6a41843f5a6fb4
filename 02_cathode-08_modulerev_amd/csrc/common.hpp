// common.hpp -- shared runtime pieces of libcathode_hip: error state, the
// library stream, launch helpers and deterministic reductions (gfx950).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <exception>
#include <stdexcept>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/cathode_hip.h"

namespace cathode {

// ------------------------------------------------------------------ errors
// The last error message, shared by every thread (set under a lock; the
// reader gets a copy that stays valid until its own next call).
void set_error(const std::string& msg);
const char* get_error();

struct hip_failure {
    hipError_t e;
};

#define CATHODE_HIP(call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ::cathode::set_error(std::string(#call) + " -> " + hipGetErrorString(e_) + " (" + \
                                 __FILE__ + ":" + std::to_string(__LINE__) + ")");          \
            throw ::cathode::hip_failure{e_};                                               \
        }                                                                                   \
    } while (0)

#define CATHODE_LAUNCH_CHECK() CATHODE_HIP(hipGetLastError())

// Every extern "C" entry point runs under one library mutex and turns HIP
// failures / host allocation failures into error codes.
std::mutex& api_mutex();
#define CATHODE_GUARD_BEGIN                                       \
    std::lock_guard<std::mutex> lock_(::cathode::api_mutex());    \
    try {
#define CATHODE_GUARD_END CATHODE_CATCH_END

// The catch clauses alone, for entry points that do not take the library
// mutex (the NDD handles, which independent threads step side by side): no
// exception crosses the C boundary.
#define CATHODE_CATCH_END                                   \
    }                                                       \
    catch (const ::cathode::hip_failure&) {                 \
        return CATHODE_ERR_HIP;                             \
    }                                                       \
    catch (const std::bad_alloc&) {                         \
        ::cathode::set_error("host allocation failed");    \
        return CATHODE_ERR_NOMEM;                           \
    }                                                       \
    catch (const std::domain_error& ex_) {                  \
        ::cathode::set_error(ex_.what());                   \
        return CATHODE_ERR_NUMERIC;                         \
    }                                                       \
    catch (const std::exception& ex_) {                     \
        ::cathode::set_error(std::string("internal error: ") + ex_.what()); \
        return CATHODE_ERR_HIP;                             \
    }

// ------------------------------------------------------------------ context
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    bool ready = false;
};
Context& ctx();
bool require_init();
void record_status(const cathode_status& s);

// ------------------------------------------------------------------ device buffers
// Device allocation, owned (hipMalloc, freed in the destructor) or borrowed
// from memory someone else holds (never freed here).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    void alloc(size_t count)
    {
        if (count <= n && p && owned) return;
        release();
        CATHODE_HIP(hipMalloc(&p, count * sizeof(T) + 256));
        n = count;
    }
    void borrow(T* q, size_t count)
    {
        release();
        p = q;
        n = count;
        owned = false;
    }
    void zero(hipStream_t s) { CATHODE_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
};

// A fixed region of device memory handed out in 256-byte aligned pieces, each
// followed by 256 bytes of slack (as DevBuf::alloc adds); nothing is
// allocated.  With base == nullptr it only counts the bytes the pieces need.
struct Arena {
    unsigned char* base = nullptr;
    size_t used = 0;
    template <class T>
    T* take(size_t count)
    {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(T) + 256;
        return base ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// BiCGStab direction weight beta = (rho1 / rho) (al / om).  After a
// breakdown (om = 0: no step was taken, or rho = 0) it is not finite; the
// direction then restarts from the residual (beta = 0, p = r).
__host__ __device__ inline double bicg_beta(double rho1, double rho, double al, double om)
{
    const double b = (rho1 / rho) * (al / om);
    return isfinite(b) ? b : 0.0;
}

// ------------------------------------------------------------------ reductions
// Block-level sum of a double held by every thread of a 256-thread block;
// fixed order (wavefront butterfly then LDS), so results are deterministic.
__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// linear thread index of a 1-, 2- or 3-D block
__device__ inline int lin_tid()
{
    return (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
}

// Result valid in the thread with lin_tid() == 0.  NT = threads per block.
template <int NT>
__device__ inline double block_sum(double v, double* lds)
{
    v = wave_sum(v);
    const int t = lin_tid();
    const int lane = t & 63, wid = t >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (t == 0) {
        for (int k = 0; k < NT / 64; k++) s += lds[k];
    }
    __syncthreads();
    return s;
}

// Block-level sum returned to every thread (same fixed order in each, so
// every block of a launch that reduces the same data gets the same value).
// lds holds NT / 64 doubles.
template <int NT>
__device__ inline double block_sum_bcast(double v, double* lds)
{
    v = wave_sum(v);
    const int t = lin_tid();
    if ((t & 63) == 0) lds[t >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < NT / 64; k++) s += lds[k];
    __syncthreads();
    return s;
}

// M sums at once, returned to every thread; lds holds M * NT / 64 doubles
template <int NT, int M>
__device__ inline void block_sum_vec_bcast(double (&v)[M], double* lds)
{
#pragma unroll
    for (int m = 0; m < M; m++) v[m] = wave_sum(v[m]);
    const int t = lin_tid();
    if ((t & 63) == 0)
#pragma unroll
        for (int m = 0; m < M; m++) lds[(t >> 6) * M + m] = v[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; m++) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; w++) s += lds[w * M + m];
        v[m] = s;
    }
    __syncthreads();
}

template <int NT>
__device__ inline double block_max_bcast(double v, double* lds)
{
    v = wave_max(v);
    const int t = lin_tid();
    if ((t & 63) == 0) lds[t >> 6] = v;
    __syncthreads();
    double s = -INFINITY;
    for (int k = 0; k < NT / 64; k++) s = fmax(s, lds[k]);
    __syncthreads();
    return s;
}

template <int NT>
__device__ inline double block_max(double v, double* lds)
{
    v = wave_max(v);
    const int t = lin_tid();
    const int lane = t & 63, wid = t >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = -INFINITY;
    if (t == 0) {
        for (int k = 0; k < NT / 64; k++) s = fmax(s, lds[k]);
    }
    __syncthreads();
    return s;
}

// M reductions at once, each a sum or (is_max[m]) a maximum, combined in the
// same order as block_sum / block_max, with one barrier; results valid in the
// thread with lin_tid() == 0.  lds holds M * NT / 64 doubles and is not
// released by a trailing barrier (one use per kernel).
template <int NT, int M>
__device__ inline void block_reduce_vec(double (&v)[M], const bool (&is_max)[M], double* lds)
{
#pragma unroll
    for (int m = 0; m < M; m++) v[m] = is_max[m] ? wave_max(v[m]) : wave_sum(v[m]);
    const int t = lin_tid();
    if ((t & 63) == 0)
#pragma unroll
        for (int m = 0; m < M; m++) lds[(t >> 6) * M + m] = v[m];
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            double s = is_max[m] ? -INFINITY : 0.0;
            for (int w = 0; w < NT / 64; w++) s = is_max[m] ? fmax(s, lds[w * M + m]) : s + lds[w * M + m];
            v[m] = s;
        }
    }
}

// Sums `nvec` interleaved partial vectors part[b*nvec + k] over blocks b in
// fixed order; out[k] = total.  One block of 256 threads.
__global__ void k_reduce_partials(const double* part, int nblocks, int nvec, const int* is_max,
                                  double* out);

void reduce_partials(const double* part, int nblocks, int nvec, const int* d_is_max, double* out,
                     hipStream_t s);

// Small device-to-host reads: a kernel on s stores n doubles from device
// memory d straight into host (pinned, hipHostMalloc) memory h; the caller
// synchronises s before reading h.  Used for the per-iteration scalars
// instead of the runtime's copy path.
void fetch_host(double* h, const double* d, size_t n, hipStream_t s);

}  // namespace cathode
