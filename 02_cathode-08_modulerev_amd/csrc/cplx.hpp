// cplx.hpp -- minimal complex value type for the device kernels
// (complex128 = cx<double>, complex64 = cx<float>), interleaved (re, im)
// like std::complex, so host arrays of std::complex / numpy complex map 1:1.
#pragma once

#include <hip/hip_runtime.h>

namespace cathode {

template <class R>
struct cx {
    R re, im;
    cx() = default;  // trivial, so cx can live in __shared__ arrays
    __host__ __device__ constexpr cx(R r) : re(r), im(0) {}
    __host__ __device__ constexpr cx(R r, R i) : re(r), im(i) {}
    template <class Q>
    __host__ __device__ explicit constexpr cx(const cx<Q>& o) : re((R)o.re), im((R)o.im) {}
};

template <class R>
__host__ __device__ inline cx<R> operator+(cx<R> a, cx<R> b) { return cx<R>(a.re + b.re, a.im + b.im); }
template <class R>
__host__ __device__ inline cx<R> operator-(cx<R> a, cx<R> b) { return cx<R>(a.re - b.re, a.im - b.im); }
template <class R>
__host__ __device__ inline cx<R> operator-(cx<R> a) { return cx<R>(-a.re, -a.im); }
template <class R>
__host__ __device__ inline cx<R> operator*(cx<R> a, cx<R> b)
{
    return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class R>
__host__ __device__ inline cx<R> operator/(cx<R> a, cx<R> b)
{
    // Smith's algorithm (scaled by the larger component of b)
    if (fabs(b.re) >= fabs(b.im)) {
        const R t = b.im / b.re, d = b.re + b.im * t;
        return cx<R>((a.re + a.im * t) / d, (a.im - a.re * t) / d);
    }
    const R t = b.re / b.im, d = b.re * t + b.im;
    return cx<R>((a.re * t + a.im) / d, (a.im * t - a.re) / d);
}
template <class R>
__host__ __device__ inline cx<R>& operator+=(cx<R>& a, cx<R> b)
{
    a.re += b.re;
    a.im += b.im;
    return a;
}
template <class R>
__host__ __device__ inline bool operator==(cx<R> a, cx<R> b) { return a.re == b.re && a.im == b.im; }
template <class R>
__host__ __device__ inline cx<R> conj(cx<R> a) { return cx<R>(a.re, -a.im); }
template <class R>
__host__ __device__ inline R abs2v(cx<R> a) { return a.re * a.re + a.im * a.im; }

typedef cx<double> cxd;
typedef cx<float> cxf;

}  // namespace cathode
