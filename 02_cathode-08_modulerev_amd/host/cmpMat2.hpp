// cmpMat2.hpp -- 2x2 complex transfer (ABCD) matrices and the S-parameter
// conversions used to normalise the microwave port power.
//
// Interface of include/cmpMat2.hpp:8-34 (CmpMat2 with members A, B, C, D,
// det(), inv(), operator*) and of InitialModule::sParamToABCD /
// ABCDToSParam (include/initialModule.hpp:1689-1739), with the reference's
// argument order and error behaviour (std::runtime_error when |S21| or the
// S-denominator falls below 1e-14).  The conversions are the standard
// two-port relations for equal real reference impedances Z0:
//   A = ((1+S11)(1-S22) + S12 S21) / (2 S21)     B = Z0 ((1+S11)(1+S22) - S12 S21) / (2 S21)
//   C = ((1-S11)(1-S22) - S12 S21) / (2 S21 Z0)  D = ((1-S11)(1+S22) + S12 S21) / (2 S21)
// and back with d = A + B/Z0 + C Z0 + D:
//   S11 = (A + B/Z0 - C Z0 - D)/d, S21 = 2/d, S12 = 2(AD - BC)/d, S22 = (-A + B/Z0 - C Z0 + D)/d.
// stub_cable_correction() applies the short-stub length change and the cable
// phase delay of include/initialModule.hpp:107-190 to measured S-parameters.
#pragma once

#include <cmath>
#include <complex>
#include <stdexcept>

namespace cathode_host {

using cplx = std::complex<double>;

struct CmpMat2 {
    cplx A, B, C, D;

    cplx det() const { return A * D - B * C; }

    CmpMat2 inv() const
    {
        const cplx d = det();
        return {D / d, -B / d, -C / d, A / d};
    }
};

inline CmpMat2 operator*(const CmpMat2& X, const CmpMat2& Y)
{
    return {X.A * Y.A + X.B * Y.C, X.A * Y.B + X.B * Y.D, X.C * Y.A + X.D * Y.C, X.C * Y.B + X.D * Y.D};
}

inline void sParamToABCD(CmpMat2& T, const cplx& S11, const cplx& S21, const cplx& S12, const cplx& S22,
                         const double Z0)
{
    if (std::abs(S21) < 1.0e-14) throw std::runtime_error("sParamToABCD: S21 is too small.");
    const cplx two_s21 = 2.0 * S21, cross = S12 * S21;
    T.A = ((1.0 + S11) * (1.0 - S22) + cross) / two_s21;
    T.B = Z0 * ((1.0 + S11) * (1.0 + S22) - cross) / two_s21;
    T.C = ((1.0 - S11) * (1.0 - S22) - cross) / (two_s21 * Z0);
    T.D = ((1.0 - S11) * (1.0 + S22) + cross) / two_s21;
}

inline void ABCDToSParam(cplx& S11, cplx& S21, cplx& S12, cplx& S22, const CmpMat2& T, const double Z0)
{
    const cplx bz = T.B / Z0, cz = T.C * Z0;
    const cplx d = T.A + bz + cz + T.D;
    if (std::abs(d) < 1.0e-14) throw std::runtime_error("ABCDToSParam: denominator is too small.");
    S11 = (T.A + bz - cz - T.D) / d;
    S21 = 2.0 / d;
    S12 = 2.0 * T.det() / d;
    S22 = (-T.A + bz - cz + T.D) / d;
}

// Lossless line of electrical length theta: [[cos, j Z0 sin], [j sin / Z0, cos]].
inline CmpMat2 line_abcd(double theta, double Z0)
{
    const cplx j(0.0, 1.0);
    return {std::cos(theta), j * Z0 * std::sin(theta), j * std::sin(theta) / Z0, std::cos(theta)};
}

// Measured S-parameters (magnitude, angle in degrees) corrected for a change
// dh of a short-circuited stub of length h0 (shunt admittance -j Y0 cot(beta h)),
// seen through a pre-line of phase phase_pre_deg, and for a cable extended by
// dl (S21, S12 gain exp(-j beta dl), S22 twice that).  beta = omega / c with
// c = 1/sqrt(eps0 epsr mu0).
struct SParams {
    cplx S11, S21, S12, S22;
};

inline SParams stub_cable_correction(SParams s, double Z0, double omega, double epsr, double h0, double dh, double dl,
                                     double phase_pre_deg)
{
    const double eps0 = 8.8541878128e-12, mu0 = 1.2566370614e-6;
    const double beta = omega * std::sqrt(eps0 * epsr * mu0);
    const double t0 = std::tan(beta * h0), t1 = std::tan(beta * (h0 + dh));
    if (std::abs(t0) < 1.0e-14 || std::abs(t1) < 1.0e-14)
        throw std::runtime_error("short-stub admittance: tan(beta*l) is too small.");
    const cplx j(0.0, 1.0);
    const cplx dY = -j / Z0 / t1 + j / Z0 / t0;
    const CmpMat2 shunt{1.0, 0.0, dY, 1.0};
    const CmpMat2 pre = line_abcd(phase_pre_deg / 180.0 * M_PI, Z0);
    CmpMat2 base;
    sParamToABCD(base, s.S11, s.S21, s.S12, s.S22, Z0);
    const CmpMat2 T = pre * shunt * pre.inv() * base;
    SParams o;
    ABCDToSParam(o.S11, o.S21, o.S12, o.S22, T, Z0);
    const cplx p = std::exp(cplx(0.0, -beta * dl));
    o.S21 *= p;
    o.S12 *= p;
    o.S22 *= p * p;
    return o;
}

}  // namespace cathode_host
