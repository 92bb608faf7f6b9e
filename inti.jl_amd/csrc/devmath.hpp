// Device math for the pair kernels: FP64 reciprocal / inverse square root
// refined from the hardware estimates, the MIXED variants that use the FP32
// transcendental unit, phase reduction in FP64, and the Hankel functions
// H0^(1), H1^(1) of real argument (series + rational modulus/phase fits,
// coefficients from gen_hankel.py).
#pragma once
#include <hip/hip_runtime.h>

#include "hankel_coeffs.h"

namespace inti {

constexpr double kTwoOverPi = 0.63661977236758134308;
constexpr double kSqrt2OverPi = 0.79788456080286535588;  // sqrt(2/pi)
constexpr double kEulerGamma = 0.57721566490153286061;
constexpr double kQuarterPi = 0.78539816339744830962;
constexpr double kLn2 = 0.69314718055994530942;

// 1/x in FP64: v_rcp_f64 estimate + two Newton steps (error ~1 ulp).
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/sqrt(x) in FP64: v_rsq_f64 estimate + two Newton steps.
__device__ __forceinline__ double rsq64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = __builtin_fma(-(x * y), y, 1.0);
    y = __builtin_fma(0.5 * y, h, y);
    h = __builtin_fma(-(x * y), y, 1.0);
    return __builtin_fma(0.5 * y, h, y);
}

// FP32 hardware estimate refined once in FP64: ~1e-14 relative, enough for
// distances that enter a phase k*d.
__device__ __forceinline__ double rsq_mixed_phase(double x) {
    double y = (double)__builtin_amdgcn_rsqf((float)x);
    double h = __builtin_fma(-(x * y), y, 1.0);
    return __builtin_fma(0.5 * y, h, y);
}

template <bool MIXED>
__device__ __forceinline__ double rsqv(double x) {
    if constexpr (MIXED) return (double)__builtin_amdgcn_rsqf((float)x);
    else return rsq64(x);
}
// inverse square root feeding a phase
template <bool MIXED>
__device__ __forceinline__ double rsqp(double x) {
    if constexpr (MIXED) return rsq_mixed_phase(x);
    else return rsq64(x);
}
template <bool MIXED>
__device__ __forceinline__ double rcpv(double x) {
    if constexpr (MIXED) return (double)__builtin_amdgcn_rcpf((float)x);
    else return rcp64(x);
}
// natural logarithm
template <bool MIXED>
__device__ __forceinline__ double logv(double x) {
    if constexpr (MIXED) return (double)(__builtin_amdgcn_logf((float)x) * 0.69314718f);
    else return log(x);
}

// Taylor coefficients of sin(2 pi f) / f and cos(2 pi f) in z = f^2:
// (-1)^k (2 pi)^(2k+1) / (2k+1)! and (-1)^k (2 pi)^(2k) / (2k)!, rounded to
// double from 50-digit values. For |f| <= 1/8 (|2 pi f| <= pi/4) the first
// omitted terms are 4.6e-17 and 2.0e-18.
constexpr double kSinRev[8] = {0x1.921fb54442d18p+2,  -0x1.4abbce625be53p+5, 0x1.466bc6775aae2p+6,
                               -0x1.32d2cce62bd86p+6, 0x1.50783487ee782p+5,  -0x1.e3074fde8871fp+3,
                               0x1.e8f434d018d63p+1,  -0x1.6fadb9f155744p-1};
constexpr double kCosRev[9] = {0x1.0000000000000p+0,  -0x1.3bd3cc9be45dep+4, 0x1.03c1f081b5ac4p+6,
                               -0x1.55d3c7e3cbffap+6, 0x1.e1f506891babbp+5,  -0x1.a6d1f2a204a8cp+4,
                               0x1.f9d38a3763cc3p+2,  -0x1.b6e24f44b128fp+0, 0x1.20c62c2f2d7f5p-2};

// (sin 2 pi u, cos 2 pi u) in FP64 for a phase u in revolutions, |u| < 2^29:
// u = q/4 + f with q = rint(4u) and |f| <= 1/8 (f is exact), the two
// polynomials above in f, then the quarter turn q mod 4.
__device__ __forceinline__ void sincos_rev64(double u, double &s, double &c) {
    const double q = __builtin_rint(4.0 * u);
    const double f = __builtin_fma(-0.25, q, u);
    const double z = f * f;
    double ps = kSinRev[7];
#pragma unroll
    for (int k = 6; k >= 0; --k) ps = __builtin_fma(ps, z, kSinRev[k]);
    double pc = kCosRev[8];
#pragma unroll
    for (int k = 7; k >= 0; --k) pc = __builtin_fma(pc, z, kCosRev[k]);
    const double S = f * ps, C = pc;
    const int qi = (int)q & 3;
    // quarter turns: (S, C) -> (C, -S) -> (-S, -C) -> (-C, S)
    const double s0 = (qi & 1) ? C : S;
    const double c0 = (qi & 1) ? S : C;
    s = (qi & 2) ? -s0 : s0;
    c = ((qi ^ (qi >> 1)) & 1) ? -c0 : c0;
}

// The same (sin 2 pi u, cos 2 pi u) with every Horner step issued as one
// three-address v_fma_f64. Where many values are live (the pullback kernels)
// the register allocator otherwise keeps the coefficients in VGPRs and
// turns each step into a copy of the coefficient plus a two-address
// v_fmac_f64 (14 extra moves per evaluation).
__device__ __forceinline__ double fma_vop3(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ void sincos_rev64_vop3(double u, double &s, double &c) {
    const double q = __builtin_rint(4.0 * u);
    const double f = __builtin_fma(-0.25, q, u);
    const double z = f * f;
    double ps = kSinRev[7];
#pragma unroll
    for (int k = 6; k >= 0; --k) ps = fma_vop3(ps, z, kSinRev[k]);
    double pc = kCosRev[8];
#pragma unroll
    for (int k = 7; k >= 0; --k) pc = fma_vop3(pc, z, kCosRev[k]);
    const double S = f * ps, C = pc;
    const int qi = (int)q & 3;
    const double s0 = (qi & 1) ? C : S;
    const double c0 = (qi & 1) ? S : C;
    s = (qi & 2) ? -s0 : s0;
    c = ((qi ^ (qi >> 1)) & 1) ? -c0 : c0;
}

// (sin 2 pi u, cos 2 pi u) for a phase u given in revolutions. FP64:
// sincos_rev64. MIXED: u is reduced to [-1/2, 1/2] in FP64 and fed to the FP32
// sin/cos unit, whose argument is in revolutions.
template <bool MIXED>
__device__ __forceinline__ void sincos_rev(double u, double &s, double &c) {
    if constexpr (MIXED) {
        float ff = (float)(u - __builtin_rint(u));
        s = (double)__builtin_amdgcn_sinf(ff);
        c = (double)__builtin_amdgcn_cosf(ff);
    } else {
        sincos_rev64(u, s, c);
    }
}

// fused multiply-add in the operands' own precision (__builtin_fma is the
// double one: on floats it would widen, fuse in FP64 and narrow again)
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T, int N>
__device__ __forceinline__ T horner(const T (&c)[N], T u) {
    T s = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) s = fma_t(s, u, c[i]);
    return s;
}

// FP64 coefficient tables of hankel01_f64, kept in memory: as inline
// constants the ~150 coefficients would be materialised in registers ahead
// of the direct-sum loop (more than half the register file); read through a
// pointer the compiler cannot prove loop-invariant, they become scalar loads
// at the point of use.
struct HkF64Table {
    double a0[18], a1[18], b0[18], b1[18];
    double p0n[7], p0d[7], q0n[7], q0d[7], p1n[7], p1d[7], q1n[7], q1d[7];
};

constexpr HkF64Table make_hk_f64_table() {
    HkF64Table t{};
    for (int i = 0; i < 18; ++i) {
        t.a0[i] = HK_F64_A0[i]; t.a1[i] = HK_F64_A1[i];
        t.b0[i] = HK_F64_B0[i]; t.b1[i] = HK_F64_B1[i];
    }
    for (int i = 0; i < 7; ++i) {
        t.p0n[i] = HK_F64_P0_NUM[i]; t.p0d[i] = HK_F64_P0_DEN[i];
        t.q0n[i] = HK_F64_Q0_NUM[i]; t.q0d[i] = HK_F64_Q0_DEN[i];
        t.p1n[i] = HK_F64_P1_NUM[i]; t.p1d[i] = HK_F64_P1_DEN[i];
        t.q1n[i] = HK_F64_Q1_NUM[i]; t.q1d[i] = HK_F64_Q1_DEN[i];
    }
    return t;
}

static __constant__ HkF64Table kHkF64 = make_hk_f64_table();

// the table in the constant address space: uniform reads are scalar loads
typedef const __attribute__((address_space(4))) HkF64Table *HkF64Ptr;
typedef const __attribute__((address_space(4))) double *ConstF64Ptr;

template <int N>
__device__ __forceinline__ double horner_p(ConstF64Ptr c, double u) {
    double s = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) s = __builtin_fma(s, u, c[i]);
    return s;
}

// H0^(1)(x) and H1^(1)(x), x > 0, FP64.
__device__ __forceinline__ void hankel01_f64(double x, double &h0r, double &h0i,
                                             double &h1r, double &h1i) {
    HkF64Ptr tb = (HkF64Ptr)&kHkF64;
    asm volatile("" : "+s"(tb));  // opaque per call: no hoisting of the table
    if (x < HK_F64_XS) {
        double u = 0.25 * x * x;
        double a0 = horner_p<18>(tb->a0, u), a1 = horner_p<18>(tb->a1, u);
        double b0 = horner_p<18>(tb->b0, u), b1 = horner_p<18>(tb->b1, u);
        double L = log(0.5 * x);
        double ix = rcp64(x);
        double J0 = a0, J1 = 0.5 * x * a1;
        h0r = J0;
        h0i = kTwoOverPi * ((L + kEulerGamma) * J0 + b0);
        h1r = J1;
        h1i = -kTwoOverPi * ix + kTwoOverPi * L * J1 - (0.5 / kPi) * x * b1;
    } else {
        double ix = rcp64(x);
        double t = (HK_F64_XS * HK_F64_XS) * ix * ix;
        double P0 = horner_p<7>(tb->p0n, t) * rcp64(horner_p<7>(tb->p0d, t));
        double Q0 = horner_p<7>(tb->q0n, t) * rcp64(horner_p<7>(tb->q0d, t)) * ix;
        double P1 = horner_p<7>(tb->p1n, t) * rcp64(horner_p<7>(tb->p1d, t));
        double Q1 = horner_p<7>(tb->q1n, t) * rcp64(horner_p<7>(tb->q1d, t)) * ix;
        double amp = kSqrt2OverPi * rsq64(x);
        // phase x - pi/4 in revolutions: sin/cos(2 pi (x/(2 pi) - 1/8))
        double s, c;
        sincos_rev64(__builtin_fma(x, kInv2Pi, -0.125), s, c);
        // H0 = amp (P0 + i Q0)(c + i s); H1 = amp (P1 + i Q1)(s - i c)
        h0r = amp * (P0 * c - Q0 * s);
        h0i = amp * (P0 * s + Q0 * c);
        h1r = amp * (P1 * s + Q1 * c);
        h1i = amp * (Q1 * s - P1 * c);
    }
}

// MIXED: FP32 series/rationals and the FP32 sin/cos unit; the phase
// x - pi/4 is reduced in FP64 first.
__device__ __forceinline__ void hankel01_mixed(double xd, double &h0r, double &h0i,
                                               double &h1r, double &h1i) {
    float x = (float)xd;
    if (x < HK_F32_XS) {
        float u = 0.25f * x * x;
        float a0 = horner(HK_F32_A0, u), a1 = horner(HK_F32_A1, u);
        float b0 = horner(HK_F32_B0, u), b1 = horner(HK_F32_B1, u);
        float L = __builtin_amdgcn_logf(0.5f * x) * 0.69314718f;
        float ix = __builtin_amdgcn_rcpf(x);
        float J0 = a0, J1 = 0.5f * x * a1;
        h0r = J0;
        h0i = 0.63661977f * ((L + 0.57721566f) * J0 + b0);
        h1r = J1;
        h1i = -0.63661977f * ix + 0.63661977f * L * J1 - 0.15915494f * x * b1;
    } else {
        float ix = __builtin_amdgcn_rcpf(x);
        float t = (HK_F32_XS * HK_F32_XS) * ix * ix;
        float P0 = horner(HK_F32_P0_NUM, t) * __builtin_amdgcn_rcpf(horner(HK_F32_P0_DEN, t));
        float Q0 = horner(HK_F32_Q0_NUM, t) * __builtin_amdgcn_rcpf(horner(HK_F32_Q0_DEN, t)) * ix;
        float P1 = horner(HK_F32_P1_NUM, t) * __builtin_amdgcn_rcpf(horner(HK_F32_P1_DEN, t));
        float Q1 = horner(HK_F32_Q1_NUM, t) * __builtin_amdgcn_rcpf(horner(HK_F32_Q1_DEN, t)) * ix;
        float amp = 0.79788456f * __builtin_amdgcn_rsqf(x);
        double u = xd * 0.15915494309189533577 - 0.125;
        float f = (float)(u - __builtin_rint(u));
        float s = __builtin_amdgcn_sinf(f), c = __builtin_amdgcn_cosf(f);
        h0r = amp * (P0 * c - Q0 * s);
        h0i = amp * (P0 * s + Q0 * c);
        h1r = amp * (P1 * s + Q1 * c);
        h1i = amp * (Q1 * s - P1 * c);
    }
}

template <bool MIXED>
__device__ __forceinline__ void hankel01(double x, double &h0r, double &h0i, double &h1r,
                                         double &h1i) {
    if constexpr (MIXED) hankel01_mixed(x, h0r, h0i, h1r, h1i);
    else hankel01_f64(x, h0r, h0i, h1r, h1i);
}

}  // namespace inti
