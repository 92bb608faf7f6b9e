// Pair kernels K(x, y) of Inti.jl src/kernels.jl for Laplace (:137-204) and
// Helmholtz (:347-413), 2D and 3D, single layer (SL), double layer (DL),
// adjoint double layer (ADL), hypersingular (HS) and the combined fields
// c0*SL + c1*DL, c0*ADL + c1*HS, written for one target against one source.
// r = x - y; nx is the target normal, ny the source normal.
#pragma once
#include "common.hpp"
#include "devmath.hpp"

namespace inti {

struct KParams {
    double k;
    double c0r, c0i, c1r, c1i;
};

template <int DIM>
__device__ __forceinline__ double dotv(const double *a, const double *b) {
    double s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) s = __builtin_fma(a[d], b[d], s);
    return s;
}

// Laplace: real kernel value (constants included).
template <int DIM, int KIND, bool MIXED>
__device__ __forceinline__ double laplace_pair(const double *r, double d2s, const double *nx,
                                               const double *ny, const KParams &p) {
    if constexpr (KIND == INTI_SL) {
        if constexpr (DIM == 2) return -0.5 * kInv2Pi * logv<MIXED>(d2s);  // -log(d)/(2pi)
        else return kInv4Pi * rsqv<MIXED>(d2s);
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        const double sg = KIND == INTI_DL ? 1.0 : -1.0;
        double rdn = dotv<DIM>(r, KIND == INTI_DL ? ny : nx);
        if constexpr (DIM == 2) return (sg * kInv2Pi) * rdn * rcpv<MIXED>(d2s);
        else {
            double id = rsqv<MIXED>(d2s);
            return (sg * kInv4Pi) * rdn * (id * id * id);
        }
    } else if constexpr (KIND == INTI_HS) {
        double nxny = dotv<DIM>(nx, ny);
        double rdnx = dotv<DIM>(r, nx), rdny = dotv<DIM>(r, ny);
        if constexpr (DIM == 2) {
            double id2 = rcpv<MIXED>(d2s);
            return kInv2Pi * id2 * (nxny - 2.0 * rdnx * rdny * id2);
        } else {
            double id = rsqv<MIXED>(d2s);
            double id2 = id * id;
            return kInv4Pi * id2 * id * (nxny - 3.0 * rdnx * rdny * id2);
        }
    } else if constexpr (KIND == INTI_COMB_SD) {
        return p.c0r * laplace_pair<DIM, INTI_SL, MIXED>(r, d2s, nx, ny, p) +
               p.c1r * laplace_pair<DIM, INTI_DL, MIXED>(r, d2s, nx, ny, p);
    } else {
        return p.c0r * laplace_pair<DIM, INTI_ADL, MIXED>(r, d2s, nx, ny, p) +
               p.c1r * laplace_pair<DIM, INTI_HS, MIXED>(r, d2s, nx, ny, p);
    }
}

// Helmholtz 3D: complex kernel value.
template <int KIND, bool MIXED>
__device__ __forceinline__ void helm3_pair(const double *r, double d2s, const double *nx,
                                           const double *ny, const KParams &p, double &kr,
                                           double &ki) {
    const double k = p.k;
    double id = rsqp<MIXED>(d2s);
    double d = d2s * id;
    double s, c;
    sincosv<MIXED>(k * d, s, c);
    double br, bi;  // bracket B; K = e^{ikd} * B * scale / (4 pi)
    double scale;
    if constexpr (KIND == INTI_SL) {
        br = 1.0; bi = 0.0; scale = id;
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        double rdn = KIND == INTI_DL ? dotv<3>(r, ny) : -dotv<3>(r, nx);
        double a = rdn * id * id;  // (1/d - ik) rdn / d^2
        br = a * id; bi = -k * a; scale = 1.0;
    } else if constexpr (KIND == INTI_HS) {
        double nxny = dotv<3>(nx, ny);
        double id2 = id * id;
        double g = dotv<3>(r, nx) * dotv<3>(r, ny) * id2 * id;
        br = id * nxny + g * (k * k * d2s - 3.0);
        bi = 3.0 * k * d * g - k * nxny;
        scale = id2;
    } else if constexpr (KIND == INTI_COMB_SD) {
        double a = dotv<3>(r, ny) * id;
        double vr = a * id, vi = -k * a;  // DL bracket / SL bracket
        br = p.c0r + p.c1r * vr - p.c1i * vi;
        bi = p.c0i + p.c1r * vi + p.c1i * vr;
        scale = id;
    } else {  // COMB_AH = c0*ADL + c1*HS
        double nxny = dotv<3>(nx, ny);
        double rdnx = dotv<3>(r, nx);
        double id2 = id * id;
        double g = rdnx * dotv<3>(r, ny) * id2 * id;
        double hr = id * nxny + g * (k * k * d2s - 3.0);  // HS bracket (x id2)
        double hi = 3.0 * k * d * g - k * nxny;
        double a = -rdnx;                                 // ADL bracket (x id2)
        double ar = a * id, ai = -k * a;
        br = p.c0r * ar - p.c0i * ai + p.c1r * hr - p.c1i * hi;
        bi = p.c0r * ai + p.c0i * ar + p.c1r * hi + p.c1i * hr;
        scale = id2;
    }
    double cs = c * scale * kInv4Pi, ss = s * scale * kInv4Pi;
    kr = cs * br - ss * bi;
    ki = cs * bi + ss * br;
}

// Helmholtz 2D: complex kernel value via H0, H1 (H2 = 2 H1 / x - H0).
template <int KIND, bool MIXED>
__device__ __forceinline__ void helm2_pair(const double *r, double d2s, const double *nx,
                                           const double *ny, const KParams &p, double &kr,
                                           double &ki) {
    const double k = p.k;
    double id = rsqp<MIXED>(d2s);
    double x = k * (d2s * id);
    double h0r, h0i, h1r, h1i;
    hankel01<MIXED>(x, h0r, h0i, h1r, h1i);
    // K = i * (Zr + i Zi) / 4  with Z the bracket below
    double zr, zi;
    if constexpr (KIND == INTI_SL) {
        zr = h0r; zi = h0i;
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        double f = k * id * (KIND == INTI_DL ? dotv<2>(r, ny) : -dotv<2>(r, nx));
        zr = f * h1r; zi = f * h1i;
    } else if constexpr (KIND == INTI_HS) {
        double ix = rcpv<MIXED>(x);
        double h2r = 2.0 * ix * h1r - h0r, h2i = 2.0 * ix * h1i - h0i;
        double fa = -k * k * id * id * dotv<2>(r, nx) * dotv<2>(r, ny);
        double fb = k * id * dotv<2>(nx, ny);
        zr = fa * h2r + fb * h1r; zi = fa * h2i + fb * h1i;
    } else if constexpr (KIND == INTI_COMB_SD) {
        double f = k * id * dotv<2>(r, ny);
        double dr = f * h1r, di = f * h1i;
        zr = p.c0r * h0r - p.c0i * h0i + p.c1r * dr - p.c1i * di;
        zi = p.c0r * h0i + p.c0i * h0r + p.c1r * di + p.c1i * dr;
    } else {
        double ix = rcpv<MIXED>(x);
        double rdnx = dotv<2>(r, nx);
        double h2r = 2.0 * ix * h1r - h0r, h2i = 2.0 * ix * h1i - h0i;
        double fa = -k * k * id * id * rdnx * dotv<2>(r, ny);
        double fb = k * id * dotv<2>(nx, ny);
        double hr = fa * h2r + fb * h1r, hi = fa * h2i + fb * h1i;
        double f = -k * id * rdnx;
        double ar = f * h1r, ai = f * h1i;
        zr = p.c0r * ar - p.c0i * ai + p.c1r * hr - p.c1i * hi;
        zi = p.c0r * ai + p.c0i * ar + p.c1r * hi + p.c1i * hr;
    }
    kr = -0.25 * zi;
    ki = 0.25 * zr;
}

// Unified entry: value of K(x, y) (zero for coincident points).
template <int PDE, int DIM, int KIND, bool MIXED>
__device__ __forceinline__ void pair_value(const double *r, const double *nx, const double *ny,
                                           const KParams &p, double &kr, double &ki) {
    double d2 = dotv<DIM>(r, r);
    bool ok = d2 > kSamePointTol2;
    double d2s = ok ? d2 : 1.0;
    if constexpr (PDE == INTI_LAPLACE) {
        kr = laplace_pair<DIM, KIND, MIXED>(r, d2s, nx, ny, p);
        ki = 0.0;
    } else if constexpr (DIM == 3) {
        helm3_pair<KIND, MIXED>(r, d2s, nx, ny, p, kr, ki);
    } else {
        helm2_pair<KIND, MIXED>(r, d2s, nx, ny, p, kr, ki);
    }
    kr = ok ? kr : 0.0;
    ki = ok ? ki : 0.0;
}

}  // namespace inti
