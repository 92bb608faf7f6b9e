// Pair kernels K(x, y) of Inti.jl src/kernels.jl for Laplace (:137-204) and
// Helmholtz (:347-413), 2D and 3D, single layer (SL), double layer (DL),
// adjoint double layer (ADL), hypersingular (HS) and the combined fields
// c0*SL + c1*DL, c0*ADL + c1*HS, written for one target against one source.
// r = x - y; nx is the target normal, ny the source normal.
//
// Every kernel is split as K = post * raw: `raw` is what the direct sum
// accumulates per pair, `post` (a complex constant of the operator: 1/(4 pi),
// i k/4, the combined-field coefficient c1, ...) multiplies the finished sum
// once per target. Coincident points (d <= 1e-14) give raw = 0: the distance
// is clamped away from zero so that every intermediate stays finite, and one
// scalar factor of raw is zeroed.
#pragma once
#include "common.hpp"
#include "devmath.hpp"

namespace inti {

// Per-operator constants seen by the pair kernels.
// internal PDE code of the 3D Helmholtz kernels with a complex wavenumber
// k + i kimag (an operator with a nonzero imaginary part is routed to it)
constexpr int kHelmCK = 2;

struct KParams {
    double k;          // wavenumber (real part)
    double kimag;      // imaginary part of the wavenumber (kHelmCK only)
    double kappa;      // k / (2 pi): phase in revolutions per unit length
    double kappa_half; // kappa / 2
    double chr, chi;   // normalised combined-field coefficient (see post_factor)
    double postr, posti;
};

constexpr double kClampD2 = 1.0e-30;  // stand-in for d^2 = 0 (masked anyway)

template <int DIM>
__device__ __forceinline__ double dotv(const double *a, const double *b) {
    double s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) s = __builtin_fma(a[d], b[d], s);
    return s;
}

__device__ __forceinline__ double mask(bool ok, double v) { return ok ? v : 0.0; }

// ------------------------------------------------------------------ Laplace
// raw values; posts: SL2 -1/(4pi), SL3 1/(4pi), DL2 1/(2pi), DL3 1/(4pi),
// ADL2 -1/(2pi), ADL3 -1/(4pi), HS2 1/(2pi), HS3 1/(4pi).
template <int DIM, int KIND, bool MIXED>
__device__ __forceinline__ double laplace_raw(const double *r, double d2s, bool ok,
                                              const double *nx, const double *ny,
                                              const KParams &p) {
    if constexpr (KIND == INTI_SL) {
        if constexpr (DIM == 2) return mask(ok, logv<MIXED>(d2s));
        else return mask(ok, rsqv<MIXED>(d2s));
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        double rdn = dotv<DIM>(r, KIND == INTI_DL ? ny : nx);
        if constexpr (DIM == 2) return mask(ok, rdn) * rcpv<MIXED>(d2s);
        else {
            double id = rsqv<MIXED>(d2s);
            return mask(ok, rdn) * (id * id * id);
        }
    } else if constexpr (KIND == INTI_HS) {
        double nxny = dotv<DIM>(nx, ny);
        double rdnx = dotv<DIM>(r, nx), rdny = dotv<DIM>(r, ny);
        if constexpr (DIM == 2) {
            double id2 = mask(ok, rcpv<MIXED>(d2s));
            return id2 * (nxny - 2.0 * rdnx * rdny * id2);
        } else {
            double id = rsqv<MIXED>(d2s);
            double id2 = id * id;
            return mask(ok, id2 * id) * (nxny - 3.0 * rdnx * rdny * id2);
        }
    } else if constexpr (KIND == INTI_COMB_SD) {  // DL + ch*SL
        return laplace_raw<DIM, INTI_DL, MIXED>(r, d2s, ok, nx, ny, p) +
               p.chr * laplace_raw<DIM, INTI_SL, MIXED>(r, d2s, ok, nx, ny, p);
    } else {  // HS + ch*ADL
        return laplace_raw<DIM, INTI_HS, MIXED>(r, d2s, ok, nx, ny, p) +
               p.chr * laplace_raw<DIM, INTI_ADL, MIXED>(r, d2s, ok, nx, ny, p);
    }
}

// ---------------------------------------------------------------- Helmholtz 3D
// raw = e^{ikd} * f * B; posts: SL, DL, HS 1/(4pi), ADL -1/(4pi),
// COMB_SD c1/(4pi) with ch = c0/c1, COMB_AH c1/(4pi) with ch = c0/c1.
template <int KIND, bool MIXED>
__device__ __forceinline__ void helm3_raw(const double *r, double d2s, bool ok, const double *nx,
                                          const double *ny, const KParams &p, double &kr,
                                          double &ki) {
    const double k = p.k;
    double id = rsqp<MIXED>(d2s);
    double d = d2s * id;
    double s, c;
    sincos_rev<MIXED>(p.kappa * d, s, c);
    double br, bi, f;
    if constexpr (KIND == INTI_SL) {
        br = 1.0; bi = 0.0; f = mask(ok, id);
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        double a = mask(ok, dotv<3>(r, KIND == INTI_DL ? ny : nx) * id * id);
        br = a * id; bi = -k * a; f = 1.0;
    } else if constexpr (KIND == INTI_HS) {
        double nxny = dotv<3>(nx, ny);
        double id2 = id * id;
        double g = dotv<3>(r, nx) * dotv<3>(r, ny) * id2 * id;
        br = id * nxny + g * (k * k * d2s - 3.0);
        bi = 3.0 * k * d * g - k * nxny;
        f = mask(ok, id2);
    } else if constexpr (KIND == INTI_COMB_SD) {  // DL + ch*SL = e*id*[ch + (a id, -k a)]
        double a = dotv<3>(r, ny) * id;
        br = p.chr + a * id;
        bi = p.chi - k * a;
        f = mask(ok, id);
    } else {  // HS + ch*ADL = e*id2*[HS bracket - ch*rdnx*(id, -k)]
        double nxny = dotv<3>(nx, ny);
        double rdnx = dotv<3>(r, nx);
        double id2 = id * id;
        double g = rdnx * dotv<3>(r, ny) * id2 * id;
        double ar = -rdnx * id, ai = k * rdnx;
        br = id * nxny + g * (k * k * d2s - 3.0) + p.chr * ar - p.chi * ai;
        bi = 3.0 * k * d * g - k * nxny + p.chr * ai + p.chi * ar;
        f = mask(ok, id2);
    }
    double cf = c * f, sf = s * f;
    if constexpr (KIND == INTI_SL) {  // B = 1: no products with 1 and 0
        kr = cf;
        ki = sf;
    } else {
        kr = cf * br - sf * bi;
        ki = cf * bi + sf * br;
    }
}

// 3D Helmholtz with a complex wavenumber k = kr + i kc (kernels.jl:322-345
// take any k::Number; FP64 only). The same raw layout as helm3_raw, raw =
// e^{i kr d} e^{-kc d} f B, with every occurrence of ik in B written as
// -kc + i kr and k^2 as kr^2 - kc^2 + 2 i kr kc:
//   DL, ADL  B = a (1/d - ik),                 a = (r.n) / d^2
//   HS       B = nx.ny (1/d - ik) + g (k^2 d^2 + 3 ik d - 3),
//                                              g = (r.nx)(r.ny) / d^3
// (G'(d) = G (ik - 1/d), G''(d) = G ((ik - 1/d)^2 + 1/d^2)).
template <int KIND>
__device__ __forceinline__ void helm3c_raw(const double *r, double d2s, bool ok, const double *nx,
                                           const double *ny, const KParams &p, double &kr,
                                           double &ki) {
    const double k = p.k, kc = p.kimag;
    const double id = rsqp<false>(d2s);
    const double d = d2s * id;
    double s, c;
    sincos_rev<false>(p.kappa * d, s, c);
    const double damp = exp(-kc * d);
    double br, bi, f;
    if constexpr (KIND == INTI_SL) {
        br = 1.0; bi = 0.0; f = mask(ok, id);
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        const double a = mask(ok, dotv<3>(r, KIND == INTI_DL ? ny : nx) * id * id);
        br = a * (id + kc); bi = -k * a; f = 1.0;
    } else if constexpr (KIND == INTI_HS || KIND == INTI_COMB_AH) {
        const double nxny = dotv<3>(nx, ny);
        const double rdnx = dotv<3>(r, nx);
        const double g = rdnx * dotv<3>(r, ny) * id * id * id;
        br = nxny * (id + kc) + g * ((k * k - kc * kc) * d2s - 3.0 * kc * d - 3.0);
        bi = -k * nxny + g * (3.0 * k * d + 2.0 * k * kc * d2s);
        if constexpr (KIND == INTI_COMB_AH) {  // + ch * (-(r.nx) (1/d - ik)), see helm3_raw
            const double ar = -rdnx * (id + kc), ai = k * rdnx;
            br += p.chr * ar - p.chi * ai;
            bi += p.chr * ai + p.chi * ar;
        }
        f = mask(ok, id * id);
    } else {  // COMB_SD: DL + ch*SL = e*id*[ch + a (1/d - ik)], a = (r.ny)/d
        static_assert(KIND == INTI_COMB_SD, "kinds of helm3c_raw");
        const double a = dotv<3>(r, ny) * id;
        br = p.chr + a * (id + kc);
        bi = p.chi - k * a;
        f = mask(ok, id);
    }
    f *= damp;
    const double cf = c * f, sf = s * f;
    kr = cf * br - sf * bi;
    ki = cf * bi + sf * br;
}

// MIXED 3D Helmholtz SL, DL, ADL and DL + ch*SL with the fewest VALU
// operations: the FP32 estimate y of 1/d is used unrefined for amplitudes
// (relative error ~1e-7); only the phase kappa*d gets one Newton step,
// kappa*d = kappa*e*(1 + h/2) with e = d2*y, h = 1 - d2*y^2 (error ~h^2);
// the phase is reduced with fract (any whole number of revolutions drops
// out), and coincident points are masked on the FP32 amplitude factor. d2 is
// the unclamped squared distance; FP32 clamping keeps every intermediate
// finite for d2 = 0.
template <int KIND, bool MASK = true>
__device__ __forceinline__ void helm3_raw_mixed(const double *r, double d2, const double *nx,
                                                const double *ny, const KParams &p, double &kr,
                                                double &ki) {
    const bool ok = !MASK || d2 > kSamePointTol2;
    const float y32 = __builtin_amdgcn_rsqf(MASK ? fmaxf((float)d2, 1.0e-30f) : (float)d2);
    const double y = (double)y32;
    // kappa d from one Newton step on the estimate e = d2 y of d:
    // d = e (1 + h / 2) with h = 1 - e y, i.e. kappa d = (kappa / 2) e (3 - e y)
    const double e = d2 * y;
    const double t = (p.kappa_half * e) * __builtin_fma(-e, y, 3.0);
    const float u = (float)__builtin_amdgcn_fract(t);
    const float ym = ok ? y32 : 0.0f;
    const double cf = (double)(__builtin_amdgcn_cosf(u) * ym);
    const double sf = (double)(__builtin_amdgcn_sinf(u) * ym);
    if constexpr (KIND == INTI_SL) {
        kr = cf; ki = sf;
    } else {
        double br, bi;
        if constexpr (KIND == INTI_COMB_SD) {
            const double a = dotv<3>(r, ny) * y;
            br = __builtin_fma(a, y, p.chr);
            bi = __builtin_fma(-p.k, a, p.chi);
        } else {  // DL / ADL: e*id*[(a id, -k a)] with a = r.n id
            const double a = dotv<3>(r, KIND == INTI_DL ? ny : nx) * y;
            br = a * y;
            bi = -p.k * a;
        }
        kr = cf * br - sf * bi;
        ki = cf * bi + sf * br;
    }
}

// ---------------------------------------------------------------- Helmholtz 2D
// K = (i/4) Z; raw chosen so that post = i/4 (SL), i k/4 (DL, HS),
// -i k/4 (ADL), i c1 k/4 (COMB_SD with ch = c0/(c1 k), COMB_AH with
// ch = c0/c1).
// The radial part: 1/d, x = k d and H0(x), H1(x); shared by K(x_i, x_j) and
// K(x_j, x_i).
struct Helm2Radial {
    double id, x, h0r, h0i, h1r, h1i;
};
template <bool MIXED>
__device__ __forceinline__ Helm2Radial helm2_radial(double d2s, const KParams &p) {
    Helm2Radial R;
    R.id = rsqp<MIXED>(d2s);
    R.x = p.k * (d2s * R.id);
    hankel01<MIXED>(R.x, R.h0r, R.h0i, R.h1r, R.h1i);
    return R;
}

// MIXED: the Hankel values are FP32 (widened into R; narrowing them back is
// exact), so the geometric factors are rounded to FP32 once and the
// combination is done in FP32
template <int KIND>
__device__ __forceinline__ void helm2_combine32(const double *r, bool ok, const double *nx,
                                                const double *ny, const KParams &p,
                                                const Helm2Radial &R, double &kr, double &ki) {
    const double id = R.id;
    const float h0r = (float)R.h0r, h0i = (float)R.h0i, h1r = (float)R.h1r, h1i = (float)R.h1i;
    float zr, zi;
    if constexpr (KIND == INTI_SL) {
        zr = h0r; zi = h0i;
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        const float f = (float)(id * dotv<2>(r, KIND == INTI_DL ? ny : nx));
        zr = f * h1r; zi = f * h1i;
    } else if constexpr (KIND == INTI_COMB_SD) {  // rdny id H1 + ch H0
        const float f = (float)(id * dotv<2>(r, ny));
        const float chr = (float)p.chr, chi = (float)p.chi;
        zr = f * h1r + chr * h0r - chi * h0i;
        zi = f * h1i + chr * h0i + chi * h0r;
    } else {
        const float ix = (float)rcpv<true>(R.x);
        const float h2r = 2.0f * ix * h1r - h0r, h2i = 2.0f * ix * h1i - h0i;
        const double rdnx = dotv<2>(r, nx);
        const float fa = (float)(-p.k * id * id * rdnx * dotv<2>(r, ny));
        const float fb = (float)(id * dotv<2>(nx, ny));
        if constexpr (KIND == INTI_HS) {
            zr = fa * h2r + fb * h1r; zi = fa * h2i + fb * h1i;
        } else {  // HS raw - ch * rdnx id H1
            const float fc = (float)(-rdnx * id);
            const float chr = (float)p.chr, chi = (float)p.chi;
            const float ar = fc * h1r, ai = fc * h1i;
            zr = fa * h2r + fb * h1r + chr * ar - chi * ai;
            zi = fa * h2i + fb * h1i + chr * ai + chi * ar;
        }
    }
    kr = (double)(ok ? zr : 0.0f);
    ki = (double)(ok ? zi : 0.0f);
}

template <int KIND, bool MIXED>
__device__ __forceinline__ void helm2_combine(const double *r, bool ok, const double *nx,
                                              const double *ny, const KParams &p,
                                              const Helm2Radial &R, double &kr, double &ki) {
    if constexpr (MIXED) {
        helm2_combine32<KIND>(r, ok, nx, ny, p, R, kr, ki);
        return;
    }
    const double k = p.k, id = R.id;
    const double h0r = R.h0r, h0i = R.h0i, h1r = R.h1r, h1i = R.h1i;
    if constexpr (KIND == INTI_SL) {
        kr = mask(ok, h0r); ki = mask(ok, h0i);
    } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
        double f = mask(ok, id * dotv<2>(r, KIND == INTI_DL ? ny : nx));
        kr = f * h1r; ki = f * h1i;
    } else if constexpr (KIND == INTI_HS) {
        double ix = rcpv<MIXED>(R.x);
        double h2r = 2.0 * ix * h1r - h0r, h2i = 2.0 * ix * h1i - h0i;
        double fa = mask(ok, -k * id * id * dotv<2>(r, nx) * dotv<2>(r, ny));
        double fb = mask(ok, id * dotv<2>(nx, ny));
        kr = fa * h2r + fb * h1r; ki = fa * h2i + fb * h1i;
    } else if constexpr (KIND == INTI_COMB_SD) {  // rdny id H1 + ch H0
        double f = id * dotv<2>(r, ny);
        double zr = f * h1r + p.chr * h0r - p.chi * h0i;
        double zi = f * h1i + p.chr * h0i + p.chi * h0r;
        kr = mask(ok, zr); ki = mask(ok, zi);
    } else {  // HS raw - ch * rdnx id H1
        double ix = rcpv<MIXED>(R.x);
        double rdnx = dotv<2>(r, nx);
        double h2r = 2.0 * ix * h1r - h0r, h2i = 2.0 * ix * h1i - h0i;
        double fa = -k * id * id * rdnx * dotv<2>(r, ny);
        double fb = id * dotv<2>(nx, ny);
        double fc = -rdnx * id;
        double ar = fc * h1r, ai = fc * h1i;
        double zr = fa * h2r + fb * h1r + p.chr * ar - p.chi * ai;
        double zi = fa * h2i + fb * h1i + p.chr * ai + p.chi * ar;
        kr = mask(ok, zr); ki = mask(ok, zi);
    }
}

template <int KIND, bool MIXED>
__device__ __forceinline__ void helm2_raw(const double *r, double d2s, bool ok, const double *nx,
                                          const double *ny, const KParams &p, double &kr,
                                          double &ki) {
    helm2_combine<KIND, MIXED>(r, ok, nx, ny, p, helm2_radial<MIXED>(d2s, p), kr, ki);
}

// ------------------------------------------------------ Helmholtz far field
// raw = e^{-i k xh.y} B, B = 1 (SL), -i k xh.n_y (DL), ch + (-i k xh.n_y)
// (FF_COMB, post c1 P, ch = c0/c1); xh is the target "coordinate", y the
// source node. Smooth: nothing to mask.
template <int DIM, int KIND, bool MIXED>
__device__ __forceinline__ void helm_ff_raw(const double *xh, const double *y, const double *ny,
                                            const KParams &p, double &kr, double &ki) {
    double s, c;
    sincos_rev<MIXED>(-p.kappa * dotv<DIM>(xh, y), s, c);
    if constexpr (KIND == INTI_FF_SL) {
        kr = c; ki = s;
    } else {
        const double bi = -p.k * dotv<DIM>(xh, ny);
        const double br = KIND == INTI_FF_COMB ? p.chr : 0.0;
        const double bim = KIND == INTI_FF_COMB ? bi + p.chi : bi;
        kr = c * br - s * bim;
        ki = c * bim + s * br;
    }
}

// raw value of the pair (post applied by the caller). MASK = false is for
// pairs known to be farther apart than the coincidence tolerance (the
// quadrature's nodes are distinct and the two nodes differ): no mask, no
// clamp, the same values.
template <int PDE, int DIM, int KIND, bool MIXED, bool MASK = true>
__device__ __forceinline__ void pair_raw(const double *r, const double *nx, const double *ny,
                                         const KParams &p, double &kr, double &ki) {
    static_assert(KIND < INTI_FF_SL, "far-field kinds go through pair_raw_xy");
    double d2 = dotv<DIM>(r, r);
    bool ok = !MASK || d2 > kSamePointTol2;
    double d2s = MASK ? fmax(d2, kClampD2) : d2;
    if constexpr (PDE == INTI_LAPLACE && DIM == 3 && MIXED) {
        // 1/d from the FP32 unit, masked in FP32: zero for coincident points,
        // and every kind below vanishes with it
        const float y32 = __builtin_amdgcn_rsqf(MASK ? fmaxf((float)d2, 1.0e-30f) : (float)d2);
        const double y = (double)(ok ? y32 : 0.0f);
        if constexpr (KIND == INTI_SL) {
            kr = y;
        } else if constexpr (KIND == INTI_DL || KIND == INTI_ADL) {
            kr = dotv<3>(r, KIND == INTI_DL ? ny : nx) * (y * y * y);
        } else if constexpr (KIND == INTI_HS) {
            const double y2 = y * y;
            kr = (y2 * y) * (dotv<3>(nx, ny) - 3.0 * dotv<3>(r, nx) * dotv<3>(r, ny) * y2);
        } else if constexpr (KIND == INTI_COMB_SD) {  // DL + ch*SL
            kr = __builtin_fma(dotv<3>(r, ny), y * y * y, p.chr * y);
        } else {  // HS + ch*ADL
            const double y2 = y * y;
            const double rdnx = dotv<3>(r, nx);
            kr = (y2 * y) * (dotv<3>(nx, ny) - 3.0 * rdnx * dotv<3>(r, ny) * y2 + p.chr * rdnx);
        }
        ki = 0.0;
    } else if constexpr (PDE == INTI_LAPLACE) {
        kr = laplace_raw<DIM, KIND, MIXED>(r, d2s, ok, nx, ny, p);
        ki = 0.0;
    } else if constexpr (PDE == kHelmCK) {
        static_assert(DIM == 3 && !MIXED, "complex wavenumbers: 3D, FP64");
        helm3c_raw<KIND>(r, d2s, ok, nx, ny, p, kr, ki);
    } else if constexpr (DIM == 3 && MIXED &&
                         (KIND == INTI_SL || KIND == INTI_DL || KIND == INTI_ADL ||
                          KIND == INTI_COMB_SD)) {
        helm3_raw_mixed<KIND, MASK>(r, d2, nx, ny, p, kr, ki);
    } else if constexpr (DIM == 3) {
        helm3_raw<KIND, MIXED>(r, d2s, ok, nx, ny, p, kr, ki);
    } else {
        helm2_raw<KIND, MIXED>(r, d2s, ok, nx, ny, p, kr, ki);
    }
}

// Pair value from the two points: the far-field kinds need the direction
// and the source node themselves, the others only r = x - y.
template <int PDE, int DIM, int KIND, bool MIXED>
__device__ __forceinline__ void pair_raw_xy(const double *x, const double *nx, const double *y,
                                            const double *ny, const KParams &p, double &kr,
                                            double &ki) {
    if constexpr (KIND >= INTI_FF_SL) {
        static_assert(PDE == INTI_HELMHOLTZ, "far fields are Helmholtz only");
        helm_ff_raw<DIM, KIND, MIXED>(x, y, ny, p, kr, ki);
    } else {
        double r[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) r[d] = x[d] - y[d];
        pair_raw<PDE, DIM, KIND, MIXED>(r, nx, ny, p, kr, ki);
    }
}

// Both directions of one unordered pair of a quadrature with itself:
// K(x_i, x_j) with target normal ni and source normal nj, and K(x_j, x_i);
// r = x_i - x_j. The shared radial parts are computed once: explicitly for
// the 2D Helmholtz kernels (the Hankel evaluation branches, so two calls
// would not be merged), as common subexpressions of two calls otherwise.
template <int PDE, int DIM, int KIND, bool MIXED, bool MASK = true>
__device__ __forceinline__ void pair_raw_sym(const double *r, const double *ni, const double *nj,
                                             const KParams &p, double &k1r, double &k1i,
                                             double &k2r, double &k2i) {
    double rn[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) rn[d] = -r[d];
    if constexpr (PDE == INTI_HELMHOLTZ && DIM == 2) {
        const double d2 = dotv<2>(r, r);
        const bool ok = !MASK || d2 > kSamePointTol2;
        const Helm2Radial R = helm2_radial<MIXED>(MASK ? fmax(d2, kClampD2) : d2, p);
        helm2_combine<KIND, MIXED>(r, ok, ni, nj, p, R, k1r, k1i);
        helm2_combine<KIND, MIXED>(rn, ok, nj, ni, p, R, k2r, k2i);
    } else {
        pair_raw<PDE, DIM, KIND, MIXED, MASK>(r, ni, nj, p, k1r, k1i);
        pair_raw<PDE, DIM, KIND, MIXED, MASK>(rn, nj, ni, p, k2r, k2i);
    }
}

// Fused value-and-accumulate of both directions of one unordered pair for
// the 3D Helmholtz DL and DL + ch*SL kinds. Their raw value times a density
// q factors as
//   raw q = (E/d) [ch q + a (1/d - i k) q],   a = (r . n_s) / d,
// with E/d = e^{ikd}/d shared by the two directions, so with ch q, k Im q
// and -k Re q staged per node beside q the source term costs two FMAs for
// w = (1/d - ik) q, two for ch q + a w and four for the complex accumulation
// (instead of forming the complex kernel value first). r = x_i - x_j; the
// direction (i <- j) adds to (a1r, a1i) with the source normal nj and q_j,
// the direction (j <- i) to (a2r, a2i) with ni and q_i. TRANS accumulates
// the transposed sums (K(x_j, x_i) q_j at i, K(x_i, x_j) q_i at j).
// The mixed symmetric form: the distance, 1/d and the phase as in
// helm3_raw_mixed (FP64 geometry and phase, FP32 transcendentals); the
// staged density terms, both directions' products and their running sums in
// FP32 (non-packed FP32 arithmetic issues at twice the FP64 rate). The
// caller adds the FP32 sums into FP64 every 64 sources, so an FP32 sum never
// holds more than 64 terms.
struct QExt32 {
    float qr, qi, cr, ci, kqi, mkqr;  // q, ch q, k Im q, -k Re q
};
__device__ __forceinline__ QExt32 qext32(double qr, double qi, const KParams &p) {
    QExt32 q;
    q.qr = (float)qr;
    q.qi = (float)qi;
    q.cr = (float)(p.chr * qr - p.chi * qi);
    q.ci = (float)(p.chr * qi + p.chi * qr);
    q.kqi = (float)(p.k * qi);
    q.mkqr = (float)(-p.k * qr);
    return q;
}
template <bool MASK, bool TRANS>
__device__ __forceinline__ void helm3_sym_fused32(const double *r, const float *ni, const float *nj,
                                                  const KParams &p, const QExt32 &qj, const QExt32 &qi,
                                                  float &a1r, float &a1i, float &a2r, float &a2i) {
    const double d2 = dotv<3>(r, r);
    const bool ok = !MASK || d2 > kSamePointTol2;
    const float y = __builtin_amdgcn_rsqf(MASK ? fmaxf((float)d2, 1.0e-30f) : (float)d2);
    const double id = (double)y;
    const double e = d2 * id;
    const double t = (p.kappa_half * e) * __builtin_fma(-e, id, 3.0);  // kappa d, one Newton step
    const float u = (float)__builtin_amdgcn_fract(t);
    const float ym = ok ? y : 0.0f;
    const float cf = __builtin_amdgcn_cosf(u) * ym;
    const float sf = __builtin_amdgcn_sinf(u) * ym;
    // the amplitudes (r . n) / d in FP32 (r rounded once; 1e-7 of the matvec)
    const float r0 = (float)r[0], r1 = (float)r[1], r2 = (float)r[2];
    const float aj = __builtin_fmaf(r0, nj[0], __builtin_fmaf(r1, nj[1], r2 * nj[2])) * y;    // K(x_i, x_j)
    const float ai = -__builtin_fmaf(r0, ni[0], __builtin_fmaf(r1, ni[1], r2 * ni[2])) * y;   // K(x_j, x_i), r -> -r
    const float b1 = TRANS ? ai : aj, b2 = TRANS ? aj : ai;
    // (i <- j): q_j
    const float w1r = __builtin_fmaf(y, qj.qr, qj.kqi), w1i = __builtin_fmaf(y, qj.qi, qj.mkqr);
    const float z1r = __builtin_fmaf(b1, w1r, qj.cr), z1i = __builtin_fmaf(b1, w1i, qj.ci);
    a1r = __builtin_fmaf(cf, z1r, __builtin_fmaf(-sf, z1i, a1r));
    a1i = __builtin_fmaf(cf, z1i, __builtin_fmaf(sf, z1r, a1i));
    // (j <- i): q_i
    const float w2r = __builtin_fmaf(y, qi.qr, qi.kqi), w2i = __builtin_fmaf(y, qi.qi, qi.mkqr);
    const float z2r = __builtin_fmaf(b2, w2r, qi.cr), z2i = __builtin_fmaf(b2, w2i, qi.ci);
    a2r = __builtin_fmaf(cf, z2r, __builtin_fmaf(-sf, z2i, a2r));
    a2i = __builtin_fmaf(cf, z2i, __builtin_fmaf(sf, z2r, a2i));
}

// One direction of the fused 3D Helmholtz DL / DL + ch*SL pair (target x,
// source y with normal ny, r = x - y), mixed precision, coincident points
// masked: acc += raw q = (E/d) [ch q + a (1/d - i k) q], a = (r . ny) / d,
// with FP32 products and an FP32 running sum (the caller adds it into FP64
// every 64 sources); the distance, the phase and the amplitude's dot
// product stay FP64.
__device__ __forceinline__ void helm3_fused_one32(const double *r, const double *ny, const KParams &p,
                                                  const QExt32 &q, float &ar, float &ai) {
    const double d2 = dotv<3>(r, r);
    const bool ok = d2 > kSamePointTol2;
    const float y = __builtin_amdgcn_rsqf(fmaxf((float)d2, 1.0e-30f));
    const double id = (double)y;
    const double e = d2 * id;
    const double t = (p.kappa_half * e) * __builtin_fma(-e, id, 3.0);  // kappa d, one Newton step
    const float u = (float)__builtin_amdgcn_fract(t);
    const float ym = ok ? y : 0.0f;
    const float cf = __builtin_amdgcn_cosf(u) * ym;
    const float sf = __builtin_amdgcn_sinf(u) * ym;
    const float b = (float)dotv<3>(r, ny) * y;
    const float wr = __builtin_fmaf(y, q.qr, q.kqi), wi = __builtin_fmaf(y, q.qi, q.mkqr);
    const float zr = __builtin_fmaf(b, wr, q.cr), zi = __builtin_fmaf(b, wi, q.ci);
    ar = __builtin_fmaf(cf, zr, __builtin_fmaf(-sf, zi, ar));
    ai = __builtin_fmaf(cf, zi, __builtin_fmaf(sf, zr, ai));
}

inline KParams basic_params(double k) {
    KParams p;
    p.k = k;
    p.kimag = 0.0;
    p.kappa = k * kInv2Pi;
    p.kappa_half = 0.5 * p.kappa;
    p.chr = p.chi = 0.0;
    p.postr = 1.0; p.posti = 0.0;
    return p;
}

// Kernel parameters of an operator and the kind actually evaluated (a
// combined kind with c1 = 0 is its first component).
KParams op_params(const Op *op, int *eff_kind);

// post factor of the four basic kinds (host and device)
__host__ __device__ inline void basic_post(int pde, int dim, int kind, double k, double &pr,
                                           double &pi) {
    pr = 0.0; pi = 0.0;
    if (kind == INTI_FF_SL || kind == INTI_FF_DL) {  // Helmholtz only
        if (dim == 3) pr = kInv4Pi;
        else pr = pi = 0.25 / sqrt(kPi * k);  // e^{i pi/4} / sqrt(8 pi k)
        return;
    }
    if (pde == INTI_LAPLACE) {
        if (dim == 2) pr = kind == INTI_SL ? -0.5 * kInv2Pi : kind == INTI_ADL ? -kInv2Pi : kInv2Pi;
        else pr = kind == INTI_ADL ? -kInv4Pi : kInv4Pi;
    } else if (dim == 3) {
        pr = kind == INTI_ADL ? -kInv4Pi : kInv4Pi;
    } else {
        pi = kind == INTI_SL ? 0.25 : kind == INTI_ADL ? -0.25 * k : 0.25 * k;
    }
}

// Full value K = post * raw for the four basic kinds (setup kernels).
template <int PDE, int DIM, int KIND>
__device__ __forceinline__ void pair_value(const double *r, const double *nx, const double *ny,
                                           const KParams &p, double &kr, double &ki) {
    static_assert(KIND <= INTI_HS || KIND == INTI_FF_SL || KIND == INTI_FF_DL,
                  "combined kinds carry their post factor in KParams");
    double rr, ri, pr, pi;
    pair_raw<PDE, DIM, KIND, false>(r, nx, ny, p, rr, ri);
    basic_post(PDE, DIM, KIND, p.k, pr, pi);
    kr = pr * rr - pi * ri;
    ki = pr * ri + pi * rr;
}

}  // namespace inti
