// Reverse mode of the direct sum y_i = sum_j K(x_i, y_j) w_j sigma_j: given
// the cotangent ubar of y (ChainRules convention, the gradient of a real
// objective J with respect to Re y + i Im y), the pullback returns
//   sigma_bar_j = w_j conj(sum_i c_i K_ij)              (= A^H ubar)
//   ybar_j      = Re(s_j sum_i c_i dK_ij/dy_j)          (source nodes)
//   nbar_j      = Re(s_j sum_i c_i dK_ij/dn_j)          (source normals)
//   wbar_j      = Re(sigma_j sum_i c_i K_ij)            (source weights)
//   xbar_i      = Re(c_i sum_j s_j dK_ij/dx_i)          (target nodes)
//   kbar        = Re(sum_ij c_i dK_ij/dk s_j)           (wavenumber)
// with c_i = conj(ubar_i), s_j = w_j sigma_j. Kernel values are recomputed,
// never stored: the source-side sums run source-major (one thread per
// source, targets staged in LDS), the target-side sums target-major, each
// with the other index range cut into chunks whose partial sums a second
// kernel adds in a fixed order (deterministic, no atomics).
//
// Kernels differentiated (src/kernels.jl:137-204, 347-413): single layer,
// double layer, c0 SL + c1 DL, and the Helmholtz far-field patterns. With
// r = x - y, d = |r|, the translation-invariant ones are
//   SL: K = g(d),            DL: K = h(d) (r . n_y),   h = -g'(d)/d,
// so dK/dy = -g'(d) r/d (SL), -(h'(d) (r.n_y) r/d + h(d) n_y) (DL),
// dK/dn_y = h(d) r, dK/dx = -dK/dy, with (E = e^{ikd}/(4 pi)):
//   Laplace 2D   g = -log(d)/(2 pi)     h = 1/(2 pi d^2)
//   Laplace 3D   g = 1/(4 pi d)         h = 1/(4 pi d^3)
//   Helmholtz 3D g = E/d, g_k = i E     h = E (1 - ikd)/d^3, h_k = E k/d,
//                h'/d = E (k^2 d^2 - 3 + 3ikd)/d^5
//   Helmholtz 2D g = (i/4) H0(kd), g_k = -(i/4) d H1(kd),
//                h = (ik/4) H1(kd)/d, h_k = (ik/4) H0(kd),
//                h'/d = (ik/4) (k H0(kd)/d^2 - 2 H1(kd)/d^3).
#include <complex>
#include <type_traits>
#include <vector>

#include "apply_sym.hpp"
#include "common.hpp"
#include "pair.hpp"

namespace inti {

namespace {

__device__ __forceinline__ double2 c2(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ double2 cm(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cs(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 ca(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {  // acc += a*b
    acc.x = __builtin_fma(a.x, b.x, __builtin_fma(-a.y, b.y, acc.x));
    acc.y = __builtin_fma(a.x, b.y, __builtin_fma(a.y, b.x, acc.y));
}

struct AdjParams {
    double k, kappa;
    double2 c0, c1, dc0, dc1;  // combined-field coefficients and d/dk
    double2 P, Pk;             // far-field prefactor and dP/dk
};

// g, g'/d, g_k (SL) and h, h'/d, h_k (DL) at distance d (d2 = d^2 > 0)
template <int PDE, int DIM, bool WANT_G, bool WANT_H>
__device__ __forceinline__ void radial(double d2, const AdjParams &p, double2 &g, double2 &gpd,
                                       double2 &gk, double2 &h, double2 &hpd, double2 &hk) {
    const double id = rsq64(d2);
    const double d = d2 * id;
    const double id2 = id * id;
    if constexpr (PDE == INTI_LAPLACE) {
        gk = hk = c2(0, 0);
        if constexpr (DIM == 2) {
            g = c2(-0.5 * kInv2Pi * log(d2), 0);
            gpd = c2(-kInv2Pi * id2, 0);
            h = c2(kInv2Pi * id2, 0);
            hpd = c2(-2.0 * kInv2Pi * id2 * id2, 0);
        } else {
            const double id3 = id2 * id;
            g = c2(kInv4Pi * id, 0);
            gpd = c2(-kInv4Pi * id3, 0);
            h = c2(kInv4Pi * id3, 0);
            hpd = c2(-3.0 * kInv4Pi * id3 * id2, 0);
        }
    } else if constexpr (DIM == 3) {
        const double k = p.k;
        double s, c;
        sincos_rev64_vop3(p.kappa * d, s, c);
        const double2 E = c2(kInv4Pi * c, kInv4Pi * s);
        const double id3 = id2 * id;
        if constexpr (WANT_G) {
            g = cs(E, id);
            gpd = cs(cm(E, c2(-1.0, k * d)), id3);
            gk = c2(-E.y, E.x);
        }
        if constexpr (WANT_H) {
            h = cs(cm(E, c2(1.0, -k * d)), id3);
            hpd = cs(cm(E, c2(k * k * d2 - 3.0, 3.0 * k * d)), id3 * id2);
            hk = cs(E, k * id);
        }
    } else {
        const double k = p.k;
        double h0r, h0i, h1r, h1i;
        hankel01<false>(k * d, h0r, h0i, h1r, h1i);
        // (i/4) z = (-z.y/4, z.x/4)
        if constexpr (WANT_G) {
            g = c2(-0.25 * h0i, 0.25 * h0r);
            gpd = c2(0.25 * k * h1i * id, -0.25 * k * h1r * id);  // -(ik/4) H1 / d
            gk = c2(0.25 * d * h1i, -0.25 * d * h1r);            // -(i/4) d H1
        }
        if constexpr (WANT_H) {
            const double q = 0.25 * k;
            h = c2(-q * h1i * id, q * h1r * id);  // (ik/4) H1 / d
            const double ar = k * h0r * id2 - 2.0 * h1r * id2 * id;
            const double ai = k * h0i * id2 - 2.0 * h1i * id2 * id;
            hpd = c2(-q * ai, q * ar);
            hk = c2(-q * h0i, q * h0r);
        }
    }
}

// The radial functions of one pair distance (shared by K(x_i, x_j) and
// K(x_j, x_i)), with the coincident-point flag.
// For c0 SL + c1 DL the coefficient products are formed here once per
// distance: g <- c0 g, gpd <- c0 gpd, gk <- dc0 g + c0 gk, h <- c1 h,
// hpd <- c1 hpd, hk <- dc1 h + c1 hk.
struct AdjRadial {
    double2 g, gpd, gk, h, hpd, hk;
    bool ok;
};
// MASK = false: the pair is known to be farther apart than the coincidence
// tolerance (distinct nodes, different nodes): no clamp, no mask.
// CHIK: the combined field of the scattering examples, c0 = -i k, c1 = 1,
// dc0/dk = -i, dc1/dk = 0 (3D Helmholtz), with the coefficient products
// written out instead of taken as general complex products, and without the
// factor 1/(4 pi) of E: every output of the pair is linear in E, and
// k_adj_sym applies it once to its sums (OSC).
template <int PDE, int DIM, int KIND, bool MASK = true, bool CHIK = false>
__device__ __forceinline__ AdjRadial adj_radial(double d2, const AdjParams &p) {
    constexpr bool WG = KIND != INTI_DL, WH = KIND != INTI_SL;
    AdjRadial R;
    R.ok = !MASK || d2 > kSamePointTol2;
    if constexpr (CHIK) {
        static_assert(PDE == INTI_HELMHOLTZ && DIM == 3 && KIND == INTI_COMB_SD, "CHIK is the 3D combined field");
        const double d2s = MASK ? fmax(d2, kClampD2) : d2;
        const double id = rsq64(d2s), d = d2s * id, id2 = id * id, id3 = id2 * id;
        double s, c;
        sincos_rev64_vop3(p.kappa * d, s, c);
        const double2 E = c2(c, s);  // e^{ikd}; the 1/(4 pi) goes to the sums
        const double k = p.k, kd = k * d, kid = k * id;
        const double2 E0 = c2(k * E.y, -k * E.x);  // -i k E
        R.g = cs(E0, id);
        R.gpd = cs(c2(-E0.x - kd * E0.y, kd * E0.x - E0.y), id3);
        R.gk = c2(E.y * id - E0.y, E0.x - E.x * id);  // dc0 E / d + c0 i E, dc0 E = -i E
        R.h = cs(c2(E.x + kd * E.y, E.y - kd * E.x), id3);
        const double A = kd * kd - 3.0, B = 3.0 * kd;
        R.hpd = cs(c2(E.x * A - E.y * B, E.x * B + E.y * A), id3 * id2);
        R.hk = cs(E, kid);  // c1 E k / d
        return R;
    }
    if constexpr (PDE == INTI_HELMHOLTZ && DIM == 3 && KIND == INTI_COMB_SD) {
        // every radial function is E = e^{ikd}/(4 pi) times a simple factor:
        // fold the coefficients into E once (E0 = c0 E, E1 = c1 E, D0 = dc0 E,
        // D1 = dc1 E) instead of into each of the six values
        const double d2s = MASK ? fmax(d2, kClampD2) : d2;
        const double id = rsq64(d2s), d = d2s * id, id2 = id * id, id3 = id2 * id;
        double s, c;
        sincos_rev64_vop3(p.kappa * d, s, c);
        const double2 E = c2(kInv4Pi * c, kInv4Pi * s);
        const double2 E0 = cm(p.c0, E), E1 = cm(p.c1, E), D0 = cm(p.dc0, E), D1 = cm(p.dc1, E);
        const double kd = p.k * d, kid = p.k * id;
        R.g = cs(E0, id);                                              // c0 E / d
        R.gpd = cs(c2(-E0.x - kd * E0.y, kd * E0.x - E0.y), id3);      // c0 E (-1 + ikd) / d^3
        R.gk = c2(D0.x * id - E0.y, D0.y * id + E0.x);                 // dc0 E / d + c0 i E
        R.h = cs(c2(E1.x + kd * E1.y, E1.y - kd * E1.x), id3);         // c1 E (1 - ikd) / d^3
        const double A = kd * kd - 3.0, B = 3.0 * kd;
        R.hpd = cs(c2(E1.x * A - E1.y * B, E1.x * B + E1.y * A), id3 * id2);  // c1 E (k^2d^2 - 3 + 3ikd) / d^5
        R.hk = c2((D1.x + kd * D1.y) * id3 + E1.x * kid,               // dc1 E (1 - ikd) / d^3 + c1 E k / d
                  (D1.y - kd * D1.x) * id3 + E1.y * kid);
        return R;
    }
    R.g = R.gpd = R.gk = R.h = R.hpd = R.hk = c2(0, 0);
    radial<PDE, DIM, WG, WH>(MASK ? fmax(d2, kClampD2) : d2, p, R.g, R.gpd, R.gk, R.h, R.hpd, R.hk);
    if constexpr (KIND == INTI_COMB_SD) {  // (3D Helmholtz returned above)
        const double2 g = R.g, h = R.h;
        R.gk = ca(cm(p.dc0, g), cm(p.c0, R.gk));
        R.g = cm(p.c0, g);
        R.gpd = cm(p.c0, R.gpd);
        R.hk = ca(cm(p.dc1, h), cm(p.c1, R.hk));
        R.h = cm(p.c1, h);
        R.hpd = cm(p.c1, R.hpd);
    }
    return R;
}

// K, dK/dk and the factored derivatives (below) of the translation-invariant
// kinds for v = x - y and the source normal ny, from the radial functions.
// SL: K = g, dK/dy = -gpd r;  DL: K = h rn, dK/dy = -(hpd rn r + h n),
// dK/dn = h r (rn = r . n_y); coincident points give zeros.
template <int PDE, int DIM, int KIND, bool MASK = true>
__device__ __forceinline__ void adj_combine(const double *v, const double *ny, const AdjRadial &R,
                                            const AdjParams &p, double2 &K, double2 &gk,
                                            double2 &a_r, double2 &a_n, double2 &b_r) {
    const double rn = KIND != INTI_SL ? dotv<DIM>(v, ny) : 0.0;
    if constexpr (KIND == INTI_SL) {
        K = R.g; gk = R.gk; a_r = cs(R.gpd, -1.0); a_n = c2(0, 0); b_r = c2(0, 0);
    } else if constexpr (KIND == INTI_DL) {
        K = cs(R.h, rn); gk = cs(R.hk, rn); a_r = cs(R.hpd, -rn); a_n = cs(R.h, -1.0); b_r = R.h;
    } else {  // c0 SL + c1 DL, coefficients folded into R
        K = make_double2(__builtin_fma(R.h.x, rn, R.g.x), __builtin_fma(R.h.y, rn, R.g.y));
        gk = make_double2(__builtin_fma(R.hk.x, rn, R.gk.x), __builtin_fma(R.hk.y, rn, R.gk.y));
        a_r = make_double2(-__builtin_fma(R.hpd.x, rn, R.gpd.x), -__builtin_fma(R.hpd.y, rn, R.gpd.y));
        b_r = R.h;
        a_n = cs(b_r, -1.0);
    }
    if constexpr (MASK) {
        const double mk = R.ok ? 1.0 : 0.0;
        K = cs(K, mk); gk = cs(gk, mk); a_r = cs(a_r, mk); a_n = cs(a_n, mk); b_r = cs(b_r, mk);
    }
}

// Value and derivatives of one pair, x the target (or far-field direction),
// in the factored form
//   dK/dy = a_r v + a_n n_y,   dK/dn_y = b_r v,
// with v = r = x - y for the translation-invariant kernels and v = xh for
// the far fields (where a_n = 0); gk = dK/dk. The sums then carry a_r v,
// a_n and b_r v, and the per-source normal is applied once at the end.
template <int PDE, int DIM, int KIND>
__device__ __forceinline__ void pair_grad(const double *x, const double *y, const double *ny,
                                          const AdjParams &p, double (&v)[DIM], double2 &K,
                                          double2 &a_r, double2 &a_n, double2 &b_r, double2 &gk) {
    if constexpr (KIND >= INTI_FF_SL) {
        const double k = p.k;
        const double xy = dotv<DIM>(x, y);
        double s, c;
        sincos_rev64_vop3(-p.kappa * xy, s, c);
        const double2 E = c2(c, s);  // e^{-ik xh.y}
        const double2 PE = cm(p.P, E);
        const double2 mixy = c2(0.0, -xy);  // -i xh.y
        const double2 Ks = PE, gkS = ca(cm(p.Pk, E), cm(PE, mixy));
        double2 Kd = c2(0, 0), gkD = c2(0, 0);
        if constexpr (KIND != INTI_FF_SL) {
            const double xn = dotv<DIM>(x, ny);
            const double2 B = c2(0.0, -k * xn);  // -ik xh.n
            Kd = cm(PE, B);
            gkD = ca(ca(cm(cm(p.Pk, E), B), cm(PE, c2(0.0, -xn))), cm(Kd, mixy));
        }
        if constexpr (KIND == INTI_FF_SL) {
            K = Ks; gk = gkS; b_r = c2(0, 0);
        } else if constexpr (KIND == INTI_FF_DL) {
            K = Kd; gk = gkD; b_r = cm(PE, c2(0.0, -k));
        } else {
            K = ca(cm(p.c0, Ks), cm(p.c1, Kd));
            gk = ca(ca(cm(p.dc0, Ks), cm(p.c0, gkS)), ca(cm(p.dc1, Kd), cm(p.c1, gkD)));
            b_r = cm(p.c1, cm(PE, c2(0.0, -k)));
        }
        a_r = cm(c2(0.0, -k), K);  // dK/dy = -ik xh K
        a_n = c2(0, 0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = x[d];
    } else {
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = x[d] - y[d];
        const AdjRadial R = adj_radial<PDE, DIM, KIND>(dotv<DIM>(v, v), p);
        adj_combine<PDE, DIM, KIND>(v, ny, R, p, K, gk, a_r, a_n, b_r);
    }
}

constexpr int kAdjBS = 256;

struct AdjArgs {
    int64_t m, n;              // targets, sources
    const double *tx, *sx, *sn, *sw;
    const double2 *c;          // conj(ubar), m
    const double2 *sig;        // sigma, n (complex; Laplace has zero imaginary parts)
    double2 *part;             // partial sums
    int split;
    int64_t chunk;
    AdjParams p;
};

// Source-major pass: NV = 3 + 2 DIM complex sums per source,
// [sum c K, sum c a_r v (DIM), sum c b_r v (DIM), sum c a_n, sum c gk]
template <int PDE, int DIM, int KIND>
__global__ __launch_bounds__(kAdjBS) void k_adj_src(AdjArgs a) {
    constexpr int NV = 3 + 2 * DIM;
    __shared__ double s_x[DIM][kAdjBS];
    __shared__ double2 s_c[kAdjBS];
    const int tid = threadIdx.x;
    const int64_t chunk = blockIdx.x % a.split, tile = blockIdx.x / a.split;
    const int64_t j = tile * kAdjBS + tid;
    const int64_t jc = j < a.n ? j : a.n - 1;
    double y[DIM], ny[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        y[d] = a.sx[d * a.n + jc];
        ny[d] = a.sn ? a.sn[d * a.n + jc] : 0.0;
    }
    double2 acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = c2(0, 0);
    const int64_t i0 = chunk * a.chunk;
    const int64_t i1 = i0 + a.chunk < a.m ? i0 + a.chunk : a.m;
    for (int64_t t0 = i0; t0 < i1; t0 += kAdjBS) {
        const int64_t i = t0 + tid;
        const bool live = i < i1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) s_x[d][tid] = live ? a.tx[d * a.m + i] : 0.0;  // unused: the loop stops at cnt
        s_c[tid] = live ? a.c[i] : c2(0, 0);
        __syncthreads();
        const int cnt = (int)(i1 - t0 < kAdjBS ? i1 - t0 : kAdjBS);
        for (int q = 0; q < cnt; ++q) {
            double x[DIM], v[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[d] = s_x[d][q];
            double2 K, a_r, a_n, b_r, gk;
            pair_grad<PDE, DIM, KIND>(x, y, ny, a.p, v, K, a_r, a_n, b_r, gk);
            const double2 cc = s_c[q];
            cfma(acc[0], cc, K);
            const double2 car = cm(cc, a_r), cbr = cm(cc, b_r);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                acc[1 + d].x = __builtin_fma(car.x, v[d], acc[1 + d].x);
                acc[1 + d].y = __builtin_fma(car.y, v[d], acc[1 + d].y);
                acc[1 + DIM + d].x = __builtin_fma(cbr.x, v[d], acc[1 + DIM + d].x);
                acc[1 + DIM + d].y = __builtin_fma(cbr.y, v[d], acc[1 + DIM + d].y);
            }
            cfma(acc[1 + 2 * DIM], cc, a_n);
            cfma(acc[2 + 2 * DIM], cc, gk);
        }
        __syncthreads();
    }
    if (j < a.n) {
        double2 *out = a.part + (chunk * a.n + j) * NV;
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = acc[q];
    }
}

// Target-major pass: DIM complex sums per target, sum_j s_j dK/dx_i
// (= -sum_j s_j (a_r r + a_n n_j)).
template <int PDE, int DIM, int KIND>
__global__ __launch_bounds__(kAdjBS) void k_adj_trg(AdjArgs a) {
    __shared__ double s_y[DIM][kAdjBS];
    __shared__ double s_n[DIM][kAdjBS];
    __shared__ double2 s_s[kAdjBS];
    const int tid = threadIdx.x;
    const int64_t chunk = blockIdx.x % a.split, tile = blockIdx.x / a.split;
    const int64_t i = tile * kAdjBS + tid;
    const int64_t ic = i < a.m ? i : a.m - 1;
    double x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = a.tx[d * a.m + ic];
    double2 acc[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) acc[d] = c2(0, 0);
    const int64_t j0 = chunk * a.chunk;
    const int64_t j1 = j0 + a.chunk < a.n ? j0 + a.chunk : a.n;
    for (int64_t t0 = j0; t0 < j1; t0 += kAdjBS) {
        const int64_t j = t0 + tid;
        const bool live = j < j1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            s_y[d][tid] = live ? a.sx[d * a.n + j] : 0.0;  // unused: the loop stops at cnt
            s_n[d][tid] = live && a.sn ? a.sn[d * a.n + j] : 0.0;
        }
        s_s[tid] = live ? cs(a.sig[j], a.sw[j]) : c2(0, 0);
        __syncthreads();
        const int cnt = (int)(j1 - t0 < kAdjBS ? j1 - t0 : kAdjBS);
        for (int q = 0; q < cnt; ++q) {
            double y[DIM], ny[DIM], v[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                y[d] = s_y[d][q];
                ny[d] = s_n[d][q];
            }
            double2 K, a_r, a_n, b_r, gk;
            pair_grad<PDE, DIM, KIND>(x, y, ny, a.p, v, K, a_r, a_n, b_r, gk);
            const double2 ss = s_s[q];
            const double2 sar = cm(ss, a_r), san = cm(ss, a_n);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {  // dK/dx = -dK/dy
                acc[d].x -= __builtin_fma(sar.x, v[d], san.x * ny[d]);
                acc[d].y -= __builtin_fma(sar.y, v[d], san.y * ny[d]);
            }
        }
        __syncthreads();
    }
    if (i < a.m) {
        double2 *out = a.part + (chunk * a.m + i) * DIM;
#pragma unroll
        for (int d = 0; d < DIM; ++d) out[d] = acc[d];
    }
}

struct AdjOut {
    int64_t m, n;
    int split, dim, cplx;
    const double2 *part, *c, *sig;
    const double *partr;  // symmetric pass: S x NR x n real sums
    const double *sw, *sn;
    double *sigma_bar, *ybar, *nbar, *wbar, *kpart, *xbar;
};

// sum the source-side partials over chunks and form the source gradients;
// per-block partial sums of kbar
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_src_reduce(AdjOut o) {
    constexpr int NV = 3 + 2 * DIM;
    __shared__ double s_k[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double kb = 0.0;
    if (j < o.n) {
        double2 S[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) S[q] = c2(0, 0);
        for (int c = 0; c < o.split; ++c) {
            const double2 *p = o.part + ((int64_t)c * o.n + j) * NV;
#pragma unroll
            for (int q = 0; q < NV; ++q) S[q] = ca(S[q], p[q]);
        }
        const double w = o.sw[j];
        const double2 sg = o.sig[j];
        const double2 s = cs(sg, w);
        if (o.sigma_bar) {
            if (o.cplx) {
                o.sigma_bar[2 * j] = w * S[0].x;
                o.sigma_bar[2 * j + 1] = -w * S[0].y;
            } else {
                o.sigma_bar[j] = w * S[0].x;
            }
        }
        if (o.wbar) o.wbar[j] = sg.x * S[0].x - sg.y * S[0].y;
        const double2 AN = S[1 + 2 * DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double nd = o.sn ? o.sn[d * o.n + j] : 0.0;
            const double2 gy = make_double2(S[1 + d].x + AN.x * nd, S[1 + d].y + AN.y * nd);
            if (o.ybar) o.ybar[d * o.n + j] = s.x * gy.x - s.y * gy.y;
            if (o.nbar) o.nbar[d * o.n + j] = s.x * S[1 + DIM + d].x - s.y * S[1 + DIM + d].y;
        }
        kb = s.x * S[NV - 1].x - s.y * S[NV - 1].y;
    }
    s_k[threadIdx.x] = kb;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s_k[threadIdx.x] += s_k[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0 && o.kpart) o.kpart[blockIdx.x] = s_k[0];
}

template <int DIM>
__global__ __launch_bounds__(256) void k_adj_trg_reduce(AdjOut o) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= o.m) return;
    double2 S[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) S[d] = c2(0, 0);
    for (int c = 0; c < o.split; ++c) {
        const double2 *p = o.part + ((int64_t)c * o.m + i) * DIM;
#pragma unroll
        for (int d = 0; d < DIM; ++d) S[d] = ca(S[d], p[d]);
    }
    const double2 cc = o.c[i];
#pragma unroll
    for (int d = 0; d < DIM; ++d) o.xbar[d * o.m + i] = cc.x * S[d].x - cc.y * S[d].y;
}

// ------------------------------------------------ symmetric pullback
// For an operator whose targets are its sources (one quadrature), both
// passes above in one: each unordered pair {i, j} is visited once, its
// radial functions computed once, and the two ordered pairs (i target, j
// source) and (j target, i source) are differentiated from them. Work and
// lane rotation as in the symmetric forward sum (tiles of 256, block (a, b),
// a <= b; lane l pairs its node with source (l + s) mod 64 of a sub-tile and
// passes the source's running sums to lane l - 1). Per node, NR = 3 DIM + 2
// real sums:
//   [S0 (2): sum_t c_t K_t,node    Y (DIM): ybar    X (DIM): xbar    N (DIM): nbar]
// where, for a pair (target t, source s) with v = x_t - x_s and
// sc = s_s c_t: alpha = Re(sc a_r), beta = Re(sc a_n), gamma = Re(sc b_r),
// g = alpha v + beta n_s; then Y_s += g, X_t -= g, N_s += gamma v, and
// kbar += Re(sc gk). The waves' sums for a sub-tile are added in LDS in wave
// order; one partial slot per (node, tile) as in the forward sum;
// deterministic.
struct AdjSymArgs {
    int64_t n;
    const double *xy, *nrm, *w;
    const double2 *c, *sig;
    double *part;   // S x NR x n
    double *kpart;  // one per block
    int64_t ntiles;
    int distinct;   // no two nodes within twice the coincidence tolerance
    int comb;       // one node gradient, source and target parts summed
    int chik;       // 3D combined field with c0 = -i k, c1 = 1, dc0 = -i, dc1 = 0
    AdjParams p;
};

template <int DIM>
__device__ __forceinline__ void adj_sym_node(const AdjSymArgs &A, int64_t j, double *x, double *nv,
                                             double2 &c, double2 &sw) {
    if (j < A.n) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            x[d] = A.xy[d * A.n + j];
            nv[d] = A.nrm ? A.nrm[d * A.n + j] : 0.0;
        }
        c = A.c[j];
        sw = cs(A.sig[j], A.w[j]);
    } else {
#pragma unroll
        for (int d = 0; d < DIM; ++d) { x[d] = kPadCoord; nv[d] = 0.0; }
        c = sw = c2(0, 0);
    }
}

// one ordered pair (target t, source s): the source's term of S0 (c_t K),
// the factors of the node-gradient term g = al v + be n_s (Y_s += g,
// X_t -= g) and ga (N_s += ga v); returns Re(sc gk)
template <int PDE, int DIM, int KIND, bool MASK>
__device__ __forceinline__ double adj_sym_terms(const double *v, const double *ns, const AdjRadial &R,
                                                const AdjParams &p, double2 ct, double2 ss, double2 &cK,
                                                double &al, double &be, double &ga) {
    double2 K, gk, a_r, a_n, b_r;
    adj_combine<PDE, DIM, KIND, MASK>(v, ns, R, p, K, gk, a_r, a_n, b_r);
    const double2 sc = cm(ss, ct);
    cK = make_double2(__builtin_fma(ct.x, K.x, -ct.y * K.y), __builtin_fma(ct.x, K.y, ct.y * K.x));
    al = sc.x * a_r.x - sc.y * a_r.y;
    be = sc.x * a_n.x - sc.y * a_n.y;
    ga = sc.x * b_r.x - sc.y * b_r.y;
    return sc.x * gk.x - sc.y * gk.y;
}

// LDS add of one double (ds_add_f64): each lane of a wave adds to its own
// address, in program order
__device__ __forceinline__ void lds_add(double *a, double v) {
    __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// COMB: the node gradient is the sum of the source and target parts (Y + X,
// asked for by passing the same output for both): NR = 2 DIM + 2 sums
// [S0 (2), G (DIM), N (DIM)] instead of 3 DIM + 2, and each pair adds its
// difference g1 - g2 once to each node.
template <int PDE, int DIM, int KIND, bool COMB, bool CHIK = false>
__global__ __launch_bounds__(kSymBS) __attribute__((amdgpu_waves_per_eu(COMB ? 3 : 1))) void k_adj_sym(AdjSymArgs A) {
    constexpr int NR = COMB ? 2 * DIM + 2 : 3 * DIM + 2;
    constexpr double OSC = CHIK ? kInv4Pi : 1.0;  // see adj_radial (CHIK)
    constexpr int NW = kSymBS / 64;
    constexpr int iY = 2, iX = COMB ? 2 : 2 + DIM, iN = COMB ? 2 + DIM : 2 + 2 * DIM;
    __shared__ double s_x[DIM][kSymBS], s_n[DIM][kSymBS];
    __shared__ double2 s_c[kSymBS], s_s[kSymBS];
    // per-wave sums of the current 64-source sub-tile: at step s lane l adds
    // to source (l + s) mod 64, so the 64 lanes of a wave add to 64 different
    // addresses and each address receives its terms in step order
    __shared__ double s_acc[NW][NR][64];
    __shared__ double s_k[kSymBS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t a, b;
    sym_item(blockIdx.x, A.ntiles, a, b);
    const bool diag = a == b;
    const int64_t n = A.n;
    const AdjParams &p = A.p;
    double xi[DIM], ni[DIM];
    double2 ci, si;
    adj_sym_node<DIM>(A, a * kSymBS + tid, xi, ni, ci, si);
    {
        double xj[DIM], nj[DIM];
        double2 cj, sj;
        adj_sym_node<DIM>(A, b * kSymBS + tid, xj, nj, cj, sj);
#pragma unroll
        for (int d = 0; d < DIM; ++d) { s_x[d][tid] = xj[d]; s_n[d][tid] = nj[d]; }
        s_c[tid] = cj;
        s_s[tid] = sj;
    }
    double own[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) own[r] = 0.0;
    double kb = 0.0;
    // unrolled: the sub-tile offsets u * 64 become immediate LDS offsets
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        __syncthreads();  // the previous sub-tile's sums have been read; s_x.. are written
        double *acc = &s_acc[wave][0][0];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r * 64 + lane] = 0.0;
        // off-diagonal tiles of distinct nodes hold no coincident pairs
        auto pairs = [&](auto mask_tag) {
            constexpr bool MASK = decltype(mask_tag)::value;
            // source js = (lane + s) mod 64 of the sub-tile, kept as the byte
            // offset o8 = 8 js (two integer operations per step)
            int o8 = lane * 8;
            for (int s = 0; s < 64; ++s) {
                const int js8 = o8;
                o8 = (o8 + 8) & 511;
                double xj[DIM], nj[DIM], v[DIM], vn[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    xj[d] = *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&s_x[d][u * 64]) + js8);
                    nj[d] = *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&s_n[d][u * 64]) + js8);
                    v[d] = xi[d] - xj[d];
                    vn[d] = -v[d];
                }
                const double2 cj =
                    *reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(&s_c[u * 64]) + 2 * js8);
                const double2 sj =
                    *reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(&s_s[u * 64]) + 2 * js8);
                const AdjRadial R = adj_radial<PDE, DIM, KIND, MASK, CHIK>(dotv<DIM>(v, v), p);
                // i target, j source: g1 = al1 v + be1 nj
                double2 cK1;
                double al1, be1, ga1;
                kb += adj_sym_terms<PDE, DIM, KIND, MASK>(v, nj, R, p, ci, sj, cK1, al1, be1, ga1);
                double *aj = reinterpret_cast<double *>(reinterpret_cast<char *>(acc) + js8);
                lds_add(aj + 0 * 64, cK1.x);
                lds_add(aj + 1 * 64, cK1.y);
#pragma unroll
                for (int d = 0; d < DIM; ++d) lds_add(aj + (iN + d) * 64, ga1 * v[d]);
                if (diag) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        const double g1 = __builtin_fma(al1, v[d], be1 * nj[d]);
                        lds_add(aj + (iY + d) * 64, g1);
                        own[iX + d] -= g1;
                    }
                } else {
                    // j target, i source: g2 = al2 vn + be2 ni = -al2 v + be2 ni
                    double2 cK2;
                    double al2, be2, ga2;
                    kb += adj_sym_terms<PDE, DIM, KIND, MASK>(vn, ni, R, p, cj, si, cK2, al2, be2, ga2);
                    own[0] += cK2.x;
                    own[1] += cK2.y;
                    if constexpr (COMB) {
                        // g1 - g2 = (al1 + al2) v + be1 nj - be2 ni
                        const double a12 = al1 + al2;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            own[iN + d] = __builtin_fma(-ga2, v[d], own[iN + d]);
                            const double df = __builtin_fma(a12, v[d], __builtin_fma(be1, nj[d], -be2 * ni[d]));
                            lds_add(aj + (iY + d) * 64, df);
                            own[iY + d] -= df;
                        }
                    } else {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            own[iN + d] = __builtin_fma(ga2, vn[d], own[iN + d]);
                            const double g1 = __builtin_fma(al1, v[d], be1 * nj[d]);
                            const double g2 = __builtin_fma(al2, vn[d], be2 * ni[d]);
                            lds_add(aj + (iY + d) * 64, g1);
                            lds_add(aj + (iX + d) * 64, -g2);
                            own[iY + d] += g2;
                            own[iX + d] -= g1;
                        }
                    }
                }
            }
        };
        if (A.distinct && !diag) pairs(std::false_type{});
        else pairs(std::true_type{});
        __syncthreads();
        // wave u adds the NW waves' sums of sub-tile u in wave order; its
        // lane l owns tile row u * 64 + l (= tid), also the row of its own
        // sums below
        if (wave == u) {
            const int64_t j = b * kSymBS + tid;
            if (j < n) {
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    double v = s_acc[0][r][lane];
#pragma unroll
                    for (int w = 1; w < NW; ++w) v += s_acc[w][r][lane];
                    A.part[(a * NR + r) * n + j] = OSC * v;
                }
            }
        }
    }
    // source tile b's sums went to slot a; target tile a's go to slot b (in
    // a diagonal block both are the same rows, and this thread wrote its
    // row's source sums above)
    const int64_t i = a * kSymBS + tid;
    if (i < n) {
        const int64_t slot = diag ? a : b;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double v = OSC * own[r];
            if (diag) v += A.part[(slot * NR + r) * n + i];
            A.part[(slot * NR + r) * n + i] = v;
        }
    }
    s_k[tid] = kb;
    __syncthreads();
    for (int off = kSymBS / 2; off > 0; off >>= 1) {
        if (tid < off) s_k[tid] += s_k[tid + off];
        __syncthreads();
    }
    if (tid == 0) A.kpart[blockIdx.x] = OSC * s_k[0];
}

// per node: sum the S slots and form the outputs
template <int DIM, bool COMB>
__global__ __launch_bounds__(256) void k_adj_sym_reduce(AdjOut o, int64_t S) {
    constexpr int NR = COMB ? 2 * DIM + 2 : 3 * DIM + 2;
    constexpr int iX = COMB ? 2 : 2 + DIM, iN = COMB ? 2 + DIM : 2 + 2 * DIM;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= o.n) return;
    double R[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) R[r] = 0.0;
    for (int64_t c = 0; c < S; ++c) {
#pragma unroll
        for (int r = 0; r < NR; ++r) R[r] += o.partr[(c * NR + r) * o.n + j];
    }
    const double w = o.sw[j];
    const double2 sg = o.sig[j];
    if (o.sigma_bar) {
        if (o.cplx) {
            o.sigma_bar[2 * j] = w * R[0];
            o.sigma_bar[2 * j + 1] = -w * R[1];
        } else {
            o.sigma_bar[j] = w * R[0];
        }
    }
    if (o.wbar) o.wbar[j] = sg.x * R[0] - sg.y * R[1];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (o.ybar) o.ybar[d * o.n + j] = R[2 + d];
        if (!COMB && o.xbar) o.xbar[d * o.n + j] = R[iX + d];
        if (o.nbar) o.nbar[d * o.n + j] = R[iN + d];
    }
}

// sigma_bar += corr^H ubar: one thread per column j gathers
// conj(val_ij) ubar_i over the column's rows (the transposed pattern, rows
// increasing) and adds the sum to sigma_bar_j: one writer per entry, a
// fixed order, deterministic. (The identity term is added on the host.)
__global__ __launch_bounds__(256) void k_adj_corr(int64_t n, const int32_t *tptr, const int32_t *trow,
                                                  const int32_t *tperm, const double *val, int cplx,
                                                  const double2 *c, double *sigma_bar) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double sr = 0.0, si = 0.0;
    for (int32_t p = tptr[j]; p < tptr[j + 1]; ++p) {
        const double2 cc = c[trow[p]];  // conj(ubar_i)
        const int32_t q = tperm[p];
        if (cplx) {
            // conj(v) ubar = conj(v c)
            const double vr = val[2 * q], vi = val[2 * q + 1];
            sr += vr * cc.x - vi * cc.y;
            si -= vr * cc.y + vi * cc.x;
        } else {
            sr += val[q] * cc.x;
        }
    }
    if (cplx) {
        sigma_bar[2 * j] += sr;
        sigma_bar[2 * j + 1] += si;
    } else {
        sigma_bar[j] += sr;
    }
}

using SrcFn = void (*)(const AdjArgs &, dim3, hipStream_t);

template <int PDE, int DIM, int KIND>
void src_launch(const AdjArgs &a, dim3 g, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_src<PDE, DIM, KIND>), g, dim3(kAdjBS), 0, s, a);
}
template <int PDE, int DIM, int KIND>
void trg_launch(const AdjArgs &a, dim3 g, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_trg<PDE, DIM, KIND>), g, dim3(kAdjBS), 0, s, a);
}

template <int PDE, int DIM>
void pick(int kind, SrcFn &fs, SrcFn &ft) {
    fs = ft = nullptr;
    switch (kind) {
        case INTI_SL: fs = src_launch<PDE, DIM, INTI_SL>; ft = trg_launch<PDE, DIM, INTI_SL>; return;
        case INTI_DL: fs = src_launch<PDE, DIM, INTI_DL>; ft = trg_launch<PDE, DIM, INTI_DL>; return;
        case INTI_COMB_SD:
            fs = src_launch<PDE, DIM, INTI_COMB_SD>; ft = trg_launch<PDE, DIM, INTI_COMB_SD>; return;
    }
    if constexpr (PDE == INTI_HELMHOLTZ) {
        switch (kind) {
            case INTI_FF_SL: fs = src_launch<PDE, DIM, INTI_FF_SL>; return;
            case INTI_FF_DL: fs = src_launch<PDE, DIM, INTI_FF_DL>; return;
            case INTI_FF_COMB: fs = src_launch<PDE, DIM, INTI_FF_COMB>; return;
        }
    }
}

using SymAdjFn = void (*)(const AdjSymArgs &, int64_t, hipStream_t);

template <int PDE, int DIM, int KIND>
void adj_sym_launch(const AdjSymArgs &a, int64_t nblocks, hipStream_t s) {
    if constexpr (PDE == INTI_HELMHOLTZ && DIM == 3 && KIND == INTI_COMB_SD) {
        if (a.chik) {
            if (a.comb)
                hipLaunchKernelGGL((k_adj_sym<PDE, DIM, KIND, true, true>), dim3((unsigned)nblocks), dim3(kSymBS), 0, s, a);
            else
                hipLaunchKernelGGL((k_adj_sym<PDE, DIM, KIND, false, true>), dim3((unsigned)nblocks), dim3(kSymBS), 0, s, a);
            return;
        }
    }
    if (a.comb) hipLaunchKernelGGL((k_adj_sym<PDE, DIM, KIND, true>), dim3((unsigned)nblocks), dim3(kSymBS), 0, s, a);
    else hipLaunchKernelGGL((k_adj_sym<PDE, DIM, KIND, false>), dim3((unsigned)nblocks), dim3(kSymBS), 0, s, a);
}

template <int PDE, int DIM>
SymAdjFn pick_sym(int kind) {
    switch (kind) {
        case INTI_SL: return adj_sym_launch<PDE, DIM, INTI_SL>;
        case INTI_DL: return adj_sym_launch<PDE, DIM, INTI_DL>;
        case INTI_COMB_SD: return adj_sym_launch<PDE, DIM, INTI_COMB_SD>;
    }
    return nullptr;
}

// partial slots of the symmetric pullback are capped: beyond this the two
// passes run
constexpr size_t kAdjSymMaxBytes = size_t(4) << 30;

struct Dev {
    void *p = nullptr;
    hipStream_t s = nullptr;
    ~Dev() {
        if (p) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    }
    bool alloc(hipStream_t st, size_t bytes) {
        s = st;
        return hipMalloc(&p, bytes ? bytes : 8) == hipSuccess;
    }
};

// chunks of the inner index so that the grid has several blocks per CU
void plan(int64_t own, int64_t inner, int num_cu, int &split, int64_t &chunk) {
    const int64_t tiles = (own + kAdjBS - 1) / kAdjBS;
    const int64_t itiles = (inner + kAdjBS - 1) / kAdjBS;
    int64_t s = (8LL * num_cu + tiles - 1) / tiles;
    if (s > itiles) s = itiles;
    if (s < 1) s = 1;
    chunk = (itiles + s - 1) / s * kAdjBS;
    split = (int)((inner + chunk - 1) / chunk);
    if (split < 1) split = 1;
}

}  // namespace

// sigma_bar_j = conj(w_j post z_j) (complex) or w_j post z_j (real), z the
// transposed sum's S partial slots added in slot order
__global__ __launch_bounds__(256) void k_adj_trans_reduce(int64_t n, int64_t S, int cplx, const double *part,
                                                          double pr, double pi, const double *w, double *sb) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    if (cplx) {
        const double2 *pp = reinterpret_cast<const double2 *>(part);
        double zr = 0.0, zi = 0.0;
        for (int64_t q = 0; q < S; ++q) {
            const double2 v = pp[q * n + j];
            zr += v.x;
            zi += v.y;
        }
        const double wj = w[j];
        sb[2 * j] = wj * (pr * zr - pi * zi);
        sb[2 * j + 1] = -wj * (pr * zi + pi * zr);
    } else {
        double z = 0.0;
        for (int64_t q = 0; q < S; ++q) z += part[q * n + j];
        sb[j] = w[j] * pr * z;
    }
}

// y += cid sigma  =>  sigma_bar += conj(cid) ubar (host arrays)
static void adjoint_identity(const Op *op, int cplx, int64_t n, const double *ubar, double *sigma_bar) {
    if (op->cid[0] == 0.0 && op->cid[1] == 0.0) return;
    for (int64_t j = 0; j < n; ++j) {
        if (cplx) {
            const double ur = ubar[2 * j], ui = ubar[2 * j + 1];
            sigma_bar[2 * j] += op->cid[0] * ur + op->cid[1] * ui;
            sigma_bar[2 * j + 1] += op->cid[0] * ui - op->cid[1] * ur;
        } else {
            sigma_bar[j] += op->cid[0] * ubar[j];
        }
    }
}

// sigma_bar alone (the product A^H ubar of adjoint solves and density
// gradients): A_ij = post K(x_i, y_j) w_j, so (A^H ubar)_j =
// conj(w_j post sum_i K(x_i, y_j) c_i) with c = conj(ubar), a sum of kernel
// values only (no derivatives). An operator on one quadrature takes the
// transposed symmetric sum (every unordered pair once), any other operator
// the transposed one-direction sum (one thread per source, the targets
// staged in LDS). Returns -1 where neither applies (the general pullback then
// runs), else a status.
static int adjoint_sigma_only(Op *op, const double *ubar, double *sigma_bar) {
    Ctx *ctx = op->ctx;
    const int64_t m = op->trg->n, n = op->src->n;
    const int cplx = op_is_complex(op) ? 1 : 0;
    const size_t cw = cplx ? 2 : 1;
    int kind = op->kind;
    const KParams kp = op_params(op, &kind);

    // symmetric plan: tiles of 256 in groups of G
    SymFn fsym = nullptr;
    int64_t T = (n + kSymBS - 1) / kSymBS, S = 0;
    if (op->trg == op->src && op->sym && kind < INTI_FF_SL && T >= 2 &&
        (op->sym == 2 || T * (T + 1) / 2 >= 2LL * ctx->num_cu)) {
        int G = 0;
        for (int g : {1, 2, 4}) {
            if (op->sym_group && g != op->sym_group) continue;
            S = (T + g - 1) / g;
            if ((size_t)S * (size_t)n * cw * sizeof(double) <= kAdjSymMaxBytes) {
                G = g;
                break;
            }
        }
        if (G && op->pde == INTI_LAPLACE && op->dim == 2) fsym = sym_trans_pick_lap2(kind, G);
        if (G && op->pde == INTI_LAPLACE && op->dim == 3) fsym = sym_trans_pick_lap3(kind, G);
        if (G && op->pde == INTI_HELMHOLTZ && op->dim == 2) fsym = sym_trans_pick_helm2(kind, G);
        if (G && op->pde == INTI_HELMHOLTZ && op->dim == 3) fsym = sym_trans_pick_helm3(kind, G);
    }
    // one-direction plan: the sum's targets are the n sources (tiles of
    // 256), its sources the m targets cut into `split` chunks
    ApplyFn fone = nullptr;
    int split = 0;
    int64_t chunk = 0;
    const int64_t ntiles = (n + kApplyBS - 1) / kApplyBS;
    if (!fsym) {
        if (op->pde == INTI_LAPLACE && op->dim == 2) fone = apply_trans_pick_lap2(kind);
        if (op->pde == INTI_LAPLACE && op->dim == 3) fone = apply_trans_pick_lap3(kind);
        if (op->pde == INTI_HELMHOLTZ && op->dim == 2) fone = apply_trans_pick_helm2(kind);
        if (op->pde == INTI_HELMHOLTZ && op->dim == 3) fone = apply_trans_pick_helm3(kind);
        if (!fone) return -1;
        const int64_t stiles = (m + kApplyBS - 1) / kApplyBS;
        int64_t sp = (8LL * ctx->num_cu + ntiles - 1) / ntiles;
        if (sp >= 8) sp = (sp + 7) / 8 * 8;
        if (sp > stiles) sp = stiles;
        if (sp < 1) sp = 1;
        chunk = (stiles + sp - 1) / sp * kApplyBS;
        split = (int)((m + chunk - 1) / chunk);
        S = split;
    }

    hipStream_t s = ctx->stream;
    std::vector<double2> hc(m);  // conj(ubar), as the correction kernel reads it
    for (int64_t i = 0; i < m; ++i) hc[i] = cplx ? make_double2(ubar[2 * i], -ubar[2 * i + 1]) : make_double2(ubar[i], 0.0);
    std::vector<double> hx;
    if (!cplx) hx.assign(ubar, ubar + m);
    Dev dc, dx, dpart, dsb;
    if (!dc.alloc(s, sizeof(double2) * m) || (!cplx && !dx.alloc(s, sizeof(double) * m)) ||
        !dpart.alloc(s, sizeof(double) * cw * (size_t)S * (size_t)n) || !dsb.alloc(s, sizeof(double) * cw * n))
        return set_error(ctx, INTI_ERR_NOMEM, "adjoint: allocation failed");
    INTI_HIP(ctx, hipMemcpyAsync(dc.p, hc.data(), sizeof(double2) * m, hipMemcpyHostToDevice, s));
    if (!cplx) INTI_HIP(ctx, hipMemcpyAsync(dx.p, hx.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    const double *xc = cplx ? (const double *)dc.p : (const double *)dx.p;  // c, unweighted
    if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
    if (fsym) {
        SymArgs a;
        a.n = n;
        a.xy = op->src->coords; a.nrm = op->src->normals; a.w = nullptr;
        a.x = xc;
        a.part = (double *)dpart.p;
        a.ntiles = T; a.ngroups = S;
        a.distinct = op->src->distinct ? 1 : 0;
        a.ch_ik = 0;  // the transposed sums are FP64 (no staged ch q)
        a.p = kp;
        fsym(a, S * (S + 1) / 2, s);
    } else {
        ApplyArgs a;
        a.m = n; a.n = m;
        a.tstride = n; a.sstride = m;
        a.tx = op->src->coords; a.tn = op->src->normals;
        a.sx = op->trg->coords; a.sn = op->trg->normals; a.sw = nullptr;
        a.x = xc; a.ldx = m;
        a.part = (double *)dpart.p; a.split = split; a.chunk = chunk;
        a.p = kp;
        fone(&a, dim3((unsigned)(ntiles * split)), s);
    }
    INTI_HIP(ctx, hipGetLastError());
    if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
    hipLaunchKernelGGL(k_adj_trans_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, S, cplx,
                       (const double *)dpart.p, kp.postr, kp.posti, op->src->weights, (double *)dsb.p);
    INTI_HIP(ctx, hipGetLastError());
    if (op->corr) {
        hipLaunchKernelGGL(k_adj_corr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, op->corr_tptr,
                           op->corr_trow, op->corr_tperm, op->corr->val, cplx, (const double2 *)dc.p,
                           (double *)dsb.p);
        INTI_HIP(ctx, hipGetLastError());
    }
    INTI_HIP(ctx, hipMemcpyAsync(sigma_bar, dsb.p, sizeof(double) * cw * n, hipMemcpyDeviceToHost, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    if (m == n) adjoint_identity(op, cplx, n, ubar, sigma_bar);
    return INTI_OK;
}

int op_adjoint(Op *op, const double *ubar, const double *sigma, double *sigma_bar, double *ybar,
               double *nbar, double *wbar, double *xbar, const double *dcoefs, double *kbar) {
    Ctx *ctx = op->ctx;
    const int kind = op->kind;
    const bool ff = kind >= INTI_FF_SL;
    if (!(kind == INTI_SL || kind == INTI_DL || kind == INTI_COMB_SD || ff))
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "adjoint: supported kinds are SL, DL, c0 SL + c1 DL and the far fields");
    if (op->precision != INTI_FP64)
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "adjoint: FP64 operators only");
    if (xbar && ff) return set_error(ctx, INTI_ERR_UNSUPPORTED, "adjoint: far-field directions carry no gradient");
    const bool geo = ybar || nbar || wbar || xbar || kbar;
    if (op->corr && geo)
        return set_error(ctx, INTI_ERR_UNSUPPORTED,
                         "adjoint: geometry/k gradients of a corrected operator are not available (the correction weights "
                         "depend on the geometry); detach the correction or request sigma_bar only");
    const int dim = op->dim;
    const int cplx = op_is_complex(op) ? 1 : 0;
    const int64_t m = op->trg->n, n = op->src->n;
    // the same output for the source and the target node gradients of an
    // operator on one quadrature: their sum, the gradient of the nodes
    const bool comb = ybar && ybar == xbar && op->trg == op->src;
    if (m == 0 || n == 0) {
        if (sigma_bar) std::fill(sigma_bar, sigma_bar + (cplx ? 2 : 1) * n, 0.0);
        if (ybar) std::fill(ybar, ybar + dim * n, 0.0);
        if (nbar) std::fill(nbar, nbar + dim * n, 0.0);
        if (wbar) std::fill(wbar, wbar + n, 0.0);
        if (xbar) std::fill(xbar, xbar + dim * m, 0.0);
        if (kbar) *kbar = 0.0;
        return INTI_OK;
    }
    if (sigma_bar && !geo) {
        const int st = adjoint_sigma_only(op, ubar, sigma_bar);
        if (st >= 0) return st;
    }
    SrcFn fs = nullptr, ft = nullptr;
    if (op->pde == INTI_LAPLACE && dim == 2) pick<INTI_LAPLACE, 2>(kind, fs, ft);
    if (op->pde == INTI_LAPLACE && dim == 3) pick<INTI_LAPLACE, 3>(kind, fs, ft);
    if (op->pde == INTI_HELMHOLTZ && dim == 2) pick<INTI_HELMHOLTZ, 2>(kind, fs, ft);
    if (op->pde == INTI_HELMHOLTZ && dim == 3) pick<INTI_HELMHOLTZ, 3>(kind, fs, ft);
    if (!fs) return set_error(ctx, INTI_ERR_UNSUPPORTED, "adjoint: no kernel for this operator");

    AdjParams p;
    p.k = op->k;
    p.kappa = op->k * kInv2Pi;
    p.c0 = make_double2(op->coefs[0], op->coefs[1]);
    p.c1 = make_double2(op->coefs[2], op->coefs[3]);
    p.dc0 = dcoefs ? make_double2(dcoefs[0], dcoefs[1]) : make_double2(0, 0);
    p.dc1 = dcoefs ? make_double2(dcoefs[2], dcoefs[3]) : make_double2(0, 0);
    p.P = p.Pk = make_double2(0, 0);
    if (ff) {
        double pr, pi;
        basic_post(op->pde, dim, INTI_FF_SL, op->k, pr, pi);
        p.P = make_double2(pr, pi);
        if (dim == 2) p.Pk = make_double2(-0.5 * pr / op->k, -0.5 * pi / op->k);
    }

    // host staging: c = conj(ubar), sigma as complex
    std::vector<double2> hc(m), hs(n);
    for (int64_t i = 0; i < m; ++i) hc[i] = cplx ? make_double2(ubar[2 * i], -ubar[2 * i + 1]) : make_double2(ubar[i], 0.0);
    for (int64_t j = 0; j < n; ++j) hs[j] = cplx ? make_double2(sigma[2 * j], sigma[2 * j + 1]) : make_double2(sigma[j], 0.0);

    hipStream_t s = ctx->stream;
    // symmetric pullback: an operator on one quadrature, every output in one
    // pass over the unordered pairs
    const int NRs = comb ? 2 * dim + 2 : 3 * dim + 2;
    const int64_t symT = (n + kSymBS - 1) / kSymBS;
    SymAdjFn fsym = nullptr;
    const bool sym_fills = op->sym == 2 || symT * (symT + 1) / 2 >= 2LL * ctx->num_cu;
    if (op->sym && sym_fills && op->trg == op->src && !ff && symT >= 2 &&
        sizeof(double) * (size_t)symT * NRs * (size_t)n <= kAdjSymMaxBytes) {
        if (op->pde == INTI_LAPLACE && dim == 2) fsym = pick_sym<INTI_LAPLACE, 2>(kind);
        if (op->pde == INTI_LAPLACE && dim == 3) fsym = pick_sym<INTI_LAPLACE, 3>(kind);
        if (op->pde == INTI_HELMHOLTZ && dim == 2) fsym = pick_sym<INTI_HELMHOLTZ, 2>(kind);
        if (op->pde == INTI_HELMHOLTZ && dim == 3) fsym = pick_sym<INTI_HELMHOLTZ, 3>(kind);
    }
    const int64_t symB = symT * (symT + 1) / 2;
    int split_s, split_t = 0;
    int64_t chunk_s, chunk_t = 0;
    plan(n, m, ctx->num_cu, split_s, chunk_s);
    if (xbar) plan(m, n, ctx->num_cu, split_t, chunk_t);
    const int NVs = 3 + 2 * dim;
    const int64_t nsb = fsym ? symB : (n + 255) / 256;  // kbar partial sums
    Dev dc, dsg, dpart, dsb, dyb, dnb, dwb, dkp, dxb;
    const size_t part_bytes = fsym ? sizeof(double) * (size_t)symT * NRs * (size_t)n
                                   : std::max(sizeof(double2) * (size_t)split_s * n * NVs,
                                              xbar ? sizeof(double2) * (size_t)split_t * m * dim : 0);
    if (!dc.alloc(s, sizeof(double2) * m) || !dsg.alloc(s, sizeof(double2) * n) || !dpart.alloc(s, part_bytes) ||
        !dsb.alloc(s, sizeof(double) * (cplx ? 2 : 1) * n) || !dyb.alloc(s, sizeof(double) * dim * n) ||
        !dnb.alloc(s, sizeof(double) * dim * n) || !dwb.alloc(s, sizeof(double) * n) ||
        !dkp.alloc(s, sizeof(double) * nsb) || !dxb.alloc(s, sizeof(double) * dim * m))
        return set_error(ctx, INTI_ERR_NOMEM, "adjoint: allocation failed");
    INTI_HIP(ctx, hipMemcpyAsync(dc.p, hc.data(), sizeof(double2) * m, hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipMemcpyAsync(dsg.p, hs.data(), sizeof(double2) * n, hipMemcpyHostToDevice, s));

    AdjArgs a;
    a.m = m; a.n = n;
    a.tx = op->trg->coords; a.sx = op->src->coords; a.sn = op->src->normals; a.sw = op->src->weights;
    a.c = (const double2 *)dc.p; a.sig = (const double2 *)dsg.p;
    a.part = (double2 *)dpart.p;
    a.p = p;
    AdjOut o;
    o.m = m; o.n = n; o.dim = dim; o.cplx = cplx;
    o.part = (const double2 *)dpart.p; o.c = (const double2 *)dc.p; o.sig = (const double2 *)dsg.p;
    o.sw = op->src->weights;
    o.sn = op->src->normals;
    o.sigma_bar = (double *)dsb.p; o.ybar = (double *)dyb.p; o.nbar = (double *)dnb.p;
    o.wbar = (double *)dwb.p; o.kpart = (double *)dkp.p; o.xbar = (double *)dxb.p;
    o.partr = (const double *)dpart.p;

    if (fsym) {
        AdjSymArgs sa;
        sa.n = n;
        sa.xy = op->src->coords; sa.nrm = op->src->normals; sa.w = op->src->weights;
        sa.c = (const double2 *)dc.p; sa.sig = (const double2 *)dsg.p;
        sa.part = (double *)dpart.p; sa.kpart = (double *)dkp.p;
        sa.ntiles = symT;
        sa.distinct = op->src->distinct ? 1 : 0;
        sa.comb = comb ? 1 : 0;
        sa.chik = op->pde == INTI_HELMHOLTZ && dim == 3 && kind == INTI_COMB_SD && p.c1.x == 1.0 && p.c1.y == 0.0 &&
                          p.c0.x == 0.0 && p.c0.y == -p.k && p.dc0.x == 0.0 && p.dc0.y == -1.0 && p.dc1.x == 0.0 &&
                          p.dc1.y == 0.0
                      ? 1
                      : 0;
        sa.p = p;
        if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
        fsym(sa, symB, s);
        INTI_HIP(ctx, hipGetLastError());
        if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
        const unsigned nb = (unsigned)((n + 255) / 256);
        if (dim == 2 && comb) hipLaunchKernelGGL((k_adj_sym_reduce<2, true>), dim3(nb), dim3(256), 0, s, o, symT);
        else if (dim == 2) hipLaunchKernelGGL((k_adj_sym_reduce<2, false>), dim3(nb), dim3(256), 0, s, o, symT);
        else if (comb) hipLaunchKernelGGL((k_adj_sym_reduce<3, true>), dim3(nb), dim3(256), 0, s, o, symT);
        else hipLaunchKernelGGL((k_adj_sym_reduce<3, false>), dim3(nb), dim3(256), 0, s, o, symT);
        INTI_HIP(ctx, hipGetLastError());
    }
    // source-major pass
    a.split = split_s; a.chunk = chunk_s;
    if (!fsym && ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
    if (!fsym) fs(a, dim3((unsigned)(((n + kAdjBS - 1) / kAdjBS) * split_s)), s);
    INTI_HIP(ctx, hipGetLastError());
    o.split = split_s;
    if (!fsym) {
        if (dim == 2) hipLaunchKernelGGL(k_adj_src_reduce<2>, dim3((unsigned)nsb), dim3(256), 0, s, o);
        else hipLaunchKernelGGL(k_adj_src_reduce<3>, dim3((unsigned)nsb), dim3(256), 0, s, o);
        INTI_HIP(ctx, hipGetLastError());
    }
    // target-major pass (reuses the partial buffer)
    if (xbar && !fsym) {
        a.split = split_t; a.chunk = chunk_t;
        ft(a, dim3((unsigned)(((m + kAdjBS - 1) / kAdjBS) * split_t)), s);
        INTI_HIP(ctx, hipGetLastError());
        o.split = split_t;
        const unsigned mb = (unsigned)((m + 255) / 256);
        if (dim == 2) hipLaunchKernelGGL(k_adj_trg_reduce<2>, dim3(mb), dim3(256), 0, s, o);
        else hipLaunchKernelGGL(k_adj_trg_reduce<3>, dim3(mb), dim3(256), 0, s, o);
        INTI_HIP(ctx, hipGetLastError());
    }
    if (!fsym && ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
    // corrected operator: sigma_bar also gets corr^H ubar and conj(identity) ubar
    if (sigma_bar && op->corr) {
        hipLaunchKernelGGL(k_adj_corr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, op->corr_tptr,
                           op->corr_trow, op->corr_tperm, op->corr->val, cplx, (const double2 *)dc.p,
                           (double *)dsb.p);
        INTI_HIP(ctx, hipGetLastError());
    }
    std::vector<double> kp(nsb);
    if (sigma_bar) INTI_HIP(ctx, hipMemcpyAsync(sigma_bar, dsb.p, sizeof(double) * (cplx ? 2 : 1) * n, hipMemcpyDeviceToHost, s));
    if (ybar) INTI_HIP(ctx, hipMemcpyAsync(ybar, dyb.p, sizeof(double) * dim * n, hipMemcpyDeviceToHost, s));
    if (nbar) INTI_HIP(ctx, hipMemcpyAsync(nbar, dnb.p, sizeof(double) * dim * n, hipMemcpyDeviceToHost, s));
    if (wbar) INTI_HIP(ctx, hipMemcpyAsync(wbar, dwb.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    // combined node gradient: the symmetric pass summed the two parts; the
    // two-pass path adds the target part on the host
    std::vector<double> xsep(comb && !fsym ? (size_t)dim * m : 0);
    if (xbar && !(comb && fsym))
        INTI_HIP(ctx, hipMemcpyAsync(comb ? xsep.data() : xbar, dxb.p, sizeof(double) * dim * m, hipMemcpyDeviceToHost, s));
    if (kbar) INTI_HIP(ctx, hipMemcpyAsync(kp.data(), dkp.p, sizeof(double) * nsb, hipMemcpyDeviceToHost, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    for (size_t q = 0; q < xsep.size(); ++q) ybar[q] += xsep[q];
    if (sigma_bar && m == n) adjoint_identity(op, cplx, n, ubar, sigma_bar);
    if (kbar) {
        double sum = 0.0;
        for (double v : kp) sum += v;
        *kbar = sum;
    }
    return INTI_OK;
}

}  // namespace inti
