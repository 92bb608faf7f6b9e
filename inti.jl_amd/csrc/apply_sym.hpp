// Symmetric direct sum for square operators whose targets are their own
// sources (one quadrature). K(x_i, x_j) and K(x_j, x_i) share the distance,
// its inverse square root and e^{ikd} (or the Hankel values), and differ only
// in which normal enters and in the sign of r; each unordered pair is
// therefore evaluated once and applied in both directions (the shared parts
// are common subexpressions of the two pair evaluations below).
//
// Work decomposition. The n nodes are cut into T tiles of 256 and the tiles
// into S groups of G. Block (a, b), a <= b, takes the target tiles of group a
// against the source tiles of group b (within a diagonal block, the tile
// pairs I <= J only). A wave holds 64 x TPL targets (TPL = 1, 2 or 4 per lane).
// For each 64-source sub-tile, lane l at step s < 32 pairs its targets with
// sources (l + s) mod 64 and (l + s + 32) mod 64, so the 64 lanes touch 64
// different sources per read, without LDS bank conflicts. The transposed
// sums of the two sources travel with the steps in two independent chains:
// after each step lane l takes lane l+1's running sums (ds_bpermute, a
// crossbar move with no LDS storage), one chain's move overlapping the other
// chain's arithmetic; after 32 steps one chain's sum for source l is in lane
// l and the other's in lane l + 32, added with one last move. The waves' sums
// are added in LDS: two waves onto zero by atomic adds (the sum of two terms
// does not depend on their order), more waves in wave order.
//
// Partial sums. A row of group c receives exactly one partial per group:
// from block (c, b), b >= c (direct part) in slot b, and from block (a, c),
// a < c (transposed part) in slot a of part[S][n]. In a diagonal block both
// land in slot a and are added by the thread that owns the row. Every slot is
// written once, in a fixed order, so the result does not depend on the
// schedule; the reduction kernel of the plain direct sum (split = S) sums the
// slots and applies the post factor, the correction and the identity.
#pragma once
#include <type_traits>

#include "apply_kernel.hpp"

namespace inti {

struct SymArgs {
    int64_t n;
    const double *xy, *nrm, *w;  // SoA coordinates / normals (stride n), weights (null: all 1)
    const double *x;             // density (complex interleaved for Helmholtz)
    double *part;                // S x n partial sums (complex interleaved)
    int64_t ntiles, ngroups;     // T, S
    int distinct;                // no two nodes within twice the coincidence tolerance
    int ch_ik;                   // combined field with ch = -i k (ch q = (k Im q, -k Re q))
    KParams p;
};

constexpr int kSymBS = 256;

// (a, b), a <= b, of block L in the row-major upper triangle of S x S.
__device__ __forceinline__ void sym_item(int64_t L, int64_t S, int64_t &a, int64_t &b) {
    auto start = [S](int64_t r) { return r * S - r * (r - 1) / 2; };  // blocks in rows < r
    const double tS = 2.0 * (double)S + 1.0;
    int64_t r = (int64_t)floor(0.5 * (tS - sqrt(fmax(tS * tS - 8.0 * (double)L, 0.0))));
    r = r < 0 ? 0 : r > S - 1 ? S - 1 : r;
    while (r > 0 && start(r) > L) --r;
    while (r + 1 < S && start(r + 1) <= L) ++r;
    a = r;
    b = r + (L - start(r));
}

// node j's geometry and q_j = w_j x_j (q_j = x_j without weights); far-away
// padding with zero weight beyond n
template <int DIM, bool CPLX>
__device__ __forceinline__ void sym_node(const SymArgs &A, int64_t j, double *y, double *nv,
                                         double &qr, double &qi) {
    if (j < A.n) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            y[d] = A.xy[d * A.n + j];
            nv[d] = A.nrm ? A.nrm[d * A.n + j] : 0.0;
        }
        const double w = A.w ? A.w[j] : 1.0;
        if constexpr (CPLX) {
            qr = w * A.x[2 * j];
            qi = w * A.x[2 * j + 1];
        } else {
            qr = w * A.x[j];
            qi = 0.0;
        }
    } else {
#pragma unroll
        for (int d = 0; d < DIM; ++d) { y[d] = kPadCoord; nv[d] = 0.0; }
        qr = qi = 0.0;
    }
}

__device__ __forceinline__ double lane_next(double v, int addr) {
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_next(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// LDS record of a source node: its coordinates, its normal when the kernel
// reads one (every kind but SL), and q (1 or 2 doubles), packed in double2
// slots; each slot is an array over the tile, so lanes reading consecutive
// nodes hit consecutive banks.
template <int PDE, int DIM, int KIND>
struct SymRec {
    static constexpr bool CPLX = PDE == INTI_HELMHOLTZ;
    static constexpr bool NRM = KIND != INTI_SL;
    static constexpr int NF = DIM + (NRM ? DIM : 0) + (CPLX ? 2 : 1);
    static constexpr int NS = (NF + 1) / 2;

    __device__ static __forceinline__ void put(double2 (*s)[kSymBS], int j, const double *y,
                                               const double *nv, double qr, double qi) {
        double f[2 * NS];
        int c = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) f[c++] = y[d];
        if constexpr (NRM) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) f[c++] = nv[d];
        }
        f[c++] = qr;
        if constexpr (CPLX) f[c++] = qi;
        if (c < 2 * NS) f[c] = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) s[k][j] = make_double2(f[2 * k], f[2 * k + 1]);
    }
    __device__ static __forceinline__ void get(const double2 (*s)[kSymBS], int j, double *y,
                                               double *nv, double &qr, double &qi) {
        double f[2 * NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const double2 v = s[k][j];
            f[2 * k] = v.x;
            f[2 * k + 1] = v.y;
        }
        int c = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) y[d] = f[c++];
        if constexpr (NRM) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) nv[d] = f[c++];
        } else {
#pragma unroll
            for (int d = 0; d < DIM; ++d) nv[d] = 0.0;
        }
        qr = f[c++];
        qi = CPLX ? f[c] : 0.0;
    }
};

// LDS record of a source node for the fused mixed 3D Helmholtz DL / DL +
// ch*SL pairs (helm3_sym_fused32), four 16-byte slots (16 KB of LDS per
// tile): the coordinates in FP64, the normal in FP32 (it enters only the
// FP32 amplitudes), and the staged density terms in FP32, (q, k Im q,
// -k Re q) and ch q (CHIK: ch = -i k, so ch q equals (k Im q, -k Re q)).
template <bool CHIK>
struct SymRecF {
    static constexpr int NS = 4;
    // y (FP64), the normal in FP32 (it only enters the FP32 amplitudes) and
    // the FP32 source terms: y0 y1 | y2 n0 n1 | n2 cr ci - | qr qi kqi mkqr
    __device__ static __forceinline__ void put(double2 (*s)[kSymBS], int j, const double *y, const double *nv,
                                               const QExt32 &q) {
        s[0][j] = make_double2(y[0], y[1]);
        s[1][j].x = y[2];
        *reinterpret_cast<float2 *>(&s[1][j].y) = make_float2((float)nv[0], (float)nv[1]);
        *reinterpret_cast<float4 *>(&s[2][j]) =
            make_float4((float)nv[2], CHIK ? 0.0f : q.cr, CHIK ? 0.0f : q.ci, 0.0f);
        *reinterpret_cast<float4 *>(&s[3][j]) = make_float4(q.qr, q.qi, q.kqi, q.mkqr);
    }
    __device__ static __forceinline__ void get(const double2 (*s)[kSymBS], int j, double *y, float *nv,
                                               QExt32 &q) {
        const double2 v0 = s[0][j], v1 = s[1][j];
        const float2 n01 = *reinterpret_cast<const float2 *>(&v1.y);
        const float4 g = *reinterpret_cast<const float4 *>(&s[2][j]);
        const float4 f = *reinterpret_cast<const float4 *>(&s[3][j]);
        y[0] = v0.x; y[1] = v0.y; y[2] = v1.x;
        nv[0] = n01.x; nv[1] = n01.y; nv[2] = g.x;
        q.qr = f.x; q.qi = f.y; q.kqi = f.z; q.mkqr = f.w;
        if constexpr (CHIK) {
            q.cr = f.z;
            q.ci = f.w;
        } else {
            q.cr = g.y;
            q.ci = g.z;
        }
    }
    __device__ static __forceinline__ void get(const double2 (*s)[kSymBS], int j, double *y, double *nv,
                                               QExt32 &q) {
        float n32[3];
        get(s, j, y, n32, q);
        nv[0] = n32[0]; nv[1] = n32[1]; nv[2] = n32[2];
    }
};

// TPL targets per lane (one LDS read of a source then serves TPL pairs): a
// block of 256 / TPL threads still covers a tile of 256 targets; thread t
// holds targets t, t + 256 / TPL, ...
// TRANS: the transposed sum z_j = sum_i K(x_i, x_j) q_i (the kernel matrix
// transposed, for A^H ubar): the same pairs with the two values exchanged.
template <int PDE, int DIM, int KIND, bool MIXED, int G, int TPL, bool TRANS = false, bool CHIK = false>
__global__ __launch_bounds__(kSymBS / TPL) void k_apply_sym(SymArgs A) {
    // mixed 3D Helmholtz DL and DL + ch*SL: value and accumulation fused,
    // the products and the 64-term running sums in FP32 (added into FP64
    // after every 32-step pass); in FP64 the fused form's larger LDS record
    // cost more than the 5 % fewer VALU instructions saved, so FP64 keeps
    // the plain pair
    constexpr bool FUSED =
        MIXED && PDE == INTI_HELMHOLTZ && DIM == 3 && (KIND == INTI_DL || KIND == INTI_COMB_SD);
    // mixed 3D Laplace (every kind): 1/d from the FP32 unit,
    // the values, their products and the 64-term running sums in FP32, added
    // into FP64 after every pass, as above (the distance and the normal dot
    // products stay FP64)
    constexpr bool FUSEDL = MIXED && PDE == INTI_LAPLACE && DIM == 3;
    using R = std::conditional_t<FUSED, SymRecF<CHIK>, SymRec<PDE, DIM, KIND>>;
    constexpr bool CPLX = PDE == INTI_HELMHOLTZ;
    constexpr int BS = kSymBS / TPL;  // threads
    constexpr int NW = BS / 64;       // waves
    // two waves add their transposed sums onto zero with LDS atomics (two
    // addends: the result does not depend on their order); more waves write
    // their own rows, added in wave order
    constexpr bool RED_ADD = NW == 2;
    __shared__ double2 s_src[R::NS][kSymBS];
    __shared__ double2 s_red[RED_ADD ? 1 : NW][kSymBS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nxt = ((lane + 1) & 63) << 2;  // ds_bpermute address of lane l+1
    const int far = ((lane + 32) & 63) << 2;  // and of lane l+32
    int64_t a, b;
    sym_item(blockIdx.x, A.ngroups, a, b);
    const bool diag = a == b;
    const int64_t n = A.n, T = A.ntiles;
    const KParams &p = A.p;
    double adr[G][TPL], adi[G][TPL];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int t = 0; t < TPL; ++t) { adr[g][t] = 0.0; adi[g][t] = 0.0; }

    for (int gj = 0; gj < G; ++gj) {
        const int64_t J = b * G + gj;
        if (J >= T) break;
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            double y[3], nv[3], qr, qi;
            sym_node<DIM, CPLX>(A, J * kSymBS + tid + t * BS, y, nv, qr, qi);
            if constexpr (FUSED) R::put(s_src, tid + t * BS, y, nv, qext32(qr, qi, p));
            else R::put(s_src, tid + t * BS, y, nv, qr, qi);
            if constexpr (RED_ADD) s_red[0][tid + t * BS] = make_double2(0.0, 0.0);
        }
        __syncthreads();
        const bool trans = !diag || gj > 0;  // some tile I < J feeds tile J
        // with one target tile per block its nodes are read once per source
        // tile, not once per 64-source sub-tile
        double xh[TPL][3], nh[TPL][3], qhr[TPL], qhi[TPL];
        if constexpr (G == 1) {
#pragma unroll
            for (int t = 0; t < TPL; ++t)
                sym_node<DIM, CPLX>(A, a * kSymBS + tid + t * BS, xh[t], nh[t], qhr[t], qhi[t]);
        }
        for (int u = 0; u < 4; ++u) {
            // Two independent rotating chains of transposed sums: at step s
            // lane l pairs its targets with source (l + s) mod 64 (chain A,
            // running sum aa) and with source (l + s + 32) mod 64 (chain B,
            // running sum ab), s < 32, and hands both sums to lane l - 1. The
            // rotation of one chain completes while the other computes, so no
            // step waits on the crossbar move of the step before. After a
            // pass A's sum of source j sits in lane j - 32 and B's in lane j;
            // the two are exchanged (register names only), so at the start of
            // every pass lane l holds A for source l and B for source l + 32.
            double aar = 0.0, aai = 0.0, abr = 0.0, abi = 0.0;
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int64_t I = a * G + gi;
                if (I >= T || (diag && gi > gj)) continue;
                double xi[TPL][3], ni[TPL][3], qir[TPL], qii[TPL];
#pragma unroll
                for (int t = 0; t < TPL; ++t) {
                    if constexpr (G == 1) {
#pragma unroll
                        for (int d = 0; d < 3; ++d) { xi[t][d] = xh[t][d]; ni[t][d] = nh[t][d]; }
                        qir[t] = qhr[t];
                        qii[t] = qhi[t];
                    } else {
                        sym_node<DIM, CPLX>(A, I * kSymBS + tid + t * BS, xi[t], ni[t], qir[t], qii[t]);
                    }
                }
                if (diag && gi == gj) {
                    // the tile against itself: every ordered pair once
                    for (int s = 0; s < 64; ++s) {
                        const int jl = u * 64 + ((lane + s) & 63);
                        double yj[3], nj[3], qjr, qji;
                        if constexpr (FUSED) {
                            QExt32 qx;
                            R::get(s_src, jl, yj, nj, qx);
                            qjr = qx.qr;
                            qji = qx.qi;
                        } else {
                            R::get(s_src, jl, yj, nj, qjr, qji);
                        }
#pragma unroll
                        for (int t = 0; t < TPL; ++t) {
                            double kr, ki;
                            if constexpr (TRANS) pair_raw_xy<PDE, DIM, KIND, MIXED>(yj, nj, xi[t], ni[t], p, kr, ki);
                            else pair_raw_xy<PDE, DIM, KIND, MIXED>(xi[t], ni[t], yj, nj, p, kr, ki);
                            if constexpr (CPLX) {
                                adr[gi][t] = __builtin_fma(kr, qjr, adr[gi][t]);
                                adr[gi][t] = __builtin_fma(-ki, qji, adr[gi][t]);
                                adi[gi][t] = __builtin_fma(kr, qji, adi[gi][t]);
                                adi[gi][t] = __builtin_fma(ki, qjr, adi[gi][t]);
                            } else {
                                adr[gi][t] = __builtin_fma(kr, qjr, adr[gi][t]);
                            }
                        }
                    }
                } else {
                    // off-diagonal tiles of a quadrature with distinct nodes hold
                    // no coincident pairs: the pair values need no mask there
                    auto pairs = [&](auto mask_tag) {
                        constexpr bool MASK = decltype(mask_tag)::value;
                        if constexpr (FUSED) {
                            // FP32 sums of this pass: s1 for the lane's
                            // targets, the chains c for the rotating sources
                            QExt32 qit[TPL];
                            float s1r[TPL], s1i[TPL], nif[TPL][3];
#pragma unroll
                            for (int t = 0; t < TPL; ++t) {
                                qit[t] = qext32(qir[t], qii[t], p);
                                s1r[t] = s1i[t] = 0.0f;
#pragma unroll
                                for (int d = 0; d < 3; ++d) nif[t][d] = (float)ni[t][d];
                            }
                            float car = 0.0f, cai = 0.0f, cbr = 0.0f, cbi = 0.0f;
                            auto one = [&](int jl, float &ajr, float &aji) {
                                double yj[3];
                                float nj[3];
                                QExt32 qj;
                                R::get(s_src, jl, yj, nj, qj);
#pragma unroll
                                for (int t = 0; t < TPL; ++t) {
                                    double r[3];
#pragma unroll
                                    for (int d = 0; d < 3; ++d) r[d] = xi[t][d] - yj[d];
                                    helm3_sym_fused32<MASK, TRANS>(r, nif[t], nj, p, qj, qit[t], s1r[t], s1i[t],
                                                                   ajr, aji);
                                }
                            };
                            for (int s = 0; s < 32; ++s) {
                                const int jl = u * 64 + ((lane + s) & 63);
                                one(jl, car, cai);
                                one(jl ^ 32, cbr, cbi);
                                car = lane_next(car, nxt);
                                cai = lane_next(cai, nxt);
                                cbr = lane_next(cbr, nxt);
                                cbi = lane_next(cbi, nxt);
                            }
#pragma unroll
                            for (int t = 0; t < TPL; ++t) {
                                adr[gi][t] += (double)s1r[t];
                                adi[gi][t] += (double)s1i[t];
                            }
                            // the FP64 chain sums do not move: lane l keeps
                            // A's sum of source l and B's of source l + 32.
                            // After the pass the FP32 chains hold A's part of
                            // source l + 32 and B's of source l in lane l.
                            aar += (double)lane_next(car, far);
                            aai += (double)lane_next(cai, far);
                            abr += (double)lane_next(cbr, far);
                            abi += (double)lane_next(cbi, far);
                        } else if constexpr (FUSEDL) {
                            float s1[TPL], qf[TPL];
#pragma unroll
                            for (int t = 0; t < TPL; ++t) {
                                s1[t] = 0.0f;
                                qf[t] = (float)qir[t];
                            }
                            const float ch = (float)p.chr;
                            float ca = 0.0f, cb = 0.0f;
                            auto one = [&](int jl, float &aj) {
                                double yj[3], nj[3], qjr, qji;
                                R::get(s_src, jl, yj, nj, qjr, qji);
                                const float qj = (float)qjr;
#pragma unroll
                                for (int t = 0; t < TPL; ++t) {
                                    double r[3];
#pragma unroll
                                    for (int d = 0; d < 3; ++d) r[d] = xi[t][d] - yj[d];
                                    const double d2 = dotv<3>(r, r);
                                    float y = __builtin_amdgcn_rsqf(MASK ? fmaxf((float)d2, 1.0e-30f) : (float)d2);
                                    if constexpr (MASK) y = d2 > kSamePointTol2 ? y : 0.0f;
                                    // k1 = K(x_i, x_j), k2 = K(x_j, x_i); r = x_i - x_j
                                    float k1 = y, k2 = y;
                                    if constexpr (KIND != INTI_SL) {
                                        // r . n in FP64: on a smooth surface it is O(|r|^2) for
                                        // close nodes, and an FP32 dot of O(|r|) terms would lose
                                        // digits in proportion to 1/|r|
                                        const float y3 = y * y * y;
                                        const float rj = (float)dotv<3>(r, nj);
                                        const float ri = (float)dotv<3>(r, ni[t]);
                                        if constexpr (KIND == INTI_DL) {  // (r . n_source) / d^3
                                            k1 = rj * y3;
                                            k2 = -ri * y3;
                                        } else if constexpr (KIND == INTI_ADL) {  // (r . n_target) / d^3
                                            k1 = ri * y3;
                                            k2 = -rj * y3;
                                        } else if constexpr (KIND == INTI_COMB_SD) {  // DL + ch SL
                                            k1 = __builtin_fmaf(rj, y3, ch * y);
                                            k2 = __builtin_fmaf(-ri, y3, ch * y);
                                        } else {
                                            // HS: (n_i . n_j - 3 (r . n_i)(r . n_j) / d^2) / d^3, the
                                            // same both ways; HS + ch ADL adds ch (r . n_target) / d^3
                                            const float nn = (float)dotv<3>(ni[t], nj);
                                            const float hs = y3 * __builtin_fmaf(-3.0f * ri * rj, y * y, nn);
                                            if constexpr (KIND == INTI_HS) {
                                                k1 = hs;
                                                k2 = hs;
                                            } else {
                                                k1 = __builtin_fmaf(ch * ri, y3, hs);
                                                k2 = __builtin_fmaf(-ch * rj, y3, hs);
                                            }
                                        }
                                        if constexpr (TRANS) {
                                            const float tk = k1; k1 = k2; k2 = tk;
                                        }
                                    }
                                    s1[t] = __builtin_fmaf(k1, qj, s1[t]);  // K(x_i, x_j) q_j
                                    aj = __builtin_fmaf(k2, qf[t], aj);     // K(x_j, x_i) q_i
                                }
                            };
                            for (int s = 0; s < 32; ++s) {
                                const int jl = u * 64 + ((lane + s) & 63);
                                one(jl, ca);
                                one(jl ^ 32, cb);
                                ca = lane_next(ca, nxt);
                                cb = lane_next(cb, nxt);
                            }
#pragma unroll
                            for (int t = 0; t < TPL; ++t) adr[gi][t] += (double)s1[t];
                            aar += (double)lane_next(ca, far);
                            abr += (double)lane_next(cb, far);
                        } else {
                            auto one = [&](int jl, double &ajr, double &aji) {
                                double yj[3], nj[3], qjr, qji;
                                R::get(s_src, jl, yj, nj, qjr, qji);
#pragma unroll
                                for (int t = 0; t < TPL; ++t) {
                                    double r[DIM];
#pragma unroll
                                    for (int d = 0; d < DIM; ++d) r[d] = xi[t][d] - yj[d];
                                    double k1r, k1i, k2r, k2i;  // K(x_i, x_j), K(x_j, x_i)
                                    if constexpr (TRANS)
                                        pair_raw_sym<PDE, DIM, KIND, MIXED, MASK>(r, ni[t], nj, p, k2r, k2i, k1r, k1i);
                                    else
                                        pair_raw_sym<PDE, DIM, KIND, MIXED, MASK>(r, ni[t], nj, p, k1r, k1i, k2r, k2i);
                                    if constexpr (CPLX) {
                                        adr[gi][t] = __builtin_fma(k1r, qjr, adr[gi][t]);
                                        adr[gi][t] = __builtin_fma(-k1i, qji, adr[gi][t]);
                                        adi[gi][t] = __builtin_fma(k1r, qji, adi[gi][t]);
                                        adi[gi][t] = __builtin_fma(k1i, qjr, adi[gi][t]);
                                        ajr = __builtin_fma(k2r, qir[t], ajr);
                                        ajr = __builtin_fma(-k2i, qii[t], ajr);
                                        aji = __builtin_fma(k2r, qii[t], aji);
                                        aji = __builtin_fma(k2i, qir[t], aji);
                                    } else {
                                        adr[gi][t] = __builtin_fma(k1r, qjr, adr[gi][t]);
                                        ajr = __builtin_fma(k2r, qir[t], ajr);
                                    }
                                }
                            };
                            for (int s = 0; s < 32; ++s) {
                                const int jl = u * 64 + ((lane + s) & 63);
                                one(jl, aar, aai);
                                one(jl ^ 32, abr, abi);
                                aar = lane_next(aar, nxt);
                                abr = lane_next(abr, nxt);
                                if constexpr (CPLX) {
                                    aai = lane_next(aai, nxt);
                                    abi = lane_next(abi, nxt);
                                }
                            }
                        }
                    };
                    if (A.distinct) pairs(std::false_type{});
                    else pairs(std::true_type{});
                    if constexpr (!FUSED && !FUSEDL) {
                        // exchange the chains' roles for the next pass
                        double t0 = aar; aar = abr; abr = t0;
                        t0 = aai; aai = abi; abi = t0;
                    }
                }
            }
            // source l: A's part in this lane, B's in lane l + 32
            double ajr = aar + lane_next(abr, far);
            double aji = CPLX ? aai + lane_next(abi, far) : 0.0;
            if (trans) {
                if constexpr (RED_ADD) {
                    double *sr = &s_red[0][u * 64 + lane].x;
                    __hip_atomic_fetch_add(sr, ajr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if constexpr (CPLX)
                        __hip_atomic_fetch_add(sr + 1, aji, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    s_red[wave][u * 64 + lane] = make_double2(ajr, aji);
                }
            }
        }
        if (trans) {
            __syncthreads();
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
                const int jt = tid + t * BS;
                const int64_t j = J * kSymBS + jt;
                if (j < n) {
                    double vr = s_red[0][jt].x, vi = s_red[0][jt].y;
                    if constexpr (!RED_ADD) {
#pragma unroll
                        for (int w = 1; w < NW; ++w) { vr += s_red[w][jt].x; vi += s_red[w][jt].y; }
                    }
                    const int64_t off = a * n + j;
                    if constexpr (CPLX) reinterpret_cast<double2 *>(A.part)[off] = make_double2(vr, vi);
                    else A.part[off] = vr;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int64_t I = a * G + gi;
        if (I >= T) continue;
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            const int64_t i = I * kSymBS + tid + t * BS;
            if (i >= n) continue;
            const int64_t off = (diag ? a : b) * n + i;
            // in a diagonal block, tile gi > 0 already holds its transposed
            // sum, written above by this same thread
            const bool add = diag && gi > 0;
            if constexpr (CPLX) {
                double2 *pp = reinterpret_cast<double2 *>(A.part) + off;
                double2 v = make_double2(adr[gi][t], adi[gi][t]);
                if (add) { const double2 o = *pp; v.x += o.x; v.y += o.y; }
                *pp = v;
            } else {
                double v = adr[gi][t];
                if (add) v += A.part[off];
                A.part[off] = v;
            }
        }
    }
}

using SymFn = void (*)(const SymArgs &, int64_t, hipStream_t);

// targets per lane (one LDS read of a source serves TPL pairs): 4 for the
// cheap Laplace pairs, 2 for 3D Helmholtz, 1 for the 2D Helmholtz kernels,
// whose Hankel evaluation already fills the register file (measured best)
template <int PDE, int DIM, int KIND, bool MIXED, int G, bool TRANS = false>
static void sym_launcher(const SymArgs &a, int64_t nblocks, hipStream_t s) {
    constexpr int TPL = PDE == INTI_LAPLACE ? 4 : DIM == 2 ? 1 : 2;
    // the fused mixed combined field with ch = -i k has its own (smaller) record
    constexpr bool CHIK_KERNEL = MIXED && !TRANS && PDE == INTI_HELMHOLTZ && DIM == 3 && KIND == INTI_COMB_SD;
    if constexpr (CHIK_KERNEL) {
        if (a.ch_ik) {
            hipLaunchKernelGGL((k_apply_sym<PDE, DIM, KIND, MIXED, G, TPL, TRANS, true>), dim3((unsigned)nblocks),
                               dim3(kSymBS / TPL), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((k_apply_sym<PDE, DIM, KIND, MIXED, G, TPL, TRANS>), dim3((unsigned)nblocks),
                       dim3(kSymBS / TPL), 0, s, a);
}

template <int PDE, int DIM, bool MIXED, int G>
static SymFn sym_kind(int kind) {
    switch (kind) {
        case INTI_SL: return sym_launcher<PDE, DIM, INTI_SL, MIXED, G>;
        case INTI_DL: return sym_launcher<PDE, DIM, INTI_DL, MIXED, G>;
        case INTI_ADL: return sym_launcher<PDE, DIM, INTI_ADL, MIXED, G>;
        case INTI_HS: return sym_launcher<PDE, DIM, INTI_HS, MIXED, G>;
        case INTI_COMB_SD: return sym_launcher<PDE, DIM, INTI_COMB_SD, MIXED, G>;
        case INTI_COMB_AH: return sym_launcher<PDE, DIM, INTI_COMB_AH, MIXED, G>;
    }
    return nullptr;
}

// G in {1, 2, 4}
template <int PDE, int DIM>
static SymFn sym_pick_pd(int kind, int mixed, int G) {
    if (mixed) {
        if (G == 1) return sym_kind<PDE, DIM, true, 1>(kind);
        if (G == 2) return sym_kind<PDE, DIM, true, 2>(kind);
        if (G == 4) return sym_kind<PDE, DIM, true, 4>(kind);
    } else {
        if (G == 1) return sym_kind<PDE, DIM, false, 1>(kind);
        if (G == 2) return sym_kind<PDE, DIM, false, 2>(kind);
        if (G == 4) return sym_kind<PDE, DIM, false, 4>(kind);
    }
    return nullptr;
}

// transposed FP64 sums of the kinds with a pullback: SL is its own
// transpose; DL and c0 SL + c1 DL need the exchanged values
template <int PDE, int DIM, int G>
static SymFn sym_trans_kind(int kind) {
    switch (kind) {
        case INTI_SL: return sym_launcher<PDE, DIM, INTI_SL, false, G>;
        case INTI_DL: return sym_launcher<PDE, DIM, INTI_DL, false, G, true>;
        case INTI_COMB_SD: return sym_launcher<PDE, DIM, INTI_COMB_SD, false, G, true>;
    }
    return nullptr;
}

template <int PDE, int DIM>
static SymFn sym_trans_pick_pd(int kind, int G) {
    if (G == 1) return sym_trans_kind<PDE, DIM, 1>(kind);
    if (G == 2) return sym_trans_kind<PDE, DIM, 2>(kind);
    if (G == 4) return sym_trans_kind<PDE, DIM, 4>(kind);
    return nullptr;
}

SymFn sym_trans_pick_lap2(int kind, int G);
SymFn sym_trans_pick_lap3(int kind, int G);
SymFn sym_trans_pick_helm2(int kind, int G);
SymFn sym_trans_pick_helm3(int kind, int G);

SymFn sym_pick_lap2(int kind, int mixed, int G);
SymFn sym_pick_lap3(int kind, int mixed, int G);
SymFn sym_pick_helm2(int kind, int mixed, int G);
SymFn sym_pick_helm3(int kind, int mixed, int G);

}  // namespace inti
