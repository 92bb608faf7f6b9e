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
// pairs I <= J only). Wave w of a block holds 64 targets. For each 64-source
// sub-tile, lane l at step s pairs its target with source (l + s) mod 64, so
// the 64 lanes touch 64 different sources, read from LDS without bank
// conflicts. The transposed sum for source (l + s) mod 64 travels with the
// step: after each step lane l takes lane l+1's running sum (ds_bpermute, a
// crossbar move with no LDS storage), and after 64 steps lane l holds the sum
// for source l over the wave's 64 targets. The four waves' sums are added in
// LDS in a fixed order.
//
// Partial sums. A row of group c receives exactly one partial per group:
// from block (c, b), b >= c (direct part) in slot b, and from block (a, c),
// a < c (transposed part) in slot a of part[S][n]. In a diagonal block both
// land in slot a and are added by the thread that owns the row. Every slot is
// written once, in a fixed order, so the result does not depend on the
// schedule; the reduction kernel of the plain direct sum (split = S) sums the
// slots and applies the post factor, the correction and the identity.
#pragma once
#include "apply_kernel.hpp"

namespace inti {

struct SymArgs {
    int64_t n;
    const double *xy, *nrm, *w;  // SoA coordinates / normals (stride n), weights
    const double *x;             // density (complex interleaved for Helmholtz)
    double *part;                // S x n partial sums (complex interleaved)
    int64_t ntiles, ngroups;     // T, S
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

// node j's geometry and q_j = w_j x_j; far-away padding with zero weight
// beyond n
template <int DIM, bool CPLX>
__device__ __forceinline__ void sym_node(const SymArgs &A, int64_t j, double *y, double *nv,
                                         double &qr, double &qi) {
    if (j < A.n) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            y[d] = A.xy[d * A.n + j];
            nv[d] = A.nrm ? A.nrm[d * A.n + j] : 0.0;
        }
        const double w = A.w[j];
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

template <int DIM>
__device__ __forceinline__ void sym_read(const double2 (*s)[kSymBS], int j, double *y, double *nv,
                                         double &qr, double &qi) {
    if constexpr (DIM == 3) {
        const double2 v0 = s[0][j], v1 = s[1][j], v2 = s[2][j], v3 = s[3][j];
        y[0] = v0.x; y[1] = v0.y; y[2] = v1.x;
        nv[0] = v1.y; nv[1] = v2.x; nv[2] = v2.y;
        qr = v3.x; qi = v3.y;
    } else {
        const double2 v0 = s[0][j], v1 = s[1][j], v2 = s[2][j];
        y[0] = v0.x; y[1] = v0.y;
        nv[0] = v1.x; nv[1] = v1.y;
        qr = v2.x; qi = v2.y;
    }
}

template <int PDE, int DIM, int KIND, bool MIXED, int G>
__global__ __launch_bounds__(kSymBS) void k_apply_sym(SymArgs A) {
    constexpr bool CPLX = PDE == INTI_HELMHOLTZ;
    constexpr int NS = DIM == 3 ? 4 : 3;
    __shared__ double2 s_src[NS][kSymBS];
    __shared__ double2 s_red[4][kSymBS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nxt = ((lane + 1) & 63) << 2;  // ds_bpermute address of lane l+1
    int64_t a, b;
    sym_item(blockIdx.x, A.ngroups, a, b);
    const bool diag = a == b;
    const int64_t n = A.n, T = A.ntiles;
    const KParams &p = A.p;
    double adr[G], adi[G];
#pragma unroll
    for (int g = 0; g < G; ++g) { adr[g] = 0.0; adi[g] = 0.0; }

    for (int gj = 0; gj < G; ++gj) {
        const int64_t J = b * G + gj;
        if (J >= T) break;
        {
            double y[3], nv[3], qr, qi;
            sym_node<DIM, CPLX>(A, J * kSymBS + tid, y, nv, qr, qi);
            if constexpr (DIM == 3) {
                s_src[0][tid] = make_double2(y[0], y[1]);
                s_src[1][tid] = make_double2(y[2], nv[0]);
                s_src[2][tid] = make_double2(nv[1], nv[2]);
                s_src[3][tid] = make_double2(qr, qi);
            } else {
                s_src[0][tid] = make_double2(y[0], y[1]);
                s_src[1][tid] = make_double2(nv[0], nv[1]);
                s_src[2][tid] = make_double2(qr, qi);
            }
        }
        __syncthreads();
        const bool trans = !diag || gj > 0;  // some tile I < J feeds tile J
        for (int u = 0; u < 4; ++u) {
            double ajr = 0.0, aji = 0.0;  // running transposed sum of source (lane + s) mod 64
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int64_t I = a * G + gi;
                if (I >= T || (diag && gi > gj)) continue;
                double xi[3], ni[3], qir, qii;
                sym_node<DIM, CPLX>(A, I * kSymBS + tid, xi, ni, qir, qii);
                if (diag && gi == gj) {
                    // the tile against itself: every ordered pair once
                    for (int s = 0; s < 64; ++s) {
                        const int jl = u * 64 + ((lane + s) & 63);
                        double yj[3], nj[3], qjr, qji;
                        sym_read<DIM>(s_src, jl, yj, nj, qjr, qji);
                        double kr, ki;
                        pair_raw_xy<PDE, DIM, KIND, MIXED>(xi, ni, yj, nj, p, kr, ki);
                        if constexpr (CPLX) {
                            adr[gi] = __builtin_fma(kr, qjr, adr[gi]);
                            adr[gi] = __builtin_fma(-ki, qji, adr[gi]);
                            adi[gi] = __builtin_fma(kr, qji, adi[gi]);
                            adi[gi] = __builtin_fma(ki, qjr, adi[gi]);
                        } else {
                            adr[gi] = __builtin_fma(kr, qjr, adr[gi]);
                        }
                    }
                } else {
#pragma unroll 2
                    for (int s = 0; s < 64; ++s) {
                        const int jl = u * 64 + ((lane + s) & 63);
                        double yj[3], nj[3], qjr, qji;
                        sym_read<DIM>(s_src, jl, yj, nj, qjr, qji);
                        double r[DIM], rn[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            r[d] = xi[d] - yj[d];
                            rn[d] = -r[d];
                        }
                        double k1r, k1i, k2r, k2i;
                        pair_raw<PDE, DIM, KIND, MIXED>(r, ni, nj, p, k1r, k1i);   // K(x_i, x_j)
                        pair_raw<PDE, DIM, KIND, MIXED>(rn, nj, ni, p, k2r, k2i);  // K(x_j, x_i)
                        if constexpr (CPLX) {
                            adr[gi] = __builtin_fma(k1r, qjr, adr[gi]);
                            adr[gi] = __builtin_fma(-k1i, qji, adr[gi]);
                            adi[gi] = __builtin_fma(k1r, qji, adi[gi]);
                            adi[gi] = __builtin_fma(k1i, qjr, adi[gi]);
                            ajr = __builtin_fma(k2r, qir, ajr);
                            ajr = __builtin_fma(-k2i, qii, ajr);
                            aji = __builtin_fma(k2r, qii, aji);
                            aji = __builtin_fma(k2i, qir, aji);
                            aji = lane_next(aji, nxt);
                        } else {
                            adr[gi] = __builtin_fma(k1r, qjr, adr[gi]);
                            ajr = __builtin_fma(k2r, qir, ajr);
                        }
                        ajr = lane_next(ajr, nxt);
                    }
                }
            }
            if (trans) s_red[wave][u * 64 + lane] = make_double2(ajr, aji);
        }
        if (trans) {
            __syncthreads();
            const int64_t j = J * kSymBS + tid;
            if (j < n) {
                const double2 v0 = s_red[0][tid], v1 = s_red[1][tid], v2 = s_red[2][tid],
                              v3 = s_red[3][tid];
                const double vr = ((v0.x + v1.x) + v2.x) + v3.x;
                const int64_t off = a * n + j;
                if constexpr (CPLX) {
                    const double vi = ((v0.y + v1.y) + v2.y) + v3.y;
                    reinterpret_cast<double2 *>(A.part)[off] = make_double2(vr, vi);
                } else {
                    A.part[off] = vr;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int64_t I = a * G + gi;
        const int64_t i = I * kSymBS + tid;
        if (I >= T || i >= n) continue;
        const int64_t off = (diag ? a : b) * n + i;
        // in a diagonal block, tile gi > 0 already holds its transposed sum,
        // written above by this same thread
        const bool add = diag && gi > 0;
        if constexpr (CPLX) {
            double2 *pp = reinterpret_cast<double2 *>(A.part) + off;
            double2 v = make_double2(adr[gi], adi[gi]);
            if (add) { const double2 o = *pp; v.x += o.x; v.y += o.y; }
            *pp = v;
        } else {
            double v = adr[gi];
            if (add) v += A.part[off];
            A.part[off] = v;
        }
    }
}

using SymFn = void (*)(const SymArgs &, int64_t, hipStream_t);

template <int PDE, int DIM, int KIND, bool MIXED, int G>
static void sym_launcher(const SymArgs &a, int64_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL((k_apply_sym<PDE, DIM, KIND, MIXED, G>), dim3((unsigned)nblocks),
                       dim3(kSymBS), 0, s, a);
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

SymFn sym_pick_lap2(int kind, int mixed, int G);
SymFn sym_pick_lap3(int kind, int mixed, int G);
SymFn sym_pick_helm2(int kind, int mixed, int G);
SymFn sym_pick_helm3(int kind, int mixed, int G);

}  // namespace inti
