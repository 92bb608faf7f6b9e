// Tiled target x source direct sum (matrix-free IntegralOperator apply) for
// NR right-hand sides at once.
//
// Layout: one thread per target, 256 targets per block; the source range is
// cut into `split` chunks so that a launch has >> 256 workgroups. Each source
// tile is read with coalesced SoA loads and staged in LDS: the geometry as a
// packed record and q_j,r = w_j sigma_j,r for the NR right-hand sides; every
// lane then reads them by broadcast. The kernel value of a pair is computed
// once and applied to all NR densities. Block b works on chunk b % split, so
// with split a multiple of 8 all blocks of one XCD (blocks are dealt to the 8
// XCDs round robin) share one source chunk in their L2.
#pragma once
#include "common.hpp"
#include "pair.hpp"

namespace inti {

struct ApplyArgs {
    int64_t m, n;
    int64_t tstride, sstride;  // SoA strides of the target / source arrays
    const double *tx, *tn;
    const double *sx, *sn, *sw;
    const double *x;  // n x nr, leading dimension ldx (elements)
    int64_t ldx;
    double *part;     // nr x split x m partial sums
    int split;
    int64_t chunk;
    KParams p;
};

// far-away padding source for the last tile (its density is zero)
constexpr double kPadCoord = 1.0e3;
constexpr int kApplyBS = 256;

template <int NR>
struct ApplyTile {
    static constexpr int TT = NR == 1 ? 256 : NR <= 4 ? 128 : 64;
};

// One right-hand side of the mixed 3D Helmholtz DL / DL + ch*SL kinds: the
// pair value and its accumulation are fused (helm3_fused_one32), with q,
// (k Im q, -k Re q) and ch q staged per source in FP32 (three double2 slots
// reserved, two used).
template <int PDE, int DIM, int KIND, bool MIXED, int NR, bool TRANS>
constexpr bool apply_fused() {
    return NR == 1 && !TRANS && MIXED && PDE == INTI_HELMHOLTZ && DIM == 3 &&
           (KIND == INTI_DL || KIND == INTI_COMB_SD);
}

// The work of one block: targets [tile*256, tile*256 + 256) against the
// sources of chunk `chunk`; partial sums to a.part. s_geo / s_q are the
// block's LDS staging buffers. Without weights (a.sw null) q_j = x_j.
// TRANS: the sum of K(source, target) instead of K(target, source), i.e.
// the kernel matrix of the operator whose targets and sources are these
// sources and targets, transposed (for A^H ubar).
template <int PDE, int DIM, int KIND, bool MIXED, int NR, bool TRANS = false>
__device__ __forceinline__ void apply_block(const ApplyArgs &a, int64_t chunk, int64_t tile,
                                            double2 *s_geo, double2 *s_q) {
    constexpr bool CPLX = PDE != INTI_LAPLACE;
    constexpr bool FUSED = apply_fused<PDE, DIM, KIND, MIXED, NR, TRANS>();
    constexpr int BS = kApplyBS;
    constexpr int TT = ApplyTile<NR>::TT;
    constexpr int NG = DIM == 3 ? 3 : 2;  // double2 slots of a geometry record
    const int tid = threadIdx.x;
    const int64_t m = a.m, n = a.n;
    const int64_t ts = a.tstride, ss = a.sstride;

    const int64_t i = tile * BS + tid;
    const int64_t ic = i < m ? i : m - 1;
    double xt[DIM], nt[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xt[d] = a.tx[d * ts + ic];
        nt[d] = a.tn ? a.tn[d * ts + ic] : 0.0;
    }
    double accr[NR], acci[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) { accr[r] = 0.0; acci[r] = 0.0; }
    const int64_t j0 = chunk * a.chunk;
    const int64_t j1 = j0 + a.chunk < n ? j0 + a.chunk : n;
    for (int64_t s0 = j0; s0 < j1; s0 += TT) {
        if (tid < TT) {
            const int64_t j = s0 + tid;
            double y[3] = {kPadCoord, kPadCoord, kPadCoord}, ny[3] = {0, 0, 0};
            if (j < j1) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    y[d] = a.sx[d * ss + j];
                    ny[d] = a.sn ? a.sn[d * ss + j] : 0.0;
                }
            }
            double2 *rec = s_geo + tid * NG;
            if constexpr (DIM == 3) {
                rec[0] = make_double2(y[0], y[1]);
                rec[1] = make_double2(y[2], ny[0]);
                rec[2] = make_double2(ny[1], ny[2]);
            } else {
                rec[0] = make_double2(y[0], y[1]);
                rec[1] = make_double2(ny[0], ny[1]);
            }
        }
        for (int q = tid; q < TT * NR; q += BS) {
            const int jj = q % TT, r = q / TT;
            const int64_t j = s0 + jj;
            double qr = 0.0, qi = 0.0;
            if (j < j1) {
                const double w = a.sw ? a.sw[j] : 1.0;
                if constexpr (CPLX) {
                    const double *xp = a.x + 2 * (j + r * a.ldx);
                    qr = w * xp[0];
                    qi = w * xp[1];
                } else {
                    qr = w * a.x[j + r * a.ldx];
                }
            }
            if constexpr (FUSED) {
                const QExt32 qx = qext32(qr, qi, a.p);
                *reinterpret_cast<float4 *>(&s_q[3 * jj]) = make_float4(qx.qr, qx.qi, qx.kqi, qx.mkqr);
                *reinterpret_cast<float2 *>(&s_q[3 * jj + 1]) = make_float2(qx.cr, qx.ci);
            } else {
                s_q[jj * NR + r] = make_double2(qr, qi);
            }
        }
        __syncthreads();
        // two pairs in flight for one right-hand side; one for many (the
        // accumulators already fill the register file) and for the 2D
        // Helmholtz kernels (the Hankel evaluation does)
        constexpr int UNR = (NR == 1 && !(CPLX && DIM == 2)) ? 2 : 1;
        if constexpr (FUSED) {
            // FP32 products and 64-term FP32 running sums added into FP64
            for (int jb = 0; jb < TT; jb += 64) {
                float sr = 0.0f, si = 0.0f;
#pragma unroll UNR
                for (int jj = jb; jj < jb + 64; ++jj) {
                    const double2 *rc = s_geo + jj * NG;
                    const double2 v0 = rc[0], v1 = rc[1], v2 = rc[2];
                    const double ys[3] = {v0.x, v0.y, v1.x}, ns[3] = {v1.y, v2.x, v2.y};
                    const float4 f = *reinterpret_cast<const float4 *>(&s_q[3 * jj]);
                    const float2 g = *reinterpret_cast<const float2 *>(&s_q[3 * jj + 1]);
                    QExt32 qx;
                    qx.qr = f.x; qx.qi = f.y; qx.kqi = f.z; qx.mkqr = f.w; qx.cr = g.x; qx.ci = g.y;
                    double r3[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) r3[d] = xt[d] - ys[d];
                    helm3_fused_one32(r3, ns, a.p, qx, sr, si);
                }
                accr[0] += (double)sr;
                acci[0] += (double)si;
            }
            __syncthreads();
            continue;
        }
#pragma unroll UNR
        for (int jj = 0; jj < TT; ++jj) {
            const double2 *rc = s_geo + jj * NG;
            double ys[DIM], ns[DIM];
            if constexpr (DIM == 3) {
                const double2 v0 = rc[0], v1 = rc[1], v2 = rc[2];
                ys[0] = v0.x; ys[1] = v0.y; ys[2] = v1.x;
                ns[0] = v1.y; ns[1] = v2.x; ns[2] = v2.y;
            } else {
                const double2 v0 = rc[0], v1 = rc[1];
                ys[0] = v0.x; ys[1] = v0.y;
                ns[0] = v1.x; ns[1] = v1.y;
            }
            double kr, ki;
            if constexpr (TRANS) pair_raw_xy<PDE, DIM, KIND, MIXED>(ys, ns, xt, nt, a.p, kr, ki);
            else pair_raw_xy<PDE, DIM, KIND, MIXED>(xt, nt, ys, ns, a.p, kr, ki);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double2 q = s_q[jj * NR + r];
                if constexpr (CPLX) {
                    accr[r] = __builtin_fma(kr, q.x, accr[r]);
                    accr[r] = __builtin_fma(-ki, q.y, accr[r]);
                    acci[r] = __builtin_fma(kr, q.y, acci[r]);
                    acci[r] = __builtin_fma(ki, q.x, acci[r]);
                } else {
                    accr[r] = __builtin_fma(kr, q.x, accr[r]);
                }
            }
        }
        __syncthreads();
    }
    if (i < m) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int64_t off = ((int64_t)r * a.split + chunk) * m + i;
            if constexpr (CPLX) reinterpret_cast<double2 *>(a.part)[off] = make_double2(accr[r], acci[r]);
            else a.part[off] = accr[r];
        }
    }
}

// ------------------------------------------- sixteen right-hand sides (MFMA)
// Y (targets x 16) += K (targets x sources) Q (sources x 16) with the kernel
// values made on the vector ALUs and the products on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), which run beside the VALU: the 16 x 4 fragment of
// K that one MFMA consumes is exactly one kernel value per lane. Lane l of a
// wavefront evaluates target l % 16 of a 16-target fragment against source
// l / 16 of the current group of 4 sources, which is the A-operand layout;
// the B operand (source l / 16, right-hand side l % 16 of Q) is one LDS read.
// A complex product takes four real MFMAs per fragment:
//   Re Y += Re K Re Q + (-Im K) Im Q,   Im Y += Re K Im Q + Im K Re Q.
// Each wavefront owns kMfmaNF fragments and a block 256 / (16 kMfmaNF)
// wavefronts, so a block still covers 256 targets and the partial-sum layout
// and grid are those of apply_block.
// Accumulator layout: element v of lane l is Y[4 v + l / 16][l % 16].
// Mixed precision: the kernel values and the weighted densities are rounded
// to FP32 and multiplied on the FP32 matrix cores (v_mfma_f32_16x16x4f32,
// twice the FP64 rate), summed in FP32 over one LDS tile of 64 sources and
// added into the FP64 accumulators; there element v of lane l holds
// Y[4 (l / 16) + v][l % 16].
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kMfmaTT = 64;  // sources per LDS tile
constexpr int kMfmaQLd = 17; // LDS row of one source's 16 densities, padded
constexpr int kMfmaNF = 4;   // 16-target fragments per wavefront

// threads of a block of the direct sum: 256 targets per block either way
template <int NR>
constexpr int apply_threads() { return NR == 16 ? 64 * (kApplyBS / 16) / kMfmaNF : kApplyBS; }

// one source's coordinates and normal from its LDS record
template <int DIM>
__device__ __forceinline__ void load_geo(const double2 *rc, double *ys, double *ns) {
    if constexpr (DIM == 3) {
        const double2 v0 = rc[0], v1 = rc[1], v2 = rc[2];
        ys[0] = v0.x; ys[1] = v0.y; ys[2] = v1.x;
        ns[0] = v1.y; ns[1] = v2.x; ns[2] = v2.y;
    } else {
        const double2 v0 = rc[0], v1 = rc[1];
        ys[0] = v0.x; ys[1] = v0.y;
        ns[0] = v1.x; ns[1] = v1.y;
    }
}

template <int PDE, int DIM, int KIND, bool MIXED>
__device__ __forceinline__ void apply_block_mfma(const ApplyArgs &a, int64_t chunk, int64_t tile,
                                                 double2 *s_geo, double2 *s_q) {
    constexpr bool CPLX = PDE != INTI_LAPLACE;
    constexpr int NR = 16, TT = kMfmaTT, NF = kMfmaNF, NT = apply_threads<16>();
    constexpr int NG = DIM == 3 ? 3 : 2;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t m = a.m, n = a.n;
    const int64_t ts = a.tstride, ss = a.sstride;
    const int64_t t0 = tile * kApplyBS + wave * 16 * NF;

    double xt[NF][DIM], nt[NF][DIM];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int64_t i = t0 + 16 * f + li;
        const int64_t ic = i < m ? i : m - 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xt[f][d] = a.tx[d * ts + ic];
            nt[f][d] = a.tn ? a.tn[d * ts + ic] : 0.0;
        }
    }
    float2 *s_q32 = reinterpret_cast<float2 *>(s_q);
    v4d accr[NF], acci[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        accr[f] = v4d{0, 0, 0, 0};
        acci[f] = v4d{0, 0, 0, 0};
    }
    const int64_t j0 = chunk * a.chunk;
    const int64_t j1 = j0 + a.chunk < n ? j0 + a.chunk : n;
    for (int64_t s0 = j0; s0 < j1; s0 += TT) {
        if (tid < TT) {
            const int64_t j = s0 + tid;
            double y[3] = {kPadCoord, kPadCoord, kPadCoord}, ny[3] = {0, 0, 0};
            if (j < j1) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    y[d] = a.sx[d * ss + j];
                    ny[d] = a.sn ? a.sn[d * ss + j] : 0.0;
                }
            }
            double2 *rec = s_geo + tid * NG;
            if constexpr (DIM == 3) {
                rec[0] = make_double2(y[0], y[1]);
                rec[1] = make_double2(y[2], ny[0]);
                rec[2] = make_double2(ny[1], ny[2]);
            } else {
                rec[0] = make_double2(y[0], y[1]);
                rec[1] = make_double2(ny[0], ny[1]);
            }
        }
#pragma unroll
        for (int q = tid; q < TT * NR; q += NT) {
            const int jj = q % TT, r = q / TT;
            const int64_t j = s0 + jj;
            double qr = 0.0, qi = 0.0;
            if (j < j1) {
                const double w = a.sw[j];
                if constexpr (CPLX) {
                    const double *xp = a.x + 2 * (j + r * a.ldx);
                    qr = w * xp[0];
                    qi = w * xp[1];
                } else {
                    qr = w * a.x[j + r * a.ldx];
                }
            }
            if constexpr (MIXED) s_q32[jj * kMfmaQLd + r] = make_float2((float)qr, (float)qi);
            else s_q[jj * kMfmaQLd + r] = make_double2(qr, qi);
        }
        __syncthreads();
        if constexpr (MIXED) {
            // complex products by Gauss's three real products:
            // P1 = Re K Re Q, P2 = Im K Im Q, P3 = (Re K + Im K)(Re Q + Im Q),
            // Re Y = P1 - P2, Im Y = P3 - P1 - P2
            v4f t1[NF], t2[NF], t3[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                t1[f] = v4f{0, 0, 0, 0};
                t2[f] = v4f{0, 0, 0, 0};
                t3[f] = v4f{0, 0, 0, 0};
            }
            for (int st = 0; st < TT / 4; ++st) {
                const int jj = 4 * st + lk;
                double ys[DIM], ns[DIM];
                load_geo<DIM>(s_geo + jj * NG, ys, ns);
                const float2 q = s_q32[jj * kMfmaQLd + li];
                const float qs = q.x + q.y;
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    double kr, ki;
                    pair_raw_xy<PDE, DIM, KIND, MIXED>(xt[f], nt[f], ys, ns, a.p, kr, ki);
                    const float fr = (float)kr;
                    t1[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(fr, q.x, t1[f], 0, 0, 0);
                    if constexpr (CPLX) {
                        const float fi = (float)ki;
                        t2[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(fi, q.y, t2[f], 0, 0, 0);
                        t3[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(fr + fi, qs, t3[f], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if constexpr (CPLX) {
                        const double p1 = t1[f][v], p2 = t2[f][v];
                        accr[f][v] += p1 - p2;
                        acci[f][v] += (double)t3[f][v] - p1 - p2;
                    } else {
                        accr[f][v] += (double)t1[f][v];
                    }
                }
            }
            __syncthreads();
            continue;
        }
        for (int st = 0; st < TT / 4; ++st) {
            const int jj = 4 * st + lk;
            double ys[DIM], ns[DIM];
            load_geo<DIM>(s_geo + jj * NG, ys, ns);
            const double2 q = s_q[jj * kMfmaQLd + li];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                double kr, ki;
                pair_raw_xy<PDE, DIM, KIND, MIXED>(xt[f], nt[f], ys, ns, a.p, kr, ki);
                accr[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(kr, q.x, accr[f], 0, 0, 0);
                if constexpr (CPLX) {
                    acci[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(kr, q.y, acci[f], 0, 0, 0);
                    accr[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ki, q.y, accr[f], 0, 0, 0);
                    acci[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(ki, q.x, acci[f], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t i = t0 + 16 * f + (MIXED ? 4 * lk + v : 4 * v + lk);
            if (i < m) {
                const int64_t off = ((int64_t)li * a.split + chunk) * m + i;
                if constexpr (CPLX) reinterpret_cast<double2 *>(a.part)[off] = make_double2(accr[f][v], acci[f][v]);
                else a.part[off] = accr[f][v];
            }
        }
    }
}

// LDS of one block: the geometry of a source tile and its densities (three
// slots per source for the fused pair)
template <int DIM, int NR, bool FUSED = false>
struct ApplyLds {
    static constexpr int NG = DIM == 3 ? 3 : 2;
    static constexpr int GEO = (NR == 16 ? kMfmaTT : ApplyTile<NR>::TT) * NG;
    static constexpr int Q = NR == 16 ? kMfmaTT * kMfmaQLd : ApplyTile<NR>::TT * NR * (FUSED ? 3 : 1);
};

template <int PDE, int DIM, int KIND, bool MIXED, int NR, bool TRANS = false>
__device__ __forceinline__ void apply_any(const ApplyArgs &a, int64_t chunk, int64_t tile, double2 *s_geo,
                                          double2 *s_q) {
    static_assert(!(TRANS && NR == 16), "the transposed sum has one right-hand side");
    if constexpr (NR == 16) apply_block_mfma<PDE, DIM, KIND, MIXED>(a, chunk, tile, s_geo, s_q);
    else apply_block<PDE, DIM, KIND, MIXED, NR, TRANS>(a, chunk, tile, s_geo, s_q);
}

template <int PDE, int DIM, int KIND, bool MIXED, int NR, bool TRANS = false>
__global__ __launch_bounds__(apply_threads<NR>()) void k_apply(ApplyArgs a) {
    using L = ApplyLds<DIM, NR, apply_fused<PDE, DIM, KIND, MIXED, NR, TRANS>()>;
    __shared__ double2 s_geo[L::GEO];
    __shared__ double2 s_q[L::Q];
    const int64_t b = blockIdx.x;
    apply_any<PDE, DIM, KIND, MIXED, NR, TRANS>(a, b % a.split, b / a.split, s_geo, s_q);
}

// ------------------------------------------------------------ batched apply
// Many independent problems (block-diagonal operator) in one persistent
// launch: the work list holds (problem, target tile, source chunk) items,
// sorted by decreasing size; block b takes items b, b + G, b + 2G, ... so
// that the large items are spread first and every block ends with small
// ones. Problem p's nodes are rows [off, off + n) of the concatenated SoA
// arrays (stride ntot); its partial sums start at part + poff.
struct BatchItem {
    int32_t p, tile, chunk, pad;
};
struct BatchProb {
    int64_t off, n, poff, chunk;
    int32_t split, pad;
    KParams kp;
};
struct BatchArgs {
    const BatchItem *items;
    int64_t nitems;
    const BatchProb *probs;
    int64_t ntot;
    const double *coords, *normals, *weights;
    const double *x;  // ntot x nr (leading dimension ldx)
    int64_t ldx;
    double *part;
};

template <int PDE, int DIM, int KIND, bool MIXED, int NR>
__global__ __launch_bounds__(apply_threads<NR>()) void k_apply_batch(BatchArgs b) {
    constexpr int CW = PDE != INTI_LAPLACE ? 2 : 1;
    using L = ApplyLds<DIM, NR, apply_fused<PDE, DIM, KIND, MIXED, NR, false>()>;
    __shared__ double2 s_geo[L::GEO];
    __shared__ double2 s_q[L::Q];
    for (int64_t it = blockIdx.x; it < b.nitems; it += gridDim.x) {
        const BatchItem w = b.items[it];
        const BatchProb P = b.probs[w.p];
        ApplyArgs a;
        a.m = P.n; a.n = P.n;
        a.tstride = b.ntot; a.sstride = b.ntot;
        a.tx = b.coords + P.off; a.tn = b.normals ? b.normals + P.off : nullptr;
        a.sx = a.tx; a.sn = a.tn; a.sw = b.weights + P.off;
        a.x = b.x + CW * P.off; a.ldx = b.ldx;
        a.part = b.part + CW * P.poff;
        a.split = P.split; a.chunk = P.chunk;
        a.p = P.kp;
        apply_any<PDE, DIM, KIND, MIXED, NR>(a, w.chunk, w.tile, s_geo, s_q);
    }
}

// Launches the kernel; with a null `a` it launches nothing and returns the
// kernel's resident blocks per CU (for the launch plan).
using ApplyFn = int (*)(const ApplyArgs *a, dim3, hipStream_t);

template <int PDE, int DIM, int KIND, bool MIXED, int NR, bool TRANS = false>
static int apply_launcher(const ApplyArgs *a, dim3 grid, hipStream_t s) {
    if (!a) {
        static int cached = 0;  // one gfx950 occupancy per kernel; a race only repeats the query
        if (cached == 0) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &nb, (const void *)k_apply<PDE, DIM, KIND, MIXED, NR, TRANS>, apply_threads<NR>(), 0) != hipSuccess)
                nb = 1;
            cached = nb < 1 ? 1 : nb;
        }
        return cached;
    }
    hipLaunchKernelGGL((k_apply<PDE, DIM, KIND, MIXED, NR, TRANS>), grid, dim3(apply_threads<NR>()), 0, s, *a);
    return 0;
}

// transposed FP64 sums (one right-hand side) of the kinds with a pullback
template <int PDE, int DIM>
static ApplyFn apply_trans_pick_pd(int kind) {
    switch (kind) {
        case INTI_SL: return apply_launcher<PDE, DIM, INTI_SL, false, 1, true>;
        case INTI_DL: return apply_launcher<PDE, DIM, INTI_DL, false, 1, true>;
        case INTI_COMB_SD: return apply_launcher<PDE, DIM, INTI_COMB_SD, false, 1, true>;
    }
    if constexpr (PDE == INTI_HELMHOLTZ) {
        switch (kind) {
            case INTI_FF_SL: return apply_launcher<PDE, DIM, INTI_FF_SL, false, 1, true>;
            case INTI_FF_DL: return apply_launcher<PDE, DIM, INTI_FF_DL, false, 1, true>;
            case INTI_FF_COMB: return apply_launcher<PDE, DIM, INTI_FF_COMB, false, 1, true>;
        }
    }
    return nullptr;
}

template <int PDE, int DIM, bool MIXED, int NR>
static ApplyFn apply_kind(int kind) {
    switch (kind) {
        case INTI_SL: return apply_launcher<PDE, DIM, INTI_SL, MIXED, NR>;
        case INTI_DL: return apply_launcher<PDE, DIM, INTI_DL, MIXED, NR>;
        case INTI_ADL: return apply_launcher<PDE, DIM, INTI_ADL, MIXED, NR>;
        case INTI_HS: return apply_launcher<PDE, DIM, INTI_HS, MIXED, NR>;
        case INTI_COMB_SD: return apply_launcher<PDE, DIM, INTI_COMB_SD, MIXED, NR>;
        case INTI_COMB_AH: return apply_launcher<PDE, DIM, INTI_COMB_AH, MIXED, NR>;
    }
    if constexpr (PDE == INTI_HELMHOLTZ) {
        switch (kind) {
            case INTI_FF_SL: return apply_launcher<PDE, DIM, INTI_FF_SL, MIXED, NR>;
            case INTI_FF_DL: return apply_launcher<PDE, DIM, INTI_FF_DL, MIXED, NR>;
            case INTI_FF_COMB: return apply_launcher<PDE, DIM, INTI_FF_COMB, MIXED, NR>;
        }
    }
    return nullptr;
}

// NR in {1, 4, 16}
template <int PDE, int DIM>
static ApplyFn apply_pick_pd(int kind, int mixed, int nr) {
    if (mixed) {
        if (nr == 1) return apply_kind<PDE, DIM, true, 1>(kind);
        if (nr == 4) return apply_kind<PDE, DIM, true, 4>(kind);
        if (nr == 16) return apply_kind<PDE, DIM, true, 16>(kind);
    } else {
        if (nr == 1) return apply_kind<PDE, DIM, false, 1>(kind);
        if (nr == 4) return apply_kind<PDE, DIM, false, 4>(kind);
        if (nr == 16) return apply_kind<PDE, DIM, false, 16>(kind);
    }
    return nullptr;
}

using BatchFn = void (*)(const BatchArgs &, dim3, hipStream_t);

template <int PDE, int DIM, int KIND, bool MIXED, int NR>
static void batch_launcher(const BatchArgs &a, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL((k_apply_batch<PDE, DIM, KIND, MIXED, NR>), grid, dim3(apply_threads<NR>()), 0, s, a);
}

// batched kinds: SL, DL and c0 SL + c1 DL (the second-kind equations)
template <int PDE, int DIM, bool MIXED, int NR>
static BatchFn batch_kind(int kind) {
    switch (kind) {
        case INTI_SL: return batch_launcher<PDE, DIM, INTI_SL, MIXED, NR>;
        case INTI_DL: return batch_launcher<PDE, DIM, INTI_DL, MIXED, NR>;
        case INTI_COMB_SD: return batch_launcher<PDE, DIM, INTI_COMB_SD, MIXED, NR>;
    }
    return nullptr;
}

template <int PDE, int DIM>
static BatchFn batch_pick_pd(int kind, int mixed, int nr) {
    if (mixed) {
        if (nr == 1) return batch_kind<PDE, DIM, true, 1>(kind);
        if (nr == 4) return batch_kind<PDE, DIM, true, 4>(kind);
        if (nr == 16) return batch_kind<PDE, DIM, true, 16>(kind);
    } else {
        if (nr == 1) return batch_kind<PDE, DIM, false, 1>(kind);
        if (nr == 4) return batch_kind<PDE, DIM, false, 4>(kind);
        if (nr == 16) return batch_kind<PDE, DIM, false, 16>(kind);
    }
    return nullptr;
}

BatchFn batch_pick_lap2(int kind, int mixed, int nr);
BatchFn batch_pick_lap3(int kind, int mixed, int nr);
BatchFn batch_pick_helm2(int kind, int mixed, int nr);
BatchFn batch_pick_helm3(int kind, int mixed, int nr);

ApplyFn apply_trans_pick_lap2(int kind);
ApplyFn apply_trans_pick_lap3(int kind);
ApplyFn apply_trans_pick_helm2(int kind);
ApplyFn apply_trans_pick_helm3(int kind);

ApplyFn apply_pick_lap2(int kind, int mixed, int nr);
ApplyFn apply_pick_lap3(int kind, int mixed, int nr);
ApplyFn apply_pick_helm2(int kind, int mixed, int nr);
ApplyFn apply_pick_helm3(int kind, int mixed, int nr);
ApplyFn apply_pick_helm3c(int kind, int nr);  // complex wavenumber, FP64

}  // namespace inti
