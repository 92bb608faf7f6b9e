"""Benchmark: pair-evaluations per second of the DIM-corrected combined-field
matvec (DL - i k SL) on one GPU, mixed precision vs FP64.

Workload: 3D Helmholtz, k = 2*pi, unit sphere as a geodesic icosphere of
frequency 23 (10,580 P2 triangles, 6-point Vioreanu-Rokhlin rule -> 63,480
nodes), complex Gaussian density from a fixed seed. The smallest icosphere
with at least the 10,240 triangles / 61,440 nodes of the stated workload.

One step = one corrected matvec y = (D - i k S + dD - i k dS) x with x, y and
the geometry resident in HBM. Printed: ONE JSON line (rank 0).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpair-evals/s of corrected DL matvec, mixed vs FP64 GPU, at rel err ≤1e-6"
# canonical per-pair counts of the combined field kernel (36 flops, 3 SFU:
# rsqrt, sin, cos), per-source / per-target factors hoisted
FLOPS_PER_PAIR = 36
PEAK_VALU_TFLOPS = 78.6  # MI355X FP64 vector (= FP32 without packed math)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


SEED = 20261202


def make_workload(inti, freq=23):
    """Unit sphere, geodesic icosphere of frequency freq (P2 triangles x
    VR4N6), complex Gaussian density: the combined-field workload."""
    Q = inti.sphere_quadrature(freq, qorder=4, base="ico")
    rng = np.random.default_rng(SEED)
    n = len(Q)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)) * np.sqrt(0.5)
    return Q, x


def circle_workload(inti):
    """Unit circle, 128 P2 elements x 5 Gauss nodes (N = 640), density
    cos(3 theta) + 0.1 N(0, 1)."""
    Q = inti.circle_quadrature(128, qorder=9)
    rng = np.random.default_rng(SEED)
    th = np.arctan2(Q.coords[1], Q.coords[0])
    sig = np.cos(3 * th) + 0.1 * rng.normal(size=len(Q))
    return Q, sig


def batch_workload(inti, P=256):
    """P ellipses a ~ U[1,3], b ~ U[0.3,1], k ~ U[1,40], U{50..2000} P2
    elements x 6 Gauss nodes. Returns (quadratures, a, b, k, rng)."""
    rng = np.random.default_rng(SEED)
    a = rng.uniform(1, 3, P)
    b = rng.uniform(0.3, 1, P)
    ks = rng.uniform(1, 40, P)
    ne = rng.integers(50, 2001, P)
    qs = [inti.ellipse_quadrature(a[p], b[p], int(ne[p]), qorder=11) for p in range(P)]
    return qs, a, b, ks, rng


def bumpy_quadrature(inti, freq=64):
    """Bumpy sphere r = 1 + 0.15 sin(5 theta) cos(4 phi), icosphere frequency
    64: 81,920 P2 triangles x VR4N6 = 491,520 nodes."""
    return inti.sphere_quadrature(freq, radial=lambda th, ph: 1.0 + 0.15 * np.sin(5 * th) * np.cos(4 * ph))


def shape_quadrature(inti, freq=29):
    """Sphere r = 1 + 0.1 sin(3 theta) cos(2 phi), icosphere frequency 29:
    16,820 P2 triangles x VR4N6 = 100,920 nodes."""
    return inti.sphere_quadrature(freq, radial=lambda th, ph: 1.0 + 0.1 * np.sin(3 * th) * np.cos(2 * ph))


def fibonacci_directions(nd=64):
    """nd unit directions on a Fibonacci lattice of the sphere (3 x nd)."""
    i = np.arange(nd) + 0.5
    pol = np.arccos(1 - 2 * i / nd)
    az = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(az) * np.sin(pol), np.sin(az) * np.sin(pol), np.cos(pol)])


def time_apply(ctx, lib, op, dx, dy, n, steps, warmup):
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)
    for _ in range(warmup):
        ctx.check(lib.inti_op_apply_device(op, 1, dx, n, dy, n, one, zero))
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.check(lib.inti_op_apply_device(op, 1, dx, n, dy, n, one, zero))
        kms.append(ctx.last_kernel_ms())  # HIP events on the context stream
    ctx.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, float(np.mean(kms))


def cpu_baseline(oracle, Q, x, k, target_seconds=15.0):
    """The oracle (FP64, OpenMP over all host threads) on a bounded row
    sample of the same direct sum; pair-evals/s."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = len(Q)
    rows = 16
    coefs = (-1j * k, 1.0)
    t = 0.0
    # grow a calibration sample to about a second, then run once at the size
    # that takes about target_seconds
    while True:
        t0 = time.perf_counter()
        oracle.matvec(1, 3, 4, k, Q, Q, x, coefs=coefs, rows=(0, rows), nthreads=threads)
        t = time.perf_counter() - t0
        if t > 1.0 or rows >= n:
            break
        rows = min(n, rows * 4)
    if rows < n:
        rows = min(n, int(rows * target_seconds / max(t, 1e-3)))
        t0 = time.perf_counter()
        oracle.matvec(1, 3, 4, k, Q, Q, x, coefs=coefs, rows=(0, rows), nthreads=threads)
        t = time.perf_counter() - t0
    rate = rows * n / t
    # one thread on a sample of about 4 s
    rows1 = max(1, min(n, int(rows * 4.0 / max(t, 1e-3) / max(threads, 1))))
    t0 = time.perf_counter()
    oracle.matvec(1, 3, 4, k, Q, Q, x, coefs=coefs, rows=(0, rows1), nthreads=1)
    t1 = time.perf_counter() - t0
    return {"value": rate / 1e9, "unit": "Gpair-evals/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 direct sum of the same combined-field operator, rows 0..{rows} "
                      f"of {n} against all {n} sources ({rows * n:.3g} pairs, {t:.1f} s)",
            "cpu_model": _cpu_model(), "host_cpus_visible": os.cpu_count(),
            "single_thread": {"value": rows1 * n / t1 / 1e9, "unit": "Gpair-evals/s",
                              "sample": f"rows 0..{rows1} ({rows1 * n:.3g} pairs, {t1:.1f} s)"}}


def _cpu_model():
    """The host CPU's model name (the `Model name` line of lscpu)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def sym_group(n, value_bytes):
    """Tiles of 256 nodes per work group of the symmetric sum, by the
    library's rule: the smallest of 1, 2, 4 whose ceil(T / G) x n partial
    sums (T = ceil(n / 256) tiles) fit 4 GiB."""
    T = -(-n // 256)
    for g in (1, 2, 4):
        if -(-T // g) * n * value_bytes <= 4 << 30:
            return g
    return None


MICRO_OPS = ["v_fma_f64", "v_fma_f32", "v_rsq_f32", "v_sin_f32", "v_cos_f32", "v_log_f32", "v_rcp_f32", "v_rsq_f64"]


def measured_roofs(ctx, lib, flops_per_pair, sfu_per_pair):
    """Issue rates of the FMA and transcendental instructions over the whole
    GPU (inti_microbench: lane-operations per second) and the two pair-rate
    roofs of a kernel with the given canonical FP64 flops and FP32
    transcendentals per pair; the smaller one binds."""
    rates = {}
    for op, name in enumerate(MICRO_OPS):
        r = C.c_double()
        ctx.check(lib.inti_microbench(ctx.handle, op, C.byref(r)))
        rates[name] = r.value
    valu_flops = 2.0 * rates["v_fma_f64"]
    sfu = min(rates["v_rsq_f32"], rates["v_sin_f32"], rates["v_cos_f32"])
    valu_roof = valu_flops / flops_per_pair
    sfu_roof = sfu / sfu_per_pair
    return {"lane_ops_per_s": {k: round(v / 1e12, 3) for k, v in rates.items()}, "unit": "T lane-ops/s",
            "fp64_fma_TFLOPs": round(valu_flops / 1e12, 2),
            "valu_roof_Gpair_s": round(valu_roof / 1e9, 1), "sfu_roof_Gpair_s": round(sfu_roof / 1e9, 1),
            "binding": "valu" if valu_roof <= sfu_roof else "sfu",
            "per_pair": {"flops": flops_per_pair, "f32_transcendentals": sfu_per_pair}}


def calibrate_copy(ctx, lib, nbytes):
    """A plain copy of known size (reads nbytes, writes nbytes) to scale the
    profiler's FETCH_SIZE / WRITE_SIZE counters."""
    a, b = C.c_void_p(), C.c_void_p()
    ctx.check(lib.inti_device_alloc(ctx.handle, nbytes, C.byref(a)))
    ctx.check(lib.inti_device_alloc(ctx.handle, nbytes, C.byref(b)))
    for _ in range(5):
        ctx.check(lib.inti_device_copy(ctx.handle, b, a, nbytes))
    log(f"bench: copied {nbytes} bytes 5 times")
    ctx.check(lib.inti_device_free(ctx.handle, a))
    ctx.check(lib.inti_device_free(ctx.handle, b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--freq", type=int, default=23)
    ap.add_argument("--workload", choices=["cfie", "batch", "bumpy", "shape", "circle", "dense"], default="cfie",
                    help="cfie (default): the headline corrected combined-field matvec; batch: 256 2D "
                         "Helmholtz problems x 16 right-hand sides; bumpy: 3D Laplace single layer on a "
                         "~491k-node bumpy sphere with refined GMRES; shape: far-field objective and its "
                         "adjoint on a perturbed sphere; circle: 2D Laplace DL on 640 nodes; dense: the "
                         "assembled corrected operator on many right-hand sides (FP64 matrix cores)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", choices=["mixed", "fp64", "gmres"], default=None,
                    help="profiling runs: time one leg only, no checks, no JSON (gmres: the batch "
                         "workload's hipGraph-captured batched GMRES, --steps Arnoldi steps)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="profiler calibration: run a plain device copy of this many bytes 5 times")
    ap.add_argument("--nprob", type=int, default=256, help="batch workload: number of problems")
    ap.add_argument("--sym", type=int, choices=[0, 1, 2], default=None,
                    help="symmetric evaluation of operators on one quadrature: 0 off, 1 automatic "
                         "(default), 2 wherever it applies")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(the direct sum would shard by target rows, the symmetric sum by blocks of tile pairs).",
              file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    inti = load_package()
    if args.sym is not None:
        inti.set_default_symmetric(args.sym)
    lib = inti._lib.load()
    ctx = inti.default_context()
    if args.copy_bytes:
        calibrate_copy(ctx, lib, args.copy_bytes)
        return
    ctx.set_profiling(True)
    run = {"cfie": run_cfie, "batch": run_batch, "bumpy": run_bumpy, "shape": run_shape,
           "circle": run_circle, "dense": run_dense}[args.workload]
    out = run(inti, lib, ctx, args)
    if out is not None:
        print(json.dumps(out), flush=True)


def run_cfie(inti, lib, ctx, args):

    k = 2 * np.pi
    Q, x = make_workload(inti, args.freq)
    n = len(Q)
    log(f"bench: nodes={n} elements={Q.ne}")
    op = inti.Helmholtz(k=k, dim=3)
    t0 = time.perf_counter()
    L64, corr = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0)
    ctx.synchronize()
    t_corr = time.perf_counter() - t0
    Lmx, _ = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0,
                                 precision="mixed", corrections=corr)
    log(f"bench: DIM correction built in {t_corr:.2f} s (nnz={corr[1].nnz})")

    nbytes = 16 * n
    dx, dy = C.c_void_p(), C.c_void_p()
    ctx.check(lib.inti_device_alloc(ctx.handle, nbytes, C.byref(dx)))
    ctx.check(lib.inti_device_alloc(ctx.handle, nbytes, C.byref(dy)))
    xc = np.ascontiguousarray(x, dtype=np.complex128)
    ctx.check(lib.inti_memcpy_h2d(ctx.handle, dx, xc.ctypes.data_as(C.c_void_p), nbytes))

    pairs = float(n) * n
    if args.only:
        L = Lmx if args.only == "mixed" else L64
        t, kms = time_apply(ctx, lib, L.iop.handle, dx, dy, n, args.steps, args.warmup)
        log(f"bench: {args.only} only: {t * 1e3:.3f} ms/apply, kernel {kms:.3f} ms, "
            f"{pairs / t / 1e9:.1f} Gpair/s")
        return
    # FP64 leg (reference for the ratio and the error)
    t64, k64 = time_apply(ctx, lib, L64.iop.handle, dx, dy, n, max(3, args.steps // 2), 1)
    y64 = np.zeros(n, dtype=np.complex128)
    ctx.check(lib.inti_memcpy_d2h(ctx.handle, y64.ctypes.data_as(C.c_void_p), dy, nbytes))
    # mixed leg: the timed region of the metric
    tmx, kmx = time_apply(ctx, lib, Lmx.iop.handle, dx, dy, n, args.steps, args.warmup)
    ymx = np.zeros(n, dtype=np.complex128)
    ctx.check(lib.inti_memcpy_d2h(ctx.handle, ymx.ctypes.data_as(C.c_void_p), dy, nbytes))
    err_mx = float(np.linalg.norm(ymx - y64) / np.linalg.norm(y64))
    err_mx_inf = float(np.abs(ymx - y64).max() / np.abs(y64).max())
    log(f"bench: fp64 {t64 * 1e3:.3f} ms/apply (kernel {k64:.3f} ms), "
        f"mixed {tmx * 1e3:.3f} ms/apply (kernel {kmx:.3f} ms), rel err {err_mx:.2e}")
    # the same matvec by the one-direction kernels (each pair evaluated once
    # per direction) in both precisions: the mixed/FP64 ratio of that algorithm
    t64p = tmxp = None
    if args.sym != 0:
        mode = args.sym if args.sym is not None else 1
        for L in (L64, Lmx):
            L.iop.set_symmetric(0)
        t64p, _ = time_apply(ctx, lib, L64.iop.handle, dx, dy, n, max(3, args.steps // 4), 1)
        tmxp, _ = time_apply(ctx, lib, Lmx.iop.handle, dx, dy, n, max(3, args.steps // 4), 1)
        for L in (L64, Lmx):
            L.iop.set_symmetric(mode)
        log(f"bench: one-direction kernels: fp64 {t64p * 1e3:.3f} ms/apply, mixed {tmxp * 1e3:.3f} ms/apply")

    cpu = None
    err64 = None
    if not args.no_cpu:
        # the CPU baseline leg: the oracle times a row sample of the same
        # operator and checks the FP64 GPU result on its first rows
        import oracle
        rows = (0, 48)
        ref = oracle.matvec(1, 3, 4, k, Q, Q, x, coefs=(-1j * k, 1.0), rows=rows)[:rows[1]]
        rp, ci, vals = corr[0].export()
        _, _, vd = corr[1].export()
        cv = vd - 1j * k * vals
        for i in range(rows[1]):
            ref[i] += np.dot(cv[rp[i]:rp[i + 1]], x[ci[rp[i]:rp[i + 1]]])
        err64 = float(np.linalg.norm(y64[:rows[1]] - ref) / np.linalg.norm(ref))
        log(f"bench: fp64 vs oracle (rows 0..{rows[1]}): {err64:.2e}")
        cpu = cpu_baseline(oracle, Q, x, k)
        log(f"bench: cpu baseline {cpu['value']:.4f} Gpair/s on {cpu['cores']} threads")

    roofs = measured_roofs(ctx, lib, FLOPS_PER_PAIR, 3)
    log(f"bench: measured roofs {roofs}")
    achieved = FLOPS_PER_PAIR * pairs / (kmx * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_cfie_mixed.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    value = pairs / tmx / 1e9 if err_mx <= 1e-6 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tmx * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 arithmetic + f32 transcendentals (mixed)",
        "data": "synthetic",
        "config": {
            "workload": f"3D Helmholtz combined field DL - i k SL, k=2pi, DIM-corrected, unit sphere "
                        f"geodesic icosphere freq {args.freq}: {Q.ne} P2 triangles x VR4N6 = {n} nodes, "
                        f"complex Gaussian density (seed 20261202)",
            "nodes": n, "elements": Q.ne, "k": k, "pairs_per_step": pairs,
            "dim_stencil": "each element's own nodes (bdim_correction's pattern when target is source)",
            "correction_nnz": corr[1].nnz, "precision": "mixed",
        },
        "fp64_value": round(pairs / t64 / 1e9, 3),
        # the same algorithm (the symmetric sum) in both precisions
        "mixed_over_fp64": round(t64 / tmx, 3),
        "one_direction_kernels": ({"mixed_value": round(pairs / tmxp / 1e9, 3),
                                   "fp64_value": round(pairs / t64p / 1e9, 3),
                                   "mixed_over_fp64": round(t64p / tmxp, 3)} if t64p else None),
        "rel_err_mixed_vs_fp64": err_mx,
        "relinf_err_mixed_vs_fp64": err_mx_inf,
        "rel_err_fp64_vs_oracle_rows": err64,
        "dim_build_s": round(t_corr, 3),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_VALU_TFLOPS, 4),
                     "traffic": traffic,
                     "kernel": ("k_apply<HELMHOLTZ,3,COMB_SD,MIXED>" if args.sym == 0 else
                                f"k_apply_sym<HELMHOLTZ,3,COMB_SD,MIXED,G={sym_group(n, 16)},TPL=2>"),
                     "kernel_ms": round(kmx, 4), "flops_per_pair": FLOPS_PER_PAIR,
                     "symmetric": args.sym != 0, "measured_roofs": roofs},
        "cpu_baseline": cpu,
    }
    ctx.check(lib.inti_device_free(ctx.handle, dx))
    ctx.check(lib.inti_device_free(ctx.handle, dy))
    return out




def _dev_buffers(ctx, lib, nbytes, host=None):
    dx, dy = C.c_void_p(), C.c_void_p()
    ctx.check(lib.inti_device_alloc(ctx.handle, nbytes, C.byref(dx)))
    ctx.check(lib.inti_device_alloc(ctx.handle, nbytes, C.byref(dy)))
    if host is not None:
        ctx.check(lib.inti_memcpy_h2d(ctx.handle, dx, host.ctypes.data_as(C.c_void_p), nbytes))
    return dx, dy


def _fetch(ctx, lib, dptr, shape, dtype):
    out = np.zeros(shape, dtype=dtype)
    ctx.check(lib.inti_memcpy_d2h(ctx.handle, out.ctypes.data_as(C.c_void_p), dptr, out.nbytes))
    return out


def run_batch(inti, lib, ctx, args):
    """256 2D Helmholtz double-layer problems on ellipses (a ~ U[1,3],
    b ~ U[0.3,1], k ~ U[1,40], U{50..2000} P2 elements x 6 Gauss nodes),
    16 complex Gaussian right-hand sides each, I/2 + D + dD per problem:
    one persistent batched apply (autotuned work items), mixed vs FP64, and
    the batched hipGraph GMRES to 1e-10."""
    P = args.nprob
    qs, a, b, ks, rng = batch_workload(inti, P)
    ops = [inti.Helmholtz(k=ks[p], dim=2) for p in range(P)]
    # the DIM corrections: one bdim_correction per problem (the reference's
    # loop), then all problems in one pass over the batch (used below)
    t0 = time.perf_counter()
    serial = []
    for p in range(P):
        Sop = inti.IntegralOperator(inti.SingleLayerKernel(ops[p]), qs[p])
        Dop = inti.IntegralOperator(inti.DoubleLayerKernel(ops[p]), qs[p])
        serial.append(inti.bdim_correction(ops[p], qs[p], qs[p], Sop, Dop,
                                           green_multiplier=np.full(len(qs[p]), -0.5)))
    ctx.synchronize()
    t_dim_serial = time.perf_counter() - t0
    Ks = [inti.DoubleLayerKernel(o) for o in ops]
    B64 = inti.BatchOperator(Ks, qs)
    B32 = inti.BatchOperator(Ks, qs, precision="mixed")
    t0 = time.perf_counter()
    built = B64.bdim_correction()
    ctx.synchronize()
    t_dim = time.perf_counter() - t0
    corr_s = [c[0] for c in built]
    corr = [c[1] for c in built]
    # the batched build against the per-problem builds: the same patterns; the
    # corrections applied to a smooth density agree to the rounding of Theta's
    # sums (different summation orders) carried through the least squares. The
    # raw weights may differ more: an element's least squares is nearly rank
    # deficient (monopoles at 1.5x the problem's half-diagonal, elements far
    # smaller), and rounding moves the weights along its near-null directions,
    # which a smooth density does not see.
    dim_pdiff, dim_wdiff = [], []
    for p in range(P):
        c = qs[p].coords
        xs = np.exp(np.sin(c[0]) + 0.5 * c[1]) * np.exp(0.3j * c[0])
        for got, ref in zip(built[p], serial[p]):
            rg, cg, vg = got.export()
            rr, cr, vr = ref.export()
            if not (np.array_equal(rg, rr) and np.array_equal(cg, cr)):
                raise SystemExit(f"bench batch: problem {p}: batched DIM pattern differs from its own build")
            rows = np.repeat(np.arange(len(rr) - 1), np.diff(rr))
            yg = np.bincount(rows, weights=(vg * xs[cg]).real, minlength=len(xs)) + \
                1j * np.bincount(rows, weights=(vg * xs[cg]).imag, minlength=len(xs))
            yr = np.bincount(rows, weights=(vr * xs[cr]).real, minlength=len(xs)) + \
                1j * np.bincount(rows, weights=(vr * xs[cr]).imag, minlength=len(xs))
            dim_pdiff.append(float(np.linalg.norm(yg - yr) / np.linalg.norm(yr)))
            dim_wdiff.append(float(np.linalg.norm(vg - vr) / np.linalg.norm(vr)))
    del serial
    dim_diff = max(dim_pdiff)
    for Bt in (B64, B32):
        Bt.set_identity(0.5)
        for p in range(P):
            Bt.attach(p, corr[p])
    ntot = B64.ntot
    sizes = np.diff(B64.offsets)
    R = 16
    pairs = float(np.sum(sizes.astype(np.float64) ** 2)) * R
    log(f"bench batch: {P} problems, ntot={ntot}, sum N_p^2 x 16 = {pairs:.3e}, DIM {t_dim:.2f} s for the batch "
        f"in one pass ({t_dim_serial:.2f} s problem by problem; on a smooth density the corrections differ by "
        f"{np.median(dim_pdiff):.1e} (median) to {dim_diff:.1e} relative, the raw weights by {np.median(dim_wdiff):.1e} (median) to "
        f"{np.max(dim_wdiff):.1e})")
    tune64 = B64.autotune(nrhs=R, reps=2)
    tune32 = B32.autotune(nrhs=R, reps=2)
    log(f"bench batch: autotune fp64 {tune64}, mixed {tune32}")
    X = (rng.normal(size=(R, ntot)) + 1j * rng.normal(size=(R, ntot))) * np.sqrt(0.5)
    X = np.ascontiguousarray(X)  # column-major ntot x R
    nbytes = X.nbytes
    dx, dy = _dev_buffers(ctx, lib, nbytes, X)
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)

    def timed(Bt, steps, warmup):
        for _ in range(warmup):
            ctx.check(lib.inti_batch_apply_device(Bt.handle, R, dx, ntot, dy, ntot, one, zero))
        ctx.synchronize()
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.check(lib.inti_batch_apply_device(Bt.handle, R, dx, ntot, dy, ntot, one, zero))
            kms.append(ctx.last_kernel_ms())
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps, float(np.mean(kms))

    steps = max(3, args.steps // 4)
    if args.only == "gmres":  # profiling run of the batched GMRES (graph capture and replay)
        t0 = time.perf_counter()
        _, info = B64.gmres(X.T, restart=100, maxiter=args.steps, rtol=1e-10)
        log(f"bench batch: gmres only: {info['iterations']} Arnoldi steps in {time.perf_counter() - t0:.2f} s")
        ctx.check(lib.inti_device_free(ctx.handle, dx))
        ctx.check(lib.inti_device_free(ctx.handle, dy))
        return None
    if args.only:  # profiling runs: one apply leg, no solver, no JSON
        t, kms = timed(B32 if args.only == "mixed" else B64, steps, 1)
        log(f"bench batch: {args.only} only: {t * 1e3:.2f} ms/apply, kernel {kms:.2f} ms")
        ctx.check(lib.inti_device_free(ctx.handle, dx))
        ctx.check(lib.inti_device_free(ctx.handle, dy))
        return None
    t64, k64 = timed(B64, steps, 1)
    y64 = _fetch(ctx, lib, dy, (R, ntot), np.complex128)
    tmx, kmx = timed(B32, steps, 1)
    ymx = _fetch(ctx, lib, dy, (R, ntot), np.complex128)
    err = float(np.linalg.norm(ymx - y64) / np.linalg.norm(y64))
    log(f"bench batch: fp64 {t64 * 1e3:.2f} ms, mixed {tmx * 1e3:.2f} ms per apply, rel err {err:.2e}")
    # FP64 batched apply vs the single-operator path on three problems
    Xh = X.T
    chk = []
    for p in (0, P // 2, P - 1):
        A = inti.IntegralOperator(Ks[p], qs[p])
        A.set_identity(0.5)
        A.attach(corr[p])
        ref = A.apply(Xh[B64.offsets[p]:B64.offsets[p + 1]])
        got = y64.T[B64.offsets[p]:B64.offsets[p + 1]]
        chk.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    # batched GMRES (FP64 operator), hipGraph-captured Arnoldi step
    t0 = time.perf_counter()
    Xs, info = B64.gmres(Xh, restart=100, maxiter=300, rtol=1e-10)
    t_gm = time.perf_counter() - t0
    rel = info["relres"]
    conv = rel.max(axis=1) <= 1e-10
    perim = np.array([q.weights.sum() for q in qs])
    ppw = sizes / (perim * ks / (2 * np.pi))  # nodes per wavelength
    log(f"bench batch: gmres {info['iterations']} steps in {t_gm:.2f} s, {int(conv.sum())}/{P} problems "
        f"converged to 1e-10; median relres {np.median(rel):.2e}; converged median ppw "
        f"{np.median(ppw[conv]) if conv.any() else 0:.1f}, unconverged median ppw "
        f"{np.median(ppw[~conv]) if (~conv).any() else 0:.1f}")
    # iterative refinement: FP64 residuals, mixed-precision batched GMRES corrections
    t0 = time.perf_counter()
    Xr, rinfo = B64.refine_solve(B32, Xh, restart=100, maxiter_inner=100, inner_rtol=1e-6, max_outer=3,
                                 rtol=1e-10)
    t_rf = time.perf_counter() - t0
    rconv = rinfo["relres"].max(axis=1) <= 1e-10
    both = conv & rconv
    # two solutions with residuals below 1e-10 differ by up to cond(A_p) x 1e-10;
    # near-resonant draws are ill-conditioned, so report the spread per problem
    diffs = np.array([float(np.linalg.norm(B64.split(Xr)[p] - B64.split(Xs)[p]) / np.linalg.norm(B64.split(Xs)[p]))
                      for p in np.flatnonzero(both)]) if both.any() else np.zeros(1)
    agree = float(diffs.max())
    log(f"bench batch: refined solve {rinfo['outer']} outer / {rinfo['inner_iterations']} inner steps in "
        f"{t_rf:.2f} s, {int(rconv.sum())}/{P} problems converged to 1e-10; vs the FP64 GMRES solutions "
        f"(problems both converged): median {np.median(diffs):.2e}, max {agree:.2e}, "
        f"{int((diffs <= 1e-10).sum())}/{len(diffs)} within 1e-10")
    # resonance check: I/2 + D alone (the exterior Dirichlet problem with a
    # double-layer density) is singular at the interior Neumann eigenvalues
    # of each ellipse; the combined field I/2 + D - i k S has no such k. The
    # same draws, right-hand sides and GMRES(100) with this operator tell
    # resonant draws from solver trouble.
    Kc = [inti.CombinedFieldKernel(o, -1j * o.k, 1.0) for o in ops]
    Bc = inti.BatchOperator(Kc, qs)
    Bc.set_identity(0.5)
    for p in range(P):
        Bc.attach(p, corr_s[p], -1j * ops[p].k)
        Bc.attach(p, corr[p], 1.0)
    Bc.autotune(nrhs=R, reps=1)
    t0 = time.perf_counter()
    _, cinfo = Bc.gmres(Xh, restart=100, maxiter=600, rtol=1e-10)
    t_cf = time.perf_counter() - t0
    cconv = cinfo["relres"].max(axis=1) <= 1e-10
    ka = ks * a  # wavenumber x semi-major axis
    log(f"bench batch: combined field I/2 + D - i k S: gmres {cinfo['iterations']} steps in {t_cf:.2f} s, "
        f"{int(cconv.sum())}/{P} problems converged to 1e-10 (of the {int((~conv).sum())} that I/2 + D left "
        f"unconverged: {int((cconv & ~conv).sum())}); median relres {np.median(cinfo['relres']):.2e}; "
        f"median k a: converged {np.median(ka[cconv]) if cconv.any() else 0:.1f}, unconverged "
        f"{np.median(ka[~cconv]) if (~cconv).any() else 0:.1f}")
    ctx.check(lib.inti_device_free(ctx.handle, dx))
    ctx.check(lib.inti_device_free(ctx.handle, dy))
    cpu = None
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        p = int(np.argmax(sizes))
        Q = qs[p]
        rows = len(Q)  # the whole largest problem, one right-hand side (a few seconds)
        t1 = time.perf_counter()
        oracle.matvec(1, 2, 1, ks[p], Q, Q, Xh[B64.offsets[p]:B64.offsets[p + 1], 0], rows=(0, rows),
                      nthreads=threads)
        tc = time.perf_counter() - t1
        cpu = {"value": rows * len(Q) / tc / 1e9, "unit": "Gpair-evals/s", "cores": threads, "kind": "port",
               "sample": f"oracle FP64 2D Helmholtz DL direct sum, rows 0..{rows} of problem {p} "
                         f"(N={len(Q)}), one right-hand side ({tc:.2f} s)"}
    return {
        "metric": "Gpair-rhs-evals/s of the batched corrected DL apply (256 2D Helmholtz problems x 16 RHS), mixed",
        "value": round(pairs / tmx / 1e9, 3), "unit": "Gpair-rhs-evals/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": round(tmx * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 arithmetic + f32 transcendentals (mixed)", "data": "synthetic",
        "config": {"workload": f"{P} 2D Helmholtz DL problems on ellipses a~U[1,3], b~U[0.3,1], k~U[1,40], "
                               f"U{{50..2000}} P2 elements x 6 Gauss nodes, 16 complex Gaussian RHS, "
                               f"I/2 + D + dD (seed 20261202)",
                   "problems": P, "ntot": int(ntot), "pair_rhs_per_step": pairs},
        "fp64_value": round(pairs / t64 / 1e9, 3), "mixed_over_fp64": round(t64 / tmx, 3),
        "rel_err_mixed_vs_fp64": err, "rel_err_batch_vs_single_fp64": max(chk),
        "kernel_ms": {"fp64": round(k64, 3), "mixed": round(kmx, 3)},
        "autotune": {"fp64": tune64, "mixed": tune32}, "dim_build_s": round(t_dim, 3), "dim_build_serial_s": round(t_dim_serial, 3),
        "dim_batch_vs_serial": {"smooth_density_rel_diff": {"median": float(np.median(dim_pdiff)), "max": dim_diff},
                                "weights_rel_diff": {"median": float(np.median(dim_wdiff)),
                                                     "max": float(np.max(dim_wdiff))}},
        "gmres": {"steps": info["iterations"], "seconds": round(t_gm, 3), "rtol": 1e-10, "restart": 100,
                  "graph": True, "ms_per_step": round(t_gm / max(info["iterations"], 1) * 1e3, 2),
                  "problems_converged": int(conv.sum()), "median_relres": float(np.median(rel)),
                  "max_relres": float(rel.max()),
                  "nodes_per_wavelength": {"min": float(ppw.min()), "median": float(np.median(ppw)),
                                           "max": float(ppw.max())}},
        "refined_solve": {"seconds": round(t_rf, 3), "outer": rinfo["outer"],
                          "inner_iterations": rinfo["inner_iterations"], "inner": "mixed batched GMRES(100)",
                          "problems_converged": int(rconv.sum()),
                          "rel_diff_vs_fp64_gmres_solution": {"median": float(np.median(diffs)), "max": agree,
                                                              "within_1e-10": int((diffs <= 1e-10).sum()),
                                                              "problems": int(len(diffs))}},
        "combined_field_gmres": {"operator": "I/2 + D - i k S (no interior resonances)", "steps": cinfo["iterations"],
                                 "seconds": round(t_cf, 3), "problems_converged": int(cconv.sum()),
                                 "of_those_unconverged_with_dl": int((cconv & ~conv).sum()),
                                 "median_relres": float(np.median(cinfo["relres"])),
                                 "median_k_times_a": {"converged": float(np.median(ka[cconv])) if cconv.any() else None,
                                                      "unconverged": float(np.median(ka[~cconv])) if (~cconv).any()
                                                      else None}},
        "cpu_baseline": cpu,
    }


def _real_harmonics(theta, phi, lmax, rng):
    """sum over l <= lmax, |m| <= l of c_lm Y_lm (real form), c ~ N(0, 1)."""
    from scipy import special
    out = np.zeros_like(theta)
    for l in range(lmax + 1):
        for m in range(0, l + 1):
            Y = special.sph_harm_y(l, m, theta, phi)
            out += rng.normal() * Y.real
            if m > 0:
                out += rng.normal() * Y.imag
    return out


def calderon_preconditioner(inti, Q, alpha=1.0, precision="mixed"):
    """Right preconditioner of the first-kind 3D Laplace single-layer
    equation: the DIM-corrected hypersingular operator with a shift,
    P = HS - alpha I. HS = -W with W the positive semi-definite hypersingular
    operator; V W = I/4 - K^2 (Calderon) is well conditioned, and the shift
    removes W's null space (the constants)."""
    op = inti.Laplace(dim=3)
    n = len(Q)
    A64 = inti.IntegralOperator(inti.AdjointDoubleLayerKernel(op), Q)
    H64 = inti.IntegralOperator(inti.HyperSingularKernel(op), Q)
    _, dH = inti.bdim_correction(op, Q, Q, A64, H64, green_multiplier=np.full(n, -0.5), derivative=True)
    P = inti.IntegralOperator(inti.HyperSingularKernel(op), Q, precision=precision)
    P.attach(dH)
    P.set_identity(-alpha)
    return P


def _sl_setup(inti, ctx, Q):
    op = inti.Laplace(dim=3)
    n = len(Q)
    t0 = time.perf_counter()
    S64 = inti.IntegralOperator(inti.SingleLayerKernel(op), Q)
    D64 = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)
    dS, _ = inti.bdim_correction(op, Q, Q, S64, D64, green_multiplier=np.full(n, -0.5))
    ctx.synchronize()
    t_dim = time.perf_counter() - t0
    S64.attach(dS)
    S32 = inti.IntegralOperator(inti.SingleLayerKernel(op), Q, precision="mixed")
    S32.attach(dS)
    return S64, S32, t_dim


def _refined(inti, S64, S32, rhs, outer_max, inner_max, tag):
    """Refinement (FP64 residuals, mixed GMRES corrections) to 1e-12, one outer
    step per call (x carried over) so that progress is logged; then the
    all-FP64 GMRES to 1e-13 in slices. Returns (x_ref, info_ref, t_ref, x64,
    info64, t64)."""
    n = len(rhs)
    t1 = time.perf_counter()
    xr = np.zeros(n)
    ir = {"outer": 0, "inner_iterations": 0, "relres": 1.0}
    for _ in range(outer_max):
        xr, step = inti.refine_solve(S64, S32, rhs, x0=xr, restart=300, maxiter_inner=inner_max,
                                     inner_rtol=1e-6, max_outer=1, rtol=1e-12)
        ir = {"outer": ir["outer"] + step["outer"],
              "inner_iterations": ir["inner_iterations"] + step["inner_iterations"], "relres": step["relres"]}
        log(f"bench {tag}: refinement {ir} at {time.perf_counter() - t1:.1f} s")
        if step["relres"] <= 1e-12 or step["outer"] == 0:
            break
    t_ref = time.perf_counter() - t1
    t1 = time.perf_counter()
    x64 = np.zeros(n)
    i64 = {"iterations": 0, "relres": 1.0}
    for _ in range(max(1, outer_max)):
        # the reference solve goes to 1e-13 so that the comparison of solutions
        # measures the refined solve, not the reference's own stopping error
        x64, step = inti.gmres(S64, rhs, x0=x64, restart=300, maxiter=inner_max, rtol=1e-13)
        i64 = {"iterations": i64["iterations"] + step["iterations"], "relres": step["relres"]}
        log(f"bench {tag}: fp64 gmres {i64} at {time.perf_counter() - t1:.1f} s")
        if step["relres"] <= 1e-13 or step["iterations"] == 0:
            break
    return xr, ir, t_ref, x64, i64, time.perf_counter() - t1


def run_bumpy(inti, lib, ctx, args):
    """3D Laplace single layer, DIM-corrected, icosphere frequency 64 (81,920
    P2 triangles x VR4N6 = 491,520 nodes): on the bumpy sphere r = 1 + 0.15
    sin(5 theta) cos(4 phi) the mixed vs FP64 apply (the metric) and a
    bounded refined solve; on the unit sphere of the same size the refined
    GMRES to 1e-10 against the all-FP64 GMRES solution. The right-hand side
    is sum_(l<=8) c_lm Y_lm, c ~ N(0, 1)."""
    freq = 64
    Q = bumpy_quadrature(inti, freq)
    n = len(Q)
    S64, S32, t_dim = _sl_setup(inti, ctx, Q)
    rng = np.random.default_rng(20261202)
    th = np.arccos(np.clip(Q.coords[2] / np.linalg.norm(Q.coords, axis=0), -1, 1))
    ph = np.arctan2(Q.coords[1], Q.coords[0])
    rhs = _real_harmonics(th, ph, 8, rng)
    log(f"bench bumpy: nodes={n}, DIM {t_dim:.2f} s")
    nbytes = 8 * n
    dx, dy = _dev_buffers(ctx, lib, nbytes, np.ascontiguousarray(rhs))
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)
    res = {}
    for name, A in (("fp64", S64), ("mixed", S32)):
        ctx.check(lib.inti_op_apply_device(A.handle, 1, dx, n, dy, n, one, zero))
        ctx.synchronize()
        t1 = time.perf_counter()
        steps = 5
        kms = []
        for _ in range(steps):
            ctx.check(lib.inti_op_apply_device(A.handle, 1, dx, n, dy, n, one, zero))
            kms.append(ctx.last_kernel_ms())
        ctx.synchronize()
        res[name] = ((time.perf_counter() - t1) / steps, float(np.mean(kms)), _fetch(ctx, lib, dy, (n,), np.float64))
    ctx.check(lib.inti_device_free(ctx.handle, dx))
    ctx.check(lib.inti_device_free(ctx.handle, dy))
    pairs = float(n) * n
    err = float(np.linalg.norm(res["mixed"][2] - res["fp64"][2]) / np.linalg.norm(res["fp64"][2]))
    log(f"bench bumpy: fp64 {res['fp64'][0] * 1e3:.1f} ms, mixed {res['mixed'][0] * 1e3:.1f} ms, err {err:.2e}")
    if args.only:  # apply legs only (profiling and launch comparisons), no solves, no JSON
        return None
    # refined solve on the bumpy sphere toward a 1e-10 FP64 residual: FP64
    # residuals, mixed GMRES(300) corrections right-preconditioned by the
    # DIM-corrected hypersingular operator with a shift (Calderon, see
    # calderon_preconditioner). At this size both this and the
    # unpreconditioned GMRES(1000) stagnate near 1e-6 (the corrected
    # first-kind operator has small and negative eigenvalues); at 12,000
    # nodes both reach 1e-10. One outer step per call so that progress is
    # logged.
    t1 = time.perf_counter()
    P32 = calderon_preconditioner(inti, Q)
    t_prec = time.perf_counter() - t1
    log(f"bench bumpy: Calderon preconditioner built in {t_prec:.2f} s")
    xb = np.zeros(n)
    ib = {"outer": 0, "inner_iterations": 0, "relres": 1.0, "restart": 300, "rtol": 1e-10,
          "preconditioner": "corrected hypersingular - I (mixed)", "preconditioner_build_s": round(t_prec, 2)}
    for _ in range(4):  # bounded: each outer step is 300 preconditioned steps (~65 s)
        xb, step = inti.refine_solve(S64, S32, rhs, x0=xb, restart=300, maxiter_inner=300, inner_rtol=1e-6,
                                     max_outer=1, rtol=1e-10, precond=P32)
        ib.update(outer=ib["outer"] + step["outer"], inner_iterations=ib["inner_iterations"] + step["inner_iterations"],
                  relres=step["relres"])
        log(f"bench bumpy: refinement {ib} at {time.perf_counter() - t1:.1f} s")
        if step["relres"] <= 1e-10 or step["outer"] == 0:
            break
    del P32
    tb = time.perf_counter() - t1
    del S64, S32
    # the full refinement on the unit sphere of the same size
    Qs = inti.sphere_quadrature(freq)
    T64, T32, t_dim_s = _sl_setup(inti, ctx, Qs)
    ths = np.arccos(np.clip(Qs.coords[2], -1, 1))
    phs = np.arctan2(Qs.coords[1], Qs.coords[0])
    rhs_s = _real_harmonics(ths, phs, 8, np.random.default_rng(20261202))
    xr, ir, t_ref, x64, i64, t_64 = _refined(inti, T64, T32, rhs_s, 6, 600, "sphere")
    sol_err = float(np.linalg.norm(xr - x64) / np.linalg.norm(x64))
    log(f"bench sphere: refined vs fp64 solution {sol_err:.2e}")
    return {
        "metric": "Gpair-evals/s of the corrected 3D Laplace SL matvec (491,520-node bumpy sphere), mixed",
        "value": round(pairs / res["mixed"][0] / 1e9, 3), "unit": "Gpair-evals/s", "n_gpus": 1, "steps": 5,
        "warmup": 1, "ms_per_step": round(res["mixed"][0] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 arithmetic + f32 rsqrt (mixed)",
        "data": "synthetic",
        "config": {"workload": f"3D Laplace SL, bumpy sphere r = 1 + 0.15 sin(5 theta) cos(4 phi), icosphere "
                               f"freq {freq}: {Q.ne} P2 triangles x VR4N6 = {n} nodes, RHS sum_(l<=8) c_lm Y_lm "
                               f"(seed 20261202), DIM-corrected", "nodes": n},
        "fp64_value": round(pairs / res["fp64"][0] / 1e9, 3),
        "mixed_over_fp64": round(res["fp64"][0] / res["mixed"][0], 3),
        "rel_err_mixed_vs_fp64": err, "kernel_ms": {k: round(v[1], 3) for k, v in res.items()},
        "dim_build_s": round(t_dim, 2),
        "bumpy_refined_solve": {"seconds": round(tb, 2), **ib},
        "unit_sphere_491520": {"dim_build_s": round(t_dim_s, 2),
                               "refined_solve": {"seconds": round(t_ref, 2), **ir, "rtol": 1e-12},
                               "fp64_gmres": {"seconds": round(t_64, 2), **i64},
                               "rel_err_refined_vs_fp64_solution": sol_err},
    }


def run_shape(inti, lib, ctx, args):
    """Shape gradient of sound-soft scattering: e^{ikz} on the sphere
    r = 1 + 0.1 sin(3 theta) cos(2 phi), k = 10 (icosphere frequency 29:
    100,920 nodes); sigma from the DIM-corrected CFIE by GMRES;
    J = sum over 64 Fibonacci directions of |u_inf|^2. The adjoint state
    (shape_gradient) gives dJ/d(nodes, normals, weights, k) through the
    solve; checked by central differences (h = 1e-6) of the whole solve along
    20 seeded random directions, with the DIM correction frozen (what the
    adjoint differentiates) and rebuilt (the full solve). The pullback of the
    N x N combined-field direct sum is timed against its forward."""
    k = 10.0
    Q = shape_quadrature(inti)
    n = len(Q)
    op = inti.Helmholtz(k=k, dim=3)
    dirs = fibonacci_directions(64)
    if args.only:  # profiling runs: a random density instead of the solve, kernels only, no JSON
        rng = np.random.default_rng(1)
        sig = rng.normal(size=n) + 1j * rng.normal(size=n)
        A = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0), Q)
        A.apply(sig)  # warm-up of both kernels
        A.pullback(sig, sig, wrt=("sigma", "nodes", "source_normals", "source_weights", "k"), dcoefs=(-1j, 0.0))
        tf, tc, ts = [], [], []
        for _ in range(3):
            A.apply(sig)
            tf.append(ctx.last_kernel_ms())
            A.pullback(sig, sig, wrt=("sigma", "nodes", "source_normals", "source_weights", "k"), dcoefs=(-1j, 0.0))
            tc.append(ctx.last_kernel_ms())
            A.pullback(sig, sig, wrt=("sigma", "source_nodes", "source_normals", "source_weights",
                                      "target_nodes", "k"), dcoefs=(-1j, 0.0))
            ts.append(ctx.last_kernel_ms())
        ctx.synchronize()
        f, c_, s_ = (float(np.median(t)) for t in (tf, tc, ts))
        log(f"bench shape: profiling legs done: forward {f:.3f} ms, combined pullback {c_:.3f} ms ({c_ / f:.3f}x), "
            f"separate outputs {s_:.3f} ms ({s_ / f:.3f}x)")
        return None
    t0 = time.perf_counter()
    J0, grad, info = shape_gradient(inti, Q.coords, Q.normals, Q.weights, Q.qtags, k, dirs)
    t_grad = time.perf_counter() - t0
    log(f"bench shape: nodes {n}, J {J0:.6e}, solve {info['solve']}, adjoint solve {info['adjoint_solve']}, "
        f"objective + gradient {t_grad:.2f} s")
    t0 = time.perf_counter()
    checks = shape_fd_check(inti, Q.coords, Q.normals, Q.weights, Q.qtags, k, dirs, grad, info["corrections"],
                            20, SEED, sigma0=info["sigma"], grad_frozen=info["grad_frozen"])
    t_fd = time.perf_counter() - t0
    # errors relative to the largest directional derivative of the same kind
    # (a random node direction can give a derivative far below the
    # gradient's size, the FD error does not shrink with it)
    scale_full, scale_frozen = {}, {}
    for w, _, fz, fl, _ in checks:
        scale_full[w] = max(scale_full.get(w, 0.0), abs(fl))
        scale_frozen[w] = max(scale_frozen.get(w, 0.0), abs(fz))
    e_full = [abs(a - fl) / scale_full[w] for w, a, _, fl, _ in checks]
    e_frozen = [abs(af - fz) / scale_frozen[w] for w, _, fz, _, af in checks]
    for (w, a, fz, fl, af), ez, el in zip(checks, e_frozen, e_full):
        log(f"bench shape: {w:8s} gradient {a:+.9e} fd(full) {fl:+.9e} ({el:.1e}); "
            f"frozen-DIM gradient {af:+.9e} fd(frozen) {fz:+.9e} ({ez:.1e})")
    # the pullback of the N x N combined-field direct sum vs its forward
    A = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0), Q)
    sig = info["sigma"] if "sigma" in info else np.exp(1j * k * Q.coords[0])
    ubar = np.conj(np.mean(sig)) * np.ones(n) + 0.0j
    nxn = {}
    legs = (("sigma_only", ("sigma",)),
            ("all_nodes_combined", ("sigma", "nodes", "source_normals", "source_weights", "k")),
            ("all_source_and_target_parts", ("sigma", "source_nodes", "source_normals", "source_weights",
                                             "target_nodes", "k")))
    for label, mode in (("symmetric", 1), ("plain", 0)):
        A.set_symmetric(mode)
        A.apply(sig)
        ctx.synchronize()
        fw = []
        for _ in range(3):
            A.apply(sig)
            fw.append(ctx.last_kernel_ms())
        t_fwd = float(np.median(fw))
        row = {"forward": round(t_fwd, 3)}
        for leg, wrt in legs:
            A.pullback(ubar, sig, wrt=wrt, dcoefs=(-1j, 0.0))
            ts = []
            for _ in range(3):
                A.pullback(ubar, sig, wrt=wrt, dcoefs=(-1j, 0.0))
                ts.append(ctx.last_kernel_ms())
            row[leg] = round(float(np.median(ts)), 3)
            row[leg + "_over_forward"] = round(float(np.median(ts)) / t_fwd, 3)
        log(f"bench shape: N x N {label} kernels (ms): {row}")
        nxn[label] = row
    sym = nxn["symmetric"]
    return {
        "metric": "adjoint / forward kernel time of the N x N combined-field pullback (all gradients: sigma, nodes, "
                  "normals, weights, k), shape gradient workload, 100,920 nodes",
        "value": sym["all_nodes_combined_over_forward"], "unit": "ratio", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": sym["all_nodes_combined"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"sound-soft 3D Helmholtz k=10, sphere r = 1 + 0.1 sin(3 theta) cos(2 phi), "
                               f"icosphere freq 29: {Q.ne} P2 triangles x VR4N6 = {n} nodes, plane wave +z, "
                               f"J = sum over 64 Fibonacci directions |u_inf|^2", "nodes": n},
        "objective": J0, "solve": info["solve"], "adjoint_solve": info["adjoint_solve"],
        "objective_and_gradient_s": round(t_grad, 2),
        "fd_checks": {"directions": 20, "h": 1e-6, "seconds": round(t_fd, 1),
                      "error_scale": "largest |directional derivative| of the same kind",
                      "worst_err_full_solve": max(e_full),
                      "worst_err_frozen_dim": max(e_frozen),
                      "by_direction": [{"wrt": w, "gradient": a, "fd_full": fl, "gradient_frozen_dim": af,
                                        "fd_frozen_dim": fz} for w, a, fz, fl, af in checks]},
        "direct_sum_NxN_kernel_ms": nxn,
    }


def shape_problem(inti, y, nv, w, qtags, k, dirs, frozen=None, rtol=1e-12, x0=None):
    """J = sum over the directions of |u_inf|^2 for the sound-soft scattering
    of e^{i k z} by the surface with quadrature (y, nv, w, qtags): sigma from
    the DIM-corrected combined field (I/2 + D - i k S) sigma = -e^{i k z},
    u_inf = F sigma with F the far field of D - i k S. frozen = (dS, dD)
    keeps those corrections instead of rebuilding them on this geometry; x0
    is GMRES's initial guess. Returns (J, u_inf, sigma, L, F, corrections,
    gmres info)."""
    Q = inti.Quadrature(y, nv, w, qtags)
    op = inti.Helmholtz(k=k, dim=3)
    if frozen is None:
        L, corr = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, c_identity=0.5)
    else:
        A = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0), Q)
        A.attach(frozen[0], -1j * k)
        A.attach(frozen[1], 1.0)
        A.set_identity(0.5)
        L, corr = A, frozen
    b = -np.exp(1j * k * y[2])
    sig, info = inti.gmres(L, b, x0=x0, restart=200, maxiter=600, rtol=rtol)
    F = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0, "FF"), inti.points(dirs),
                              inti.Quadrature(y, nv, w, np.arange(len(w), dtype=np.int64).reshape(1, -1)))
    u = F.apply(sig)
    return float(np.sum(np.abs(u) ** 2)), u, sig, L, F, corr, info


def shape_gradient(inti, y, nv, w, qtags, k, dirs, dim_derivative=True):
    """dJ/d(nodes, normals, weights, k) of shape_problem by the adjoint state:
    with L sigma = b, L^H lam = F^H ubar (ubar = 2 u_inf),
        dJ = Re<ubar, dF sigma> + Re<lam, db> - Re<lam, dL sigma>,
    dF sigma and dL sigma from the pullbacks of the far-field operator and of
    the combined-field direct sum (node gradient: source and target parts in
    one pass), db of b = -e^{i k z}, and the k-derivative of the correction's
    coefficient -i k. dim_derivative: the corrections' own dependence on the
    geometry and k (traces, Theta, the per-element least squares, the
    monopoles) by bdim_pullback with the cotangents -lam_i conj(c sigma_j)
    of their entries; without it the correction matrices are held fixed.
    Returns (J, grad dict, info)."""
    n = len(w)
    J, u, sig, L, F, (dS, dD), sinfo = shape_problem(inti, y, nv, w, qtags, k, dirs)
    ubar = 2.0 * u
    gF = F.pullback(ubar, sig, wrt=("sigma", "source_nodes", "source_normals", "source_weights", "k"),
                    dcoefs=(-1j, 0.0))
    lam, ainfo = inti.gmres_adjoint(L, gF["sigma"], restart=100, maxiter=600, rtol=1e-12)
    Q = inti.Quadrature(y, nv, w, qtags)
    op = inti.Helmholtz(k=k, dim=3)
    P = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0), Q)
    gL = P.pullback(lam, sig, wrt=("nodes", "source_normals", "source_weights", "k"), dcoefs=(-1j, 0.0))
    e = np.exp(1j * k * y[2])
    lc = np.conj(lam)
    gnodes = gF["source_nodes"] - gL["nodes"]
    gnodes[2] += np.real(lc * (-1j * k * e))
    dSs = dS.to_scipy() @ sig
    gk = gF["k"] - gL["k"] + float(np.real(np.sum(lc * (-1j * y[2] * e)))) - float(np.real(np.vdot(lam, -1j * dSs)))
    grad = {"nodes": gnodes, "normals": gF["source_normals"] - gL["source_normals"],
            "weights": gF["source_weights"] - gL["source_weights"], "k": gk}
    grad_frozen = {key: np.copy(v) for key, v in grad.items()}
    if dim_derivative:
        # entries (i, j) of dS and dD enter L with c_S = -i k and c_D = 1
        rp, ci, _ = dS.export()
        rows = np.repeat(np.arange(n), np.diff(rp))
        lr, sc = lam[rows], np.conj(sig[ci])
        gb = inti.bdim_pullback(op, Q, -lr * np.conj(-1j * k) * sc, -lr * sc, rp, ci,
                                green_multiplier=np.full(n, -0.5))
        for name in ("nodes", "normals", "weights"):
            grad[name] = grad[name] + gb[name]
        grad["k"] += gb["k"]
    return J, grad, {"solve": sinfo, "adjoint_solve": ainfo, "corrections": (dS, dD), "sigma": sig,
                     "grad_frozen": grad_frozen}


def shape_fd_check(inti, y, nv, w, qtags, k, dirs, grad, corr, ndir, seed, h=1e-6, sigma0=None,
                   rtol=1e-14, grad_frozen=None):
    """Central differences of J along ndir seeded random directions (nodes,
    normals, weights and k in turn), with the DIM correction frozen at the
    base geometry (what the adjoint-state gradient differentiates) and
    rebuilt on each perturbed geometry (the full corrected solve). The
    perturbed solves start from the base solution sigma0 and run to rtol:
    GMRES's stopping error divided by 2h is the FD noise floor (1e-12 / 2e-6
    would be 5e-7 of J). Returns a list of (which, adjoint, fd_frozen,
    fd_full, adjoint_frozen), the last from grad_frozen (the gradient with
    the corrections held fixed) when given."""
    rng = np.random.default_rng(seed)
    base = [y, nv, w, k]
    names = ("nodes", "normals", "weights", "k")
    out = []
    for t in range(ndir):
        which = t % 4
        v = rng.normal(size=np.shape(base[which])) if which != 3 else 1.0
        pred = float(np.sum(grad[names[which]] * v)) if which != 3 else float(grad["k"])
        gf = grad_frozen if grad_frozen is not None else grad
        predf = float(np.sum(gf[names[which]] * v)) if which != 3 else float(gf["k"])
        fds = []
        for frozen in (corr, None):
            vals = []
            for sgn in (1.0, -1.0):
                q = list(base)
                q[which] = base[which] + sgn * h * v
                vals.append(shape_problem(inti, q[0], q[1], q[2], qtags, q[3], dirs, frozen=frozen,
                                          rtol=rtol, x0=sigma0)[0])
            fds.append((vals[0] - vals[1]) / (2 * h))
        out.append((names[which], pred, fds[0], fds[1], predf))
    return out


def run_circle(inti, lib, ctx, args):
    """2D Laplace double layer on the unit circle, 128 P2 elements x 5 Gauss
    nodes (N = 640), sigma = cos(3 theta) + 0.1 N(0, 1): bdim_correction +
    assemble_matrix + one corrected matvec on the GPU, and the same three
    steps on the CPU oracle (the reference CPU path, restated) as the CPU
    baseline leg."""
    Q, sig = circle_workload(inti)
    n = len(Q)
    op = inti.Laplace(dim=2)
    times = []
    for it in range(12):  # 2 untimed warm-up pipelines (first-call scratch allocations), median of 10
        t0 = time.perf_counter()
        S, D = inti.single_double_layer(op=op, target=Q, source=Q, compression={"method": "none"},
                                        correction={"method": "dim"})
        y = D @ sig
        if it >= 2:
            times.append(time.perf_counter() - t0)
    t_gpu = float(np.median(times))
    Dm = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)
    Dm.attach(inti.bdim_correction(op, Q, Q, inti.IntegralOperator(inti.SingleLayerKernel(op), Q), Dm,
                                   green_multiplier=np.full(n, -0.5))[1])
    dx, dy = _dev_buffers(ctx, lib, 8 * n, np.ascontiguousarray(sig))
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)
    for _ in range(10):
        ctx.check(lib.inti_op_apply_device(Dm.handle, 1, dx, n, dy, n, one, zero))
    ctx.synchronize()
    t0 = time.perf_counter()
    reps = 200
    for _ in range(reps):
        ctx.check(lib.inti_op_apply_device(Dm.handle, 1, dx, n, dy, n, one, zero))
    ctx.synchronize()
    t_mv = (time.perf_counter() - t0) / reps
    ymv = _fetch(ctx, lib, dy, (n,), np.float64)
    ctx.check(lib.inti_device_free(ctx.handle, dx))
    ctx.check(lib.inti_device_free(ctx.handle, dy))
    cpu_leg = {}
    if not args.no_cpu:
        # the CPU baseline leg: the oracle runs the same pipeline and checks
        # the GPU results
        import oracle
        times = []
        for it in range(12):  # warmed up and timed as the GPU pipeline
            t0 = time.perf_counter()
            rp, ci, vS, vD, diag = oracle.bdim(0, 2, 0.0, Q, Q, np.full(n, -0.5))
            Am = oracle.assemble(0, 2, 1, 0.0, Q, Q).real
            yo = Am @ sig + oracle.csr_matvec(n, n, rp, ci, vD, sig).real
            if it >= 2:
                times.append(time.perf_counter() - t0)
        t_cpu = float(np.median(times))
        # the oracle's own rounding spread: the same algorithm built with FMA
        # contraction (the DIM least squares here have R-diagonal ratios ~5e8)
        rpf, cif, _, vDf, _ = oracle.bdim(0, 2, 0.0, Q, Q, np.full(n, -0.5), variant="fma")
        yf = Am @ sig + oracle.csr_matvec(n, n, rpf, cif, vDf, sig).real
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu_leg = {
            "pipeline_s": {"gpu_bdim_assemble_matvec": round(t_gpu, 4),
                           "cpu_oracle_bdim_assemble_matvec": round(t_cpu, 4)},
            "rel_err_gpu_vs_oracle": float(np.linalg.norm(ymv - yo) / np.linalg.norm(yo)),
            "rel_err_gpu_dense_vs_oracle": float(np.linalg.norm(y - yo) / np.linalg.norm(yo)),
            "oracle_rounding_spread": float(np.linalg.norm(yf - yo) / np.linalg.norm(yo)),
            "dim_rdiag_ratio": float(diag[0]),
            "cpu_baseline": {"value": round(t_cpu, 4), "unit": "s per bdim + assemble + matvec",
                             "cores": threads, "kind": "port",
                             "sample": "the whole 640-node pipeline, median of 10 after 2 warm-ups"},
        }
    return {
        "metric": "Gpair-evals/s of corrected DL matvec, 2D Laplace unit circle N=640, FP64",
        "value": round(float(n) * n / t_mv / 1e9, 3), "unit": "Gpair-evals/s", "n_gpus": 1, "steps": reps,
        "warmup": 10, "ms_per_step": round(t_mv * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2D Laplace DL, unit circle, 128 P2 elements x 5 Gauss nodes = 640 nodes, "
                               "sigma = cos(3 theta) + 0.1 N(0,1) (seed 20261202)", "nodes": n},
        "gpu_pipeline_s": round(t_gpu, 4),
        "cpu_baseline": None, **cpu_leg,
    }


def run_dense(inti, lib, ctx, args):
    """The assembled DIM-corrected 3D Helmholtz double layer (k = 2 pi, unit
    sphere, icosphere frequency 12: 17,280 nodes, a 4.8 GB complex matrix in
    HBM) applied to R = 16 and R = 64 complex right-hand sides by the FP64
    matrix-core GEMM, against the matrix-free 16-RHS direct sum of the same
    operator. 8 m n R flops per product; the matrix is read once."""
    Q = inti.sphere_quadrature(12)
    n = len(Q)
    op = inti.Helmholtz(k=2 * np.pi, dim=3)
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), Q)
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)
    _, dD = inti.bdim_correction(op, Q, Q, Sop, Dop, green_multiplier=np.full(n, -0.5))
    t0 = time.perf_counter()
    Dd = inti.DenseOperator(Dop, corrections=[(dD, 1.0)])
    ctx.synchronize()
    t_asm = time.perf_counter() - t0
    Dop.attach(dD)
    rng = np.random.default_rng(20261202)
    res = {}
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)
    for R in (16, 64):
        X = np.ascontiguousarray((rng.normal(size=(R, n)) + 1j * rng.normal(size=(R, n))) * np.sqrt(0.5))
        dx, dy = _dev_buffers(ctx, lib, X.nbytes, X)
        for _ in range(2):
            ctx.check(lib.inti_dense_gemm(ctx.handle, 1, n, n, R, Dd.device_pointer, n, dx, n, dy, n, one, zero))
        ctx.synchronize()
        kms = []
        t1 = time.perf_counter()
        reps = 10
        for _ in range(reps):
            ctx.check(lib.inti_dense_gemm(ctx.handle, 1, n, n, R, Dd.device_pointer, n, dx, n, dy, n, one, zero))
            kms.append(ctx.last_kernel_ms())
        ctx.synchronize()
        t_call = (time.perf_counter() - t1) / reps
        Yg = _fetch(ctx, lib, dy, (R, n), np.complex128)
        # matrix-free apply of the same corrected operator on the same columns
        for _ in range(1):
            ctx.check(lib.inti_op_apply_device(Dop.handle, R, dx, n, dy, n, one, zero))
        ctx.synchronize()
        t1 = time.perf_counter()
        kmf = []
        for _ in range(reps):
            ctx.check(lib.inti_op_apply_device(Dop.handle, R, dx, n, dy, n, one, zero))
            kmf.append(ctx.last_kernel_ms())
        ctx.synchronize()
        t_mf = (time.perf_counter() - t1) / reps
        Ym = _fetch(ctx, lib, dy, (R, n), np.complex128)
        ctx.check(lib.inti_device_free(ctx.handle, dx))
        ctx.check(lib.inti_device_free(ctx.handle, dy))
        k = float(np.mean(kms))
        flops = 8.0 * n * n * R
        res[R] = {"gemm_kernel_ms": round(k, 3), "gemm_call_ms": round(t_call * 1e3, 3),
                  "tflops": round(flops / (k * 1e-3) / 1e12, 2),
                  "matrix_GBps": round(16.0 * n * n / (k * 1e-3) / 1e9, 1),
                  "matrix_free_ms": round(t_mf * 1e3, 3),
                  # the library's events time the last launch of 16 columns
                  "matrix_free_kernel_ms_per_16": round(float(np.sum(kmf)) / reps, 3),
                  "rel_diff_vs_matrix_free": float(np.linalg.norm(Yg - Ym) / np.linalg.norm(Ym))}
        log(f"bench dense: R={R} {res[R]}")
    r16 = res[16]
    return {
        "metric": "TFLOP/s of the dense corrected-operator product on 16 complex RHS (FP64 matrix cores)",
        "value": r16["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": 10, "warmup": 2,
        "ms_per_step": r16["gemm_kernel_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"assembled DIM-corrected 3D Helmholtz DL, k=2pi, unit sphere icosphere freq 12: "
                               f"{n} nodes ({16.0 * n * n / 1e9:.2f} GB complex matrix), 16 and 64 complex RHS "
                               f"(seed 20261202)", "nodes": n},
        "assembly_s": round(t_asm, 3), "by_rhs": res,
        "roofline": {"R16": {"bound": "hbm", "achieved": r16["matrix_GBps"], "peak": 8000.0, "unit": "GB/s",
                             "frac": round(r16["matrix_GBps"] / 8000.0, 4)},
                     "R64": {"bound": "mfma", "achieved": res[64]["tflops"], "peak": 78.6, "unit": "TFLOP/s",
                             "frac": round(res[64]["tflops"] / 78.6, 4)}},
    }


if __name__ == "__main__":
    main()
