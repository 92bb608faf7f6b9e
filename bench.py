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


def make_workload(inti, freq):
    Q = inti.sphere_quadrature(freq, qorder=4, base="ico")
    rng = np.random.default_rng(20261202)
    n = len(Q)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)) * np.sqrt(0.5)
    return Q, x


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
    while True:
        t0 = time.perf_counter()
        oracle.matvec(1, 3, 4, k, Q, Q, x, coefs=coefs, rows=(0, rows), nthreads=threads)
        t = time.perf_counter() - t0
        if t > target_seconds / 4 or rows >= n:
            break
        rows = min(n, int(rows * max(2.0, target_seconds / 4 / max(t, 1e-3))))
    rate = rows * n / t
    return {"value": rate / 1e9, "unit": "Gpair-evals/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 direct sum of the same combined-field operator, rows 0..{rows} "
                      f"of {n} against all {n} sources ({rows * n:.3g} pairs, {t:.1f} s)"}


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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="profiling runs: time one leg only, no checks, no JSON")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="profiler calibration: run a plain device copy of this many bytes 5 times")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the direct sum would shard by target rows; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    inti = load_package()
    lib = inti._lib.load()
    ctx = inti.default_context()
    if args.copy_bytes:
        calibrate_copy(ctx, lib, args.copy_bytes)
        return
    ctx.set_profiling(True)

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

    # FP64 GPU vs the oracle on a row sample (parity of the measured result)
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

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(oracle, Q, x, k)
        log(f"bench: cpu baseline {cpu['value']:.4f} Gpair/s on {cpu['cores']} threads")

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
            "correction_nnz": corr[1].nnz, "precision": "mixed",
        },
        "fp64_value": round(pairs / t64 / 1e9, 3),
        "mixed_over_fp64": round(t64 / tmx, 3),
        "rel_err_mixed_vs_fp64": err_mx,
        "relinf_err_mixed_vs_fp64": err_mx_inf,
        "rel_err_fp64_vs_oracle_rows": err64,
        "dim_build_s": round(t_corr, 3),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_VALU_TFLOPS, 4),
                     "traffic": traffic,
                     "kernel": "k_apply<HELMHOLTZ,3,COMB_SD,MIXED>", "kernel_ms": round(kmx, 4),
                     "flops_per_pair": FLOPS_PER_PAIR},
        "cpu_baseline": cpu,
    }
    ctx.check(lib.inti_device_free(ctx.handle, dx))
    ctx.check(lib.inti_device_free(ctx.handle, dy))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
