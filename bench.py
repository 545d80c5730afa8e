"""Benchmark of the line-opacity hot path (Korg.jl's line_absorption!) on one MI355X.

Workload (the headline configuration): solar stand-in atmosphere (56 layers), 4000-7000 A at
0.01 A (300,001 points), 300,000 seeded synthetic lines; alpha is 56 x 300,001 Float32 in HBM.
A step is one full line_absorption! pass (setup, binning, profile accumulation) over
device-resident inputs.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "korg.jl_amd"))

METRIC = "spectra/s + Voigt evals/s, FP32 vs FP64 GPU at ≤1e-4 flux err, vs CPU"
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X published dense FP32 matrix peak

# Far-wing kernel (k_wing_mfma): the wing profile of a line at layer l and point j is the
# series sum_k c_k(l) R_j^k, R_j = 1/(lambda_j - lambda_0)^2, k = 1..4; the R^3, R^4 terms are
# evaluated only inside the line's tail radius (beyond it they are below 1e-6 of the leading
# term).  Algorithmic cost: 2 FMA = 4 flops per (layer, far-zone point) for R, R^2, plus 4 flops
# per tail evaluation.  Issued work is counted separately: one v_mfma_f32_32x32x2f32 is
# 32*32*2*2 = 4096 flops; layer padding to 64 and 32-point block alignment make the issued
# flops exceed the algorithmic ones.
WING_FLOPS_PER_EVAL = 4.0
WING_FLOPS_PER_TAIL_EVAL = 4.0
MFMA_FLOPS = 4096.0
# The reference's line_profile + voigt_hjerting per evaluation (src/line_absorption.jl:245-310,
# division/exp/sqrt counted as one op): A 19, B 27, C 57, D 28.
REF_FLOPS = (19.0, 27.0, 57.0, 28.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--cpu-fraction", type=float, default=0.5,
                    help="fraction of the lines timed on the CPU oracle (every k-th line)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flux", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (PMC byte-counter scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the line-opacity "
              "path shards by wavelength tiles or by star, which this build does not run.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from korg_line import Context, DeviceProblem, rt, workloads
    from korg_line.hipmem import synchronize as device_sync

    t = time.time()
    P, atm, wls = workloads.solar_wide()
    log(f"workload built in {time.time() - t:.1f}s: {len(P.lines)} lines, {P.n_wl} points, "
        f"{P.n_layers} layers")
    ctx = Context(0)
    D = DeviceProblem(P)
    a32 = D.new_alpha("float32")
    device_sync()

    def run(alpha, precision, steps, warmup):
        for _ in range(warmup):
            ctx.line_absorption_device(alpha, D, precision, synchronize=True)
        device_sync()
        ctx.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.line_absorption_device(alpha, D, precision, synchronize=True)
            stats.append(ctx.stats())
        ctx.synchronize()
        device_sync()
        return (time.perf_counter() - t0) / steps, stats

    log("timing the Float32 form ...")
    t32, st32 = run(a32, "fp32", args.steps, args.warmup)
    log(f"  {t32 * 1e3:.3f} ms/step")
    log("timing the Float64 form ...")
    a64 = D.new_alpha("float64")
    t64, st64 = run(a64, "fp64", args.fp64_steps, 1)
    log(f"  {t64 * 1e3:.3f} ms/step")

    # branch mix (untimed) and accuracy of one fresh pass of each form
    ctx.set_tuning(count_branches=1)
    z32 = D.new_alpha("float32")
    z64 = D.new_alpha("float64")
    ctx.line_absorption_device(z64, D, "fp64")
    ctx.line_absorption_device(z32, D, "fp32")
    sb = ctx.stats()
    ctx.set_tuning()
    nA, nB, nC, nD = sb["n_evals_branch"]
    E = sb["n_evals"]
    lines64 = z64.numpy().T  # (layers, points)
    lines32 = z32.numpy().astype(np.float64).T
    m = np.abs(lines64) >= 1e-8 * np.abs(lines64).max()
    alpha_err = float(np.max(np.abs(lines32[m] - lines64[m]) / np.abs(lines64[m])))
    flux_err = None
    if not args.no_flux:
        cntm = np.array([np.interp(wls.all_wls, P.cntm_knots, P.cntm_values[l])
                         for l in range(P.n_layers)])
        aref = np.array([np.interp(5000e-8, P.cntm_knots, P.cntm_values[l])
                         for l in range(P.n_layers)])
        tau = 10.0 ** atm.lgTau5
        f64 = rt.normalized_flux(cntm + lines64, cntm, aref, tau, atm.temps, wls.all_wls)
        f32 = rt.normalized_flux(cntm + lines32, cntm, aref, tau, atm.temps, wls.all_wls)
        flux_err = float(np.max(np.abs(f32 - f64)))
    log(f"alpha max rel err {alpha_err:.3e}, flux max abs err {flux_err}")

    wing_ms = float(np.mean([s["profile_ms"] for s in st32]))
    near_ms = float(np.mean([s["core_ms"] for s in st32]))
    setup_ms = float(np.mean([s["setup_ms"] for s in st32]))
    bin_ms = float(np.mean([s["bin_ms"] for s in st32]))
    e_near = st32[-1]["n_evals_near"]
    e_far = E - e_near
    e_tail = st32[-1]["n_evals_tail"]
    n_mfma = st32[-1]["n_wing_mfma"]
    wing_flops = WING_FLOPS_PER_EVAL * e_far + WING_FLOPS_PER_TAIL_EVAL * e_tail
    achieved = wing_flops / (wing_ms * 1e-3) / 1e12
    issued = MFMA_FLOPS * n_mfma / (wing_ms * 1e-3) / 1e12
    ref_flops = sum(c * f for c, f in zip((nA, nB, nC, nD), REF_FLOPS))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(P, args.cpu_fraction)

    out = {
        "metric": METRIC,
        "value": 1.0 / t32,
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "solar_wide: 300,000 seeded lines, 4000-7000 A at 0.01 A (300,001 "
                        "points), 56-layer solar stand-in atmosphere, FP32 alpha resident in HBM",
            "precision": "fp32 profiles + fp64 setup/window/branch/offset fix-ups",
            "parallelism": "single GPU",
        },
        "voigt_evals_per_s": E / t32,
        "n_evals": E,
        "evals_by_branch": {"A_far_wing": nA, "B_core": nB, "C_core": nC, "D_damped_wing": nD},
        "fp64_spectra_per_s": 1.0 / t64,
        "fp64_ms_per_step": t64 * 1e3,
        "fp32_over_fp64": t64 / t32,
        "alpha_max_rel_err_fp32_vs_fp64": alpha_err,
        "flux_max_abs_err_fp32_vs_fp64": flux_err,
        "evals_far_wing_mfma": e_far,
        "evals_near_zone_exact": e_near,
        "evals_far_with_tail_terms": e_tail,
        "kernel_ms": {"setup": setup_ms, "binning": bin_ms, "k_wing_mfma": wing_ms,
                      "k_near": near_ms},
        "reference_equivalent_tflops": ref_flops / t32 / 1e12,
        "roofline": {
            "bound": "mfma",
            "kernel": "k_wing_mfma",
            "achieved": achieved,
            "peak": PEAK_FP32_MFMA_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_MFMA_TFLOPS,
            "traffic": load_traffic(),
            "flops_per_launch": wing_flops,
            "issued_mfma_tflops": issued,
            "issued_frac": issued / PEAK_FP32_MFMA_TFLOPS,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def copy_calibration(mib, reps=5):
    """Plain device-to-device copies of a known size: FETCH_SIZE/WRITE_SIZE per copy dispatch
    against mib MiB give the counters' scale."""
    import numpy as np
    from korg_line.hipmem import D2D, DeviceArray, hip, synchronize
    n = mib * (1 << 20) // 4
    src = DeviceArray.from_numpy(np.arange(n, dtype=np.float32))
    dst = DeviceArray.zeros((n,), np.float32)
    import ctypes
    for _ in range(reps):
        rc = hip().hipMemcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), n * 4, D2D)
        assert rc == 0
    synchronize()
    log(f"copied {mib} MiB x {reps}")


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic_k_wing_mfma_solar_wide.json")
    if os.path.exists(p):
        try:
            return json.load(open(p))["bytes_per_launch"]
        except Exception:
            return None
    return None


def cpu_baseline(P, fraction):
    """The Float64 CPU restatement with the reference's threading (line chunks, private alpha
    per thread, summed) on a sample of every k-th line, scaled to one full spectrum by the exact
    evaluation counts."""
    import numpy as np
    from oracle import oracle
    from korg_line.problem import Problem

    k = max(1, int(round(1 / fraction)))
    idx = np.arange(0, len(P.lines), k)
    S = Problem(P.all_wls, P.lines.take(idx), P.species_mass, P.species_is_mol, P.n_div_U,
                P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode, P.cntm_values,
                P.cntm_knots, P.cutoff_threshold)
    threads = min(16, os.cpu_count() or 1)
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    lb, ub, e_full = oracle.line_absorption(a, P, mode=oracle.WINDOWS_ONLY, return_windows=True,
                                            nthreads=threads)
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    t0 = time.perf_counter()
    _, _, e_sample = oracle.line_absorption(a, S, nthreads=threads, mode=oracle.KORG_CHUNKS,
                                            return_windows=True)
    dt = time.perf_counter() - t0
    t_full = dt * e_full / e_sample
    log(f"cpu baseline: {len(idx)} lines ({e_sample:.3g} evals) in {dt:.1f}s on {threads} "
        f"threads -> {t_full:.1f}s per spectrum")
    return {"value": 1.0 / t_full, "unit": "spectra/s", "cores": threads, "kind": "port",
            "evals_per_s": e_sample / dt,
            "sample": f"every {k}th line of the same 300,000-line list ({len(idx)} lines, "
                      f"{e_sample:.3g} of {e_full:.3g} evaluations), full grid, Float64, "
                      f"reference thread schedule; time scaled by the evaluation ratio"}


if __name__ == "__main__":
    main()
