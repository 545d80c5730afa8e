"""Benchmark of the line-opacity hot path (Korg.jl's line_absorption!) on one MI355X.

Workload (the headline configuration): solar stand-in atmosphere (56 layers), 4000-7000 A at
0.01 A (300,001 points), 300,000 seeded synthetic lines; alpha is 56 x 300,001 Float32 in HBM.
A step is one full line_absorption! pass (setup, binning, profile accumulation) over
device-resident inputs, replayed from a HIP graph under the autotuned tiling; per-kernel times
come from HIP events around the same kernels in direct calls.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

``--workload fit`` measures the fit step (one JSON line of its own): 64 stars drawn as for the
batched form, 5000-6000 A at 0.01 A (100,001 points), 100,000 shared lines; per step, the line
opacity in Float32 with Float64 fix-ups, radiative transfer, chi^2 against a noisy copy of the
flux and its gradient through the line-opacity adjoint, all in one HIP graph replay.

``--workload batched`` measures the batched form instead (one JSON line of its own): 256 stars,
Teff ~ U[4000, 6500] K, log g ~ U[1, 5], [M/H] ~ U[-2.5, 0.5] (scaled solar stand-ins), each with
a window of U[50, 1000] A inside 4000-7000 A at 0.01 A and its own list of U[5k, 150k] lines;
FP32 alpha; the tuning from the autotuner (its results file ships in the package); the whole
pipeline captured in a HIP graph, a step being one replay over all 256 stars.
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
PEAK_BF16_MFMA_TFLOPS = 2500.0  # MI355X published dense bfloat16 matrix peak
PEAK_FP32_TFLOPS = 157.3        # MI355X published FP32 vector peak
PEAK_FP64_MFMA_TFLOPS = 78.6    # MI355X published dense FP64 matrix peak
DTYPE = "fp32+bf16x3"           # the arithmetic the Float32 form computes in
PRECISION = ("fp32 exact Hunger evaluation in each line's near zone + far wing as a 4-term "
             "series on the bfloat16 matrix cores (hi/lo splits of both operands, three products, "
             "fp32 accumulate), summed where a 64-point sub-tile is more than 8 half-widths from "
             "the line centre as a 12-term expansion about the sub-tile centre (far field, same "
             "bfloat16 products) + fp64 setup/window/zone/branch/offset fix-ups")

# Far-wing kernel (k_wing_bf16): the wing profile of a line at layer l and point j is the
# series sum_k c_k(l) R_j^k, R_j = 1/(lambda_j - lambda_0)^2, k = 1..4.  Algorithmic cost: 2 FMA =
# 4 flops per (layer, far-zone point) for R, R^2, plus 4 flops per evaluation inside the line's
# tail radius for R^3, R^4 (beyond it they are below 1e-5 of the leading term).  Issued work:
# v_mfma_f32_32x32x16_bf16, 32*32*16*2 = 32768 flops each, three per group of 4 lines (bfloat16
# high/low splits of both operands) with all four terms everywhere and layers padded to 64.
WING_FLOPS_PER_EVAL = 4.0
WING_FLOPS_PER_TAIL_EVAL = 4.0
MFMA_FLOPS = 32768.0
# Far-field kernel (k_wing_ff): three v_mfma_f32_32x32x16_bf16 per (group of 4 lines, 32-column
# block of the 96 (sub-tile, term) columns, 32-layer half) of a 512-point super-tile's list: 4.5
# per list entry and 64-layer block.
FF_MFMA_PER_ENTRY = 4.5
F64_MFMA_FLOPS = 2048.0  # v_mfma_f64_16x16x4_f64
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
    ap.add_argument("--cpu-fraction", type=float, default=1.0,
                    help="fraction of the lines timed on the CPU oracle (every k-th line; 1: the "
                         "whole workload)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flux", action="store_true")
    ap.add_argument("--no-hydrogen", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (PMC byte-counter scale)")
    ap.add_argument("--near-serial", action="store_true",
                    help="run the near zone after the far wing on one stream in every pass (for "
                         "kernel traces whose durations are the kernels' own)")
    ap.add_argument("--workload", choices=("solar_wide", "batched", "fit"), default="solar_wide")
    ap.add_argument("--stars", type=int, default=256, help="batched workload: number of stars")
    ap.add_argument("--tuning-cache", default=None,
                    help="batched workload: autotuner results file (default: the one shipped "
                         "in the package; a file without this shape's entry is timed and appended)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the line-opacity "
              "path shards by wavelength tiles or by star, which this build does not run.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "fit":
        bench_fit(args)
        return

    import numpy as np
    from korg_line import Context, DeviceProblem, FluxTarget, workloads
    from korg_line.hipmem import synchronize as device_sync

    t = time.time()
    P, atm, wls = workloads.solar_wide()
    t_flatten = time.time() - t
    log(f"workload built in {t_flatten:.1f}s: {len(P.lines)} lines, {P.n_wl} points, "
        f"{P.n_layers} layers")
    ctx = Context(0)
    D = DeviceProblem(P)
    a32 = D.new_alpha("float32")
    a64 = D.new_alpha("float64")
    device_sync()
    cache = args.tuning_cache or TUNING_CACHE
    tune32 = ctx.autotune_batched([a32], [D], "fp32", cache_path=cache)
    tune64 = ctx.autotune_batched([a64], [D], "fp64", cache_path=cache)
    log(f"tuning: fp32 {tune32}, fp64 {tune64}")

    def use(tuning, near_concurrent=0):
        t = {k: v for k, v in tuning.items() if k != "count_branches"}
        t["near_concurrent"] = 1 if args.near_serial else near_concurrent
        ctx.set_tuning(**t)

    def direct(alpha, precision, steps, warmup):
        """Direct calls (events around every kernel): the per-kernel device times."""
        for _ in range(warmup):
            ctx.line_absorption_device(alpha, D, precision, synchronize=True)
        stats = []
        for _ in range(steps):
            ctx.line_absorption_device(alpha, D, precision, synchronize=True)
            stats.append(ctx.stats())
        device_sync()
        return stats

    def replay(alpha, precision, steps, warmup):
        """The timed steps: replays of the whole pass captured in a HIP graph."""
        with ctx.plan_batched([alpha], [D], precision) as plan:
            for _ in range(warmup):
                plan.launch()
            device_sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                plan.launch()
            device_sync()
            return (time.perf_counter() - t0) / steps

    log("timing the Float32 form ...")
    # per-kernel times: the near zone after the wing, on one stream, so that each kernel's events
    # time it alone; the timed steps then run the default schedule (near zone beside the wing)
    use(tune32, near_concurrent=1)
    st32 = direct(a32, "fp32", args.steps, args.warmup)
    use(tune32)
    t32 = replay(a32, "fp32", args.steps, args.warmup)
    direct32 = float(np.mean([s["total_ms"] for s in st32]))
    log(f"  {t32 * 1e3:.3f} ms/step (graph replay); direct calls {direct32:.3f} ms device time")
    log("timing the Float64 form ...")
    use(tune64)
    st64 = direct(a64, "fp64", 1, 1)
    t64 = replay(a64, "fp64", args.fp64_steps, 1)
    log(f"  {t64 * 1e3:.3f} ms/step")
    # the Float32 form's algorithm in Float64 (series far wing on the Float64 matrix cores, exact
    # near zones): the like-for-like counterpart of the Float32 form
    log("timing the Float64 series form ...")
    ctx.set_tuning()
    as64 = D.new_alpha("float64")
    sts = direct(as64, "fp64_series", 1, 1)
    ts64 = replay(as64, "fp64_series", max(args.fp64_steps, 3), 1)
    log(f"  {ts64 * 1e3:.3f} ms/step")
    use(tune32)

    # branch mix (untimed) and accuracy of one fresh pass of each form
    z32 = D.new_alpha("float32")
    z64 = D.new_alpha("float64")
    use(tune64)
    ctx.line_absorption_device(z64, D, "fp64")
    use(tune32)
    ctx.line_absorption_device(z32, D, "fp32")
    ctx.set_tuning(**{**{k: v for k, v in tune32.items() if k != "count_branches"},
                      "count_branches": 1})
    zb = D.new_alpha("float32")
    ctx.line_absorption_device(zb, D, "fp32")
    sb = ctx.stats()
    use(tune32)
    nA, nB, nC, nD = sb["n_evals_branch"]
    E = sb["n_evals"]
    lines64 = z64.numpy().T  # (layers, points)
    lines32 = z32.numpy().astype(np.float64).T
    m = np.abs(lines64) >= 1e-8 * np.abs(lines64).max()
    alpha_err = float(np.max(np.abs(lines32[m] - lines64[m]) / np.abs(lines64[m])))
    zs = D.new_alpha("float64")
    ctx.set_tuning()
    ctx.line_absorption_device(zs, D, "fp64_series")
    use(tune32)
    liness = zs.numpy().T
    alpha_err_series = float(np.max(np.abs(liness[m] - lines64[m]) / np.abs(lines64[m])))
    alpha_err_32_series = float(np.max(np.abs(lines32[m] - liness[m]) / np.abs(liness[m])))
    del lines64, lines32, liness, zs
    flux_err = None
    if not args.no_flux:
        # the whole spectrum on the device: continuum on the grid, lines added, Korg's planar
        # transfer (korg_continuum_batched, korg_flux_batched), once per form
        tau, aref = workloads.fit_reference(atm)
        fn = {}
        for prec, dt, tune in (("fp64", "float64", tune64), ("fp32", "float32", tune32)):
            use(tune)
            tot = D.new_alpha(dt)
            ctx.continuum_batched([tot], [D])
            ctx.line_absorption_device(tot, D, prec)
            ft = FluxTarget(P, tau, aref)
            ctx.flux_batched([tot], [D], [ft])
            fn[prec] = ft.flux() / ft.cntm()
            del tot
        use(tune32)
        flux_err = float(np.max(np.abs(fn["fp32"] - fn["fp64"])))
    log(f"alpha max rel err {alpha_err:.3e}, flux max abs err {flux_err}")

    wing_ms = float(np.mean([s["profile_ms"] for s in st32]))  # far field + wing tiles
    far_ms = float(np.mean([s["far_ms"] for s in st32]))
    near_ms = float(np.mean([s["core_ms"] for s in st32]))
    setup_ms = float(np.mean([s["setup_ms"] for s in st32]))
    bin_ms = float(np.mean([s["bin_ms"] for s in st32]))
    e_near = st32[-1]["n_evals_near"]
    e_far = E - e_near
    e_tail = st32[-1]["n_evals_tail"]
    n_mfma = st32[-1]["n_wing_mfma"]
    n_lblocks = -(-P.n_layers // 64)
    ff_mfma = FF_MFMA_PER_ENTRY * st32[-1]["n_far_entries"] * n_lblocks
    ff_issued = MFMA_FLOPS * ff_mfma / (far_ms * 1e-3) / 1e12 if far_ms > 0 else 0.0
    wing_flops = WING_FLOPS_PER_EVAL * e_far + WING_FLOPS_PER_TAIL_EVAL * e_tail
    achieved = wing_flops / (wing_ms * 1e-3) / 1e12
    issued = MFMA_FLOPS * n_mfma / (wing_ms * 1e-3) / 1e12
    ref_flops = sum(c * f for c, f in zip((nA, nB, nC, nD), REF_FLOPS))
    # the near zone holds every B and C evaluation; the rest of it is A (counted) and D
    near_A = sb["n_evals_near_A"]
    near_D = sb["n_evals_near"] - nB - nC - near_A
    near_flops = sum(c * f for c, f in zip((near_A, nB, nC, near_D), REF_FLOPS))
    near_achieved = near_flops / (near_ms * 1e-3) / 1e12
    traffic = load_traffic()
    rooflines = {
        "k_near2": {
            "bound": "valu", "achieved": near_achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": near_achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic.get("k_near2"), "kernel_ms": near_ms, "flops_per_launch": near_flops,
            "evals_by_branch": {"A": near_A, "B": nB, "C": nC, "D": near_D},
            "note": "exact Hunger evaluations of the near zones (k_near2: lanes over 8 layers x 8 "
                    "points), the reference's per-branch op counts (A 19, B 27, C 57, D 28) over "
                    "the FP32 vector peak (which counts packed FP32; this build issues none)"},
        "far_wing": {
            "bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_MFMA_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_MFMA_TFLOPS,
            "traffic": None, "kernel_ms": wing_ms,
            "flops_per_launch": wing_flops, "issued_mfma_tflops": issued,
            "issued_frac": issued / PEAK_BF16_MFMA_TFLOPS,
            "note": "k_wing_ff + k_wing_bf16: the reference-equivalent work of every far-zone "
                    "evaluation, 4 flops (R, R^2) + 4 per tail evaluation (R^3, R^4), over the "
                    "bfloat16 matrix peak; the far field does most of it as 12 terms per 64 points, "
                    "so this algorithmic rate can exceed what the instructions issue"},
        "k_wing_ff": {
            "bound": "mfma", "achieved": ff_issued, "peak": PEAK_BF16_MFMA_TFLOPS,
            "unit": "TFLOP/s", "frac": ff_issued / PEAK_BF16_MFMA_TFLOPS,
            "traffic": traffic.get("k_wing_ff"), "kernel_ms": far_ms,
            "far_entries": st32[-1]["n_far_entries"],
            "note": "issued v_mfma_f32_32x32x16_bf16 (4.5 per (line, 512-point super-tile) entry "
                    "and 64-layer block) over the bfloat16 matrix peak; bound by the vector work "
                    "that builds the expansion coefficients"},
        "k_wing_bf16": {
            "bound": "mfma", "achieved": None, "peak": PEAK_BF16_MFMA_TFLOPS,
            "unit": "TFLOP/s", "frac": None,
            "traffic": traffic.get("k_wing_bf16"), "kernel_ms": wing_ms - far_ms,
            "note": "the tiles' remaining (line, tile) pairs term by term and each tile's start "
                    "from the far-field sums; latency-bound on the lines' coefficient loads"},
    }
    s_mfma = sts[-1]["n_wing_mfma"]
    s_wing_ms = sts[-1]["profile_ms"]
    s_ach = s_mfma * F64_MFMA_FLOPS / (s_wing_ms * 1e-3) / 1e12
    rooflines["k_wing_f64"] = {
        "bound": "mfma", "achieved": s_ach, "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
        "frac": s_ach / PEAK_FP64_MFMA_TFLOPS, "traffic": None, "kernel_ms": s_wing_ms,
        "note": "the Float64 series form's far wing: v_mfma_f64_16x16x4_f64 (2048 flops each, 8 "
                "per line and 32-point tile) over the FP64 matrix peak"}
    dominant = "k_near2" if near_ms >= wing_ms else "far_wing"
    log(f"k_near2 {near_ms:.3f} ms ({near_achieved:.1f} TFLOP/s), far wing {wing_ms:.3f} ms "
        f"(k_wing_ff {far_ms:.3f} ms; {achieved:.1f} TFLOP/s algorithmic)")

    hyd = None if args.no_hydrogen else hydrogen_lines_bench(ctx, D, atm)
    e2e = None if args.no_cpu else end_to_end_host(ctx, P, t_flatten)

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
        "dtype": DTYPE,
        "data": "synthetic",
        "config": {
            "workload": "solar_wide: 300,000 seeded lines, 4000-7000 A at 0.01 A (300,001 "
                        "points), 56-layer solar stand-in atmosphere, FP32 alpha resident in HBM",
            "precision": PRECISION,
            "parallelism": "single GPU",
            "tuning": tune32,
            "step": "one HIP graph replay of the whole pass (setup, binning, profiles)",
        },
        "direct_call_ms": direct32,
        "voigt_evals_per_s": E / t32,
        "n_evals": E,
        "evals_by_branch": {"A_far_wing": nA, "B_core": nB, "C_core": nC, "D_damped_wing": nD},
        "fp64_spectra_per_s": 1.0 / t64,
        "fp64_ms_per_step": t64 * 1e3,
        "fp32_over_fp64": t64 / t32,
        "fp64_series_ms_per_step": ts64 * 1e3,
        "fp32_over_fp64_series": ts64 / t32,
        "alpha_max_rel_err_fp64_series_vs_fp64": alpha_err_series,
        "alpha_max_rel_err_fp32_vs_fp64_series": alpha_err_32_series,
        "fp64_series_kernels_ms": {"setup": sts[-1]["setup_ms"], "wing_f64_mfma": sts[-1]["profile_ms"],
                                   "near_f64": sts[-1]["core_ms"]},
        "alpha_max_rel_err_fp32_vs_fp64": alpha_err,
        "flux_max_abs_err_fp32_vs_fp64": flux_err,
        "flux_scheme": "normalized flux of continuum + lines by the device's Korg planar transfer "
                       "(korg_flux_batched: anchored tau, linear source, Korg's E2), both forms",
        "evals_far_wing_mfma": e_far,
        "evals_near_zone_exact": e_near,
        "evals_far_with_tail_terms": e_tail,
        "kernel_ms": {"setup": setup_ms, "binning": bin_ms, "far_wing": wing_ms,
                      "k_wing_ff": far_ms, "k_wing_bf16": wing_ms - far_ms, "k_near2": near_ms},
        "reference_equivalent_tflops": ref_flops / t32 / 1e12,
        "roofline": {"kernel": dominant, **{k: v for k, v in rooflines[dominant].items()
                                            if k in ("bound", "achieved", "peak", "unit", "frac",
                                                     "traffic")}},
        "rooflines": rooflines,
        "end_to_end_host_f64": e2e,
        "cpu_baseline": cpu,
    }
    if hyd is not None:
        out["hydrogen_lines"] = hyd
    print(json.dumps(out), flush=True)
    ctx.close()


def hydrogen_lines_bench(ctx, D, atm, reps=5):
    """synthesize's hydrogen lines (korg_hydrogen_line_absorption) on the same grid and layers:
    the 84 tabulated transitions, 150 A windows, MHD as synthesize chooses it, into a scratch
    Float32 alpha; host-timed over ``reps`` calls between device synchronizations."""
    import numpy as np
    from korg_line.atmosphere import partition_function
    from korg_line.hipmem import synchronize as device_sync
    U = partition_function("H_I")(np.log(atm.temps))
    a = D.new_alpha("float32")
    ctx.set_hydrogen_tables()
    ctx.hydrogen_line_absorption(a, D, U)  # warm-up
    device_sync()
    t = time.perf_counter()
    for _ in range(reps):
        ctx.hydrogen_line_absorption(a, D, U, synchronize=False)
    device_sync()
    ms = (time.perf_counter() - t) / reps * 1e3
    log(f"hydrogen lines: {ms:.3f} ms per call")
    return {"ms": ms, "transitions": 84, "window_A": 150.0,
            "scheme": "Stehle-Hutcheon Stark tables (Lyman/Balmer/Paschen to n = 30), Barklem "
                      "resonance for H alpha-gamma, MHD; Float64, host-timed"}


TUNING_CACHE = os.path.join(ROOT, "korg.jl_amd", "korg_line", "tuning_mi355x.txt")


def bench_batched(args):
    """The batched forward over many stars as one graph replay per step."""
    import numpy as np
    from korg_line import Context, DeviceProblem, FluxTarget, workloads
    from korg_line.hipmem import synchronize as device_sync

    t = time.time()
    stars = workloads.batched_stars(n_stars=args.stars, seed=3)
    probs = [P for P, _, _ in stars]
    n_lines = sum(len(P.lines) for P in probs)
    n_pts = sum(P.n_wl for P in probs)
    L = probs[0].n_layers
    log(f"{len(probs)} stars built in {time.time() - t:.1f}s: {n_lines} lines, {n_pts} points, "
        f"{L} layers ({n_pts * L * 4 / 1e9:.2f} GB of FP32 alpha)")
    ctx = Context(0)
    dprobs = [DeviceProblem(P) for P in probs]
    a32 = [d.new_alpha("float32") for d in dprobs]
    device_sync()
    t = time.time()
    tuning = ctx.autotune_batched(a32, dprobs, "fp32", cache_path=args.tuning_cache or TUNING_CACHE)
    log(f"tuning {tuning} ({time.time() - t:.1f}s)")
    t = time.time()
    plan = ctx.plan_batched(a32, dprobs, "fp32")
    log(f"plan captured in {time.time() - t:.2f}s: {plan.info()}")

    def replay(steps):
        stats = []
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.launch(synchronize=True)
            stats.append(ctx.stats())
        return (time.perf_counter() - t0) / steps, stats

    replay(args.warmup)
    device_sync()
    t32, st32 = replay(args.steps)
    device_sync()
    log(f"  FP32 graph replay: {t32 * 1e3:.2f} ms/step ({len(probs) / t32:.1f} spectra/s)")
    E = st32[-1]["n_evals"]
    stream_ms = float(np.mean([s["total_ms"] for s in st32]))

    # the all-FP64 form of the same batch (direct call), and its accuracy against it
    ctx.set_tuning()
    a64 = [d.new_alpha("float64") for d in dprobs]
    ctx.line_absorption_batched(a64, dprobs, "fp64")
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        ctx.line_absorption_batched(a64, dprobs, "fp64")
    t64 = (time.perf_counter() - t0) / args.fp64_steps
    log(f"  FP64 direct: {t64 * 1e3:.1f} ms/step")
    # accuracy on fresh results of every 16th star (alpha and normalized flux)
    ctx.set_tuning(**{k: v for k, v in tuning.items() if k != "count_branches"})
    sample = list(range(0, len(probs), 16))
    z32 = [dprobs[k].new_alpha("float32") for k in sample]
    z64 = [dprobs[k].new_alpha("float64") for k in sample]
    ctx.line_absorption_batched(z32, [dprobs[k] for k in sample], "fp32")
    ctx.line_absorption_batched(z64, [dprobs[k] for k in sample], "fp64")
    alpha_err = flux_err = 0.0
    for k, b32, b64 in zip(sample, z32, z64):
        l64 = b64.numpy().T
        l32 = b32.numpy().astype(np.float64).T
        m = np.abs(l64) >= 1e-8 * np.abs(l64).max()
        if m.any():
            alpha_err = max(alpha_err, float(np.max(np.abs(l32[m] - l64[m]) / np.abs(l64[m]))))
    # normalized flux of the sample on the device: continuum + lines, Korg's planar transfer
    sd = [dprobs[k] for k in sample]
    fts = {}
    for prec, dt in (("fp64", "float64"), ("fp32", "float32")):
        tot = [d.new_alpha(dt) for d in sd]
        ctx.continuum_batched(tot, sd)
        ctx.line_absorption_batched(tot, sd, prec)
        fts[prec] = [FluxTarget(stars[k][0], *workloads.fit_reference(stars[k][1])) for k in sample]
        ctx.flux_batched(tot, sd, fts[prec])
        del tot
    for a, b in zip(fts["fp32"], fts["fp64"]):
        flux_err = max(flux_err, float(np.max(np.abs(a.flux() / a.cntm() - b.flux() / b.cntm()))))
    log(f"  sample of {len(sample)} stars: alpha max rel err {alpha_err:.3e}, flux {flux_err:.3e}")

    # the rows on either side of the line opacity, over the whole batch: the continuum on the grid
    # (an HBM write of every alpha, korg_continuum_batched) and the emergent flux of continuum +
    # lines (korg_flux_batched); host-timed over 3 synchronous calls each
    ctx.continuum_batched(a32, dprobs)
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.continuum_batched(a32, dprobs)
    tc = (time.perf_counter() - t0) / 3
    cbytes = 4.0 * L * n_pts
    fts = [FluxTarget(P, 10.0 ** atm.lgTau5, atm.continuum(np.array([5000e-8]))[:, 0])
           for P, atm, _ in stars]
    ctx.line_absorption_batched(a32, dprobs, "fp32")
    ctx.flux_batched(a32, dprobs, fts)
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.flux_batched(a32, dprobs, fts)
    tf = (time.perf_counter() - t0) / 3
    log(f"  continuum on the grid: {tc * 1e3:.2f} ms ({cbytes / tc / 1e9:.0f} GB/s); flux of "
        f"{len(probs)} spectra: {tf * 1e3:.2f} ms")

    cpu = None if args.no_cpu else cpu_baseline_batched(stars, E / len(probs))
    out = {
        "metric": METRIC,
        "value": len(probs) / t32,
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "data": "synthetic",
        "config": {
            "workload": f"batched: {len(probs)} stars (Teff U[4000,6500], log g U[1,5], "
                        f"[M/H] U[-2.5,0.5]), windows U[50,1000] A in 4000-7000 A at 0.01 A, "
                        f"U[5k,150k] lines each ({n_lines} lines, {n_pts} points), 56 layers, "
                        "FP32 alpha in HBM, one HIP graph replay per step",
            "precision": PRECISION,
            "parallelism": "single GPU",
            "tuning": tuning,
        },
        "voigt_evals_per_s": E / t32,
        "n_evals": E,
        "replay_stream_ms": stream_ms,
        "fp64_spectra_per_s": len(probs) / t64,
        "fp64_ms_per_step": t64 * 1e3,
        "fp32_over_fp64": t64 / t32,
        "alpha_max_rel_err_fp32_vs_fp64": alpha_err,
        "flux_max_abs_err_fp32_vs_fp64": flux_err,
        "accuracy_sample": f"every 16th star ({len(sample)} stars)",
        "continuum_on_grid": {"ms": tc * 1e3, "bytes_written": cbytes, "GB_per_s": cbytes / tc / 1e9,
                              "frac_of_hbm_peak": cbytes / tc / 8e12},
        "flux": {"ms": tf * 1e3, "spectra_per_s": len(probs) / tf,
                 "scheme": "Korg's planar default (anchored tau, linear source, E2), Float64"},
        "plan": plan.info(),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    ctx.close()


def bench_fit(args):
    """The fit step of 64 stars as one graph replay per step (forward Float32 + adjoint)."""
    import numpy as np
    from korg_line import Context, DeviceProblem, FitTarget, workloads
    from korg_line.hipmem import synchronize as device_sync

    n_stars = 64 if args.stars == 256 else args.stars
    t = time.time()
    probs, atms, ll, wls = workloads.fit_stars(n_stars)
    L = probs[0].n_layers
    log(f"{n_stars} stars built in {time.time() - t:.1f}s: {len(ll)} shared lines, "
        f"{probs[0].n_wl} points, {L} layers")
    ctx = Context(0)
    cache = args.tuning_cache or TUNING_CACHE
    dprobs = [DeviceProblem(P) for P in probs]
    a32 = [d.new_alpha("float32") for d in dprobs]
    tuning = ctx.autotune_batched(a32, dprobs, "fp32", cache_path=cache)
    refs = [workloads.fit_reference(a) for a in atms]
    rng = np.random.default_rng(7)
    targets = [FitTarget(P, tr, ar, np.ones(P.n_wl), 0.01) for P, (tr, ar) in zip(probs, refs)]
    # data: the model flux with lines 10% deeper, plus noise -- an iterate away from the optimum,
    # where the gradient is dominated by the systematic residuals (at the optimum it is a sum of
    # noise and any flux error is amplified by sqrt(points))
    ctx.fit_gradient_batched(a32, dprobs, targets, "fp32")
    for tg in targets:
        f = tg.flux()
        tg.set_data(1 - 1.1 * (1 - f) + 0.01 * rng.standard_normal(f.size))
    device_sync()
    t = time.time()
    plan = ctx.plan_fit(a32, dprobs, targets, "fp32")
    log(f"fit plan captured in {time.time() - t:.2f}s: {plan.info()}")
    for _ in range(args.warmup):
        plan.launch()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.launch()
    device_sync()
    tstep = (time.perf_counter() - t0) / args.steps
    log(f"  fit step (graph replay): {tstep * 1e3:.1f} ms ({n_stars / tstep:.1f} gradients/s)")
    g32 = [tg.grads() for tg in targets]
    c32 = np.array([tg.chi2() for tg in targets])
    t64 = gerr = cerr = None
    gerr_q = {}
    if args.fp64_steps > 0:
        # the same step with the Float64 forward (the all-Float64 chain)
        ctx.set_tuning()
        a64 = [d.new_alpha("float64") for d in dprobs]
        t64s = [FitTarget(P, tr, ar, tg.t["data"].numpy(), 0.01)
                for P, (tr, ar), tg in zip(probs, refs, targets)]
        t0 = time.perf_counter()
        ctx.fit_gradient_batched(a64, dprobs, t64s, "fp64")
        t64 = time.perf_counter() - t0
        c64 = np.array([tg.chi2() for tg in t64s])
        # per star and quantity: |g32 - g64| / |g64| (2-norms over layers; n/U over all species)
        for ga, gb in zip(g32, [tg.grads() for tg in t64s]):
            for q in ("T", "n_e", "xi", "n_HI", "n_div_U"):
                nb = np.linalg.norm(gb[q])
                if nb > 0:
                    gerr_q[q] = max(gerr_q.get(q, 0.0), float(np.linalg.norm(ga[q] - gb[q]) / nb))
        gerr = max(gerr_q.values())
        cerr = float(np.max(np.abs(c32 - c64) / np.abs(c64)))
        log(f"  Float64 step (direct): {t64 * 1e3:.1f} ms; gradient error {gerr:.2e} "
            f"({', '.join(f'{k} {v:.1e}' for k, v in gerr_q.items())}), chi2 {cerr:.2e}")
    cpu = None if args.no_cpu else cpu_baseline_fit(probs[0], refs[0], targets[0].t["data"].numpy())
    out = {
        "metric": METRIC,
        "value": n_stars / tstep,
        "unit": "gradient-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tstep * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "data": "synthetic",
        "config": {
            "workload": f"fit step: {n_stars} stars (Teff U[4000,6500], log g U[1,5], [M/H] "
                        "U[-2.5,0.5]), 5000-6000 A at 0.01 A (100,001 points), 100,000 shared "
                        "lines, 56 layers; per step: line opacity, radiative transfer (fp64), "
                        "chi^2, line-opacity adjoint, one HIP graph replay",
            "precision": PRECISION + "; adjoint: far zones by fp32 series moments on the matrix "
                         "cores (v_mfma_f32_16x16x4_f32, fp64 totals per 64 points), near zones "
                         "by exact fp32 evaluation of H and its partials, fp64 chain rule",
            "parallelism": "single GPU",
            "tuning": tuning,
        },
        "fp64_forward_step_ms": t64 * 1e3 if t64 is not None else None,
        "gradient_max_rel_err_vs_fp64_chain": gerr,
        "gradient_rel_err_by_quantity": gerr_q,
        "chi2_max_rel_err_vs_fp64_chain": cerr,
        "reference_gradient": "the all-Float64 device chain (Korg's ForwardDiff needs Julia, "
                              "which is absent); the adjoint itself is pinned by central "
                              "differences and the Float64 oracle",
        "plan": plan.info(),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    ctx.close()


def cpu_baseline_fit(P, ref, data):
    """The Float64 oracle chain (line opacity, transfer, its reverse mode, line adjoint) for one
    star of the fit workload."""
    import numpy as np
    from oracle import oracle
    threads = min(16, os.cpu_count() or 1)
    tau_ref, alpha_ref = ref
    t0 = time.perf_counter()
    al = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(al, P, nthreads=threads, mode=oracle.KORG_CHUNKS)
    ac = np.array([np.interp(P.all_wls, P.cntm_knots, P.cntm_values[l]) for l in range(P.n_layers)])
    F = oracle.rt_flux(ac + al, P.all_wls, P.temps, tau_ref, alpha_ref)
    Fc = oracle.rt_flux(ac, P.all_wls, P.temps, tau_ref, alpha_ref)
    r = (F / Fc - data) / 0.01
    G = oracle.rt_flux_vjp(ac + al, P.all_wls, P.temps, tau_ref, alpha_ref, 2 * r / 0.01 / Fc)
    oracle.line_absorption_vjp(np.asfortranarray(G), P, nthreads=threads)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: one star's gradient in {dt:.1f}s on {threads} threads")
    return {"value": 1.0 / dt, "unit": "gradient-evals/s", "cores": threads, "kind": "port",
            "sample": "one star of the workload, whole: Float64 oracle line opacity (reference "
                      "thread schedule), transfer and its reverse mode, line adjoint"}


def cpu_baseline_batched(stars, evals_per_star):
    """The Float64 CPU restatement on whole stars, smallest first, for about 20 s of work; the
    rate in evaluations per second gives spectra/s at the batch's mean evaluations per star."""
    import numpy as np
    from oracle import oracle
    threads = min(16, os.cpu_count() or 1)
    order = np.argsort([len(P.lines) * P.n_wl for P, _, _ in stars])
    done_evals, dt, n = 0.0, 0.0, 0
    for k in order:
        P = stars[k][0]
        a = np.zeros((P.n_layers, P.n_wl), order="F")
        t0 = time.perf_counter()
        _, _, e = oracle.line_absorption(a, P, nthreads=threads, mode=oracle.KORG_CHUNKS,
                                         return_windows=True)
        dt += time.perf_counter() - t0
        done_evals += e
        n += 1
        if dt > 20.0:
            break
    rate = done_evals / dt
    log(f"cpu baseline: {n} stars ({done_evals:.3g} evals) in {dt:.1f}s on {threads} threads")
    return {"value": rate / evals_per_star, "unit": "spectra/s", "cores": threads, "kind": "port",
            "evals_per_s": rate,
            "sample": f"the {n} smallest stars of the batch, whole, Float64, reference thread "
                      f"schedule; spectra/s = evaluations/s / mean evaluations per star"}


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
    """HBM bytes per launch of the solar_wide kernels, from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE summary (scaled by a device copy's readings)."""
    p = os.path.join(ROOT, "profiles", "traffic_solar_wide.json")
    try:
        return {k: v["bytes_per_launch"] for k, v in json.load(open(p))["kernels"].items()}
    except (OSError, KeyError, ValueError):
        return {}


def end_to_end_host(ctx, P, t_flatten, reps=3):
    """The call Korg makes (src/synthesize.jl:305-306 into korg_line_absorption): host Float64
    alpha and host inputs, staged to the device, the Float32 form, alpha copied back -- the
    PCIe-inclusive rate, next to the host-side flattening of the arguments."""
    import numpy as np
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    ctx.line_absorption(a, P, "fp32")  # warm-up (staging buffers grown)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.line_absorption(a, P, "fp32")
    ms = (time.perf_counter() - t0) / reps * 1e3
    log(f"end to end (host f64 alpha, copies included): {ms:.1f} ms per call")
    return {"ms": ms, "spectra_per_s": 1e3 / ms, "flatten_ms": t_flatten * 1e3,
            "alpha_bytes_each_way": 8 * P.n_layers * P.n_wl,
            "includes": "host->device copies of the inputs and of Float64 alpha, the Float32 "
                        "pass, device->host copy of alpha (korg_line_absorption); flatten_ms: "
                        "the host building of the flattened arguments (line list defaults, n/U, "
                        "continuum knots), once"}


def cpu_baseline(P, fraction):
    """The Float64 CPU restatement with the reference's threading (line chunks, private alpha
    per thread, summed) on a sample of every k-th line, scaled to one full spectrum by the exact
    evaluation counts."""
    import numpy as np
    from oracle import oracle
    from korg_line.problem import Problem

    k = max(1, int(round(1 / fraction)))
    idx = np.arange(0, len(P.lines), k)
    S = P if k == 1 else Problem(P.all_wls, P.lines.take(idx), P.species_mass, P.species_is_mol,
                                 P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi,
                                 P.cntm_mode, P.cntm_values, P.cntm_knots, P.cutoff_threshold)
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
    sample = ("the whole workload (300,000 lines, full grid), Float64, reference thread "
              "schedule" if k == 1 else
              f"every {k}th line of the same 300,000-line list ({len(idx)} lines, "
              f"{e_sample:.3g} of {e_full:.3g} evaluations), full grid, Float64, reference "
              f"thread schedule; time scaled by the evaluation ratio")
    return {"value": 1.0 / t_full, "unit": "spectra/s", "cores": threads, "kind": "port",
            "evals_per_s": e_sample / dt, "seconds": dt, "sample": sample}


if __name__ == "__main__":
    main()
