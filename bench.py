"""Headline benchmark: kernel pair-evaluations per second of the fused Convolve +
ConvolveGradient(X, alpha) on one MI355X.

Workload (raw kernel sums at scale): N = M = 1,000,000 points in 3D, targets X and
sources Y i.i.d. uniform in [0, 1]^3, weights W (M x 3) and alpha (N x 3) i.i.d. N(0, 1)
(separate seeds), sigma = 0.05.  One step = one call of the fused pair-sum in the headline
precision (default bf16mfma: the distances' cross term on the matrix cores, FP32 exp2 and
accumulation with FP64 folds; Hilbert sort of the targets + the pair-sum kernel) with every
input already resident in HBM; value = N*M / (seconds per step).  FP64 inputs and outputs,
as at the plugin boundary.

Beside the headline the run reports:
  * the same step in the other FP32 forms (fp32c: FP32 VALU distances; fp32x: exact FP32
    differences everywhere) and in plain FP64 (the reference's arithmetic on the GPU);
  * each FP32 form's error against the FP64 GPU result over all N rows, and
    `error_bound_met`: the headline's norm-wise error <= 1e-6;
  * `roofline`: the pair-sum kernel's duration from HIP events on its stream,
    29 algorithmic flops per pair (d=3, Dv=3) against the FP32 peak of 157.3 TFLOP/s;
  * `cpu_baseline`: the FP64 oracle (a C restatement of ExactKernel) timed on a bounded
    sample of target rows against all M sources on the host's cores, whose outputs also
    give each FP32 form's error against the FP64 oracle on those rows.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--m M] [--precision P]

Other workloads (one JSON line each, their own metric; the default is the headline):
    --workload shooting      landmark geodesic shooting, 1,000 control points, sigma 0.2,
                             10 steps, Euler and RK2, GPU FP64/FP32C vs the CPU oracle
    --workload registration  varifold registration gradient: icosphere template with
                             20,480 triangles (radius 1), bump-displaced target, 4,096
                             control points on a 16^3 grid, 20 steps, checkpoints every 5,
                             adjoint gradient vs central differences (eps 1e-4);
                             with --trace-only, W + K gradients in --precision alone
    --workload atlas         batched atlas gradient: 64 subjects of 2,000-40,000 triangles,
                             2,048 shared control points, eager and HIP-graph replay
    --workload varifold      varifold match + gradient between two sets of 1,000,000
                             oriented sphere samples, sigma 0.02, BF16MFMA and FP32C vs FP64
    --workload curves        polyline varifold match + gradient between two closed 2D
                             curves of 1,000,000 segments, sigma 0.02, FP32C vs FP64
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

FLOPS_PER_PAIR = 29           # fused Convolve (Dv=3) + ConvolveGradient(X, alpha), d = 3
FP32_PEAK = 157.3e12          # MI355X FP32 vector peak, FLOP/s
KERNELS = {
    "bf16mfma": "k_pairsum_bf<CONVGRAD1, 3, 3, 4> + k_reduce_splits",
    "fp32c": "k_pairsum<float, CONVGRAD1, 3, 3, 0, 2, ADAPTIVE> + k_reduce_splits",
    "fp32x": "k_pairsum<float, CONVGRAD1, 3, 3, 0, 2, EXACT> + k_reduce_splits",
}
FORMS = {
    "bf16mfma": "squared distances' cross term -2 x.y + |y|^2 on the matrix cores "
                "(v_mfma_f32_16x16x32_bf16 over exact three-part BF16 splits of the FP32 "
                "group-centred coordinates, FP32 accumulation), target norm applied exactly, "
                "FP32 exp2 and accumulation with FP64 folds per 512 sources",
    "fp32c": "FP32 exp2/FMA on FP64-re-centred coordinates (expanded distance), FP64 folds "
             "per 64-source tile, exact differences in wide wavefronts",
    "fp32x": "FP32 exp2/FMA with exact (hi + lo) coordinate differences in every wavefront",
}
# HBM bytes per launch of the pair-sum dispatch at N = M = 1M, sigma 0.05: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE, each scaled by the same counter on a 24 MB device copy.
# Counter runs cannot be made from inside this process, so the figures are the profiled ones.
PROFILED = {
    # k_pairsum_bf: FETCH_SIZE 7,688,425 KiB x 1.997 (the copy kernel's 24 MB read counts
    # 11,735 KiB) = 15.72 GB, WRITE_SIZE 796,878 KiB x 1.0 = 0.82 GB (the FP64 split partials)
    "bf16mfma": {"traffic": 15.72e9 + 0.82e9,
                 "source": "profiles/k_pairsum_bf_convgrad1_N1M_sigma0.05.md",
                 "note": "VALU-issue bound with matrix-core co-issue: two "
                         "v_mfma_f32_16x16x32_bf16 per 256 pairs form the distances' cross "
                         "term and alpha.w, the VALU keeps exp2 and the FP32 accumulation "
                         "(8.8 VALU instructions per pair against 14.7 for fp32c); 16.5 GB per "
                         "launch is 69 GB/s, under 1% of HBM bandwidth; the peak is the FP32 "
                         "vector peak of the 29 algorithmic flops per pair"},
    # k_pairsum<float, ..., R=2>: FETCH_SIZE 2,920,758 KiB x 1.997 = 5.97 GB, WRITE_SIZE
    # 421,875 KiB = 0.43 GB
    "fp32c": {"traffic": 5.97e9 + 0.43e9,
              "source": "profiles/k_pairsum_convgrad1_N1M_sigma0.05.md",
              "note": "VALU-issue bound, no matrix-core work (14.7 VALU instructions per pair); "
                      "6.4 GB per launch is 21 GB/s, under 1% of HBM bandwidth"},
}
METRIC = "kernel pair-evals/s (Convolve+Gradient), FP32-comp vs FP64, at ≤1e-6 rel err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--m", type=int, default=1_000_000)
    ap.add_argument("--sigma", type=float, default=0.05)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate wall time for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="bf16mfma", choices=["bf16mfma", "fp32c", "fp32x"],
                    help="the headline leg's precision")
    ap.add_argument("--workload", default="kernel",
                    choices=["kernel", "shooting", "registration", "atlas", "varifold",
                             "curves"])
    ap.add_argument("--subjects", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true",
                    help="workload legs: skip the autotuned repeat (for kernel traces of the "
                         "default plans)")
    ap.add_argument("--trace-only", action="store_true",
                    help="registration leg: only --warmup untimed and --steps timed gradients "
                         "in --precision at the default plans, no other precision and no "
                         "checks (for kernel traces)")
    args = ap.parse_args()
    if args.steps < 1 or args.fp64_steps < 1 or args.warmup < 0:
        ap.error("--steps and --fp64-steps must be at least 1 (the FP64 step is the error "
                 "reference), --warmup at least 0")
    return args


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path (targets would be "
              "sharded across GPUs with the sources replicated); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import ctypes
    from deformetrica_code_amd import _lib, device

    if device.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload != "kernel":
        print(json.dumps(WORKLOADS[args.workload](args)), flush=True)
        return
    L = _lib.load()
    N, M, sigma = args.n, args.m, args.sigma

    rng_x = np.random.default_rng(20261203)
    rng_a = np.random.default_rng(20261204)
    X = rng_x.uniform(0.0, 1.0, size=(N, 3))
    Y = rng_x.uniform(0.0, 1.0, size=(M, 3))
    W = rng_x.normal(size=(M, 3))
    A = rng_a.normal(size=(N, 3))
    stream = device.Stream()
    dX, dY, dW, dA = (device.DeviceArray.from_numpy(a, stream) for a in (X, Y, W, A))
    conv = device.DeviceArray((N, 3))
    grad = device.DeviceArray((N, 3))
    ws = device.Workspace()
    # one plain device copy of known size (the sources, 24 MB): the reference kernel
    # (k_copy) for scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run
    copy_dst = device.DeviceArray((M, 3))
    _lib.check(L.dfm_memcpy(ctypes.c_void_p(copy_dst.ptr), ctypes.c_void_p(dY.ptr), M * 3 * 8, 2,
                            ctypes.c_void_p(stream.handle)))
    del copy_dst

    def step(prec):
        device.conv_grad1(dX, dA, dY, dW, sigma, prec, conv=conv, grad=grad, stream=stream,
                          workspace=ws)

    def timed(prec, steps):
        """Wall time per step over `steps` calls, and the pair-sum kernel's average duration
        from HIP events recorded on its stream."""
        stream.synchronize()
        L.dfm_profile_read(None, None)
        L.dfm_profile_work(None, None)
        tot, cnt = ctypes.c_double(), ctypes.c_long()
        L.dfm_profile_enable(1)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(prec)
        stream.synchronize()
        t1 = time.perf_counter()
        L.dfm_profile_enable(0)
        _lib.check(L.dfm_profile_read(ctypes.byref(tot), ctypes.byref(cnt)))
        wp, wf = ctypes.c_double(), ctypes.c_double()
        _lib.check(L.dfm_profile_work(ctypes.byref(wp), ctypes.byref(wf)))
        _WORK.update(pairs=wp.value / steps, flops=wf.value / steps)
        return (t1 - t0) / steps, tot.value / 1e3 / max(cnt.value, 1)

    # ---- the headline precision, then the other FP32 forms and plain FP64
    head = args.precision
    others = [p for p in ("fp32c", "bf16mfma", "fp32x") if p != head]
    for _ in range(args.warmup):
        step(head)
    s_per_step, kern_s = timed(head, args.steps)
    head_work = dict(_WORK)
    res = {head: (s_per_step, kern_s, conv.numpy(stream), grad.numpy(stream))}
    print("%s: %.4f s/step, kernel %.4f s" % (head, s_per_step, kern_s), file=sys.stderr,
          flush=True)
    for prec in others:
        step(prec)
        s_o, k_o = timed(prec, max(1, args.steps // 4))
        res[prec] = (s_o, k_o, conv.numpy(stream), grad.numpy(stream))
        print("%s: %.4f s/step, kernel %.4f s" % (prec, s_o, k_o), file=sys.stderr, flush=True)
    step("fp64")
    s64, kern64 = timed("fp64", args.fp64_steps)
    c64 = conv.numpy(stream)
    g64 = grad.numpy(stream)
    print("fp64: %.4f s/step" % s64, file=sys.stderr, flush=True)

    def nrel(ref, got):
        return float(np.linalg.norm(ref - got) / np.linalg.norm(ref))

    def maxrms(ref, got):
        """largest element error relative to the RMS of the output"""
        return float(np.max(np.abs(ref - got)) / np.sqrt(np.mean(ref ** 2)))

    def errors(cref, gref, rows=None):
        out = {}
        for prec in [head] + others:
            c, g = res[prec][2], res[prec][3]
            if rows is not None:
                c, g = c[rows], g[rows]
            key = "" if prec == head else prec + "_"
            out.update({key + "conv_normwise": nrel(cref, c), key + "grad_normwise": nrel(gref, g),
                        key + "conv_max_over_rms": maxrms(cref, c),
                        key + "grad_max_over_rms": maxrms(gref, g)})
        return out

    err_fp64gpu = {"rows": N, **errors(c64, g64)}

    # ---- CPU baseline: the FP64 oracle on a bounded sample of target rows
    cpu = None
    err_oracle = None
    if not args.no_cpu:
        from oracle import oracle
        threads = max(1, min(16, os.cpu_count() or 1))
        probe = threads * 2
        oracle.convolve_grad1_threaded(X[:threads], A[:threads], Y, W, sigma, threads)  # warm
        ta = time.perf_counter()
        oracle.convolve_grad1_threaded(X[:probe], A[:probe], Y, W, sigma, threads)
        tb = time.perf_counter()
        per_row = (tb - ta) / probe
        # at least 4,096 rows for the error check, more if the time budget allows
        rows = int(max(4096, probe, min(N, args.cpu_seconds / max(per_row, 1e-9))))
        rows = min(N, (rows // threads) * threads or probe)
        idx = np.random.default_rng(7).choice(N, size=rows, replace=False)
        idx.sort()
        ta = time.perf_counter()
        oc, og = oracle.convolve_grad1_threaded(X[idx], A[idx], Y, W, sigma, threads)
        tb = time.perf_counter()
        cpu = {
            "value": rows * M / (tb - ta), "unit": "pair-evals/s", "cores": threads,
            "kind": "port",
            "sample": "%d random target rows x all %d sources, FP64 Convolve then "
                      "ConvolveGradient(X, alpha) (restated ExactKernel loops, libm exp, "
                      "without the reference's per-pair heap allocations: faster than the "
                      "reference), %.1f s on %d threads" % (rows, M, tb - ta, threads),
            "scan": _cpu_scan(oracle, sigma, threads),
        }
        err_oracle = {"rows": rows, **errors(oc, og, idx),
                      "fp64_gpu_conv_normwise": nrel(oc, c64[idx]),
                      "fp64_gpu_grad_normwise": nrel(og, g64[idx])}
    within = (err_fp64gpu["conv_normwise"] <= 1e-6 and err_fp64gpu["grad_normwise"] <= 1e-6
              and (err_oracle is None or (err_oracle["conv_normwise"] <= 1e-6
                                          and err_oracle["grad_normwise"] <= 1e-6)))

    pairs = float(N) * float(M)
    achieved = FLOPS_PER_PAIR * pairs / kern_s
    profiled = PROFILED[head] if (N, M, sigma) == (1000000, 1000000, 0.05) else None
    out = {
        "metric": METRIC,
        "value": pairs / s_per_step,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "raw kernel sums: fused Convolve + ConvolveGradient(X, alpha), "
                        "N=M=%d points U[0,1]^3, d=3, Dv=3, W, alpha ~ N(0,1), sigma=%g; "
                        "%s" % (N, sigma, FORMS[head]),
            "N": N, "M": M, "d": 3, "dv": 3, "sigma": sigma, "precision": head,
            "parallelism": "single GPU",
        },
        "error_bound_met": within,
        "error_bound": "normwise relative error <= 1e-6 of Convolve and of the gradient, "
                       "against the FP64 GPU form (all rows) and the FP64 oracle (sampled rows)",
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK,
            "traffic": profiled["traffic"] if profiled else None,
            "traffic_unit": "bytes per launch",
            "traffic_source": profiled["source"] if profiled else None,
            "bound_note": profiled["note"] if profiled else None,
            "kernel": KERNELS[head],
            "kernel_ms": kern_s * 1e3,
            "flops_per_pair": FLOPS_PER_PAIR,
            "library_count": {"pairs_per_step": head_work["pairs"],
                              "flops_per_step": head_work["flops"]},
        },
        "cpu_baseline": cpu,
        "fp64_plain_gpu": {"s_per_step": s64, "kernel_s": kern64, "pair_evals_per_s": pairs / s64,
                           "speedup_over_fp64": s64 / s_per_step},
        "error_vs_fp64_gpu": err_fp64gpu,
        "error_vs_fp64_oracle": err_oracle,
    }
    for prec in others:
        out[prec] = {"s_per_step": res[prec][0], "kernel_s": res[prec][1],
                     "pair_evals_per_s": pairs / res[prec][0], "kernel": KERNELS[prec],
                     "form": FORMS[prec]}
    print(json.dumps(out), flush=True)


def _cpu_scan(oracle, sigma, threads):
    """The oracle's fused sum at N = M in {2^12, 2^14, 2^16}, on one thread and on
    `threads` threads (2^16 on one thread is skipped: ~70 s), in pair-evals/s; 1e12-pair
    times are these rates extrapolated (cost proportional to N * M)."""
    rng = np.random.default_rng(20261205)
    out = []
    for n in (1 << 12, 1 << 14, 1 << 16):
        X = rng.uniform(0.0, 1.0, size=(n, 3))
        W = rng.normal(size=(n, 3))
        A = rng.normal(size=(n, 3))
        for t in (1, threads):
            if n == 1 << 16 and t == 1:
                continue
            ta = time.perf_counter()
            oracle.convolve_grad1_threaded(X, A, X, W, sigma, t)
            dt = time.perf_counter() - ta
            out.append({"n": n, "threads": t, "pair_evals_per_s": n * n / dt,
                        "extrapolated_1e12_pairs_s": 1e12 * dt / (n * n)})
    return out


# ---------------------------------------------------------------------------------------
# other workloads


def _line(metric, value, unit, hib, ms, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "higher_is_better": hib,
           "ms_per_step": ms, "vs_baseline": None, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def _nrel(ref, got):
    ref, got = np.asarray(ref), np.asarray(got)
    return float(np.linalg.norm(ref - got) / max(np.linalg.norm(ref), 1e-300))


_WORK = {}  # algorithmic work per call of the last _profiled run: pairs, flops


def _profiled(L, fn, reps):
    """Wall seconds per call (timed with profiling off: the per-launch HIP events are not
    part of the work), pair-sum kernel seconds per call (HIP events, a second pass of `reps`
    calls) and launches per call; the library's count of algorithmic pairs and flops per
    call goes to _WORK."""
    import ctypes
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    t1 = time.perf_counter()
    L.dfm_profile_read(None, None)
    L.dfm_profile_work(None, None)
    tot, cnt = ctypes.c_double(), ctypes.c_long()
    pairs, flops = ctypes.c_double(), ctypes.c_double()
    L.dfm_profile_enable(1)
    for _ in range(reps):
        fn()
    L.dfm_profile_enable(0)
    L.dfm_profile_read(ctypes.byref(tot), ctypes.byref(cnt))
    L.dfm_profile_work(ctypes.byref(pairs), ctypes.byref(flops))
    _WORK.update(pairs=pairs.value / reps, flops=flops.value / reps)
    return (t1 - t0) / reps, tot.value / 1e3 / reps, cnt.value // reps


def _roof(kernel_s, work=None):
    """Roofline of a workload's pair sums: the library's algorithmic pairs and flops per call
    (dfm_profile_work: sum over launches of pairs x dfm_flops_per_pair(op)) over the summed
    pair-sum kernel time, against the FP32 vector peak."""
    w = dict(work or _WORK)
    return {"bound": "valu", "pairs_per_call": w["pairs"], "flops_per_call": w["flops"],
            "flops_per_pair_mean": w["flops"] / max(w["pairs"], 1.0),
            "kernel_ms": kernel_s * 1e3, "pair_evals_per_s": w["pairs"] / kernel_s,
            "achieved": w["flops"] / kernel_s / 1e12, "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": w["flops"] / kernel_s / FP32_PEAK}


def _tuned(L, fn, reps, explore=64):
    """Explore plans with the autotuner over `explore` calls, then time `reps` calls with
    the tuned plans (autotuner back to heuristics afterwards)."""
    L.dfm_autotune_reset()
    L.dfm_autotune(1)
    for _ in range(explore):
        fn()
    L.dfm_autotune(2)
    try:
        return _profiled(L, fn, reps)
    finally:
        L.dfm_autotune(0)
        L.dfm_autotune_reset()


def workload_shooting(args):
    """Landmark geodesic shooting: 1,000 control points U[0,1]^3, momenta N(0, 0.01^2),
    sigma 0.2, 10 steps (11 time points); Euler (Diffeos::Shoot) and RK2 (Heun)."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.diffeos import HIPDiffeos
    from oracle import deformations as dfo
    rng = np.random.default_rng(20261205)
    n, sigma, T = 1000, 0.2, 11
    p0 = rng.uniform(size=(n, 3))
    a0 = rng.normal(scale=0.01, size=(n, 3))
    L = _lib.load()
    res = {}
    for scheme in ("euler", "heun"):
        ta = time.perf_counter()
        P, A = dfo.shoot(p0, a0, sigma, T, scheme=scheme)
        cpu_s = time.perf_counter() - ta
        row = {"cpu_oracle_ms": cpu_s * 1e3}
        for prec in ("fp64", "fp32c"):
            df = HIPDiffeos(3, prec)
            df.SetKernelWidth(sigma)
            df.SetNumberOfTimePoints(T)
            df.SetShootingScheme(scheme)
            df.SetStartPositions(p0)
            df.SetStartMomentas(a0)
            df.Update()
            wall, kern, launches = _profiled(L, df.Update, 50)
            roof = _roof(kern)
            twall, tkern, _ = (None, None, None) if args.no_tune else _tuned(L, df.Update, 50)
            row[prec] = {"ms": wall * 1e3, "pairsum_kernel_ms": kern * 1e3,
                         "pairsum_launches": launches, "roofline": roof,
                         "autotuned_ms": twall and twall * 1e3,
                         "autotuned_pairsum_kernel_ms": tkern and tkern * 1e3,
                         "rel_err_final_positions_vs_oracle":
                             _nrel(P[-1] - p0, df.GetTrajectoryPositions()[-1] - p0),
                         "rel_err_final_momenta_vs_oracle": _nrel(A[-1], df.GetTrajectoryMomentas()[-1])}
        res[scheme] = row
    e = res["euler"]
    return _line("geodesic shooting time (1,000 control points, 10 steps)", e["fp32c"]["ms"], "ms",
                 False, e["fp32c"]["ms"],
                 {"workload": "landmark shooting: 1,000 control points U[0,1]^3, momenta "
                              "N(0, 0.01^2), sigma 0.2, 10 steps, host arrays in and out"},
                 dtype="f32", schemes=res,
                 cpu_baseline={"value": e["cpu_oracle_ms"], "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "whole workload, Euler, FP64 oracle (one thread)"})


def _grid(lo, hi, counts):
    axes = [np.linspace(lo, hi, c) for c in counts]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)


def workload_registration(args):
    """Varifold registration gradient at the scale of one subject."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.atlas import HIPDeterministicAtlas
    from deformetrica_code_amd.device import DeviceArray
    from deformetrica_code_amd.meshes import bump_field, icosphere
    rng = np.random.default_rng(20261206)
    V, F = icosphere(5, 1.0)
    target = V + bump_field(V, 8, 0.1, 0.3, rng)
    cp = _grid(-1.5, 1.5, (16, 16, 16))
    mom = rng.normal(scale=0.01, size=(1, cp.shape[0], 3))
    L = _lib.load()
    kw = dict(kind="varifold", deformation_width=0.3, attachment_width=0.1, data_sigma_squared=1.0,
              number_of_time_points=21, checkpoint_every=5)
    if args.trace_only:
        A = HIPDeterministicAtlas(F, V.shape[0], cp.shape[0], [(target, F)],
                                  precision=args.precision, **kw)
        d = [DeviceArray.from_numpy(x, A.stream) for x in (cp, mom, V)]
        for _ in range(args.warmup):
            A.gradient_device(*d)
        A.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            A.gradient_device(*d)
        A.stream.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        return _line("registration gradient time (varifold, 20,480 triangles, 4,096 control "
                     "points, 20 steps)", ms, "ms", False, ms,
                     {"workload": "registration leg, --trace-only", "precision": args.precision},
                     dtype="f32", steps=args.steps, warmup=args.warmup)
    out = {}
    for prec in ("fp32c", "bf16mfma", "fp64"):
        A = HIPDeterministicAtlas(F, V.shape[0], cp.shape[0], [(target, F)], precision=prec, **kw)
        dcp, dmom, dV = (DeviceArray.from_numpy(x, A.stream) for x in (cp, mom, V))
        run = lambda: (A.gradient_device(dcp, dmom, dV), A.stream.synchronize())
        run()
        wall, kern, launches = _profiled(L, run, 1 if prec == "fp64" else 3)
        roof = _roof(kern)
        tuned, graph_ms, table = (None, None), None, None
        if prec != "fp64" and not args.no_tune:
            # explore plans, time the tuned eager gradient, then capture it and replay
            L.dfm_autotune_reset()
            L.dfm_autotune(1)
            for _ in range(64):
                run()
            L.dfm_autotune(2)
            tuned = _profiled(L, run, 3)[:2]
            table = _lib.autotune_report().splitlines()
            g = A.capture(dcp, dmom, dV)
            g.launch()
            A.stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                g.launch()
            A.stream.synchronize()
            graph_ms = (time.perf_counter() - t0) / 5 * 1e3
            L.dfm_autotune(0)
            L.dfm_autotune_reset()
        out[prec] = {"ms": wall * 1e3, "pairsum_kernel_ms": kern * 1e3, "pairsum_launches": launches,
                     "roofline": roof, "graph_ms": graph_ms, "autotune_table": table,
                     "autotuned_ms": tuned[0] and tuned[0] * 1e3,
                     "grads": [A.grad_momenta.numpy(A.stream), A.grad_cp.numpy(A.stream),
                               A.grad_template.numpy(A.stream)],
                     "atlas": A}
    # central differences of the FP64 log-likelihood along random directions (eps = 1e-4)
    A64 = out["fp64"]["atlas"]
    f = lambda c, m, v: float(np.sum(A64.ComputeLogLikelihood(c, m, v)[0]))
    fd_err = {}
    drng = np.random.default_rng(5)
    for name, g in zip(("momenta", "control_points", "template"), out["fp64"]["grads"]):
        D = drng.normal(size=g.shape)
        D /= np.linalg.norm(D)
        e = 1e-4

        def at(h):
            if name == "momenta":
                return cp, mom + h * D, V
            if name == "control_points":
                return cp + h * D, mom, V
            return cp, mom, V + h * D

        fd = (f(*at(e)) - f(*at(-e))) / (2 * e)
        fd_err[name] = abs(float(np.sum(g * D)) - fd) / max(abs(fd), 1e-300)
    errs = {prec: {k: _nrel(a, b) for k, a, b in zip(("momenta", "control_points", "template"),
                                                       out["fp64"]["grads"], out[prec]["grads"])}
            for prec in ("fp32c", "bf16mfma")}
    # the line reports the faster FP32 precision (default plans), the other beside it
    best = min(("fp32c", "bf16mfma"), key=lambda k: out[k]["ms"])
    other = "bf16mfma" if best == "fp32c" else "fp32c"
    r, o = out[best], out[other]
    return _line("registration gradient time (varifold, 20,480 triangles, 4,096 control points, 20 steps)",
                 r["ms"], "ms", False, r["ms"],
                 {"workload": "icosphere template 20,480 triangles radius 1, target displaced by 8 "
                              "Gaussian bumps (amplitude 0.1, width 0.3), 16^3 control grid over "
                              "[-1.5,1.5]^3, deformation sigma 0.3, varifold sigma 0.1, 21 time "
                              "points, checkpoints every 5",
                  "precision": best},
                 dtype="f32", pairsum_kernel_ms=r["pairsum_kernel_ms"],
                 pairsum_launches=r["pairsum_launches"], roofline=r["roofline"],
                 fp64_roofline=out["fp64"]["roofline"], autotuned_ms=r["autotuned_ms"],
                 autotuned_graph_ms=r["graph_ms"], autotune_table=r["autotune_table"],
                 fp64_ms=out["fp64"]["ms"],
                 gradient_rel_err_vs_fp64=errs[best],
                 fp64_gradient_vs_central_differences=fd_err,
                 **{other: {"ms": o["ms"], "pairsum_kernel_ms": o["pairsum_kernel_ms"],
                            "roofline": o["roofline"], "autotuned_ms": o["autotuned_ms"],
                            "autotuned_graph_ms": o["graph_ms"],
                            "gradient_rel_err_vs_fp64": errs[other]}})


def workload_atlas(args):
    """Batched deterministic-atlas gradient over many subjects of uneven size, in DFM_FP32C
    and DFM_BF16MFMA (the batch's Convolve / Grad1 / fused sums with the distances' cross
    term on the matrix cores), each against an FP64 gradient of the same problem."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.atlas import HIPDeterministicAtlas
    from deformetrica_code_amd.device import DeviceArray
    from deformetrica_code_amd.meshes import bump_field, decimated_icosphere, icosphere
    S = args.subjects
    rng = np.random.default_rng(20261207)
    V, F = icosphere(5, 1.0)
    targets = []
    for s in range(S):
        tv, tf = decimated_icosphere(int(rng.integers(2000, 40001)), rng)
        targets.append((tv + bump_field(tv, 8, 0.1, 0.3, rng), tf))
    cp = _grid(-1.5, 1.5, (16, 16, 8))
    mom = rng.normal(scale=0.01, size=(S, cp.shape[0], 3))
    L = _lib.load()
    kw = dict(kind="varifold", deformation_width=0.25, attachment_width=0.1,
              data_sigma_squared=1.0, number_of_time_points=11)

    def grads(A):
        return [a.numpy(A.stream) for a in (A.grad_momenta, A.grad_cp, A.grad_template)]

    A64 = HIPDeterministicAtlas(F, V.shape[0], cp.shape[0], targets, precision="fp64",
                                checkpoint_every=5, **kw)
    d64 = [DeviceArray.from_numpy(x, A64.stream) for x in (cp, mom, V)]
    A64.gradient_device(*d64)
    A64.stream.synchronize()
    g64 = grads(A64)
    del A64, d64

    def measure(prec):
        A = HIPDeterministicAtlas(F, V.shape[0], cp.shape[0], targets, precision=prec,
                                  checkpoint_every=5, **kw)
        dcp, dmom, dV = (DeviceArray.from_numpy(x, A.stream) for x in (cp, mom, V))
        run = lambda: (A.gradient_device(dcp, dmom, dV), A.stream.synchronize())
        run()
        wall, kern, launches = _profiled(L, run, 3)
        roof = _roof(kern)
        err = {k: _nrel(ref, got) for k, got, ref in zip(("momenta", "control_points", "template"),
                                                         grads(A), g64)}
        # explore plans, then capture the gradient with the tuned plans
        L.dfm_autotune_reset()
        L.dfm_autotune(1)
        for _ in range(64):  # every candidate of every key timed twice (at most 30 candidates)
            run()
        L.dfm_autotune(2)
        twall, tkern, _ = _profiled(L, run, 3)
        troof = _roof(tkern)
        g = A.capture(dcp, dmom, dV)
        g.launch()
        A.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            g.launch()
        A.stream.synchronize()
        graph_ms = (time.perf_counter() - t0) / 3 * 1e3
        report = _lib.autotune_report()
        # the same gradient with the whole trajectory resident (checkpoint every step: nothing
        # recomputed in the backward), tuned plans, graph replay
        A1 = HIPDeterministicAtlas(F, V.shape[0], cp.shape[0], targets, precision=prec,
                                   checkpoint_every=1, **kw)
        g1 = A1.capture(dcp, dmom, dV)
        g1.launch()
        A1.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            g1.launch()
        A1.stream.synchronize()
        resident_ms = (time.perf_counter() - t0) / 3 * 1e3
        grad_diff = float(np.max(np.abs(A1.grad_momenta.numpy(A1.stream)
                                        - A.grad_momenta.numpy(A.stream))))
        L.dfm_autotune(0)
        L.dfm_autotune_reset()
        return {"graph_ms": graph_ms, "eager_ms": wall * 1e3, "pairsum_kernel_ms": kern * 1e3,
                "pairsum_launches": launches, "roofline": roof,
                "autotuned_eager_ms": twall * 1e3, "autotuned_pairsum_kernel_ms": tkern * 1e3,
                "autotuned_roofline": troof, "gradient_rel_err_vs_fp64": err,
                "autotune_table": report.splitlines(), "device_bytes": A.device_bytes(),
                "loglik": [float(v) for v in A.loglik.numpy(A.stream)],
                "resident_trajectory": {"checkpoint_every": 1, "graph_ms": resident_ms,
                                        "device_bytes": A1.device_bytes(),
                                        "max_abs_grad_momenta_diff_vs_k5": grad_diff}}

    res = {prec: measure(prec) for prec in ("fp32c", "bf16mfma")}
    best = min(res, key=lambda k: res[k]["graph_ms"])  # the line reports the faster precision
    b = res.pop(best)
    other = next(iter(res))
    faces = [t[1].shape[0] for t in targets]
    return _line("batched atlas gradient time (%d subjects)" % S, b["graph_ms"], "ms", False,
                 b["graph_ms"],
                 {"workload": "%d subjects, decimated icospheres of %d-%d triangles (mean %d) "
                              "displaced by 8 Gaussian bumps each, template icosphere 20,480 "
                              "triangles, 2,048 control points (16x16x8 grid), momenta N(0, 0.01^2), "
                              "deformation sigma 0.25, varifold sigma 0.1, 11 time points, "
                              "checkpoints every 5" % (S, min(faces), max(faces), int(np.mean(faces))),
                  "precision": best},
                 dtype="f32", graph="replay of the autotuned gradient", **b, **{other: res[other]})


def workload_varifold(args):
    """Varifold match + gradient between two sets of 1M oriented sphere samples."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.device import DeviceArray, Stream, Workspace
    from deformetrica_code_amd.meshes import sphere_samples
    from deformetrica_code_amd.surfaces import samples_match_gradient
    n = args.n
    rng = np.random.default_rng(20261208)
    cs, Ns = sphere_samples(n, rng)
    ct, Nt = sphere_samples(n, rng)
    L = _lib.load()
    st = Stream()
    ws = Workspace()
    d = [DeviceArray.from_numpy(v, st) for v in (cs, Ns, ct, Nt)]
    res = {}
    for prec, reps in (("bf16mfma", 2), ("fp32c", 2), ("fp64", 1)):
        run = lambda: (samples_match_gradient("varifold", 0.02, *d, precision=prec, stream=st,
                                              workspace=ws), st.synchronize())
        out = run()[0]
        wall, kern, launches = _profiled(L, run, reps)
        res[prec] = {"ms": wall * 1e3, "pairsum_kernel_ms": kern * 1e3, "launches": launches,
                     "roofline": _roof(kern),
                     "match": float(out[0].numpy(st)[0]), "grad_c": out[1].numpy(st),
                     "grad_n": out[2].numpy(st)}
    f64 = res["fp64"]

    def errs(r):
        return {"match_rel_err_vs_fp64": abs(r["match"] - f64["match"]) / abs(f64["match"]),
                "grad_centres_rel_err_vs_fp64": _nrel(f64["grad_c"], r["grad_c"]),
                "grad_normals_rel_err_vs_fp64": _nrel(f64["grad_n"], r["grad_n"])}

    r, c = res["bf16mfma"], res["fp32c"]
    pairs = r["roofline"]["pairs_per_call"]
    return _line("varifold data term pair-evals/s (1M x 1M samples, match + gradient)",
                 pairs / (r["ms"] / 1e3), "pair-evals/s", True, r["ms"],
                 {"workload": "two sets of %d oriented samples of the unit sphere (normals radial "
                              "x log-normal(-9, 0.3) areas), varifold kernel, sigma 0.02: "
                              "|S|^2 + |T|^2 - 2<S,T> and d/d centres, d/d normals; "
                              "DFM_BF16MFMA (distances' cross term on the matrix cores)" % n},
                 dtype="f32", roofline=r["roofline"], fp64_roofline=f64["roofline"],
                 fp64_ms=f64["ms"], speedup_over_fp64=f64["ms"] / r["ms"], **errs(r),
                 fp32c={"ms": c["ms"], "pair_evals_per_s": pairs / (c["ms"] / 1e3),
                        "roofline": c["roofline"], **errs(c)})


def workload_curves(args):
    """Polyline varifold match + gradient between two closed 2D curves of 1M segments each
    (NonOrientedPolyLine::ComputeMatch / ComputeMatchGradient at scale)."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.curves import cell_incidence, curve_match, curve_match_gradient
    from deformetrica_code_amd.device import DeviceArray, Stream, Workspace
    n = args.n
    t = np.linspace(0.0, 2 * np.pi, n, endpoint=False)
    rs = 1.0 + 0.2 * np.sin(5 * t)
    rt = 1.05 + 0.15 * np.sin(5 * t + 0.3) + 0.02 * np.cos(17 * t)
    VS = np.stack([rs * np.cos(t), rs * np.sin(t)], axis=1)
    VT = np.stack([rt * np.cos(t + 0.01), rt * np.sin(t + 0.01)], axis=1)
    E = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1).astype(np.int32)
    off, inc = cell_incidence(E, n)
    L = _lib.load()
    st = Stream()
    ws = Workspace()
    dV, dW = DeviceArray.from_numpy(VS, st), DeviceArray.from_numpy(VT, st)
    dE = DeviceArray.from_numpy(E, st, dtype=np.int32)
    dOff = DeviceArray.from_numpy(off, st, dtype=np.int32)
    dInc = DeviceArray.from_numpy(inc, st, dtype=np.int32)
    sigma = 0.02
    res = {}
    for prec, reps in (("fp32c", 2), ("fp32x", 1), ("fp64", 1)):
        def run():
            m = curve_match("varifold", sigma, dV, dE, dW, dE, prec, st, ws)
            g = curve_match_gradient("varifold", sigma, dV, dE, dOff, dInc, dW, dE, prec, st, ws)
            st.synchronize()
            return m, g
        m, g = run()
        wall, kern, launches = _profiled(L, run, reps)
        res[prec] = {"ms": wall * 1e3, "pairsum_kernel_ms": kern * 1e3, "launches": launches,
                     "roofline": _roof(kern), "match": float(m.numpy(st)[0]), "grad": g.numpy(st)}
    r = res["fp32c"]
    pairs = r["roofline"]["pairs_per_call"]
    return _line("polyline varifold pair-evals/s (2D, 1M x 1M segments, match + gradient)",
                 pairs / (r["ms"] / 1e3), "pair-evals/s", True, r["ms"],
                 {"workload": "two closed 2D polylines of %d segments each (radius 1 + 0.2 "
                              "sin 5t and a perturbed copy), NonOrientedPolyLine varifold, "
                              "sigma 0.02: match and d match / d points" % n},
                 dtype="f32", roofline=r["roofline"],
                 fp64_ms=res["fp64"]["ms"], speedup_fp32c_over_fp64=res["fp64"]["ms"] / r["ms"],
                 match_rel_err_vs_fp64=abs(r["match"] - res["fp64"]["match"]) / abs(res["fp64"]["match"]),
                 grad_rel_err_vs_fp64=_nrel(res["fp64"]["grad"], r["grad"]),
                 fp32x={"ms": res["fp32x"]["ms"],
                        "match_rel_err_vs_fp64": abs(res["fp32x"]["match"] - res["fp64"]["match"])
                        / abs(res["fp64"]["match"]),
                        "grad_rel_err_vs_fp64": _nrel(res["fp64"]["grad"], res["fp32x"]["grad"])},
                 fp32c_pairsum_kernel_ms=r["pairsum_kernel_ms"], launches=r["launches"],
                 note="the gradient at a point is the difference of its two segments' kernel "
                      "sums, which differ by about h/sigma (segment length over kernel width, "
                      "1/3000 here): the FP32 sums' ~1e-7 error reaches the gradient amplified "
                      "about sigma/h-fold, in FP32X too; the match is not affected")


WORKLOADS = {"shooting": workload_shooting, "registration": workload_registration,
             "atlas": workload_atlas, "varifold": workload_varifold, "curves": workload_curves}


if __name__ == "__main__":
    main()
