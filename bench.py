"""Headline benchmark: kernel pair-evaluations per second of the fused Convolve +
ConvolveGradient(X, alpha) on one MI355X.

Workload (raw kernel sums at scale): N = M = 1,000,000 points in 3D, targets X and
sources Y i.i.d. uniform in [0, 1]^3, weights W (M x 3) and alpha (N x 3) i.i.d. N(0, 1)
(separate seeds), sigma = 0.05.  One step = one call of the FP32-compensated fused
pair-sum (Morton sort of the targets + the pair-sum kernel) with every input already
resident in HBM; value = N*M / (seconds per step).  FP64 inputs and outputs, as at the
plugin boundary.

Beside the headline the run reports:
  * the same step in plain FP64 (the reference's arithmetic on the GPU) and the speed-up;
  * the FP32-compensated result's error against the FP64 GPU result over all N rows;
  * `roofline`: the pair-sum kernel's duration from HIP events on its stream,
    29 algorithmic flops per pair (d=3, Dv=3) against the FP32 peak of 157.3 TFLOP/s;
  * `cpu_baseline`: the FP64 oracle (a C restatement of ExactKernel) timed on a bounded
    sample of target rows against all M sources on the host's cores, whose outputs also
    give the FP32-compensated path's error against the FP64 oracle on those rows.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--m M]
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
    return ap.parse_args()


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
        tot, cnt = ctypes.c_double(), ctypes.c_long()
        L.dfm_profile_enable(1)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(prec)
        stream.synchronize()
        t1 = time.perf_counter()
        L.dfm_profile_enable(0)
        _lib.check(L.dfm_profile_read(ctypes.byref(tot), ctypes.byref(cnt)))
        return (t1 - t0) / steps, tot.value / 1e3 / max(cnt.value, 1)

    # ---- FP32-compensated (headline)
    for _ in range(args.warmup):
        step("fp32c")
    s_per_step, kern_s = timed("fp32c", args.steps)
    c32 = conv.numpy(stream)
    g32 = grad.numpy(stream)
    print("fp32c: %.4f s/step, kernel %.4f s" % (s_per_step, kern_s), file=sys.stderr, flush=True)

    # ---- FP32 with exact differences in every wavefront
    step("fp32x")
    s_x, kern_x = timed("fp32x", max(1, args.steps // 4))
    cx = conv.numpy(stream)
    gx = grad.numpy(stream)
    print("fp32x: %.4f s/step, kernel %.4f s" % (s_x, kern_x), file=sys.stderr, flush=True)

    # ---- plain FP64 form on the GPU
    step("fp64")
    s64, kern64 = timed("fp64", args.fp64_steps)
    c64 = conv.numpy(stream)
    g64 = grad.numpy(stream)
    print("fp64: %.4f s/step" % s64, file=sys.stderr, flush=True)

    def nrel(ref, got):
        return float(np.linalg.norm(ref - got) / np.linalg.norm(ref))

    def maxrms(ref, got):
        """largest element error relative to max(|element|, RMS of the output)"""
        scale = np.maximum(np.abs(ref), np.sqrt(np.mean(ref ** 2)))
        return float(np.max(np.abs(ref - got) / scale))

    err_fp64gpu = {
        "rows": N,
        "conv_normwise": nrel(c64, c32), "grad_normwise": nrel(g64, g32),
        "conv_max_over_rms": maxrms(c64, c32), "grad_max_over_rms": maxrms(g64, g32),
        "fp32x_conv_normwise": nrel(c64, cx), "fp32x_grad_normwise": nrel(g64, gx),
        "fp32x_conv_max_over_rms": maxrms(c64, cx), "fp32x_grad_max_over_rms": maxrms(g64, gx),
    }

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
        rows = int(max(probe, min(N, args.cpu_seconds / max(per_row, 1e-9))))
        rows = (rows // threads) * threads or probe
        idx = np.random.default_rng(7).choice(N, size=rows, replace=False)
        idx.sort()
        ta = time.perf_counter()
        oc, og = oracle.convolve_grad1_threaded(X[idx], A[idx], Y, W, sigma, threads)
        tb = time.perf_counter()
        cpu = {
            "value": rows * M / (tb - ta), "unit": "pair-evals/s", "cores": threads,
            "kind": "port",
            "sample": "%d random target rows x all %d sources, FP64 Convolve then "
                      "ConvolveGradient(X, alpha) (restated ExactKernel loops, libm exp), "
                      "%.1f s on %d threads" % (rows, M, tb - ta, threads),
        }
        err_oracle = {
            "rows": rows,
            "conv_normwise": nrel(oc, c32[idx]), "grad_normwise": nrel(og, g32[idx]),
            "conv_max_over_rms": maxrms(oc, c32[idx]), "grad_max_over_rms": maxrms(og, g32[idx]),
            "fp64_gpu_conv_normwise": nrel(oc, c64[idx]),
            "fp64_gpu_grad_normwise": nrel(og, g64[idx]),
        }

    pairs = float(N) * float(M)
    achieved = FLOPS_PER_PAIR * pairs / kern_s
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
                        "FP32 exp/FMA with FP64 re-centring and FP64 tile folds" % (N, sigma),
            "N": N, "M": M, "d": 3, "dv": 3, "sigma": sigma, "precision": "fp32c",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK,
            "traffic": None,
            "kernel": "k_pairsum<float, CONVGRAD1, 3, 3, 0, 4>",
            "kernel_ms": kern_s * 1e3,
            "flops_per_pair": FLOPS_PER_PAIR,
        },
        "cpu_baseline": cpu,
        "fp64_plain_gpu": {"s_per_step": s64, "kernel_s": kern64, "pair_evals_per_s": pairs / s64,
                           "speedup_fp32c_over_fp64": s64 / s_per_step},
        "fp32x_exact_differences": {"s_per_step": s_x, "kernel_s": kern_x,
                                    "pair_evals_per_s": pairs / s_x},
        "error_vs_fp64_gpu": err_fp64gpu,
        "error_vs_fp64_oracle": err_oracle,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
