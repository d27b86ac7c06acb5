#!/usr/bin/env python3
"""Benchmark: MCTDHF right-hand-side evaluations per second on one MI355X.

One step = one full MCTDHX right-hand side (Method_TimeDerivative) of the 1D FEDVR MCTDHF
system with N_g = 4096 grid points (819 elements x 6 GLL points on [-200, 200]), N = 4
spin-polarised fermions in M = 8 orbitals (70 configurations), soft-Coulomb external
potential and interaction, seeded synthetic state resident in HBM.

Prints one JSON line (rank 0).  Extra fields: `roofline` for the dominant kernel (the
mean-field contraction W = V S, MFMA-bound), `cpu_baseline` (the FP64 oracle, a C restatement
of the reference algorithm, timed on this host's cores), and the per-quantity error of the
GPU result against that FP64 oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MCTDHF RHS evals/s (M=8, 4096 pts): split-FP32 MFMA vs FP64, rel err ≤1e-9"
PEAK = {"fp64": 78.6, "split_fp32": 157.3, "bf16x3": 2500.0}  # dense MFMA TFLOP/s, MI355X


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--workload", default="mctdhf_1d_4096")
    p.add_argument("--precision", default="fp64", choices=["fp64"])
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--profile-iters", type=int, default=50)
    return p.parse_args()


def cpu_baseline(wl, state, budget_s: float):
    """FP64 oracle (C + OpenMP restatement of the reference) on this host: evals/s of the RHS."""
    import oracle

    bdir = tempfile.mkdtemp(prefix="cody_oracle_")
    try:
        so = oracle.build(native=True, build_dir=bdir)
    except Exception:
        so = None
    lib_path = so if so and os.path.exists(so) else None
    lib = oracle.load(lib_path) if lib_path else oracle.load()
    oracle._LIB = lib
    from tests.common import oracle_system

    s, x, wts = oracle_system(wl)
    d = s.time_derivative(state)  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        s.time_derivative(state)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 200:
            break
    threads = lib.orc_num_threads()
    return {
        "value": n / el,
        "unit": "RHS evals/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"{n} full RHS evaluations of the same {wl.nG}-point, M={wl.M}, N={wl.N} state "
                  f"({el:.1f} s, OpenMP threads={threads}, no BLAS; built -O3 -march={'native' if lib_path else 'x86-64-v3'})",
    }, d, s, wts


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent MCTDHF instances shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)
    import torch

    from codyfortran_amd import workloads as wlm

    wl = wlm.WORKLOADS[args.workload]
    ctx, x, w = wlm.make_context(wl)
    state = wlm.seeded_state(x, w, wl.N, wl.M, seed=1234)
    dev = torch.device("cuda:0")
    d_state = torch.from_numpy(state.view(np.float64).copy()).to(dev)
    d_out = torch.zeros_like(d_state)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    for _ in range(args.warmup):
        ctx.time_derivative_dev(d_state.data_ptr(), d_out.data_ptr(), 0.0, sptr)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        ctx.time_derivative_dev(d_state.data_ptr(), d_out.data_ptr(), 0.0, sptr)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(stop)
    ms_per_step = ms_total / args.steps
    value = 1000.0 / ms_per_step

    # per-launch breakdown on the context's own stream (HIP events around each launch)
    ctx.upload_state(state)
    kernels = ctx.profile_rhs(args.profile_iters)
    mf_ms = kernels.get("k_meanfield_f64")
    flops = 2.0 * wl.nG ** 2 * 2 * wl.pairs  # pair symmetry exploited: 2 N_g^2 2P
    achieved = flops / (mf_ms * 1e-3) / 1e12 if mf_ms else None
    roof = {
        "bound": "mfma",
        "kernel": "k_meanfield_f64",
        "achieved": achieved,
        "peak": PEAK[args.precision],
        "unit": "TFLOP/s",
        "frac": achieved / PEAK[args.precision] if achieved else None,
        "traffic": None,
        "flop_per_launch": flops,
        "avg_launch_ms": mf_ms,
    }

    gpu_d = d_out.cpu().numpy().view(np.complex128)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else args.precision,
        "data": "synthetic: seeded complex Gaussian orbitals x exp(-x^2/8), MGS-orthonormalised in "
                "the FEDVR weights; complex Gaussian CI vector, normalised",
        "config": {"workload": wl.name, "nG": wl.nG, "elements": wl.nE, "gll_points": wl.nLoc,
                   "N": wl.N, "M": wl.M, "nC": wl.nC, "pairs": wl.pairs, "precision": args.precision,
                   "regularization": wl.reg},
        "roofline": roof,
        "kernels_ms": kernels,
        "wall_s_timed": wall,
    }
    if not args.no_cpu_baseline:
        cb, d_orc, s_orc, wts = cpu_baseline(wl, state, args.cpu_seconds)
        result["cpu_baseline"] = cb
        from tests.common import state_rel

        result["rel_err_vs_fp64_oracle"] = {"rhs": state_rel(gpu_d, d_orc, wl.nC, wts, wl.M)}
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
