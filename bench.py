#!/usr/bin/env python3
"""Benchmark: MCTDHF right-hand-side evaluations per second on one MI355X.

One step = one full MCTDHX right-hand side (Method_TimeDerivative) of the 1D FEDVR MCTDHF
system with N_g = 4096 grid points (819 elements x 6 GLL points on [-200, 200]), N = 4
spin-polarised fermions in M = 8 orbitals (70 configurations), soft-Coulomb external
potential and interaction, seeded synthetic state resident in HBM.

Prints one JSON line (rank 0).  Extra fields: `roofline` for the dominant kernel (the
mean-field contraction W = V S, MFMA-bound), `cpu_baseline` (the FP64 oracle, a C restatement
of the reference algorithm, timed on this host's cores), and the per-quantity error (mean fields
W, right-hand side, energy) of every form against that FP64 oracle.  The headline is the fastest
form whose errors are all <= 1e-9.
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
# the CPU baseline's OpenMP threads stay on neighbouring cores (read when the runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
sys.path.insert(0, ROOT)

METRIC = "MCTDHF RHS evals/s (M=8, 4096 pts): split-FP32 MFMA vs FP64, rel err ≤1e-9"
# The bound of each contraction form: FP64 and split-FP32 are MFMA-bound (useful FLOP against the
# dense FP64 / FP32 MFMA peak); the 3xBF16 and int8-sliced forms do their products faster than
# HBM delivers their operand bytes (int8: 15 digit products of 2 N_g^2 Pc ops take ~8 us at
# 5.0 POP/s, the 84 MB of V digits ~10.5 us at 8 TB/s), so they are HBM-bound: algorithmic bytes
# = V and S operands read once plus W^T written once, against 8 TB/s.
BOUND = {"fp64": "mfma", "split_fp32": "mfma", "bf16x3": "hbm", "int8_sliced": "hbm", "bf16_sliced": "hbm"}
PEAK = {"fp64": 78.6, "split_fp32": 157.3, "bf16x3": 8000.0, "int8_sliced": 8000.0,
        "bf16_sliced": 8000.0}  # TFLOP/s or GB/s
PREC_CODE = {"fp64": 0, "split_fp32": 1, "bf16x3": 2, "int8_sliced": 3, "bf16_sliced": 4}
# bytes per operand element of the HBM-bound forms: three BF16 terms; five int8 digits (the
# BF16-sliced form reads the same digits and widens them to BF16 in registers)
OPERAND_BYTES = {"bf16x3": 6.0, "int8_sliced": 5.0, "bf16_sliced": 5.0}
# (the int8-sliced form's right-hand side runs its contraction with the orbital correlation of each
# 16-point tile in the epilogue: k_mf_i8_orb)
KERNEL = {"fp64": "k_mf_f64_w", "split_fp32": "k_mf_f32x3", "bf16x3": "k_mf_bf16x3", "int8_sliced": "k_mf_i8_orb",
          "bf16_sliced": "k_mf_i8_orb<bf16>"}
TOL = 1e-9
# forms whose contraction skips the k blocks of all-zero pair sources (cody_set_source_skip)
SKIPPING = ("fp64", "int8_sliced", "bf16_sliced")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--workload", default="mctdhf_1d_4096")
    p.add_argument("--precision", default="auto",
                   choices=["auto", "fp64", "int8_sliced", "bf16_sliced", "split_fp32", "bf16x3"],
                   help="auto: time every form, report the fastest whose W, RHS and energy errors "
                        "vs the FP64 oracle are all <= 1e-9")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--profile-iters", type=int, default=50)
    p.add_argument("--batch", type=int, default=64,
                   help="independent instances of the batched measurement (0: skip it)")
    p.add_argument("--batch-steps", type=int, default=20)
    p.add_argument("--batch-int8-only", action="store_true",
                   help="time the batch with the int8-sliced contraction only (a profiling run of that form)")
    p.add_argument("--prop-steps", type=int, default=1000,
                   help="steps of each propagation measurement (0: skip them)")
    p.add_argument("--adjoint-steps", type=int, default=1000,
                   help="steps of the laser-gradient measurement (0: skip it)")
    p.add_argument("--adjoint-fd", type=int, default=4,
                   help="laser coefficients whose gradient is checked by central differences")
    p.add_argument("--adjoint-only", action="store_true",
                   help="run only the laser-gradient measurement (no finite-difference checks unless "
                        "--adjoint-fd > 0): the command of the adjoint kernel profile")
    p.add_argument("--conservation", action="store_true",
                   help="run only the conservation study (energy and norm drift against the rho^-1 "
                        "regularisation, dt and the SIL tolerance) and print it as one JSON line")
    p.add_argument("--pmc-probe", action="store_true",
                   help="short run for hardware-counter passes: a 256 MiB device copy (the unit "
                        "of the byte counters) and a few launches of the FP64 and int8 "
                        "contraction kernels, dense and with the zero-source skip; prints no JSON line")
    p.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                   help="turn the FETCH_SIZE and WRITE_SIZE passes over --pmc-probe (rocprofv3 output "
                        "folders) into profiles/pmc_traffic_<workload>.json and .md; needs no GPU")
    p.add_argument("--no-dense", action="store_true",
                   help="time the forms only with the zero-source skip (no dense repeat)")
    return p.parse_args()


def _propagate(ctx, wl, w, state, steps, dt, step_fn, chunk=50):
    """Runs `steps` steps of `step_fn(t0, t1) -> converged` from `state`, timing only the
    steps, in chunks of `chunk` with a check of the state between chunks; stops at the first
    chunk whose state is non-finite, whose CI norm left [0.5, 2] times its start value, or whose
    energy moved by more than its start value."""
    import numpy as np
    from codyfortran_amd.workloads import split_state

    ctx.upload_state(state)
    e0 = ctx.energy(state)
    n0 = float(np.linalg.norm(state[:wl.nC]))
    t, done, wall, ok = 0.0, 0, 0.0, True
    out = {"dt": dt, "energy_start": e0}
    while done < steps:
        k = min(chunk, steps - done)
        t0 = time.perf_counter()
        for _ in range(k):
            ok &= bool(step_fn(t, t + dt))
            t += dt
        wall += time.perf_counter() - t0
        done += k
        s1 = ctx.download_state()
        nrm = float(np.linalg.norm(s1[:wl.nC]))
        e1 = ctx.energy(s1) if np.all(np.isfinite(s1)) else float("nan")
        if not (np.isfinite(e1) and 0.5 < nrm / n0 < 2.0 and abs(e1 - e0) <= abs(e0)):
            out.update({"diverged": True, "diverged_within_steps": [done - k + 1, done],
                        "steps_run": done})
            return out
    _, orbs = split_state(s1, wl.nC, wl.nG, wl.M)
    ov = orbs.conj().T @ (orbs * w[:, None])
    out.update({"diverged": False, "steps_run": done, "ms_per_step": 1000.0 * wall / done,
                "all_converged": bool(ok), "ci_norm_drift": abs(nrm - n0),
                "orbital_orthonormality_error": float(np.max(np.abs(ov - np.eye(wl.M)))),
                "energy_end": e1, "energy_drift_rel": abs(e1 - e0) / abs(e0)})
    return out


def propagation_measurement(args, wlm, wl, state, prec_code):
    """The workload's propagation, device-resident between upload and download:
    - the reference's Strang split step A(dt/2) B(dt) A(dt/2) (SIL with Krylov dimension
      wl.krylov_dim in the FEDVR metric on the one-body part, RK4 on the rest) at the
      workload's dt, from the seeded state and from a smooth seeded state;
    - RK4 on the full right-hand side from the smooth state at dt = 0.002, below the explicit
      stability limit 2.83/lambda_max of the 4096-point FEDVR one-body operator
      (lambda_max ~ 8e2), which the Strang split does not need.
    The seeded state has i.i.d. grid values (energy 666): the RK4 part of the split is far
    outside its stability region at dt = 0.02 and the state diverges in the first step, on the
    GPU and in the oracle alike.  The reference split step's energy change does not shrink with
    dt: its nonlinear part evolves C with the full h1 without the -P h phi term of the orbitals,
    so A + B is not the full right-hand side.  The consistent split (B = full minus linear) is
    second order; full-RHS RK4 conserves the energy to the level set by the rho^-1
    regularisation."""
    ctx, x, w = wlm.make_context(wl)
    ctx.set_precision(prec_code)
    smooth = wlm.smooth_state(x, w, wl.N, wl.M, seed=1234)
    kd = wl.krylov_dim

    def strang(t0, t1):
        return ctx.strang(t0, t1, krylov_dim=kd, tol=1e-10, max_restarts=8, metric=True)

    def rk4(t0, t1):
        ctx.rk4(t0, t1, full=True)
        return True

    ctx.upload_state(smooth)
    strang(0.0, wl.dt)  # warm-up: buffers and graphs
    rk4(0.0, 0.002)
    scheme = "Strang: SIL(dt/2) RK4(dt, nonlinear part) SIL(dt/2)"
    out = {"steps": args.prop_steps, "krylov_dim": kd,
           "strang_seeded_state": dict(scheme=scheme, **_propagate(ctx, wl, w, state, args.prop_steps, wl.dt, strang)),
           "strang_smooth_state": dict(scheme=scheme, **_propagate(ctx, wl, w, smooth, args.prop_steps, wl.dt, strang))}
    # the consistent split (nonlinear part = full minus linear part): second order in dt
    ctx.set_split(True)
    strang(0.0, wl.dt)  # warm-up of the graphs of this split
    out["strang_consistent_smooth_state"] = dict(
        scheme="Strang: SIL(dt/2) RK4(dt, full minus linear part) SIL(dt/2)",
        **_propagate(ctx, wl, w, smooth, args.prop_steps, wl.dt, strang))
    ctx.set_split(False)
    out["rk4_full_rhs_smooth_state"] = dict(scheme="RK4 on the full RHS",
                                            **_propagate(ctx, wl, w, smooth, args.prop_steps, 0.002, rk4))
    ctx.close()
    return out


def adjoint_measurement(args, wlm, wl, checks=True):
    """Gradient of the final dipole <x>(T) in 32 laser coefficients: E(t) = sum_k c_k
    sin(k pi t / T) sin^2(pi t / T), T = 20, c_k ~ N(0, 0.01^2) (seeded), args.adjoint_steps
    Strang steps of dt = 0.02 from the smooth seeded state, consistent split, FP64; the
    linear half-steps are the fixed-degree Chebyshev expansion, checkpoints every 50 steps.
    Reported: forward and gradient wall time, the gradient's agreement with central differences
    (h = 1e-6) of the same forward map for the first args.adjoint_fd coefficients and along a
    random unit direction, and with a central difference (h = 1e-4) along the same direction of
    the Lanczos (SIL) Strang map."""
    n = args.adjoint_steps
    T = 20.0
    dt = 0.02  # a run of fewer than 1000 steps covers the start of the pulse
    rng = np.random.default_rng(2024)
    c = 0.01 * rng.standard_normal(32)
    ctx, x, w = wlm.make_context(wl)
    ctx.set_split(True)
    state = wlm.smooth_state(x, w, wl.N, wl.M, seed=1234)
    emax = float(np.sum(np.abs(c))) * 1.05
    lo, hi, K = ctx.set_chebyshev(dt, emax)

    def forward(cc, sil=False):
        ctx.set_laser(cc, T)
        ctx.upload_state(state)
        for m in range(n):
            if sil:
                ctx.strang(m * dt, (m + 1) * dt, krylov_dim=wl.krylov_dim, tol=1e-10, max_restarts=8, metric=True)
            else:
                ctx.strang_cheb(m * dt, (m + 1) * dt)
        return ctx.dipole(ctx.download_state())

    ctx.set_laser(c, T)
    ctx.upload_state(state)
    ctx.dipole_gradient(0.0, dt, 2, 50)  # warm-up: buffers, the segment tape and graphs
    t0 = time.perf_counter()
    J_fwd = forward(c)
    t_fwd = time.perf_counter() - t0
    ctx.set_laser(c, T)
    ctx.upload_state(state)
    t0 = time.perf_counter()
    J, g = ctx.dipole_gradient(0.0, dt, n, 50)
    t_grad = time.perf_counter() - t0
    # the same gradient with the whole tape kept in HBM by the forward pass (no recompute)
    ctx.upload_state(state)
    ctx.dipole_gradient(0.0, dt, n, n)  # allocates the n-step tape (a warm run)
    ctx.upload_state(state)
    t0 = time.perf_counter()
    J_tape, g_tape = ctx.dipole_gradient(0.0, dt, n, n)
    t_tape = time.perf_counter() - t0
    tape_info = {"gradient_s": t_tape, "gradient_over_forward": t_tape / t_fwd,
                 "max_abs_diff_vs_checkpointed_over_grad_norm": float(np.max(np.abs(g_tape - g)) / np.linalg.norm(g)),
                 "J_diff": abs(J_tape - J)}
    if not checks:
        ctx.close()
        return {"workload": wl.name, "steps": n, "dt": dt, "T": T, "coefficients": len(c), "checkpoint_every": 50,
                "split": "consistent", "chebyshev": {"degree": K, "lo": lo, "hi": hi}, "J": J,
                "grad_norm": float(np.linalg.norm(g)), "forward_s": t_fwd, "gradient_s": t_grad,
                "gradient_over_forward": t_grad / t_fwd, "full_tape_in_hbm": tape_info}
    h = 1e-6
    fd = []
    for k in range(min(args.adjoint_fd, len(c))):
        e = np.zeros(len(c))
        e[k] = h
        fd.append((forward(c + e) - forward(c - e)) / (2 * h))
    fd = np.array(fd)
    u = rng.standard_normal(len(c))
    u /= np.linalg.norm(u)
    d_fd = (forward(c + h * u) - forward(c - h * u)) / (2 * h)
    hs = 1e-4
    d_sil = (forward(c + hs * u, sil=True) - forward(c - hs * u, sil=True)) / (2 * hs)
    ctx.close()
    gn = float(np.linalg.norm(g))
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's stored-trajectory adjoint on the same map, a bounded sample of two steps
        import oracle
        from tests.common import oracle_system

        so, _, _ = oracle_system(wl)
        so.set_laser(c, T)
        m = 2
        t0 = time.perf_counter()
        so.dipole_gradient(state, 0.0, dt, m, lo, hi, K, consistent=True)
        el = time.perf_counter() - t0
        threads = int(oracle._LIB.orc_num_threads()) if oracle._LIB is not None else None
        cpu = {"value": el / m, "unit": "s per step of the gradient", "cores": threads, "kind": "port",
               "sample": f"forward and adjoint of {m} Chebyshev-Strang steps of the same system on the oracle "
                         f"(C + OpenMP, {threads} threads), {el:.1f} s"}
    return {
        "workload": wl.name, "steps": n, "dt": dt, "T": T, "coefficients": len(c), "checkpoint_every": 50,
        "split": "consistent", "chebyshev": {"degree": K, "lo": lo, "hi": hi},
        "J": J, "J_forward_only": J_fwd, "grad_norm": gn,
        "forward_s": t_fwd, "gradient_s": t_grad, "gradient_over_forward": t_grad / t_fwd,
        "fd_components": len(fd),
        "fd_max_abs_err_over_grad_norm": float(np.max(np.abs(g[: len(fd)] - fd)) / gn) if len(fd) else None,
        "directional_fd": {"adjoint": float(g @ u), "fd": d_fd, "rel_err": abs(float(g @ u) - d_fd) / gn},
        "directional_fd_sil_strang": {"fd": d_sil, "rel_err": abs(float(g @ u) - d_sil) / gn, "h": hs},
        "gpu_s_per_step": t_grad / n,
        "full_tape_in_hbm": tape_info,
        "cpu_baseline": cpu,
    }


def conservation_measurement(args, wlm, wl):
    """Where norm and energy conservation to 1e-10 hold at the workload's size, from the smooth
    seeded state over t = 2 (device-resident, FP64 contraction):
    - RK4 on the full right-hand side for regularisations eps = 1e-8, 1e-10, 1e-12 (the rho^-1
      regularisation of S_Method_Mb_OrbBased.f90:737-744 changes the equations of motion by
      O(eps), so the energy is conserved only up to a floor set by eps) and steps dt = 0.002,
      0.001, 0.0005 (below the explicit stability limit ~3.4e-3 of the 4096-point grid);
    - the consistent Strang split (SIL Krylov 16, tolerance 1e-12) for dt = 0.04 ... 0.005
      (second order in dt);
    - the consistent split at dt = 0.02 for SIL tolerances 1e-8, 1e-10, 1e-12 (the SIL acts on
      the orbitals only: its tolerance shows in their orthonormality, the CI norm is the RK4
      part's)."""
    import dataclasses

    T = 2.0
    out = {"t_end": T, "state": "smooth seeded state (seed 1234)", "rk4_full": [], "strang_consistent_dt": [],
           "strang_consistent_tol": []}

    def run(ctx, x, w, steps, dt, fn):
        smooth = wlm.smooth_state(x, w, wl.N, wl.M, seed=1234)
        return _propagate(ctx, wl, w, smooth, steps, dt, fn, chunk=max(50, steps // 4))

    for eps in (1e-8, 1e-10, 1e-12):
        ctx, x, w = wlm.make_context(dataclasses.replace(wl, reg=eps))
        for dt in (0.002, 0.001, 0.0005):
            def rk4(t0, t1):
                ctx.rk4(t0, t1, full=True)
                return True
            r = run(ctx, x, w, int(round(T / dt)), dt, rk4)
            out["rk4_full"].append({"eps": eps, **r})
            print(f"conservation: rk4 eps {eps:g} dt {dt:g}: energy drift {r.get('energy_drift_rel')}", flush=True)
        ctx.close()
    ctx, x, w = wlm.make_context(wl)
    ctx.set_split(True)
    for dt in (0.04, 0.02, 0.01, 0.005):
        r = run(ctx, x, w, int(round(T / dt)), dt,
                lambda t0, t1: ctx.strang(t0, t1, krylov_dim=16, tol=1e-12, max_restarts=8, metric=True))
        out["strang_consistent_dt"].append({"eps": wl.reg, "tol": 1e-12, **r})
        print(f"conservation: consistent split dt {dt:g}: energy drift {r.get('energy_drift_rel')}", flush=True)
    for tol in (1e-8, 1e-10, 1e-12):
        r = run(ctx, x, w, int(round(T / 0.02)), 0.02,
                lambda t0, t1: ctx.strang(t0, t1, krylov_dim=16, tol=tol, max_restarts=8, metric=True))
        out["strang_consistent_tol"].append({"eps": wl.reg, "tol": tol, **r})
        print(f"conservation: consistent split tol {tol:g}: orthonormality {r.get('orbital_orthonormality_error')}",
              flush=True)
    ctx.close()
    return out


def propagated_energy_by_form(args, wlm, wl, forms):
    """The propagated quantity of the precision comparison: the energy after args.prop_steps
    Strang steps (consistent split, dt = wl.dt, Krylov dimension wl.krylov_dim) from the smooth
    seeded state, for each contraction form, relative to the FP64 form's."""
    ctx, x, w = wlm.make_context(wl)
    smooth = wlm.smooth_state(x, w, wl.N, wl.M, seed=1234)
    ctx.set_split(True)
    energy = {}
    for f in forms:
        ctx.set_precision(PREC_CODE[f])
        ctx.upload_state(smooth)
        t = 0.0
        for _ in range(args.prop_steps):
            ctx.strang(t, t + wl.dt, krylov_dim=wl.krylov_dim, tol=1e-10, max_restarts=8, metric=True)
            t += wl.dt
        energy[f] = ctx.energy(ctx.download_state())
    ctx.close()
    ref = energy["fp64"]
    return {"steps": args.prop_steps, "t_end": args.prop_steps * wl.dt, "energy_fp64": ref,
            "energy_rel_err_vs_fp64": {f: abs(e - ref) / abs(ref) for f, e in energy.items() if f != "fp64"}}


def pmc_traffic(kname: str, workload: str, skip: bool = False):
    """HBM bytes per launch of `kname` (the contraction kernel of the form, dense or with the
    zero-source skip) from the committed hardware-counter summary (rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE passes over `bench.py --pmc-probe`, scaled by a 256 MiB copy); None when no summary
    exists for this workload and mode."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.json")
    if not os.path.exists(path) or "<bf16>" in kname:  # (the counters were taken of the int8 kernel only)
        return None
    with open(path) as f:
        d = json.load(f)
    mode = "skip" if skip else "dense"
    for v in d.get("launches", []):
        if v["kernel"].split("<")[0] == kname and v["mode"] == mode and v["role"] == "contraction alone":
            return v["hbm_bytes"]
    return None


# The probe's segments, in launch order.  A segment holds the contraction kernel launched alone
# (cody_profile_contraction: pair sources by k_pair_src_t, or k_pair_src_i8 for the int8 form, the
# kernel's operands, one warm-up, then three launches) and three full right-hand sides; those two
# kernels run nowhere else in the probe (the right-hand side forms its pair sources in its front
# launch), so the summary splits the counter rows at their dispatches.
PMC_SEGMENTS = [("fp64", "dense"), ("fp64", "skip"), ("int8_sliced", "dense"), ("int8_sliced", "skip")]
# the launches of a right-hand side besides the contraction kernel
RHS_LAUNCHES = ("k_rhs_front", "k_rhs_wsum", "k_rhs_qform", "k_rhs_mid", "k_rhs_gram", "k_rhs_post",
                "k_rhs_final", "k_ci_apply_project")


def pmc_probe(torch, wlm, wl):
    """Launches counted by `rocprofv3 --pmc` runs: three plain copies of 256 MiB (the unit of the
    byte counters), then the segments of PMC_SEGMENTS."""
    dev = torch.device("cuda:0")
    a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)  # 256 MiB
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    ctx, x, w = wlm.make_context(wl)
    st = wlm.seeded_state(x, w, wl.N, wl.M, seed=1234)
    ctx.upload_state(st)
    for prec, mode in PMC_SEGMENTS:
        ctx.set_precision(PREC_CODE[prec])
        ctx.set_source_skip(mode == "skip")
        ctx.profile_contraction(3)
        for _ in range(3):
            ctx.time_derivative(st)
        torch.cuda.synchronize()
    ctx.close()
    print("pmc probe: 3 copies of 256 MiB; segments " + ", ".join(f"{p}/{m}" for p, m in PMC_SEGMENTS) +
          ", each the contraction kernel alone (4 launches) and 3 full right-hand sides", flush=True)


def _pmc_rows(folder: str, counter: str):
    """(dispatch id, kernel name, grid size, value) of every dispatch in a rocprofv3 --pmc output
    folder (its rocpd SQLite database, or CSV output), in dispatch order."""
    import csv
    import glob
    import re
    import sqlite3

    rows = {}

    def add(did, kname, grid, val):
        name = re.sub(r"^(void )?(cody::)?", "", kname).split("(")[0].strip()
        prev = rows.get(did)  # (a counter summed over the dispatch's hardware instances)
        rows[did] = (did, name, grid, val + (prev[3] if prev else 0.0))

    for fn in glob.glob(os.path.join(folder, "**", "*.db"), recursive=True):
        db = sqlite3.connect(fn)
        for did, kname, grid, cname, val in db.execute(
                "select dispatch_id, kernel_name, grid_size, counter_name, value from counters_collection"):
            if cname == counter:
                add(int(did), kname, int(grid or 0), float(val))
        db.close()
    for fn in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter:
                    add(int(r["Dispatch_Id"]), r["Kernel_Name"], int(r.get("Grid_Size", 0) or 0),
                        float(r["Counter_Value"]))
    return [rows[k] for k in sorted(rows)]


def pmc_summary(fetch_dir: str, write_dir: str, wl) -> None:
    """profiles/pmc_traffic_<workload>.json and .md from the FETCH_SIZE and WRITE_SIZE passes over
    --pmc-probe: per segment and kernel the median counter per dispatch, in HBM bytes after the
    calibration by the 256 MiB copy."""
    per = {}
    calib = {}
    for counter, folder in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _pmc_rows(folder, counter)
        copies = [r for r in rows if "copyBuffer" in r[1]]
        gmax = max((r[2] for r in copies), default=0)
        big = [r for r in copies if r[2] == gmax]  # the 256 MiB copies
        if len(big) != 3:
            raise SystemExit(f"pmc summary: {counter}: the three 256 MiB calibration copies are missing in {folder}")
        calib[counter] = float(np.median([r[3] for r in big]))
        seps = [r[0] for r in rows if r[1] in ("k_pair_src_t", "k_pair_src_i8")]
        if len(seps) != len(PMC_SEGMENTS):
            raise SystemExit(f"pmc summary: {counter}: {len(seps)} segment starts, expected {len(PMC_SEGMENTS)}")
        bounds = seps + [1 << 62]
        for i, (prec, mode) in enumerate(PMC_SEGMENTS):
            seg = [r for r in rows if bounds[i] <= r[0] < bounds[i + 1] and "rocclr" not in r[1]]
            kern = KERNEL[prec]
            names = []
            for r in seg:
                if r[1] not in names:
                    names.append(r[1])
            for name in names:
                vals = [r[3] for r in seg if r[1] == name]
                base = name.split("<")[0]
                if base == kern:
                    # warm-ups, then 3 alone, then one per right-hand side
                    roles = [("contraction alone", vals[-6:-3]), ("in the right-hand side", vals[-3:])]
                elif base in RHS_LAUNCHES:
                    roles = [("in the right-hand side", vals[-3:])]
                else:  # (operand set-up of the precision or the profile)
                    roles = []
                for role, v in roles:
                    if v:
                        per.setdefault((name, prec, mode, role), {})[counter] = float(np.median(v))
    copy_kib = 256 * 1024.0
    ff, wf = copy_kib / calib["FETCH_SIZE"], copy_kib / calib["WRITE_SIZE"]
    nGp = 256 * ((wl.nG + 255) // 256)
    Pc = 16 * ((2 * wl.pairs + 15) // 16)
    launches = []
    for (name, prec, mode, role), c in per.items():
        fk, wk = c.get("FETCH_SIZE"), c.get("WRITE_SIZE")
        hb = None if fk is None or wk is None else 1024.0 * (fk * ff + wk * wf)
        launches.append({"kernel": name, "form": prec, "mode": mode, "role": role, "fetch_kib": fk,
                         "write_kib": wk, "hbm_bytes": hb})
    out = {"workload": wl.name, "copy_kib": copy_kib, "fetch_factor": ff, "write_factor": wf,
           "nGp": nGp, "Pc": Pc, "launches": launches}
    base = os.path.join(ROOT, "profiles", f"pmc_traffic_{wl.name}")
    with open(base + ".json", "w") as f:
        json.dump(out, f, indent=1)
    lines = [f"| {v['kernel']} | {v['form']} | {v['mode']} | {v['role']} | {v['fetch_kib']:.0f} | "
             f"{v['write_kib']:.0f} | {v['hbm_bytes']:,.0f} |" for v in launches if v["hbm_bytes"] is not None]
    with open(base + ".table.md", "w") as f:
        f.write(f"Calibration: 256 MiB copy reads FETCH_SIZE {calib['FETCH_SIZE']:.0f} KiB and WRITE_SIZE "
                f"{calib['WRITE_SIZE']:.0f} KiB: read bytes = {ff:.3f} x FETCH_SIZE KiB x 1024, write bytes = "
                f"{wf:.3f} x WRITE_SIZE KiB x 1024.\n\n"
                "| kernel | form | mode | role | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes per launch |\n"
                "|---|---|---|---|---:|---:|---:|\n" + "\n".join(lines) + "\n")
    print(f"wrote {base}.json and {base}.table.md ({len(launches)} entries)")


def batched_measurement(args, torch, wlm, check: bool):
    """Independent seeded instances (N_g in {511, 1026, 2046, 4096}, M in {4, 6, 8, 12}, N = M/2,
    softening a ~ U(0.5, 2)) replayed as one captured graph per batched right-hand side.  Each
    context's contraction k split is picked by the project autotuner first."""
    from codyfortran_amd import hip

    dev = torch.device("cuda:0")
    ws = wlm.batched_workloads(n=args.batch)
    ctxs, states, d_s, d_o, splits = [], [], [], [], {}
    for i, w in enumerate(ws):
        ctx, x, wts = wlm.make_context(w)
        st = wlm.seeded_state(x, wts, w.N, w.M, seed=500 + i)
        ctx.upload_state(st)
        splits.setdefault(f"{w.nG}x{w.M}", ctx.autotune())
        ctxs.append((ctx, wts))
        states.append(st)
        d_s.append(torch.from_numpy(st.view(np.float64).copy()).to(dev))
        d_o.append(torch.zeros_like(d_s[-1]))
    b = hip.Batch([c for c, _ in ctxs])
    stream = torch.cuda.Stream(device=dev)
    ps, po = [t.data_ptr() for t in d_s], [t.data_ptr() for t in d_o]

    def timed():
        for _ in range(3):
            b.time_derivative_dev(ps, po, 0.0, stream.cuda_stream)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.batch_steps):
            b.time_derivative_dev(ps, po, 0.0, stream.cuda_stream)
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / args.batch_steps

    refs = None
    if check:
        from tests.common import oracle_system

        refs = []
        for w, st in zip(ws, states):
            s, _, _ = oracle_system(w)
            refs.append(s.time_derivative(st))
            del s  # frees the oracle's dense interaction matrix

    def worst_err():
        from tests.common import state_rel

        return max(state_rel(o.cpu().numpy().view(np.complex128), r, w.nC, wts, w.M)
                   for (ctx, wts), w, o, r in zip(ctxs, ws, d_o, refs))

    if args.batch_int8_only:
        for ctx, _ in ctxs:
            ctx.set_precision(hip.PREC_INT8_SLICED)
        pick8, _ = b.autotune(ps, po)
        ms8 = timed()
        b.close()
        for c, _ in ctxs:
            c.close()
        return {"instances": len(ws), "value": len(ws) * 1000.0 / ms8, "unit": "instance RHS evals/s",
                "ms_per_batched_rhs": ms8, "form": "int8_sliced", "dtype": "f64 (int8_sliced contraction)",
                "schedule": {"chains": hip.Batch.SCHEDULES[pick8][0],
                             "split_threshold": hip.Batch.SCHEDULES[pick8][1]}}
    ms_default = timed()
    pick, cand_ms = b.autotune(ps, po)
    ms = timed()
    flops = sum(2.0 * w.nG ** 2 * 2 * w.pairs for w in ws)
    out = {"instances": len(ws), "value": len(ws) * 1000.0 / ms, "unit": "instance RHS evals/s",
           "ms_per_batched_rhs": ms, "contraction_flop_per_batched_rhs": flops,
           "contraction_tflops_aggregate": flops / (ms * 1e-3) / 1e12, "dtype": "f64",
           "k_split_by_shape": splits,
           "schedule": {"form": "one HIP graph, 4 parallel chains, grouped launches per orbital count",
                        "ms_before_schedule_autotune": ms_default,
                        "autotuned": {"chains": hip.Batch.SCHEDULES[pick][0],
                                      "split_threshold": hip.Batch.SCHEDULES[pick][1],
                                      "candidates_ms": {f"{c}x{t}": round(m, 4) for (c, t), m
                                                        in zip(hip.Batch.SCHEDULES, cand_ms)}}}}
    if check:
        out["max_rhs_rel_err_vs_fp64_oracle"] = worst_err()
    # the same batch with the int8-sliced contraction in every instance
    for ctx, _ in ctxs:
        ctx.set_precision(hip.PREC_INT8_SLICED)
    pick8, cand8 = b.autotune(ps, po)
    ms8 = timed()
    out["int8_sliced"] = {"value": len(ws) * 1000.0 / ms8, "ms_per_batched_rhs": ms8,
                          "dtype": "f64 (int8_sliced contraction)",
                          "schedule": {"chains": hip.Batch.SCHEDULES[pick8][0],
                                       "split_threshold": hip.Batch.SCHEDULES[pick8][1]}}
    if check:
        out["int8_sliced"]["max_rhs_rel_err_vs_fp64_oracle"] = worst_err()
    # the batch's figure: the faster form whose every instance is within 1e-9 of the FP64 oracle
    # (without the oracle check, the FP64 form)
    fp64 = {k: out[k] for k in ("value", "ms_per_batched_rhs", "dtype", "max_rhs_rel_err_vs_fp64_oracle") if k in out}
    out["fp64"] = fp64
    i8 = out["int8_sliced"]
    if check and i8["max_rhs_rel_err_vs_fp64_oracle"] <= TOL and i8["value"] > fp64["value"]:
        out.update(value=i8["value"], ms_per_batched_rhs=i8["ms_per_batched_rhs"], dtype=i8["dtype"],
                   max_rhs_rel_err_vs_fp64_oracle=i8["max_rhs_rel_err_vs_fp64_oracle"],
                   contraction_tflops_aggregate=flops / (i8["ms_per_batched_rhs"] * 1e-3) / 1e12)
        out["form"] = "int8_sliced"
    else:
        out["form"] = "fp64"
    b.close()
    for c, _ in ctxs:
        c.close()
    return out


def cpu_baseline(wl, state, budget_s: float):
    """FP64 oracle (C + OpenMP restatement of the reference) on this host: evals/s of the RHS.

    The oracle's mean-field sweep leaves out the source points where every pair source is exactly
    zero, as the reference's convolution does (ConvolutionIntegral_Apply collects |src| > 0 first);
    that is the reported value.  The same evaluation with every source point visited is reported
    beside it (`dense`), and on one thread (`single_thread`)."""
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
    orbs = state[wl.nC:].reshape(wl.M, wl.nG).T
    oracle.set_compaction(True)
    active = s.active_sources(orbs)

    def timed(budget, cap):
        s.time_derivative(state)  # warm-up
        times, t0 = [], time.perf_counter()
        while True:
            t1 = time.perf_counter()
            s.time_derivative(state)
            times.append(time.perf_counter() - t1)
            el = time.perf_counter() - t0
            if (el >= budget and len(times) >= 20) or len(times) >= cap:
                return times, el

    d = s.time_derivative(state)
    s.time_derivative(state)
    times, el = timed(budget_s, 200)
    n, med = len(times), float(np.median(times))
    threads = lib.orc_num_threads()
    oracle.set_compaction(False)
    dtimes, del_ = timed(budget_s / 3, 60)
    oracle.set_compaction(True)
    # the same evaluation on one thread (a bounded sample: a few evaluations)
    lib.orc_set_num_threads(1)
    n1, t1 = 0, time.perf_counter()
    while True:
        s.time_derivative(state)
        n1 += 1
        el1 = time.perf_counter() - t1
        if el1 >= budget_s / 3 or n1 >= 50:
            break
    lib.orc_set_num_threads(threads)
    env = f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', 'unset')}, nproc={os.cpu_count()}"
    return {
        "value": 1.0 / med,
        "unit": "RHS evals/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"median of {n} timed full RHS evaluations (after 3 warm-up) of the same {wl.nG}-point, "
                  f"M={wl.M}, N={wl.N} state ({el:.1f} s, OpenMP threads={threads}, {env}, no BLAS; built -O3 "
                  f"-march={'native' if lib_path else 'x86-64-v3'}); zero sources left out as in the "
                  f"reference's convolution: {active} of {wl.nG} source points visited",
        "active_source_fraction": active / wl.nG,
        "mean_value": n / el,
        "dense": {"value": 1.0 / float(np.median(dtimes)), "unit": "RHS evals/s", "cores": int(threads),
                  "sample": f"median of {len(dtimes)} evaluations visiting all {wl.nG} source points "
                            f"({del_:.1f} s)"},
        "single_thread": {"value": n1 / el1, "unit": "RHS evals/s", "cores": 1,
                          "sample": f"{n1} evaluations of the same state on one OpenMP thread ({el1:.1f} s)"},
    }, d, s, wts


def small_workload_measurement(args, torch, wlm):
    """The 401-point, M=4, N=2 system: FP64 GPU and CPU-oracle RHS evals/s and the RHS error."""
    wl = wlm.WORKLOADS["mctdhf_1d_401"]
    ctx, x, w = wlm.make_context(wl)
    state = wlm.seeded_state(x, w, wl.N, wl.M, seed=1234)
    dev = torch.device("cuda:0")
    d_state = torch.from_numpy(state.view(np.float64).copy()).to(dev)
    d_out = torch.zeros_like(d_state)
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    ctx.set_precision(0)
    for _ in range(args.warmup):
        ctx.time_derivative_dev(d_state.data_ptr(), d_out.data_ptr(), 0.0, sptr)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    steps = max(args.steps, 1)
    start.record(stream)
    for _ in range(steps):
        ctx.time_derivative_dev(d_state.data_ptr(), d_out.data_ptr(), 0.0, sptr)
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    out = {"workload": wl.name, "nG": wl.nG, "N": wl.N, "M": wl.M, "precision": "fp64",
           "gpu_evals_per_s": 1000.0 / ms, "gpu_ms_per_eval": ms}
    if not args.no_cpu_baseline:
        from tests.common import oracle_system, state_rel

        s, _, wts = oracle_system(wl)
        d = s.time_derivative(state)
        n, t0 = 0, time.perf_counter()
        while True:
            s.time_derivative(state)
            n += 1
            el = time.perf_counter() - t0
            if el >= 2.0 or n >= 2000:
                break
        g = d_out.cpu().numpy().view(np.complex128)
        out["cpu_evals_per_s"] = n / el
        out["cpu_sample"] = f"{n} evaluations on the oracle's OpenMP threads ({el:.1f} s)"
        out["rhs_rel_err_vs_fp64_oracle"] = state_rel(g, d, wl.nC, wts, wl.M)
    ctx.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent MCTDHF instances shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)
    import torch

    from codyfortran_amd import workloads as wlm


    wl = wlm.WORKLOADS[args.workload]
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], wl)
        return
    if args.pmc_probe:
        pmc_probe(torch, wlm, wl)
        return
    if args.conservation:
        print(json.dumps({"conservation": conservation_measurement(args, wlm, wl)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_measurement(args, wlm, wl, checks=args.adjoint_fd > 0)}), flush=True)
        return
    ctx, x, w = wlm.make_context(wl)
    state = wlm.seeded_state(x, w, wl.N, wl.M, seed=1234)
    dev = torch.device("cuda:0")
    d_state = torch.from_numpy(state.view(np.float64).copy()).to(dev)
    d_out = torch.zeros_like(d_state)
    # the library's work is enqueued on an explicit (default-priority) torch stream, so the torch
    # events bracketing the timed region are recorded on the stream the kernels run on
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    assert sptr != 0

    flops = 2.0 * wl.nG ** 2 * 2 * wl.pairs  # pair symmetry exploited: 2 N_g^2 2P
    nGp = 256 * ((wl.nG + 255) // 256)
    Pc = 16 * ((2 * wl.pairs + 15) // 16)
    # source points of the seeded state where some pair source conj(phi_l) phi_k w is non-zero:
    # the only columns of V the zero-source skip contracts (the reference's convolution leaves the
    # others out too)
    _, orbs0 = wlm.split_state(state, wl.nC, wl.nG, wl.M)
    src = np.abs(orbs0)[:, :, None] * np.abs(orbs0)[:, None, :] * w[:, None, None]
    n_src = int(np.count_nonzero(src.reshape(wl.nG, -1).any(axis=1)))

    def time_form(prec, skip):
        ctx.set_precision(PREC_CODE[prec])
        ctx.set_source_skip(skip)
        for _ in range(args.warmup):
            ctx.time_derivative_dev(d_state.data_ptr(), d_out.data_ptr(), 0.0, sptr)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        # the timed region: K right-hand sides on `stream`, bracketed by device-wide synchronisation
        t0 = time.perf_counter()
        start.record(stream)
        for _ in range(args.steps):
            ctx.time_derivative_dev(d_state.data_ptr(), d_out.data_ptr(), 0.0, sptr)
        stop.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = start.elapsed_time(stop) / args.steps
        out = d_out.cpu().numpy().view(np.complex128).copy()
        w_field = ctx.meanfield_w(orbs0)
        energy = ctx.energy(state)
        ctx.upload_state(state)  # (the profiles below run on the context's own state)
        kernels = ctx.profile_rhs(args.profile_iters)
        kname = KERNEL[prec]
        kms = ctx.profile_contraction(args.profile_iters)
        cols = n_src if (skip and prec in SKIPPING) else nGp  # contraction columns the form needs
        useful = 2.0 * wl.nG * (n_src if (skip and prec in SKIPPING) else wl.nG) * 2 * wl.pairs
        if BOUND[prec] == "mfma":
            work, unit = useful, "TFLOP/s"
            ach = useful / (kms * 1e-3) / 1e12 if kms else None
        else:
            eb = OPERAND_BYTES[prec]
            work = eb * nGp * cols + eb * Pc * cols + 8.0 * Pc * nGp  # V and S operands once, W^T once
            if prec in ("int8_sliced", "bf16_sliced"):  # S's exponents (one per line and 64-point block); the epilogue's
                # correlation: phi, h phi read, c written once, two overlap partials per 16-point tile,
                # the canonical rho2
                npr = wl.M * (wl.M - 1) // 2
                work += 4.0 * Pc * cols / 64
                work += 3 * 16.0 * wl.M * wl.nG + 2 * 16.0 * wl.M ** 2 * (nGp // 16) + 16.0 * npr * npr
            unit = "GB/s"
            ach = work / (kms * 1e-3) / 1e9 if kms else None
        traffic = pmc_traffic(kname, wl.name, skip and prec in SKIPPING)
        roof = {"bound": BOUND[prec], "kernel": kname, "achieved": ach, "peak": PEAK[prec], "unit": unit,
                "frac": ach / PEAK[prec] if ach else None, "traffic": traffic,
                ("flop_per_launch" if BOUND[prec] == "mfma" else "algorithmic_bytes_per_launch"): work,
                "useful_flop_per_launch": useful, "avg_launch_ms": kms,
                "traffic_source": (f"profiles/pmc_traffic_{wl.name}.json (HBM bytes per launch)"
                                   if traffic else None)}
        return {"precision": prec, "zero_source_skip": bool(skip and prec in SKIPPING), "ms_per_step": ms,
                "value": 1000.0 / ms, "wall_s_timed": wall, "roofline": roof, "kernels_ms": kernels,
                "dstate": out, "W": w_field, "energy": energy}

    forms = (["fp64", "int8_sliced", "bf16_sliced", "split_fp32", "bf16x3"] if args.precision == "auto"
             else [args.precision])
    results = {f: time_form(f, True) for f in forms}
    # the exact forms also without the zero-source skip (every source point contracted)
    dense = {} if args.no_dense else {f + "_dense": time_form(f, False) for f in forms if f in SKIPPING}
    ctx.set_source_skip(True)
    ctx.set_precision(0)

    errs = {}
    cb = None
    if not args.no_cpu_baseline:
        cb, d_orc, s_orc, wts = cpu_baseline(wl, state, args.cpu_seconds)
        from tests.common import rel, state_rel

        _, orbs = wlm.split_state(state, wl.nC, wl.nG, wl.M)
        w_orc = s_orc.meanfield_w(orbs)
        e_orc = s_orc.energy(state)
        errs = {f: {"rhs": state_rel(r["dstate"], d_orc, wl.nC, wts, wl.M),
                    "W": rel(r["W"], w_orc),
                    "energy": abs(r["energy"] - e_orc) / abs(e_orc)} for f, r in {**results, **dense}.items()}
    # a form qualifies when every reported quantity is within TOL of the FP64 oracle
    ok = {f: bool(max(errs[f].values()) <= TOL) if f in errs else f.startswith("fp64")
          for f in list(forms) + list(dense)}
    valid = [f for f in forms if ok[f]]
    best = max(valid, key=lambda f: results[f]["value"]) if valid else forms[0]
    b = results[best]
    result = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else f"f64 ({best} contraction)",
        "data": "synthetic: seeded complex Gaussian orbitals x exp(-x^2/8), MGS-orthonormalised in "
                "the FEDVR weights; complex Gaussian CI vector, normalised",
        "config": {"workload": wl.name, "nG": wl.nG, "elements": wl.nE, "gll_points": wl.nLoc,
                   "N": wl.N, "M": wl.M, "nC": wl.nC, "pairs": wl.pairs, "precision": best,
                   "regularization": wl.reg, "zero_source_skip": b["zero_source_skip"],
                   "launches": "kernel by kernel on one stream (HIP graphs for propagator steps)"},
        "zero_source_skip": {
            "what": "the contraction leaves out the k blocks where every pair source conj(phi_l) phi_k w is "
                    "exactly zero, as the reference's convolution leaves out zero sources "
                    "(M_Utils_ConvolutionIntegral.f90:84-93); the skipped products are exact zeros, so the "
                    "results equal the dense contraction's up to summation order; cpu_baseline.value uses "
                    "the same compaction",
            "source_points_nonzero": n_src, "source_points": wl.nG,
            "dense": {f: {"value": r["value"], "ms_per_step": r["ms_per_step"], "meets_1e-9": ok[f],
                          "rel_err": errs.get(f), "roofline": r["roofline"]} for f, r in dense.items()},
        },
        "roofline": b["roofline"],
        "cpu_baseline": cb,
        "rel_err_vs_fp64_oracle": errs,
        "forms": {f: {"value": r["value"], "ms_per_step": r["ms_per_step"], "rel_err": errs.get(f),
                      "meets_1e-9": ok[f], "zero_source_skip": r["zero_source_skip"], "roofline": r["roofline"],
                      "kernels_ms": r["kernels_ms"]} for f, r in {**results, **dense}.items()},
    }
    if args.prop_steps > 0:
        result["propagation"] = propagation_measurement(args, wlm, wl, state, PREC_CODE[best])
        if "fp64" in forms and len(forms) > 1:
            result["propagation"]["by_contraction_form"] = propagated_energy_by_form(args, wlm, wl, forms)
    if args.batch > 0:
        result["batched"] = batched_measurement(args, torch, wlm, check=not args.no_cpu_baseline)
    if wl.name != "mctdhf_1d_401":
        result["small_workload"] = small_workload_measurement(args, torch, wlm)
    if args.adjoint_steps > 0:
        result["adjoint"] = adjoint_measurement(args, wlm, wl)
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
