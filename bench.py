"""Benchmark of the DNS + droplet time step on one MI355X.

Workload (the configuration the headline metric is quoted on): 256^3 periodic box,
deterministic forcing, 1,000,000 droplets with log-normal radii (median 25 um,
sigma_ln 0.4, clipped to 5-60 um), hydrodynamic interactions with a cutoff of 50
radii, collision detection and statistics.  A "step" is one full time step of the
flow and of every droplet.  Synthetic inputs: seeded initial flow and droplets,
generated on the host before the timed region; all state is resident in HBM.

Prints ONE JSON line.  value = FP32-mixed steps/s; the FP64 path is timed on the
same workload for the speed-up, and the two are compared after 10 steps for the
velocity error.  The CPU baseline runs the FP64 oracle (the project's CPU
restatement of the reference algorithm) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mpi_2d_domain_decomp_turb_particles_amd as dd  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = "256^3 DNS (deterministic forcing) + 1M log-normal droplets, HI 50 radii, collisions"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_params(precision: int, N: int = 256, npart: int = 1_000_000):
    # parameter.d_256_det: rnu 0.002, dt 0.001, eta/tau_K 0.01392/0.09694, force 1.242477/0.391356
    return dd.default_params(
        N=N, precision=precision, dt=0.001, rnu=0.002, forcing=1, force=(1.242477, 0.391356),
        etk_DNS=0.01392, tauk_DNS=0.09694, npart=npart, nset=10,
        radius=[5.0, 10.0, 15.0, 20.0, 25.0, 30.0, 35.0, 40.0, 50.0, 60.0],
        hdi_included=1, hdi_trunc=50.0, nonoverlap=1,
    )


def make_solver(precision, N, npart, seed_flow=123, seed_drop=290710):
    s = dd.Solver(workload_params(precision, N, npart))
    s.init_flow(dd.SPEC_K4_GAUSS, seed_flow, 4.0, 1.0)
    s.init_droplets(seed=seed_drop, radii=dd.RADII_LOGNORMAL, median_um=25.0, sigma_ln=0.4, rmin_um=5.0, rmax_um=60.0)
    return s


def kernel_bytes(N: int, prec: int):
    """Compulsory HBM bytes per launch of each flow kernel (every input read once,
    every output written once).  C = one complex field N*N*(N/2), R = one real field."""
    cs = 16 if prec == dd.FP64 else 8
    rs = cs // 2
    C = N * N * (N // 2) * cs
    R = N * N * N * rs
    return {
        "y_inv": 12 * C,            # 6 fields read + written
        "x_fused": 6 * C + 3 * C + 3 * R,   # read u,w; write u x w and physical u
        "y_fwd": 6 * C,
        "pre": 0,
        "z_fused": 21 * C,          # read N, u, o (9C); write u, o, 6 inverse outputs (12C)
        "sym": 0,
    }


def measured_traffic(kernel, N, prec):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    key = f"{kernel}/N{N}/{'fp64' if prec == dd.FP64 else 'fp32'}"
    ent = tab.get(key)
    return None if ent is None else ent.get("bytes_per_launch")


def timed_run(s, steps, warmup):
    if warmup:
        s.step(warmup)
    t0 = time.perf_counter()
    st = s.step(steps)
    t1 = time.perf_counter()
    return (t1 - t0), s.timing(), st


def cpu_baseline(N, npart, state, steps=2):
    """FP64 oracle on the same workload: one untimed step (droplet history start),
    then `steps` timed steps.  Returns steps/s and the thread count used."""
    from oracle import oracle as O
    p = workload_params(dd.FP64, N, npart)
    o = O.Oracle(p)
    o.set_flow(*state["flow"])
    d = state["drop"]
    o.set_droplets(d["yp"], d["fvp"], d["fdp"], d["pertvel"])
    o.set_new_particle(1)
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return steps / dt, dt


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--npart", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--only-fp32", action="store_true",
                    help="time the FP32-mixed path only (for profiler runs)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the whole periodic domain lives on a single MI355X; --gpus {a.gpus} is not supported", file=sys.stderr)
        return 2
    N, npart = a.N, a.npart

    # ---- FP32-mixed (the value)
    log(f"set-up N={N} droplets={npart}")
    s32 = make_solver(dd.FP32_MIXED, N, npart)
    log("FP32-mixed timed run")
    init_state = {"flow": s32.get_flow(), "drop": s32.get_droplets()}
    el32, t32, st = timed_run(s32, a.steps, a.warmup)
    cc = s32.collision_counts()
    steps_per_s = a.steps / el32
    kms = list(t32.flow_kernel_ms)
    names = ["y_inv", "x_fused", "y_fwd", "pre", "z_fused", "sym"]
    per_launch = {n: kms[i] / a.steps for i, n in enumerate(names)}
    kb = kernel_bytes(N, dd.FP32_MIXED)
    dom = max(("y_inv", "x_fused", "y_fwd", "z_fused"), key=lambda n: per_launch[n])
    achieved = kb[dom] / (per_launch[dom] * 1e-3) / 1e9
    flow_bytes = sum(kb.values())
    flow_ms = t32.flow_ms / a.steps
    copy_ms, copy_bytes = s32.profile_copy(20)
    s32.close()
    if a.only_fp32:
        print(json.dumps({"metric": "fp32-mixed steps/s (profiling run)", "value": steps_per_s,
                          "flow_kernel_ms_per_launch": per_launch, "phase_ms_per_step": {
                              "flow": flow_ms, "cells": t32.cells_ms / a.steps, "interp": t32.interp_ms / a.steps,
                              "hdi": t32.hdi_ms / a.steps, "advance": t32.advance_ms / a.steps,
                              "colldet": t32.colldet_ms / a.steps, "overlap": t32.overlap_ms / a.steps}}))
        return 0

    # ---- FP64 plain path, same initial state
    log(f"FP32-mixed {steps_per_s:.3f} steps/s; FP64 timed run")
    s64 = dd.Solver(workload_params(dd.FP64, N, npart))
    s64.set_flow(*init_state["flow"])
    d = init_state["drop"]
    s64.set_droplets(d["yp"], d["fvp"], d["fdp"], d["pertvel"])
    el64, t64, _ = timed_run(s64, a.steps, a.warmup)
    fp64_sps = a.steps / el64
    s64.close()

    # ---- accuracy: 10 steps FP32-mixed vs FP64 from the same state
    log(f"FP64 {fp64_sps:.3f} steps/s; accuracy run")
    vel_err = None
    try:
        ra = dd.Solver(workload_params(dd.FP32_MIXED, N, npart))
        rb = dd.Solver(workload_params(dd.FP64, N, npart))
        for r in (ra, rb):
            r.set_flow(*init_state["flow"])
            r.set_droplets(d["yp"], d["fvp"], d["fdp"], d["pertvel"])
            r.step(10)
        va, _ = ra.get_flow()
        vb, _ = rb.get_flow()
        num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(va, vb))
        den = sum(float(np.sum(y ** 2)) for y in vb)
        vel_err = (num / den) ** 0.5
        ra.close()
        rb.close()
    except dd.DDNSError as e:
        print(f"accuracy run failed: {e}", file=sys.stderr)

    # ---- CPU baseline (oracle, FP64) on a bounded sample
    cpu = None
    if not a.no_cpu:
        log("CPU baseline")
        try:
            sps, secs = cpu_baseline(N, npart, init_state, a.cpu_steps)
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            cpu = {"value": sps, "unit": "steps/s", "cores": threads, "kind": "port",
                   "sample": f"{a.cpu_steps} full steps of the same workload after one untimed step ({secs:.1f} s), FP64 oracle"}
        except Exception as e:  # the baseline is reported, never required
            print(f"cpu baseline failed: {e}", file=sys.stderr)

    out = {
        "metric": "steps/s at 256^3+1M droplets, FP32-mixed vs FP64 on 1 GPU, vel err <=1e-5",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el32 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (f32 flow + HI, f64 droplets)",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "N": N, "droplets": npart, "parallelism": "single GPU"},
        "fp64_steps_per_s": fp64_sps,
        "speedup_fp32_mixed_vs_fp64": steps_per_s / fp64_sps,
        "vel_rel_l2_10_steps_fp32_vs_fp64": vel_err,
        "droplet_updates_per_s": steps_per_s * npart,
        "pairs_tested_per_s": cc.pairs_tested / (t32.colldet_ms / a.steps * 1e-3) if t32.colldet_ms > 0 else None,
        "gmres_iters_last_step": cc.gmres_iters,
        "flow_ms_per_step": flow_ms,
        "flow_algorithmic_GBps": flow_bytes / (flow_ms * 1e-3) / 1e9,
        "phase_ms_per_step": {
            "flow": flow_ms, "cells": t32.cells_ms / a.steps, "interp": t32.interp_ms / a.steps,
            "hdi": t32.hdi_ms / a.steps, "advance": t32.advance_ms / a.steps,
            "colldet": t32.colldet_ms / a.steps, "overlap": t32.overlap_ms / a.steps,
        },
        "flow_kernel_ms_per_launch": per_launch,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, N, dd.FP32_MIXED),
                     "algorithmic_bytes": kb[dom]},
        "hbm_copy_GBps": copy_bytes / (copy_ms * 1e-3) / 1e9,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
