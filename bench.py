#!/usr/bin/env python3
"""Throughput of the FastEddy compressible-LES RK3 step on one MI355X.

Workload (default, --workload dry): the dry convective boundary layer at production size, 512x512x128
cells at dx=dy=dz=10 m, dt=0.02 s, FP32 state (rho, rho*u, rho*v, rho*w,
rho*theta, rho*e), seeded uniform +/-0.25 K theta perturbations below 400 m,
geostrophic wind (10,0) m/s, surface heat flux 0.35 K m/s, f=1e-4 1/s,
5th-order upwind advection, 1.5-order TKE closure, RK3.  Synthetic initial
state from the reference's own initial-condition rules; inputs resident in
HBM before the timed region.  --workload moist runs the 256x256x160 moist
shallow-cumulus case (dx=25 m, dt=0.05 s, N_prog = 8 with rho*q_v and rho*q_l).

value = whole-job cell-steps/s: interior cells x RK3 steps / wall seconds of
exactly --steps steps, bracketed by device synchronisation.

roofline: the stage kernel (the fused tendency + RK3 update launch, three per
step) is the dominant kernel.  Its algorithmic HBM bytes per launch are
8 B/cell x N_prog for the first stage (read phi^n, write phi*) and
12 B/cell x N_prog for stages 2 and 3 (read the stage input and phi^n, write
the output): 32 x N_prog = 192 B per cell-step in total.  achieved = those
bytes over the kernel's own HIP-event durations inside the timed region.

cpu_baseline: the repository's FP32 CPU oracle (a restatement of the
reference step, OpenMP over i within each phase) timed on this host for six
RK3 steps of the same workload (after one untimed step; about 10 s on 16 cores).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "fasteddy-model_amd"))

METRIC = "cell-steps/s & % HBM roofline: plain vs batched/learned-SGS at fixed error"
PEAK_HBM = 8.0e12


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch measured with rocprofv3 PMC counters, if a summary is committed."""
    p = os.path.join(HERE, "profiles", f"traffic_{kernel_tag}.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh).get("bytes_per_launch")


def cpu_baseline(params, threads, nsteps=6):
    """The FP32 CPU oracle on the full workload: one untimed RK3 step (first touch), then
    nsteps timed RK3 steps."""
    import fe_py
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle_py
    m = fe_py.Model(params, os.path.join("/tmp", "fe_bench_cpu"), device=False)
    try:
        o = oracle_py.Oracle(params, 32, threads=threads)
        o.load_from_model(m)
    finally:
        m.close()
    o.steps(1)
    t0 = time.perf_counter()
    o.steps(nsteps)
    dt = time.perf_counter() - t0
    o.close()
    cells = params["Nx"] * params["Ny"] * params["Nz"]
    return {"value": cells * nsteps / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} RK3 steps of the full {params['Nx']}x{params['Ny']}x{params['Nz']} workload after one "
                      f"untimed step (FP32 CPU oracle, OpenMP x{threads}); {dt:.1f} s timed"}


def moist_initial_state(m, seed=2024, nh=3):
    """Shallow-cumulus moisture: q_v = 17 g/kg at the ground, 13 g/kg at 1.5 km, 4 g/kg at 3 km,
    seeded uniform +/-0.025 g/kg below 1.6 km, no liquid (theta perturbations of +/-0.1 K below
    1.6 km come from the case's own initial-condition rules)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    hydro, mq, zpos = m.array("hydroFlds"), m.array("moistScalars"), m.array("zPos")
    _, nxh, nyh, nzh = hydro.shape
    sl = (slice(nh, nxh - nh), slice(nh, nyh - nh), slice(nh, nzh - nh))
    z = zpos[sl].astype(np.float64)
    qv = np.interp(z, [0.0, 1500.0, 3000.0, 1e6], [17.0, 13.0, 4.0, 4.0])
    qv = qv + np.where(z < 1600.0, rng.uniform(-0.025, 0.025, size=qv.shape), 0.0)
    mq[0][sl] = (hydro[0][sl].astype(np.float64) * qv).astype(mq.dtype)
    mq[1][sl] = 0.0
    if m.L.fe_upload_state() != 0:
        raise RuntimeError("fe_upload_state failed")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", choices=("dry", "moist"), default="dry",
                    help="dry: the 512x512x128 CBL (the headline line); moist: the 256x256x160 shallow-cumulus case")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is not part of this build)", file=sys.stderr)
        sys.exit(2)

    import fe_py
    moist = args.workload == "moist"
    base = fe_py.CFG2 if moist else fe_py.CFG1
    params = dict(base, Nx=args.nx or base["Nx"], Ny=args.nx or base["Ny"], Nz=args.nz or base["Nz"])
    cells = params["Nx"] * params["Ny"] * params["Nz"]
    nprog = 8 if moist else 6
    m = fe_py.Model(params, os.path.join("/tmp", "fe_bench"))
    if moist:
        moist_initial_state(m)
    L = m.L
    L.fe_copy_probe.restype = C.c_double
    L.fe_copy_probe.argtypes = [C.c_long, C.c_int]
    L.fe_timing_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]
    try:
        variant = L.fe_kernel_info().decode()
        m.device_steps(args.warmup)
        L.fe_timing_enable(args.steps * 4 + 8)
        t0 = time.perf_counter()
        m.device_steps(args.steps)          # synchronises at the end
        wall = time.perf_counter() - t0
        ms = (C.c_double * 3)()
        cnt = (C.c_int * 3)()
        L.fe_timing_read(ms, cnt, 3)
        L.fe_timing_enable(0)
        # one cuda_timeIntCommence batch through the C-ABI: the steps plus the batch-end copy of
        # state, p' and surface fields back into the host arrays (PCIe-inclusive rate; not `value`)
        nb = 50
        L.fe_set_int(b"NtBatch", nb)
        t0 = time.perf_counter()
        m.commence()
        wall_batch = time.perf_counter() - t0
        copy_bw = L.fe_copy_probe(1 << 28, 20)
    finally:
        m.close()

    value = cells * args.steps / wall
    ms_per_step = wall / args.steps * 1e3
    # algorithmic bytes: stage 1 reads phi^n and writes phi* (8 B x N_prog per cell),
    # stages 2-3 read the stage input and phi^n and write the output (12 B x N_prog per cell)
    n_plain, n_phys = cnt[1], cnt[2]
    stage_ms = ms[1] + ms[2]
    alg_bytes = cells * nprog * 4 * (2 * (n_plain // 2) + 3 * (n_plain - n_plain // 2) + 3 * n_phys)
    achieved = alg_bytes / (stage_ms * 1e-3) if stage_ms > 0 else 0.0
    roof = {
        "bound": "hbm",
        "achieved": achieved / 1e9,
        "peak": PEAK_HBM / 1e9,
        "unit": "GB/s",
        "frac": achieved / PEAK_HBM,
        "traffic": None if moist else traffic_from_profiles("stage_kernel"),
        "kernel": "fe::stage_kernel (fused tendency + RK3 stage update, 3 launches/step)",
        "avg_launch_ms": {"stage_plain": ms[1] / max(n_plain, 1), "stage_physics": ms[2] / max(n_phys, 1),
                          "surface_pass": ms[0] / max(cnt[0], 1)},
        "step_frac": value * 32 * nprog / PEAK_HBM,
        "copy_GBps": copy_bw,
    }
    pcie = {"value": cells * nb / wall_batch, "unit": "cell-steps/s", "ntbatch": nb,
            "note": "one cuda_timeIntCommence batch including the batch-end device-to-host copy"}
    cpu = None
    if not args.no_cpu:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
            cpu = cpu_baseline(params, threads)
        except Exception as exc:  # the baseline is reported, never required
            cpu = {"error": str(exc)}
    line = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": (f"plain moist shallow cumulus {params['Nx']}x{params['Ny']}x{params['Nz']} dx=25m dt=0.05s "
                                "FP32 RK3 (prognostic TKE, q_v, q_l, saturation adjustment)") if moist else
                               f"plain dry CBL {params['Nx']}x{params['Ny']}x{params['Nz']} dx=10m dt=0.02s FP32 RK3 (prognostic TKE)",
                   "cells": cells, "n_prog": nprog, "kernel_variant": variant},
        "roofline": roof,
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
