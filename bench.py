#!/usr/bin/env python3
"""Benchmark of the Kirchhoff prestack time-migration hot path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
3-D prestack KTM at one rank's share -- 1,000,000 seeded shot/receiver pairs
over 10 x 10 km (offsets U(0,4 km), azimuths U(0,2 pi)), 1500 samples at
4 ms, traces = Ricker diffractions of 200 scatterers + noise at 10 dB,
v_rms = 1500 + 350 t + 0.02 x, image 400 x 400 x 1500 at 25 m, aperture
2500 m, triangle anti-alias filter (trangeMax = 8).

A step is one trace block (ntblock = 50,000 traces, the reference's read
block, Kirchhoff.c:63): trace preparation + migration of the block into the
whole resident image.  Inputs are resident in HBM before the timed region.

The JSON line reports G contributions/s (contribution = one evaluation of
the inner loop of Kirchhoff.c:1376-1396 for a (trace, column, sample) whose
column is in aperture), the HBM roofline of the migration kernel at the
algorithmic 12 B per contribution (three FP32 taps, Kirchhoff.c:1392), and
the CPU oracle timed on this host on a bounded trace sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FP32 = 12.0   # three FP32 trace taps per contribution
BYTES_HALF = 6.0    # three 16-bit taps
METRIC = "G trace×image-pt contributions/s, %HBM roofline; BF16-trace vs FP32 at err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed trace blocks")
    ap.add_argument("--warmup", type=int, default=1, help="untimed trace blocks")
    ap.add_argument("--ntr", type=int, default=1_000_000, help="survey traces")
    ap.add_argument("--block", type=int, default=50_000, help="traces per step (ntblock)")
    ap.add_argument("--ncols", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy 3x (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration():
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 1 << 28  # 2^28 floats = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}), flush=True)


def cpu_baseline(sv, traces_np, target_s, nthreads):
    """Oracle FP32 (the reference algorithm, strict IEEE, -O3) on the first S
    traces over the whole image, jobs = one slab per thread as the reference's
    MPI decomposition.  S is calibrated so the run takes about target_s."""
    import oracle
    cfg = sv.cfg.replace(jobs_x=nthreads, jobs_y=1)
    n_cal = 16
    t0 = time.perf_counter()
    _, _, nc = oracle.migrate(cfg, sv.ns, sv.dt, traces_np[:n_cal], sv.coords[:n_cal], sv.vel,
                              nthreads=nthreads)
    el = time.perf_counter() - t0
    rate = nc / max(el, 1e-9)
    per_trace = nc / n_cal
    S = int(max(n_cal, min(traces_np.shape[0], target_s * rate / max(per_trace, 1))))
    log(f"cpu baseline: calibration {nc:.3e} contrib in {el:.2f}s; sample {S} traces")
    t0 = time.perf_counter()
    _, _, nc = oracle.migrate(cfg, sv.ns, sv.dt, traces_np[:S], sv.coords[:S], sv.vel,
                              nthreads=nthreads)
    el = time.perf_counter() - t0
    return {"value": nc / el / 1e9, "unit": "G contributions/s", "cores": nthreads,
            "kind": "port",
            "sample": f"first {S} of {sv.ntr} traces over the whole {sv.cfg.end_iline}x"
                      f"{sv.cfg.end_xline}x{sv.ns} image, {nc:.3e} contributions in {el:.1f} s, "
                      f"{nthreads} job slabs (oracle/ktm_oracle.c, gcc -O3 -march=x86-64-v3, strict IEEE)"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU decomposition is "
                                   "described in DESIGN.md (image slabs need no collective)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from random_amd import PREP_REFERENCE, TRACE_F32, surveys, synth
    from random_amd.ktm import Plan

    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    sv = surveys.area3d(ntr=args.ntr, ncols=args.ncols)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    traces, coords = synth.traces_gpu(plan, sv, device=dev)
    km, fold, _ = plan.new_image()
    counter = torch.zeros(1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t_setup:.1f}s: {sv.ntr} traces x {sv.ns}, image "
        f"{plan.image_shape}, tm={plan.setup['trange_max']}")

    nblocks = (sv.ntr + args.block - 1) // args.block
    steps = args.steps
    warm = args.warmup

    def run_block(b):
        i0 = (b % nblocks) * args.block
        sl = slice(i0, min(i0 + args.block, sv.ntr))
        prepared, geom = plan.prep(traces[sl], coords[sl], PREP_REFERENCE, TRACE_F32)
        e_mig0.record()
        plan.migrate_block(prepared, geom, km, fold, counter=counter)
        e_mig1.record()

    e_mig0 = torch.cuda.Event(enable_timing=True)
    e_mig1 = torch.cuda.Event(enable_timing=True)
    for b in range(warm):
        run_block(b)
        torch.cuda.synchronize()
        log(f"warmup block {b}: migrate {e_mig0.elapsed_time(e_mig1):.1f} ms")
    plan.reset(km, fold)
    counter.zero_()
    torch.cuda.synchronize()

    mig_ms = []
    e_start = torch.cuda.Event(enable_timing=True)
    e_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e_start.record()
    for b in range(steps):
        run_block(b)
        e_mig1.synchronize()
        mig_ms.append(e_mig0.elapsed_time(e_mig1))
        log(f"step {b}: migrate kernel {mig_ms[-1]:.1f} ms")
    e_end.record()
    torch.cuda.synchronize()
    total_ms = e_start.elapsed_time(e_end)
    n_contrib = int(counter.item())
    gcs = n_contrib / (total_ms * 1e-3) / 1e9
    per_launch = n_contrib / steps
    achieved = per_launch * BYTES_FP32 / (np.mean(mig_ms) * 1e-3) / 1e9

    result = {
        "metric": METRIC, "value": round(gcs, 3), "unit": "G contributions/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": round(total_ms / steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded survey, Ricker diffractions + 10 dB noise, generated on GPU)",
        "config": {"workload": "3-D prestack KTM, one rank's share: 1e6 traces x 1500 samples at "
                               "4 ms, image 400x400x1500 at 25 m, aperture 2500 m, trangeMax 8; "
                               "step = one 50,000-trace block (prep + migration)",
                   "traces": sv.ntr, "ns": sv.ns, "image": list(plan.image_shape[:2]) + [sv.ns],
                   "block_traces": args.block, "contributions": n_contrib},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_migrate", "bytes_per_contribution": BYTES_FP32,
                     "avg_launch_ms": round(float(np.mean(mig_ms)), 2)},
    }
    if not args.no_cpu and not args.profile:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
        nthreads = max(1, min(nthreads, 64))
        sample = max(64, min(sv.ntr, 4000))
        tr_np = traces[:sample].cpu().numpy()
        result["cpu_baseline"] = cpu_baseline(sv, tr_np, args.cpu_seconds, nthreads)
        result["cpu_baseline"]["speedup_gpu_over_cpu"] = round(gcs / result["cpu_baseline"]["value"], 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
