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
# HBM bytes per launch of k_migrate on this workload, from rocprofv3 --pmc
# FETCH_SIZE (465,868,004 KiB, scaled x2.0) + WRITE_SIZE (1,882,132 KiB, x1.0),
# the scale factors being what the same counters read for a 1 GiB device copy
# (profiles/k_migrate_area3d_1e6.md)
TRAFFIC = 956_024_975_360


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
    ap.add_argument("--cpu-stride", type=int, default=100,
                    help="CPU baseline on every N-th trace (100: the 1 %% subset)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the ingest, velocity-gradient, adjoint/LSM, post-stack and EOM legs")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3")
    ap.add_argument("--profile-leg", default="fp32",
                    help="with --profile: the precision leg (a PRECISION_LEGS label) to run")
    ap.add_argument("--leg-steps", type=int, default=3, help="timed blocks per precision leg")
    ap.add_argument("--err-traces", type=int, default=10000,
                    help="traces of the FP64-oracle error check on the 50x50x1500 sub-volume")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy 3x (HBM counter calibration) and exit")
    ap.add_argument("--leg", choices=("vgrad", "adjoint", "poststack", "eom", "valu", "ingest",
                                      "job"),
                    help="only run this extra leg, print its JSON and exit (for profiling)")
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


def measured_peaks(dev, reps=5):
    """(HBM GB/s of a 1 GiB device copy, counting read + write; LDS read GB/s
    of the ktm_probe_lds kernel at full occupancy), measured on this GPU."""
    import torch
    from random_amd import _lib
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = reps * 2 * n * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    nblocks, iters = 8192, 2048       # 8 workgroups of 16 KB LDS per CU, 256 CUs x 4
    out = torch.empty(nblocks * 256, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()
    _lib.check(L.ktm_probe_lds(out.data_ptr(), nblocks, iters, stream), "ktm_probe_lds")
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        _lib.check(L.ktm_probe_lds(out.data_ptr(), nblocks, iters, stream), "ktm_probe_lds")
    e1.record()
    torch.cuda.synchronize()
    lds_gbs = reps * nblocks * 256 * iters * 64 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    return copy_gbs, lds_gbs


VALU_MODES = {0: "v_fma_f32", 1: "v_add_f32", 2: "v_sqrt_f32", 3: "v_cvt_u32_f32/v_cvt_f32_u32",
              4: "v_fma_f32 dependent chain", 5: "migration tap mix (2 fma, 2 sqrt, add, min, cvt, "
              "lshl_add, add)", 6: "1 v_sqrt_f32 : 4 v_fma_f32", 7: "1 v_sqrt_f32 : 1 v_fma_f32",
              8: "v_min_f32", 9: "v_rndne_f32", 10: "v_lshl_add_u32",
              11: "v_cmp_le_f32 (SGPR out; hazard nops: upper bound)", 12: "v_cndmask_b32",
              13: "v_add_u32", 14: "v_fma_f64", 15: "v_readlane_b32 (hazard nops: upper bound)",
              16: "v_rsq_f64", 17: "v_lshlrev_b64", 18: "v_mov_b32 (hazard nops: upper bound)"}


def valu_probe(dev, modes=tuple(VALU_MODES), waves_per_simd=(8,), target_ms=40.0):
    """VALU issue rate of each probe mode (ktm_probe_valu): wave-instructions
    per second over the whole chip and per CU-cycle at the shader clock the
    probe itself recorded (clock64 against the 100 MHz wall clock)."""
    import numpy as np
    import torch

    from random_amd import _lib
    L = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    ncu = torch.cuda.get_device_properties(dev).multi_processor_count
    out = {}
    for wps in waves_per_simd:
        nblocks = ncu * wps           # 4 waves per workgroup: wps waves per SIMD
        buf = torch.zeros(3 * nblocks, dtype=torch.int32, device=dev)
        for mode in modes:
            ops = L.ktm_probe_valu_ops(mode)
            iters = 256
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            for _ in range(6):        # grow iters until one launch takes ~target_ms
                torch.cuda.synchronize()
                e0.record()
                _lib.check(L.ktm_probe_valu(mode, buf.data_ptr(), nblocks, iters, stream),
                           "ktm_probe_valu")
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1)
                if ms >= 0.5 * target_ms:
                    break
                iters = min(1 << 24, int(iters * min(64.0, target_ms / max(ms, 1e-3))))
            w = buf.cpu().numpy().view(np.uint32).reshape(nblocks, 3).astype(np.float64)
            ghz = float(np.median(w[:, 0] / w[:, 1]) * 0.1)
            insts = float(nblocks) * 4 * iters * ops
            rate = insts / (ms * 1e-3)
            out[f"{mode}@{wps}"] = {"mode": VALU_MODES[mode], "waves_per_simd": wps,
                                    "wave_insts": insts, "ms": round(ms, 3),
                                    "wave_insts_per_s": rate, "shader_clock_ghz": round(ghz, 3),
                                    "per_cu_cycle": round(rate / (ncu * ghz * 1e9), 4)}
            log(f"valu probe mode {mode} ({VALU_MODES[mode]}), {wps} waves/SIMD: {ms:.2f} ms, "
                f"{rate:.4g} wave-insts/s, clock {ghz:.3f} GHz, "
                f"{out[f'{mode}@{wps}']['per_cu_cycle']:.3f} per CU-cycle")
    return out


# VALU instructions of one k_migrate launch on this workload (the first
# 50,000-trace block) by class, from two rocprofv3 --pmc passes
# (profiles/k_migrate_area3d_1e6.md).  Which instructions each counter
# counts was read off ktm_probe_valu, whose instruction streams are known:
# FMA_F32 = v_fma_f32/v_fmamk, ADD_F32 = v_add/v_sub_f32, TRANS_F32 =
# v_sqrt_f32, CVT = conversions, INT32 = v_add_u32 and v_lshl_add_u32;
# v_min/v_med3, v_rndne, v_cmp, v_cndmask, v_readlane are in no class.
KMIG_COUNTS = {
    "exact": {"kernel": "k_migrate<12, 2, 8, false, false>", "contributions": 1806379368000,
              "launch_ms": 942.06, "clock_ghz": 2.381, "valu": 488628380575,
              "fma_f32": 94969484010, "add_f32": 97736566828, "mul_f32": 5946021325,
              "trans_f32": 58496827853, "cvt": 30618638965, "int32": 61308780164,
              "fma_f64": 5207598064, "add_f64": 370325576, "mul_f64": 1532187535,
              "trans_f64": 744491152, "int64": 6386285760, "salu": 128120225353,
              "lds": 36591172744},
    "fast_kahan": {"kernel": "k_migrate<12, 2, 8, true, true>", "contributions": 1806379368000,
                   "launch_ms": 748.03, "clock_ghz": 2.377, "valu": 371632398974,
                   "fma_f32": 57804779776, "add_f32": 93506701952, "mul_f32": 2618300416,
                   "trans_f32": 57806059776, "cvt": 28921269888, "int32": 81815267442,
                   "fma_f64": 23040000, "add_f64": 0, "mul_f64": 15360000, "trans_f64": 3840000,
                   "int64": 5722754464, "salu": 44813163964, "lds": 62309817190},
}
# the probe mode whose measured rate prices each class (SIMD cycles per
# wave-instruction); INT32 at the v_add_u32 rate and the unclassified rest
# at the v_min_f32 rate (v_med3/v_rndne/v_cmp/v_cndmask measure the same)
CLASS_MODE = {"fma_f32": 0, "add_f32": 1, "mul_f32": 0, "trans_f32": 2, "cvt": 3, "int32": 13,
              "fma_f64": 14, "add_f64": 14, "mul_f64": 14, "trans_f64": 16, "int64": 17,
              "other": 8}


def valu_roofline(counts, probe, contributions, launch_ms, ncu, clock_ghz):
    """The VALU bound of one migration launch: the SIMD cycles its
    instructions occupy at the probe's measured rates (instructions per class
    scaled to this launch's contributions) over the SIMD cycles the launch
    had (4 SIMDs per CU at the kernel's shader clock).  frac <= 1 whenever
    the kernel issues no faster than the probe's independent streams."""
    scale = contributions / counts["contributions"]
    cyc = {c: 4.0 / probe[f"{m}@8"]["per_cu_cycle"] for c, m in CLASS_MODE.items()}
    n = {c: counts[c] * scale for c in CLASS_MODE if c != "other"}
    n["other"] = counts["valu"] * scale - sum(n.values())
    busy = {c: n[c] * cyc[c] for c in n}
    total = sum(busy.values())
    secs = launch_ms * 1e-3
    avail = 4.0 * ncu * clock_ghz * 1e9 * secs
    f64 = sum(busy[c] for c in ("fma_f64", "add_f64", "mul_f64", "trans_f64"))
    issue_peak = probe["0@8"]["wave_insts_per_s"]
    return {"bound": "valu", "unit": "G SIMD-cycles/s", "achieved": round(total / secs / 1e9, 1),
            "peak": round(4.0 * ncu * clock_ghz, 1), "frac": round(total / avail, 4),
            "kernel": counts["kernel"], "avg_launch_ms": round(launch_ms, 2),
            "shader_clock_ghz": clock_ghz,
            "cycles_per_wave_instruction": {c: round(v, 2) for c, v in cyc.items()},
            "valu_insts_per_contribution": round(counts["valu"] / counts["contributions"], 5),
            "busy_share": {c: round(b / total, 4) for c, b in busy.items()},
            "f64_exact_path_share": round(f64 / total, 4),
            "issue_rate_frac_of_fma_peak": round(counts["valu"] * scale / secs / issue_peak, 4)}


# ---- CPU-baseline leg: the only code in this file that runs the oracle
# (oracle/), as the CPU timing beside the GPU figures and as the checker of
# the GPU precision modes; every GPU figure comes from libktm.so. ----

def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sv, traces, nthreads, stride=100, ref_stride=1000):
    """Oracle FP32 (the reference algorithm restated, strict IEEE) on the
    deterministic 1 % subset -- every stride-th trace of the survey -- over the
    whole image, jobs = one image slab per thread (the reference's MPI
    decomposition, Kirchhoff.c:280-295,347-350), built -O3 -march=native on
    this host; the same restatement built with the reference's own flags
    (-O0 -ffast-math, Kirchhoff.c:3) timed on every ref_stride-th trace.
    Quotes the faster rate."""
    import numpy as np

    import oracle
    cfg = sv.cfg.replace(jobs_x=nthreads, jobs_y=1)
    t0 = time.perf_counter()
    oracle.build_native()
    log(f"cpu baseline: -march=native oracle built in {time.perf_counter() - t0:.1f}s")
    runs = {}
    for name, st, kw in (("O3_march_native", stride, {"native": True}),
                         ("O0_ffast_math", ref_stride, {"refflags": True})):
        idx = np.arange(0, sv.ntr, st)
        tr = traces[idx].cpu().numpy()
        co = np.ascontiguousarray(sv.coords[idx])
        t0 = time.perf_counter()
        _, _, nc = oracle.migrate(cfg, sv.ns, sv.dt, tr, co, sv.vel, nthreads=nthreads, **kw)
        el = time.perf_counter() - t0
        runs[name] = {"G_contributions_per_s": round(nc / el / 1e9, 4), "traces": int(idx.size),
                      "trace_stride": st, "contributions": int(nc), "seconds": round(el, 1)}
        log(f"cpu baseline {name}: {idx.size} traces, {nc:.3e} contributions in {el:.1f}s")
    best = max(runs, key=lambda k: runs[k]["G_contributions_per_s"])
    b = runs[best]
    return {"value": b["G_contributions_per_s"], "unit": "G contributions/s", "cores": nthreads,
            "kind": "port", "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"every {b['trace_stride']}th trace of the {sv.ntr}-trace survey "
                      f"({b['traces']} traces, {b['contributions']:.3e} contributions) over the "
                      f"whole {sv.cfg.end_iline}x{sv.cfg.end_xline}x{sv.ns} image in "
                      f"{b['seconds']} s on {nthreads} threads, one image slab each "
                      f"(oracle/ktm_oracle.c, build {best})",
            "builds": runs, "quoted_build": best}


def cpu_baseline_subvolume(sv, cfg, tr_np, co_np):
    """CPU-baseline leg, checker part: the oracle on a sub-volume, against which
    the GPU paths are measured -- its FP32 mode (the reference algorithm: the
    GPU's FP32 reference path must equal it bit for bit), FP64_DATA (the
    reference's FP32 traveltime index, FP64 data path: what storage and
    accumulation error is measured against), FP64_FULL (every float of the
    loop widened, the traveltime too), and the FP64 post-stack image of the
    FP64_DATA sums.  Returns a dict of arrays and counts."""
    import oracle
    nthreads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    out = {}
    t0 = time.perf_counter()
    out["km32"], out["fold32"], out["n32"] = oracle.migrate(cfg, sv.ns, sv.dt, tr_np, co_np, sv.vel,
                                                            mode=oracle.FP32, nthreads=nthreads)
    out["km64"], out["fold64"], out["n64"] = oracle.migrate(cfg, sv.ns, sv.dt, tr_np, co_np, sv.vel,
                                                            mode=oracle.FP64_DATA,
                                                            nthreads=nthreads)
    out["km64f"], out["fold64f"], _ = oracle.migrate(cfg, sv.ns, sv.dt, tr_np, co_np, sv.vel,
                                                     mode=oracle.FP64_FULL, nthreads=nthreads)
    _, aa = oracle.alias_filter(cfg, sv.ns, sv.dt)
    out["img64"] = oracle.poststack(out["km64"], out["fold64"], sv.dt, aa)
    out["img64f"] = oracle.poststack(out["km64f"], out["fold64f"], sv.dt, aa)
    log(f"oracle on the sub-volume (FP32, FP64_DATA, FP64_FULL, post-stack): {out['n64']:.3e} "
        f"contributions, {time.perf_counter() - t0:.1f}s")
    return out


def cpu_baseline_eom(cfg, sv, tr, coords, nthreads):
    """CPU-baseline leg for equivalent-offset migration: the oracle's csp()
    over the given traces, its CSP grid split into one slab per thread (the
    reference's decomposition, EOM.c), -O3 -march=native.  Returns
    (samples, seconds)."""
    import numpy as np

    import oracle
    oracle.build_native()
    c = cfg.replace(jobs_x=nthreads, jobs_y=1)
    t0 = time.perf_counter()
    _, _, ncpu = oracle.eom_all(c, sv.ns, np.float32(sv.dt), tr, coords, cfg.velocity(sv.ns),
                                nthreads=nthreads, native=True)
    return ncpu, time.perf_counter() - t0

# ---- end of the CPU-baseline leg ----


PRECISION_LEGS = (
    # label, prep mode, trace storage, accumulation | index flags
    ("fp32", "REFERENCE", "F32", "PLAIN"),
    ("fp32_closed_kahan", "CLOSED", "F32", "KAHAN"),
    ("fp16", "CLOSED", "F16", "KAHAN"),
    ("bf16", "CLOSED", "BF16", "KAHAN"),
    ("bf16_plain", "CLOSED", "BF16", "PLAIN"),
    ("fp32_closed_fast", "CLOSED", "F32", "PLAIN|FAST"),
    ("fp32_closed_kahan_fast", "CLOSED", "F32", "KAHAN|FAST"),
    ("fp16_fast", "CLOSED", "F16", "KAHAN|FAST"),
    ("bf16_fast", "CLOSED", "BF16", "KAHAN|FAST"),
)


def leg_args(prep, store, acc):
    """(prep mode, trace dtype, accum flags) of a PRECISION_LEGS row."""
    import random_amd as ra
    flags = 0
    for f in acc.split("|"):
        flags |= {"PLAIN": ra.ACCUM_PLAIN, "KAHAN": ra.ACCUM_KAHAN, "FAST": ra.INDEX_FAST}[f]
    return (getattr(ra, f"PREP_{prep}"), getattr(ra, f"TRACE_{store}"), flags)


def precision_errors(sv, plan, traces, coords, ntr, dev):
    """Every path on the 50 x 50 x ns sub-volume at columns 175..224 of both
    axes, migrating the first ntr traces of the survey: the FP32 reference
    path against the oracle's FP32 mode (bit for bit, fold and contribution
    count), and every path's image sum km and post-stack image against the
    FP64 oracle with the reference's FP32 traveltime index (FP64_DATA) and
    with an FP64 traveltime (FP64_FULL)."""
    import numpy as np

    from random_amd.ktm import Plan
    cfg = sv.cfg.replace(start_iline=175, end_iline=225, start_xline=175, end_xline=225,
                         jobs_x=10, jobs_y=2)
    sub = Plan(cfg, sv.ns, sv.dt, sv.vel)
    ref = cpu_baseline_subvolume(sv, cfg, traces[:ntr].cpu().numpy(), sv.coords[:ntr])
    out = {"error_check": f"50x50x{sv.ns} sub-volume (columns 175-224 x 175-224), first {ntr} "
                          f"traces, {ref['n64']:.3e} contributions; FP64_DATA = oracle with the "
                          f"reference FP32 traveltime index and an FP64 data path, FP64_FULL = "
                          f"every float of the loop widened (FP64 traveltime); image = post-stack "
                          f"output (Kirchhoff.c:1423-1479), km = the migrated sum before it"}
    n64 = np.linalg.norm(ref["km64"])
    m64 = float(np.abs(ref["km64"]).max())
    n64f = np.linalg.norm(ref["km64f"])
    ni64 = np.linalg.norm(ref["img64"])
    ni64f = np.linalg.norm(ref["img64f"])
    for label, prep, store, acc in PRECISION_LEGS:
        mode, dtype, accum = leg_args(prep, store, acc)
        km, fold, n = sub.migrate(traces[:ntr], coords[:ntr], mode=mode, dtype=dtype, accum=accum)
        img = sub.poststack(km, fold).cpu().numpy().astype(np.float64)
        k = km.cpu().numpy()
        d = k.astype(np.float64) - ref["km64"]
        e = {"rel_l2_vs_fp64": float(np.linalg.norm(d) / n64),
             "max_abs_vs_fp64_rel_to_max": float(np.abs(d).max() / m64),
             "rel_l2_vs_fp64_full": float(np.linalg.norm(k - ref["km64f"]) / n64f),
             "image_rel_l2_vs_fp64": float(np.linalg.norm(img - ref["img64"]) / ni64),
             "image_rel_l2_vs_fp64_full": float(np.linalg.norm(img - ref["img64f"]) / ni64f)}
        if label == "fp32":
            f = fold.cpu().numpy()
            e["vs_oracle_fp32"] = {
                "km_bit_identical_fraction": float(np.mean(k.view(np.uint32) ==
                                                           ref["km32"].view(np.uint32))),
                "fold_equal": bool(np.array_equal(f, ref["fold32"])),
                "contributions": n, "contributions_oracle": ref["n32"],
                "contributions_equal": bool(n == ref["n32"])}
        out.setdefault(label, {}).update(e)
        log(f"{label}: {e}")
    return out


def index_mismatch(sv, plan, geom, dev, nsamp=1 << 22, seed=5):
    """How often the fast form's floor(q) differs from the reference's
    traveltime index, on (trace, column, sample) triples drawn from the
    workload: random traces of a block, random image columns within the CMP
    aperture, uniform samples; q and the reference index come from the
    traveltime self-test kernel (ktm_selftest_traveltime)."""
    import numpy as np
    import torch

    from random_amd import _lib
    rng = np.random.default_rng(seed)
    g = geom.cpu().numpy().view(np.float32).reshape(-1, 8)
    ns, dt = sv.ns, np.float32(sv.dt)
    ncol = sv.cfg.end_iline - sv.cfg.start_iline
    d = sv.cfg.iline_distance
    ap2 = np.float32(sv.cfg.aperture) ** 2
    rows = []
    have = 0
    while have < nsamp:
        i = rng.integers(0, g.shape[0], nsamp)
        X = rng.integers(0, ncol, nsamp)
        Y = rng.integers(0, ncol, nsamp)
        IX = (X * d).astype(np.float32)
        IY = (Y * d).astype(np.float32)
        gi = g[i]
        ok = (IX - gi[:, 4]) ** 2 + (IY - gi[:, 5]) ** 2 < ap2
        t = rng.integers(0, ns, nsamp)
        dis2 = (IX - gi[:, 0]) ** 2 + (IY - gi[:, 1]) ** 2
        dir2 = (IX - gi[:, 2]) ** 2 + (IY - gi[:, 3]) ** 2
        v2 = sv.vel[X, Y, t] * sv.vel[X, Y, t]
        rows.append(np.stack([dis2, dir2, t.astype(np.float32), v2], 1)[ok])
        have += int(ok.sum())
    inp = np.ascontiguousarray(np.concatenate(rows)[:nsamp], np.float32)
    d_in = torch.from_numpy(inp).to(dev)
    n = inp.shape[0]
    qf = torch.empty(n, dtype=torch.float32, device=dev)
    qr = torch.empty_like(qf)
    tk = torch.empty(n, dtype=torch.int32, device=dev)
    tr = torch.empty_like(tk)
    _lib.check(_lib.lib().ktm_selftest_traveltime(d_in.data_ptr(), n, float(dt), ns, qf.data_ptr(),
                                                  qr.data_ptr(), tk.data_ptr(), tr.data_ptr(),
                                                  torch.cuda.current_stream().cuda_stream),
               "ktm_selftest_traveltime")
    q = qf.cpu().numpy()
    t_fast = np.where(q < ns, np.floor(q), ns).astype(np.int64)
    t_ref = tr.cpu().numpy().astype(np.int64)
    t_ker = tk.cpu().numpy().astype(np.int64)
    mism = t_fast != t_ref
    out = {"samples": int(n), "fast_form_mismatch_rate": float(mism.mean()),
           "exact_path_mismatch_rate": float(np.mean(t_ker != t_ref))}
    log(f"index mismatch: {out}")
    return out


def ingest_leg(sv, plan, traces, block, dev, reps=5):
    """Device SEG-Y ingest of one block of raw IBM trace records (as read from
    a file, Kirchhoff.c:1281-1290): fused byte swap + IBM decode + header
    words + FP32 reference prep (ktm_ingest), timed beside the decode alone
    and the prep of already-decoded traces."""
    import numpy as np
    import torch

    from random_amd import _lib
    from random_amd._lib import PREP_REFERENCE, SEGY_IBM, TRACE_F32
    from random_amd.ktm import segy_decode
    n, ns = min(block, traces.shape[0]), plan.ns
    rec = 240 + 4 * ns
    hdr = np.zeros((n, 240), np.uint8)
    hdr[:, 72:88] = sv.coords[:n].astype(">i4").view(np.uint8).reshape(n, 16)
    d_hdr = torch.from_numpy(hdr).to(dev)
    raw = torch.empty(n * rec, dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().ktm_segy_encode(traces[:n].data_ptr(), n, ns, SEGY_IBM, d_hdr.data_ptr(),
                                          raw.data_ptr(), torch.cuda.current_stream().cuda_stream),
               "ktm_segy_encode")
    dec, co = segy_decode(raw, n, ns, SEGY_IBM)

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_ingest = timed(lambda: plan.ingest(raw, n, SEGY_IBM, PREP_REFERENCE, TRACE_F32))
    ms_decode = timed(lambda: segy_decode(raw, n, ns, SEGY_IBM))
    ms_prep = timed(lambda: plan.prep(dec, co, PREP_REFERENCE, TRACE_F32))
    b_ingest = n * (rec + 4 * plan.row + 32)          # records in, prepared rows + geometry out
    b_decode = n * (rec + 4 * ns + 16)
    b_prep = n * (4 * ns + 16 + 4 * plan.row + 32)
    out = {"traces": n, "record_bytes": rec, "format": "IBM (SEG-Y format 1)"}
    for k, ms, b in (("ingest_fused", ms_ingest, b_ingest), ("decode_only", ms_decode, b_decode),
                     ("prep_from_fp32", ms_prep, b_prep)):
        gbs = b / (ms * 1e-3) / 1e9
        out[k] = {"ms": round(ms, 3), "bytes": int(b), "GB_per_s": round(gbs, 1),
                  "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"ingest {k}: {ms:.3f} ms, {gbs:.0f} GB/s")
    out["file"] = file_leg(plan, raw, hdr, n, dev)
    return out


def job_only_leg(dev, n=200_000):
    """The job leg alone on the first n traces of the 3-D survey."""
    import numpy as np

    from random_amd import surveys, synth
    from random_amd.ktm import Plan
    sv = surveys.area3d()
    sv.coords = np.ascontiguousarray(sv.coords[:n])
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    traces, _ = synth.traces_gpu(plan, sv, device=dev)
    del plan
    return job_leg(sv, traces, dev, ntr=n)


def ingest_only_leg(dev, n=50_000):
    """The ingest leg alone on the first block of the 3-D survey."""
    import numpy as np

    from random_amd import surveys, synth
    from random_amd.ktm import Plan
    sv = surveys.area3d()
    sv.coords = np.ascontiguousarray(sv.coords[:n])
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    traces, _ = synth.traces_gpu(plan, sv, device=dev)
    return ingest_leg(sv, plan, traces, n, dev)


def file_leg(plan, raw, hdr, n, dev):
    """File I/O of the same block, reported apart from the in-memory figures:
    the records written as an IBM SEG-Y file (3600-byte file header, ns and
    dt in every trace header), then read back three ways: the host reader
    (ktm_segy_read: fread + IBM decode on one core, as Kirchhoff.c:1281-1290
    does), the raw records into pinned memory, and those copied to the device
    and ingested (ktm_ingest).  The file was just written, so the reads come
    from the page cache, not the disk."""
    import os
    import shutil
    import tempfile

    import numpy as np
    import torch

    from random_amd import _lib
    from random_amd._lib import PREP_REFERENCE, SEGY_IBM, TRACE_F32
    ns = plan.ns
    rec = 240 + 4 * ns
    host = raw.cpu().numpy().reshape(n, rec).copy()
    host[:, 114:116] = np.array([ns], dtype=">u2").view(np.uint8)
    host[:, 116:118] = np.array([round(plan.dt * 1e6)], dtype=">u2").view(np.uint8)
    d = tempfile.mkdtemp(prefix="ktm_io_")
    path = os.path.join(d, "block.sgy")
    try:
        t0 = time.perf_counter()
        with open(path, "wb") as f:
            f.write(bytes(3600))
            f.write(host.tobytes())
            f.flush()
            os.fsync(f.fileno())
        t_write = time.perf_counter() - t0
        nbytes = 3600 + host.nbytes
        samples = np.empty((n, ns), np.float32)
        coords = np.empty((n, 4), np.int32)
        L = _lib.lib()
        t0 = time.perf_counter()
        _lib.check(L.ktm_segy_read(path.encode(), 0, n, 0, samples.ctypes.data, coords.ctypes.data,
                                   None), "ktm_segy_read")
        t_host = time.perf_counter() - t0
        pinned = torch.empty(host.nbytes, dtype=torch.uint8, pin_memory=True)
        view = pinned.numpy()
        t0 = time.perf_counter()
        with open(path, "rb") as f:
            f.seek(3600)
            got = f.readinto(memoryview(view))
        t_read = time.perf_counter() - t0
        assert got == host.nbytes
        d_raw = torch.empty(host.nbytes, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d_raw.copy_(pinned, non_blocking=True)
        torch.cuda.synchronize()
        t_h2d = time.perf_counter() - t0
        t0 = time.perf_counter()
        prepared, geom, dco = plan.ingest(d_raw, n, SEGY_IBM, PREP_REFERENCE, TRACE_F32)
        torch.cuda.synchronize()
        t_ing = time.perf_counter() - t0
        # the coordinates through the file: host reader, device ingest, headers written
        want = host[:, 72:88].copy().view(">i4").astype(np.int32)
        ok = bool(np.array_equal(coords, want) and np.array_equal(dco.cpu().numpy(), want))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    gb = nbytes / 1e9
    res = {"file_bytes": int(nbytes),
           "write_s": round(t_write, 3), "write_GB_per_s": round(gb / t_write, 2),
           "host_read_decode_s": round(t_host, 3),
           "host_read_decode_GB_per_s": round(gb / t_host, 2),
           "raw_read_pinned_s": round(t_read, 3), "raw_read_GB_per_s": round(gb / t_read, 2),
           "h2d_s": round(t_h2d, 4), "h2d_GB_per_s": round(host.nbytes / 1e9 / t_h2d, 1),
           "device_ingest_s": round(t_ing, 4),
           "file_to_prepared_on_device_s": round(t_read + t_h2d + t_ing, 3),
           "coords_round_trip_equal": ok,
           "note": "page-cache reads of a just-written file (fsync'd), one host thread"}
    log(f"file: {gb:.2f} GB written {t_write:.2f} s; host read+decode {t_host:.2f} s; raw read "
        f"{t_read:.3f} s, H2D {t_h2d * 1e3:.1f} ms, ingest {t_ing * 1e3:.1f} ms")
    return res


def vgrad_leg(dev, reps=3, fd_nodes=10):
    """Velocity gradient of the gather-flatness objective on the 600 x 240
    line (2001 samples, 2000 columns x 64 offset classes, 100 x 200 spline
    nodes, time chunks of 250 samples): objective alone, objective +
    gradient, and central differences (h = 1 m/s) at seeded nodes."""
    import numpy as np
    import torch

    from random_amd import surveys, synth
    from random_amd.ktm import Plan
    from random_amd.vgrad import VelocityGradient, central_differences, linear_nodes
    sv = surveys.line2d_vgrad()
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    tr, co = synth.traces_gpu(plan, sv, device=dev)
    vg = VelocityGradient(plan, tr, co, nt_nodes=100, nx_nodes=200, chunk=250)
    V = linear_nodes(sv.ns, sv.dt, 100, 200, 1500.0, 400.0, scale=0.97)
    vg(V)

    def timed(grad):
        t = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            vg(V, grad=grad)
            t.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(t))

    ms_j, ms_g = timed(False), timed(True)
    per_pass = vg.pairs * sv.ns
    J, g, _ = vg(V)
    g = g.cpu().numpy()
    rng = np.random.default_rng(20240042)
    nodes = [(int(rng.integers(2, 98)), int(rng.integers(2, 198))) for _ in range(fd_nodes)]
    fd = central_differences(vg, V, nodes, h=1.0)
    ad = np.array([g[i, j] for i, j in nodes])
    rel = np.abs(ad - fd) / np.abs(ad)
    log(f"vgrad: J {ms_j:.1f} ms, J+grad {ms_g:.1f} ms, {per_pass:.3e} contributions per pass, "
        f"FD max rel {rel.max():.2e}")
    return {"workload": "2-D line 600 shots x 240 receivers (144,000 traces), 2001 samples, "
                        "2000 columns x 64 offset classes, v on 100 x 200 B-spline nodes, "
                        "time chunks of 250 samples",
            "contributions_per_pass": int(per_pass), "ms_objective": round(ms_j, 1),
            "ms_objective_and_gradient": round(ms_g, 1),
            "G_contributions_per_s_objective": round(per_pass / (ms_j * 1e-3) / 1e9, 1),
            "fd_nodes": fd_nodes, "fd_h_m_per_s": 1.0,
            "fd_max_rel_err": float(rel.max()), "fd_median_rel_err": float(np.median(rel))}


def adjoint_leg(dev, reps=3, niter=20):
    """The adjoint pair and least-squares migration on the 200 x 240 line
    (48,000 traces x 1501 samples, 1600 columns): one application of L^T
    (linear migration: closed-form filter, compensated sums) and of L
    (demigration), the dot-product test on seeded Gaussian d and m, and
    niter CGLS iterations (time-gain preconditioned) from data modelled by L
    on 50 point scatterers."""
    import numpy as np
    import torch

    from random_amd import surveys
    from random_amd.ktm import Plan
    from random_amd.lsm import Operator, cgls, dot64, dot_test, time_gain
    sv = surveys.line2d_lsm()
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    co = torch.from_numpy(sv.coords).to(dev)
    op = Operator(plan, co)
    g = torch.Generator(device=dev).manual_seed(1)
    d = torch.randn((sv.ntr, sv.ns), generator=g, device=dev)
    m = torch.randn(plan.image_shape, generator=g, device=dev)
    n = plan.migrate(d, co)[2]
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    ms_a, ms_f = [], []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        e0.record()
        op.adjoint(d)
        e1.record()
        op.forward(m)
        e2.record()
        torch.cuda.synchronize()
        ms_a.append(e0.elapsed_time(e1))
        ms_f.append(e1.elapsed_time(e2))
    ms_a, ms_f = float(np.median(ms_a[1:])), float(np.median(ms_f[1:]))
    _, _, rel = dot_test(op, d, m)
    m_true = torch.from_numpy(surveys.point_reflectivity(sv, plan.image_shape)).to(dev)
    d_obs = op.forward(m_true)
    w = time_gain(op, plan.image_shape, nprobe=4, seed=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m_est, hist = cgls(op, d_obs, niter=niter, precond=w)
    torch.cuda.synchronize()
    ms_cg = (time.perf_counter() - t0) * 1e3
    mig = op.adjoint(d_obs)
    corr = dot64(m_est, m_true) / (dot64(m_est, m_est) * dot64(m_true, m_true)) ** 0.5
    corr0 = dot64(mig, m_true) / (dot64(mig, mig) * dot64(m_true, m_true)) ** 0.5
    log(f"adjoint: L^T {ms_a:.2f} ms, L {ms_f:.2f} ms, {n:.3e} contributions, dot rel {rel:.2e}, "
        f"CGLS {niter} it {ms_cg:.0f} ms, residual {hist[-1][0] / hist[0][0]:.3f}")
    # the pair with the relaxed traveltime index (INDEX_FAST on both sides)
    opf = Operator(plan, co, fast=True)
    ms_af, ms_ff = [], []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        e0.record()
        opf.adjoint(d)
        e1.record()
        opf.forward(m)
        e2.record()
        torch.cuda.synchronize()
        ms_af.append(e0.elapsed_time(e1))
        ms_ff.append(e1.elapsed_time(e2))
    ms_af, ms_ff = float(np.median(ms_af[1:])), float(np.median(ms_ff[1:]))
    _, _, rel_f = dot_test(opf, d, m)
    d_obs_f = opf.forward(m_true)
    wf = time_gain(opf, plan.image_shape, nprobe=4, seed=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m_f, hist_f = cgls(opf, d_obs_f, niter=niter, precond=wf)
    torch.cuda.synchronize()
    ms_cg_f = (time.perf_counter() - t0) * 1e3
    corr_f = dot64(m_f, m_true) / (dot64(m_f, m_f) * dot64(m_true, m_true)) ** 0.5
    log(f"adjoint, fast index: L^T {ms_af:.2f} ms, L {ms_ff:.2f} ms, dot rel {rel_f:.2e}, "
        f"CGLS {ms_cg_f:.0f} ms, residual {hist_f[-1][0] / hist_f[0][0]:.3f}")
    return {"workload": "2-D line 200 shots x 240 receivers (48,000 traces), 1501 samples, "
                        "1600 columns; 50 point scatterers",
            "contributions_per_application": int(n),
            "ms_migration_adjoint": round(ms_a, 2), "ms_demigration": round(ms_f, 2),
            "G_contributions_per_s_migration": round(n / (ms_a * 1e-3) / 1e9, 1),
            "G_contributions_per_s_demigration": round(n / (ms_f * 1e-3) / 1e9, 1),
            "dot_test_rel": float(rel), "cgls_iterations": niter,
            "cgls_ms": round(ms_cg, 1),
            "cgls_relative_residual": round(hist[-1][0] / hist[0][0], 4),
            "correlation_with_truth": {"migration": round(corr0, 4), "lsm": round(corr, 4)},
            "fast_index_pair": {"ms_migration_adjoint": round(ms_af, 2),
                                "ms_demigration": round(ms_ff, 2), "dot_test_rel": float(rel_f),
                                "cgls_ms": round(ms_cg_f, 1),
                                "cgls_relative_residual": round(hist_f[-1][0] / hist_f[0][0], 4),
                                "correlation_with_truth_lsm": round(corr_f, 4)}}


def poststack_leg(dev, plan=None, reps=3):
    """The steps after the migration on the 3-D image (400 x 400 x 1500,
    one job slab): the post-stack filter of Kirchhoff.c:1423-1479 (derivative,
    fold, low-pass as an FP64 circular convolution by FFT) and the image<job>.su
    trace records (big-endian headers and samples, Kirchhoff.c:1437-1598),
    on a seeded image and fold."""
    import numpy as np
    import torch

    from random_amd import _lib, surveys
    from random_amd.ktm import Plan
    if plan is None:
        sv = surveys.area3d(ntr=1000)
        plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    g = torch.Generator(device=dev).manual_seed(7)
    km = torch.randn(plan.image_shape, generator=g, device=dev)
    fold = torch.randint(1, 200, plan.image_shape, generator=g, device=dev).float()
    nx, ny, nh, ns = plan.setup["nx"], plan.setup["ny"], plan.nh, plan.ns
    nrec = nx * ny * nh
    rec = torch.empty(nrec * (240 + 4 * ns), dtype=torch.uint8, device=dev)
    L = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    img = plan.poststack(km, fold)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    ms_p, ms_r = [], []
    for _ in range(reps):
        torch.cuda.synchronize()
        e0.record()
        img = plan.poststack(km, fold)
        e1.record()
        _lib.check(L.ktm_image_records(plan._h, img.data_ptr(), 0, rec.data_ptr(), stream),
                   "ktm_image_records")
        e2.record()
        torch.cuda.synchronize()
        ms_p.append(e0.elapsed_time(e1))
        ms_r.append(e1.elapsed_time(e2))
    ms_p, ms_r = float(np.median(ms_p)), float(np.median(ms_r))
    ntraces = int(np.prod(plan.image_shape[:3]))
    # the FFT form: two traces per complex transform of N >= 2 ns - 1 points
    # (256 ... 8192 and 3072), forward and inverse, 5 N log2 N flop each (the
    # radix-2 count)
    nfft = next(n for n in (256, 512, 1024, 2048, 3072, 4096, 8192) if n >= 2 * ns - 1)
    flop = (ntraces / 2.0) * 2 * 5.0 * nfft * np.log2(nfft)
    bytes_p = 3 * 4 * ntraces * ns            # km and fold in, image out
    bytes_r = 4 * nrec * ns + rec.numel()     # image in, records out
    log(f"poststack: {ms_p:.2f} ms ({bytes_p / ms_p / 1e6:.0f} GB/s, {flop / ms_p / 1e9:.1f} "
        f"TFLOP/s FP64 at 5 N log2 N); records {ms_r:.2f} ms ({bytes_r / ms_r / 1e6:.0f} GB/s)")
    return {"workload": f"3-D image {plan.image_shape[0]} x {plan.image_shape[1]} x {ns}, "
                        f"{ntraces} image traces",
            "ms_poststack": round(ms_p, 2),
            "poststack_method": f"FFT of {nfft} points in FP64, two traces per complex transform",
            "poststack_bytes": int(bytes_p),
            "poststack_GB_per_s": round(bytes_p / (ms_p * 1e-3) / 1e9, 1),
            "poststack_hbm_frac": round(bytes_p / (ms_p * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
            "poststack_TFLOP_per_s_fp64_fft_count": round(flop / (ms_p * 1e-3) / 1e12, 2),
            "ms_image_records": round(ms_r, 2),
            "image_records_bytes": int(bytes_r),
            "image_records_GB_per_s": round(bytes_r / (ms_r * 1e-3) / 1e9, 1),
            "image_records_hbm_frac": round(bytes_r / (ms_r * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}


def host_leg(sv, traces, ntr=100_000):
    """The host-array path (ktm_migrate, the one-call driver): the first ntr
    traces of the 3-D survey from pageable host memory, plan set-up (the
    0.96 GB velocity model uploaded), blocks of ntblock traces copied in and
    migrated into the whole image, the post-stack filter, and km, fold and
    image copied back -- the PCIe-inclusive rate beside the kernels' own."""
    from random_amd.ktm import migrate_host
    tr = traces[:ntr].cpu().numpy()
    co = sv.coords[:ntr]
    km, fold, image, st = migrate_host(sv.cfg, sv.ns, sv.dt, tr, co, sv.vel)
    n = int(st["contributions"])
    img_bytes = km.nbytes + fold.nbytes + image.nbytes
    del km, fold, image
    log(f"host path: {ntr} traces, {st['total_ms']:.0f} ms total, {st['migrate_ms']:.0f} ms in "
        f"the migration kernels")
    return {"traces": ntr, "contributions": n, "ms_total": round(st["total_ms"], 1),
            "ms_migration_kernels": round(st["migrate_ms"], 1),
            "G_contributions_per_s_total": round(n / (st["total_ms"] * 1e-3) / 1e9, 1),
            "G_contributions_per_s_kernels": round(n / (st["migrate_ms"] * 1e-3) / 1e9, 1),
            "host_bytes_in": int(tr.nbytes + co.nbytes + sv.vel.nbytes),
            "host_bytes_out": img_bytes}


def job_leg(sv, traces, dev, ntr=200_000, job=0):
    """kirchhoff(job) end to end, the reference's entry point: the first ntr
    traces of the 3-D survey written as an IBM SEG-Y file (encoded on the
    device), main()'s setup (ktm_globals_init: the file counted, the
    velocity model uploaded; not timed), then one kirchhoff(job) call -- the
    blocks read from the file, copied, decoded and migrated, the post-stack
    filter, and image<job>.su encoded and written -- against the device time
    of its migration launches (ktm_last_stats).  The blocks' input overlaps
    the previous block's migration, so what the wall time adds to the
    kernels is the first block's input and the output."""
    import ctypes as C
    import os
    import shutil
    import tempfile

    import numpy as np
    import torch

    from random_amd import _lib
    from random_amd._lib import SEGY_IBM
    n, ns = min(ntr, traces.shape[0]), sv.ns
    rec = 240 + 4 * ns
    hdr = np.zeros((n, 240), np.uint8)
    hdr[:, 72:88] = sv.coords[:n].astype(">i4").view(np.uint8).reshape(n, 16)
    hdr[:, 114:116] = np.array([ns], dtype=">u2").view(np.uint8)
    hdr[:, 116:118] = np.array([round(sv.dt * 1e6)], dtype=">u2").view(np.uint8)
    d_hdr = torch.from_numpy(hdr).to(dev)
    raw = torch.empty(n * rec, dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().ktm_segy_encode(traces[:n].data_ptr(), n, ns, SEGY_IBM, d_hdr.data_ptr(),
                                          raw.data_ptr(), torch.cuda.current_stream().cuda_stream),
               "ktm_segy_encode")
    host = raw.cpu().numpy()
    del raw, d_hdr
    d = tempfile.mkdtemp(prefix="ktm_job_")
    path = os.path.join(d, "input.sgy")
    img_dir = os.path.join(d, "image_dir")
    L = _lib.lib()
    try:
        with open(path, "wb") as f:
            f.write(bytes(3600))
            f.write(host.tobytes())
        del host
        c = sv.cfg.to_c()
        vel = np.ascontiguousarray(sv.vel, np.float32)
        _lib.check(L.ktm_globals_init(C.byref(c), path.encode(), None,
                                      vel.ctypes.data_as(C.c_void_p), vel.shape[2],
                                      np.float32(sv.dt), img_dir.encode()), "ktm_globals_init")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.kirchhoff(job)
        wall = time.perf_counter() - t0
        _lib.check(L.ktm_last_status(), "kirchhoff")
        st = _lib.KtmStats()
        _lib.check(L.ktm_last_stats(C.byref(st)), "ktm_last_stats")
        out_bytes = os.path.getsize(os.path.join(img_dir, f"image{job}.su"))
    finally:
        L.ktm_globals_free()
        shutil.rmtree(d, ignore_errors=True)
    blocks = -(-n // sv.cfg.ntblock)
    log(f"job: kirchhoff({job}) on {n} traces ({blocks} blocks) {wall:.2f} s, migration kernels "
        f"{st.migrate_ms / 1e3:.2f} s")
    return {"traces": n, "blocks": int(blocks), "input_file_bytes": 3600 + n * rec,
            "image_file_bytes": int(out_bytes), "contributions": int(st.contributions),
            "wall_s": round(wall, 3), "migration_kernels_s": round(st.migrate_ms / 1e3, 3),
            "kernel_share_of_wall": round(st.migrate_ms / 1e3 / wall, 4),
            "G_contributions_per_s_end_to_end": round(st.contributions / wall / 1e9, 1),
            "note": "page-cache reads of a just-written file; image written to the page cache"}


def eom_leg(dev, reps=3, cpu_stride=10):
    """Equivalent-offset migration (EOM.c csp()) of the 40-shot line:
    4800 traces x 1001 samples into 800 CSPs x 160 offset bins, plus the post
    filter; the CPU oracle timed on the first traces of the same survey."""
    import numpy as np
    import torch

    from random_amd import surveys
    from random_amd.eom import EomPlan
    sv, cfg = surveys.line2d_eom()
    tr = surveys.traces_numpy(sv)
    plan = EomPlan(cfg, sv.ns, sv.dt)
    d_tr, d_co = torch.from_numpy(tr).to(dev), torch.from_numpy(sv.coords).to(dev)
    plan.migrate(d_tr, d_co)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    ms_b, ms_p = [], []
    for _ in range(reps):
        csp, fold = plan.new_gathers(dev)
        counter = torch.zeros(1, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        e0.record()
        plan.block(d_tr, d_co, csp, fold, -1, counter)
        e1.record()
        plan.post(csp, fold)
        e2.record()
        torch.cuda.synchronize()
        ms_b.append(e0.elapsed_time(e1))
        ms_p.append(e1.elapsed_time(e2))
    n = int(counter.item())
    ms_b, ms_p = float(np.median(ms_b)), float(np.median(ms_p))
    nthreads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    ncpu, el = cpu_baseline_eom(cfg, sv, tr[::cpu_stride], sv.coords[::cpu_stride], nthreads)
    log(f"eom: binning {ms_b:.2f} ms, post {ms_p:.2f} ms, {n:.3e} samples; cpu {ncpu:.3e} in {el:.2f}s")
    return {"workload": "40-shot line, 4800 traces x 1001 samples -> 800 CSPs x 160 "
                        "offset bins of 12.5 m, aperture 2000 m, v = 2000 m/s",
            "evaluated_samples": n, "ms_binning": round(ms_b, 2), "ms_post": round(ms_p, 2),
            "G_samples_per_s": round(n / (ms_b * 1e-3) / 1e9, 2),
            "cpu_oracle_G_samples_per_s": round(ncpu / el / 1e9, 4),
            "cpu_sample": f"every {cpu_stride}th trace ({tr[::cpu_stride].shape[0]} traces), "
                          f"{nthreads} threads (one CSP slab each), -O3 -march=native, "
                          f"{el:.2f} s"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.leg:
        import torch
        fn = {"vgrad": vgrad_leg, "adjoint": adjoint_leg, "poststack": poststack_leg,
              "eom": eom_leg, "ingest": ingest_only_leg, "job": job_only_leg,
              "valu": lambda d: valu_probe(d, waves_per_simd=(1, 2, 4, 8))}[args.leg]
        print(json.dumps(fn(torch.device("cuda:0"))), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only; across GPUs each would migrate its "
                                   "own image slabs (job ids), which needs no collective",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from random_amd import (ACCUM_KAHAN, ACCUM_PLAIN, INDEX_FAST, PREP_CLOSED, PREP_REFERENCE,
                            TRACE_BF16, TRACE_F16, TRACE_F32, surveys, synth)
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
    e_mig0 = torch.cuda.Event(enable_timing=True)
    e_mig1 = torch.cuda.Event(enable_timing=True)

    def run_mode(steps, warm, mode, dtype, accum, label):
        """warm + steps trace blocks into a fresh image; returns (G contrib/s of
        the timed steps incl. prep, per-launch migrate ms list, contributions)."""
        km, fold, comp = plan.new_image(kahan=(accum & ~INDEX_FAST) == ACCUM_KAHAN)
        counter = torch.zeros(1, dtype=torch.int64, device=dev)

        def run_block(b):
            i0 = (b % nblocks) * args.block
            sl = slice(i0, min(i0 + args.block, sv.ntr))
            prepared, geom = plan.prep(traces[sl], coords[sl], mode, dtype)
            e_mig0.record()
            plan.migrate_block(prepared, geom, km, fold, comp, dtype=dtype, mode=mode,
                               accum=accum, counter=counter)
            e_mig1.record()

        for b in range(warm):
            run_block(b)
            torch.cuda.synchronize()
            log(f"{label} warmup block {b}: migrate {e_mig0.elapsed_time(e_mig1):.1f} ms")
        plan.reset(km, fold, comp)
        counter.zero_()
        mig_ms = []
        e_start = torch.cuda.Event(enable_timing=True)
        e_end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e_start.record()
        for b in range(steps):
            run_block(b)
            e_mig1.synchronize()
            mig_ms.append(e_mig0.elapsed_time(e_mig1))
            log(f"{label} step {b}: migrate kernel {mig_ms[-1]:.1f} ms")
        e_end.record()
        torch.cuda.synchronize()
        total_ms = e_start.elapsed_time(e_end)
        n = int(counter.item())
        # fold starts at 1 and counts the contributions that add into the image
        # (T < ns, offset class in range: the "effective" ones)
        run_mode.effective = int((fold.double() - 1.0).sum().item())
        del km, fold, comp
        return n / (total_ms * 1e-3) / 1e9, mig_ms, n, total_ms

    steps, warm = args.steps, args.warmup
    if args.profile and args.profile_leg != "fp32":
        row = [r for r in PRECISION_LEGS if r[0] == args.profile_leg][0]
        mode, dtype, accum = leg_args(*row[1:4])
        g, ms, n, tot = run_mode(steps, warm, mode, dtype, accum, row[0])
        print(json.dumps({"leg": row[0], "value": round(g, 3), "avg_launch_ms": float(np.mean(ms)),
                          "contributions": n}), flush=True)
        return
    gcs, mig_ms, n_contrib, total_ms = run_mode(steps, warm, PREP_REFERENCE, TRACE_F32,
                                                ACCUM_PLAIN, "fp32")
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
                   "block_traces": args.block, "contributions": n_contrib,
                   "effective_contributions": run_mode.effective},
    }
    avg_ms = float(np.mean(mig_ms))
    if args.profile:
        result["roofline"] = {"kernel": KMIG_COUNTS["exact"]["kernel"], "avg_launch_ms": avg_ms}
    else:
        ncu = torch.cuda.get_device_properties(dev).multi_processor_count
        probe = valu_probe(dev, modes=sorted(set(CLASS_MODE.values())))
        result["roofline"] = valu_roofline(KMIG_COUNTS["exact"], probe, per_launch, avg_ms, ncu,
                                           KMIG_COUNTS["exact"]["clock_ghz"])
        copy_gbs, lds_gbs = measured_peaks(dev)
        traffic_gbs = TRAFFIC / (avg_ms * 1e-3) / 1e9
        # the kernel reads one 4-byte word of LDS per contribution (the three
        # taps pre-combined into F(T) when the trace is staged)
        lds_read = per_launch * 4 / (avg_ms * 1e-3) / 1e9
        result["roofline"].update({
            "traffic": TRAFFIC,
            "hbm_traffic_GB_per_s": round(traffic_gbs, 1),
            "hbm_traffic_frac": round(traffic_gbs / HBM_PEAK_GBS, 4),
            "hbm_copy_peak_measured": round(copy_gbs, 1),
            "hbm_traffic_frac_of_copy_peak": round(traffic_gbs / copy_gbs, 4),
            "algorithmic_bytes_rate": {
                "bytes_per_contribution": BYTES_FP32, "GB_per_s": round(achieved, 1),
                "note": "the three FP32 taps of Kirchhoff.c:1392 per contribution; the kernel "
                        "reads one pre-combined tap per contribution from LDS, so this rate "
                        "is not HBM traffic and exceeds the HBM peak"},
            "lds": {"read_bytes_per_contribution": 4, "achieved": round(lds_read, 1),
                    "peak_measured": round(lds_gbs, 1), "frac": round(lds_read / lds_gbs, 4),
                    "unit": "GB/s"}})
    if not args.profile and args.leg_steps > 0:
        legs = {}
        for label, prep, store, acc in PRECISION_LEGS:
            if label == "fp32":
                continue
            mode, dtype, accum = leg_args(prep, store, acc)
            g, ms, n, _ = run_mode(args.leg_steps, 1, mode, dtype, accum, label)
            legs[label] = {"value": round(g, 3), "speedup_vs_fp32": round(g / gcs, 3),
                           "avg_launch_ms": round(float(np.mean(ms)), 2),
                           "trace_bytes_resident": int(sv.ntr * plan.row *
                                                       (2 if store in ("BF16", "F16") else 4))}
            if label == "bf16_fast":
                rf = valu_roofline(KMIG_COUNTS["fast_kahan"], probe, n / len(ms),
                                   float(np.mean(ms)), ncu, KMIG_COUNTS["fast_kahan"]["clock_ghz"])
                legs[label]["valu_roofline_frac"] = rf["frac"]
        blk, geom0 = plan.prep(traces[:args.block], coords[:args.block], PREP_CLOSED, TRACE_F32)
        legs["index_mismatch"] = index_mismatch(sv, plan, geom0, dev)
        del blk, geom0
        if not args.no_cpu and args.err_traces > 0:
            for k, v in precision_errors(sv, plan, traces, coords, args.err_traces, dev).items():
                if isinstance(v, dict):
                    legs.setdefault(k, {}).update(v)
                else:
                    legs[k] = v
        result["mixed_precision"] = legs
    if not args.profile and not args.no_extra:
        result["ingest"] = ingest_leg(sv, plan, traces, args.block, dev)
        result["host_path"] = host_leg(sv, traces)
        result["job"] = job_leg(sv, traces, dev)
        result["velocity_gradient"] = vgrad_leg(dev)
        result["adjoint_lsm"] = adjoint_leg(dev)
        result["poststack"] = poststack_leg(dev, plan)
        result["eom"] = eom_leg(dev)
    if not args.no_cpu and not args.profile:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
        nthreads = max(1, min(nthreads, 64))
        result["cpu_baseline"] = cpu_baseline(sv, traces, nthreads, stride=args.cpu_stride)
        result["cpu_baseline"]["speedup_gpu_over_cpu"] = round(gcs / result["cpu_baseline"]["value"], 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
