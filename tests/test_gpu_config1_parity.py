"""GPU-vs-oracle parity at the 3-D workload's real shape: the first 10,000
traces of the 1e6-trace survey (Ricker diffractions + noise, generated on
the GPU) migrated into the whole 400 x 400 x 1500 image at 25 m (aperture
2500 m, trangeMax 8) with the FP32 reference path, and the 50 x 50 x 1500
sub-volume at columns 175-224 of both axes compared bit for bit with the
oracle's FP32 mode (Kirchhoff.c:1104-1404) on the same columns.

The image is split into 16 x 16 job slabs of 25 x 25 columns so that the
sub-volume is exactly the slabs (7, 7), (7, 8), (8, 7) and (8, 8): the oracle
migrates those four jobs (Kirchhoff.c:1341-1347) and the GPU migrates (a)
every slab at once into the whole image and (b) the four jobs alone, whose
contribution counter gives the sub-volume's count."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import oracle
from random_amd import surveys, synth
from random_amd.ktm import Plan

pytestmark = pytest.mark.gpu

NTR = 10_000
X0, NSUB = 175, 50


@pytest.fixture(scope="module")
def subvolume(cuda):
    sv = surveys.area3d()                     # the full 1e6-trace, 400 x 400 survey
    sv.coords = np.ascontiguousarray(sv.coords[:NTR])
    cfg = sv.cfg.replace(jobs_x=16, jobs_y=16)
    plan = Plan(cfg, sv.ns, sv.dt, sv.vel)
    assert plan.setup["nx"] == 25 and plan.setup["ny"] == 25
    assert plan.image_shape == (400, 400, 1, sv.ns)
    traces, coords = synth.traces_gpu(plan, sv, device=cuda)
    print(f"case: 3-D workload, first {NTR} traces into the whole 400 x 400 x {sv.ns} image",
          flush=True)
    km, fold, n_all = plan.migrate(traces, coords, block=5_000)
    jobs = [(X0 // 25 + a) * 16 + (X0 // 25 + b) for a in (0, 1) for b in (0, 1)]
    kmj, foldj, _ = plan.new_image(device=cuda)
    counter = torch.zeros(1, dtype=torch.int64, device=cuda)
    prepared, geom = plan.prep(traces, coords)
    for job in jobs:
        plan.migrate_block(prepared, geom, kmj, foldj, job=job, counter=counter)
    torch.cuda.synchronize()
    tr = traces.cpu().numpy()

    def job_oracle(job):
        return job, oracle.migrate_job(cfg, sv.ns, sv.dt, job, tr, sv.coords, sv.vel)

    with ThreadPoolExecutor(max_workers=4) as ex:
        res = dict(ex.map(job_oracle, jobs))
    km_o = np.zeros((NSUB, NSUB, 1, sv.ns), np.float32)
    fold_o = np.zeros_like(km_o)
    n_o = 0
    for job, (k, f, n) in res.items():
        xs, ys = job // 16 - X0 // 25, job % 16 - X0 // 25
        km_o[25 * xs:25 * xs + 25, 25 * ys:25 * ys + 25] = k
        fold_o[25 * xs:25 * xs + 25, 25 * ys:25 * ys + 25] = f
        n_o += n
    sl = (slice(X0, X0 + NSUB), slice(X0, X0 + NSUB))
    return {"km": km[sl].cpu().numpy(), "fold": fold[sl].cpu().numpy(),
            "kmj": kmj[sl].cpu().numpy(), "foldj": foldj[sl].cpu().numpy(),
            "n_jobs": int(counter.item()), "n_all": n_all, "km_o": km_o, "fold_o": fold_o,
            "n_o": n_o}


def test_subvolume_bit_identical_to_the_oracle(subvolume):
    s = subvolume
    assert s["n_o"] > 1e9 and np.abs(s["km_o"]).max() > 0
    same = s["km"].view(np.uint32) == s["km_o"].view(np.uint32)
    print(f"bit-identical fraction {same.mean():.6f} of {same.size} samples", flush=True)
    assert same.mean() == 1.0
    assert np.array_equal(s["fold"], s["fold_o"])


def test_subvolume_contribution_count(subvolume):
    s = subvolume
    assert s["n_jobs"] == s["n_o"], (s["n_jobs"], s["n_o"])
    assert s["n_all"] > s["n_jobs"]


def test_four_jobs_alone_equal_the_whole_image_on_them(subvolume):
    s = subvolume
    assert np.array_equal(s["kmj"].view(np.uint32), s["km"].view(np.uint32))
    assert np.array_equal(s["foldj"], s["fold"])
