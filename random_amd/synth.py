"""GPU generation of the seeded synthetic surveys of ``random_amd.surveys``."""
from __future__ import annotations

import numpy as np

from ._lib import check, lib
from .ktm import Plan, _ptr, _stream
from .surveys import RICKER_F0, Survey


def traces_gpu(plan: Plan, sv: Survey, device="cuda", noise: bool = True, coords=None):
    """(traces [ntr][ns] float32, coords [ntr][4] int32) on the device.

    Signal: ``ktm_synth_traces``; noise: ``ktm_synth_noise`` with
    sigma^2 = mean(signal^2) / 10^(snr_db/10) over the whole survey."""
    import torch
    if coords is None:
        coords = torch.from_numpy(sv.coords).to(device)
    ntr = coords.shape[0]
    geom = plan.geometry(coords)
    scat = torch.from_numpy(np.ascontiguousarray(sv.scatterers, np.float32)).to(device)
    traces = torch.empty((ntr, sv.ns), dtype=torch.float32, device=device)
    check(lib().ktm_synth_traces(plan._h, _ptr(geom), ntr, _ptr(scat), scat.shape[0], RICKER_F0,
                                 sv.v0, sv.vt, sv.vx, _ptr(traces), _stream()), "ktm_synth_traces")
    if noise:
        ss = 0.0
        for i in range(0, ntr, 65536):
            ss += float((traces[i:i + 65536].double() ** 2).sum().item())
        p = ss / (ntr * sv.ns)
        sigma = float(np.sqrt(p / 10.0 ** (sv.snr_db / 10.0)))
        check(lib().ktm_synth_noise(_ptr(traces), traces.numel(), sigma, sv.seed ^ 0x5EED,
                                    _stream()), "ktm_synth_noise")
    return traces, coords
