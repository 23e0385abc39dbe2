"""Measured projections of a (moving) phantom, made on the GPU with the library's operators.

p_i = P_i(W(phantom; alpha_i v + beta_i f)) for every view i, optionally with Poisson noise:
the forward model the reconstruction inverts (src/AUMISART.cu:193-203 warps the image with the
same DVF form before projecting).  All arithmetic runs in libmcsart_hip; nothing here computes.
"""
from __future__ import annotations

import numpy as np

from .configs import ScanData, ScanSpec
from .ops import DeviceArray, host_deform, host_projection
from .synth import poisson_noise


def measure(spec: ScanSpec, data: ScanData, photons_seed: int, log=None) -> np.ndarray:
    """(V, nb, na) float32 projections of ``data.phantom`` warped per view, plus Poisson noise
    at ``spec.noise_photons`` photons per ray when it is set.  ``log(msg)``, when given, is
    called every 100 views and around the noise step (large scans take minutes)."""
    g = spec.geometry
    d_ph = DeviceArray.from_host(data.phantom)
    d_w = DeviceArray.empty(g.vol_shape)
    d_f = [DeviceArray.from_host(f) for f in data.fields] if data.fields else None
    d_p = DeviceArray.empty(g.det_shape)
    out = np.empty((spec.n_views,) + g.det_shape, np.float32)
    try:
        for i, a in enumerate(spec.angles):
            src = d_ph
            if d_f is not None:
                host_deform(d_w, d_ph, g.nx, g.ny, g.nz, float(data.alpha[i]),
                            float(data.beta[i]), *d_f)
                src = d_w
            host_projection(d_p, src, float(a), g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi,
                            g.nx, g.ny, g.nz)
            out[i] = d_p.to_host()
            if log and (i + 1) % 100 == 0:
                log(f"[measure] {spec.name}: {i + 1}/{spec.n_views} views projected")
    finally:
        for x in [d_ph, d_w, d_p] + (d_f or []):
            x.close()
    if spec.noise_photons:
        if log:
            log(f"[measure] {spec.name}: Poisson noise on {out.size} rays")
        out = poisson_noise(out, photons_seed, photons=spec.noise_photons, dx_mm=spec.vx_mm,
                            log=log)
        if log:
            log(f"[measure] {spec.name}: noise done")
    return out
