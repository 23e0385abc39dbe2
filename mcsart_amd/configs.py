"""The scan workloads the parity tests and the benchmark run.

Sizes follow the published workload definitions of the McSART benchmark suite: a 64^3 static
case, a 128^3 motion-compensated case, a 256x256x128 clinical case, a 16-scan uneven batch and a
served Poisson stream.  Geometry: circular orbit, SAD 1000 mm, SDD 1500 mm, centred detector,
phi_i = 2 pi i / V.  Units are converted to the reference's voxel units (SO = SAD/dx,
da = pitch/dx, db = pitch/dz).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .geometry import Geometry, uniform_angles
from .synth import (SplitMix64, Voxel, breathing_trace, ellipsoid_phantom, motion_fields,
                    seed_for)


@dataclass
class ScanSpec:
    name: str
    nx: int
    ny: int
    nz: int
    vx_mm: float
    vz_mm: float
    na: int
    nb: int
    pitch_mm: float
    n_views: int
    motion: bool
    op_iter: List[int]
    lam: float = 1.0
    nonneg: bool = True
    period_s: float = 4.0
    noise_photons: Optional[float] = None
    seed: int = 0

    @property
    def geometry(self) -> Geometry:
        return Geometry.from_mm(self.nx, self.ny, self.nz, self.vx_mm, self.vz_mm, self.na,
                                self.nb, self.pitch_mm, self.pitch_mm)

    @property
    def voxel(self) -> Voxel:
        return Voxel(self.vx_mm, self.vz_mm)

    @property
    def angles(self) -> np.ndarray:
        return uniform_angles(self.n_views)


def c1_static64() -> ScanSpec:
    """64^3 at 4 mm, detector 96x96 at 4 mm, 60 views, static SART, 5 iterations, nonneg."""
    return ScanSpec("static64", 64, 64, 64, 4.0, 4.0, 96, 96, 4.0, 60, False, [1] * 5,
                    seed=seed_for(1))


def c2_mcsart128() -> ScanSpec:
    """128^3 at 2 mm, detector 256x192 at 1.5 mm, 200 views (30 ms apart), 10 outer iterations
    of (image sweep, motion sweep), Poisson noise at 1e5 photons."""
    return ScanSpec("mcsart128", 128, 128, 128, 2.0, 2.0, 256, 192, 1.5, 200, True, [1, 0] * 10,
                    noise_photons=1e5, seed=seed_for(2))


def c3_clinical() -> ScanSpec:
    """256x256x128 at 1.5x1.5x2 mm, detector 512x384 at 0.78 mm, 656 views, 20 outer iterations."""
    return ScanSpec("clinical256", 256, 256, 128, 1.5, 2.0, 512, 384, 0.78, 656, True,
                    [1, 0] * 20, noise_photons=1e5, seed=seed_for(3))


_C4_VOLS = [(128, 128, 128, 2.0, 2.0), (192, 192, 128, 1.5, 2.0), (256, 256, 128, 1.5, 2.0),
            (320, 320, 160, 1.2, 1.6)]
_C4_DETS = [(256, 192), (384, 288), (512, 384)]


def c4_batch(n_jobs: int = 16, n_outer: int = 10) -> List[ScanSpec]:
    """16 uneven scans: volume from the four shapes, V ~ U{200..900}, detector from three
    shapes, breathing period ~ U(3, 6) s; pitch chosen so the transaxial detector FOV at the
    isocentre equals the volume FOV."""
    jobs = []
    for j in range(n_jobs):
        seed = seed_for(4, j)
        rng = SplitMix64(seed ^ 0x5A5A)
        vi = int(rng.integers(1, 0, len(_C4_VOLS) - 1)[0])
        di = int(rng.integers(1, 0, len(_C4_DETS) - 1)[0])
        V = int(rng.integers(1, 200, 900)[0])
        period = float(rng.uniform(1, 3.0, 6.0)[0])
        nx, ny, nz, vx, vz = _C4_VOLS[vi]
        na, nb = _C4_DETS[di]
        pitch = nx * vx * 1.5 / na   # magnification 1.5: FOV at iso = volume FOV
        jobs.append(ScanSpec(f"batch{j}", nx, ny, nz, vx, vz, na, nb, pitch, V, True,
                             [1, 0] * n_outer, period_s=period, noise_photons=1e5, seed=seed))
    return jobs


def c5_large() -> ScanSpec:
    """512x512x256 at 1 mm, detector 1024x768 at 0.39 mm, 900 views."""
    return ScanSpec("large512", 512, 512, 256, 1.0, 1.0, 1024, 768, 0.39, 900, True, [1, 0] * 10,
                    noise_photons=1e5, seed=seed_for(5))


def scaled(spec: ScanSpec, n_views: Optional[int] = None, op_iter=None) -> ScanSpec:
    """Same geometry with fewer views / sweeps (for bounded CPU work)."""
    d = dict(spec.__dict__)
    if n_views is not None:
        d["n_views"] = n_views
    if op_iter is not None:
        d["op_iter"] = list(op_iter)
    return ScanSpec(**d)


@dataclass
class ScanData:
    spec: ScanSpec
    phantom: np.ndarray
    alpha: np.ndarray          # breathing trace (reference `volumes`)
    beta: np.ndarray           # its derivative (reference `flows`)
    fields: Optional[list]     # [ax, ay, az, bx, by, bz] or None


def make_scan_inputs(spec: ScanSpec) -> ScanData:
    """Phantom, trace and v/f fields of a scan (no projections)."""
    g = spec.geometry
    vox = spec.voxel
    phantom = ellipsoid_phantom(g.nx, g.ny, g.nz, vox, spec.seed)
    if spec.motion:
        alpha, beta = breathing_trace(spec.n_views, spec.seed + 101, period_s=spec.period_s)
        vx, vy, vz = motion_fields(g.nx, g.ny, g.nz, vox, spec.seed + 202)
        fx, fy, fz = motion_fields(g.nx, g.ny, g.nz, vox, spec.seed + 303, scale=0.3)
        fields = [vx, vy, vz, fx, fy, fz]
    else:
        alpha = np.zeros(spec.n_views, np.float32)
        beta = np.zeros(spec.n_views, np.float32)
        fields = None
    return ScanData(spec, phantom, alpha, beta, fields)
