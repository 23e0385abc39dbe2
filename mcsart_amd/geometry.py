"""Scan geometry in the reference's units and parameter conventions.

The reference works in voxel units: SO, SD, da and db are pre-divided by the voxel size, and the
detector offsets it passes to its operators are *internal* offsets, ``user_offset - (n/2 - 0.5)``
(``src/entry_AUMISART.cu:76-94``).  All float parameters are rounded to float32 because the
reference reads them as ``float``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, asdict

import numpy as np


def _f32(x) -> float:
    return float(np.float32(x))


@dataclass(frozen=True)
class Geometry:
    nx: int
    ny: int
    nz: int
    na: int
    nb: int
    da: float
    db: float
    ai: float   # internal offset (already shifted by -(na/2 - 0.5))
    bi: float   # internal offset (already shifted by -(nb/2 - 0.5))
    SO: float
    SD: float

    @staticmethod
    def from_user(nx, ny, nz, na, nb, da=1.0, db=1.0, ai=None, bi=None, SO=None, SD=None):
        """Build from GEO_PARA-style values: ``ai``/``bi`` are the user offsets (default 0),
        converted like ``src/entry_AUMISART.cu:76-94`` in float32 arithmetic."""
        f = np.float32
        ai_i = (f(0.0) if ai is None else f(ai)) - (f(na) / f(2) - f(0.5))
        bi_i = (f(0.0) if bi is None else f(bi)) - (f(nb) / f(2) - f(0.5))
        return Geometry(int(nx), int(ny), int(nz), int(na), int(nb), _f32(da), _f32(db),
                        float(ai_i), float(bi_i), _f32(SO), _f32(SD))

    @staticmethod
    def from_mm(nx, ny, nz, vx_mm, vz_mm, na, nb, pitch_a_mm, pitch_b_mm, sad_mm=1000.0,
                sdd_mm=1500.0):
        """Physical scanner description -> reference units (SO=SAD/dx, da=pitch/dx, db=pitch/dz).
        Anisotropic z is expressed exactly by passing db in z-voxel units."""
        return Geometry.from_user(nx, ny, nz, na, nb, da=pitch_a_mm / vx_mm,
                                  db=pitch_b_mm / vz_mm, SO=sad_mm / vx_mm, SD=sdd_mm / vx_mm)

    @property
    def nvox(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def ndet(self) -> int:
        return self.na * self.nb

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def det_shape(self):
        return (self.nb, self.na)

    def validate(self) -> None:
        if min(self.nx, self.ny, self.nz, self.na, self.nb) <= 0:
            raise ValueError("all sizes must be positive")
        if self.nx != self.ny:
            raise ValueError("nx == ny is required (src/cu_deform.cu:85, src/cu_add.cu:58)")
        if not (self.SO > 0 and self.SD > self.SO):
            raise ValueError("need 0 < SO < SD")
        if not (self.da > 0 and self.db > 0):
            raise ValueError("detector pitch must be positive")

    def asdict(self):
        return asdict(self)


def uniform_angles(n_views: int, span: float = 2 * math.pi) -> np.ndarray:
    """phi_i = span*i/V in float32 radians."""
    return (np.arange(n_views, dtype=np.float64) * (span / n_views)).astype(np.float32)
