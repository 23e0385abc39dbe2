"""Seeded synthetic scans: ellipsoid phantoms, breathing traces and 5DCT motion fields.

Random numbers come from a counter-based SplitMix64 stream implemented here with numpy uint64
arithmetic, so data are bit-reproducible from the seed alone and do not depend on numpy's own
generators.  Seeds follow ``base + 1000*config + job`` with base 20261203.

Conventions (the reference's, ``src/cu_deform.cu:74-89``, ``src/AUMISART.cu:160-172``):
  * volumes are float32 arrays shaped ``(nz, ny, nx)`` (MATLAB column-major ``(nx, ny, nz)``);
  * field ``alpha_x`` displaces array dim 2 (index iy) and ``alpha_y`` displaces dim 1 (ix); SI
    motion is ``alpha_z`` (the rotation axis), AP motion is ``alpha_x``, LR (``alpha_y``) is 0;
  * displacements are in voxel units; the breathing trace alpha(t) is the reference's
    ``volumes`` and beta = d(alpha)/dt its ``flows``.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

SEED_BASE = 20261203
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def seed_for(config: int, job: int = 0) -> int:
    return SEED_BASE + 1000 * int(config) + int(job)


class SplitMix64:
    """Counter-based SplitMix64: value k of stream `seed` is mix(seed + (k+1)*golden)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def next_u64(self, n: int) -> np.ndarray:
        k = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            z = self.state + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller on pairs of uniforms."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))   # 1-u1 in (0, 1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def _uniform_at(self, k0: int, n: int) -> np.ndarray:
        """Uniforms of counters k0+1 .. k0+n, without advancing the stream."""
        k = np.arange(k0 + 1, k0 + 1 + n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = self.state + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal_range(self, n: int, s: int, e: int) -> np.ndarray:
        """Entries s .. e-1 of what normal(n) would return now, without advancing the stream
        (normal(n) takes m = ceil(n/2) pairs: entry i < m is r_i cos, entry i >= m is
        r_(i-m) sin)."""
        m = (n + 1) // 2
        out = np.empty(e - s, dtype=np.float64)
        for lo, hi, first in ((s, min(e, m), True), (max(s, m), e, False)):
            if lo >= hi:
                continue
            j0, j1 = (lo, hi) if first else (lo - m, hi - m)
            u1 = self._uniform_at(self.counter + j0, j1 - j0)
            u2 = self._uniform_at(self.counter + m + j0, j1 - j0)
            r = np.sqrt(-2.0 * np.log1p(-u1))
            out[lo - s:hi - s] = r * (np.cos if first else np.sin)(2 * np.pi * u2)
        return out

    def skip_normal(self, n: int) -> None:
        """Advance the stream past what normal(n) consumes."""
        self.counter += 2 * ((n + 1) // 2)

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Uniform integers in [lo, hi]."""
        return (lo + np.floor(self.uniform(n) * (hi - lo + 1))).astype(np.int64)


def _axes_mm(n, d):
    return (np.arange(n, dtype=np.float64) - n / 2 + 0.5) * d


@dataclass
class Voxel:
    dx: float   # in-plane voxel size (mm), dims 1 and 2
    dz: float   # axial voxel size (mm)


def ellipsoid_phantom(nx, ny, nz, vox: Voxel, seed: int, n_ellipsoids: int = 10) -> np.ndarray:
    """Additive axis-aligned ellipsoids: centres U(+-0.35 FOV) per axis, semi-axes
    U(0.05, 0.40) FOV, attenuation U(0.01, 0.03) /mm; voxelised by the voxel centre."""
    rng = SplitMix64(seed)
    fov = np.array([nx * vox.dx, ny * vox.dx, nz * vox.dz])
    xs, ys, zs = _axes_mm(nx, vox.dx), _axes_mm(ny, vox.dx), _axes_mm(nz, vox.dz)
    img = np.zeros((nz, ny, nx), dtype=np.float64)
    for _ in range(n_ellipsoids):
        c = rng.uniform(3, -0.35, 0.35) * fov
        a = rng.uniform(3, 0.05, 0.40) * fov
        mu = rng.uniform(1, 0.01, 0.03)[0]
        qx = ((xs - c[0]) / a[0]) ** 2
        qy = ((ys - c[1]) / a[1]) ** 2
        qz = ((zs - c[2]) / a[2]) ** 2
        # every term is >= 0, so only voxels with each term <= 1 can be inside: evaluate the
        # ellipsoid's bounding box (the same voxels and sums as over the whole grid)
        bx, by, bz = (np.nonzero(q <= 1.0)[0] for q in (qx, qy, qz))
        if bx.size == 0 or by.size == 0 or bz.size == 0:
            continue
        sx = slice(bx[0], bx[-1] + 1)
        sy = slice(by[0], by[-1] + 1)
        sz = slice(bz[0], bz[-1] + 1)
        inside = (qz[sz, None, None] + qy[None, sy, None] + qx[None, None, sx]) <= 1.0
        img[sz, sy, sx] += mu * inside
    return img.astype(np.float32)


def breathing_trace(n_views: int, seed: int, period_s: float = 4.0, dt_s: float = 0.03,
                    amplitude: float = 1.0, noise_frac: float = 0.05):
    """alpha(t) = A sin(2 pi t / T) + N(0, (noise_frac*A)^2), beta = d(alpha)/dt (central
    differences, one-sided at the ends), in 1/s.  Returns float32 (alpha, beta)."""
    rng = SplitMix64(seed)
    t = dt_s * np.arange(n_views, dtype=np.float64)
    alpha = amplitude * np.sin(2 * np.pi * t / period_s) + noise_frac * amplitude * rng.normal(n_views)
    beta = np.gradient(alpha, dt_s) if n_views > 1 else np.zeros_like(alpha)
    return alpha.astype(np.float32), beta.astype(np.float32)


def _blob_field(nx, ny, nz, vox: Voxel, rng: SplitMix64, n_blobs=8):
    xs, ys, zs = _axes_mm(nx, vox.dx), _axes_mm(ny, vox.dx), _axes_mm(nz, vox.dz)
    fov = np.array([nx * vox.dx, ny * vox.dx, nz * vox.dz])
    out = np.zeros((nz, ny, nx), dtype=np.float64)
    for _ in range(n_blobs):
        sigma = rng.uniform(1, 15.0, 30.0)[0]
        c = rng.uniform(3, -0.35, 0.35) * fov
        amp = rng.uniform(1, 0.5, 1.0)[0]
        gx = np.exp(-0.5 * ((xs - c[0]) / sigma) ** 2)
        gy = np.exp(-0.5 * ((ys - c[1]) / sigma) ** 2)
        gz = np.exp(-0.5 * ((zs - c[2]) / sigma) ** 2)
        gzy = amp * gz[:, None] * gy[None, :]   # (amp gz) gy, then times gx per voxel

        def add_slab(z0, gzy=gzy, gx=gx):   # elementwise: slabs in threads, same sums
            out[z0:z0 + 8] += gzy[z0:z0 + 8, :, None] * gx[None, None, :]

        list(_pool().map(add_slab, range(0, nz, 8)))
    return out


_POOL = None


def _pool():
    """Worker threads for elementwise synthesis over volume slabs (numpy releases the GIL)."""
    global _POOL
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        _POOL = ThreadPoolExecutor(max_workers=max(1, min(16, n or 1)))
    return _POOL


def motion_fields(nx, ny, nz, vox: Voxel, seed: int, si_mm=12.0, ap_mm=4.0, scale=1.0):
    """One 5DCT coefficient field (x, y, z components) as float32 arrays in voxel units: each
    component a sum of 8 Gaussian blobs (sigma U(15, 30) mm, centres in the inner 70%, amplitude
    U(0.5, 1)); SI (alpha_z) normalised to si_mm/dz at its maximum, AP (alpha_x) to ap_mm/dx,
    LR (alpha_y) zero; then multiplied by `scale`."""
    rng = SplitMix64(seed)
    fz = _blob_field(nx, ny, nz, vox, rng)
    fxx = _blob_field(nx, ny, nz, vox, rng)
    fz *= (si_mm / vox.dz) / max(fz.max(), 1e-30)
    fxx *= (ap_mm / vox.dx) / max(fxx.max(), 1e-30)
    fx = (scale * fxx).astype(np.float32)
    fy = np.zeros((nz, ny, nx), dtype=np.float32)
    fzz = (scale * fz).astype(np.float32)
    return fx, fy, fzz


def poisson_inversion(lam: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Poisson draws by inversion of the CDF, all rays at once: the smallest n with
    u <= sum_{k<=n} e^-lam lam^k / k! (at most 1000), each term formed as term *= lam / n."""
    term = np.array([math.exp(-x) for x in lam], dtype=np.float64)
    cdf = term.copy()
    n = np.zeros(lam.size, dtype=np.float64)
    active = np.nonzero(u > cdf)[0]
    while active.size:
        n[active] += 1.0
        term[active] *= lam[active] / n[active]
        cdf[active] += term[active]
        active = active[(u[active] > cdf[active]) & (n[active] < 1000.0)]
    return n


def poisson_noise(proj: np.ndarray, seed: int, photons=1e5, dx_mm=1.0, log=None,
                  block=1 << 21) -> np.ndarray:
    """counts ~ Poisson(I0 exp(-p dx)), p~ = -ln(max(counts, 1)/I0)/dx.  The Poisson draw uses the
    normal approximation N(lam, lam) rounded to an integer (lam >= ~600 for the configs' data);
    exact inversion is used where lam < 50.  Rays are drawn in blocks of `block` (the same
    numbers as one draw over all of them); `log(msg)`, when given, reports each block."""
    rng = SplitMix64(seed)
    p = np.asarray(proj, dtype=np.float64)
    lam = photons * np.exp(-p * dx_mm)
    flat = lam.ravel()
    n = flat.size
    counts = np.empty(n, dtype=np.float64)

    def draw(s):   # a block's normals depend on its counter range only: blocks run in threads
        e = min(n, s + block)
        counts[s:e] = np.rint(flat[s:e] + np.sqrt(flat[s:e]) * rng.normal_range(n, s, e))
        return e

    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                         else (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for e in ex.map(draw, range(0, n, block)):
            if log and e < n:
                log(f"[poisson_noise] {e}/{n} rays")
    rng.skip_normal(n)
    small = np.nonzero(flat < 50.0)[0]
    if small.size:
        u = rng.uniform(small.size)
        sb = 1 << 22   # each draw depends only on its own ray: blocks change nothing
        for s in range(0, small.size, sb):
            e = min(small.size, s + sb)
            counts[small[s:e]] = poisson_inversion(flat[small[s:e]], u[s:e])
            if log:
                log(f"[poisson_noise] low-count rays: {e}/{small.size}")
    counts = np.maximum(counts.reshape(p.shape), 1.0)
    return (-np.log(counts / photons) / dx_mm).astype(np.float32)
