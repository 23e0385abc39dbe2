"""Shared helpers for the test-suite."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def smooth_random(shape, seed, scale=1.0):
    """Seeded smooth positive field (sum of a few separable cosines), float32."""
    from mcsart_amd.synth import SplitMix64
    rng = SplitMix64(seed)
    grids = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
    out = np.zeros(shape)
    for _ in range(4):
        k = rng.uniform(len(shape), 0.5, 3.0)
        ph = rng.uniform(len(shape), 0, 2 * np.pi)
        term = np.ones(shape)
        for d in range(len(shape)):
            term = term * np.cos(2 * np.pi * k[d] * grids[d] + ph[d])
        out += term
    return (scale * (out - out.min() + 0.1)).astype(np.float32)


def smooth_random_large(shape, seed, scale=1.0):
    """As smooth_random, built by float32 broadcasting of 1-D factors (no full-size coordinate
    grids), for volumes of 10^7..10^8 voxels."""
    from mcsart_amd.synth import SplitMix64
    rng = SplitMix64(seed)
    axes = [np.linspace(0, 1, n, dtype=np.float32) for n in shape]
    out = np.zeros(shape, np.float32)
    for _ in range(4):
        k = rng.uniform(len(shape), 0.5, 3.0)
        ph = rng.uniform(len(shape), 0, 2 * np.pi)
        f = [np.cos(2 * np.pi * k[d] * axes[d] + ph[d]).astype(np.float32) for d in range(3)]
        out += f[0][:, None, None] * (f[1][:, None] * f[2][None, :])[None, :, :]
    out -= out.min() - np.float32(0.1)
    return (np.float32(scale) * out).astype(np.float32)


def random_fields(shape, seed, amp=2.5):
    """Six seeded smooth coefficient fields (ax, ay, az, bx, by, bz) with values in +-amp."""
    fs = []
    for c in range(6):
        f = smooth_random(shape, seed + 17 * c)
        f = (f - f.mean()) / max(np.abs(f - f.mean()).max(), 1e-6) * amp
        fs.append(f.astype(np.float32))
    return fs
