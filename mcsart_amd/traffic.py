"""Algorithmic bytes of the culled kernels, from the per-view set sizes of mcsart_view_voxels.

Culling (mcsart_job_set_culling) makes the per-voxel work of three kernels depend on the view:
with R_v the voxels of view v's read set, S_v the same with the volume's edge shell, H_v the
voxels where its ratio backprojection can be nonzero, U_v = |H_v u S_{v+1}| and L_v the number
of volume layouts its rays read (1 or 2), one launch must move (4-byte floats, N voxels, D
detector cells):

* forward warp of view v : fields (24) and the x sample (4) read, W(x) written per layout (4),
  for the read set only                                      -> (28 + 4 L_v) R_v
* warp-add of view v     : fields read everywhere (the displacement decides), the u sample and
  x read and x written where the sample can be nonzero; H_v stands for that set (it is H_v moved
  by the displacements) -> 24 N + 12 H_v; view 0 of a sweep is uncut (it clamps every voxel)
                                                             -> 36 N
* motion step of view v (the motion sweep's ratio backprojection fused with the update of view v
  and the prep of view v + 1): (r, Pq) read (8 D); accumulators read where u can be nonzero or
  g is needed (8 U_v; none on the sweep's first view), written where u can be nonzero (8 H_v;
  everywhere, 8 N, on the first view, which starts them); constants (16) read and g written per
  layout (4) where view v + 1 reads them, its shell included -> 8 U_v + 8 H_v
  + (16 + 4 L_{v+1}) S_{v+1} + 8 D; the last view prepares nothing -> 16 H_v + 8 D.

Uncut, the same kernels move 28N + 4N per layout, 36N and 32N + 4N per layout + 8D.
"""
from __future__ import annotations

import numpy as np


def _layouts(bits) -> float:
    return float(bin(int(bits) & 3).count("1"))


def warp_bytes(voxels, views=None) -> float:
    """Mean per-launch bytes of the forward warp over `views` (default: all)."""
    V = voxels.shape[0]
    views = range(V) if views is None else views
    return float(np.mean([(28 + 4 * _layouts(voxels[v, 4])) * voxels[v, 0] for v in views]))


def warp_add_bytes(voxels, N, views=None) -> float:
    """Mean per-launch bytes of the warp-add over `views` (default: all; view 0 uncut)."""
    V = voxels.shape[0]
    views = range(V) if views is None else views
    return float(np.mean([36.0 * N if v == 0 else 24.0 * N + 12.0 * voxels[v, 2]
                          for v in views]))


def motion_step_bytes(voxels, N, D, v, first=False) -> float:
    """Bytes of the motion step of view v (1 <= v < V)."""
    V = voxels.shape[0]
    H, U = float(voxels[v, 2]), float(voxels[v, 3])
    if v + 1 >= V:   # last view: no prep (U_v = H_v)
        return (8.0 * N if first else 16.0 * H) + 8.0 * D
    Sn, Ln = float(voxels[v + 1, 1]), _layouts(voxels[v + 1, 4])
    acc = 8.0 * N if first else 8.0 * U + 8.0 * H
    return acc + (16 + 4 * Ln) * Sn + 8.0 * D


def motion_sweep_bytes(voxels, N, D) -> float:
    """Mean per-launch bytes of the motion step over a motion sweep (views 1..V-1: the first
    starts the accumulators, the last prepares nothing)."""
    V = voxels.shape[0]
    return float(np.mean([motion_step_bytes(voxels, N, D, v, first=(v == 1))
                          for v in range(1, V)]))


def motion_middle_bytes(voxels, N, D) -> float:
    """Mean per-launch bytes of the motion step over views 1..V-2 in its middle form (the
    launches mcsart_job_kernel_time times)."""
    V = voxels.shape[0]
    return float(np.mean([motion_step_bytes(voxels, N, D, v) for v in range(1, V - 1)]))
