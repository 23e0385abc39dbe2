"""The allocation-free projection (mcsart_projection_noalloc) against mcsart_projection.

Both run the same projector; the no-alloc variant writes the transposed copy its x-branch rays
read into the library's fixed device array instead of an allocated scratch.  Bar: rel-L2 <= 1e-5
against mcsart_projection (the per-projection tolerance against the FP64 oracle), on two calls
in a row, the second with a smaller volume and detector (the array keeps the first call's
contents), and each against the FP64 oracle.
"""
import numpy as np
import pytest

import oracle
from mcsart_amd._lib import check, load
from mcsart_amd.geometry import Geometry
from mcsart_amd.ops import DeviceArray, host_projection
from _util import rel_l2, smooth_random

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _noalloc(g, d_out, d_img, angle):
    lib = load()
    check(lib.mcsart_projection_noalloc(None, d_out.ptr, d_img.ptr, float(angle), g.SO, g.SD,
                                        g.da, g.na, g.ai, g.db, g.nb, g.bi, g.nx, g.ny, g.nz),
          "mcsart_projection_noalloc")
    check(lib.mcsart_device_synchronize(), "mcsart_device_synchronize")


def test_noalloc_projection_matches_allocating_entry_point(gpu_lib):
    cases = [
        # larger first, then smaller: the second call finds the first call's data in the array
        (Geometry.from_user(96, 96, 64, 160, 96, da=1.0, db=1.0, SO=300.0, SD=450.0),
         [0.0, 0.3, 0.7853982, 2.2]),
        (Geometry.from_user(40, 40, 24, 64, 40, da=1.0, db=1.0, SO=120.0, SD=200.0),
         [0.7853982, 1.1, 0.0]),
    ]
    for k, (g, angles) in enumerate(cases):
        img = smooth_random(g.vol_shape, 91 + k)
        d_img = DeviceArray.from_host(img)
        d_a = DeviceArray.zeros(g.det_shape)
        d_b = DeviceArray.zeros(g.det_shape)
        for ang in angles:
            print(f"noalloc case {k} ({g.nx}x{g.ny}x{g.nz}, det {g.na}x{g.nb}) angle {ang}",
                  flush=True)
            _noalloc(g, d_a, d_img, ang)
            host_projection(d_b, d_img, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi,
                            g.nx, g.ny, g.nz)
            a, b = d_a.to_host(), d_b.to_host()
            ref = oracle.projection(img, g, np.float32(ang))
            assert rel_l2(a, b) <= TOL, (k, ang)
            assert rel_l2(a, ref) <= TOL and rel_l2(b, ref) <= TOL, (k, ang)
        for d in (d_img, d_a, d_b):
            d.close()


def test_noalloc_projection_rejects_oversize(gpu_lib):
    lib = load()
    nmax = int(lib.mcsart_projection_noalloc_max_voxels())
    assert nmax == 1 << 23
    # 256 x 256 x 129 voxels: one slice over the limit; nothing is launched, so the pointers are
    # never dereferenced
    dummy = DeviceArray.zeros((4,))
    rc = lib.mcsart_projection_noalloc(None, dummy.ptr, dummy.ptr, 0.0, 700.0, 1000.0, 1.0, 64,
                                       -31.5, 1.0, 48, -23.5, 256, 256, 129)
    assert rc == -1   # MCSART_EINVAL
    assert b"exceeds the scratch" in lib.mcsart_last_error()
    dummy.close()
