"""Culling geometry (ConeCull, mcsart_amd/csrc/footprint.hpp), checked on the CPU.

tests/cull_check.cpp replays every cell the three projector kernels read, from the kernels' own
per-ray constants, and checks that each lies in the view's read set (the voxels the warp and the
motion step still write); it also checks that every voxel whose backprojection can be nonzero lies
in the view's u support (where the warp-add still samples), and that every window of the rows
projector covers its footprint (its last weight boundary is the footprint's end).  A set narrowed by one voxel must fail
the same replay, so the replay does see the sets' edges.  The GPU side of the same property (the
sweeps give the same results bit for bit with culling on, off, and with every skipped write made
as a NaN) is in tests/test_gpu_cull.py.
"""
import os
import shutil
import subprocess

import pytest

from mcsart_amd import configs
from mcsart_amd.geometry import Geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cull_check(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if hipcc is None:
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("cull") / "cull_check")
    cmd = [hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I",
           os.path.join(ROOT, "mcsart_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cull_check.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return out


def _run(binary, g, n_uniform, shrink=0.0, extra=()):
    args = [repr(v) for v in (g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD)]
    r = subprocess.run([binary, *args, str(n_uniform), str(shrink), *map(str, extra)],
                       capture_output=True, text=True, timeout=600)
    last = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr
    return r.returncode, last


def _fields(line):
    """The summary line's "name value" pairs."""
    tok = line.split()
    return {tok[i]: float(tok[i + 1]) for i in range(0, len(tok) - 1, 2)}


# exact multiples of 45 degrees (branch ties) and their float32 neighbours
TIES = ("0.7853982", "0.78539824", "2.3561945", "3.1415925", "1e-6", "4.712389")

GEOMETRIES = {
    "64^3": configs.c1_static64().geometry,
    "128^3": configs.c2_mcsart128().geometry,
    "offset truncated detector": Geometry.from_user(48, 48, 40, 160, 50, da=0.4, db=0.6, ai=23.5,
                                                    bi=-7.25, SO=120, SD=200),
    "odd sizes": Geometry.from_user(37, 37, 21, 61, 29, da=1.1, db=0.9, ai=0.3, bi=0.7, SO=90,
                                    SD=150),
    "coarse detector (wide windows)": Geometry.from_user(40, 40, 32, 20, 16, da=4.0, db=4.0,
                                                         SO=100, SD=160),
    "2x2x2": Geometry.from_user(2, 2, 2, 4, 3, da=1.0, db=1.0, SO=30, SD=50),
    "three slices": Geometry.from_user(64, 64, 3, 96, 80, da=1.0, db=1.0, SO=200, SD=300),
    "source close to the volume": Geometry.from_user(64, 64, 32, 200, 150, da=1.5, db=1.5, SO=50,
                                                     SD=120),
    # the orbit just outside the volume's corners (the least SO allows) and a detector just past
    # the source: the fan reaches a third of the volume, planes near a corner lie behind the source
    "source at the corners' circle": Geometry.from_user(64, 64, 32, 64, 80, da=0.2, db=1.0,
                                                        SO=45.3, SD=47.0),
}


@pytest.mark.parametrize("name", sorted(GEOMETRIES))
def test_read_set_and_support_cover_every_read(cull_check, name):
    rc, line = _run(cull_check, GEOMETRIES[name], 16, 0.0, TIES)
    assert rc == 0, f"{name}: {line}"


def test_read_set_at_256x256x128(cull_check):
    g = configs.c3_clinical().geometry
    rc, line = _run(cull_check, g, 8, 0.0, ("0.3", "0.7853982", "2.0"))
    assert rc == 0, line
    assert 0.55 < _fields(line)["kept"] < 0.75, line   # the fan misses a third of this volume
    # the image projector's rays share their z-rows on every view here (share_misses == 0 is
    # part of the exit status)
    assert _fields(line)["shared_views"] == 11, line


def test_replay_catches_rows_too_far_apart_to_share(cull_check, monkeypatch):
    """With odd sizes adjacent detector rows' footprints lie ~0.7 voxel apart: two rays may share
    their z-rows (rows_share holds for R = 2), four may not.  Forced to share at R = 4 as well,
    the replay finds rays whose windows start more than one row apart."""
    g = GEOMETRIES["odd sizes"]
    rc, line = _run(cull_check, g, 4, 0.0)
    assert rc == 0 and _fields(line)["shared_views"] == 4, line
    monkeypatch.setenv("CULL_CHECK_FORCE_SHARE", "1")
    rc, line = _run(cull_check, g, 4, 0.0)
    assert rc == 1 and _fields(line)["share_misses"] > 0, line


@pytest.mark.parametrize("name", ["128^3", "offset truncated detector", "odd sizes",
                                  "coarse detector (wide windows)"])
def test_set_sizes_match_brute_force(cull_check, name):
    """mcsart_view_voxels (per column: z ranges from binary searches) equals the replay's voxel
    by voxel counts of the read set, the read set with the edge shell, the voxels with cells and
    rows, and their union with the next view's read set and shell."""
    import numpy as np
    from mcsart_amd.driver import view_voxels
    g = GEOMETRIES[name]
    args = [repr(v) for v in (g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD)]
    extra = ("0.7853982", "2.3561945", "1e-6")
    r = subprocess.run([cull_check, *args, "8", "0", *extra], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-500:]
    brute = np.array([[int(x) for x in line.split()[2:]] for line in r.stdout.splitlines()
                      if line.startswith("count")])
    ang = np.concatenate([2 * np.pi * np.arange(8) / 8, np.array(extra, float)]).astype(np.float32)
    assert np.array_equal(view_voxels(g, ang)[:, :4], brute)


@pytest.mark.parametrize("name", ["128^3", "offset truncated detector"])
def test_replay_catches_a_set_one_voxel_too_narrow(cull_check, name):
    rc, line = _run(cull_check, GEOMETRIES[name], 16, 1.0)
    assert rc == 1, line
    f = _fields(line)
    assert f["bad_read"] > 0 and f["bad_support"] > 0 and f["window_misses"] == 0, line
