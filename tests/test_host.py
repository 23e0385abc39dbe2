"""Host-side checks that need no GPU: the C-ABI library loads and exports every declared symbol,
the entry_AUMISART parameter parsing, geometry conversion and the seeded synthetic data."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from mcsart_amd import _lib, configs
from mcsart_amd.driver import MexBadInput, parse_geo, parse_iter
from mcsart_amd.synth import (SplitMix64, breathing_trace, ellipsoid_phantom, motion_fields,
                              poisson_noise, seed_for, Voxel)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mcsart_hip.h")).read()
    return sorted(set(re.findall(r"\b(mcsart_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes signature table covers exactly the header
    assert sorted(_lib.SIGNATURES) == declared


def test_status_strings_without_device():
    lib = _lib.load()
    assert lib.mcsart_status_string(0) == b"ok"
    assert lib.mcsart_status_string(-1) == b"invalid argument"
    assert lib.mcsart_version() >= 10000


def test_job_bytes_accounting():
    from mcsart_amd.driver import job_bytes
    g = configs.c2_mcsart128().geometry
    b = job_bytes(g, 200, 20)
    N, D = g.nvox, g.ndet
    # projections + image + 6 fields + 7 pristine + 5 work volumes (pairs in two layouts, u) +
    # the image sweep's warped image in two layouts + 2 motion accumulators + 4 motion constants
    # (Ga, Gb, H and x) + (r, n) pairs + norms
    floor = 4 * (200 * D + N + 6 * N + 7 * N + 5 * N + 2 * N + 2 * N + 4 * N + 2 * D + 20 * 200)
    assert floor <= b <= floor + (1 << 20)


def test_geometry_from_mm_matches_reference_units():
    g = configs.c3_clinical().geometry
    assert g.SO == np.float32(1000 / 1.5) and g.SD == np.float32(1000.0)
    assert g.da == np.float32(0.78 / 1.5) and g.db == np.float32(0.78 / 2.0)
    assert g.ai == -(512 / 2 - 0.5) and g.bi == -(384 / 2 - 0.5)


def test_parse_geo_aliases_and_defaults():
    # src/entry_AUMISART.cu:34-108: nv/nu aliases, da/db default 1, ai/bi shifted, SI and DI
    g = parse_geo(dict(nx=8, ny=8, nz=4, nv=10, nu=6, SI=50.0, DI=25.0, ai=1.0))
    assert (g.na, g.nb, g.da, g.db) == (10, 6, 1.0, 1.0)
    assert g.SO == 50.0 and g.SD == 75.0
    assert g.ai == np.float32(1.0) - np.float32(4.5) and g.bi == -2.5
    with pytest.raises(MexBadInput):
        parse_geo(dict(nx=8, ny=8, nz=4, na=10, nb=6))        # no SO
    with pytest.raises(MexBadInput):
        parse_geo(dict(ny=8, nz=4, na=10, nb=6, SO=1, SD=2))  # no nx


def test_parse_iter_required_fields():
    base = dict(n_iter=2, op_iter=[1, 0], n_views=3, outIter=1, volumes=[0, 0, 0],
                flows=[0, 0, 0], err_weights=[0, 0, 0], angles=[0, 1, 2])
    base["lambda"] = 0.5
    p = parse_iter(base)
    assert p["op_iter"].dtype == np.int32 and p["lam"] == 0.5
    for k in ("op_iter", "n_views", "lambda", "outIter", "volumes", "angles"):
        bad = dict(base)
        del bad[k]
        with pytest.raises(MexBadInput):
            parse_iter(bad)


def test_splitmix_reproducible_and_uniform():
    a = SplitMix64(123).uniform(10000)
    b = SplitMix64(123).uniform(10000)
    assert np.array_equal(a, b)
    assert 0.49 < a.mean() < 0.51 and a.min() >= 0 and a.max() < 1
    z = SplitMix64(5).normal(20000)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    # known first value of the stream (pins the mixing constants)
    assert int(SplitMix64(0).next_u64(1)[0]) == 0xE220A8397B1DCDAF


def test_synthetic_scan_properties():
    spec = configs.c2_mcsart128()
    vox = spec.voxel
    ph = ellipsoid_phantom(32, 32, 32, Voxel(8.0, 8.0), seed_for(2))
    assert ph.dtype == np.float32 and ph.min() >= 0 and 0.01 <= ph.max() <= 0.3
    a, b = breathing_trace(200, 7)
    assert a.shape == (200,) and np.abs(a).max() < 1.5
    assert abs(float(np.corrcoef(b[1:-1], np.gradient(np.sin(2 * np.pi * 0.03 * np.arange(200) / 4.0), 0.03)[1:-1])[0, 1])) > 0.3
    fx, fy, fz = motion_fields(32, 32, 32, Voxel(8.0, 8.0), 9)
    assert np.isclose(fz.max(), 12.0 / 8.0) and np.isclose(fx.max(), 4.0 / 8.0)
    assert np.all(fy == 0)
    p = poisson_noise(np.full((4, 5), 2.0, np.float32), 3, photons=1e5, dx_mm=1.0)
    assert np.abs(p - 2.0).max() < 0.1


def test_batch_mix_is_seeded():
    j1 = configs.c4_batch()
    j2 = configs.c4_batch()
    assert [(j.nx, j.nz, j.na, j.n_views) for j in j1] == [(j.nx, j.nz, j.na, j.n_views) for j in j2]
    assert len(j1) == 16 and all(200 <= j.n_views <= 900 for j in j1)
    assert len({(j.nx, j.nz) for j in j1}) >= 2


def test_ctypes_structs_match_header(tmp_path):
    """The Python mirrors of the header's structs have the C layout (sizes and field offsets
    printed by a program compiled against include/mcsart_hip.h)."""
    import ctypes
    import shutil
    import subprocess
    from mcsart_amd.driver import CGeom, COpts, CTuning
    from mcsart_amd.serve import CRequest, CRequestInfo, CServerOpts
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = [(CGeom, "mcsart_geom"), (COpts, "mcsart_opts"), (CTuning, "mcsart_tuning"),
               (CServerOpts, "mcsart_server_opts"), (CRequest, "mcsart_request"),
               (CRequestInfo, "mcsart_request_info")]
    c_names = {"lam": "lambda"}
    lines = []
    for py, cname in structs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {c_names.get(f, f)}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"mcsart_hip.h\"\n"
                   "int main(void) {\n" + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([cc, "-I", os.path.join(root, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = []
    for py, _ in structs:
        want.append(ctypes.sizeof(py))
        want.extend(getattr(py, f).offset for f, _ in py._fields_)
    assert got == want


def _scan_dicts(V=6, motion=False):
    geo = dict(nx=16, ny=16, nz=8, nv=24, nu=20, da=1.5, db=1.25, SI=90.0, DI=60.0, ai=0.5,
               dx=2.0)
    it = dict(n_iter=3, op_iter=[1, 0, 1], n_views=V, **{"lambda": 0.8}, outIter=1,
              volumes=np.linspace(0, 1, V), flows=np.zeros(V), err_weights=np.ones(V),
              angles=np.linspace(0, 3, V))
    return geo, it


def test_recon_cli_dry_run(tmp_path):
    """The file driver parses the sidecar with the gateway's rules (aliases nv/nu, SI/DI,
    offsets, defaults) and checks the raw file sizes, without touching the GPU."""
    from mcsart_amd import scanio
    from mcsart_amd.driver import parse_geo
    geo, it = _scan_dicts()
    proj = np.zeros((6, 20, 24), np.float32)
    js = scanio.write_scan(str(tmp_path / "s"), geo, it, proj)
    r = scanio.run_cli(js, str(tmp_path / "o"), dry_run=True)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    g = parse_geo(geo)
    assert (got["nx"], got["na"], got["nb"]) == (16, 24, 20)
    assert np.float32(got["ai"]) == np.float32(g.ai) and np.float32(got["bi"]) == np.float32(g.bi)
    assert np.float32(got["SD"]) == np.float32(150.0) and got["dx"] == 2.0
    assert got["op_iter"] == [1, 0, 1] and got["motion"] == 0
    assert got["defaulted"] == 16   # bi only
    # a missing required field: the gateway's message, non-zero exit
    del geo["nz"]
    js = scanio.write_scan(str(tmp_path / "s2"), geo, it, proj)
    r = scanio.run_cli(js, str(tmp_path / "o2"), dry_run=True)
    assert r.returncode != 0 and "nz" in r.stderr
    # a projection file of the wrong size
    geo, it = _scan_dicts()
    js = scanio.write_scan(str(tmp_path / "s3"), geo, it, proj[:5])
    r = scanio.run_cli(js, str(tmp_path / "o3"), dry_run=True)
    assert r.returncode != 0 and "bytes" in r.stderr


def test_gateway_parser_rules():
    """mcsart_parse_geo / mcsart_parse_iter: per-gateway fields, the projection gateway's
    'bi > -1' rule, default bits, and the reference's check order for ITER_PARA."""
    from mcsart_amd.driver import MexBadInput, parse_geo_params, parse_iter
    base = dict(nx=8, ny=8, nz=4, na=10, nb=12, SO=100.0, SD=150.0)
    p = parse_geo_params(base)
    assert p.defaulted == 1 | 2 | 4 | 8 | 16 and p.dx == 1.0
    assert p.ai == np.float32(-(10 / 2 - 0.5)) and p.bi == np.float32(-(12 / 2 - 0.5))
    for gw, bi_user, want in (("projection", -2.0, -2.0), ("projection", 0.0, -5.5),
                              ("backprojection", -2.0, -7.5), ("aumisart", -2.0, -7.5)):
        q = parse_geo_params(dict(base, bi=bi_user, angle=0.5), gw)
        assert q.bi == np.float32(want), (gw, bi_user)
    with pytest.raises(MexBadInput, match="angle"):
        parse_geo_params(base, "backprojection")
    assert parse_geo_params(dict(nx=3, ny=4, nz=5, ind=2), "diff").ind == 2
    it = dict(op_iter=[1], n_views=2, outIter=1, volumes=[0, 0], flows=[0, 0],
              err_weights=[1, 1], angles=[0, 1])
    with pytest.raises(MexBadInput, match="lambda"):
        parse_iter(it)
    it["lambda"] = 1.0
    p = parse_iter(it)
    assert p["n_iter"] == 1 and not p["motion"]
    del it["op_iter"]
    del it["n_views"]   # op_iter is checked first (src/entry_AUMISART.cu:119-128)
    with pytest.raises(MexBadInput, match="op_iter"):
        parse_iter(it)


def test_poisson_noise_blocks_draw_the_same_numbers():
    """Blocked Poisson noise (progress on large scans) equals the one-draw form: the normal
    sequence is rebuilt from the counter-based stream, so the synthetic data do not depend on
    the block size."""
    from mcsart_amd.synth import SplitMix64, poisson_noise
    rng = SplitMix64(123)
    for n in (1, 2, 7, 10, 11):
        ref = SplitMix64(99).normal(n)
        r = SplitMix64(99)
        for b in (1, 2, 3, 5, n):
            got = np.concatenate([r.normal_range(n, s, min(n, s + b)) for s in range(0, n, b)])
            assert np.array_equal(got, ref), (n, b)
    proj = rng.uniform(3 * 5 * 7).reshape(3, 5, 7) * 8.0   # includes lam < 50 rays
    a = poisson_noise(proj, 5, photons=1e4, dx_mm=1.0, block=1 << 24)
    for b in (1, 4, 13, 105):
        assert np.array_equal(poisson_noise(proj, 5, photons=1e4, dx_mm=1.0, block=b), a), b


def test_poisson_inversion_matches_scalar_loop():
    """The vectorised Poisson inversion for low-count rays performs, ray by ray, the scalar
    loop's float64 operations (term *= lam / n; cdf += term), so its draws are identical."""
    import math
    from mcsart_amd.synth import SplitMix64, poisson_inversion
    rng = SplitMix64(7)
    lam = np.concatenate([rng.uniform(200, 0.0, 50.0), [0.0, 1e-9, 49.999]])
    u = rng.uniform(lam.size)
    ref = np.empty(lam.size)
    for j, lk in enumerate(lam):
        term = math.exp(-lk)
        cdf, n = term, 0
        while u[j] > cdf and n < 1000:
            n += 1
            term *= lk / n
            cdf += term
        ref[j] = n
    assert np.array_equal(poisson_inversion(lam, u), ref)
