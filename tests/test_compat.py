"""The reference-side binding compiled and run: tests/compat_check.cpp includes
include/mcsart_compat.h, calls every host_* shim and both host_AUMISART overloads with the
reference's argument types (include/cu_*.h, include/AUMISART.h:17,19) and links libmcsart_hip.
The build compiles it to tests/compat_check.bin; here it is compiled again with -Werror (CPU) and
its outputs are checked against the FP64 oracle (GPU).
"""
import os
import subprocess

import numpy as np
import pytest

from _util import random_fields, rel_l2, smooth_random

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "compat_check.bin")


def test_compat_header_compiles_and_links(tmp_path):
    from mcsart_amd import build
    r = subprocess.run(build.compat_check_command(str(tmp_path / "compat_check")),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    # every host_* shim resolved against the library's exported C ABI
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(tmp_path / "compat_check")],
                        capture_output=True, text=True)
    for sym in ("mcsart_projection", "mcsart_backprojection", "mcsart_deform", "mcsart_deform2",
                "mcsart_update_udvf", "mcsart_add", "mcsart_add2", "mcsart_initial",
                "mcsart_initial2", "mcsart_division", "mcsart_diff", "mcsart_multiple",
                "mcsart_aumisart", "mcsart_aumisart_motion"):
        assert f" {sym}\n" in nm.stdout, sym


@pytest.mark.gpu
def test_compat_binding_matches_oracle(tmp_path, gpu_lib):
    import oracle
    from mcsart_amd.geometry import Geometry, uniform_angles
    assert os.path.exists(BIN), "tests/compat_check.bin is built by the library build"
    g = Geometry.from_user(32, 32, 24, 48, 40, da=1.0, db=1.0, SO=100.0, SD=160.0)
    V, op = 8, [1, 0, 1]
    ang = uniform_angles(V)
    vol = np.linspace(-0.4, 0.5, V).astype(np.float32)
    flo = np.linspace(0.2, -0.1, V).astype(np.float32)
    x = smooth_random(g.vol_shape, 81, scale=0.05)
    fields = [0.5 * f for f in random_fields(g.vol_shape, 82)]
    proj = np.stack([oracle.projection(oracle.deform(x, vol[i], flo[i], fields), g, ang[i])
                     for i in range(V)]).astype(np.float32)
    with open(tmp_path / "params.txt", "w") as f:
        f.write(f"{g.nx} {g.ny} {g.nz} {g.na} {g.nb} {V} {len(op)} {g.da!r} {g.db!r} {g.ai!r} "
                f"{g.bi!r} {g.SO!r} {g.SD!r} 1.0 " + " ".join(map(str, op)) + "\n")
    for name, a in (("img", x), ("proj", proj), ("angles", ang), ("volumes", vol),
                    ("flows", flo), ("fields", np.stack(fields))):
        np.ascontiguousarray(a, np.float32).tofile(tmp_path / f"{name}.bin")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr

    def out(name, shape):
        return np.fromfile(tmp_path / f"out_{name}.bin", np.float32).reshape(shape)

    vs, ds = g.vol_shape, g.det_shape
    assert rel_l2(out("projection", ds), oracle.projection(x, g, ang[0])) <= 1e-5
    assert rel_l2(out("backprojection", vs), oracle.backprojection(proj[0], g, ang[0])) <= 1e-5
    assert rel_l2(out("deform", vs), oracle.deform(x, vol[1] - vol[0], flo[1] - flo[0],
                                                   fields)) <= 1e-5
    assert rel_l2(out("deform2", vs), oracle.deform2(x, 0.5, 0.25, fields)) <= 1e-5
    i2 = oracle.initial2(x, -0.3, 0.1)
    assert rel_l2(out("initial2", vs), i2) <= 1e-5
    w = oracle.deform(x, vol[1] - vol[0], flo[1] - flo[0], fields)
    a2 = oracle.add2(w, i2, x, 0.7, 2)
    assert rel_l2(out("add2", vs), a2) <= 1e-5
    # division maps the 0/0 and x/0 of the last planes to 0: compare where the divisor is not tiny
    dv = out("division", vs)
    m = np.abs(i2) > 1e-6
    np.testing.assert_allclose(dv[m], (a2 / np.where(m, i2, 1))[m], rtol=1e-4, atol=1e-6)
    assert np.all(dv[~m & (i2 == 0)] == 0)
    dz = oracle.diff_backward(x, 3)
    np.testing.assert_allclose(out("multiple", vs), oracle.multiple(dz, x, 0.5, 1), rtol=1e-5,
                               atol=1e-9)
    upd = out("update_udvf", (6,) + vs)
    ref_upd = oracle.update_udvf(fields, x, out("division", vs), 0.4, -0.2)
    for c in range(6):
        assert rel_l2(upd[c], ref_upd[c]) <= 1e-5, c
    # the drivers: overload 1 (static) and overload 2 (motion), schedule (image, motion, image)
    z = np.zeros(vs)
    r_static = oracle.aumisart(z, proj, g, ang, vol, flo, op)
    assert rel_l2(out("static_img", vs), r_static[0]) <= 1e-3
    np.testing.assert_allclose(out("static_norm", (len(op) * V,)), r_static[1], rtol=1e-3,
                               atol=1e-6)
    r64 = oracle.aumisart(z, proj, g, ang, vol, flo, op, fields=fields)
    r32 = oracle.aumisart(z, proj, g, ang, vol, flo, op, fields=fields, dtype=np.float32)
    # the first image sweep is well conditioned; after the motion sweep the reference's own FP32
    # envelope applies
    n = out("motion_norm", (len(op) * V,))
    np.testing.assert_allclose(n[:V], r64[1][:V], rtol=1e-3, atol=1e-6)
    assert rel_l2(out("motion_img", vs), r64[0]) <= max(3 * rel_l2(r32[0], r64[0]), 1e-3)
    assert rel_l2(out("motion_ax", vs), r64[2]) <= max(3 * rel_l2(r32[2], r64[2]), 1e-3)
