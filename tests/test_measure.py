"""Host-side checks of the measurement code: the algorithmic byte counts the roofline figures
divide by, the kernel-name patterns that pick a kernel's counters out of a profile, the launch
counts per scan, and the counter-table parsers (on small synthetic rocprofv3 CSV files)."""
import csv
import importlib.util
import json
import os

import pytest

from mcsart_amd import configs, profile_summary as ps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_profile_byte_counts():
    N, D = 1000, 10
    ab = ps.algorithmic_bytes
    assert ab("k_project_rows<0, 2, 2, 1>", N, D) == 4 * N + 4 * D      # x in, projection out
    assert ab("k_project_rows<1, 2, 3, 4>", N, D) == 4 * N + 12 * D     # x, p in; (r, P1) out
    assert ab("k_project<2, 3, 3>", N, D) == 8 * N + 12 * D             # (g, Q) pairs
    assert ab("k_backproject<0, 4, true, 0>", N, D) == 4 * N + 4 * D    # plain
    assert ab("k_backproject<1, 4, true, 0>", N, D) == 4 * N + 8 * D    # ratio, u stored
    assert ab("k_backproject<1, 5, false, 1>", N, D) == 36 * N + 8 * D  # motion step, one layout
    assert ab("k_backproject<1, 4, true, 2>", N, D) == 12 * N + 8 * D   # static add with xT
    # the warps carry their lane depth as a second template argument
    assert ab("k_warp_tile<true, 32>", N, D) == 32 * N
    assert ab("k_warp_add_tile<false, 32>", N, D) == 36 * N
    assert ab("k_motion_begin", N, D) == 56 * N
    assert ab("k_no_such_kernel", N, D) is None


def test_bench_byte_counts_agree_with_profiles(bench):
    """For one output layout, the bench's per-kind formulas are the profile's per-kernel ones."""
    spec = configs.c2_mcsart128()
    g = spec.geometry
    N, D = g.nvox, g.ndet
    one = [1] * spec.n_views   # every view: natural layout only
    assert bench.algorithmic_bytes(spec, "proj_image") == ps.algorithmic_bytes(
        "k_project_rows<1, 2, 2, 1>", N, D)
    assert bench.algorithmic_bytes(spec, "proj_motion") == ps.algorithmic_bytes(
        "k_project_rows<2, 2, 2, 1>", N, D)
    assert bench.algorithmic_bytes(spec, "back") == ps.algorithmic_bytes(
        "k_backproject<1, 4, true, 0>", N, D)
    assert bench.algorithmic_bytes(spec, "warp", one) == ps.algorithmic_bytes(
        "k_warp_tile<true, 32>", N, D)
    assert bench.algorithmic_bytes(spec, "warp_add") == ps.algorithmic_bytes(
        "k_warp_add_tile<true, 32>", N, D)
    assert bench.algorithmic_bytes(spec, "motion", one) == ps.algorithmic_bytes(
        "k_backproject<1, 4, true, 1>", N, D)
    # both layouts on every view: one more output volume per launch
    both = [3] * spec.n_views
    assert bench.algorithmic_bytes(spec, "warp", both) == 36 * N
    assert bench.algorithmic_bytes(spec, "motion", both) == 40 * N + 8 * D


def test_culled_byte_counts(bench):
    """The culled kernels' bytes (mcsart_amd/traffic.py): uncut sets give the uncut formulas, and
    a small scan with distinct set sizes gives the hand-computed per-view sums."""
    import numpy as np
    from mcsart_amd import traffic
    spec = configs.c2_mcsart128()
    g = spec.geometry
    N, D = g.nvox, g.ndet
    full = np.tile(np.array([N, N, N, N, 1], np.int64), (spec.n_views, 1))
    assert bench.algorithmic_bytes(spec, "warp", None, full) == 32 * N
    assert bench.algorithmic_bytes(spec, "warp_add", None, full) == 36 * N
    assert bench.algorithmic_bytes(spec, "motion", None, full) == 36 * N + 8 * D
    N, D = 1000, 10
    # per view: read set R, with shell S, support H, union U, layouts (1 natural, 2 transposed,
    # 3 both)
    vox = np.array([[600, 700, 500, 650, 1], [620, 720, 510, 640, 3], [630, 730, 520, 660, 2],
                    [640, 740, 530, 530, 1]], np.int64)
    assert traffic.warp_bytes(vox) == (32 * 600 + 36 * 620 + 32 * 630 + 32 * 640) / 4
    assert traffic.warp_add_bytes(vox, N) == (36 * N + 3 * 24 * N + 12 * (510 + 520 + 530)) / 4
    first = 8 * N + (16 + 4 * 1) * 730 + 8 * D             # view 1 preparing view 2 (1 layout)
    middle = 8 * 660 + 8 * 520 + (16 + 4 * 1) * 740 + 8 * D   # view 2 preparing view 3
    last = 16 * 530 + 8 * D                                 # view 3: no prep
    assert traffic.motion_sweep_bytes(vox, N, D) == (first + middle + last) / 3
    assert traffic.motion_middle_bytes(vox, N, D) == (
        8 * 640 + 8 * 510 + (16 + 4 * 1) * 730 + 8 * D + middle) / 2


def test_launches_per_scan(bench):
    spec = configs.c2_mcsart128()   # [1, 0] x 10 over 200 views, motion overload
    assert bench.launches_per_scan(spec) == (10 * 200, 10 * 199)


def test_profiled_picks_the_kernel_of_each_kind(bench, tmp_path, monkeypatch):
    kernels = [
        {"kernel": "k_backproject<1, 4, true, 1>", "launches": 199, "traffic_bytes": 11},
        {"kernel": "k_backproject<1, 4, true, 0>", "launches": 200, "traffic_bytes": 22},
        {"kernel": "k_backproject<1, 5, true, 0>", "launches": 3, "traffic_bytes": 99},
        {"kernel": "k_project_rows<1, 2, 2, 1>", "launches": 200, "traffic_bytes": 33},
        {"kernel": "k_project_rows<2, 2, 2, 1>", "launches": 199, "traffic_bytes": 44},
        {"kernel": "k_warp_tile<true, 32>", "launches": 400, "traffic_bytes": 55},
        {"kernel": "k_warp_add_tile<true, 32>", "launches": 200, "traffic_bytes": 66,
         "valu_busy": 0.4},
    ]
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "mcsart128_kernels.json").write_text(json.dumps({"kernels": kernels}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    want = {"motion": 11, "backprojector": 22, "projector": 33, "projector_motion": 44,
            "warp": 55, "warp_add": 66}
    for kind, v in want.items():   # the most-launched variant of the kind
        assert bench.profiled("mcsart128", kind) == v, kind
    assert bench.profiled("mcsart128", "warp_add", "valu_busy") == 0.4
    assert bench.profiled("mcsart128", "warp", "valu_busy") is None
    assert bench.profiled("clinical256", "warp") is None   # no summary for that workload


def _write_pmc(d, rows):
    d.mkdir()
    with open(d / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for i, (k, c, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})


def test_pmc_bytes_and_copy_calibration(tmp_path):
    """Counter values are KiB per dispatch; the run's last copy kernel is the calibration."""
    k = "void mcsart::(anonymous namespace)::k_warp_tile<true, 32>(float*, float const*)"
    _write_pmc(tmp_path / "f", [(k, "FETCH_SIZE", 100.0), (k, "FETCH_SIZE", 300.0),
                                ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 10.0),
                                ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 20.0)])
    per, calib = ps.load_pmc(str(tmp_path / "f"))
    assert per["k_warp_tile<true, 32>"] == 200.0 * 1024
    assert calib == 20.0 * 1024


def test_lds_table(tmp_path):
    k = "void mcsart::(anonymous namespace)::k_backproject<1, 4, true, 0>(mcsart::GeomArgs)"
    vals = {"SQ_LDS_BANK_CONFLICT": 50.0, "SQ_LDS_IDX_ACTIVE": 150.0, "SQ_INSTS_LDS": 7.0,
            "SQ_WAIT_INST_LDS": 20.0, "SQ_ACTIVE_INST_VALU": 300.0, "SQ_WAVE_CYCLES": 1000.0,
            "SQ_BUSY_CYCLES": 5.0, "SQ_WAIT_INST_ANY": 250.0}
    _write_pmc(tmp_path / "l", [(k, c, v) for c, v in vals.items()])
    md = ps.lds_markdown(str(tmp_path / "l"), "t", ["cmd"])
    row = [r for r in md.splitlines() if r.startswith("| `k_backproject")]
    assert row == ["| `k_backproject<1, 4, true, 0>` | 1 | 7 | 0.500 | 0.020 | 0.250 | 0.300 |"]


def test_occupancy_table(tmp_path):
    """MeanOccupancyPerCU averaged per (kernel, grid size): launches of different scan sizes
    stay apart."""
    k = "void mcsart::(anonymous namespace)::k_warp_tile<true, 32>(float*, float const*)"
    path = tmp_path / "o"
    path.mkdir()
    with open(path / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Grid_Size",
                                          "Workgroup_Size", "Counter_Name", "Counter_Value"])
        w.writeheader()
        rows = [(k, 2097152, "MeanOccupancyPerCU", 14.0), (k, 2097152, "MeanOccupancyPerCU", 16.0),
                (k, 2097152, "SQ_BUSY_CU_CYCLES", 99.0), (k, 8388608, "MeanOccupancyPerCU", 20.5),
                ("__amd_rocclr_copyBuffer", 1024, "MeanOccupancyPerCU", 1.0)]
        for i, (n, grid, c, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": n, "Grid_Size": grid,
                        "Workgroup_Size": 1024, "Counter_Name": c, "Counter_Value": v})
    md = ps.occupancy_markdown(str(path), "t", ["cmd"])
    rows = [r for r in md.splitlines() if r.startswith("| `")]
    assert rows == ["| `k_warp_tile<true, 32>` | 2097152 | 1024 | 32768 | 2 | 15.0 |",
                    "| `k_warp_tile<true, 32>` | 8388608 | 1024 | 131072 | 1 | 20.5 |"]


def test_per_view_trace_grouping(tmp_path):
    """mcsart_amd.profile_views puts each launch of a sweep trace on the view it serves (the
    image sweep's kernels on views 0 .. V-1, the motion sweep's on 1 .. V-1) and groups the motion
    step by the layouts of the view after it: with every launch's duration set to its view's
    index, each group's mean is the mean of the indices in it."""
    import numpy as np
    from mcsart_amd import profile_views as pv
    from mcsart_amd.driver import view_voxels
    spec = configs.c2_mcsart128()
    V = spec.n_views
    names = {"warp": "k_warp_tile<true, 32>", "projector": "k_project_rows<1, 2, 2, 2, true, true>",
             "backprojector": "k_backproject<1, 6, true, 0>", "warp_add": "k_warp_add_tile<true, 32>",
             "projector_motion": "k_project_rows<2, 2, 2, 1, true, false>",
             "motion": "k_backproject<1, 5, true, 1>"}
    offset = {k: 1000 * (i + 1) for i, k in enumerate(names)}
    order = [(k, v) for v in range(V) for k in ("warp", "projector", "backprojector", "warp_add")]
    order += [(k, v) for v in range(1, V) for k in ("projector_motion", "motion")]
    with open(tmp_path / "run_kernel_trace.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 0
        for k, v in order:
            ns = 1000 * (offset[k] + v)
            w.writerow([f"void mcsart::(anonymous namespace)::{names[k]}(float const*, int)", t,
                        t + ns])
            t += ns + 500
    res = pv.summarise(spec, pv.read_trace(str(tmp_path)))
    lay = view_voxels(spec.geometry, spec.angles)[:, 4]
    deg = np.degrees(spec.angles.astype(np.float64)) % 180.0
    for k in names:
        views = np.arange(0 if k in ("warp", "projector", "backprojector", "warp_add") else 1, V)
        served = np.minimum(views + 1, V - 1) if k == "motion" else views
        assert res[k]["launches"] == len(views)
        assert res[k]["mean_us"] == pytest.approx(offset[k] + views.mean())
        for code, name in pv.LAYOUTS.items():
            sel = views[lay[served] == code]
            if sel.size:
                assert res[k]["by_layout"][name] == (pytest.approx(offset[k] + sel.mean()), sel.size)
        lo = views[deg[views] < 15]
        assert res[k]["by_angle_15deg"][0] == pytest.approx(offset[k] + lo.mean())
    pv.write_md(spec, res, str(tmp_path / "views.md"), ["cmd"])
    assert "| motion | %d |" % (V - 1) in (tmp_path / "views.md").read_text()
