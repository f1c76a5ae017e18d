"""Host-side logic without a GPU: the profile summary of bench.py (level-0 selection, including
two levels that share the largest grid, and the copy calibration of the traffic counters), the
warm-start network's initial weights, and the random-ellipsoid workload generator's ranges."""
import csv
import json
import os

import numpy as np
import pytest

import bench
from __graft_entry__ import load_package


def _write(path, rows, fields):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_summary_splits_levels_sharing_the_largest_grid(tmp_path):
    """Levels 0 and 1 of the half sweep launched with the same grid (the tiles aim every level at
    the same block count): the summary keeps the slow group as level 0; traffic uses the largest
    counter value of each kernel scaled by the copy calibration."""
    name = "void wl::k_gs_half<float>(wl::Geo, wl::ZT, wl::LRows<float>, float const*, float*, int)"
    stats = tmp_path / "x_kernel_stats.csv"
    trace = tmp_path / "x_kernel_trace.csv"
    _write(stats, [dict(Name=name, Calls=8, AverageNs=90000, TotalDurationNs=720000, Percentage=100)],
           ["Name", "Calls", "AverageNs", "TotalDurationNs", "Percentage"])
    rows = []
    for k in range(8):
        dur = 170000 if k % 2 == 0 else 17000  # levels 0 and 1 interleaved, same grid
        rows.append(dict(Kernel_Name=name, Grid_Size_X=262144, Grid_Size_Y=4, Grid_Size_Z=1,
                         Start_Timestamp=1000 * k, End_Timestamp=1000 * k + dur))
    rows.append(dict(Kernel_Name=name, Grid_Size_X=65536, Grid_Size_Y=4, Grid_Size_Z=1,
                     Start_Timestamp=0, End_Timestamp=5000))
    _write(trace, rows, ["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp",
                         "End_Timestamp"])
    fetch = tmp_path / "f.csv"
    write = tmp_path / "w.csv"
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(fetch, [dict(Kernel_Name="__amd_rocclr_copyBuffer", Counter_Name="FETCH_SIZE", Counter_Value=524288),
                   dict(Kernel_Name=name, Counter_Name="FETCH_SIZE", Counter_Value=1000),
                   dict(Kernel_Name=name, Counter_Name="FETCH_SIZE", Counter_Value=300000)], fields)
    _write(write, [dict(Kernel_Name="__amd_rocclr_copyBuffer", Counter_Name="WRITE_SIZE", Counter_Value=1048576),
                   dict(Kernel_Name=name, Counter_Name="WRITE_SIZE", Counter_Value=130000)], fields)
    md, js = tmp_path / "s.md", tmp_path / "t.json"
    bench.summarize(str(stats), str(fetch), str(write), str(md), str(js), "cmd")
    text = md.read_text()
    line = [ln for ln in text.splitlines() if ln.startswith("| gs_half@L0")][0]
    assert "| 4 | 170.00" in line and "split from 4" in line
    t = json.loads(js.read_text())
    # the copy reads half of what it moves on FETCH_SIZE: scale 2; WRITE_SIZE scale 1
    assert t["_fetch_scale"] == pytest.approx(2.0) and t["_write_scale"] == pytest.approx(1.0)
    assert t["gs_half@L0"] == pytest.approx((300000 * 2.0 + 130000 * 1.0) * 1024)
    # kernel-trace-only summaries
    md2 = tmp_path / "s2.md"
    bench.summarize(str(stats), "-", "-", str(md2), "-", "cmd")
    assert "gs_half@L0" in md2.read_text() and "HBM traffic" not in md2.read_text()


def test_summary_level0_over_all_instances_of_a_key(tmp_path):
    """Two instances of one key (the prolongation with and without the folded coarse increment):
    the one on the largest grid is level 0, the other is listed as coarser-only, and the key's
    traffic is the larger one, not their mean."""
    hot = "void wl::k_prolong_inc<float, true>(wl::Geo, wl::Geo, wl::ZT, int, int)"
    cold = "void wl::k_prolong_inc<float, false>(wl::Geo, wl::Geo, wl::ZT, int, int)"
    stats = tmp_path / "y_kernel_stats.csv"
    trace = tmp_path / "y_kernel_trace.csv"
    _write(stats, [dict(Name=hot, Calls=2, AverageNs=30000, TotalDurationNs=60000, Percentage=90),
                   dict(Name=cold, Calls=2, AverageNs=4000, TotalDurationNs=8000, Percentage=10)],
           ["Name", "Calls", "AverageNs", "TotalDurationNs", "Percentage"])
    rows = [dict(Kernel_Name=hot, Grid_Size_X=262144, Grid_Size_Y=4, Grid_Size_Z=1, Start_Timestamp=0,
                 End_Timestamp=30000) for _ in range(2)]
    rows += [dict(Kernel_Name=cold, Grid_Size_X=4096, Grid_Size_Y=4, Grid_Size_Z=1, Start_Timestamp=0,
                  End_Timestamp=4000) for _ in range(2)]
    _write(trace, rows, ["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp",
                         "End_Timestamp"])
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fetch, [dict(Kernel_Name="__amd_rocclr_copyBuffer", Counter_Name="FETCH_SIZE", Counter_Value=524288),
                   dict(Kernel_Name=hot, Counter_Name="FETCH_SIZE", Counter_Value=80000),
                   dict(Kernel_Name=cold, Counter_Name="FETCH_SIZE", Counter_Value=2000)], fields)
    _write(write, [dict(Kernel_Name="__amd_rocclr_copyBuffer", Counter_Name="WRITE_SIZE", Counter_Value=1048576),
                   dict(Kernel_Name=hot, Counter_Name="WRITE_SIZE", Counter_Value=40000),
                   dict(Kernel_Name=cold, Counter_Name="WRITE_SIZE", Counter_Value=1000)], fields)
    md, js = tmp_path / "s.md", tmp_path / "t.json"
    bench.summarize(str(stats), str(fetch), str(write), str(md), str(js), "cmd")
    text = md.read_text()
    assert "| prolong_inc@L0 | wl::k_prolong_inc<float, true> | 2 | 30.00 |" in text
    assert "| prolong_inc (coarser levels only) | wl::k_prolong_inc<float, false> |" in text
    t = json.loads(js.read_text())
    assert t["prolong_inc@L0"] == pytest.approx((80000 * 2.0 + 40000 * 1.0) * 1024)


def test_summary_labels_the_coarse_tail(tmp_path):
    """The coarse tail is one workgroup: its only grid is its largest, yet it never runs level 0,
    so the kernel-trace table names it as the coarse-level kernel."""
    tail = "void wl::k_mg_tail_lds<float>(wl::TailDesc<float>, float const*, int, int, int)"
    stats = tmp_path / "z_kernel_stats.csv"
    trace = tmp_path / "z_kernel_trace.csv"
    _write(stats, [dict(Name=tail, Calls=3, AverageNs=30000, TotalDurationNs=90000, Percentage=100)],
           ["Name", "Calls", "AverageNs", "TotalDurationNs", "Percentage"])
    _write(trace, [dict(Kernel_Name=tail, Grid_Size_X=1024, Grid_Size_Y=1, Grid_Size_Z=1, Start_Timestamp=0,
                        End_Timestamp=30000) for _ in range(3)],
           ["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp", "End_Timestamp"])
    md = tmp_path / "s.md"
    bench.summarize(str(stats), "-", "-", str(md), "-", "cmd")
    text = md.read_text()
    assert "| mg_tail (coarse levels, one workgroup) | wl::k_mg_tail_lds<float> | 3 | 30.00 |" in text
    assert "mg_tail@L0" not in text


def test_initial_network_weights():
    wl = load_package()
    w = wl.init_weights(3)
    sizes = [27 * ci * co for ci, co in wl.learned.LAYERS]
    assert w.size == sum(sizes) and w.dtype == np.float32
    first = w[:sizes[0]].reshape(27, 8, 16)
    assert not np.any(first[:, 3:, :]) and np.all(np.abs(first[:, :3, :]).sum(axis=(0, 1)) > 0)
    last = w[-sizes[-1]:].reshape(27, 16, 8)
    assert not np.any(last[:, :, 1:]) and np.any(last[:, :, 0])
    assert np.array_equal(wl.init_weights(3), w) and not np.array_equal(wl.init_weights(4), w)


def test_random_ellipsoid_workload_ranges(monkeypatch):
    """The C4 generator: 1-3 ellipsoids, semi-axes in [8,24], centres in the middle half, Re in
    [500, 5000] (ν = 32/Re), for seeds 0..49 (checked on the arguments it passes on)."""
    wl = load_package()
    made = []

    class FakeSim:
        def __init__(self, n, uBC, L, ν, body, T):
            made.append((n, L, ν, body))

    monkeypatch.setattr(wl.api, "Simulation", FakeSim)
    for seed in range(50):
        wl.learned.random_ellipsoid_simulation(seed)
    for n, L, nu, body in made:
        parts = body.parts()
        assert 1 <= len(parts) <= 3 and L == 32 and 500 <= 32 / nu <= 5000
        for p in parts:
            assert all(8 <= a <= 24 for a in p.axes)
            assert all(n[d] / 4 <= p.center[d] <= 3 * n[d] / 4 for d in range(3))
