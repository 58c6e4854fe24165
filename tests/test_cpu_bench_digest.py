"""The profile digest of bench.py (host logic, no GPU): the split of
dispatches that left at once on a device stop flag, the selection of the
roofline kernel's fused launches, and the per-kernel table over synthetic
rocprofv3 CSV files."""
import csv
import os

import numpy as np

import bench


def test_split_early_separates_two_clusters():
    rng = np.random.default_rng(1)
    work = list(11.0 + 0.5 * rng.standard_normal(40))
    dead = list(1.7 + 0.1 * rng.standard_normal(60))
    hi, n = bench.split_early(work + dead)
    assert n == 60 and sorted(hi) == sorted(work)


def test_split_early_keeps_one_cluster_and_rare_outliers():
    rng = np.random.default_rng(2)
    one = list(5.0 + 0.3 * rng.standard_normal(100))
    assert bench.split_early(one) == (sorted(one), 0)
    # a few slow dispatches (under a tenth) are not a cluster of their own
    hi, n = bench.split_early([5.0] * 95 + [40.0] * 5)
    assert n == 0 and len(hi) == 100


def test_fused_launches_skip_the_first_after_the_marker():
    K, M = "k_mg_down<float, 2>", "k_pcg_start<float, double>"
    seq = []
    t = 0
    for step in range(3):  # each step: start, unfused first launch, 2 fused working, 2 early exits
        for k, v in ((M, 5.0), (K, 6.0), ("k_mg_down2<float>", 8.0), (K, 11.0), (K, 12.0), (K, 1.8), (K, 1.9)):
            seq.append((t, k, v))
            t += 1
    work, nearly, first = bench.fused_fine_down_values(seq)
    assert first == 3 and nearly == 6 and sorted(work) == [11.0] * 3 + [12.0] * 3
    # a solve whose start rides on the right-side kernel: that kernel marks the first launch
    seq = [(t, "k_prhs<false, float>" if k == M else k, v) for t, k, v in seq]
    work, nearly, first = bench.fused_fine_down_values(seq)
    assert first == 3 and nearly == 6 and sorted(work) == [11.0] * 3 + [12.0] * 3


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_profile_table_means_bytes_and_traffic_line(tmp_path):
    K, M = "void k_mg_down<float, 2>(args)", "void k_pcg_start<float, double>(args)"
    d = str(tmp_path)
    tr = [("Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp", "End_Timestamp")]
    seq = [(M, 1000), (K, 6000), (K, 10000), (K, 2000), (M, 1000), (K, 6000), (K, 12000), (K, 2000)]
    rows, t = [], 0
    for name, dur in seq:
        rows.append((name, 1024, 256, 1, t, t + dur))
        t += dur + 100
    _write(f"{d}/tr/t_kernel_trace.csv", tr[0], rows)
    ch = ("Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value")
    for cname, fused_val in (("FETCH_SIZE", 8000.0), ("WRITE_SIZE", 12000.0)):
        vals = {M: 10.0}
        crow = []
        for q, (name, dur) in enumerate(seq):
            v = vals.get(name, fused_val if dur >= 10000 else (3000.0 if dur == 6000 else 0.5))
            crow.append((q, name, 1024 * 256, cname, v))
        _write(f"{d}/{cname}/p_counter_collection.csv", ch, crow)
        # a 1 GiB copy calibrates the counter (FETCH counts 64 B per 128 B here, WRITE 1:1)
        big = 524288.0 if cname == "FETCH_SIZE" else 1048576.0
        _write(f"{d}/copy_{cname}/p_counter_collection.csv", ch, [(0, "__amd_rocclr_copyBuffer", 1 << 28, cname, big)])
    out = f"{d}/table.md"
    bench.profile_table(out, f"{d}/tr", f"{d}/FETCH_SIZE", f"{d}/WRITE_SIZE", f"{d}/copy_FETCH_SIZE",
                        f"{d}/copy_WRITE_SIZE", "t", roofline_bytes=27.76e6)
    text = open(out).read()
    # fused launches: 10 and 12 us, mean 11; the first-V-cycle launches (6 us) and the early exits are left out
    assert "2 dispatches that did the work (2 early exits and 2 unfused first-V-cycle launches left out), mean 11.00 us" in text
    # traffic: FETCH 8000 KiB x 2 + WRITE 12000 KiB x 1
    assert f"traffic_bytes_per_launch: {8000 * 1024 * 2 + 12000 * 1024:.0f}" in text
