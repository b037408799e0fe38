"""bench.py's host-side contract: the one-GPU guard and the HBM-traffic
calibration of the roofline line (no GPU needed)."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def write_counters(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, kib in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": kib})


def test_pmc_traffic_calibrates_on_the_copy_kernel(tmp_path):
    # k_copy16 reads and writes CALIBRATION_BYTES once each.  The synthetic
    # counters report half of its reads (FETCH factor 2) and all of its writes
    # (WRITE factor 1), so the advance kernel's calibrated traffic is 2 x its
    # FETCH plus its WRITE, divided by runs x working launches.
    cb = bench.CALIBRATION_BYTES
    fetch_copy = cb / 2 / 1024  # KiB as the counter reports it
    write_copy = cb / 1024
    copy = "(anonymous namespace)::k_copy16(HIP_vector_type<unsigned int, 4u>*, HIP_vector_type<unsigned int, 4u> const*, unsigned long)"
    adv = "void (anonymous namespace)::k_adv_dense<2, 256, false, 16, 1>(cgpk::KP)"
    other = "void (anonymous namespace)::k_adv_push<2, 256>(cgpk::KP)"
    out = "(anonymous namespace)::k_output(unsigned int const*, int)"
    (tmp_path / "fetch").mkdir()
    (tmp_path / "write").mkdir()
    write_counters(tmp_path / "fetch" / "f_counter_collection.csv", "FETCH_SIZE",
                   [(copy, fetch_copy), (adv, 100.0), (adv, 300.0), (other, 5000.0), (out, 12345.0)])
    write_counters(tmp_path / "write" / "w_counter_collection.csv", "WRITE_SIZE",
                   [(copy, write_copy), (adv, 40.0), (other, 7000.0), (out, 777.0)])
    traffic, factors = bench.pmc_traffic(str(tmp_path), "k_adv_dense<", runs=2, launches_per_run=1.0)
    assert factors["fetch_size_factor"] == pytest.approx(2.0)
    assert factors["write_size_factor"] == pytest.approx(1.0)
    assert traffic == pytest.approx((2.0 * (100.0 + 300.0) + 40.0) * 1024.0 / 2.0)


def test_pmc_traffic_without_calibration_rows_is_none(tmp_path):
    write_counters(tmp_path / "x_counter_collection.csv", "FETCH_SIZE",
                   [("void k_adv_dense<1, 256, false, 16, 1>(KP)", 10.0)])
    assert bench.pmc_traffic(str(tmp_path), "k_adv_dense<", 1, 1.0) == (None, None)


def test_byte_models_from_work_counters():
    # One launch's counters by hand: 10 push items, 100 edges (40 with a
    # weight), 70 (edge, BFS word) reads, 90 (edge, SSSP query) pairs in 30
    # (edge, SSSP word) pairs, 5 next words, 7 lowered values, 12 (item, SSSP
    # query) sources; lane-per-query: 3 items, 50 edges, 400 pairs, 9 lowered,
    # 2 next words, 20 (item, query) sources.  W = 4 mask words.
    w = [0] * 24
    w[bench.PUSH_ITEMS], w[bench.PUSH_EDGES], w[bench.PUSH_WEDGES] = 10, 100, 40
    w[bench.BFS_WORDS], w[bench.SSSP_PAIRS], w[bench.SSSP_WORDS] = 70, 90, 30
    w[bench.NEW_WORDS], w[bench.IMPROVED], w[bench.SSSP_SRC] = 5, 7, 12
    w[bench.DENSE_ITEMS], w[bench.DENSE_EDGES], w[bench.DENSE_PAIRS] = 3, 50, 400
    w[bench.DENSE_IMP], w[bench.DENSE_NEW], w[bench.DENSE_SRC] = 9, 2, 20
    push, dense = bench.bytes_builder(w, 4)
    assert push == 10 * 40 + 100 * 4 + 40 * 4 + 70 * 8 + 90 * 4 + 7 * 4 + 5 * 16
    assert dense == 3 * 16 + 50 * 8 + 400 * 4 + 9 * 4 + 2 * 16
    push, dense = bench.bytes_scope_formula(w, 4)
    assert push == 10 * 40 + 100 * 4 + 40 * 4 + (70 + 30) * 8 + 5 * 16 + 7 * 4 + 12 * 4
    assert dense == 3 * 40 + 50 * 16 + 2 * 16 + 9 * 4 + 20 * 4
