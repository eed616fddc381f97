"""bench.py --summarize on synthetic rocprofv3 CSV files: the dispatch
picks, the copy calibration and the per-launch traffic that the committed
profile summaries (and the benchmark's roofline.traffic) are made of."""
import csv
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

KERNEL = "void k_join_x<8, false>(JoinArgs, unsigned short const*)"


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_summary_picks_calibrates_and_sums(tmp_path, capsys):
    kt, pf, pw = (str(tmp_path / n) for n in ("kt", "pf", "pw"))
    # kernel trace: five dispatches of the picked kernel lasting 1..5 ms,
    # another kernel in between
    trace = []
    for i in range(5):
        start = 1_000_000_000 * (i + 1)
        trace.append([10 + 2 * i, KERNEL, start, start + (i + 1) * 1_000_000])
        trace.append([11 + 2 * i, "k_other(int)", start, start + 7_000_000])
    _write(os.path.join(kt, "run_kernel_trace.csv"), ["Dispatch_Id", "Kernel_Name", "Start_Timestamp",
                                                       "End_Timestamp"], trace)
    # counters: three 1 GiB calibration copies first, then the kernel's
    # dispatches; FETCH_SIZE reads half of a copy's bytes (factor 2),
    # WRITE_SIZE all of them (factor 1)
    kib = bench.COPY_BYTES / 1024
    fetch = [[d, "__amd_rocclr_copyBuffer", "FETCH_SIZE", kib / 2] for d in (1, 2, 3)]
    write = [[d, "__amd_rocclr_copyBuffer", "WRITE_SIZE", kib] for d in (1, 2, 3)]
    for i in range(5):
        fetch.append([10 + 2 * i, KERNEL, "FETCH_SIZE", 1000.0 * (i + 1)])
        write.append([10 + 2 * i, KERNEL, "WRITE_SIZE", 100.0 * (i + 1)])
        fetch.append([11 + 2 * i, "k_other(int)", "FETCH_SIZE", 9e9])
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(os.path.join(pf, "run_counter_collection.csv"), hdr, fetch)
    _write(os.path.join(pw, "run_counter_collection.csv"), hdr, write)

    bench.summarize([kt, pf, pw], ["k_join_x<8, false>@-3:"])
    out = capsys.readouterr().out
    lines = {l.split(":")[0]: l for l in out.splitlines() if ":" in l}
    # the last three dispatches: 3, 4, 5 ms; FETCH 3000..5000 KiB, WRITE 300..500 KiB
    assert "avg_duration_ms: 4.0000  (dispatches: 3)" in out
    assert "copy calibration factor 2.0000" in lines["FETCH_SIZE"]
    assert "copy calibration factor 1.0000" in lines["WRITE_SIZE"]
    expect = 4000.0 * 1024 * 2.0 + 400.0 * 1024 * 1.0
    got = float(lines["traffic_bytes_per_launch"].split(":")[1])
    assert abs(got - expect) <= 1e-6 * expect


def test_read_traffic_parses_committed_profiles():
    for prec in ("fp32", "bf16", "learned"):
        tr, src = bench.read_traffic(prec)
        assert tr is not None and tr > 1e9, prec
        assert src.endswith(f"join_short_{prec}_jmerge5M_pmc.txt")
    assert bench.read_traffic("no_such_path") == (None, None)
