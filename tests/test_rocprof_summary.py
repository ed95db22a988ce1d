"""The rocprofv3 CSV reductions behind the committed profile summaries, on small synthetic CSVs."""
import csv
import json

from iqcp_amd import rocprof_summary as rs

CC_COLS = ["Dispatch_Id", "Kernel_Name", "VGPR_Count", "LDS_Block_Size", "Counter_Name", "Counter_Value"]


def write_counters(d, rows):
    d.mkdir(parents=True)
    with open(d / "p_counter_collection.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(CC_COLS)
        w.writerows(rows)


def write_stats(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        w.writerows(rows)


def test_names():
    assert rs.short_name("void qcg::jk_kernel<1, 0, 0, 0, double, 3>(int const*, int)") == \
        "jk_kernel<1, 0, 0, 0, double, 3>"
    assert rs.bench_label("jk_kernel<1, 0, 0, 0, double, 3>") == "jk_kernel<1, 0, 0, 0, double>"
    assert rs.bench_label("jk_kernel_mp<0, 0, 0, 0, 3>") == "jk_kernel_mp<0, 0, 0, 0>"
    assert rs.bench_label("jk_wave_kernel<2, 1, 2, 0, double>") == "jk_wave_kernel<2, 1, 2, 0, double>"
    assert rs.bench_label("fill_kernel") is None


def test_traffic_calibration_and_tiles(tmp_path):
    k_mp = "void qcg::jk_kernel_mp<0, 0, 0, 0, 3>(int)"
    k_64 = "void qcg::jk_kernel<1, 0, 0, 0, double, 3>(int)"
    copy = "__amd_rocclr_copyBuffer"
    # copy probe: FETCH_SIZE reads half of the 1 GiB (two counter rows per dispatch), WRITE_SIZE all of it
    write_counters(tmp_path / "cf", [[1, copy, 8, 0, "FETCH_SIZE", 262144.0], [1, copy, 8, 0, "FETCH_SIZE", 262144.0]])
    write_counters(tmp_path / "cw", [[1, copy, 8, 0, "WRITE_SIZE", 1048576.0]])
    write_counters(tmp_path / "f", [[1, k_mp, 64, 512, "FETCH_SIZE", 100.0], [2, k_mp, 64, 512, "FETCH_SIZE", 300.0],
                                    [3, k_64, 96, 256, "FETCH_SIZE", 50.0]])
    write_counters(tmp_path / "w", [[1, k_mp, 64, 512, "WRITE_SIZE", 1000.0], [2, k_mp, 64, 512, "WRITE_SIZE", 1000.0],
                                    [3, k_64, 96, 256, "WRITE_SIZE", 10.0]])
    tiles = {"fp64": [[[256, 1], [64, 1]] for _ in range(21)], "mixed": [[[128, 2], [64, 1]] for _ in range(21)]}
    tiles["mixed"][20] = [[64, 8], [64, 1]]   # (ss|ss) is the last class
    tiles["fp64"][19] = [[128, 1], [64, 1]]   # (ps|ss) is next to last
    (tmp_path / "tiles.json").write_text(json.dumps(tiles))
    rs.main(["traffic", "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"),
             "--copy-fetch", str(tmp_path / "cf"), "--copy-write", str(tmp_path / "cw"),
             "--tiles", str(tmp_path / "tiles.json"), "--json", str(tmp_path / "t.json"),
             "--md", str(tmp_path / "t.md"), "--command", "rocprofv3 --pmc FETCH_SIZE -- python3 bench.py"])
    t = json.loads((tmp_path / "t.json").read_text())
    assert abs(t["calibration"]["fetch_factor"] - 2.0) < 1e-12 and t["calibration"]["write_factor"] == 1.0
    mp = t["kernels"]["jk_kernel_mp<0, 0, 0, 0>"]
    assert mp["fetch_bytes"] == round(200.0 * 1024 * 2) and mp["write_bytes"] == 1000 * 1024
    assert mp["tile"] == [64, 8] and mp["dispatches"] == 2 and mp["lds_rows"] == "3"
    assert t["kernels"]["jk_kernel<1, 0, 0, 0, double>"]["tile"] == [128, 1]
    md = (tmp_path / "t.md").read_text()
    assert "rocprofv3 --pmc FETCH_SIZE" in md and "`jk_kernel_mp<0, 0, 0, 0>` | (64, 8) | 2 |" in md


def test_counters_states_and_f64(tmp_path):
    k = "void qcg::jk_kernel<0, 0, 0, 0, double, 3>(int)"
    write_counters(tmp_path / "s", [
        [1, k, 64, 184, "SQ_WAVES", 1000.0], [1, k, 64, 184, "SQ_WAVE_CYCLES", 1e6],
        [1, k, 64, 184, "SQ_WAIT_ANY", 5e5], [1, k, 64, 184, "SQ_WAIT_INST_ANY", 2e5],
        [1, k, 64, 184, "SQ_ACTIVE_INST_VALU", 1e5], [1, k, 64, 184, "SQ_ACTIVE_INST_SCA", 5e4],
        [1, k, 64, 184, "SQ_ACTIVE_INST_LDS", 2.5e4], [1, k, 64, 184, "SQ_INSTS_VALU", 6.144e6]])
    write_counters(tmp_path / "d", [
        [1, k, 64, 184, "SQ_INSTS_VALU_FMA_F64", 1e6], [1, k, 64, 184, "SQ_INSTS_VALU_MUL_F64", 5e5],
        [1, k, 64, 184, "SQ_INSTS_VALU_ADD_F64", 5e5], [1, k, 64, 184, "SQ_INSTS_VALU_TRANS_F64", 1e4]])
    write_stats(tmp_path / "ks.csv", [[k, 10, 1e6, 100000.0, 50.0, 90000, 110000, 1.0]])
    rs.main(["counters", "--pmc", str(tmp_path / "s"), "--pmc", str(tmp_path / "d"),
             "--stats", str(tmp_path / "ks.csv"), "--md", str(tmp_path / "c.md")])
    row = [l for l in (tmp_path / "c.md").read_text().splitlines() if l.startswith("| `jk_kernel")][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    # dispatches, waves, VGPRs, LDS, avg µs, wait counter, wait issue, VALU, scalar, LDS, issue rate
    assert cells[1:11] == ["1", "1000", "64", "184", "100.0", "0.50", "0.20", "0.10", "0.05", "0.03"]
    assert cells[11] == f"{6.144e6 * 4 / (1024 * 2.4e9 * 1e-4):.2f}"
    fl = 64 * (2e6 + 5e5 + 5e5) / 1e-4
    assert cells[12] == "2.01 M" and cells[13] == f"{fl / 1e12:.2f}" and cells[14] == f"{fl / 78.6e12:.3f}"


def test_mfma_states_and_rate(tmp_path):
    k = "void qcg::xc_v_kernel<float, 5>(int, int)"
    other = "void qcg::jk_kernel<0, 0, 0, 0, double, 1>(int)"
    write_counters(tmp_path / "s", [
        [1, k, 120, 0, "SQ_WAVES", 4096.0], [1, k, 120, 0, "SQ_WAVE_CYCLES", 1e6],
        [1, k, 120, 0, "SQ_WAIT_ANY", 3e5], [1, k, 120, 0, "SQ_WAIT_INST_ANY", 2e5],
        [1, k, 120, 0, "SQ_ACTIVE_INST_VALU", 1e5], [1, k, 120, 0, "SQ_ACTIVE_INST_LDS", 5e4],
        [1, other, 64, 0, "SQ_WAVES", 10.0]])
    write_counters(tmp_path / "m", [
        [1, k, 120, 0, "SQ_INSTS_VALU_MFMA_F32", 9.6e7], [1, k, 120, 0, "SQ_INSTS_VALU_MFMA_F64", 0.0],
        [1, k, 120, 0, "SQ_VALU_MFMA_BUSY_CYCLES", 3.072e9]])
    write_stats(tmp_path / "ks.csv", [[k, 10, 2e7, 2000000.0, 40.0, 1900000, 2100000, 1.0]])
    rs.main(["mfma", "--pmc", str(tmp_path / "s"), "--pmc", str(tmp_path / "m"), "--stats", str(tmp_path / "ks.csv"),
             "--match", "xc_", "--md", str(tmp_path / "x.md")])
    rows = [l for l in (tmp_path / "x.md").read_text().splitlines() if l.startswith("| `")]
    assert len(rows) == 1 and "jk_kernel" not in rows[0]
    cells = [c.strip() for c in rows[0].strip("|").split("|")]
    # dispatches, VGPRs, avg µs, four wave states, MFMA instructions, busy, TFLOP/s, fraction of peak
    assert cells[1:8] == ["1", "120", "2000.0", "0.30", "0.20", "0.10", "0.05"]
    assert cells[8] == "96.00 M (F32)"
    assert cells[9] == f"{3.072e9 / (1024 * 2.4e9 * 2e-3):.2f}"
    fl = 9.6e7 * 2048 / 2e-3
    assert cells[10] == f"{fl / 1e12:.1f}" and cells[11] == f"{fl / 157.3e12:.3f}"
