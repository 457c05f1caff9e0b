"""CPU checks of bench.py's contract pieces that need no GPU: the one-GPU
guard, the selection of one, two or three steps per launch for the lattices a
leg creates (and its restoration of the environment), the roofline
arithmetic (bytes a launch must move, not per update), the traffic file's
source stamp (a stale file is refused), and the segmentation of a profiled
run's counter file into workloads."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_pass_mode_sets_and_restores_the_switch(monkeypatch):
    monkeypatch.delenv("FFB_TWO_STEP", raising=False)
    monkeypatch.delenv("FFB_STEPS_PER_PASS", raising=False)
    forms = ("fp32", "fp64", "fp16dev", "bf16dev", "fp32dev", "e3m12dev")
    assert [bench.steps_per_pass(p) for p in forms] == [5, 4, 3, 3, 5, 2]
    with bench.pass_mode(1):
        assert os.environ["FFB_TWO_STEP"] == "0" and bench.steps_per_pass("fp32") == 1
    assert "FFB_TWO_STEP" not in os.environ and "FFB_STEPS_PER_PASS" not in os.environ
    with bench.pass_mode(2):
        assert os.environ["FFB_STEPS_PER_PASS"] == "2" and bench.steps_per_pass("fp64") == 2
    monkeypatch.setenv("FFB_TWO_STEP", "1")
    with bench.pass_mode(None):
        assert bench.steps_per_pass("e3m12dev") == 2 and bench.steps_per_pass("fp32") == 5
    with bench.pass_mode(1):
        assert os.environ["FFB_TWO_STEP"] == "0"
    assert os.environ["FFB_TWO_STEP"] == "1" and "FFB_STEPS_PER_PASS" not in os.environ
    monkeypatch.setenv("FFB_STEPS_PER_PASS", "4")  # four steps only where a four-step kernel exists
    assert bench.steps_per_pass("fp64") == 4 and bench.steps_per_pass("bf16dev") == 3
    monkeypatch.setenv("FFB_STEPS_PER_PASS", "5")  # five steps only for the 4-byte forms
    assert bench.steps_per_pass("fp32dev") == 5 and bench.steps_per_pass("fp64") == 4
    assert bench.steps_per_pass("e3m12dev") == 3
    monkeypatch.setenv("FFB_TWO_STEP", "0")  # one launch per step wins over the steps per pass
    monkeypatch.setenv("FFB_STEPS_PER_PASS", "3")
    assert bench.steps_per_pass("fp32") == 1


def test_batch_steps_per_pass_selection(monkeypatch):
    monkeypatch.delenv("FFB_BATCH_PAIRS", raising=False)
    monkeypatch.delenv("FFB_BATCH_STEPS_PER_PASS", raising=False)
    assert [bench.batch_steps_per_pass(p) for p in ("fp32", "fp64", "fp32dev")] == [5, 3, 5]
    assert bench.batch_steps_per_pass("e3m12dev") == 3 and bench.batch_steps_per_pass("fp16dev") == 3
    monkeypatch.setenv("FFB_BATCH_STEPS_PER_PASS", "2")
    assert bench.batch_steps_per_pass("fp32") == 2 and bench.batch_steps_per_pass("fp16dev") == 2
    monkeypatch.setenv("FFB_BATCH_PAIRS", "0")
    assert bench.batch_steps_per_pass("fp32") == 1


def test_roofline_counts_one_read_and_write_per_launch():
    """A pair launch of the FP32 4096^2 cavity must move 18*4*4096^2 bytes,
    not twice that: its fraction is that over the launch time and 8 TB/s, at
    most 1 for any launch not faster than the peak allows."""
    must = 18 * 4 * 4096 * 4096
    run = {"achieved_gbs": must / 0.2e-3 / 1e9, "bytes_per_launch": must, "steps_per_launch": 2,
           "kernel_ms": 0.2}
    r = bench.roofline(run, "no_such_workload")
    assert abs(r["frac"] - must / 0.2e-3 / 8e12) < 1e-4 and r["frac"] < 1
    rates = bench.hbm_rates(160000.0, "fp32", 2)
    assert abs(rates["min_bytes_gbs"] - 160000e6 * 72 / 2 / 1e9) < 0.1
    assert abs(rates["effective_gbs"] - 2 * rates["min_bytes_gbs"]) < 0.2


def test_traffic_file_is_stamped_and_a_stale_one_refused(monkeypatch, tmp_path):
    stamp = bench.kernel_source_hash()
    good = tmp_path / "good.txt"
    good.write_text(f"# kernel sources {stamp}\nsl4096_fp32 700000000 610000000 6 ffb::step2_kernel<float>\n")
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(good))
    t = bench.traffic_from_profiles("sl4096_fp32")
    assert t["bytes"] == 1.31e9 and t["kernel"] == "ffb::step2_kernel<float>"
    assert bench.traffic_from_profiles("no_such_workload") is None
    stale = tmp_path / "stale.txt"
    stale.write_text("# kernel sources 0123456789abcdef\nsl4096_fp32 700000000 610000000 6 k\n")
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(stale))
    assert bench.traffic_from_profiles("sl4096_fp32") is None
    assert "not used" in bench.traffic_status()


def test_committed_traffic_file_matches_the_sources():
    """profiles/traffic.txt is re-taken whenever a kernel source changes."""
    stamp, table = bench._traffic_table()
    assert stamp == bench.kernel_source_hash(), bench.traffic_status()
    names = [n for n, _ in bench._profile_workloads()]
    assert set(names) <= set(table)
    pair, single = table["sl4096_fp32"]["bytes"], table["sl4096_fp32_single_pass"]["bytes"]
    assert 18 * 4 * 4096 * 4096 < pair < 2 * single


def test_profile_segments_follow_the_markers():
    m1, m2 = bench.MARK_START // 16, bench.MARK_END // 16
    disp = [("init_kernel", 100, 1.0), ("copy_kernel<0>", 2097152, 5), ("copy_kernel<1>", m1, 1),
            ("copy_kernel<3>", m1, 1), ("step2_kernel<a>", 10, 100.0), ("step2_kernel<a>", 10, 102.0),
            ("copy_kernel<1>", m2, 1), ("batch_pair_kernel", 9, 5.0), ("batch_pair_kernel", 9, 5.0),
            ("copy_kernel<1>", m1, 1), ("batch_count_kernel", 1, 0.1), ("batch_pair_kernel", 9, 7.0),
            ("copy_kernel<2>", m2, 1)]
    assert bench._segments(disp) == [("step2_kernel<a>", 101.0, 2), ("batch_pair_kernel", 7.0, 1)]


def test_valu_record_from_counters(monkeypatch, tmp_path):
    """--record-valu: per workload the mean SQ_INSTS_VALU of its kernel's
    dispatches and busy = insts * 4 / (1024 SIMDs * GRBM_GUI_ACTIVE / 8)."""
    m1, m2 = bench.MARK_START // 16, bench.MARK_END // 16
    monkeypatch.setattr(bench, "_profile_workloads", lambda: [("w1", None), ("w2", None)])
    rows, did = [], [0]

    def dispatch(kernel, grid, vals):
        did[0] += 1
        for c in bench.VALU_COUNTERS:
            rows.append(f"{did[0]},{kernel},{grid},{c},{vals.get(c, 0.0)}")

    cnt = {"SQ_INSTS_VALU": 1.0e8, "SQ_ACTIVE_INST_VALU": 2.0e7, "SQ_WAVES": 1000.0,
           "GRBM_GUI_ACTIVE": 8 * 500000.0, "SQ_BUSY_CYCLES": 1.0}
    dispatch("copy_kernel<1>", m1, {})
    dispatch("stepk_kernel<a>", 10, cnt)
    dispatch("stepk_kernel<a>", 10, {**cnt, "SQ_INSTS_VALU": 1.2e8})
    dispatch("copy_kernel<1>", m2, {})
    dispatch("copy_kernel<1>", m1, {})
    dispatch("batch_pair_kernel", 9, cnt)
    dispatch("copy_kernel<1>", m2, {})
    path = tmp_path / "sq.csv"
    path.write_text("Dispatch_Id,Kernel_Name,Grid_Size,Counter_Name,Counter_Value\n" + "\n".join(rows) + "\n")
    monkeypatch.setattr(bench, "VALU_FILE", str(tmp_path / "valu.txt"))
    assert bench.record_valu(str(path)) == 0
    v = bench.valu_from_profiles("w1")
    assert v["wave_insts_per_launch"] == 1.1e8 and v["kernel"] == "stepk_kernel<a>"
    assert abs(v["busy"] - round(1.1e8 * 4 / (1024 * 500000.0), 3)) < 1e-9
    run = {"achieved_gbs": 3000.0, "bytes_per_launch": 1, "steps_per_launch": 5, "kernel_ms": 0.4}
    assert bench.roofline(run, "w1")["limiter"] == "neither"  # busy 0.859, no traffic record
    assert bench.roofline(run, "no_such_workload").get("limiter") is None
    text = (tmp_path / "valu.txt").read_text().replace(bench.kernel_source_hash(), "0123456789abcdef")
    (tmp_path / "valu.txt").write_text(text)
    assert bench.valu_from_profiles("w1") is None  # stamped with other sources: refused


def test_committed_valu_file_matches_the_sources():
    """profiles/valu.txt is re-taken with profiles/traffic.txt."""
    assert bench.valu_from_profiles("sl4096_fp32") is not None
