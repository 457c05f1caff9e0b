"""CPU checks of bench.py's contract pieces that need no GPU: the one-GPU
guard, the selection of one launch per step or step pairs for the lattices a
leg creates (and its restoration of the environment), and the traffic lookup
from profiles/traffic.txt."""
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
    assert bench.uses_pairs("fp32") and bench.uses_pairs("fp64") and bench.uses_pairs("fp16dev")
    assert not bench.uses_pairs("e3m12dev")
    with bench.pass_mode(False):
        assert os.environ["FFB_TWO_STEP"] == "0" and not bench.uses_pairs("fp32")
    assert "FFB_TWO_STEP" not in os.environ
    monkeypatch.setenv("FFB_TWO_STEP", "1")
    with bench.pass_mode(None):
        assert bench.uses_pairs("e3m12dev")
    with bench.pass_mode(False):
        assert os.environ["FFB_TWO_STEP"] == "0"
    assert os.environ["FFB_TWO_STEP"] == "1"


def test_batch_pair_selection(monkeypatch):
    monkeypatch.delenv("FFB_BATCH_PAIRS", raising=False)
    assert bench.batch_uses_pairs("fp32") and bench.batch_uses_pairs("fp64")
    assert not bench.batch_uses_pairs("e3m12dev") and not bench.batch_uses_pairs("fp16dev")
    monkeypatch.setenv("FFB_BATCH_PAIRS", "0")
    assert not bench.batch_uses_pairs("fp32")


def test_traffic_lines_for_both_paths():
    for workload in ("sl4096_fp32", "sl4096_fp64", "sl4096_fp32_single_pass", "sl4096_fp64_single_pass"):
        t = bench.traffic_from_profiles(workload)
        assert t is not None and t > 1e9, workload
    # the pair moves fewer bytes than two single steps, more than one step's minimum
    pair, single = bench.traffic_from_profiles("sl4096_fp32"), bench.traffic_from_profiles("sl4096_fp32_single_pass")
    assert 18 * 4 * 4096 * 4096 < pair < 2 * single
    assert bench.traffic_from_profiles("no_such_workload") is None
