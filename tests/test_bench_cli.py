"""bench.py's command-line contract that does not need a GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300)


def test_sparse_state_needs_a_directed_workload():
    r = run_bench("--workload", "plain", "--scale-down", "6", "--sparse-state")
    assert r.returncode != 0 and "directed LP" in r.stderr


def test_shard_index_is_checked():
    r = run_bench("--workload", "plain", "--scale-down", "6", "--shard", "3/2")
    assert r.returncode != 0 and "0 <= I < N" in r.stderr


def test_driver_flags_are_accepted_beyond_the_workload_batches():
    """--steps 20 --warmup 5 needs 25 insertion snapshots; the plain workload
    has 3, so the batches replay from a rebuilt initial state.  Without a
    device the run must get past argument checks to the first device call."""
    r = run_bench("--workload", "plain", "--scale-down", "6", "--steps", "20", "--warmup", "5")
    assert "insertion batches" not in r.stderr
    assert "workload plain:" in r.stderr
    if r.returncode != 0:       # no GPU here: the first device call is what stops it
        assert "hipGetDevice" in r.stderr or "device" in r.stderr


def test_zero_steps_is_refused():
    r = run_bench("--workload", "plain", "--scale-down", "6", "--steps", "0")
    assert r.returncode != 0 and "--steps must be >= 1" in r.stderr
