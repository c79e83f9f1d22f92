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


def test_more_steps_than_insertion_batches_is_refused_before_any_device_call():
    r = run_bench("--workload", "plain", "--scale-down", "6", "--steps", "500", "--warmup", "1")
    assert r.returncode != 0 and "insertion batches; warmup+steps=501" in r.stderr
    assert "device " not in r.stderr
