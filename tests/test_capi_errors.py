"""Error paths of the C ABI that need no GPU: bad arguments are refused with
an error code and message, and a server whose creation fails before any of
its threads starts (here: no device, so its pipeline cannot be created) is
torn down without aborting the process.  The failure after the finisher
thread has started is covered on the device by
test_gpu_server_teardown.py.  Each case runs in a child process so that an
abort would be reported as a failure rather than end the test run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

CHILD = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, {pkg!r})
from treesvd_amd import capi
L = capi.lib()
cfg = capi.ServerConfig(0, 0, 0)
h = ctypes.c_void_p()
rc = L.tsvd_server_create(None, ctypes.byref(cfg), ctypes.byref(h))
print("null-config", rc, L.tsvd_last_error().decode())
print("destroy-null", L.tsvd_server_destroy(None))
t = ctypes.c_int64()
print("submit-null", L.tsvd_server_submit(None, None, None, 0, ctypes.byref(t)))
print("wait-null", L.tsvd_server_wait(None, 0, None, None))
try:
    capi.Server(100, False, np.arange(8), 0.2, 1e-6, 4, 2, 4)
    print("created")
except capi.TreeSVDError as e:
    print("create-failed", e)
print("alive")
"""


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible: creation would succeed")
def test_server_error_paths_without_a_device():
    code = CHILD.format(pkg=os.path.join(ROOT, "tree-embedding_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    out = r.stdout
    assert r.returncode == 0, (r.returncode, out, r.stderr[-2000:])
    lines = {ln.split(" ", 1)[0]: ln for ln in out.splitlines()}
    assert lines["null-config"].split()[1] != "0" and "null argument" in lines["null-config"]
    assert lines["destroy-null"].split()[1] == "0"
    assert lines["submit-null"].split()[1] != "0"
    assert lines["wait-null"].split()[1] != "0"
    assert "create-failed" in lines and "device" in lines["create-failed"]
    assert "alive" in lines
