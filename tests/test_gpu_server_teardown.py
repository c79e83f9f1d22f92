"""A server whose creation fails after its finisher thread has started (the
overlapped form: finisher running, worker not yet started) is torn down by
its destructor without aborting the process.  The failure is forced from the
host by TSVD_FAIL_AFTER_FINISHER in a child process, so that an abort would be
reported as a failure rather than end the test run; the same child then
creates a working server, which shows the device and library are usable
after the teardown."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

CHILD = r"""
import os, sys
import numpy as np
sys.path.insert(0, {pkg!r})
from treesvd_amd import capi
capi.lib()
os.environ["TSVD_FAIL_AFTER_FINISHER"] = "1"
try:
    capi.Server(100, False, np.arange(8), 0.2, 1e-6, 4, 2, 4)
    print("created", flush=True)
except capi.TreeSVDError as e:
    print("create-failed", e, flush=True)
del os.environ["TSVD_FAIL_AFTER_FINISHER"]
s = capi.Server(100, False, np.arange(8), 0.2, 1e-6, 4, 2, 4)
s.close()
print("alive", flush=True)
"""


def test_server_torn_down_after_finisher_started():
    code = CHILD.format(pkg=os.path.join(ROOT, "tree-embedding_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    out = r.stdout
    assert r.returncode == 0, (r.returncode, out, r.stderr[-2000:])
    assert "create-failed" in out and "after its finisher started" in out, out
    assert "alive" in out, out
