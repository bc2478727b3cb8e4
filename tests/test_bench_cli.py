"""bench.py's command line on the CPU: the one-GPU contract.

bench.py times the iteration on one GPU; asked for more it must refuse with
a message and a nonzero status before it touches a device.
"""

import os
import subprocess
import sys

from conftest import ROOT


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""
