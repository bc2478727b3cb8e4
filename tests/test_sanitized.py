"""The oracle and the host preprocessing under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

Builds oracle/liborc_san.so and tests/sanitize/_build/libpreprocess_san.so
(gcc, -fsanitize=address,undefined, no recovery), then runs, in one child
Python process that preloads the gcc ASan runtime:
  - the oracle's golden and mathematical tests over the sanitizer build
    (ORC_LIB selects it): every branch the fixtures reach, bit for bit;
  - tests/sanitize/san_preprocess.py: split stencils, alias order and both
    sweep-plan drivers on every fixture.
Any sanitizer report aborts the child, and the test fails with its output.
"""

import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

TESTS = os.path.join(ROOT, "tests")


def _asan_runtime():
    gcc = shutil.which("gcc")
    if not gcc:
        return None
    path = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_and_preprocessing_under_sanitizers():
    rt = _asan_runtime()
    if rt is None or shutil.which("make") is None:
        pytest.skip("gcc's ASan runtime or make is not available")
    for d, target in ((os.path.join(ROOT, "oracle"), "liborc_san.so"), (os.path.join(TESTS, "sanitize"), "all")):
        r = subprocess.run(["make", "-s", "-C", d, target], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ)
    env.update(LD_PRELOAD=rt, ORC_LIB=os.path.join(ROOT, "oracle", "liborc_san.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-p", "no:xdist",
           os.path.join(TESTS, "test_oracle_golden.py"), os.path.join(TESTS, "test_oracle_math.py"),
           os.path.join(TESTS, "sanitize", "san_preprocess.py")]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, out
    assert "passed" in r.stdout and "failed" not in r.stdout, out
