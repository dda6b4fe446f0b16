"""The host-side code under AddressSanitizer and UndefinedBehaviorSanitizer: the
host tables (Background, RECFAST, ionization history, their Dual instantiations,
the threaded batch, argument checks) and the CPU oracle (RHS, Jacobian, initial
conditions, adaptive and replayed source columns, plin, C_l with exact and spline
j_l, the momentum-bin bound).  Built with g++/gcc from tests/sanitize/Makefile;
any report aborts the driver.  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def test_host_code_is_clean_under_asan_and_ubsan():
    if not shutil.which("g++"):
        pytest.skip("no g++")
    subprocess.check_call(["make", "-s", "-C", HERE], timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(HERE, "_build", "driver")], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitized host run: ok" in r.stdout
