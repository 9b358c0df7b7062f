"""CPU emulation of libddafft under AddressSanitizer (tests/emul): every
kernel and every host-side launch sequence of the library runs on host
threads, checked against the oracle, so an out-of-bounds or null access is a
CPU error report instead of a GPU fault.  It takes several minutes, so it runs
only when DDAFFT_EMUL=1 is set (the pre-flight before GPU runs)."""

import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.skipif(os.environ.get("DDAFFT_EMUL") != "1", reason="set DDAFFT_EMUL=1 to run")


def test_emulated_library_under_asan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "emul"), "-s"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan + " " + ubsan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emul", "run_emul.py"), "quick"], env=env,
                       capture_output=True, text=True, timeout=1800)
    print(r.stdout[-4000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "ALL OK" in r.stdout
