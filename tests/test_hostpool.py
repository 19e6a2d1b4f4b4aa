"""The host thread pool that packs and unpacks large calls (sapporo2_amd/csrc/hostpool.h),
built on the CPU with ThreadSanitizer: every element of every job is visited exactly once,
with spinning and with sleeping workers, with two caller threads at once, and in a forked
child."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "hostpool_check.cpp")
INC = os.path.join(ROOT, "sapporo2_amd", "csrc")


@pytest.fixture(scope="module")
def harness():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "hostpool_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-I", INC, SRC, "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("threads,spin,mode", [("4", "300", ""), ("4", "0", ""), ("1", "300", ""),
                                               ("8", "50", ""), ("4", "300", "fork")])
def test_hostpool_visits_each_element_once(harness, threads, spin, mode):
    env = dict(os.environ, SAPPORO2_AMD_HOST_THREADS=threads, SAPPORO2_AMD_HOST_SPIN_US=spin,
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([harness, "150", mode], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "ok" in r.stdout and "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
