"""Host build of the library's lane bodies under AddressSanitizer + UndefinedBehaviorSanitizer
(tests/cpu_lanes/lanes_host.cpp): the fused filter + smoother pass with host loops and
sequential scans against the oracle's FP64 C filter + smoother, the per-series status bits
on forced failures (negative noise variance, NaN observation, non-PD smoother), and the
plan construction's refusals and slot coverage.  Host code only: g++ on the .cpp, gcc on
the oracle's C file, no GPU code is built."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def lanes_host(tmp_path_factory):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    d = tmp_path_factory.mktemp("lanes_host")
    obj = str(d / "kfrts.o")
    exe = str(d / "lanes_host")
    subprocess.run(["gcc", "-O1", "-g", "-std=c99"] + SAN + ["-c", "-o", obj, os.path.join(ROOT, "oracle", "kfrts.c")],
                   check=True)
    subprocess.run(["g++", "-O1", "-g", "-std=c++17"] + SAN +
                   ["-o", exe, os.path.join(ROOT, "tests", "cpu_lanes", "lanes_host.cpp"), obj, "-lm"], check=True)
    return exe


def test_lane_bodies_vs_oracle_and_status_bits_under_sanitizers(lanes_host):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([lanes_host], capture_output=True, text=True, timeout=600, env=env)
    print(res.stdout)
    print(res.stderr[-4000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "LANES_HOST_OK" in res.stdout
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr
