"""The continuous-batching core (sprout_amd/csrc/pool_core.h) on the host, driven by a
backend that only records its calls (tests/pool_core_check.cpp): requests from four threads,
a serving loop and two waiters that drive rounds themselves, budgets from one scene to all of
them, one to four streams. Every request completes; each scene's requests run in arrival
order, each in consecutive rounds for exactly its substeps; launches hold only resident
active scenes of one key; the resident scenes fit the budget after every restore; a refused
request fails alone. The same check runs under ThreadSanitizer when ROCm's clang is present."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "pool_core_check.cpp")
CASES = [(1, 24, 10, 3, 6), (2, 40, 8, 1, 4), (3, 64, 10, 4, 1000), (4, 30, 12, 2, 5), (5, 8, 40, 3, 4),
         (6, 5, 30, 2, 4)]


def build(tmp_path, cxx, extra):
    exe = os.path.join(tmp_path, "check")
    subprocess.run([cxx, "-std=c++17", "-pthread", *extra, "-o", exe, SRC], check=True)
    return exe


@pytest.mark.parametrize("case", CASES)
def test_rounds_complete_in_order_within_budget(tmp_path, case):
    exe = build(str(tmp_path), "g++", ["-O2", "-Wall"])
    r = subprocess.run([exe, *map(str, case)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("ok")


CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG), reason="no clang with ThreadSanitizer here")
def test_rounds_have_no_data_race(tmp_path):
    exe = build(str(tmp_path), CLANG, ["-O1", "-g", "-fsanitize=thread"])
    for case in CASES[:3]:
        r = subprocess.run([exe, *map(str, case)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
