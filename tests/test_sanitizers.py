"""The C oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(plus float-to-integer overflow checks) and run over random small scenes and
edge cases: non-cubic 1-7 voxel volumes, plateaus of zero normals, NaN and
infinite LUT entries, unbound samplers, cameras inside the box, zero-length
light vectors, pixel sub-ranges, both precisions, 1-3 OpenMP threads
(oracle/sanitize_main.c, built by `make -C oracle sanitize`).  Any
sanitizer report fails the run."""

import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    # verify_asan_link_order=0: the environment may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([os.path.join(ORACLE, "_san", "ovr_sanitize"), "1500"], capture_output=True, text=True,
                         env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr, res.stderr
    assert "1500 scenes ok" in res.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_layout_and_brick_rules_under_sanitizers(tmp_path):
    """The texel-layout, transparent-brick, LUT-prefix, matrix-inverse and
    cost-ordered-dispatch rules the kernels and the host side use
    (pyvr_amd/csrc/hvr_rules.h),
    compiled natively with the sanitizers and brute-forced by
    tests/native/rules_check.cpp."""
    root = os.path.dirname(ORACLE)
    exe = tmp_path / "rules_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
                    "-ffp-contract=off", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "pyvr_amd", "csrc"),
                    os.path.join(root, "tests", "native", "rules_check.cpp"), "-o", str(exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "runtime error" not in res.stderr and "rules ok" in res.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_frame_copy_pool_under_thread_sanitizer(tmp_path):
    """The worker pool that copies whole frames out of pinned memory
    (pyvr_amd/csrc/hvr_copy.h), compiled natively with ThreadSanitizer and
    checked against memcpy by tests/native/copy_check.cpp."""
    root = os.path.dirname(ORACLE)
    exe = tmp_path / "copy_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(root, "pyvr_amd", "csrc"),
                    os.path.join(root, "tests", "native", "copy_check.cpp"), "-o", str(exe), "-pthread"], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ThreadSanitizer" not in res.stderr and "copy ok" in res.stdout
