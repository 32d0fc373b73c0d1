"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
LeakSanitizer (CPU): the FP64 oracle's round trips (every switch
combination's path) and the CLI's .dat/.bin writer and parameter dump, built
by tests/native/Makefile with gcc/g++ (no GPU code)."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

NATIVE = os.path.join(ROOT, "tests", "native")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not available")
    r = subprocess.run(["make", "-C", NATIVE], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return os.path.join(NATIVE, "_build")


def run_clean(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=ENV)
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-3000:]
    assert "runtime error" not in text and "Sanitizer" not in text, text[-3000:]
    return r.stdout


def test_oracle_round_trips_are_sanitizer_clean(built):
    out = run_clean([os.path.join(built, "oracle_check")])
    assert out.count(": ok") == 4


def test_cli_writer_is_sanitizer_clean(built, tmp_path):
    # the number formatting against printf, 15 writer cases (3 formats x 5
    # sizes) and 3 of the overlapped fetch/write helper (no failure, failure
    # at the first and third item)
    out = run_clean([os.path.join(built, "cli_writer_check"), str(tmp_path)])
    assert "format_g20 against printf: ok" in out
    assert out.count(": ok") == 19
