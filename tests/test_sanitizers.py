"""CPU builds under sanitizers: the oracle (every function, degenerate and
small shapes) with AddressSanitizer + UndefinedBehaviorSanitizer, and the
stage layer's bounded-buffer pipes with ThreadSanitizer and with
AddressSanitizer. Host code only, built with gcc/g++ (no device code)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
ASAN_UBSAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
              "-fno-omit-frame-pointer", "-g", "-O1"]


def _build_run(tmp_path, compiler, sources, flags, includes, out_name, expect):
    if shutil.which(compiler) is None:
        pytest.skip("%s not available" % compiler)
    exe = str(tmp_path / out_name)
    cmd = [compiler] + flags + ["-I" + i for i in includes] + sources + ["-o", exe, "-lm",
                                                                           "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0 and expect in r.stdout, r.stdout + r.stderr


def test_oracle_under_asan_ubsan(tmp_path):
    srcs = [os.path.join(SAN, "oracle_driver.c"), os.path.join(ROOT, "oracle", "tdvref.c"),
            os.path.join(ROOT, "oracle", "tdvref_io.c")]
    _build_run(tmp_path, "gcc", srcs, ["-std=c11", "-ffp-contract=off"] + ASAN_UBSAN,
               [os.path.join(ROOT, "oracle")], "oracle_asan", "oracle sanitize ok")


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_pipes_under_sanitizers(tmp_path, san):
    flags = ["-std=c++17", "-fsanitize=" + san, "-fno-sanitize-recover=all", "-g", "-O1"]
    _build_run(tmp_path, "g++", [os.path.join(SAN, "pipes_driver.cpp")], flags,
               [os.path.join(ROOT, "include")], "pipes_" + san.split(",")[0],
               "pipes sanitize ok")
