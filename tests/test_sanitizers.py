"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the CPU code: the oracle
(oracle/mpf_oracle.c) driven through every path the tests use by tests/sanitize_driver.c, and the
output writers of include/openphase_amd/io.h through tests/io_driver.cpp on the golden states, whose
bytes must still equal the reference's. Host code only: the device code is not built with
sanitizers."""
import os
import subprocess
from pathlib import Path

import pytest

from _formats import build_driver, write_files
from _golden import IO_CASES
from test_io_formats import IOCase, _check
import oracle as O

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


def _env():
    env = dict(os.environ)
    # an LD_PRELOAD library ahead of the sanitizer runtime is not an error for these runs
    env["ASAN_OPTIONS"] = "verify_asan_link_order=0:detect_leaks=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    env["OMP_NUM_THREADS"] = "2"
    return env


def _build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and ("asan" in r.stderr or "ubsan" in r.stderr):
        pytest.skip("no sanitizer runtime in this toolchain: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr


def test_oracle_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "san_driver"
    _build(["gcc", "-std=c11", "-O1", "-fopenmp", *SAN, "-I", str(ROOT / "oracle"), str(ROOT / "tests" / "sanitize_driver.c"),
            str(ROOT / "oracle" / "mpf_oracle.c"), "-lm", "-o", str(exe)])
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=_env(), timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert "sanitize driver ok" in r.stdout


@pytest.mark.parametrize("name", IO_CASES)
def test_writers_clean_under_sanitizers(name, tmp_path):
    try:
        exe = build_driver(tmp_path, SAN, "io_driver_san")
    except subprocess.CalledProcessError as err:
        pytest.skip(f"no sanitizer build: {err}")
    case = IOCase(name)
    orc = O.Oracle(case.cfg)
    lists, flags, rec, vol = case.run(orc)
    orc.close()
    got = write_files(tmp_path, case.cfg, lists, flags, case.phase, case.quat, rec, vol, exe=exe, env=_env())
    _check(got, case.expect, name + " (sanitized writers)")
