"""AddressSanitizer runs of the host code, on the CPU.

  * the oracle (C) built with gcc -fsanitize=address (oracle/Makefile `asan`)
    runs the oracle's own tests (bandpass, walk, dedup, percentile,
    characterisation, adjoint, raw signal, the C linker, threads) under gcc's
    libasan;
  * the library's host code (parameters, column names, TIFF/ETS decoding and
    its error paths, the C ABI's argument checks) built host-only with
    -fsanitize=address runs the ABI and ingest tests under clang's ASan
    runtime (no device is used: the device code is not in that build).
Any AddressSanitizer report fails the run.  Leaks are not checked (the
Python interpreter itself keeps memory until exit)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pytest_under(preload: str, env_extra: dict, files: list) -> None:
    env = dict(os.environ)
    env["LD_PRELOAD"] = preload + ((":" + env["LD_PRELOAD"]) if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env.update(env_extra)
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
           *[os.path.join(ROOT, "tests", f) for f in files]]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900, cwd=ROOT)
    out = r.stdout[-4000:] + r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr, out
    assert r.returncode == 0, out


def test_oracle_under_asan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    lib = os.path.join(ROOT, "oracle", "_build_asan", "liboracle_asan.so")
    rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                        text=True, check=True).stdout.strip()
    if not os.path.isabs(rt):
        pytest.skip("gcc's libasan is not installed")
    _pytest_under(rt, {"GT_TEST_ORACLE_LIB": lib},
                  ["test_oracle.py", "test_oracle_char.py", "test_adjoint_oracle.py",
                   "test_raw_oracle.py", "test_link_oracle.py"])


def test_host_library_under_asan():
    from tracking_optimizations_amd import _build
    try:
        rt = _build.asan_runtime()
    except FileNotFoundError:
        pytest.skip("clang's ASan runtime is not installed")
    lib = _build.build_asan()
    _pytest_under(rt, {"GT_TEST_LIB": lib}, ["test_abi.py", "test_ingest.py"])
