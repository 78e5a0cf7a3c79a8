"""The host-side helpers (pn_amd/csrc/pn1_host.c) under AddressSanitizer
and UndefinedBehaviorSanitizer: tests/host_sanitize_check.c drives input
validation, the IC generators, the .dat reader (good, short and missing
files), snapshots and checkpoints (round trip, too small a capacity, a
truncated file).  No device is involved."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_host_helpers_clean_under_asan_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "host_check")
    r = subprocess.run([gcc, "-O1", "-g", "-std=gnu11", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                        "-I", os.path.join(ROOT, "include"),
                        os.path.join(HERE, "host_sanitize_check.c"),
                        os.path.join(ROOT, "pn_amd", "csrc", "pn1_host.c"), "-lm", "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower():
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
