"""The single-pass pair adjoint, checked on the host.

tests/adj_rev_check.hip compiles the device functions of
pn_amd/csrc/pn1_adj_rev.hpp for the CPU (hipcc, host pass) and compares the
hand-written reverse sweep of one pair with forward-mode duals of the
term-by-term pair function adj_pair_phi, on 2000 random pairs (close pairs,
stars near the BH, c = 1 and 100).  The GPU tests then pin the whole VJP
against the complex-step oracle.  Measured: 2.5e-15 relative.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    return None


def test_reverse_sweep_matches_duals(tmp_path):
    cc = hipcc()
    if cc is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "adj_rev_check")
    subprocess.run([cc, "-O2", "-std=c++17", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "pn_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "adj_rev_check.hip"), "-o", exe],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([exe, "2000"], check=True, capture_output=True, text=True).stdout
    vals = dict(line.split() for line in out.strip().splitlines() if line.split()[0] in
                ("worst", "masked_nonzero"))
    assert int(vals["masked_nonzero"]) == 0, out
    assert float(vals["worst"]) <= 1e-13, out
