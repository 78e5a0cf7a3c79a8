"""Host-side parts of libpn1 (no GPU): exports, validation, initial
conditions, file formats, and the failure paths of the drop-in shapes."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob
from conftest import GOLDEN, ROOT
from pn_amd import lib as pn

HEADER = os.path.join(ROOT, "include", "pn1.h")
STICKY9 = os.path.join(ROOT, "pn_amd", "sticky9")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|char\s*\*|const char\s*\*)\s*\*?\s*(\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert {"force", "energy", "initial_energy", "pn1_create", "pn1_step"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", pn.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = pn.load()
    for n in names:
        assert hasattr(L, n)


def test_library_is_gfx950_hip_code():
    """libpn1.so embeds a gfx950 code object (our own build)."""
    blob = open(pn.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def _case(n=17, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    v = rng.normal(0, 0.3, (n, 3))
    m = np.full(n, 1.0 / (n - 1))
    m[0] = 0.05
    return x, v, m


@pytest.mark.parametrize("mutate,needle", [
    (lambda x, v, m: (x, v, m, 0.0), "eps"),
    (lambda x, v, m: (x, v, m, -1e-3), "eps"),
    (lambda x, v, m: (x, v, np.where(np.arange(len(m)) == 0, 0.0, m), 1e-3), "black hole mass"),
    (lambda x, v, m: (np.where(np.arange(len(m))[:, None] == 3, x[0], x), v, m, 1e-3), "exactly on"),
    (lambda x, v, m: (np.where(np.arange(len(m))[:, None] == 2, np.nan, x), v, m, 1e-3), "non-finite"),
    (lambda x, v, m: (x, np.where(np.arange(len(m))[:, None] == 2, np.inf, v), m, 1e-3), "non-finite"),
])
def test_check_input_rejects(mutate, needle):
    x, v, m, eps = mutate(*_case())
    rc, msg = pn.check_input(x, v, m, eps)
    assert rc == pn.PN1_EINVAL and needle in msg, msg


def test_check_input_accepts():
    x, v, m = _case()
    assert pn.check_input(x, v, m, 1e-3) == (pn.PN1_OK, "")


def test_uniform_sphere_ic():
    x, v, m = pn.ic_uniform_sphere(2048, 2, R=1.0, Q=0.5, m_bh=0.01, eps=1e-3)
    x2, v2, m2 = pn.ic_uniform_sphere(2048, 2, R=1.0, Q=0.5, m_bh=0.01, eps=1e-3)
    assert np.array_equal(x, x2) and np.array_equal(v, v2)
    x3, _, _ = pn.ic_uniform_sphere(2048, 3, R=1.0, Q=0.5, m_bh=0.01, eps=1e-3)
    assert not np.array_equal(x, x3)
    # FP32-exact values (the reference reads its input as float), BH at rest at 0
    assert np.array_equal(x, x.astype(np.float32).astype(np.float64))
    assert np.array_equal(v, v.astype(np.float32).astype(np.float64))
    assert np.all(x[0] == 0) and np.all(v[0] == 0) and m[0] == 0.01
    assert np.allclose(m[1:], 1.0 / 2048, rtol=0, atol=0)
    r = np.linalg.norm(x[1:], axis=1)
    # sampled in r <= R, then shifted by the star centre of mass (~R/sqrt(N))
    assert r.max() <= 1.0 + 5.0 / np.sqrt(2048)
    assert np.abs(x[1:].mean(0)).max() < 1e-6 and np.abs(v[1:].mean(0)).max() < 1e-6
    # virial ratio K/|W| with W the softened potential including the BH
    e = ob.orc_energy_terms(x, v, m, 1e-3, 1e12)
    assert abs(e[0] / abs(e[1]) - 0.5) < 1e-5


def test_plummer_ic():
    x, v, m = pn.ic_plummer(20000, 3, a=1.0, m_bh=0.05)
    r = np.linalg.norm(x[1:], axis=1)
    rmax = (0.999 ** (-2.0 / 3.0) - 1.0) ** -0.5
    assert r.max() <= rmax * (1 + 1e-6)
    # median radius of a Plummer sphere of scale radius 1: (2^(2/3) - 1)^(-1/2)
    assert abs(np.median(r) / (2 ** (2 / 3) - 1) ** -0.5 - 1.0) < 0.03
    vesc = np.sqrt(2.0) * (1 + r * r) ** -0.25
    assert np.all(np.linalg.norm(v[1:], axis=1) <= vesc * (1 + 1e-6) + 1e-6)
    assert np.array_equal(x, x.astype(np.float32).astype(np.float64))


def test_read_dat_matches_float_parse():
    path = os.path.join(GOLDEN, "init10k_head.dat")
    lines = open(path).read().split("\n")
    rows = np.array([[np.float32(t) for t in ln.split()] for ln in lines if ln.strip()],
                    dtype=np.float64)
    n = rows.shape[0] // 2 + 1
    x, v = pn.read_dat(path, n)
    assert np.all(x[0] == 0) and np.all(v[0] == 0)
    assert np.array_equal(x[1:], rows[: n - 1]) and np.array_equal(v[1:], rows[n - 1:])


def test_read_dat_errors(tmp_path):
    with pytest.raises(pn.Pn1Error) as ei:
        pn.read_dat(str(tmp_path / "missing.dat"), 5)
    assert ei.value.code == pn.PN1_EIO
    short = tmp_path / "short.dat"
    short.write_text("1 2 3 \n4 5 6 \n")
    with pytest.raises(pn.Pn1Error) as ei:
        pn.read_dat(str(short), 3)          # needs 2 + 2 lines
    assert ei.value.code == pn.PN1_EIO


def test_snapshot_format(tmp_path):
    x, v, m = _case(5)
    p = tmp_path / "snap.dat"
    pn.write_snapshot(str(p), x, v)
    pn.write_snapshot(str(p), x, v, append=True)
    want = "".join("%f %f %f \n" % tuple(r) for r in np.vstack([x, v]))
    assert p.read_text() == want * 2


def test_checkpoint_round_trip(tmp_path):
    x, v, m = _case(33)
    p = tmp_path / "state.ckpt"
    pn.write_checkpoint(str(p), x, v, m, t=0.125, step=128, eps=1e-4, c=50.0, dt=2.0**-12)
    d = pn.read_checkpoint(str(p))
    assert np.array_equal(d["x"], x) and np.array_equal(d["v"], v) and np.array_equal(d["m"], m)
    assert (d["t"], d["step"], d["eps"], d["c"], d["dt"]) == (0.125, 128, 1e-4, 50.0, 2.0**-12)
    (tmp_path / "bad.ckpt").write_bytes(b"not a checkpoint")
    with pytest.raises(pn.Pn1Error):
        pn.read_checkpoint(str(tmp_path / "bad.ckpt"))


def _run_py(code):
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          cwd=ROOT, timeout=120)


def test_legacy_force_rejects_bad_input_before_device():
    """Invalid input fails loudly (stderr + exit 1) like the reference's
    own failures, before any device work."""
    code = ("import numpy as np; from pn_amd import lib as pn\n"
            "x = np.zeros((3, 3)); v = np.zeros((3, 3)); m = np.array([0.0, 0.5, 0.5])\n"
            "x[1, 0] = 1; x[2, 1] = 1\n"
            "pn.legacy_force(x, v, m, 1e-3)\n")
    r = _run_py(code)
    assert r.returncode == 1 and "force: black hole mass" in r.stderr, r.stderr


def _has_gpu():
    try:
        return os.path.exists("/dev/kfd")
    except OSError:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_no_device_fails_loudly():
    with pytest.raises(pn.Pn1Error) as ei:
        pn.Context(16)
    assert ei.value.code == pn.PN1_ENODEV
    code = ("import numpy as np; from pn_amd import lib as pn\n"
            "x = np.random.default_rng(0).uniform(-1, 1, (4, 3)); v = np.zeros((4, 3))\n"
            "m = np.full(4, 0.25)\n"
            "pn.legacy_force(x, v, m, 1e-3)\n")
    r = _run_py(code)
    assert r.returncode == 1 and "gfx950 device" in r.stderr, r.stderr


def test_capacity_bounds_rejected_before_device():
    """Capacities outside [1, PN1_MAX_CAPACITY] are refused with EINVAL before
    any device is touched (the kernels' 32-bit state offsets stop there)."""
    for cap in (0, -5, pn.PN1_MAX_CAPACITY + 1, 2**31 - 1):
        with pytest.raises(pn.Pn1Error) as ei:
            pn.Context(cap)
        assert ei.value.code == pn.PN1_EINVAL, cap
        with pytest.raises(pn.Pn1Error) as ei:
            pn.Ensemble(cap, 4)
        assert ei.value.code == pn.PN1_EINVAL, cap
    assert 6 * pn.PN1_MAX_CAPACITY < 2**31


def test_sticky9_missing_input_and_usage(tmp_path):
    r = subprocess.run([STICKY9, "--in", str(tmp_path / "none.dat"), "--out",
                        str(tmp_path / "o.dat")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stderr.startswith("data_input:")
    r = subprocess.run([STICKY9, "--no-such-option"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_status_codes_match_header():
    text = open(HEADER).read()
    for name, val in (("PN1_OK", pn.PN1_OK), ("PN1_EINVAL", pn.PN1_EINVAL),
                      ("PN1_ENOMEM", pn.PN1_ENOMEM), ("PN1_EDEVICE", pn.PN1_EDEVICE),
                      ("PN1_ENODEV", pn.PN1_ENODEV), ("PN1_EIO", pn.PN1_EIO),
                      ("PN1_FP64", pn.FP64), ("PN1_MIXED", pn.MIXED)):
        assert int(re.search(rf"#define {name}\s+\(?(-?\d+)\)?", text).group(1)) == val
    sh = re.search(r"#define PN1_MAX_CAPACITY\s+\(1 << (\d+)\)", text)
    assert 1 << int(sh.group(1)) == pn.PN1_MAX_CAPACITY
