"""The oracle against the golden vectors written from the compiled reference."""
import numpy as np
import pytest

import oracle_bind as ob
from conftest import c_values, energy_atol, load_fixture, rel_err


def test_force_c1_bitwise(force_fixture):
    d = force_fixture
    a = ob.orc_force(d["x"], d["v"], d["m"], float(d["eps"]), 1.0)
    assert np.array_equal(a.view(np.uint64), d["a_c1"].view(np.uint64)), d["name"]


def test_force_rows_bitwise(force_fixture):
    """orc_force_rows (the O(n)-per-star rows the n = 270000 GPU test checks
    against) equals the golden vectors' star rows bit for bit at c = 1 and
    orc_force's at the fixture's other c, rows taken in a shuffled order."""
    d = force_fixture
    n = d["m"].shape[0]
    if n < 2:
        pytest.skip("no star rows")
    rows = np.random.default_rng(n).permutation(np.arange(1, n))
    for c in [1.0] + c_values(d):
        full = d["a_c1"] if c == 1.0 else ob.orc_force(d["x"], d["v"], d["m"],
                                                       float(d["eps"]), c)
        got = ob.orc_force_rows(d["x"], d["v"], d["m"], float(d["eps"]), c, rows)
        assert np.array_equal(got.view(np.uint64), full[rows].view(np.uint64)), (d["name"], c)


@pytest.mark.parametrize("name", ["rand_n257", "c1_n1025", "close_n65"])
def test_force_omp_matches_serial(name):
    """The OpenMP force (the multi-core CPU baseline): star rows bitwise the
    golden vectors', the BH's row to the reordering of its a1c pair sum."""
    d = load_fixture(name)
    for c in [1.0] + c_values(d):
        full = d["a_c1"] if c == 1.0 else ob.orc_force(d["x"], d["v"], d["m"],
                                                       float(d["eps"]), c)
        a, ran = ob.orc_force_omp(d["x"], d["v"], d["m"], float(d["eps"]), c, 4)
        assert ran >= 1
        assert np.array_equal(a[1:].view(np.uint64), full[1:].view(np.uint64)), (name, c)
        assert np.abs(a[0] - full[0]).max() <= 1e-12 * np.abs(full[0]).max(), (name, c)


def test_initial_energy_c1_bitwise_and_text(force_fixture):
    d = force_fixture
    E0, txt = ob.orc_initial_energy_text(d["x"], d["v"], d["m"], float(d["eps"]), 1.0)
    assert E0 == float(d["E0_c1"])
    assert txt == str(d["init_txt_c1"])


def test_energy_text_c1(force_fixture):
    d = force_fixture
    txt = ob.orc_energy_text(0.25, d["x"], d["v"], d["m"], float(d["E0_c1"]) * (1.0 + 1e-6),
                             float(d["eps"]), 1.0)
    assert txt == str(d["energy_txt_c1"])


def test_force_speed_of_light(force_fixture):
    """At c != 1 the reference is run on v/c, m/c^2 and rescaled; that adds a
    few roundings per term, so agreement is to ~1e-13 rather than bitwise."""
    d = force_fixture
    for c in c_values(d):
        a = ob.orc_force(d["x"], d["v"], d["m"], float(d["eps"]), float(c))
        err = rel_err(a, d[f"a_c{c}"])
        assert err.max() < 1e-12, (d["name"], c, err.max())


def test_energy_speed_of_light(force_fixture):
    d = force_fixture
    for c in c_values(d):
        e = ob.orc_energy_terms(d["x"], d["v"], d["m"], float(d["eps"]), float(c))
        ref = d[f"E_c{c}"]           # kin, pot, pot_pn, total
        mine = np.array([e[0], e[1], e[2], e[4]])
        assert np.all(np.abs(mine - ref) <= energy_atol(d, ref)), (d["name"], c)


@pytest.mark.parametrize("name,tol", [("traj_rand_n65_c1", 0.0), ("traj_c1_n1025_c100", 1e-12)])
def test_q1_trajectory(name, tol):
    """The oracle's stage driver against a stage loop around the reference's
    force(): bitwise at c = 1 (same force, same stage arithmetic)."""
    d = load_fixture(name)
    x, v = ob.orc_q1_steps(d["x"], d["v"], d["m"], float(d["eps"]), float(d["c"]),
                           float(d["dt"]), int(d["steps"]))
    if tol == 0.0:
        assert np.array_equal(x, d["xK"]) and np.array_equal(v, d["vK"])
    else:
        assert np.abs(x - d["xK"]).max() <= tol * np.abs(d["xK"]).max()
        assert np.abs(v - d["vK"]).max() <= tol * np.abs(d["vK"]).max()
