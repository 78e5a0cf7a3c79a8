"""Property-based checks (hypothesis) of the oracle's invariants over
randomly drawn small systems: rotation covariance, translation invariance,
exact velocity parity and star-permutation invariance of the force, and
rotation invariance of the energy.  These hold for the reference's
equations whatever the state, so a broken term shows up for some draw."""
import numpy as np
from hypothesis import given, settings, strategies as st

import oracle_bind as ob


@st.composite
def systems(draw):
    n = draw(st.integers(2, 9))
    seed = draw(st.integers(0, 2**31 - 1))
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    v = rng.normal(0, draw(st.floats(0.01, 2.0)), (n, 3))
    m = rng.uniform(0.05, 1.0, n)
    eps = draw(st.sampled_from([1e-3, 1e-2, 0.1]))
    c = draw(st.sampled_from([1.0, 3.0, 50.0]))
    # keep every star off the BH (the reference's 1/|d|^2 term is unsoftened)
    d = np.linalg.norm(x[1:] - x[0], axis=1)
    x[1:][d < 1e-3] += 0.01
    return x, v, m, eps, c, rng


# the same 40 draws on every run (no example database): a fresh machine
# checks what earlier runs checked
SET = settings(max_examples=40, deadline=None, derandomize=True, database=None)


def rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def ok(a, b):
    """Per-particle relative agreement, with an absolute floor for particles
    whose acceleration nearly cancels."""
    scale = np.abs(b).max()
    return (np.linalg.norm(a - b, axis=1) <= 1e-11 * np.maximum(np.linalg.norm(b, axis=1),
                                                                  1e-3 * scale)).all()


@SET
@given(systems())
def test_rotation(sys_):
    x, v, m, eps, c, rng = sys_
    R = rotation(rng)
    a = ob.orc_force(x, v, m, eps, c)
    ar = ob.orc_force(x @ R.T, v @ R.T, m, eps, c)
    assert ok(ar, a @ R.T)
    e = ob.orc_energy_terms(x, v, m, eps, c)
    er = ob.orc_energy_terms(x @ R.T, v @ R.T, m, eps, c)
    assert np.abs(er - e).max() <= 1e-10 * max(1.0, np.abs(e).max())


@SET
@given(systems())
def test_translation(sys_):
    x, v, m, eps, c, rng = sys_
    shift = rng.uniform(-1, 1, 3)
    assert ok(ob.orc_force(x + shift, v, m, eps, c), ob.orc_force(x, v, m, eps, c))


@SET
@given(systems())
def test_velocity_parity(sys_):
    x, v, m, eps, c, _ = sys_
    assert np.array_equal(ob.orc_force(x, -v, m, eps, c), ob.orc_force(x, v, m, eps, c))


@SET
@given(systems())
def test_star_permutation(sys_):
    x, v, m, eps, c, rng = sys_
    n = len(m)
    perm = np.concatenate([[0], 1 + rng.permutation(n - 1)])
    a = ob.orc_force(x, v, m, eps, c)
    assert ok(ob.orc_force(x[perm], v[perm], m[perm], eps, c), a[perm])
