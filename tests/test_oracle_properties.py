"""Pins of the oracle that do not come from the reference binary: closed
forms at n = 1 and n = 2, invariants of the equations, the exact 1/c^2
structure, and the one-rsqrt-per-pair decomposition used by the kernels."""
import numpy as np
import pytest

import hoisted_model as hm
import oracle_bind as ob
from conftest import rel_err


def rand_state(n, seed, eps=1e-3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    v = rng.normal(0, 0.3, (n, 3))
    m = rng.uniform(0.5, 1.5, n) / max(n - 1, 1)
    m[0] = 0.05
    x[0] = rng.uniform(-0.05, 0.05, 3)
    v[0] = rng.normal(0, 0.01, 3)
    return x, v, m, eps


# ---------------------------------------------------------------- closed forms
def test_black_hole_alone():
    """n = 1: no force, and each step of the stage scheme advances x by
    v dt / 2 (its stage weights sum to 1/2, Limit-sticky9.c:884-891)."""
    x = np.array([[0.25, -0.5, 1.0]])
    v = np.array([[0.125, 0.375, -0.0625]])
    m = np.array([0.05])
    assert np.array_equal(ob.orc_force(x, v, m, 1e-3, 1.0), np.zeros((1, 3)))
    dt, K = 2.0**-10, 16
    xk, vk = ob.orc_q1_steps(x, v, m, 1e-3, 1.0, dt, K)
    assert np.array_equal(vk, v)
    assert np.allclose(xk, x + K * v * dt / 2, rtol=0, atol=1e-15)
    e = ob.orc_energy_terms(x, v, m, 1e-3, 1.0)
    assert e[0] == 0.5 * m[0] * (v[0] @ v[0]) and e[1] == 0.0 and e[4] == e[0]


def two_body_closed_form(x, v, m, eps, c):
    """BH + one star from the term table (Newton, a1, a1c with its j = i
    term and the 'remove double-counting' term, the star's BH terms; the
    star's three-body sum holds only its own j = i term, which the
    reference removes exactly)."""
    lam = 1.0 / (c * c)
    m0, m1 = m
    d = x[1] - x[0]
    D2 = d @ d
    e = 1.0 / np.sqrt(D2 + eps * eps)
    E3 = e**3
    v1, v0 = v[1], v[0]
    V2 = v1 @ v1
    vd = v1 @ d
    # star
    newton1 = -m0 * E3 * d
    B = m0 * E3 * (-d) * (4 * m0 * e - V2) + 4 * m0 * E3 * (v1 @ -d) * v1
    S = 5 * m0 * m1 * e**4 * d - m1 * E3 * (4 * V2 * d - 7 * vd * v1)
    a1 = newton1 + lam * (B + S)
    # black hole
    newton0 = m1 * E3 * d
    A1 = m1 * E3 * d * (5 * m0 * e - 2 * V2 + 1.5 * vd**2 / D2) + 3 * m1 * E3 * vd * v1
    C0 = 4 * m1**2 * e**4 * d
    Vt = (m1**2 * E3 / m0) * (4 * V2 * d - 7 * vd * v1)
    C2self = m1**2 * E3 * d * (4 * e + 1.25 / eps) - Vt
    Q2 = m1**2 * E3 * d * (4 * e + 1.25 * eps) - Vt
    a0 = newton0 + lam * (A1 + C0 + C2self + Q2)
    # energy: the three-body term holds only its removed self pair
    kin = 0.5 * (m0 * (v0 @ v0) + m1 * V2)
    pot = -m0 * m1 * e
    potpn = (0.375 * m1 * V2**2 + 1.5 * m0 * e * m1 * V2 + 0.5 * m0**2 * e**2 * m1
             + 0.5 * m0 * e * m1 * (3 * V2 - 7 * (v0 @ v1) - (d @ v0) * vd * e**2))
    return np.vstack([a0, a1]), np.array([kin, pot, lam * potpn, 0.0, kin + pot + lam * potpn])


@pytest.mark.parametrize("c", [1.0, 10.0, 100.0])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_two_body_closed_form(seed, c):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (2, 3))
    v = rng.normal(0, 0.3, (2, 3))
    m = np.array([0.05, 0.7])
    eps = 0.05
    a_ref, e_ref = two_body_closed_form(x, v, m, eps, c)
    a = ob.orc_force(x, v, m, eps, c)
    assert rel_err(a, a_ref).max() < 1e-13
    e = ob.orc_energy_terms(x, v, m, eps, c)
    assert np.all(np.abs(e - e_ref) <= 1e-13 * np.abs(e_ref).max())


# ---------------------------------------------------------------- invariants
def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


@pytest.mark.parametrize("n", [17, 65])
def test_rotation_covariance(n):
    x, v, m, eps = rand_state(n, 5)
    R = rotation(n)
    a = ob.orc_force(x, v, m, eps, 10.0)
    ar = ob.orc_force(x @ R.T, v @ R.T, m, eps, 10.0)
    assert rel_err(ar, a @ R.T).max() < 1e-12
    e = ob.orc_energy_terms(x, v, m, eps, 10.0)
    er = ob.orc_energy_terms(x @ R.T, v @ R.T, m, eps, 10.0)
    assert np.abs(er - e).max() < 1e-12 * np.abs(e).max()


def test_translation_invariance():
    x, v, m, eps = rand_state(33, 6)
    shift = np.array([0.5, -0.25, 0.125])
    a = ob.orc_force(x, v, m, eps, 10.0)
    at = ob.orc_force(x + shift, v, m, eps, 10.0)
    assert rel_err(at, a).max() < 1e-12


def test_velocity_parity_is_exact():
    """Every velocity term is a product of two velocities, so flipping all
    velocities changes no bit of the forces or the energy."""
    x, v, m, eps = rand_state(33, 7)
    assert np.array_equal(ob.orc_force(x, -v, m, eps, 3.0), ob.orc_force(x, v, m, eps, 3.0))
    assert np.array_equal(ob.orc_energy_terms(x, -v, m, eps, 3.0),
                          ob.orc_energy_terms(x, v, m, eps, 3.0))


def test_star_permutation():
    x, v, m, eps = rand_state(33, 8)
    perm = np.concatenate([[0], 1 + np.random.default_rng(1).permutation(32)])
    a = ob.orc_force(x, v, m, eps, 10.0)
    ap = ob.orc_force(x[perm], v[perm], m[perm], eps, 10.0)
    assert rel_err(ap, a[perm]).max() < 1e-12


def test_newtonian_momentum():
    """With the 1PN terms switched off (c huge) sum_i m_i a_i = 0."""
    x, v, m, eps = rand_state(65, 9)
    a = ob.orc_force(x, v, m, eps, 1e30)
    p = (m[:, None] * a).sum(0)
    assert np.abs(p).max() < 1e-14 * (m[:, None] * np.abs(a)).sum()


def test_inverse_c_squared_structure():
    """a(c) = a_N + a_PN / c^2 and E(c) = E_N + E_PN / c^2 exactly: with c a
    power of two, 4 (a(2) - a_N) and 16 (a(4) - a_N) are the same a_PN."""
    x, v, m, eps = rand_state(33, 10)
    aN = ob.orc_force(x, v, m, eps, 2.0**60)
    p2 = 4.0 * (ob.orc_force(x, v, m, eps, 2.0) - aN)
    p4 = 16.0 * (ob.orc_force(x, v, m, eps, 4.0) - aN)
    assert rel_err(p4, p2).max() < 1e-12
    eN = ob.orc_energy_terms(x, v, m, eps, 2.0**60)
    e2 = 4.0 * (ob.orc_energy_terms(x, v, m, eps, 2.0) - eN)
    e4 = 16.0 * (ob.orc_energy_terms(x, v, m, eps, 4.0) - eN)
    assert np.abs(e4 - e2).max() < 1e-12 * np.abs(e2).max()


# ---------------------------------------------------------------- decomposition
@pytest.mark.parametrize("n,c", [(2, 1.0), (3, 100.0), (17, 1.0), (17, 50.0), (129, 100.0)])
def test_decomposition_matches_oracle(n, c):
    """The kernels' algebra (one rsqrt per pair, separable O(n) sums; see
    tests/hoisted_model.py) against the oracle."""
    x, v, m, eps = rand_state(n, 20 + n)
    a = ob.orc_force(x, v, m, eps, c)
    ah = hm.force(x, v, m, eps, c)
    tol = 1e-12 if c == 1.0 else 1e-13
    assert rel_err(ah, a).max() < tol
    e = ob.orc_energy_terms(x, v, m, eps, c)
    eh = hm.energy_terms(x, v, m, eps, c)
    scale = (m * m).sum() / eps
    assert np.all(np.abs(eh - e) <= 4e-15 * scale + 1e-13 * np.abs(e).max())


@pytest.mark.parametrize("seed,eps", [(40, 1e-3), (41, 1e-4), (42, 0.1)])
def test_xi_identity(seed, eps):
    """The BH's a1c needs Xi_i = sum' m_j w^3 (D2_i - D2_j); the kernels
    accumulate only W3 = sum' m_j w^3 and rebuild Xi as -2 d_i.N - phi +
    eps^2 W3.  The two agree to rounding, in 80-bit arithmetic on a system
    with stars far out (|d| ~ 30), one next to the BH and one close pair."""
    x, v, m, _ = rand_state(65, seed, eps)
    x[1] = x[0] + [1e-3, 0.0, 0.0]
    x[2] = [30.0, -4.0, 2.0]
    x[3] = [-25.0, 11.0, 7.0]
    x[5] = x[4] + [0.0, 2e-4, 0.0]
    L = np.longdouble
    xl, vl, ml = x.astype(L), v.astype(L), m.astype(L)
    S = hm.pair_sums(xl, vl, ml, L(eps))
    d, D2, e = hm.prep(xl, vl, ml, L(eps))
    xi = hm.xi_from_sums(d, S["Nst"], S["phi"], S["W3"], L(eps))
    # scale: the magnitudes that cancel on either side
    scale = (2 * np.linalg.norm(d, axis=1) * np.linalg.norm(S["Nst"], axis=1)
             + S["phi"] + np.abs(S["Xi_direct"]))
    assert (np.abs(xi - S["Xi_direct"]) / scale).astype(float).max() < 1e-16
    # and a planted mistake (the factor 2 dropped) fails it
    bad = -np.einsum("ij,ij->i", d, S["Nst"]) - S["phi"] + L(eps) ** 2 * S["W3"]
    assert (np.abs(bad - S["Xi_direct"]) / scale).astype(float).max() > 1e-3


def test_oracle_against_extended_precision():
    """The oracle (FP64, reference operation order) against the same
    formulas in 80-bit arithmetic: at c = 50 its stars are good to 1e-13."""
    x, v, m, eps = rand_state(65, 30)
    L = np.longdouble
    ex = hm.force(x.astype(L), v.astype(L), m.astype(L), L(eps), L(50.0))
    a = ob.orc_force(x, v, m, eps, 50.0)
    assert rel_err(a.astype(L), ex).astype(float).max() < 1e-13
