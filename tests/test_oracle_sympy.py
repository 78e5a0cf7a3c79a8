"""Symbolic restatement of the force at n = 2 and 3 (BH + 1 or 2 stars).

The force is written here in sympy from the term table of the reference's
formulas (Newtonian N; BH 1PN A1; a1c prefix C0, pair sum C2 and its Q2
removal CQ; star terms B, S; the ac pair sum X and its exact self removal
XR; Limit-sticky9.c:39-422), independently of the oracle's code: sums over
stars include j = i, 1PN parts carry lam = 1/c^2.  It pins
  - the oracle's force values (to 1e-13 relative);
  - the complex-step derivatives used to check the GPU adjoint: every entry
    of the Jacobian-vector product d(u . a)/d(x, v, m0) is compared with
    the exact symbolic derivative (evaluated at 30 digits).
"""
import numpy as np
import pytest
import sympy as sp

import oracle_bind as ob

EPS, C = 0.05, 3.0


def dot(a, b):
    return sum(p * q for p, q in zip(a, b))


def vsum(*vs):
    return [sum(c) for c in zip(*vs)]


def scale(s, v):
    return [s * c for c in v]


def sym_force(X, V, M, eps, lam, plant=None):
    """a_p (lists of 3 sympy expressions) for particles 0..n-1, BH = 0.
    `plant` swaps in a deliberate mistake (for the test that shows the
    comparison catches one)."""
    n = len(M)
    m0 = M[0]
    x0 = X[0]
    d = [[X[i][k] - x0[k] for k in range(3)] for i in range(n)]
    e = [1 / sp.sqrt(dot(d[i], d[i]) + eps**2) for i in range(n)]

    def rr(i, j):
        return [X[j][k] - X[i][k] for k in range(3)]

    def w(i, j):
        r = rr(i, j)
        return 1 / sp.sqrt(dot(r, r) + eps**2)

    a = []
    for i in range(n):                        # N: all j incl. BH and self
        acc = [0, 0, 0]
        for j in range(n):
            acc = vsum(acc, scale(M[j] * w(i, j) ** 3, rr(i, j)))
        a.append(acc)
    stars = range(1, n)
    # BH: A1 + a1c (C0 + C2 + CQ)
    pn0 = [0, 0, 0]
    for i in stars:
        vi, di = V[i], d[i]
        vd = dot(vi, di)
        br = 5 * m0 * e[i] - 2 * dot(vi, vi) + sp.Rational(3, 2) * vd**2 / dot(di, di)
        pn0 = vsum(pn0, scale(M[i] * e[i] ** 3 * br, di), scale(3 * M[i] * e[i] ** 3 * vd, vi))
        pn0 = vsum(pn0, scale(4 * M[i] ** 2 * e[i] ** 4, di))                          # C0
        for j in stars:                                                              # C2
            wij = w(i, j)
            vj, dj = V[j], d[j]
            br2 = (4 * e[j] + sp.Rational(5, 4) * wij - sp.Rational(1, 4) * dot(dj, dj) * wij**3
                   + sp.Rational(1, 4) * dot(di, di) * wij**3)
            t1 = scale(M[i] * e[i] ** 3 * M[j] * br2, di)
            c35 = sp.Rational(36, 10) if plant == "c2_coef" else sp.Rational(7, 2)
            t2 = scale(-c35 * wij**3 * e[i] * M[i] * M[j], rr(i, j))
            cv = M[i] * e[i] ** 3 * M[j] / m0
            t3 = scale(-cv, vsum(scale(4 * dot(vi, vj), di), scale(-3 * dot(vi, di), vj),
                                 scale(-4 * dot(vj, di), vi)))
            pn0 = vsum(pn0, t1, t2, t3)
        Vi = scale(M[i] ** 2 * e[i] ** 3 / m0, vsum(scale(4 * dot(vi, vi), di),
                                                     scale(-7 * dot(vi, di), vi)))
        pn0 = vsum(pn0, scale(M[i] ** 2 * e[i] ** 3 * (4 * e[i] + sp.Rational(5, 4) * eps), di),
                   scale(-1, Vi))                                                    # CQ
    a[0] = vsum(a[0], scale(lam, pn0))
    # stars: B + S + X - XR
    for i in stars:
        vi, di = V[i], d[i]
        b = scale(-1, di)
        pn = vsum(scale(m0 * e[i] ** 3 * (4 * m0 * e[i] - dot(vi, vi)), b),
                  scale(4 * m0 * e[i] ** 3 * dot(vi, b), vi))                       # B
        pn = vsum(pn, scale(5 * m0 * M[i] * e[i] ** 4, di),
                  scale(-M[i] * e[i] ** 3, vsum(scale(4 * dot(vi, vi), di),
                                                scale(-7 * dot(vi, di), vi))))       # S
        for j in stars:                                                              # X
            vj = V[j]
            r = rr(i, j)
            wij = w(i, j)
            u = [vj[k] - vi[k] for k in range(3)]
            cj = scale(-1, d[j])
            bb, r2 = dot(b, b), dot(r, r)
            x1 = scale(M[j] * m0 * e[i] ** 3 * (4 * wij + sp.Rational(5, 4) * e[j]
                                                 + sp.Rational(1, 4) * (bb - r2) * e[j] ** 3), b)
            x2 = scale(M[j] * m0 * wij**3 * (4 * e[i] + sp.Rational(5, 4) * e[j]
                                              + sp.Rational(1, 4) * (r2 - bb) * e[j] ** 3), r)
            x3 = scale(-sp.Rational(7, 2) * M[j] * m0 * e[j] ** 3 * (wij - e[i]), cj)
            x4 = scale(-M[j] * e[i] ** 3, vsum(scale(4 * dot(vi, vj), b), scale(-3 * dot(vj, b), vi),
                                               scale(-4 * dot(vi, b), vj)))
            x5 = scale(M[j] * wij**3 * (dot(vi, vi) - 2 * dot(u, u)
                                         + sp.Rational(3, 2) * dot(vj, r) ** 2 * wij**2), r)
            x6 = scale(M[j] * wij**3 * dot(r, [4 * vi[k] - 3 * vj[k] for k in range(3)]), u)
            if plant == "drop_x6":
                x6 = [0, 0, 0]
            if plant == "x4_sign":
                x4 = scale(-1, x4)
            pn = vsum(pn, x1, x2, x3, x4, x5, x6)
        bb = dot(b, b)
        xr = vsum(scale(M[i] * e[i] ** 3 * m0 * (4 / eps + sp.Rational(5, 4) * e[i]
                                                  + sp.Rational(1, 4) * bb * e[i] ** 3), b),
                  scale(-sp.Rational(7, 2) * M[i] * e[i] ** 3 * m0 * (1 / eps - e[i]), b),
                  scale(-M[i] * e[i] ** 3, vsum(scale(4 * dot(vi, vi), b),
                                                scale(-7 * dot(vi, b), vi))))       # XR
        a[i] = vsum(a[i], scale(lam, vsum(pn, scale(-1, xr))))
    return a


def state(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, (n, 3))
    v = rng.normal(0, 0.4, (n, 3))
    m = np.concatenate([[0.3], rng.uniform(0.05, 0.2, n - 1)])
    return x, v, m


@pytest.mark.parametrize("n", [2, 3])
def test_symbolic_force_matches_oracle_and_jacobians_match_complex_step(n):
    x, v, m = state(n, 40 + n)
    u = np.random.default_rng(50 + n).normal(size=(n, 3))
    Xs = [[sp.Symbol(f"x{p}_{k}") for k in range(3)] for p in range(n)]
    Vs = [[sp.Symbol(f"v{p}_{k}") for k in range(3)] for p in range(n)]
    m0s = sp.Symbol("m0")
    M = [m0s] + [sp.Float(float(mm), 30) for mm in m[1:]]
    lam = sp.Float(1.0 / C**2, 30)
    eps = sp.Float(EPS, 30)
    a = sym_force(Xs, Vs, M, eps, lam)
    subs = {m0s: sp.Float(float(m[0]), 30)}
    for p in range(n):
        for k in range(3):
            subs[Xs[p][k]] = sp.Float(float(x[p, k]), 30)
            subs[Vs[p][k]] = sp.Float(float(v[p, k]), 30)
    # values against the oracle
    a_sym = np.array([[float(a[p][k].evalf(30, subs=subs)) for k in range(3)] for p in range(n)])
    a_orc = ob.orc_force(x, v, m, EPS, C)
    assert np.abs(a_sym - a_orc).max() <= 1e-13 * np.abs(a_orc).max(), (a_sym, a_orc)
    # d(u . a)/d(coordinate) against the complex-step oracle
    phi = sum(float(u[p, k]) * a[p][k] for p in range(n) for k in range(3))
    coords = [("m0",)] + [(kind, p, k) for p in range(n) for k in range(3) for kind in ("x", "v")]
    want = ob.cs_force_vjp(x, v, m, u, EPS, C, coords)
    got = []
    for co in coords:
        s = m0s if co[0] == "m0" else (Xs if co[0] == "x" else Vs)[co[1]][co[2]]
        got.append(float(sp.diff(phi, s).evalf(30, subs=subs)))
    got = np.array(got)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), (got, want)


@pytest.mark.parametrize("plant", ["c2_coef", "drop_x6", "x4_sign"])
def test_symbolic_comparison_catches_planted_mistakes(plant):
    """A wrong coefficient, a dropped term or a flipped sign in the symbolic
    restatement must break the value comparison with the oracle."""
    n = 3
    x, v, m = state(n, 40 + n)
    Xs = [[sp.Float(float(x[p, k]), 30) for k in range(3)] for p in range(n)]
    Vs = [[sp.Float(float(v[p, k]), 30) for k in range(3)] for p in range(n)]
    M = [sp.Float(float(mm), 30) for mm in m]
    a = sym_force(Xs, Vs, M, sp.Float(EPS, 30), sp.Float(1.0 / C**2, 30), plant=plant)
    a_sym = np.array([[float(a[p][k].evalf(30)) for k in range(3)] for p in range(n)])
    a_orc = ob.orc_force(x, v, m, EPS, C)
    assert np.abs(a_sym - a_orc).max() > 1e-9 * np.abs(a_orc).max()
