"""The complex-step oracle (oracle/ls9_oracle_cs.c), the checker of the
adjoint, checked on the CPU:
  - its real part is the FP64 oracle's force (same source, other type);
  - its derivatives agree with central finite differences of the FP64
    oracle (h = 1e-6 relative; FD truncation and rounding ~1e-8), for
    positions, velocities and the BH mass, for the force and for a short
    trajectory of the stage driver;
  - the trajectory loss's host gradient agrees with finite differences.
Finite differences are an independent check: they use only the FP64
oracle, which is pinned bit for bit to the compiled reference."""
import numpy as np
import pytest

import oracle_bind as ob
from pn_amd import lib as pn
from pn_amd import workloads

EPS = 1e-2


def cluster(n_stars, seed, m_bh=0.05):
    return pn.ic_plummer(n_stars, seed, a=1.0, m_bh=m_bh)


def coords_of(n, rng, k):
    out = [("m0",)]
    for _ in range(k):
        out.append(("x" if rng.random() < 0.5 else "v", int(rng.integers(0, n)),
                    int(rng.integers(0, 3))))
    out += [("x", 0, 1), ("v", 0, 2)]          # the BH's own coordinates
    return out


def fd(fn, x, v, m, coord, h):
    x, v, m = x.copy(), v.copy(), m.copy()
    arr = {"x": x, "v": v, "m0": m}[coord[0]]
    idx = (0,) if coord[0] == "m0" else (coord[1], coord[2])
    base = arr[idx]
    step = h * max(1.0, abs(base))
    arr[idx] = base + step
    fp = fn(x, v, m)
    arr[idx] = base - step
    fm = fn(x, v, m)
    return (fp - fm) / (2 * step)


@pytest.mark.parametrize("c", [1.0, 7.0])
def test_cs_real_part_is_the_oracle(c):
    x, v, m = cluster(40, 1)
    a = ob.orc_force(x, v, m, EPS, c)
    acs = ob.orc_force_cs(x, v, m, EPS, c)
    assert np.abs(acs.imag).max() == 0.0
    assert np.abs(acs.real - a).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("c", [1.0, 7.0])
def test_cs_force_derivative_vs_finite_differences(c):
    x, v, m = cluster(32, 2)
    rng = np.random.default_rng(3)
    u = rng.normal(size=x.shape)
    coords = coords_of(len(m), rng, 10)
    g = ob.cs_force_vjp(x, v, m, u, EPS, c, coords)
    phi = lambda xx, vv, mm: float(np.sum(u * ob.orc_force(xx, vv, mm, EPS, c)))
    gfd = np.array([fd(phi, x, v, m, co, 1e-6) for co in coords])
    scale = np.abs(g).max()
    assert np.abs(g - gfd).max() <= 1e-6 * scale, (g, gfd)


def test_cs_trajectory_derivative_vs_finite_differences():
    x, v, m = cluster(16, 4, m_bh=0.02)
    eps, c, dt, steps = 1e-2, 10.0, 2.0**-8, 6
    xt, vt = ob.orc_q1_steps(x, v, m, eps, c, dt, steps)
    _, gx, gv = workloads.radial_loss(xt, vt, r_cut=1.0)
    rng = np.random.default_rng(5)
    coords = coords_of(len(m), rng, 8)
    g = ob.cs_trajectory_grad(x, v, m, eps, c, dt, steps, (gx, gv), coords)

    def loss(xx, vv, mm):
        a, b = ob.orc_q1_steps(xx, vv, mm, eps, c, dt, steps)
        return float(np.sum(gx * a) + np.sum(gv * b))   # linearised loss

    gfd = np.array([fd(loss, x, v, m, co, 1e-6) for co in coords])
    assert np.abs(g - gfd).max() <= 1e-6 * np.abs(g).max(), (g, gfd)


def test_radial_loss_gradient_vs_finite_differences():
    rng = np.random.default_rng(6)
    x = rng.normal(0, 0.3, (60, 3))
    v = rng.normal(0, 1, (60, 3))
    loss, gx, gv = workloads.radial_loss(x, v, r_cut=0.5)
    assert loss > 0.0
    h = 1e-6
    for arr, g in ((x, gx), (v, gv)):
        for p in (0, 5, 33):
            for k in range(3):
                arr[p, k] += h
                lp = workloads.radial_loss(x, v, 0.5)[0]
                arr[p, k] -= 2 * h
                lm = workloads.radial_loss(x, v, 0.5)[0]
                arr[p, k] += h
                assert abs((lp - lm) / (2 * h) - g[p, k]) <= 1e-7 * max(1.0, abs(g[p, k]))
