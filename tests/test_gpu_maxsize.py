"""The HIP path at the reference's largest size and at the headline size,
checked row by row against the oracle.

The reference caps n at NMAX = 270000 (Limit-sticky9.c:11, the static
arrays of main(), Limit-sticky9.c:742-751).  A full oracle force() at that
size is ~1 h of CPU, so the check takes a sample of star rows: each row of
the oracle's force() is an O(n) sum (orc_force_rows, bitwise those rows of
the full call), and the sample covers the 256-star block edges of the pair
kernels, the last star, the stars closest to the BH (where the 1PN terms
are largest) and seeded random stars.  The BH's own row sums over every
pair; it is checked against the FP64 form (mixed) and through Newton's
third law (the FP64 form's Newtonian part).

Tolerances:
  FP64 rows: ||a - a_ref|| / ||a_ref|| <= 1e-11.  The HIP path sums each
    row's n terms in a different order than the reference's one serial
    loop (a j split into partial sums, fixed-order combination).  Measured
    worst row of the sample on MI355X: 2.9e-13 at n = 270000, c = 1 (a star
    near the BH, whose 1PN terms cancel against each other), 3.6e-14 at
    c = 50, 2.7e-14 at the headline size.
  Mixed: against FP64 (every star) and the oracle rows: max <= 1e-6,
    BH <= 1e-6, at the initial state and one FP64 step on.  Measured on
    MI355X: 3.9e-7 (NMAX, one step on).
"""
import numpy as np
import pytest

import oracle_bind as ob
from conftest import rel_err
from pn_amd import lib as pn

pytestmark = pytest.mark.gpu

NMAX = 270000


def sample_rows(x, seed, k=32):
    n = x.shape[0]
    edges = [1, 2, 63, 64, 65, 255, 256, 257, 511, 512, n - 257, n - 256, n - 255, n - 2, n - 1]
    d = np.linalg.norm(x[1:] - x[0], axis=1)
    near = 1 + np.argsort(d)[:8]
    rnd = np.random.default_rng(seed).integers(1, n, k)
    rows = np.unique(np.concatenate([edges, near, rnd]))
    return rows[(rows >= 1) & (rows < n)].astype(np.int32)


def check_rows(a, x, v, m, eps, c, seed):
    rows = sample_rows(x, seed)
    ref = ob.orc_force_rows(x, v, m, eps, c, rows)
    err = np.linalg.norm(a[rows] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"n={len(m)} c={c}: {len(rows)} oracle rows, max rel {err.max():.3e} "
          f"(row {rows[err.argmax()]})", flush=True)
    assert err.max() <= 1e-11, (rows[err.argmax()], err.max())


def mixed_vs_fp64(a32, a64, label):
    err = rel_err(a32, a64)
    print(f"{label}: mixed star p99.9 {np.quantile(err[1:], 0.999):.3e} max {err[1:].max():.3e} "
          f"BH {err[0]:.3e}", flush=True)
    assert err[1:].max() <= 1e-6
    assert err[0] <= 1e-6


def mixed_rows_vs_oracle(a32, x, v, m, eps, c, seed, label):
    rows = sample_rows(x, seed)
    ref = ob.orc_force_rows(x, v, m, eps, c, rows)
    err = np.linalg.norm(a32[rows] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"{label}: mixed vs {len(rows)} oracle rows max rel {err.max():.3e}", flush=True)
    assert err.max() <= 1e-6, (rows[err.argmax()], err.max())


@pytest.fixture(scope="module")
def nmax():
    x, v, m = pn.ic_plummer(NMAX - 1, 11, a=1.0, m_bh=0.05)
    assert len(m) == NMAX
    ctx = pn.Context(NMAX)
    yield ctx, x, v, m
    ctx.close()


@pytest.mark.parametrize("c", [1.0, 50.0])
def test_nmax_fp64_rows_vs_oracle(nmax, c):
    ctx, x, v, m = nmax
    print(f"case: n = {NMAX} Plummer, FP64, c = {c}", flush=True)
    ctx.set_params(1e-4, c, 2.0**-12, pn.FP64)
    ctx.upload(x, v, m)
    a = ctx.force()
    assert np.isfinite(a).all()
    check_rows(a, x, v, m, 1e-4, c, seed=int(c))


def test_nmax_mixed_and_momentum(nmax):
    ctx, x, v, m = nmax
    print(f"case: n = {NMAX} Plummer, mixed vs FP64 and oracle rows, c = 50", flush=True)
    ctx.set_params(1e-4, 50.0, 2.0**-12, pn.FP64)
    ctx.upload(x, v, m)
    a64 = ctx.force()
    ctx.set_params(1e-4, 50.0, 2.0**-12, pn.MIXED)
    a32 = ctx.force()
    mixed_vs_fp64(a32, a64, f"n={NMAX} c=50")
    mixed_rows_vs_oracle(a32, x, v, m, 1e-4, 50.0, 11, f"n={NMAX} c=50")
    # the same after one FP64 step (positions with all 53 bits: the
    # double-single lo parts are then not zero)
    ctx.set_params(1e-4, 50.0, 2.0**-12, pn.FP64)
    ctx.step(1)
    xs, vs = ctx.download()
    a64 = ctx.force()
    ctx.set_params(1e-4, 50.0, 2.0**-12, pn.MIXED)
    ctx.upload(xs, vs, m)
    a32 = ctx.force()
    mixed_vs_fp64(a32, a64, f"n={NMAX} c=50, one step on")
    mixed_rows_vs_oracle(a32, xs, vs, m, 1e-4, 50.0, 12, f"n={NMAX} c=50, one step on")
    ctx.upload(x, v, m)
    # Newtonian part alone (c -> infinity), FP64: the total momentum change
    # vanishes, BH row included
    ctx.set_params(1e-4, 1e30, 2.0**-12, pn.FP64)
    a = ctx.force()
    p = (m[:, None] * a).sum(0)
    assert np.abs(p).max() < 1e-12 * (m[:, None] * np.abs(a)).sum(), p


def test_nmax_legacy_shape(nmax):
    """The reference's own force() prototype at n = NMAX equals the context
    API's FP64 force."""
    ctx, x, v, m = nmax
    print(f"case: n = {NMAX} legacy force()", flush=True)
    pn.set_c(50.0)
    pn.set_mode(pn.FP64)
    try:
        a = pn.legacy_force(x, v, m, 1e-4)
    finally:
        pn.set_c(1.0)
    ctx.set_params(1e-4, 50.0, 2.0**-12, pn.FP64)
    ctx.upload(x, v, m)
    ref = ctx.force()
    assert rel_err(a, ref).max() <= 1e-13


def test_headline_fp64_rows_vs_oracle():
    """N = 131072 Plummer + BH 0.05, c = 50, eps = 1e-4 (the benchmark's
    workload): FP64 rows against the oracle."""
    x, v, m = pn.ic_plummer(131072, 3, a=1.0, m_bh=0.05)
    print("case: n = 131073 Plummer, FP64 rows, c = 50", flush=True)
    ctx = pn.Context(len(m))
    try:
        ctx.set_params(1e-4, 50.0, 2.0**-12, pn.FP64)
        ctx.upload(x, v, m)
        a = ctx.force()
    finally:
        ctx.close()
    check_rows(a, x, v, m, 1e-4, 50.0, seed=3)
