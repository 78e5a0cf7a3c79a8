"""Adjoint of the force and of the stage driver (pn1_force_vjp,
pn1_adjoint_forward / _backward) against the complex-step oracle and
finite differences.

Complex step (oracle/ls9_oracle_cs.c) differentiates the oracle's own code
exactly to rounding.  The GPU evaluates the same formulas in another
arrangement (contracted with the cotangent first, FP64 rsqrt, other
summation order): the VJP and the trajectory gradient are held to 1e-12 of
the largest gradient component compared (measured: 1e-14 to 1e-16).  At the
C4 size (N = 16384, 256 steps) the oracle is too slow, and the gradient is
checked against central finite differences of GPU trajectories
(h = 1e-6 relative), whose own truncation and rounding set the bound 1e-5.
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from conftest import GOLDEN
from pn_amd import lib as pn
from pn_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pn.Context(capacity=20000)
    yield c
    c.close()


def all_coords(n):
    out = [("m0",)]
    for p in range(n):
        for k in range(3):
            out += [("x", p, k), ("v", p, k)]
    return out


def sample_coords(n, k, seed):
    rng = np.random.default_rng(seed)
    out = [("m0",), ("x", 0, 0), ("v", 0, 1)]
    for _ in range(k):
        out.append(("x" if rng.random() < 0.5 else "v", int(rng.integers(1, n)),
                    int(rng.integers(0, 3))))
    return out


def pick(gx, gv, gm0, coords):
    return np.array([gm0 if co[0] == "m0" else (gx if co[0] == "x" else gv)[co[1], co[2]]
                     for co in coords])


CASES = [  # (name, n_stars, m_bh, eps, c, coords)
    ("plummer_n16_c1", 16, 0.05, 1e-2, 1.0, None),
    ("plummer_n16_c30", 16, 0.05, 1e-3, 30.0, None),
    ("plummer_n256_c100", 256, 0.02, 1e-3, 100.0, 48),
    ("plummer_n1024_c5", 1024, 0.05, 1e-3, 5.0, 24),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_force_vjp_vs_complex_step(ctx, case):
    name, ns, mbh, eps, c, k = case
    x, v, m = pn.ic_plummer(ns, 11, a=1.0, m_bh=mbh)
    u = np.random.default_rng(12).normal(size=x.shape)
    ctx.set_params(eps, c, 2.0**-10, pn.FP64)
    ctx.upload(x, v, m)
    gx, gv, gm0 = ctx.force_vjp(u)
    coords = all_coords(len(m)) if k is None else sample_coords(len(m), k, 13)
    want = ob.cs_force_vjp(x, v, m, u, eps, c, coords)
    got = pick(gx, gv, gm0, coords)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"{name}: max |g - g_cs| / max |g_cs| = {err:.3e}")
    assert err <= 1e-12, err
    # translation invariance: the position gradient sums to zero
    assert np.abs(gx.sum(0)).max() <= 1e-12 * np.abs(gx).max() * len(m)


def test_force_vjp_mode_independent(ctx):
    """The adjoint is FP64 whatever the context's mode."""
    x, v, m = pn.ic_plummer(300, 14, a=1.0, m_bh=0.02)
    u = np.random.default_rng(15).normal(size=x.shape)
    out = []
    for mode in (pn.FP64, pn.MIXED):
        ctx.set_params(1e-3, 50.0, 2.0**-10, mode)
        ctx.upload(x, v, m)
        out.append(ctx.force_vjp(u))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_force_vjp_chunking_invariance(ctx):
    """The single-pass kernel's j chunks (automatic, or set by an explicit j
    split: 1, 3, 7, 64 chunks of whole 64-record tiles, ragged last tiles)
    only reorder the sums: every chunking agrees to rounding, and each is
    bitwise reproducible."""
    x, v, m = pn.ic_plummer(1000, 16, a=1.0, m_bh=0.02)
    u = np.random.default_rng(17).normal(size=x.shape)
    ctx.set_params(1e-3, 50.0, 2.0**-10, pn.FP64)
    ctx.upload(x, v, m)
    out = {}
    for split in (0, 1, 3, 7, 64):
        ctx.set_tuning(split, 0, -1)
        a = ctx.force_vjp(u)
        b = ctx.force_vjp(u)
        for p, q in zip(a, b):
            assert np.array_equal(np.asarray(p), np.asarray(q)), split
        out[split] = a
    ctx.set_tuning(0, 0, -1)
    ref = out[0]
    scale = max(np.abs(ref[0]).max(), np.abs(ref[1]).max(), abs(ref[2]))
    for split, (gx, gv, g0) in out.items():
        err = max(np.abs(gx - ref[0]).max(), np.abs(gv - ref[1]).max(), abs(g0 - ref[2]))
        print(f"jsplit {split}: max diff / scale {err / scale:.2e}")
        assert err <= 1e-13 * scale, split


def run_adjoint(ctx, x, v, m, eps, c, dt, steps, interval, r_cut):
    ctx.set_params(eps, c, dt, pn.FP64)
    ctx.upload(x, v, m)
    ctx.adjoint_forward(steps, interval)
    xt, vt = ctx.download()
    loss, lx, lv = workloads.radial_loss(xt, vt, r_cut)
    gx0, gv0, gm0 = ctx.adjoint_backward(lx, lv)
    return loss, (lx, lv), (gx0, gv0, gm0), (xt, vt)


def test_adjoint_trajectory_vs_complex_step(ctx):
    x, v, m = pn.ic_plummer(96, 21, a=1.0, m_bh=0.02)
    eps, c, dt, steps = 1e-3, 100.0, 2.0**-10, 24
    loss, lgrad, (gx0, gv0, gm0), (xt, vt) = run_adjoint(ctx, x, v, m, eps, c, dt, steps, 5,
                                                          r_cut=0.6)
    # the forward part is the stage driver itself
    xo, vo = ob.orc_q1_steps(x, v, m, eps, c, dt, steps)
    assert np.abs(xt - xo).max() <= 1e-12 and np.abs(vt - vo).max() <= 1e-11
    coords = sample_coords(len(m), 20, 22)
    want = ob.cs_trajectory_grad(x, v, m, eps, c, dt, steps, lgrad, coords)
    got = pick(gx0, gv0, gm0, coords)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"trajectory n=97, {steps} steps: loss {loss:.6e}, rel err {err:.3e}")
    assert err <= 1e-12, err


def test_adjoint_checkpoint_interval_invariance(ctx):
    """Recompute from checkpoints is exact: every interval gives the same bits."""
    x, v, m = pn.ic_plummer(200, 23, a=1.0, m_bh=0.02)
    res = [run_adjoint(ctx, x, v, m, 1e-3, 100.0, 2.0**-10, 12, k, 0.5)[2] for k in (1, 5, 12, 50)]
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(np.asarray(a), np.asarray(b))


def test_adjoint_leaves_initial_state(ctx):
    x, v, m = pn.ic_plummer(64, 24, a=1.0, m_bh=0.02)
    run_adjoint(ctx, x, v, m, 1e-3, 100.0, 2.0**-10, 7, 3, 0.5)
    xd, vd = ctx.download()
    assert np.array_equal(xd, x) and np.array_equal(vd, v)


def test_adjoint_rejects_changed_state(ctx):
    x, v, m = pn.ic_plummer(64, 25, a=1.0, m_bh=0.02)
    ctx.set_params(1e-3, 100.0, 2.0**-10, pn.FP64)
    ctx.upload(x, v, m)
    ctx.adjoint_forward(4, 2)
    ctx.set_params(2e-3, 100.0, 2.0**-10, pn.FP64)
    with pytest.raises(pn.Pn1Error, match="changed"):
        ctx.adjoint_backward(np.zeros_like(x), np.zeros_like(v))


def test_refused_allocation_leaves_context_usable():
    """An adjoint run whose checkpoint segment cannot be allocated (10^9
    stage states of 6 x 4097 doubles, ~590 TB) returns PN1_ENOMEM; the
    context keeps working (the next launches do not see the refused
    allocation's error), and a backward pass has nothing to reverse until a
    forward run completes."""
    x, v, m = pn.ic_plummer(4096, 26, a=1.0, m_bh=0.02)
    c = pn.Context(len(m))
    try:
        c.set_params(1e-3, 100.0, 2.0**-10, pn.FP64)
        c.upload(x, v, m)
        c.adjoint_forward(2, 2)                 # a complete run first
        a0 = c.force()
        with pytest.raises(pn.Pn1Error) as ei:
            c.adjoint_forward(10**9, 10**9)
        assert ei.value.code == pn.PN1_ENOMEM
        assert np.array_equal(c.force(), a0)
        with pytest.raises(pn.Pn1Error, match="no pn1_adjoint_forward"):
            c.adjoint_backward(np.zeros_like(x), np.zeros_like(v))
        c.adjoint_forward(2, 2)
        gx0, gv0, gm0 = c.adjoint_backward(np.zeros_like(x), np.ones_like(v))
        assert np.isfinite(gx0).all() and np.isfinite(gv0).all() and np.isfinite(gm0)
        c.step(2)
    finally:
        c.close()


def test_adjoint_c4_vs_finite_differences(ctx):
    """The differentiation workload: N = 16384 Plummer (seed 4) + BH 0.02, c = 100,
    eps = 1e-3, dt = 2^-10, 256 steps, checkpoints every 16, loss = mean
    squared radial velocity within r < 0.2; 64 random coordinates (48 over
    all stars, 16 among the stars starting within r < 0.3, whose gradients
    are largest) and m0 against central differences of the same GPU
    trajectories."""
    W = workloads.ADJOINT
    x, v, m = pn.ic_plummer(W["n_stars"], W["seed"], a=1.0, m_bh=W["m_bh"])
    loss, lgrad, (gx0, gv0, gm0), _ = run_adjoint(ctx, x, v, m, W["eps"], W["c"], W["dt"],
                                                  W["steps"], W["interval"], W["r_cut"])
    lx, lv = lgrad
    assert loss > 0.0
    r = np.linalg.norm(x[1:] - x[0], axis=1)
    rng = np.random.default_rng(26)
    inner = 1 + np.flatnonzero(r < 0.3)
    picks = list(rng.choice(np.arange(1, len(m)), 48, replace=False)) + \
        list(rng.choice(inner, 16, replace=False))
    coords = [("m0",)] + [("x" if rng.random() < 0.5 else "v", int(p), int(rng.integers(0, 3)))
                          for p in picks]

    def traj_loss(xx, vv, mm):
        ctx.upload(xx, vv, mm)
        ctx.step(W["steps"])
        xt, vt = ctx.download()
        return float(np.sum(lx * xt) + np.sum(lv * vt))     # linearised loss

    h = 1e-6
    got, want = [], []
    for co in coords:
        xx, vv, mm = x.copy(), v.copy(), m.copy()
        arr = {"x": xx, "v": vv, "m0": mm}[co[0]]
        idx = (0,) if co[0] == "m0" else (co[1], co[2])
        step = h * max(1.0, abs(arr[idx]))
        base = arr[idx]
        arr[idx] = base + step
        fp = traj_loss(xx, vv, mm)
        arr[idx] = base - step
        fm = traj_loss(xx, vv, mm)
        want.append((fp - fm) / (2 * step))
        got.append(gm0 if co[0] == "m0" else (gx0 if co[0] == "x" else gv0)[co[1], co[2]])
    got, want = np.array(got), np.array(want)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"C4: loss {loss:.6e}, |g| max {np.abs(want).max():.3e}, FD rel err {err:.3e}")
    assert err <= 1e-5, (got, want)


def test_adjoint_c4_segment_vs_oracle_complex_step(ctx):
    """The differentiation workload at full size (N = 16384 Plummer, seed 4,
    + BH 0.02, c = 100, eps = 1e-3, dt = 2^-10) over 8 steps: the GPU adjoint
    of the radial-velocity loss (r < 0.2) against complex step through the
    oracle's own stage driver (oracle/ls9_oracle_cs.c) at m0 and 16
    coordinates of (x0, v0), the committed fixture
    tests/golden/c4_grad_8steps.npz (tests/golden/make_c4_grad.py): max
    difference <= 1e-10 of the largest component."""
    z = np.load(os.path.join(GOLDEN, "c4_grad_8steps.npz"))
    W = workloads.ADJOINT
    steps = int(z["steps"])
    x, v, m = pn.ic_plummer(W["n_stars"], W["seed"], a=1.0, m_bh=W["m_bh"])
    print(f"case: C4 segment, N = {W['n_stars']}, {steps} steps", flush=True)
    ctx.set_params(W["eps"], W["c"], W["dt"], pn.FP64)
    ctx.upload(x, v, m)
    ctx.adjoint_forward(steps, 16)
    xt, vt = ctx.download()
    loss, lx, lv = workloads.radial_loss(xt, vt, W["r_cut"])
    gx0, gv0, gm0 = ctx.adjoint_backward(lx, lv)
    assert abs(loss - float(z["loss"])) <= 1e-12 * abs(float(z["loss"]))
    names = {0: "m0", 1: "x", 2: "v"}
    coords = [(names[int(k)],) if int(k) == 0 else (names[int(k)], int(p), int(a))
              for k, p, a in zip(z["kind"], z["particle"], z["axis"])]
    got = pick(gx0, gv0, gm0, coords)
    want = z["grad"]
    err = np.abs(got - want) / np.abs(want).max()
    print(f"C4 segment: loss {loss:.9e}, |g| max {np.abs(want).max():.3e}, max diff "
          f"{err.max():.3e} of it (coordinate {coords[int(err.argmax())]})", flush=True)
    assert err.max() <= 1e-10, (got, want)


def test_torch_autograd_wrapper(ctx):
    """A loss written in torch, differentiated through the library's adjoint
    (tests/torch_adjoint.py), against the numpy path and torch.autograd's
    finite-difference gradcheck on a small system."""
    import torch
    import torch_adjoint as ta

    x, v, m = pn.ic_plummer(12, 31, a=1.0, m_bh=0.02)
    eps, c, dt, steps = 1e-2, 20.0, 2.0**-9, 10
    ctx.set_params(eps, c, dt, pn.FP64)
    x0 = torch.tensor(x, requires_grad=True)
    v0 = torch.tensor(v, requires_grad=True)
    m0 = torch.tensor(m[0], dtype=torch.float64, requires_grad=True)
    xt, vt = ta.Trajectory.apply(x0, v0, m0, ctx, m, steps, 3)
    loss = ta.radial_loss_torch(xt, vt, 0.8)
    loss.backward()
    # numpy path: same loss, gradient by hand, adjoint through the C ABI
    _, _, (gx0, gv0, gm0), _ = run_adjoint(ctx, x, v, m, eps, c, dt, steps, 3, 0.8)
    assert np.allclose(x0.grad.numpy(), gx0, rtol=1e-12, atol=1e-14)
    assert np.allclose(v0.grad.numpy(), gv0, rtol=1e-12, atol=1e-14)
    assert abs(m0.grad.item() - gm0) <= 1e-12 * max(1.0, abs(gm0))

    def f(xx, vv, mm):
        a, b = ta.Trajectory.apply(xx, vv, mm, ctx, m, steps, 3)
        return ta.radial_loss_torch(a, b, 0.8)

    assert torch.autograd.gradcheck(f, (x0.detach().requires_grad_(),
                                        v0.detach().requires_grad_(),
                                        m0.detach().requires_grad_()),
                                    eps=1e-6, atol=1e-6, rtol=1e-4)
