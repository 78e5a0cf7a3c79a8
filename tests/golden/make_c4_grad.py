"""Writes tests/golden/c4_grad_8steps.npz: gradients of the differentiation
workload's loss over a short segment, by complex step through the oracle's
stage driver.

Workload (pn_amd/workloads.py: ADJOINT): a seeded Plummer sphere of 16384
stars (seed 4) + BH of mass 0.02, c = 100, eps = 1e-3, dt = 2^-10; the loss
is the mean squared radial velocity of the stars within r < 0.2 of the BH
after STEPS = 8 steps of the reference's stage driver (Limit-sticky9.c:
841-893, on the force of Limit-sticky9.c:21-424 with 1/c^2 on its 1PN
terms).  For m0 and 16 coordinates of (x0, v0) (8 of stars inside the loss
region at the end, 8 seeded random ones, x and v alternating), d loss /
d coordinate = (dL/dx_T, dL/dv_T) . Im(x_T, v_T) / h, where (x_T, v_T) is
the oracle's trajectory run in complex arithmetic (oracle/ls9_oracle_cs.c)
from the initial state with i h added to that coordinate (h = 1e-30); the
loss gradient is evaluated at the oracle's real final state.

    python tests/golden/make_c4_grad.py [processes]

Slow: one complex-valued 8-step run per coordinate (about 20 min each on
one core at N = 16384), spread over `processes` (default 6) processes.
Needs the oracle built (make -C oracle) and libpn1.so for the seeded
initial conditions (host code only).
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

import oracle_bind as ob  # noqa: E402
from pn_amd import lib as pn  # noqa: E402
from pn_amd import workloads  # noqa: E402

W = workloads.ADJOINT
STEPS = 8
OUT = os.path.join(HERE, "c4_grad_8steps.npz")


def initial():
    return pn.ic_plummer(W["n_stars"], W["seed"], a=1.0, m_bh=W["m_bh"])


def final_and_loss():
    x, v, m = initial()
    xt, vt = ob.orc_q1_steps(x, v, m, W["eps"], W["c"], W["dt"], STEPS)
    return xt, vt, workloads.radial_loss(xt, vt, W["r_cut"])


def coords(xt):
    n = xt.shape[0]
    r = np.linalg.norm(xt[1:] - xt[0], axis=1)
    inside = 1 + np.flatnonzero(r < W["r_cut"])
    rng = np.random.default_rng(4)
    picks = list(rng.choice(inside, 8, replace=False)) + list(rng.integers(1, n, 8))
    out = [("m0",)]
    for q, p in enumerate(picks):
        out.append(("x" if q % 2 == 0 else "v", int(p), int(rng.integers(0, 3))))
    return out


def one(args):
    co, gx, gv = args
    x, v, m = initial()
    t0 = time.time()
    xt, vt = ob.orc_q1_steps_cs(*ob._perturbed(x, v, m, co), W["eps"], W["c"], W["dt"], STEPS)
    g = float(np.sum(gx * xt.imag) + np.sum(gv * vt.imag)) / ob.CS_H
    print(f"{co}: {g:.15e} ({time.time() - t0:.0f} s)", flush=True)
    return g


def main(procs):
    xt, vt, (loss, gx, gv) = final_and_loss()
    cs = coords(xt)
    print(f"loss after {STEPS} steps {loss:.15e}; {len(cs)} coordinates", flush=True)
    with mp.Pool(procs) as pool:
        grads = pool.map(one, [(co, gx, gv) for co in cs], chunksize=1)
    kind = np.array([{"m0": 0, "x": 1, "v": 2}[co[0]] for co in cs], dtype=np.int32)
    part = np.array([co[1] if len(co) > 1 else 0 for co in cs], dtype=np.int32)
    axis = np.array([co[2] if len(co) > 1 else 0 for co in cs], dtype=np.int32)
    np.savez(OUT, steps=STEPS, loss=loss, kind=kind, particle=part, axis=axis,
             grad=np.array(grads), x_final_checksum=float(np.abs(xt).sum()))
    print(f"wrote {OUT}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 6)
