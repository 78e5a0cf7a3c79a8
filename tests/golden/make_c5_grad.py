"""Writes tests/golden/c5_grad_m0_k<K>.npz: for clusters K of the
batched-ensemble workload (pn_amd/workloads.py: ensemble(), seed 5), the
derivative dL_T/dm0 of the cluster's radial-velocity loss after T steps with
respect to its BH mass, for every T = 1 .. 128, by complex step through the
oracle's stage driver (oracle/ls9_oracle_cs.c, the reference's formulas
Limit-sticky9.c:21-424, 841-893).

One complex-valued run carries the tangent d(x_T, v_T)/dm0 of every step in
its imaginary parts (m0 + i h, h = 1e-30), so the whole series costs one run.
A second run starts from the same state with every star's position moved up
by one ulp: the spread between the two series at a step T is the
conditioning of dL_T/dm0 itself, i.e. how far any implementation's value at
T can be pinned.  Also stored: the loss and the closest star-BH approach up
to each step.

    python tests/golden/make_c5_grad.py K [K ...]

Slow: two complex-valued 128-step runs per cluster (in parallel, about 15 min
at N = 2284 on one core each).  Needs the oracle built (make -C oracle) and
libpn1.so for the seeded initial conditions (host code only).
"""
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

import oracle_bind as ob  # noqa: E402
from pn_amd import workloads  # noqa: E402

W = workloads.ENSEMBLE
STEPS = 128
R_CUT = 0.2


def series(k, ulp):
    """(dL_T/dm0, L_T, closest approach up to T) for T = 1 .. STEPS."""
    x, v, m = workloads.ensemble()[k]
    if ulp:
        x = x.copy()
        x[1:] = np.nextafter(x[1:], np.inf)
    eps, c, dt = W["eps"], W["c"], W["dt"]
    xc, vc, mc = ob._perturbed(x, v, m, ("m0",))
    g, loss, dmin = np.empty(STEPS), np.empty(STEPS), np.empty(STEPS)
    d = np.inf
    for s in range(STEPS):
        xc, vc = ob.orc_q1_steps_cs(xc, vc, mc, eps, c, dt, 1)
        xr, vr = xc.real.copy(), vc.real.copy()
        loss[s], gx, gv = workloads.radial_loss(xr, vr, R_CUT)
        g[s] = (np.sum(gx * xc.imag) + np.sum(gv * vc.imag)) / ob.CS_H
        d = min(d, float(np.linalg.norm(xr[1:] - xr[0], axis=1).min()))
        dmin[s] = d
        if (s + 1) % 16 == 0:
            print(f"cluster {k}{' (+1 ulp)' if ulp else ''}: step {s + 1}, dL/dm0 {g[s]:.6e}",
                  flush=True)
    return g, loss, dmin


def one(k):
    with mp.Pool(2) as pool:
        (g, loss, dmin), (gu, _, _) = pool.starmap(series, [(k, False), (k, True)])
    x, v, m = workloads.ensemble()[k]
    out = os.path.join(HERE, f"c5_grad_m0_k{k}.npz")
    np.savez(out, cluster=k, n_stars=len(m) - 1, m0=m[0], r_cut=R_CUT, grad_m0=g,
             grad_m0_ulp=gu, loss=loss, closest_bh_approach=dmin)
    spread = np.abs(g - gu) / np.abs(g)
    print(f"cluster {k} N={len(m) - 1}: dL/dm0 after {STEPS} steps {g[-1]:.6e} (+1 ulp: "
          f"{gu[-1]:.6e}); closest star-BH approach {dmin[-1]:.3e} at step "
          f"{int(np.argmax(dmin == dmin[-1])) + 1}; 1-ulp spread by step 16/64/96/128: "
          f"{spread[15]:.1e} {spread[63]:.1e} {spread[95]:.1e} {spread[-1]:.1e}", flush=True)


if __name__ == "__main__":
    for a in sys.argv[1:]:
        one(int(a))
