"""Writes the golden fixtures in tests/golden/ from the compiled reference.

Run in a container that has the reference sources (it builds
oracle/_ref/libls9ref.so from /root/reference/Limit-sticky9.c with
oracle/Makefile, then calls that library):

    python tests/golden/make_golden.py

Each fixture is an .npz of inputs and the reference's outputs:
  x, v, m, eps             inputs (AoS [n][3], index 0 = BH)
  a_c1, E0_c1, init_txt_c1, energy_txt_c1
                           force(), initial_energy()'s *init_ene and printed
                           text, energy()'s printed text (t = 0.25,
                           init_ene = E0 (1 + 1e-6)) -- the reference as is
  a_c<C>, E_c<C>           force and energy components at speed of light C,
                           from the reference on rescaled inputs:
                           a_C(x, v, m) = C^2 a_ref(x, v/C, m/C^2),
                           E_C(x, v, m) = C^4 E_ref(x, v/C, m/C^2)
                           (every 1PN term has one more power of m or v^2
                           than the Newtonian ones; the reference is C = 1)
Trajectory fixtures (traj_*.npz) hold the state after K steps of a stage
loop written here from Limit-sticky9.c:841-893 around the reference's own
force(); the stage arithmetic is numpy float64 in the reference's
operation order.
"""
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle_bind as ob  # noqa: E402
from pn_amd import lib as pn  # noqa: E402

REF_SRC = "/root/reference/Limit-sticky9.c"
REF_DAT = "/root/reference/Init10k-sphere.dat"
NUM = r"[-+]?\d+\.\d+e[-+]\d+"


def build_ref():
    subprocess.run(["make", "-s", "-C", ob.ORACLE_DIR, "ref"], check=True)


def parse_block(txt):
    """pot, kin, pot_pn, total from the printed energy block."""
    m = re.search(rf"pot = \s*({NUM}) kin = \s*({NUM}) \n pot_pn = \s*({NUM}) \n  total= \s*({NUM})",
                  txt)
    return np.array([float(g) for g in m.groups()])


def rescaled(c, x, v, m, eps):
    """Reference force and energy at speed of light c via the rescaling."""
    vr, mr = v / c, m / (c * c)
    a = c * c * ob.ref_force(x, vr, mr, eps)
    E0, txt = ob.ref_initial_energy_text(x, vr, mr, eps)
    pot, kin, potpn, total = parse_block(txt)
    c4 = c ** 4
    # {kin, pot, pot_pn, total}; total from the full-precision *init_ene
    return a, c4 * np.array([kin, pot, potpn, E0])


def random_case(n, seed, spread_mass=True, eps=1e-3, vscale=0.3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 3))
    v = rng.normal(0.0, vscale, (n, 3))
    m = np.full(n, 1.0 / max(n - 1, 1))
    if spread_mass:
        m[1:] *= rng.uniform(0.5, 1.5, n - 1)
    m[0] = 0.05
    x[0] = rng.uniform(-0.05, 0.05, 3)     # BH off the origin ...
    v[0] = rng.normal(0.0, 0.01, 3)        # ... and moving (v0 enters the energy)
    return x, v, m, eps


def c1_case():
    """1024 stars, seeded uniform sphere R = 1, virial 0.5, m_BH = 0.01, eps = 1e-3."""
    x, v, m = pn.ic_uniform_sphere(1024, 1, R=1.0, Q=0.5, m_bh=0.01, eps=1e-3)
    return x, v, m, 1e-3


def dat_case(nstars=64):
    """The head of the reference's own input file with the reference's
    masses (Limit-sticky9.c:813-816: m0 = 100/n, m_i = 1/(n-1)), eps = 1e-3."""
    path = os.path.join(HERE, "init10k_head.dat")
    n = nstars + 1
    x, v = pn.read_dat(path, n)
    m = np.full(n, 1.0 / (n - 1))
    m[0] = 100.0 / n
    return x, v, m, 1e-3


def write_dat_head(nstars=64):
    with open(REF_DAT) as f:
        lines = f.readlines()
    half = len(lines) // 2
    with open(os.path.join(HERE, "init10k_head.dat"), "w") as f:
        f.writelines(lines[:nstars])
        f.writelines(lines[half:half + nstars])


def force_case(name, x, v, m, eps, cs=(100.0,)):
    a1 = ob.ref_force(x, v, m, eps)
    E0, txt0 = ob.ref_initial_energy_text(x, v, m, eps)
    etxt = ob.ref_energy_text(0.25, x, v, m, E0 * (1.0 + 1e-6), eps)
    out = dict(x=x, v=v, m=m, eps=np.float64(eps), a_c1=a1, E0_c1=np.float64(E0),
               init_txt_c1=np.array(txt0), energy_txt_c1=np.array(etxt))
    for c in cs:
        a, e = rescaled(c, x, v, m, eps)
        out[f"a_c{int(c)}"] = a
        out[f"E_c{int(c)}"] = e
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: n={len(m)} |a|max={np.abs(a1).max():.3e} E0={E0:.15e}", flush=True)


def harness_steps(force_fn, x, v, dt, steps):
    """Stage loop of Limit-sticky9.c:841-893, numpy float64 arithmetic."""
    x, v = x.copy(), v.copy()
    for _ in range(steps):
        hx0, hv0 = x.copy(), v.copy()
        a = force_fn(x, v)
        x1 = x + 0.5 * v * dt
        v1 = v + 0.5 * a * dt
        hx1, hv1 = x1, v1
        a = force_fn(x1, v1)
        x1 = x + 0.5 * v1 * dt       # v1 of stage 1 (statement order)
        v1 = v + 0.5 * a * dt
        hx2, hv2 = x1, v1
        a = force_fn(x1, v1)
        x1 = x + v1 * dt
        v1 = v + a * dt
        hx3, hv3 = x1, v1
        x = (hx0 + hx3 + 2.0 * (hx1 + hx2)) / 6.0
        v = (hv0 + hv3 + 2.0 * (hv1 + hv2)) / 6.0
    return x, v


def traj_case(name, x, v, m, eps, c, dt, steps):
    vr_m = m / (c * c)

    def f(xs, vs):
        return c * c * ob.ref_force(xs, vs / c, vr_m, eps)

    xK, vK = harness_steps(f, x, v, dt, steps)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), x=x, v=v, m=m, eps=np.float64(eps),
                        c=np.float64(c), dt=np.float64(dt), steps=np.int64(steps), xK=xK, vK=vK)
    print(f"{name}: n={len(m)} c={c} dt={dt} steps={steps}", flush=True)


def main():
    if not os.path.exists(REF_SRC):
        sys.exit("reference sources not present: fixtures cannot be regenerated here")
    build_ref()
    write_dat_head()
    for n, seed in [(2, 11), (3, 12), (17, 13), (65, 14), (257, 15)]:
        force_case(f"rand_n{n}", *random_case(n, seed), cs=(100.0, 50.0))
    # close pairs: two stars 3e-4 apart and one 2e-4 from the BH, eps = 1e-4
    x, v, m, _ = random_case(65, 16)
    x[7] = x[6] + np.array([3e-4, -1e-4, 2e-4])
    x[9] = x[0] + np.array([2e-4, 0.0, -1e-4])
    force_case("close_n65", x, v, m, 1e-4, cs=(50.0,))
    force_case("dat_n65", *dat_case())
    force_case("c1_n1025", *c1_case(), cs=(100.0,))
    traj_case("traj_rand_n65_c1", *random_case(65, 17), c=1.0, dt=2.0**-10, steps=16)
    traj_case("traj_c1_n1025_c100", *c1_case(), c=100.0, dt=2.0**-10, steps=64)


if __name__ == "__main__":
    main()
