"""Seeded synthetic workloads (initial conditions from libpn1's generators).

    plummer_cluster()  N = 131072 Plummer stars + BH of mass 0.05 (c = 50,
                       eps = 1e-4, dt = 2^-12)
    ensemble()         256 Plummer clusters, star counts log-uniform in
                       [256, 8192], BH masses uniform in [0.005, 0.05]
                       (c = 100, eps = 1e-3, dt = 2^-10)
    ADJOINT            N = 16384 Plummer stars + BH of mass 0.02, c = 100,
                       eps = 1e-3, dt = 2^-10, 256 steps, checkpoints every
                       16 steps; loss radial_loss(r_cut = 0.2)
"""
import numpy as np

from . import lib as pn

PLUMMER = dict(n_stars=131072, seed=3, m_bh=0.05, c=50.0, eps=1e-4, dt=2.0**-12)
ENSEMBLE = dict(n_clusters=256, seed=5, n_min=256, n_max=8192, m_bh_min=0.005,
                m_bh_max=0.05, c=100.0, eps=1e-3, dt=2.0**-10)
# seed 4 (closest star-BH approach in 256 steps: 4.7e-3).  With seed 7 one
# star falls to 2.75e-6 of the BH at step 165, where the reference's
# unsoftened 1/|d|^2 BH term (Limit-sticky9.c:74) makes the derivative of
# the step map ~1e16 (finite differences then depend on h).
ADJOINT = dict(n_stars=16384, seed=4, m_bh=0.02, c=100.0, eps=1e-3, dt=2.0**-10, steps=256,
               interval=16, r_cut=0.2)


def plummer_cluster(n_stars=PLUMMER["n_stars"], seed=PLUMMER["seed"], m_bh=PLUMMER["m_bh"]):
    return pn.ic_plummer(n_stars, seed, a=1.0, m_bh=m_bh)


def ensemble_shape(n_clusters=ENSEMBLE["n_clusters"], seed=ENSEMBLE["seed"],
                   n_min=ENSEMBLE["n_min"], n_max=ENSEMBLE["n_max"],
                   m_bh_min=ENSEMBLE["m_bh_min"], m_bh_max=ENSEMBLE["m_bh_max"]):
    """Star counts and BH masses of the clusters (numpy PCG64 from `seed`)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, n_clusters)
    sizes = np.rint(np.exp(np.log(n_min) + u * (np.log(n_max) - np.log(n_min)))).astype(int)
    m_bh = rng.uniform(m_bh_min, m_bh_max, n_clusters)
    return sizes, m_bh


def ensemble(n_clusters=ENSEMBLE["n_clusters"], seed=ENSEMBLE["seed"], **kw):
    """List of (x, v, m) states, cluster k a Plummer sphere of sizes[k]
    stars (scale radius 1, total star mass 1) from generator seed
    1000 * seed + k."""
    sizes, m_bh = ensemble_shape(n_clusters, seed, **kw)
    return [pn.ic_plummer(int(ns), 1000 * seed + k, a=1.0, m_bh=float(mb))
            for k, (ns, mb) in enumerate(zip(sizes, m_bh))]


def interactions(clusters):
    """Pair interactions of one force evaluation: sum of n^2 over clusters
    (n counts the BH), the convention of the single-cluster figures."""
    return float(sum(float(len(m)) ** 2 for _, _, m in clusters))


def radial_loss(x, v, r_cut=0.2):
    """Mean squared radial velocity of the stars within r_cut of the BH,
    relative to the BH: v_r = (v_i - v_0) . (x_i - x_0) / |x_i - x_0|.
    Returns (loss, dL/dx, dL/dv) (the BH's rows carry minus the sum of the
    stars', as the loss depends on relative coordinates only).  The set of
    stars inside r_cut is taken as fixed (the loss is smooth away from
    stars exactly at r_cut)."""
    r = x[1:] - x[0]
    w = v[1:] - v[0]
    R = np.sqrt((r * r).sum(1))
    sel = R < r_cut
    k = int(sel.sum())
    gx = np.zeros_like(x)
    gv = np.zeros_like(v)
    if k == 0:
        return 0.0, gx, gv
    vr = (w * r).sum(1) / R
    loss = float((vr[sel] ** 2).sum() / k)
    coef = np.where(sel, 2.0 * vr / k, 0.0)
    gv[1:] = coef[:, None] * r / R[:, None]
    gx[1:] = coef[:, None] * (w - vr[:, None] * r / R[:, None]) / R[:, None]
    gx[0] = -gx[1:].sum(0)
    gv[0] = -gv[1:].sum(0)
    return loss, gx, gv
