"""Ensemble API (pn1_ens_*): many clusters in one persistent pair launch.

Every cluster of an ensemble must come out as if it had been run alone:
  - with one j chunk per star (chunk >= the largest cluster) and a context
    whose j range is not split (jsplit 1), the ensemble and the context
    evaluate the same pair sums in the same order, so forces, energies and
    whole steps agree bit for bit, in both precision modes;
  - with the automatic chunk (256 to 512 records, several chunks
    per star for the larger clusters), FP64 forces agree with the oracle
    to the same 1e-12 per-particle bound as the context API;
  - the persistent kernel hands items to workgroups in whatever order they
    ask, but every item writes its own slot: repeated evaluations are
    bitwise identical.
Cluster sizes cover the empty cluster (BH only), one star, sizes either
side of the 64-lane wave and 256-star block, and both sides of a chunk.
"""
import numpy as np
import pytest

import oracle_bind as ob
from conftest import energy_atol, rel_err
from pn_amd import lib as pn
from pn_amd import workloads

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1000, 1537]
EPS, C, DT = 1e-3, 20.0, 2.0**-10


def make_clusters(sizes=SIZES, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for k, ns in enumerate(sizes):
        m0 = float(rng.uniform(0.005, 0.05))
        if ns < 4:
            # the samplers centre the stars on the BH: with one star that
            # puts it on the BH, which pn1 rejects
            x = np.vstack([np.zeros(3), rng.uniform(-1, 1, (ns, 3))])
            v = np.vstack([np.zeros(3), rng.normal(0, 0.3, (ns, 3))])
            m = np.concatenate([[m0], np.full(ns, 1.0 / max(ns, 1))])
            out.append((x, v, m))
        elif k % 2:
            out.append(pn.ic_plummer(ns, seed + k, a=1.0, m_bh=m0))
        else:
            out.append(pn.ic_uniform_sphere(ns, seed + k, R=1.0, Q=0.5, m_bh=m0, eps=EPS))
    return out


@pytest.fixture(scope="module")
def clusters():
    return make_clusters()


@pytest.fixture(scope="module")
def ens(clusters):
    n = sum(len(m) for _, _, m in clusters)
    e = pn.Ensemble(capacity=n + 64, max_clusters=len(clusters) + 4)
    yield e
    e.close()


@pytest.fixture(scope="module")
def ctx():
    c = pn.Context(capacity=2048)
    yield c
    c.close()


def single(ctx, cl, mode, what, steps=0):
    ctx.set_tuning(1, 0, -1)
    ctx.set_params(EPS, C, DT, mode)
    ctx.upload(*cl)
    if what == "force":
        return ctx.force()
    if what == "energy":
        return ctx.energy()
    ctx.step(steps)
    return ctx.download()


@pytest.mark.parametrize("mode", [pn.FP64, pn.MIXED])
def test_force_bitwise_equals_single_cluster(ens, ctx, clusters, mode):
    ens.set_tuning(4096, 0, -1)
    ens.set_params(EPS, C, DT, mode)
    ens.upload(clusters)
    assert ens.get_tuning()[0] == 4096
    got = ens.force()
    for k, cl in enumerate(clusters):
        want = single(ctx, cl, mode, "force")
        diff = np.abs(got[k] - want).max()
        print(f"mode {mode} cluster {k} n={len(cl[2])}: max |diff| {diff:.3e}")
        assert np.array_equal(got[k], want), (mode, k, diff)


def test_energy_bitwise_equals_single_cluster(ens, ctx, clusters):
    ens.set_tuning(4096, 0, -1)
    ens.set_params(EPS, C, DT, pn.FP64)
    ens.upload(clusters)
    got = ens.energy()
    assert got.shape == (len(clusters), 5)
    for k, cl in enumerate(clusters):
        want = single(ctx, cl, pn.FP64, "energy")
        assert np.array_equal(got[k], want), (k, got[k], want)


@pytest.mark.parametrize("mode", [pn.FP64, pn.MIXED])
def test_steps_bitwise_equal_single_cluster(ens, ctx, clusters, mode):
    ens.set_tuning(4096, 0, -1)
    ens.set_params(EPS, C, DT, mode)
    ens.upload(clusters)
    ens.step(3)
    xs, vs = ens.download()
    for k, cl in enumerate(clusters):
        x, v = single(ctx, cl, mode, "step", steps=3)
        assert np.array_equal(xs[k], x) and np.array_equal(vs[k], v), (mode, k)


def test_default_chunk_fp64_vs_oracle(ens, clusters):
    ens.set_tuning(0, 0, -1)
    ens.set_params(EPS, C, DT, pn.FP64)
    ens.upload(clusters)
    L = ens.get_tuning()[0]
    assert 256 <= L <= 512 and L % 64 == 0
    got = ens.force()
    for k, (x, v, m) in enumerate(clusters):
        want = ob.orc_force(x, v, m, EPS, C)
        err = rel_err(got[k], want)
        print(f"cluster {k} n={len(m)}: max rel {err.max():.3e}")
        assert err.max() <= 1e-12, (k, err.max())


def test_default_chunk_energy_and_steps_vs_oracle(ens, clusters):
    ens.set_tuning(0, 0, -1)
    ens.set_params(EPS, C, DT, pn.FP64)
    ens.upload(clusters)
    e = ens.energy()
    for k, (x, v, m) in enumerate(clusters):
        want = ob.orc_energy_terms(x, v, m, EPS, C)
        tol = energy_atol({"m": m, "eps": EPS}, want)
        assert np.abs(e[k] - want).max() <= tol, (k, e[k], want)
    ens.step(2)
    xs, vs = ens.download()
    for k, (x, v, m) in enumerate(clusters):
        if len(m) > 600:
            continue     # oracle time: the smaller clusters cover the path
        xo, vo = ob.orc_q1_steps(x, v, m, EPS, C, DT, 2)
        assert np.abs(xs[k] - xo).max() <= 1e-13 * max(1.0, np.abs(xo).max()), k
        assert np.abs(vs[k] - vo).max() <= 1e-13 * max(1.0, np.abs(vo).max()), k


@pytest.mark.parametrize("mode", [pn.FP64, pn.MIXED])
def test_repeat_evaluations_bitwise_identical(ens, clusters, mode):
    ens.set_tuning(128, 0, -1)
    ens.set_params(EPS, C, DT, mode)
    ens.upload(clusters)
    a1 = np.concatenate(ens.force())
    for _ in range(3):
        a2 = np.concatenate(ens.force())
        assert np.array_equal(a1, a2)


def test_mixed_accuracy(ens, clusters):
    ens.set_tuning(0, 0, -1)
    ens.set_params(EPS, C, DT, pn.FP64)
    ens.upload(clusters)
    a64 = ens.force()
    ens.set_params(EPS, C, DT, pn.MIXED)
    a32 = ens.force()
    for k in range(len(clusters)):
        if len(clusters[k][2]) < 2:
            continue
        err = rel_err(a32[k], a64[k])
        print(f"cluster {k} n={len(clusters[k][2])}: mixed star max {err[1:].max():.3e} "
              f"BH {err[0]:.3e}")
        assert err[1:].max() <= 1e-6 and err[0] <= 1e-6, (k, err[1:].max(), err[0])


HARD_K = 30.0          # PN1_HARD_K: a star whose sum m_j w^2 exceeds 30 |a| gets an FP64 row


def lattice_cluster(q, h, seed, bh_dist=40.0):
    """(2q+1)^3 equal stars on a jittered cubic lattice of spacing h around
    a star at its exact centre; the BH (mass 0.01) bh_dist away.  The
    centre star's star terms cancel to its neighbours' jitter and the far
    BH pulls little: its terms cancel far beyond HARD_K."""
    rng = np.random.default_rng(seed)
    g = np.arange(-q, q + 1) * h
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.argsort(np.abs(pts).sum(1), kind="stable")]     # centre first
    pts[1:] += rng.uniform(-0.02 * h, 0.02 * h, pts[1:].shape)
    ns = len(pts)
    x = np.vstack([[bh_dist, 0.0, 0.0], pts])
    v = np.vstack([np.zeros(3), rng.normal(0, 0.2, (ns, 3))])
    m = np.concatenate([[0.01], np.full(ns, 1.0 / ns)])
    return x, v, m


def cancel_ratio(x, m, a):
    """sum_j m_j / (r_ij^2 + eps^2) over the other stars, over |a_i|, per star."""
    xs = x[1:]
    d = xs[None, :, :] - xs[:, None, :]
    w2 = 1.0 / ((d * d).sum(-1) + EPS * EPS)
    np.fill_diagonal(w2, 0.0)
    return (w2 * m[None, 1:]).sum(1) / np.linalg.norm(a[1:], axis=1)


def test_hard_stars_of_many_clusters(ens, ctx, clusters):
    """Stars whose pair terms cancel (HARD_K) take FP64 rows of their own;
    the ensemble lists them for all its clusters in one pass.  Five lattice
    clusters, each with such stars, between ordinary ones: the mixed forces
    of every star are within 1e-6 of FP64, and with one chunk per cluster
    (an unsplit single context) bitwise the single-cluster results."""
    lat = [lattice_cluster(3, 0.1 + 0.01 * k, 60 + k) for k in range(5)]
    mix = [clusters[8], lat[0], clusters[10], lat[1], lat[2], clusters[12], lat[3], lat[4]]
    ens.set_tuning(4096, 0, -1)
    ens.set_params(EPS, C, DT, pn.FP64)
    ens.upload(mix)
    a64 = ens.force()
    for k, (x, _, m) in enumerate(mix):
        if k in (1, 3, 4, 6, 7):
            kr = cancel_ratio(x, m, a64[k])
            print(f"lattice cluster {k}: {int((kr > HARD_K).sum())} stars with "
                  f"sum m w^2 / |a| > {HARD_K:g} (max {kr.max():.3e})", flush=True)
            assert (kr > 3 * HARD_K).any()
    ens.set_params(EPS, C, DT, pn.MIXED)
    ens.upload(mix)
    ens.force()                       # thresholds and hard flags from this one
    a32 = ens.force()
    for k, cl in enumerate(mix):
        err = rel_err(a32[k], a64[k])
        print(f"cluster {k} n={len(cl[2])}: mixed vs FP64 star max {err[1:].max():.3e} "
              f"BH {err[0]:.3e}", flush=True)
        assert err[1:].max() <= 1e-6 and err[0] <= 1e-6, (k, err[1:].max(), err[0])
        ctx.set_tuning(1, 0, -1)
        ctx.set_params(EPS, C, DT, pn.MIXED)
        ctx.upload(*cl)
        ctx.force()
        assert np.array_equal(a32[k], ctx.force()), k


def test_more_hard_stars_than_split_rows():
    """1100 lattice clusters of 27 stars, each with a centre star whose terms
    cancel (sum m w^2 / |a| ~ 250): more hard stars than the 1024 whose rows
    are split in segments, so every row is taken whole.  Mixed within 1e-6
    of FP64 for every star, and repeated evaluations bitwise identical (the
    split must not depend on the order in which the clusters' workgroups
    list their hard stars)."""
    lat = [lattice_cluster(1, 0.1, 1000 + k) for k in range(1100)]
    x0, _, m0 = lat[0]
    kr = cancel_ratio(x0, m0, ob.orc_force(*lat[0], EPS, C))
    assert kr.max() > 3 * HARD_K
    n = sum(len(m) for _, _, m in lat)
    with pn.Ensemble(capacity=n, max_clusters=len(lat)) as e:
        e.set_params(EPS, C, DT, pn.FP64)
        e.upload(lat)
        a64 = np.concatenate(e.force())
        e.set_params(EPS, C, DT, pn.MIXED)
        e.upload(lat)
        e.force()
        a32 = np.concatenate(e.force())
        err = rel_err(a32, a64)
        print(f"{len(lat)} lattice clusters: mixed vs FP64 max {err.max():.3e}", flush=True)
        assert err.max() <= 1e-6
        for _ in range(2):
            assert np.array_equal(np.concatenate(e.force()), a32)


def test_graph_replay_equals_eager_profiled(ens, clusters):
    """pn1_ens_step replays a captured graph; with profiling on it launches
    eagerly.  Both must give the same bits."""
    ens.set_tuning(0, 0, -1)
    ens.set_params(EPS, C, DT, pn.MIXED)
    ens.upload(clusters)
    ens.step(4)
    xg, vg = ens.download()
    ens.upload(clusters)
    ens.profile_enable(True)
    ens.step(4)
    k, ms = ens.profile_read()
    ens.profile_enable(False)
    xe, ve = ens.download()
    assert k == 12 and ms > 0.0
    for a, b in zip(xg + vg, xe + ve):
        assert np.array_equal(a, b)


def test_upload_rejects_bad_layouts(ens, clusters):
    x, v, m = clusters[3]
    with pytest.raises(pn.Pn1Error):
        ens.upload([])
    with pytest.raises(pn.Pn1Error):
        ens.upload([(x, v, m)] * (ens.max_clusters + 1))
    bad = m.copy()
    bad[0] = 0.0
    with pytest.raises(pn.Pn1Error, match="cluster 1"):
        ens.upload([(x, v, m), (x, v, bad)])
    xl, vl, ml = clusters[-1]
    big = ens.capacity // len(ml) + 1
    assert big <= ens.max_clusters
    with pytest.raises(pn.Pn1Error, match="capacity"):
        ens.upload([(xl, vl, ml)] * big)
    # the previous good state is still usable after a rejected upload
    ens.upload(clusters[:2])
    assert len(ens.force()) == 2


# ---------------------------------------------------------------- adjoint
def test_force_vjp_bitwise_equals_single_cluster(ens, ctx, clusters):
    """With one chunk per cluster (and an unsplit single-context j range)
    the ensemble VJP runs each cluster's sums in the same order."""
    rng = np.random.default_rng(41)
    us = [rng.normal(size=(len(m), 3)) for _, _, m in clusters]
    ens.set_tuning(4096, 0, -1)
    ens.set_params(EPS, C, DT, pn.FP64)
    ens.upload(clusters)
    gxs, gvs, gm0 = ens.force_vjp(us)
    for k, cl in enumerate(clusters):
        ctx.set_tuning(1, 0, -1)
        ctx.set_params(EPS, C, DT, pn.FP64)
        ctx.upload(*cl)
        gx, gv, g0 = ctx.force_vjp(us[k])
        assert np.array_equal(gxs[k], gx) and np.array_equal(gvs[k], gv), k
        assert gm0[k] == g0, (k, gm0[k], g0)


def ens_loss_grads(ens, r_cut):
    xs, vs = ens.download()
    out = [workloads.radial_loss(x, v, r_cut) for x, v in zip(xs, vs)]
    return [o[1] for o in out], [o[2] for o in out]


@pytest.mark.parametrize("chunk", [4096, 0])
def test_adjoint_equals_single_cluster(ens, ctx, clusters, chunk):
    """Trajectory gradient of every cluster (its own radial loss) through the
    ensemble adjoint: bitwise the single-context result with matching
    shapes, within 1e-12 with the automatic chunk."""
    sub = clusters[:9]            # up to 257 stars: keeps the test short
    steps, interval = 6, 4
    ens.set_tuning(chunk, 0, -1)
    ens.set_params(EPS, C, DT, pn.MIXED)      # the adjoint runs FP64 regardless
    ens.upload(sub)
    ens.adjoint_forward(steps, interval)
    lx, lv = ens_loss_grads(ens, 0.8)
    gx0s, gv0s, gm0 = ens.adjoint_backward(lx, lv)
    for k, cl in enumerate(sub):
        ctx.set_tuning(1, 0, -1)
        ctx.set_params(EPS, C, DT, pn.FP64)
        ctx.upload(*cl)
        ctx.adjoint_forward(steps, interval)
        xt, vt = ctx.download()
        _, gx, gv = workloads.radial_loss(xt, vt, 0.8)
        gx0, gv0, g0 = ctx.adjoint_backward(gx, gv)
        if chunk:
            assert np.array_equal(gx0s[k], gx0) and np.array_equal(gv0s[k], gv0), k
            assert gm0[k] == g0, k
        else:
            scale = max(np.abs(gx0).max(), np.abs(gv0).max(), abs(g0), 1e-300)
            assert np.abs(gx0s[k] - gx0).max() <= 1e-12 * scale, k
            assert np.abs(gv0s[k] - gv0).max() <= 1e-12 * scale, k
            assert abs(gm0[k] - g0) <= 1e-12 * scale, k
