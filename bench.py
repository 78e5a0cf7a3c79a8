"""Headline benchmark: interactions/s of the force path at N = 131072.

Workload (one MI355X): a seeded Plummer sphere of 131072 stars (scale radius
1, total star mass 1) plus a central black hole of mass 0.05, speed of light
c = 50, Plummer softening eps = 1e-4, RK step dt = 2^-12.  A step is one step
of the reference's stage driver (Limit-sticky9.c:841-893): three force
evaluations of n^2 interactions each (n = 131073 counts the black hole), plus
the O(n) closed forms and the stage update, with the state resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--ens-steps K]

Prints ONE JSON line.  `value` is the mixed-precision form's interactions/s
(star-star pairs in FP32 on double-single positions, everything else FP64);
the FP64 form is measured the same way and reported beside it, with the
per-particle acceleration error of the mixed form against the FP64 form at
the same state (the initial state and two FP64 steps on).  The timed steps
replay the captured step graph, as pn1_step runs them; the pair kernel's
average launch duration comes from a separate pass of 3 steps after them,
launched one by one with HIP events around every pair launch on the
library's stream.  Algorithmic work: 90 flops per interaction (FMA = 2,
rsqrt = 1; Newtonian 20 + three-body 1PN 70).

CPU legs (skipped with --no-cpu), on the host's cores beside the GPU legs:
the reference's own force() (compiled from its sources into oracle/_ref)
on one core for the first 24576 stars (the bounded sample, `cpu_baseline`)
and at n = 4097, 16385, 65537 for an n^2 fit and its projection to the full
n; the oracle restatement over OpenMP threads (`cpu_openmp`); and, as the
checker, the oracle's force rows at a sample of stars against both GPU
forms (`accuracy_vs_oracle_rows`).

The line also carries an "ensemble" object: the batched workload of 256
seeded Plummer clusters (star counts log-uniform in [256, 8192]) advanced
together through the ensemble API for --ens-steps steps (default 128), its
interactions/s (sum of n^2 per force evaluation), pair-kernel roofline and
mixed-vs-FP64 error, against the same clusters stepped one at a time, and
the adjoint of every cluster's BH-mass gradient over the same steps,
an "adjoint" object: the differentiation workload (N = 16384, 256 steps,
checkpoints every 16) forward and backward wall times and the force VJP's
pair rate (FP64), and a "shard_of_8" object: one GPU's share of a run whose
stars are split across 8 GPUs (16384 stars against all 131073 particles,
pn1_force_range), its pair kernel time and fraction of the full launch.
"""
import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pn_amd import lib as pn  # noqa: E402
from pn_amd import workloads  # noqa: E402

METRIC = ("interactions/s & % vector peak at N=128k, mixed vs FP64 GPU @ acc rel err"
          "≤1e-6")
N_STARS, SEED, M_BH, C_LIGHT, EPS, DT = 131072, 3, 0.05, 50.0, 1e-4, 2.0**-12
FLOPS_PER_INTERACTION = 90
PEAK_FP32 = 157.3e12   # MI355X vector FP32
PEAK_FP64 = 78.6e12    # MI355X vector FP64
CPU_SAMPLE_STARS = 24576
CPU_FIT_SIZES = (4097, 16385, 65537)
CPU_OMP_SAMPLE_STARS = 49152
ENS_TRAFFIC = "ensemble_pair_traffic_c5.json"


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def run_form(ctx, x, v, m, mode, steps, warmup, autotune=True, prof_steps=3):
    """`steps` timed steps (graph-replayed, as pn1_step runs them) after
    `warmup`; then, outside the timed region, `prof_steps` steps launched
    one by one with HIP events around every pair launch on the library's
    stream, for the pair kernel's average duration."""
    ctx.set_params(EPS, C_LIGHT, DT, mode)
    ctx.upload(x, v, m)
    tuned = ctx.autotune(2) if autotune else None     # before the timed region
    e0 = ctx.energy()
    if warmup:
        ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    t1 = time.perf_counter()
    e1 = ctx.energy()
    ctx.profile_enable(True)
    ctx.step(prof_steps)
    ctx.sync()
    launches, pair_ms = ctx.profile_read()
    ctx.profile_enable(False)
    n = len(m)
    wall = t1 - t0
    return dict(
        interactions_per_s=3.0 * n * n * steps / wall,
        ms_per_step=1e3 * wall / steps,
        pair_launches=launches,
        pair_avg_ms=pair_ms / max(launches, 1),
        energy_drift=float((e1[4] - e0[4]) / abs(e0[4])),
        shape=ctx.get_tuning(),
        autotune_best_ms=tuned[2] if tuned else None,
    )


def roofline(n, avg_ms, peak, traffic=None, valu=None):
    flops = FLOPS_PER_INTERACTION * float(n) * n
    achieved = flops / (avg_ms * 1e-3) / 1e12
    out = {"bound": "valu", "achieved": round(achieved, 3), "peak": peak / 1e12,
           "unit": "TFLOP/s", "frac": round(achieved * 1e12 / peak, 4), "traffic": traffic}
    if valu:
        # the FLOPs of the VALU instructions the kernel issues (rocprofv3 counters,
        # profiles/pair_kernel_valu_n131073.txt), beside the algorithmic figure
        issued = valu["valu_flops_f32"] + valu["valu_flops_f64"]
        out["valu_counters"] = {"tflops": round(issued / (avg_ms * 1e-3) / 1e12, 3),
                                "flops_per_interaction": round(issued / (float(n) * n), 2),
                                "instructions_per_interaction":
                                    round(64.0 * valu["valu_wave_instructions"] / (float(n) * n), 2)}
    return out


def sources_sha256():
    """Hash of the library's sources (pn_amd/csrc/*) and of its build recipe
    (pn_amd/build.py: compiler flags), which every committed counter summary
    records: counters measured on another build are not reported as this
    build's."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "pn_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".hpp", ".inc", ".h", ".c")):
            with open(os.path.join(d, name), "rb") as f:
                h.update(name.encode() + b"\0" + f.read())
    with open(os.path.join(ROOT, "pn_amd", "build.py"), "rb") as f:
        h.update(b"build.py\0" + f.read())
    return h.hexdigest()


def load_profile(name):
    """A committed counter summary, if it was measured on these sources."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    return d if d.get("sources_sha256") == sources_sha256() else None


def load_valu(kind, shape):
    """Measured VALU work per launch (rocprofv3 SQ_INSTS_VALU_* counters), if
    profiled on these sources at this launch shape."""
    d = load_profile("pair_kernel_valu_n131073.json")
    if d is None:
        return None
    want = d.get("shape", {}).get(kind, {})
    if any(shape.get(k) != v for k, v in want.items()):
        return None
    return d.get(kind)


def load_traffic(kind, name="pair_kernel_traffic_n131073.json", shape=None):
    """HBM bytes per pair launch from a committed rocprofv3 --pmc summary,
    if it was measured on these sources at this launch shape (the partial
    sums written scale with the j split / chunk length); else None."""
    d = load_profile(name)
    if d is None:
        return None
    if shape is not None and "chunk" in shape and "by_chunk" in d:
        row = d["by_chunk"].get(kind, {}).get(str(shape["chunk"]))
        return row["total"] if row else None
    want = d.get("shape", {}).get(kind)
    if shape is not None and want is not None:
        if any(shape.get(k) != v for k, v in want.items()):
            return None
    return d.get(kind)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    return ob


def _cpu_name():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def cpu_force_timer(x, v, m):
    """(kind, t(n)): the wall time of one call of the reference's own force()
    (compiled from its sources into oracle/_ref) on the first n particles of
    the workload, one core; the oracle restatement if that library is
    absent.  The reference has no speed of light: it runs on v/c, m/c^2
    (the same work)."""
    ob = _oracle()
    kind = "reference" if ob.have_ref() else "port"

    def t(n):
        xs, vs, ms = x[:n].copy(), v[:n].copy() / C_LIGHT, m[:n].copy() / C_LIGHT**2
        t0 = time.perf_counter()
        if kind == "reference":
            ob.ref_force(xs, vs, ms, EPS)
        else:
            ob.orc_force(xs, vs, ms, EPS, 1.0)
        dt = time.perf_counter() - t0
        print(f"# cpu {kind} force() n={n}: {dt:.2f} s", file=sys.stderr, flush=True)
        return dt
    return kind, t


def cpu_baseline(x, v, m, fit):
    """One force() call on the first CPU_SAMPLE_STARS stars (the bounded
    sample), beside `fit`: the n^2 fit over CPU_FIT_SIZES, measured
    concurrently (cpu_fit) on another core."""
    kind, t = cpu_force_timer(x, v, m)
    n = CPU_SAMPLE_STARS + 1
    dt = t(n)
    return {"value": float(n) * n / dt, "unit": "interactions/s", "cores": 1, "kind": kind,
            "sample": f"one force() call, first {CPU_SAMPLE_STARS} stars + BH of the workload "
                      f"(n={n}, {n * n:.3e} interactions, {dt:.1f} s, gcc -O2, 1 thread, "
                      f"{_cpu_name()})",
            "n2_fit": fit}


def cpu_fit(x, v, m, out):
    """t(n) = a n^2 + b n + c through one force() call at each of
    CPU_FIT_SIZES (one core), and its value at the workload's full n."""
    kind, t = cpu_force_timer(x, v, m)
    ns = np.array(CPU_FIT_SIZES, dtype=float)
    ts = np.array([t(int(k)) for k in CPU_FIT_SIZES])
    a, b, c = np.linalg.solve(np.stack([ns**2, ns, np.ones_like(ns)], 1), ts)
    nf = float(N_STARS + 1)
    tf = a * nf * nf + b * nf + c
    out.update({"kind": kind, "cores": 1, "sizes": [int(k) for k in CPU_FIT_SIZES],
                "seconds": [float(q) for q in ts],
                "interactions_per_s": [float(k * k / q) for k, q in zip(ns, ts)],
                "fit": "t(n) = a n^2 + b n + c through the three calls",
                "a_ns_per_interaction": float(a * 1e9), "b": float(b), "c": float(c),
                "full_n": int(nf), "full_n_seconds": float(tf),
                "full_n_interactions_per_s": float(nf * nf / tf)})


def oracle_rows_check(label, a, x, v, m, seed, k=40):
    """The oracle's force rows (orc_force_rows: the FP64 CPU restatement,
    each row O(n)) at a sample of stars: the near-BH stars, the 256-star
    block edges and seeded random stars; max per-star relative error of `a`
    there."""
    ob = _oracle()
    n = len(m)
    d = np.linalg.norm(x[1:] - x[0], axis=1)
    rows = np.unique(np.concatenate([[1, 2, 255, 256, 257, n - 1], 1 + np.argsort(d)[:8],
                                     np.random.default_rng(seed).integers(1, n, k)]))
    rows = rows[(rows >= 1) & (rows < n)].astype(np.int32)
    ref = ob.orc_force_rows(x, v, m, EPS, C_LIGHT, rows)
    err = np.linalg.norm(a[rows] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"# {label}: {len(rows)} oracle rows, max rel {err.max():.3e}", file=sys.stderr,
          flush=True)
    return float(err.max()), int(len(rows))


def cpu_openmp(x, v, m, reserve=0):
    """Beside the one-core baseline: the oracle's force() restatement with its
    O(n^2) loops over OpenMP threads (one per core of this process's CPU
    share, at most 16, less `reserve` cores busy with other calls), on the
    first CPU_OMP_SAMPLE_STARS stars.  Not the reference (which is serial);
    its star rows are bitwise the serial ones."""
    ob = _oracle()
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores) - reserve)
    n = CPU_OMP_SAMPLE_STARS + 1
    xs, vs, ms = x[:n].copy(), v[:n].copy(), m[:n].copy()
    t0 = time.perf_counter()
    _, ran = ob.orc_force_omp(xs, vs, ms, EPS, C_LIGHT, threads)
    dt = time.perf_counter() - t0
    print(f"# cpu OpenMP force() n={n}: {dt:.2f} s on {ran} threads", file=sys.stderr,
          flush=True)
    return {"value": float(n) * n / dt, "unit": "interactions/s", "cores": int(ran),
            "kind": "port",
            "sample": f"one force() call of the oracle restatement over {ran} OpenMP threads, "
                      f"first {CPU_OMP_SAMPLE_STARS} stars + BH of the workload (n={n}, "
                      f"{n * n:.3e} interactions, {dt:.1f} s, gcc -O2 -fopenmp)"}


def time_steps(obj, steps, warmup, eager=False):
    """Wall time of `steps` steps after `warmup` (graph-replayed steps, or
    with eager=True the same launches one by one)."""
    obj.profile_enable(eager)
    obj.step(warmup)
    obj.sync()
    t0 = time.perf_counter()
    obj.step(steps)
    obj.sync()
    wall = time.perf_counter() - t0
    obj.profile_enable(False)
    return wall


def run_ensemble(steps, warmup, chunk=0, tile=0, variant=-1, eager=False, adj_steps=0,
                 autotune=True):
    """The batched workload: 256 seeded Plummer clusters (star counts
    log-uniform in [256, 8192], BH masses uniform in [0.005, 0.05], c = 100,
    eps = 1e-3, dt = 2^-10), all advanced by one pn1_ens_step per step (one
    persistent pair launch per force evaluation, graph-captured steps),
    against the same clusters advanced one after another through a single
    context (pn1_step, also graph-captured)."""
    W = workloads.ENSEMBLE
    cl = workloads.ensemble()
    inter = workloads.interactions(cl)
    ntot = sum(len(m) for _, _, m in cl)
    res = {}
    with pn.Ensemble(capacity=ntot, max_clusters=len(cl)) as ens:
        ens.set_tuning(chunk, tile, variant)
        shapes = {}
        for mode, name in ((pn.FP64, "fp64"), (pn.MIXED, "mixed")):
            ens.set_tuning(chunk, tile, variant)
            ens.set_params(W["eps"], W["c"], W["dt"], mode)
            ens.upload(cl)
            if autotune:
                ens.autotune(4)                     # before the timed region
            shapes[name] = ens.get_tuning()
            e0 = ens.energy()
            wall = time_steps(ens, steps, warmup, eager)
            e1 = ens.energy()
            # pair-kernel time: HIP events around each persistent pair launch
            # of 3 steps (9 launches), after one untimed step
            ens.upload(cl)                          # keeps the tuned shape
            ens.step(1)
            ens.profile_enable(True)
            ens.step(3)
            k, ms = ens.profile_read()
            ens.profile_enable(False)
            drift = np.abs((e1[:, 4] - e0[:, 4]) / np.abs(e0[:, 4]))
            res[name] = dict(interactions_per_s=3.0 * inter * steps / wall,
                             ms_per_step=1e3 * wall / steps, pair_avg_ms=ms / max(k, 1),
                             energy_drift_max=float(drift.max()),
                             energy_drift_median=float(np.median(drift)))
        ens.set_params(W["eps"], W["c"], W["dt"], pn.FP64)
        ens.upload(cl)
        a64 = np.concatenate(ens.force())
        ens.set_params(W["eps"], W["c"], W["dt"], pn.MIXED)
        a32 = np.concatenate(ens.force())
        # the BH-mass gradient of every cluster's radial loss (FP64 adjoint)
        adj = None
        if adj_steps > 0:
            ens.upload(cl)
            t0 = time.perf_counter()
            ens.adjoint_forward(adj_steps, 16)
            xs, vs = ens.download()
            t1 = time.perf_counter()
            lg = [workloads.radial_loss(x, v, 0.2) for x, v in zip(xs, vs)]
            _, _, gm0 = ens.adjoint_backward([g[1] for g in lg], [g[2] for g in lg])
            t2 = time.perf_counter()
            adj = {"steps": adj_steps, "checkpoint_interval": 16,
                   "loss": "per cluster: mean squared radial velocity within r < 0.2 of its BH",
                   "forward_s": t1 - t0, "backward_s": t2 - t1,
                   "grad_m0_median_abs": float(np.median(np.abs(gm0))),
                   "grad_m0_max_abs": float(np.abs(gm0).max()),
                   # clusters where a star passes close to the BH, whose 1PN term
                   # has an unsoftened 1/|d|^2 (Limit-sticky9.c:74): the step
                   # map's derivative there is huge, and so is the gradient
                   "clusters_grad_m0_above_1e3": int((np.abs(gm0) > 1e3).sum())}
    err = rel_err(a32, a64)
    sep_wall = 0.0
    with pn.Context(capacity=max(len(m) for _, _, m in cl)) as ctx:
        ctx.set_params(W["eps"], W["c"], W["dt"], pn.MIXED)
        for x, v, m in cl:
            ctx.upload(x, v, m)
            sep_wall += time_steps(ctx, steps, warmup, eager)
    sep = 3.0 * inter * steps / sep_wall
    mixed = res["mixed"]
    return {
        "workload": f"{len(cl)} Plummer clusters, N log-uniform in [{W['n_min']}, {W['n_max']}] "
                    f"(seed {W['seed']}), BH mass U[{W['m_bh_min']}, {W['m_bh_max']}], "
                    f"c={W['c']:g}, eps={W['eps']:g}, dt=2^-10, {steps} steps, mixed",
        "particles": ntot, "interactions_per_force": inter,
        "interactions_per_s": mixed["interactions_per_s"],
        "ms_per_step": mixed["ms_per_step"],
        "separate_interactions_per_s": sep,
        "ensemble_over_separate": mixed["interactions_per_s"] / sep,
        "roofline": roofline_flops(inter, mixed["pair_avg_ms"], PEAK_FP32,
                                   load_traffic("mixed", ENS_TRAFFIC,
                                                {"chunk": shapes["mixed"][0],
                                                 "variant": shapes["mixed"][2]})),
        "fp64": {"interactions_per_s": res["fp64"]["interactions_per_s"],
                 "roofline": roofline_flops(inter, res["fp64"]["pair_avg_ms"], PEAK_FP64,
                                            load_traffic("fp64", ENS_TRAFFIC,
                                                         {"chunk": shapes["fp64"][0],
                                                          "variant": shapes["fp64"][2]})),
                 "energy_drift_max": res["fp64"]["energy_drift_max"]},
        "mixed_accuracy_vs_fp64": {"rel_err_max": float(err.max()),
                                   "rel_err_p999": float(np.quantile(err, 0.999)),
                                   "energy_drift_max": mixed["energy_drift_max"],
                                   "energy_drift_median": mixed["energy_drift_median"]},
        "tuning": {"mixed": list(shapes["mixed"]), "fp64": list(shapes["fp64"]),
                   "chosen_by": "pn1_ens_autotune" if autotune else "automatic chunk"},
        "adjoint_bh_mass": adj,
    }


def run_adjoint(steps, interval, vjp_reps=5):
    """The differentiation workload (workloads.ADJOINT: N = 16384 Plummer,
    seed 4, + BH 0.02, c = 100, eps = 1e-3, dt = 2^-10): `steps` steps forward
    with a checkpoint every `interval`, the radial-velocity loss at r < 0.2,
    and the discrete adjoint back to d loss / d (x0, v0, m0).  Also the force
    VJP alone (pn1_force_vjp: one O(n^2) reverse-sweep pass over both sides of
    every pair) against one FP64 force evaluation at the same state."""
    W = workloads.ADJOINT
    x, v, m = pn.ic_plummer(W["n_stars"], W["seed"], a=1.0, m_bh=W["m_bh"])
    n = len(m)
    with pn.Context(capacity=n, device=0) as ctx:
        ctx.set_params(W["eps"], W["c"], W["dt"], pn.FP64)
        ctx.upload(x, v, m)
        t0 = time.perf_counter()
        ctx.adjoint_forward(steps, interval)
        xt, vt = ctx.download()
        t1 = time.perf_counter()
        loss, lx, lv = workloads.radial_loss(xt, vt, W["r_cut"])
        gx0, gv0, gm0 = ctx.adjoint_backward(lx, lv)
        t2 = time.perf_counter()
        u = np.random.default_rng(1).normal(size=x.shape)
        ctx.force_vjp(u)
        ctx.force()
        ta = time.perf_counter()
        for _ in range(vjp_reps):
            ctx.force_vjp(u)
        tb = time.perf_counter()
        for _ in range(vjp_reps):
            ctx.force()
        tc = time.perf_counter()
    vjp_ms = 1e3 * (tb - ta) / vjp_reps
    force_ms = 1e3 * (tc - tb) / vjp_reps
    return {
        "workload": f"Plummer N={W['n_stars']} (seed {W['seed']}) + BH m={W['m_bh']}, "
                    f"c={W['c']:g}, eps={W['eps']:g}, dt=2^-10, {steps} steps, checkpoint every "
                    f"{interval}; loss: mean squared radial velocity within r < {W['r_cut']}",
        "forward_s": t1 - t0, "backward_s": t2 - t1,
        "backward_over_forward": (t2 - t1) / (t1 - t0),
        "loss": loss, "grad_m0": gm0,
        "grad_x0_max": float(np.abs(gx0).max()), "grad_v0_max": float(np.abs(gv0).max()),
        "vjp_ms": vjp_ms, "force_ms": force_ms,
        "vjp_pairs_per_s": float(n) * n / (vjp_ms * 1e-3),
        "vjp_over_force_time": vjp_ms / force_ms,
        "dtype": "f64",
    }


def run_shard(ctx, x, v, m, shards=8, reps=5):
    """One GPU's share of a run whose stars are split across `shards` GPUs
    (every particle's state all-gathered first): the pair kernel over the
    first n/shards stars, in whole 256-star blocks, against all n particles,
    timed alone on this GPU through pn1_force_range (HIP events around the
    pair launch)."""
    n = len(m)
    i1 = 1 + ((n - 1) // shards + 255) // 256 * 256
    out = {"shards": shards, "i_stars": i1 - 1, "j_particles": n}
    for mode, name in ((pn.MIXED, "mixed"), (pn.FP64, "fp64")):
        ctx.set_params(EPS, C_LIGHT, DT, mode)
        ctx.upload(x, v, m)
        ctx.force_range(1, i1)                      # warm-up
        ctx.profile_enable(True)
        for _ in range(reps):
            ctx.force_range(1, i1)
        k, ms = ctx.profile_read()
        ctx.profile_enable(False)
        pair_ms = ms / max(k, 1)
        out[name] = {"pair_ms": pair_ms,
                     "interactions_per_s": (i1 - 1) * n / (pair_ms * 1e-3)}
    return out


def run_drift(x, v, m, steps, every=100, track=10):
    """Relative energy drift (E_t - E_0) / |E_0| of both forms over `steps`
    steps of the headline workload, sampled every `every` steps, and an
    encounter log: every `track` steps the state is downloaded, and each time
    the fastest star's speed jumps by more than 20% the step, the star, its
    speed and its closest sampled approach to the BH are recorded (the
    energy jumps of this workload are such ejections out of close BH
    passages that dt does not resolve).  A third run, FP64 from the same
    state with every star's position moved by one ulp, shows how far those
    unresolved encounters alone spread the drift."""
    out = {"workload": "Plummer N=131072 + BH m=0.05, c=50, eps=1e-4, dt=2^-12",
           "steps": steps, "every": every, "track": track}
    n = len(m)
    xp = x.copy()
    xp[1:] = np.nextafter(xp[1:], np.inf)
    with pn.Context(capacity=n, device=0) as ctx:
        for mode, name, x0 in ((pn.FP64, "fp64", x), (pn.MIXED, "mixed", x),
                               (pn.FP64, "fp64_one_ulp", xp)):
            ctx.set_params(EPS, C_LIGHT, DT, mode)
            ctx.upload(x0, v, m)
            e0 = ctx.energy()
            series, events, dmin = [], [], {}
            vmax_prev = None
            t0 = time.perf_counter()
            for k in range(0, steps, track):
                ctx.step(min(track, steps - k))
                t = k + min(track, steps - k)
                xt, vt = ctx.download()
                d = np.linalg.norm(xt[1:] - xt[0], axis=1)
                sp = np.linalg.norm(vt[1:], axis=1)
                i, j = int(np.argmin(d)), int(np.argmax(sp))
                dmin[i + 1] = min(dmin.get(i + 1, np.inf), float(d[i]))
                if vmax_prev is not None and sp[j] > 1.2 * vmax_prev:
                    events.append({"step": t, "star": j + 1, "speed": float(sp[j]),
                                   "closest_sampled_approach": dmin.get(j + 1)})
                vmax_prev = float(sp[j])
                if t % every == 0 or t == steps:
                    e = ctx.energy()
                    series.append(float((e[4] - e0[4]) / abs(e0[4])))
                    print(f"# {name} step {t}: drift {series[-1]:+.6e}", file=sys.stderr,
                          flush=True)
            out[name] = {"drift": series, "final": series[-1], "E0": float(e0[4]),
                         "ejections": events, "wall_s": time.perf_counter() - t0}
    out["mixed_over_fp64_final"] = out["mixed"]["final"] / out["fp64"]["final"]
    out["mixed_minus_fp64_final"] = out["mixed"]["final"] - out["fp64"]["final"]
    out["fp64_one_ulp_minus_fp64_final"] = out["fp64_one_ulp"]["final"] - out["fp64"]["final"]
    return out


def roofline_flops(inter, avg_ms, peak, traffic=None):
    achieved = FLOPS_PER_INTERACTION * inter / (avg_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": peak / 1e12,
            "unit": "TFLOP/s", "frac": round(achieved * 1e12 / peak, 4), "traffic": traffic}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--jsplit", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--ens-steps", type=int, default=128,
                    help="steps of the batched-ensemble workload (0 skips it)")
    ap.add_argument("--ens-chunk", type=int, default=0)
    ap.add_argument("--autotune", action="store_true",
                    help="pick the headline pair kernels' launch shapes with pn1_autotune "
                         "(default: the automatic plan, whose HBM traffic is profiled)")
    ap.add_argument("--no-ens-autotune", action="store_true",
                    help="use the ensemble's automatic chunk instead of pn1_ens_autotune")
    ap.add_argument("--adj-steps", type=int, default=256,
                    help="steps of the differentiation workload (0 skips it)")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel of the ensemble leg directly instead of "
                         "replaying captured step graphs (for profilers)")
    ap.add_argument("--drift-steps", type=int, default=0,
                    help="only measure the energy drift of both forms over this many steps "
                         "of the headline workload, print it as JSON and exit")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB between two device buffers with torch's copy "
                         "kernel, 3 times (a workload of known HBM traffic for calibrating "
                         "profiler byte counters), and exit")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        src = torch.zeros(args.calib_copy << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": args.calib_copy << 20, "reps": 3}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this path runs on one GPU; there is no multi-GPU mode "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    x, v, m = pn.ic_plummer(N_STARS, SEED, a=1.0, m_bh=M_BH)
    n = len(m)
    if args.drift_steps:
        print(json.dumps(run_drift(x, v, m, args.drift_steps)), flush=True)
        return
    # the CPU baseline's calls run on their own cores beside the GPU legs
    fit, sample, threads = {}, {}, []
    if not args.no_cpu:
        old_stack = threading.stack_size(256 << 20)
        threads = [threading.Thread(target=cpu_fit, args=(x, v, m, fit), daemon=True),
                   threading.Thread(target=lambda: sample.update(cpu_baseline(x, v, m, fit)),
                                    daemon=True)]
        for t in threads:
            t.start()
        threading.stack_size(old_stack)
    with pn.Context(capacity=n, device=0) as ctx:
        ctx.set_tuning(args.jsplit, args.tile, args.variant)
        # accuracy of the mixed form: at the initial state (FP32-exact
        # positions) and after two FP64 steps (positions with all 53 bits)
        states = {"initial": (x, v)}
        ctx.set_params(EPS, C_LIGHT, DT, pn.FP64)
        ctx.upload(x, v, m)
        ctx.step(2)
        states["two_steps"] = ctx.download()
        forces = {}
        for name, (xs, vs) in states.items():
            for mode in (pn.FP64, pn.MIXED):
                ctx.set_params(EPS, C_LIGHT, DT, mode)
                ctx.upload(xs, vs, m)
                forces[name, mode] = ctx.force()
        err = rel_err(forces["initial", pn.MIXED], forces["initial", pn.FP64])
        err2 = rel_err(forces["two_steps", pn.MIXED], forces["two_steps", pn.FP64])
        print(f"# mixed vs FP64: star max {err[1:].max():.3e} (initial), "
              f"{err2[1:].max():.3e} (two steps on)", file=sys.stderr, flush=True)
        tune = args.jsplit == 0 and args.variant == -1 and args.autotune
        fp64 = run_form(ctx, x, v, m, pn.FP64, args.steps, args.warmup, tune)
        print(f"# FP64: {fp64['ms_per_step']:.2f} ms/step, pair {fp64['pair_avg_ms']:.3f} ms",
              file=sys.stderr, flush=True)
        mixed = run_form(ctx, x, v, m, pn.MIXED, args.steps, args.warmup, tune)
        print(f"# mixed: {mixed['ms_per_step']:.2f} ms/step, pair {mixed['pair_avg_ms']:.3f} ms",
              file=sys.stderr, flush=True)
        jsplit, tile, variant = mixed["shape"]
        shard = run_shard(ctx, x, v, m)
        for name, form in (("mixed", mixed), ("fp64", fp64)):
            shard[name]["fraction_of_full_launch"] = shard[name]["pair_ms"] / form["pair_avg_ms"]

    out = {
        "metric": METRIC,
        "value": mixed["interactions_per_s"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairs on double-single f32 positions; f64 BH-star 1PN, RK update, energy",
        "data": "synthetic: seeded Plummer sphere (AHW 1974 sampling, seed 3), FP32-exact",
        "config": {"workload": "Plummer N=131072 + BH m=0.05, c=50, eps=1e-4, dt=2^-12, "
                               "3 force evaluations per step",
                   "n": n, "interactions_per_step": 3 * n * n, "jsplit": jsplit, "tile": tile,
                   "variant": variant, "fp64_shape": list(fp64["shape"]),
                   "shape_chosen_by": "pn1_autotune" if tune else "fixed / automatic plan"},
        "roofline": roofline(n, mixed["pair_avg_ms"], PEAK_FP32,
                             load_traffic("mixed", shape={"jsplit": jsplit, "variant": variant,
                                                          "tile": tile}),
                             load_valu("mixed", {"jsplit": jsplit, "variant": variant,
                                                 "tile": tile})),
        "fp64": {"interactions_per_s": fp64["interactions_per_s"],
                 "ms_per_step": fp64["ms_per_step"],
                 "roofline": roofline(n, fp64["pair_avg_ms"], PEAK_FP64,
                                      load_traffic("fp64", shape={"jsplit": fp64["shape"][0],
                                                                  "variant": fp64["shape"][2]}),
                                      load_valu("fp64", {"jsplit": fp64["shape"][0],
                                                         "variant": fp64["shape"][2]})),
                 "energy_drift": fp64["energy_drift"]},
        "mixed_over_fp64": mixed["interactions_per_s"] / fp64["interactions_per_s"],
        "mixed_accuracy_vs_fp64": {"star_rel_err_max": float(err[1:].max()),
                                   "star_rel_err_p999": float(np.quantile(err[1:], 0.999)),
                                   "bh_rel_err": float(err[0]),
                                   "two_steps_on": {
                                       "star_rel_err_max": float(err2[1:].max()),
                                       "star_rel_err_p999": float(np.quantile(err2[1:], 0.999)),
                                       "bh_rel_err": float(err2[0])},
                                   "energy_drift_mixed": mixed["energy_drift"]},
        "pair_kernel_avg_ms": {"mixed": mixed["pair_avg_ms"], "fp64": fp64["pair_avg_ms"],
                               "launches": mixed["pair_launches"]},
        # one GPU's i shard of an 8-GPU run at this size (stars split, j all-gathered)
        "shard_of_8": shard,
        "flops_per_interaction": FLOPS_PER_INTERACTION,
        # the literature's Newtonian-only convention (20 flops per interaction)
        "tflops_newtonian_convention": {"mixed": 20.0 * mixed["interactions_per_s"] / 1e12,
                                        "fp64": 20.0 * fp64["interactions_per_s"] / 1e12},
    }
    if args.ens_steps > 0:
        print("# ensemble leg", file=sys.stderr, flush=True)
        out["ensemble"] = run_ensemble(args.ens_steps, args.warmup, chunk=args.ens_chunk,
                                       tile=args.tile, variant=args.variant, eager=args.eager,
                                       adj_steps=args.ens_steps if args.adj_steps > 0 else 0,
                                       autotune=not args.no_ens_autotune and args.ens_chunk == 0)
    if args.adj_steps > 0:
        print("# adjoint leg", file=sys.stderr, flush=True)
        out["adjoint"] = run_adjoint(args.adj_steps, min(16, args.adj_steps))
    if not args.no_cpu:
        # the checker: the oracle's rows against both forms
        acc = {}
        for name, (xs, vs) in states.items():
            for mode, label in ((pn.MIXED, "mixed"), (pn.FP64, "fp64")):
                e, k = oracle_rows_check(f"{label} {name}", forces[name, mode], xs, vs, m,
                                         seed=len(acc))
                acc[f"{label}_{name}"] = e
        acc["rows_per_state"] = k
        out["accuracy_vs_oracle_rows"] = acc
        threads[1].join()
        out["cpu_baseline"] = sample
        out["cpu_openmp"] = cpu_openmp(x, v, m, reserve=1 if threads[0].is_alive() else 0)
        threads[0].join()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
