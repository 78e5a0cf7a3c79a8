"""Launch-shape tuner for the pair kernels (runs on the GPU through the C ABI).

    python -m pn_amd.tune [--workload plummer|sphere|ensemble] [--nstars N] [--seed S]
                          [--m0 M] [--eps E] [--c C] [--mode fp64|mixed|both]
                          [--reps R] [--json OUT]

For each launch shape (j split S, mixed-form compensation interval `tile`,
kernel variant) it times the pair kernel with HIP events over `reps` force
evaluations of the same state, and for the mixed form measures the
per-particle acceleration error against the FP64 form.  It prints one row
per shape and the fastest shape per mode whose mixed error stays within
1e-6 (max over stars), and can write everything as JSON.  The library's
automatic choice (pn1_set_tuning(ctx, 0, 0, -1)) is listed first.  With
--workload ensemble it sweeps the ensemble API's work-item chunk length
on the batched workload (256 Plummer clusters) instead.
"""
import argparse
import itertools
import json
import sys

import numpy as np

from . import lib as pn


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def time_shape(ctx, x, v, m, eps, c, mode, shape, reps):
    ctx.set_tuning(*shape)
    ctx.set_params(eps, c, 2.0**-12, mode)
    ctx.upload(x, v, m)
    ctx.profile_enable(True)
    ctx.force()                       # warm-up
    ctx.profile_enable(True)
    a = None
    for _ in range(reps):
        a = ctx.force()
    k, ms = ctx.profile_read()
    ctx.profile_enable(False)
    used = ctx.get_tuning()
    return a, ms / max(k, 1), used


def sweep(ctx, x, v, m, eps, c, modes, reps=3, jsplits=(8, 16, 32, 64, 128),
          tiles=(8, 16, 32, 64), variants=(0, 1), out=sys.stdout):
    n = len(m)
    rows = []
    a64, _, _ = time_shape(ctx, x, v, m, eps, c, pn.FP64, (0, 0, -1), 1)
    for mode in modes:
        name = "fp64" if mode == pn.FP64 else "mixed"
        shapes = [(0, 0, -1)]
        tl = (0,) if mode == pn.FP64 else tiles
        shapes += [(s, t, var) for s, t, var in itertools.product(jsplits, tl, variants)]
        for shape in shapes:
            a, ms, used = time_shape(ctx, x, v, m, eps, c, mode, shape, reps)
            err = rel_err(a, a64) if mode == pn.MIXED else np.zeros(n)
            row = dict(mode=name, requested=list(shape), jsplit=used[0], tile=used[1],
                       variant=used[2], ms=ms, interactions_per_s=float(n) * n / (ms * 1e-3),
                       err_p999=float(np.quantile(err[1:], 0.999)) if n > 1 else 0.0,
                       err_max=float(err[1:].max()) if n > 1 else 0.0, err_bh=float(err[0]))
            rows.append(row)
            print(f"{name:5s} S={used[0]:4d} tile={used[1]:3d} variant={used[2]} "
                  f"{'(auto)' if shape == (0, 0, -1) else '      '} {ms:9.3f} ms "
                  f"{row['interactions_per_s']:.3e} int/s  err p99.9 {row['err_p999']:.2e} "
                  f"max {row['err_max']:.2e}", file=out, flush=True)
    best = {}
    for name in {r["mode"] for r in rows}:
        ok = [r for r in rows if r["mode"] == name and r["err_max"] <= 1e-6]
        if ok:
            best[name] = min(ok, key=lambda r: r["ms"])
            b = best[name]
            print(f"best {name}: S={b['jsplit']} tile={b['tile']} variant={b['variant']} "
                  f"{b['ms']:.3f} ms", file=out, flush=True)
    return rows, best


def ens_sweep(ens, clusters, eps, c, mode, reps=3, chunks=(64, 128, 256, 512, 1024, 2048),
              tiles=(16, 32), variants=(0, 1), out=sys.stdout, shapes=None):
    """Ensemble form: times the persistent pair kernel per work-item chunk
    length (and, mixed, compensation interval and variant)."""
    inter = sum(float(len(m)) ** 2 for _, _, m in clusters)
    rows = []
    name = "fp64" if mode == pn.FP64 else "mixed"
    tl = (0,) if mode == pn.FP64 else tiles
    for shape in shapes or [(0, 0, -1)] + list(itertools.product(chunks, tl, variants)):
        ens.set_tuning(*shape)
        ens.set_params(eps, c, 2.0**-10, mode)
        ens.upload(clusters)
        ens.profile_enable(True)
        ens.force()
        ens.profile_enable(True)
        for _ in range(reps):
            ens.force()
        k, ms = ens.profile_read()
        ens.profile_enable(False)
        used = ens.get_tuning()
        ms /= max(k, 1)
        row = dict(mode=name, requested=list(shape), chunk=used[0], tile=used[1],
                   variant=used[2], ms=ms, interactions_per_s=inter / (ms * 1e-3))
        rows.append(row)
        print(f"{name:5s} chunk={used[0]:5d} tile={used[1]:3d} variant={used[2]} "
              f"{'(auto)' if shape == (0, 0, -1) else '      '} {ms:9.3f} ms "
              f"{row['interactions_per_s']:.3e} int/s", file=out, flush=True)
    b = min(rows, key=lambda r: r["ms"])
    print(f"best {name}: chunk={b['chunk']} tile={b['tile']} variant={b['variant']} "
          f"{b['ms']:.3f} ms", file=out, flush=True)
    return rows, b


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=("plummer", "sphere", "ensemble"), default="plummer")
    ap.add_argument("--nstars", type=int, default=131072)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--m0", type=float, default=0.05)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--c", type=float, default=50.0)
    ap.add_argument("--mode", choices=("fp64", "mixed", "both"), default="both")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--json", default=None)
    ap.add_argument("--shape", default=None, metavar="A,TILE,VARIANT",
                    help="time only this shape (A = j split, or chunk length with "
                         "--workload ensemble); for profiling one launch shape")
    args = ap.parse_args(argv)
    if args.workload == "ensemble":
        from . import workloads
        W = workloads.ENSEMBLE
        clusters = workloads.ensemble()
        ntot = sum(len(m) for _, _, m in clusters)
        modes = {"fp64": (pn.FP64,), "mixed": (pn.MIXED,), "both": (pn.FP64, pn.MIXED)}[args.mode]
        print(f"# ensemble of {len(clusters)} Plummer clusters ({ntot} particles), "
              f"eps={W['eps']} c={W['c']}, {args.reps} timed force evaluations per shape",
              flush=True)
        result = {}
        one = [tuple(int(t) for t in args.shape.split(","))] if args.shape else None
        with pn.Ensemble(ntot, len(clusters)) as ens:
            for mode in modes:
                rows, best = ens_sweep(ens, clusters, W["eps"], W["c"], mode, reps=args.reps,
                                       shapes=one)
                result["fp64" if mode == pn.FP64 else "mixed"] = dict(rows=rows, best=best)
        if args.json:
            with open(args.json, "w") as f:
                json.dump({"workload": vars(args), **result}, f, indent=1)
        return
    if args.workload == "plummer":
        x, v, m = pn.ic_plummer(args.nstars, args.seed, a=1.0, m_bh=args.m0)
    else:
        x, v, m = pn.ic_uniform_sphere(args.nstars, args.seed, R=1.0, Q=0.5, m_bh=args.m0,
                                       eps=args.eps)
    modes = {"fp64": (pn.FP64,), "mixed": (pn.MIXED,), "both": (pn.FP64, pn.MIXED)}[args.mode]
    print(f"# {args.workload} N={args.nstars} seed={args.seed} m0={args.m0} eps={args.eps} "
          f"c={args.c}, {args.reps} timed force evaluations per shape", flush=True)
    with pn.Context(len(m)) as ctx:
        rows, best = sweep(ctx, x, v, m, args.eps, args.c, modes, reps=args.reps)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"workload": vars(args), "rows": rows, "best": best}, f, indent=1)


if __name__ == "__main__":
    main()
