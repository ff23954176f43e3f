"""Train the learned preconditioner on the GPU and save its weights.

Dataset: pressure systems (Laplacian, mu = -1, Dirichlet wall at x = 0) on
n^3 hex meshes with interior vertices perturbed by +-20% (one mesh seed per
system), plus graded slabs that carry the near-wall structure of the 192^3
graded mesh (gnn.graded_slab); the (residual, error) pairs are IC(0)-PCG
iterates of A x = 0 from seeded random starts after the listed iteration
counts (exact errors e = -x).  Training: Adam on the energy-norm loss
(z - e)^T A (z - e) / (e^T A e), mini-batches of 8 pairs, cosine-decayed
learning rate.  All systems stay resident on the GPU during training.

    python -m freecappuccino_amd.train_gnn --n 64 --systems 32 --graded 16 --steps 3000 --out weights.npy

With --stream COUNT the systems are streamed instead (stream_train): each new one is
assembled and sampled on the GPU, joins a window of resident systems, the oldest leaves,
and Adam steps follow -- the form that scales to thousands of systems:

    python -m freecappuccino_amd.train_gnn --stream 2000 --window 8 --steps-per-system 2 --out w.npy --log w.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import fcap, gnn
from . import workloads as W


def stream_train(net: gnn.Network, ctx: fcap.Context, n: int = 64, count: int = 24, window: int = 8,
                 steps_per_system: int = 1, batch: int = 8, lr: float = 2e-3, lr_final: float = 2e-4,
                 iters=gnn.SAMPLE_ITERS, seed: int = 0, graded_every: int = 0, graded_n: int = 192,
                 progress=None) -> dict:
    """Streams `count` pressure systems through a window of `window` resident ones.

    Each new system is meshed on the host (n^3 hex box, interior vertices perturbed +-20%,
    one mesh seed per system; every `graded_every`-th one a graded slab instead, 0: none),
    assembled on the GPU (Laplacian, mu = -1, Dirichlet x = 0, 8 blobs), and IC(0)-PCG
    generates its (residual, error) pairs; the oldest system then leaves with its pairs, and
    `steps_per_system` Adam steps of `batch` pairs drawn from the resident systems follow,
    the learning rate cosine-decayed over the stream.  Returns the per-phase wall times
    (seconds, summed over the stream) and the loss history."""
    resident = []
    t = {"mesh": 0.0, "assemble": 0.0, "sample": 0.0, "train": 0.0}
    history = []
    t_start = time.perf_counter()
    try:
        for k in range(count):
            graded = graded_every > 0 and k % graded_every == graded_every - 1
            sd = seed * 100_000 + k
            t0 = time.perf_counter()
            mesh = gnn.graded_slab(graded_n, max(8, graded_n // 6), sd) if graded else W.case_mesh(4, n, seed=sd)
            t1 = time.perf_counter()
            s = fcap.System(mesh, ctx)
            su0 = W.blob_rhs(s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol"), 8, seed=10_000 + sd)
            s.assemble_laplacian(np.full(mesh.n_cells, -1.0), None, su0)
            t2 = time.perf_counter()
            net.add_samples(s, iters, seed=sd)
            resident.append(s)
            if len(resident) > window:
                old = resident.pop(0)
                net.drop_system(old)
                old.close()
            t3 = time.perf_counter()
            frac = k / max(1, count)
            rate = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
            losses = net.train(steps_per_system, batch=batch, lr=float(rate), seed=seed * 7919 + k)
            t4 = time.perf_counter()
            history.extend(float(x) for x in losses)
            t["mesh"] += t1 - t0
            t["assemble"] += t2 - t1
            t["sample"] += t3 - t2
            t["train"] += t4 - t3
            if progress and (k + 1) % progress == 0:
                print(f"system {k + 1}/{count}: loss {np.mean(history[-progress * steps_per_system:]):.4f} "
                      f"(lr {rate:.2e}), {time.perf_counter() - t_start:.1f} s", flush=True)
    finally:
        net.clear_samples()
        for s in resident:
            s.close()
    return {"seconds": t, "wall_s": time.perf_counter() - t_start, "loss": history, "count": count,
            "window": window, "steps_per_system": steps_per_system, "batch": batch, "n": n,
            "graded_every": graded_every, "graded_n": graded_n, "seed": seed, "lr": lr, "lr_final": lr_final,
            "iters": list(iters)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64, help="cells per direction of the perturbed training meshes")
    ap.add_argument("--systems", type=int, default=32, help="perturbed n^3 systems")
    ap.add_argument("--graded", type=int, default=16, help="graded slabs (gnn.graded_slab)")
    ap.add_argument("--graded-n", type=int, default=192, help="cells along x of a graded slab")
    ap.add_argument("--iters", type=str, default=",".join(str(k) for k in gnn.SAMPLE_ITERS),
                    help="IC(0)-PCG iteration counts whose (r, e) pairs are recorded")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", type=str, default=None, help="start from these weights (.npy)")
    ap.add_argument("--out", type=str, default=gnn.DEFAULT_WEIGHTS)
    ap.add_argument("--log", type=str, default=None, help="write the loss history here (JSON)")
    ap.add_argument("--stream", type=int, default=0, metavar="COUNT",
                    help="stream COUNT systems through a window of --window resident ones instead of "
                         "training on a fixed set (--steps-per-system Adam steps per new system)")
    ap.add_argument("--window", type=int, default=8)
    ap.add_argument("--steps-per-system", type=int, default=1)
    ap.add_argument("--graded-every", type=int, default=0,
                    help="with --stream: every k-th streamed system is a graded slab (0: none)")
    ap.add_argument("--smooth", type=int, default=0,
                    help="damped-Jacobi steps after the network taken into the loss (the solver applies 3)")
    ap.add_argument("--omega", type=float, default=0.8, help="their damping (the solver's: 0.8)")
    ap.add_argument("--coarse", action="store_true",
                    help="take the solver's coarse correction (after the Jacobi steps) into the loss")
    args = ap.parse_args(argv)

    ctx = fcap.Context(0)
    init = np.load(args.init, allow_pickle=False) if args.init else gnn.init_params(args.seed)
    net = gnn.Network(init, ctx)
    if args.smooth:
        net.set_training_smoothing(args.smooth, args.omega)
    if args.coarse:
        net.set_training_coarse(True)
    command = " ".join(["python -m freecappuccino_amd.train_gnn", *(argv if argv is not None else sys.argv[1:])])
    iters = [int(x) for x in args.iters.split(",")]
    if args.stream:
        res = stream_train(net, ctx, n=args.n, count=args.stream, window=args.window,
                           steps_per_system=args.steps_per_system, batch=args.batch, lr=args.lr,
                           lr_final=args.lr_final, iters=iters, seed=args.seed, graded_every=args.graded_every,
                           graded_n=args.graded_n, progress=max(1, args.stream // 20))
        net.save(args.out)
        if args.log:
            with open(args.log, "w") as f:
                json.dump(dict(res, command=command, smooth=args.smooth, omega=args.omega, coarse=args.coarse), f)
        h = res["loss"]
        print(f"saved {args.out}: {args.stream} systems in {res['wall_s']:.1f} s, first-100 mean loss "
              f"{np.mean(h[:100]):.4f}, last-100 mean loss {np.mean(h[-100:]):.4f}", flush=True)
        net.close()
        ctx.close()
        return
    t0 = time.time()
    systems = gnn.build_dataset(net, args.n, args.systems, iters=iters, seed0=args.seed * 100000, ctx=ctx,
                                n_graded=args.graded, graded_n=args.graded_n)
    print(f"{args.systems} perturbed {args.n}^3 systems + {args.graded} graded slabs, {net.num_samples} pairs, "
          f"{time.time() - t0:.1f} s", flush=True)
    history = []
    done = 0
    while done < args.steps:
        # cosine-decayed learning rate, in blocks of 100 steps
        k = min(100, args.steps - done)
        frac = done / max(1, args.steps)
        lr = args.lr_final + 0.5 * (args.lr - args.lr_final) * (1 + np.cos(np.pi * frac))
        losses = net.train(k, batch=args.batch, lr=float(lr), seed=args.seed * 7919 + done)
        history.extend(float(x) for x in losses)
        done += k
        print(f"step {done}: loss {losses.mean():.4f} (lr {lr:.2e}), {time.time() - t0:.1f} s", flush=True)
    net.save(args.out)
    if args.log:
        with open(args.log, "w") as f:
            json.dump({"command": command, "n": args.n, "systems": args.systems, "graded": args.graded,
                       "graded_n": args.graded_n, "iters": iters, "steps": args.steps, "batch": args.batch,
                       "lr": args.lr, "lr_final": args.lr_final, "seed": args.seed, "init": args.init,
                       "smooth": args.smooth, "omega": args.omega, "coarse": args.coarse, "seconds": time.time() - t0,
                       "loss": history}, f)
    print(f"saved {args.out}: first-100 mean loss {np.mean(history[:100]):.4f}, "
          f"last-100 mean loss {np.mean(history[-100:]):.4f}", flush=True)
    net.clear_samples()
    for s in systems:
        s.close()
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
