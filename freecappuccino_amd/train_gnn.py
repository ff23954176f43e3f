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
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import fcap, gnn


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
    args = ap.parse_args(argv)

    ctx = fcap.Context(0)
    init = np.load(args.init, allow_pickle=False) if args.init else gnn.init_params(args.seed)
    net = gnn.Network(init, ctx)
    iters = [int(x) for x in args.iters.split(",")]
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
            json.dump({"n": args.n, "systems": args.systems, "graded": args.graded, "graded_n": args.graded_n,
                       "iters": iters, "steps": args.steps, "batch": args.batch, "lr": args.lr,
                       "lr_final": args.lr_final, "seconds": time.time() - t0, "loss": history}, f)
    print(f"saved {args.out}: first-100 mean loss {np.mean(history[:100]):.4f}, "
          f"last-100 mean loss {np.mean(history[-100:]):.4f}", flush=True)
    net.clear_samples()
    for s in systems:
        s.close()
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
