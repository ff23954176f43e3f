"""Benchmark of the pressure-solve hot path on one MI355X.

A step is one pass of the hot path over the resident system: face-loop
assembly of the Laplacian into the CSR matrix (fvm_laplacian) followed by the
Krylov solve to 1e-8 FP64 relative residual ||su - A fi||_2 / ||su||_2.

Default workload: 192^3 hex box (7,077,888 cells), interior vertices
perturbed by +-20% (seeded), spacing graded along x by ratio 1.05 toward the
wall (aspect ratio capped at 100), mu = -1, Dirichlet wall at x = 0, RHS = 32
seeded Gaussian blobs with zero mean.  Three solvers run on the same system:
  * ic_pcg_fp64  -- IC(0)-PCG in FP64 (the reference's DIC-PCG, level-scheduled sweeps);
  * ic_pcg_mixed -- FP32 DIC-PCG inside FP64 iterative refinement;
  * fcg_learned  -- flexible CG in FP32 preconditioned by the trained BF16
                    message-passing network, inside FP64 iterative refinement,
                    IC(0)-PCG fallback on stagnation.
`value` is the time to 1e-8 of fcg_learned (ms per step, lower is better);
every solver's time, iterations, FP64 residual and error against the FP64
IC(0)-PCG solution are in `solvers`.

The roofline object is the dominant kernel of the headline solver: one hidden
message-passing layer of the network (about 4/5 of a learned-FCG iteration),
timed alone with HIP events on the library's stream; its algorithmic bytes per
cell are the cell's own BF16 row in and out, its ELL slots and count
(gnn_layer_bytes).  The FP64 SpMV on the same matrix is reported beside it
(`spmv`), and so is the CPU baseline: the oracle's sequential DIC-PCG (the
reference algorithm restated in C) timed on one host core over a bounded
number of iterations and scaled to the GPU's FP64 IC(0)-PCG iteration count.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--case 5] [--n 192]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METHODS = ("ic_pcg_fp64", "ic_pcg_mixed", "fcg_learned")

# HBM bytes per launch of the hidden network layer, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (x2.0 and x1.0, the scales a 1 GiB device copy gives) on the
# 192^3 graded box; summary and commands in profiles/gnn_layer_hex192_graded.md
LAYER_TRAFFIC = {(5, 192): 1_171_501_056 + 452_984_832}


def spmv_bytes(n, nnz, value_bytes=8, index_bytes=4):
    """Algorithmic bytes of one SpMV: values + column indices + row pointers + x once + y."""
    return nnz * (value_bytes + index_bytes) + index_bytes * (n + 1) + 2 * value_bytes * n


def gnn_layer_bytes(n, width=6):
    """Algorithmic bytes of one message-passing layer: the cell's own BF16 row in (64) and out
    (64), its ELL slots (neighbour index + packed features, 8 per slot) and count (4); the
    neighbours' rows are the other cells' own rows and are not counted again."""
    return n * (64 + 64 + 8 * width + 4)


def gnn_bytes(n, width=6, layers=4):
    """One application: RMS pass (r, a_ii), input pass (r, a_ii in; u, h0 out), layers 1..L-1,
    and the last layer (reads u, writes z instead of a row)."""
    return n * (8 + 8 + 4 + 64) + (layers - 1) * gnn_layer_bytes(n, width) + n * (64 + 8 * width + 4 + 4 + 4)


def cpu_baseline(mesh, a, su, gpu_iters, budget_iters):
    """The oracle's sequential DIC-PCG on one host core, `budget_iters` iterations, scaled."""
    from oracle import oracle as O
    csr = O.csr_build(mesh.n_cells, mesh.owner[:mesh.n_inner] + 1, mesh.neighbour + 1)
    t0 = time.perf_counter()
    r = O.iccg(csr, a, su, np.zeros(mesh.n_cells), sor=1e-30, nsw=budget_iters)
    dt = time.perf_counter() - t0
    per_iter = dt / max(r.iters, 1)
    return {"value": per_iter * gpu_iters * 1e3, "unit": "ms", "cores": 1, "kind": "port",
            "sample": f"oracle DIC-PCG (C restatement of src/iccg.f90, gcc -O2, 1 thread) for {r.iters} "
                      f"iterations on the same system ({dt:.1f} s incl. pivots), per-iteration time x "
                      f"{gpu_iters} GPU FP64 IC(0)-PCG iterations; host "
                      f"{platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--case", type=int, default=5, choices=(1, 2, 5),
                    help="1: 40^3 box; 2: 128^3 perturbed; 5: 192^3 perturbed, x-graded 1.05")
    ap.add_argument("--n", type=int, default=None, help="cells per direction (overrides the case's size)")
    ap.add_argument("--methods", type=str, default=",".join(METHODS))
    ap.add_argument("--weights", type=str, default=None, help="network weights (.npy); default: the packaged ones")
    ap.add_argument("--cpu-iters", type=int, default=40, help="iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--profile-spmv", action="store_true",
                    help="only assemble and run --spmv-reps FP64 SpMVs plus a 1 GiB device copy "
                         "(the calibration kernel of the HBM counters), then exit")
    ap.add_argument("--profile-gnn", type=int, default=0, metavar="REPS",
                    help="only assemble and run REPS applications of the network, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    methods = [m for m in args.methods.split(",") if m]
    for m in methods:
        if m not in METHODS:
            ap.error(f"unknown method {m}")

    from freecappuccino_amd import fcap, gnn
    from freecappuccino_amd import workloads as W

    args.n = args.n or {1: 40, 2: 128, 5: 192}[args.case]
    mesh = W.case_mesh(args.case, args.n)
    ctx = fcap.Context(0)
    s = fcap.System(mesh, ctx)
    n, nnz = mesh.n_cells, mesh.nnz
    xc, yc, zc, vol = s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol")
    nblobs = 32 if args.case == 5 else 16
    su0 = W.blob_rhs(xc, yc, zc, vol, nblobs, seed=2)
    mu = np.full(n, -1.0)
    s.assemble_laplacian(mu, None, su0)
    sizes = s.sizes()
    if args.profile_spmv:
        ms = s.spmv_bench(64, args.spmv_reps)
        copy_bytes = 1 << 30
        copy_ms = ctx.copy_bench(copy_bytes, 3)
        print(json.dumps({"spmv_fp64_ms": ms, "spmv_bytes": spmv_bytes(n, nnz),
                          "spmv_GBps": spmv_bytes(n, nnz) / ms / 1e6, "copy_bytes": copy_bytes,
                          "copy_ms": copy_ms, "copy_GBps": 2 * copy_bytes / copy_ms / 1e6,
                          "cells": n, "nnz": nnz}))
        s.close()
        ctx.close()
        return
    if args.profile_gnn:
        net = gnn.Network.load(args.weights or gnn.DEFAULT_WEIGHTS, ctx)
        ms = net.bench(s, args.profile_gnn)
        b = gnn_bytes(n)
        print(json.dumps({"gnn_ms": ms, "gnn_bytes": b, "gnn_GBps": b / ms / 1e6,
                          "layer_bytes": gnn_layer_bytes(n), "cells": n, "reps": args.profile_gnn}))
        net.close()
        s.close()
        ctx.close()
        return

    net = None
    layer_ms = None
    if "fcg_learned" in methods:
        net = gnn.Network.load(args.weights or gnn.DEFAULT_WEIGHTS, ctx)
        net.attach(s)
    code = {"ic_pcg_fp64": fcap.METHOD_ICCG, "ic_pcg_mixed": fcap.METHOD_ICCG_MIXED,
            "fcg_learned": fcap.METHOD_FCG_LEARNED}

    def step(method):
        s.set_solution(np.zeros(n))
        t0 = time.perf_counter()
        s.assemble_laplacian_dev()
        rc, info = s.solve(method, fcap.STOP_L2, 1e-8, 20000)
        return rc, info, (time.perf_counter() - t0) * 1e3, s.last_ms(0)

    results, sols = {}, {}
    for name in methods:
        for _ in range(args.warmup):
            step(code[name])
        times, infos = [], []
        for _ in range(args.steps):
            rc, info, wall_ms, asm_ms = step(code[name])
            times.append(wall_ms)
            infos.append((rc, info, asm_ms))
        rc, info, asm_ms = infos[-1]
        results[name] = {"ms": float(np.median(times)), "status": fcap.STATUS_NAMES.get(rc, rc),
                         "iters": info["iters"], "outer": info["outer"], "fallback_iters": info["fallback_iters"],
                         "rel_l2": info["rel_l2"], "assembly_ms": asm_ms,
                         "solve_device_ms": info["solve_ms"] + info["setup_ms"]}
        sols[name] = s.solution()
        print(f"{name}: {results[name]}", file=sys.stderr, flush=True)
    if net is not None:
        layer_ms = net.bench_layer(s, args.spmv_reps)
        net.detach(s)
        net.close()
    if "ic_pcg_fp64" in sols:
        x64 = sols["ic_pcg_fp64"]
        for name in results:
            results[name]["err_vs_ic_pcg_fp64"] = float(np.linalg.norm(sols[name] - x64) / np.linalg.norm(x64))
            results[name]["speedup_vs_ic_pcg_fp64"] = results["ic_pcg_fp64"]["ms"] / results[name]["ms"]

    spmv_ms = s.spmv_bench(64, args.spmv_reps)
    spmv32_ms = s.spmv_bench(32, args.spmv_reps)
    b64 = spmv_bytes(n, nnz, 8, 4)
    b32 = spmv_bytes(n, nnz, 4, 4)
    gbs = b64 / (spmv_ms * 1e-3) / 1e9
    head = "fcg_learned" if "fcg_learned" in results else min(results, key=lambda k: results[k]["ms"])
    value = results[head]["ms"]
    workload = (f"hex box {args.n}^3 perturbed +-20% (seeded)"
                + (", x-graded 1.05 toward the wall (aspect <= 100)" if args.case == 5 else "")
                + f", Laplacian mu=-1, Dirichlet x=0, {nblobs} Gaussian blobs; "
                "face-loop assembly + solve to 1e-8 FP64 rel. L2 residual")
    line = {
        "metric": "pressure-solve time to 1e-8 rel. residual, learned-FCG vs IC-PCG; SpMV GB/s",
        "value": round(value, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(value, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": {"ic_pcg_fp64": "f64", "ic_pcg_mixed": "f32+f64", "fcg_learned": "bf16+f32+f64"}[head],
        "data": "synthetic",
        "config": {"workload": workload, "cells": n, "nnz": nnz, "levels_fwd": sizes["nlev_fwd"],
                   "solver": head},
        "solvers": results,
        "spmv": {"fp64_ms": spmv_ms, "fp64_GBps": gbs, "fp32_ms": spmv32_ms,
                 "fp32_GBps": b32 / (spmv32_ms * 1e-3) / 1e9, "bytes_fp64": b64, "bytes_fp32": b32},
    }
    if layer_ms is not None:
        lb = gnn_layer_bytes(n)
        lgbs = lb / (layer_ms * 1e-3) / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": round(lgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(lgbs / HBM_PEAK_GBS, 4),
                            "traffic": LAYER_TRAFFIC.get((args.case, args.n)),
                            "kernel": "k_gnn_layer_quad<false,false,6> (hidden message-passing layer)",
                            "launch_ms": layer_ms, "bytes_per_launch": lb}
    else:
        line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_spmv<double,double,0>",
                            "launch_ms": spmv_ms, "bytes_per_launch": b64}
    if not args.no_cpu and "ic_pcg_fp64" in results:
        a = s.get("a")
        su = s.get("su")
        line["cpu_baseline"] = cpu_baseline(mesh, a, su, results["ic_pcg_fp64"]["iters"], args.cpu_iters)
    s.close()
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
