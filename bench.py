"""Benchmark of the pressure-solve hot path on one MI355X.

A step is one pass of the hot path over the resident system: face-loop
assembly of the Laplacian into the CSR matrix (fvm_laplacian) followed by the
Krylov solve to 1e-8 FP64 relative residual ||su - A fi||_2 / ||su||_2.

Default workload: 192^3 hex box (7,077,888 cells), interior vertices
perturbed by +-20% (seeded), spacing graded along x by ratio 1.05 toward the
wall (aspect ratio capped at 100), mu = -1, Dirichlet wall at x = 0, RHS = 32
seeded Gaussian blobs with zero mean.  Four solves run on the same system:
  * ic_pcg_fp64  -- IC(0)-PCG in FP64 (the reference's DIC-PCG; triangular sweeps
                    with one launch per sweep, each row waiting for the rows it reads);
  * ic_pcg_mixed -- FP32 DIC-PCG inside FP64 iterative refinement (same sweeps);
  * fcg_learned  -- flexible CG in FP32 preconditioned by the trained BF16
                    message-passing network followed by three damped-Jacobi
                    steps (omega 0.8) and a coarse correction over 4,096-cell
                    aggregates of the mesh, inside FP64 iterative refinement,
                    IC(0)-PCG fallback on stagnation;
  * fcg_learned_every_iter -- fcg_learned with the network on every iteration
                    (by default it runs on every second one, the Jacobi guess on
                    the others);
  * fcg_learned_no_coarse -- fcg_learned without the coarse correction (and
                    with the network on every iteration);
  * fcg_jacobi_poly -- the same flexible CG without the network
                    (FC_METHOD_FCG_JACOBI): the Jacobi guess, the same three
                    damped-Jacobi steps and the same coarse correction on every
                    iteration, which shows what the trained network itself buys;
  * ic_pcg_fp64_levels -- ic_pcg_fp64 with one sweep launch per level (the same
                    bits, the schedule before the dependency-driven one).
`value` is the time to 1e-8 of fcg_learned (ms per step, lower is better);
every solver's time, iterations, FP64 residual (the solver's own and one
recomputed on the host from the resident CSR arrays, `rel_l2_host`) and error
against the FP64 IC(0)-PCG solution are in `solvers`; `target` states the 2x
goal of the learned solve against FP64 IC(0)-PCG and whether it is met, and
`sweeps` times one FP64 DIC sweep pair alone against its compulsory bytes.

The roofline object is the dominant kernel of the headline solver: one hidden
message-passing layer of the network (about 4/5 of a learned-FCG iteration),
timed alone with HIP events on the library's stream; its algorithmic bytes per
cell are the cell's own BF16 row in and out, its ELL slots and count
(gnn_layer_bytes).  The FP64 SpMV on the same matrix is reported beside it
(`spmv`), and so is the CPU baseline: the oracle's sequential DIC-PCG (the
reference algorithm restated in C) timed on one pinned host core over a bounded
number of iterations, its pivots timed apart, scaled to the GPU's FP64 IC(0)-PCG
iteration count, and converted to the reference's own flang build through the
ratio in oracle/calibration.json.

--case 3 times configuration 3 instead (DILU-BiCGStab, FP64 and FP32-in-FP64, on
the 96^3 permuted graded Kuhn-tet convection-diffusion system; roofline of the
DILU sweep pair), and --case 4 configuration 4 (training streamed over --systems
64^3 systems, batch 8; value in systems per second, with the projected time for
the 2,000 systems the configuration names).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--case 5] [--n 192]
"""
from __future__ import annotations

import argparse
import inspect
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METHODS = ("ic_pcg_fp64", "ic_pcg_mixed", "fcg_learned", "fcg_learned_every_iter", "fcg_learned_no_coarse",
           "fcg_jacobi_poly", "ic_pcg_fp64_levels")

# HBM bytes per launch of the hidden network layer, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (x2.0 and x1.0, the scales a 1 GiB device copy gives) on the
# 192^3 graded box; summary and commands in profiles/gnn_layer_hex192_graded.md
LAYER_TRAFFIC = {(5, 192): 1_171_501_056 + 452_984_832}
# the same for the hidden layer over 256-cell clusters (k_gnn_layer_cl): FETCH_SIZE 431,862 KiB
# (x 2.0) + WRITE_SIZE 442,368 KiB (x 1.0); summary in profiles/gnn_layer_hex192_graded.md
LAYER_TRAFFIC_CL = {(5, 192): 431_862 * 1024 * 2 + 442_368 * 1024}


def spmv_bytes(n, nnz, value_bytes=8, index_bytes=4):
    """Algorithmic bytes of one SpMV: values + column indices + row pointers + x once + y."""
    return nnz * (value_bytes + index_bytes) + index_bytes * (n + 1) + 2 * value_bytes * n


def gnn_layer_bytes(n, width=6):
    """Algorithmic bytes of one message-passing layer: the cell's own BF16 row in (64) and out
    (64), its ELL slots (neighbour index + packed features, 8 per slot) and count (4); the
    neighbours' rows are the other cells' own rows and are not counted again."""
    return n * (64 + 64 + 8 * width + 4)


def gnn_layer_bytes_cl(n, clusters, halo_total, width=6, cells_per_cluster=256):
    """Algorithmic bytes of one hidden layer over the cluster layout: the cell's own BF16 row in
    (64) and out (64); per position of the padded cluster layout one 8-byte word per slot pair
    (the two slots' LDS rows and BF16 gates); per cluster its halo count (4) and halo positions
    (4 each).  The halo rows are other cells' own rows and are not counted again."""
    n_pad = clusters * cells_per_cluster
    return n * (64 + 64) + n_pad * 8 * (width // 2) + 4 * clusters + 4 * halo_total


def gnn_bytes(n, width=6, layers=4):
    """One application: RMS pass (r, a_ii), input pass (r, a_ii in; u, h0 out), layers 1..L-1,
    and the last layer (reads u, writes z instead of a row)."""
    return n * (8 + 8 + 4 + 64) + (layers - 1) * gnn_layer_bytes(n, width) + n * (64 + 8 * width + 4 + 4 + 4)


def calcp_step_bytes(n, nif, nnz, nb):
    """Algorithmic bytes per launch of each kernel of the pressure-correction step: every index,
    face value and cell value the kernel needs read once, every output written once (4-byte
    indices, 8-byte values; a face's geometry is xf, yf, zf, arx, ary, arz, facint = 56 bytes)."""
    off = nnz - n
    return {
        # own, nei, ij, ji, geometry; a(ij), a(ji), flmass out; 12 cell fields + 4 gradients (3 each)
        "k_calcp_faces": nif * (16 + 56 + 24) + n * 24 * 8,
        # ptr, diag, a and su out; per off-diagonal slot kface (face and side) + a; per face flmass
        "k_calcp_rows": n * (8 + 16) + off * (4 + 8) + nif * 8,
        # own, nei, geometry, (x, y, z)-products out; u, centre and previous gradient per cell
        "k_grad_faces": nif * (8 + 56 + 24) + n * 7 * 8,
        # ptr, diag, vol, gradient out; kface per slot; products per face; boundary faces
        # (owner list entry, S, value)
        "k_grad_rows": n * (8 + 8 + 24) + off * 4 + nif * 24 + nb * (4 + 24 + 8),
        # own, nei, ij, a(ij), flmass in and out; pp per cell
        "k_correct_faces": nif * (12 + 8 + 16) + n * 8,
        # pp, grad pp, vol, 1/a_P (3), u, v, w, p in and out
        "k_correct_cells": n * (8 + 24 + 8 + 24 + 64),
        # own, nei, geometry, fmcor out, flmass in and out; centre, vol, den, apu, grad pp per cell
        "k_nonorth_faces": nif * (8 + 56 + 8 + 16) + n * 9 * 8,
        # ptr, diag, su out; kface per slot; fmcor per face
        "k_face_source_rows": n * (8 + 8) + off * 4 + nif * 8,
    }


def _inlet_outlet(m):
    """The mesh with its x = 0 patch as inlet, x = 1 as outlet and walls elsewhere."""
    import dataclasses
    types = ["inlet", "outlet", "wall", "wall", "wall", "wall"]
    return dataclasses.replace(m, patches=[(t, nf, st) for t, (_, nf, st) in zip(types, m.patches)])


def _calcp_inputs(api, m, geo):
    """Fields of the pressure-correction step (cells, and cells + boundary values extrapolated
    from the owners), the pressure correction pp and the inlet mass fluxes."""
    from freecappuccino_amd import workloads as W
    n, nif = m.n_cells, m.n_inner
    bown = m.owner[nif:]
    u, v, w = W.taylor_green(geo.xc, geo.yc, geo.zc)
    rng = np.random.default_rng(7)
    f = {"den": np.ones(n), "u": u, "v": v, "w": w, "p": np.sin(3.0 * geo.xc) * geo.yc + geo.zc}
    for k in ("apu", "apv", "apw"):
        f[k] = (0.5 + rng.random(n)) / geo.vol
    pp = 1e-2 * np.cos(2.0 * geo.yc) * geo.xc
    bnd = api.boundary_groups(m)
    s_in = slice(int(bnd[1]), int(bnd[1]) + int(bnd[0]))
    o = m.owner[s_in]
    fmi = -(u[o] * geo.arx[s_in] + v[o] * geo.ary[s_in] + w[o] * geo.arz[s_in])
    fb = {k: np.concatenate([f[k], f[k][bown]]) for k in ("u", "v", "w", "p")}
    return f, fb, pp, np.concatenate([pp, pp[bown]]), fmi


def _calcp_pass(api, m, geo, csr, f, fb, pp, ppb, fmi, timed, **kw):
    """One pass of the step's calls; wall ms per call name are appended to timed."""
    def clock(name, fn, *a):
        t0 = time.perf_counter()
        r = fn(*a, **kw)
        timed.setdefault(name, []).append((time.perf_counter() - t0) * 1e3)
        return r

    g = {}
    for k, src in (("dudx", "u"), ("dvdx", "v"), ("dwdx", "w"), ("dpdx", "p")):
        g[k] = np.stack(clock("grad_gauss", api.grad_gauss, m, geo, fb[src]), axis=1)
    out = clock("calcp_assemble", api.calcp_assemble, m, geo, csr, dict(f, **g))
    su = clock("adjust_mass_flow", api.adjust_mass_flow, m, geo, float(fmi.sum()), f["den"], fb["u"], fb["v"],
               fb["w"], out["su"], fmi)[0]
    gpp = np.stack(clock("grad_gauss", api.grad_gauss, m, geo, ppb), axis=1)
    fl = clock("calcp_correct", api.calcp_correct, m, csr, out["a"], pp, gpp, geo.vol, f["apu"], f["apv"],
               f["apw"], 0.2, 0, out["flmass"], f["u"], f["v"], f["w"], f["p"])[0]
    # the library's call takes the CSR structure (it checks the face slots), the oracle's does not
    with_csr = [csr] if "csr" in inspect.signature(api.calcp_nonorth).parameters else []
    su2 = clock("calcp_nonorth", api.calcp_nonorth, m, geo, *with_csr, f["den"], f["apu"], gpp, fl)[0]
    return float(np.abs(su).sum()), float(np.abs(su2).sum())


def _assemble_device_resident(fcap, ctx, m, geo, csr, f, reps):
    """Wall ms of fc_calcp_assemble with every floating-point input and output in device memory
    (hipMalloc'd once; index arrays stay on the host): the call copies them on the device."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    n, nif = m.n_cells, m.n_inner
    host = [np.ascontiguousarray(getattr(geo, k), np.float64)
            for k in ("xc", "yc", "zc", "vol", "xf", "yf", "zf", "arx", "ary", "arz", "facint")]
    host += [np.ascontiguousarray(f[k], np.float64) for k in fcap.CALCP_FIELDS]
    host += [np.ascontiguousarray(np.asarray(f[k]).reshape(n, 3), np.float64) for k in fcap.CALCP_GRADS]
    outs = [np.zeros(m.nnz), np.zeros(n), np.zeros(nif)]
    dev = []
    try:
        for a in host + outs:
            p = C.c_void_p()
            if hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) != 0:
                raise RuntimeError("hipMalloc failed")
            dev.append(p)
            hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1)
        dp = [C.cast(p, C.POINTER(C.c_double)) for p in dev]
        ip = lambda x: np.ascontiguousarray(x, np.int32).ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
        own1 = (m.owner[:nif].astype(np.int64) + 1).astype(np.int32)
        nei1 = (m.neighbour.astype(np.int64) + 1).astype(np.int32)
        z = np.zeros(1)
        zi = np.zeros(1, np.int32)
        times = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            # 11 geometry arrays, 8 fields, 4 gradients, then the outputs a, su, flmass
            rc = ctx.L.fc_calcp_assemble(ctx.h, n, nif, ip(own1), ip(nei1), *dp[:23], 0, 0, ip(zi), ip(zi),
                                         z.ctypes.data_as(C.POINTER(C.c_double)), ip(csr.ioffset), ip(csr.ja),
                                         ip(csr.diag), ip(csr.ij_slot), ip(csr.ji_slot), m.nnz, dp[23], dp[24],
                                         dp[25], *[z.ctypes.data_as(C.POINTER(C.c_double))] * 3)
            hip.hipDeviceSynchronize()
            times.append((time.perf_counter() - t0) * 1e3)
            ctx.check(rc)
        return float(np.median(times[1:]))
    finally:
        for p in dev:
            hip.hipFree(p)


def _assemble_resident_system(fcap, ctx, m, f, reps):
    """fc_sys_calcp_assemble on a resident system of the mesh (geometry, structure and slot
    maps already on the device): wall ms with host-array fields, wall ms with the fields and
    flmass in device memory, and the device ms of the launches (HIP events)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    n = m.n_cells
    host = [np.ascontiguousarray(f[k], np.float64) for k in fcap.CALCP_FIELDS]
    host += [np.ascontiguousarray(np.asarray(f[k]).reshape(n, 3), np.float64) for k in fcap.CALCP_GRADS]
    s = fcap.System(m, ctx)
    dev = []
    try:
        t_host, t_dev, t_kern = [], [], []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            s.calcp_assemble(*host)
            t_host.append((time.perf_counter() - t0) * 1e3)
        for a in host + [np.zeros(max(m.n_inner, 1))]:
            p = C.c_void_p()
            if hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) != 0:
                raise RuntimeError("hipMalloc failed")
            dev.append(p)
            hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1)
        ptrs = [p.value for p in dev]
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            s.calcp_assemble(*ptrs[:12], flmass=ptrs[12])   # returns after a stream synchronize
            t_dev.append((time.perf_counter() - t0) * 1e3)
            t_kern.append(s.last_ms(0))
        # the Gauss gradient of u (boundary values zero) into a device (3, numCells) array
        p = C.c_void_p()
        if hip.hipMalloc(C.byref(p), 8 * s.ntotal) != 0:
            raise RuntimeError("hipMalloc failed")
        dev.append(p)
        u_full = np.concatenate([host[4], np.zeros(s.ntotal - n)])
        hip.hipMemcpy(p, u_full.ctypes.data, u_full.nbytes, 1)
        t_grad, t_grad_kern = [], []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            s.grad_gauss(p.value, out=ptrs[8])
            t_grad.append((time.perf_counter() - t0) * 1e3)
            t_grad_kern.append(s.last_ms(0))
        # the corrections with the system's matrix and pp = its (zero) solution, fields in place
        t_cor, t_cor_kern = [], []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            ctx.check(ctx.L.fc_sys_calcp_correct(s.h, None, C.c_void_p(ptrs[11]), *[C.c_void_p(x) for x in ptrs[1:4]],
                                                 0.2, 1, *[C.c_void_p(x) for x in (ptrs[12], ptrs[4], ptrs[5],
                                                                                   ptrs[6], ptrs[7])]))
            t_cor.append((time.perf_counter() - t0) * 1e3)
            t_cor_kern.append(s.last_ms(0))
        # the corrector's source into the system's su, flmass in place
        t_no, t_no_kern = [], []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            ctx.check(ctx.L.fc_sys_calcp_nonorth(s.h, *[C.c_void_p(x) for x in (ptrs[0], ptrs[1], ptrs[11],
                                                                                ptrs[12])]))
            t_no.append((time.perf_counter() - t0) * 1e3)
            t_no_kern.append(s.last_ms(0))
        return {"host_fields_wall_ms": float(np.median(t_host[1:])),
                "nonorth_device_wall_ms": float(np.median(t_no[1:])),
                "nonorth_device_ms": float(np.median(t_no_kern[1:])),
                "correct_device_wall_ms": float(np.median(t_cor[1:])),
                "correct_device_ms": float(np.median(t_cor_kern[1:])),
                "device_fields_wall_ms": float(np.median(t_dev[1:])),
                "device_ms": float(np.median(t_kern[1:])),
                "grad_gauss_device_wall_ms": float(np.median(t_grad[1:])),
                "grad_gauss_device_ms": float(np.median(t_grad_kern[1:]))}
    finally:
        for p in dev:
            hip.hipFree(p)
        s.close()


def profile_calcp(mesh, ctx, reps, cpu_n):
    """The pressure-correction step's calls on the workload mesh with inlet and outlet patches:
    Gauss gradients, assembly, adjustMassFlow, corrections and the non-orthogonal corrector's
    source, each through the host-array C ABI (wall times include the PCIe copies; the kernels'
    own times come from a kernel trace of the same command).  The oracle runs the same calls on
    a cpu_n^3 mesh of the same family for the CPU baseline."""
    from freecappuccino_amd import fcap
    from freecappuccino_amd import workloads as W

    m = _inlet_outlet(mesh)
    geo = fcap.mesh_geometry(m, ctx)
    csr = fcap.create_csr_matrix_from_mesh_data(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1, ctx)
    data = _calcp_inputs(fcap, m, geo)
    timed = {}
    for _ in range(reps):
        check = _calcp_pass(fcap, m, geo, csr, *data, timed, ctx=ctx)
        print(f"calcp step pass: {({k: round(v[-1], 1) for k, v in timed.items()})}", file=sys.stderr, flush=True)
    f = dict(data[0])
    for k, src in (("dudx", "u"), ("dvdx", "v"), ("dwdx", "w"), ("dpdx", "p")):
        f[k] = np.stack(fcap.grad_gauss(m, geo, data[1][src], ctx=ctx), axis=1)
    resident = _assemble_device_resident(fcap, ctx, m, geo, csr, f, reps)
    print(f"calcp_assemble with device-resident fields: {resident:.1f} ms", file=sys.stderr, flush=True)
    on_system = _assemble_resident_system(fcap, ctx, m, f, reps)
    print(f"sys_calcp_assemble on the resident system: {on_system}", file=sys.stderr, flush=True)
    res = {"cells": m.n_cells, "inner_faces": m.n_inner, "nnz": m.nnz, "reps": reps,
           "wall_ms_incl_pcie": {k: float(np.median(v)) for k, v in timed.items()},
           "calcp_assemble_wall_ms_device_resident": resident,
           "sys_calcp_assemble": on_system,
           "kernel_bytes": calcp_step_bytes(m.n_cells, m.n_inner, m.nnz, m.n_boundary), "check": check}
    if cpu_n:
        from oracle import oracle as O
        cm = _inlet_outlet(W.case_mesh(5, cpu_n))
        cgeo = O.mesh_geometry(cm)
        ccsr = O.csr_build(cm.n_cells, cm.owner[:cm.n_inner] + 1, cm.neighbour + 1)
        cdata = _calcp_inputs(O, cm, cgeo)
        ct = {}
        _calcp_pass(O, cm, cgeo, ccsr, *cdata, ct)
        total = sum(sum(v) for v in ct.values())
        res["cpu_baseline"] = {
            "value": total * m.n_cells / cm.n_cells, "unit": "ms", "cores": 1, "kind": "port",
            "per_pass_ms": {k: float(sum(v)) for k, v in ct.items()},
            "sample": f"the oracle (C restatement, gcc -O2, 1 thread) on one pass of the same calls over a "
                      f"{cpu_n}^3 mesh of the same family ({cm.n_cells} cells, {total / 1e3:.2f} s), "
                      f"scaled by cells to {m.n_cells}"}
    return res


def _momentum_inputs(api, m, geo, ctx_kw):
    from freecappuccino_amd import workloads as W
    n = m.n_cells
    u, v, w = W.taylor_green(geo.xc, geo.yc, geo.zc)
    flux = api.face_flux_central(m, geo, u, v, w, **ctx_kw)
    rng = np.random.default_rng(11)
    sp = [rng.random(n) * geo.vol * 10 for _ in range(3)]
    src = [W.blob_rhs(geo.xc, geo.yc, geo.zc, geo.vol, 8, seed=20 + c) for c in range(3)]
    phi0 = [0.1 * rng.standard_normal(n) for _ in range(3)]
    return np.full(n, 1e-3), flux, sp, src, phi0, [0.7, 0.7, 0.7]


def profile_momentum(mesh, ctx, reps, cpu_n, sor=1e-8):
    """calcuvw's three component solves (BiCGStab-DILU, FP64) on one convection-diffusion
    assembly of the workload mesh: a central-difference flux of a Taylor-Green field, vis = 1e-3,
    random sp, blob sources, urf 0.7.  Device times from the solver's own events; the oracle's
    BiCGStab on a cpu_n^3 mesh of the same family, scaled by cells x iterations, is the CPU
    baseline."""
    from freecappuccino_amd import fcap
    from freecappuccino_amd import workloads as W
    geo = fcap.mesh_geometry(mesh, ctx)
    vis, flux, sp, src, phi0, urf = _momentum_inputs(fcap, mesh, geo, {"ctx": ctx})
    s = fcap.System(mesh, ctx)
    runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        s.assemble_convdiff(vis, flux, 0.7, None)
        rc, _, _, infos = s.solve_momentum(sp, src, urf, phi0, fcap.METHOD_BICGSTAB, sor, 2000)
        wall = (time.perf_counter() - t0) * 1e3
        runs.append({"wall_ms_incl_pcie": wall, "status": fcap.STATUS_NAMES.get(rc, rc),
                     "iters": [x["iters"] for x in infos],
                     "solve_device_ms": [x["solve_ms"] + x["setup_ms"] for x in infos]})
        print(f"momentum pass: {runs[-1]}", file=sys.stderr, flush=True)
    s.close()
    res = {"cells": mesh.n_cells, "nnz": mesh.nnz, "sor": sor, "runs": runs}
    if cpu_n:
        from oracle import oracle as O
        cm = W.case_mesh(5, cpu_n)
        cgeo = O.mesh_geometry(cm)
        csr = O.csr_build(cm.n_cells, cm.owner[:cm.n_inner] + 1, cm.neighbour + 1)
        cvis, cflux, csp, csrc, cphi, curf = _momentum_inputs(O, cm, cgeo, {})
        t0 = time.perf_counter()
        a_base = O.convdiff(cm, cgeo, csr, cvis, cflux, 0.7)
        per_iter, its = [], []
        for c in range(3):
            a_c, su_c, _ = O.momentum_rows(csr, a_base, csp[c], csrc[c], cphi[c], curf[c])
            t1 = time.perf_counter()
            r = O.bicgstab(csr, a_c, su_c, cphi[c], sor=sor, nsw=2000)
            per_iter.append((time.perf_counter() - t1) / max(r.iters, 1))
            its.append(r.iters)
        dt = time.perf_counter() - t0
        gpu_iters = runs[-1]["iters"]
        scale = mesh.n_cells / cm.n_cells
        res["cpu_baseline"] = {
            "value": 1e3 * scale * sum(p * g for p, g in zip(per_iter, gpu_iters)), "unit": "ms", "cores": 1,
            "kind": "port",
            "sample": f"the oracle's assembly, rows and BiCGStab-DILU (C restatement, gcc -O2, 1 thread) on a "
                      f"{cpu_n}^3 mesh of the same family ({cm.n_cells} cells, iterations {its}, {dt:.1f} s), "
                      f"time per iteration scaled by cells and by the GPU's iterations {gpu_iters}"}
    return res


def profile_io(mesh, nthreads, workdir):
    """Host-side rows of the step: the polyMesh reader on the workload mesh written in both
    layouts (native 1-based, OpenFOAM ASCII), and a restart file of the mesh's fields written
    and read back (FIELDS of freecappuccino_amd.restart, the reference's record list)."""
    import shutil
    import tempfile

    from freecappuccino_amd import mesh as M
    from freecappuccino_amd import polymesh, restart
    res = {"cells": mesh.n_cells, "faces": mesh.n_faces, "threads": nthreads}
    d = tempfile.mkdtemp(prefix="fcap_io_", dir=workdir)
    try:
        for fmt in ("native", "foam"):
            pm = os.path.join(d, fmt)
            M.write_polymesh(mesh, pm, fmt=fmt)
            nbytes = sum(os.path.getsize(os.path.join(pm, f)) for f in os.listdir(pm))
            for nt in sorted({1, nthreads}):
                t0 = time.perf_counter()
                got = polymesh.read_polymesh(pm, nthreads=nt)
                dt = time.perf_counter() - t0
                assert got.n_cells == mesh.n_cells and got.n_faces == mesh.n_faces
                res[f"polymesh_{fmt}_{nt}t"] = {"s": dt, "MB": nbytes / 1e6, "MB_per_s": nbytes / dt / 1e6}
                print(f"polyMesh {fmt} {nt} threads: {dt:.2f} s", file=sys.stderr, flush=True)
            shutil.rmtree(pm)
        n, nif, ntot = mesh.n_cells, mesh.n_inner, mesh.n_cells + mesh.n_boundary
        sizes = {"flmass": nif, "uu": n, "vv": n, "ww": n, "uv": n, "uw": n, "vw": n}
        rng = np.random.default_rng(5)
        fields = {k: rng.standard_normal(sizes.get(k, ntot)) for k in restart.FIELDS}
        nbytes = 8 * sum(a.size for a in fields.values())
        path = os.path.join(d, "restart.bin")
        t0 = time.perf_counter()
        restart.write(path, fields, itime=100, time=1.5)
        tw = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, _, back = restart.read(path)
        tr = time.perf_counter() - t0
        assert all(np.array_equal(back[k], fields[k]) for k in restart.FIELDS)
        res["restart"] = {"MB": nbytes / 1e6, "write_s": tw, "read_s": tr, "write_MB_per_s": nbytes / tw / 1e6,
                          "read_MB_per_s": nbytes / tr / 1e6}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(mesh, a, su, gpu_iters, budget_iters):
    """The oracle's sequential DIC-PCG (src/iccg.f90 restated in C) on one pinned host core.

    The pivots (the DIC factorisation) are timed on their own, then `budget_iters` iterations
    of the whole solver; the per-iteration time excludes the pivots and is scaled to the
    GPU's FP64 IC(0)-PCG iteration count.  oracle/calibration.json holds the ratio of the
    reference's own flang -O2 build to this restatement, measured on one core of the build
    container, which turns the figure into a reference-equivalent one."""
    from oracle import oracle as O
    csr = O.csr_build(mesh.n_cells, mesh.owner[:mesh.n_inner] + 1, mesh.neighbour + 1)
    allowed = os.sched_getaffinity(0)
    core = max(allowed)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        O.dic_pivots(csr, a)
        t_piv = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = O.iccg(csr, a, su, np.zeros(mesh.n_cells), sor=1e-300, nsw=budget_iters)
        t_all = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, allowed)
    per_iter = max(t_all - t_piv, 0.0) / max(r.iters, 1)
    value = (t_piv + per_iter * gpu_iters) * 1e3
    out = {"value": value, "unit": "ms", "cores": 1, "kind": "port",
           "sample": f"oracle DIC-PCG (C restatement of src/iccg.f90, gcc -O2 -ffp-contract=off, 1 thread pinned "
                     f"to core {core}): pivots timed alone, then {r.iters} iterations on the same system "
                     f"({t_all:.1f} s); value = pivots + per-iteration time x {gpu_iters} (the GPU FP64 "
                     f"IC(0)-PCG iteration count)",
           "pivots_ms": t_piv * 1e3, "ms_per_iter": per_iter * 1e3, "sample_iters": r.iters,
           "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "allowed_cores": len(allowed), "core": core}
    cal_path = os.path.join(ROOT, "oracle", "calibration.json")
    if os.path.exists(cal_path):
        with open(cal_path) as f:
            cal = json.load(f)
        ratio = cal["ratio_reference_over_oracle"]
        out["reference_equivalent_ms"] = value * ratio
        out["calibration"] = {"ratio_reference_over_oracle": ratio, "measured_on": cal["cpu_model"],
                              "mesh": cal["mesh"], "file": "oracle/calibration.json"}
    return out


class OracleSolve(threading.Thread):
    """The FP64 oracle's DIC-PCG (src/iccg.f90 restated in C) run to convergence on the bench
    system, on one pinned host core, beside the GPU work: the checker that states each GPU
    solution's error against the FP64 oracle.  It stops on the reference's own L1 rule at
    rsm < 1e-10, two decades tighter than the GPU solves' 1e-8, so its own error does not
    limit the comparison."""

    SOR = 1e-10

    def __init__(self, mesh, a, su):
        super().__init__(daemon=True)
        self.mesh, self.a, self.su = mesh, a, su
        self.x = None
        self.info = {}

    def run(self):
        from oracle import oracle as O
        allowed = os.sched_getaffinity(0)
        core = min(allowed)
        os.sched_setaffinity(0, {core})          # this thread only
        m = self.mesh
        csr = O.csr_build(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1)
        t0 = time.perf_counter()
        r = O.iccg(csr, self.a, self.su, np.zeros(m.n_cells), sor=self.SOR, nsw=20000)
        self.x = r.fi
        self.info = {"method": f"oracle DIC-PCG (C restatement of src/iccg.f90), L1 rsm < {self.SOR:g}, "
                               f"1 thread pinned to core {core}, beside the GPU runs",
                     "iters": r.iters, "seconds": time.perf_counter() - t0}


def host_rel_residual(s, x):
    """||su - A x||_2 / ||su||_2 in FP64 on the host, from the resident system's own CSR arrays."""
    import scipy.sparse as sp
    n = s.sizes()["n"]
    A = sp.csr_matrix((s.get("a"), s.get("ja") - 1, s.get("ioffset") - 1), shape=(n, n))
    su = s.get("su")
    return float(np.linalg.norm(su - A @ x) / np.linalg.norm(su))


def sweep_pair_bytes(n, nnz):
    """Compulsory bytes of one forward + backward sweep pair: per row and triangle its packed row
    id, count, pivot, input and output (4 + 4 + 8 + 8 + 8), and per strict-triangle entry its
    column and value (4 + 8); the strict entries of both triangles are nnz - n."""
    return 2 * 32 * n + 12 * (nnz - n)


def bench_case3(args, ctx):
    """Configuration 3: DILU-BiCGStab on the momentum-like convection-diffusion system of the
    permuted, graded Kuhn-tet box, FP64 and FP32-in-FP64 to 1e-8."""
    from freecappuccino_amd import fcap
    from freecappuccino_amd import workloads as W
    m = W.case_mesh(3, args.n)
    s = fcap.System(m, ctx)
    geo = fcap.mesh_geometry(m, ctx)
    u, v, w = W.taylor_green(geo.xc, geo.yc, geo.zc)
    flux = fcap.face_flux_central(m, geo, u, v, w, ctx)
    vis = np.full(m.n_cells, 1e-3)
    su = W.blob_rhs(geo.xc, geo.yc, geo.zc, geo.vol, 16, seed=5) / geo.vol.mean()
    s.assemble_convdiff(vis, flux, 0.7, su)
    results = {}
    for name, method in (("bicgstab_fp64", fcap.METHOD_BICGSTAB), ("bicgstab_mixed", fcap.METHOD_BICGSTAB_MIXED)):
        times = []
        for k in range(args.warmup + args.steps):
            s.set_solution(np.zeros(m.n_cells))
            t0 = time.perf_counter()
            rc, info = s.solve(method, fcap.STOP_L2, 1e-8, 20000)
            if k >= args.warmup:
                times.append((time.perf_counter() - t0) * 1e3)
        x = s.solution()
        results[name] = {"ms": float(np.median(times)), "status": fcap.STATUS_NAMES.get(rc, rc), "iters": info["iters"],
                         "outer": info["outer"], "rel_l2": info["rel_l2"], "rel_l2_host": host_rel_residual(s, x),
                         "solve_device_ms": info["solve_ms"] + info["setup_ms"]}
        results[name]["x"] = x
        print(f"{name}: { {k: v for k, v in results[name].items() if k != 'x'} }", file=sys.stderr, flush=True)
    x64 = results["bicgstab_fp64"].pop("x")
    xm = results["bicgstab_mixed"].pop("x")
    results["bicgstab_mixed"]["err_vs_fp64"] = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
    n, nnz = m.n_cells, m.nnz
    pair_ms = s.sweep_bench(dilu=True, reps=args.spmv_reps)
    pb = sweep_pair_bytes(n, nnz)
    gbs = pb / (pair_ms * 1e-3) / 1e9
    sizes = s.sizes()
    value = results["bicgstab_mixed"]["ms"]
    line = {
        "metric": "momentum-like solve time to 1e-8 rel. residual, DILU-BiCGStab (configuration 3)",
        "value": round(value, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(value, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"hex box {args.n}^3, each hex split into 6 Kuhn tets, graded 1.1 toward y=0 "
                               f"(aspect <= 100), cells permuted (seeded); upwind Taylor-Green convection, "
                               f"vis 1e-3, urf 0.7, 16 blobs; DILU-BiCGStab, FP32 Krylov in FP64 refinement",
                   "cells": n, "nnz": nnz, "levels_fwd": sizes["nlev_fwd"]},
        "solvers": results,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "DILU forward + backward sweep pair (src/bicgstab.f90:107-124)",
                     "launch_ms": pair_ms, "bytes_per_launch": pb},
    }
    s.close()
    return line


def bench_case4(args, ctx):
    """Configuration 4: streamed training of the network on 64^3 perturbed systems, batch 8."""
    from freecappuccino_amd import gnn, train_gnn
    net = gnn.Network(gnn.init_params(0), ctx)
    warm = train_gnn.stream_train(net, ctx, n=args.n, count=args.warmup, window=8, seed=99)
    res = train_gnn.stream_train(net, ctx, n=args.n, count=args.systems, window=8, seed=0)
    net.close()
    t = res["seconds"]
    per_sys = res["wall_s"] / res["count"]
    h = res["loss"]
    k = max(1, len(h) // 4)
    line = {
        "metric": "training throughput of the learned preconditioner (configuration 4): streamed systems per second",
        "value": round(1.0 / per_sys, 4), "unit": "systems/s", "n_gpus": 1, "steps": args.systems,
        "warmup": args.warmup, "ms_per_step": round(per_sys * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16+f32", "data": "synthetic",
        "config": {"workload": f"{args.systems} streamed hex boxes {args.n}^3 perturbed +-20% (one seed each), "
                               f"Laplacian mu=-1, Dirichlet x=0, 8 blobs; IC(0)-PCG (r, e) pairs after "
                               f"{list(gnn.SAMPLE_ITERS)} iterations; window of 8 resident systems, 1 Adam step "
                               f"of a batch of 8 pairs per new system",
                   "cells_per_system": args.n ** 3},
        "phases_ms_per_system": {k2: v / res["count"] * 1e3 for k2, v in t.items()},
        "projected_2000_systems_s": per_sys * 2000,
        "loss_first_quarter": float(np.mean(h[:k])), "loss_last_quarter": float(np.mean(h[-k:])),
        "warmup_wall_s": warm["wall_s"],
    }
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--case", type=int, default=5, choices=(1, 2, 3, 4, 5),
                    help="1: 40^3 box; 2: 128^3 perturbed; 3: 96^3 permuted graded Kuhn tets, DILU-BiCGStab; "
                         "4: streamed training on 64^3 perturbed systems; 5: 192^3 perturbed, x-graded 1.05")
    ap.add_argument("--systems", type=int, default=24,
                    help="--case 4: systems streamed in the timed run (configuration 4 names 2,000)")
    ap.add_argument("--n", type=int, default=None, help="cells per direction (overrides the case's size)")
    ap.add_argument("--methods", type=str, default=",".join(METHODS))
    ap.add_argument("--weights", type=str, default=None, help="network weights (.npy); default: the packaged ones")
    ap.add_argument("--cpu-iters", type=int, default=80, help="iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--network-layout", choices=("clusters", "cells"), default="clusters",
                    help="layout of the network's inference layers (fc_set_network_layout)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch each solver iteration's kernels one by one instead of replaying a "
                         "captured HIP graph (for profilers that do not follow graph launches)")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--profile-spmv", action="store_true",
                    help="only assemble and run --spmv-reps FP64 SpMVs plus a 1 GiB device copy "
                         "(the calibration kernel of the HBM counters), then exit")
    ap.add_argument("--profile-gnn", type=int, default=0, metavar="REPS",
                    help="only assemble and run REPS applications of the network, then exit")
    ap.add_argument("--profile-calcp", type=int, default=0, metavar="REPS",
                    help="only run REPS passes of the pressure-correction step's calls (gradients, "
                         "assembly, adjustMassFlow, corrections, non-orthogonal source), then exit")
    ap.add_argument("--profile-momentum", type=int, default=0, metavar="REPS",
                    help="only run REPS passes of calcuvw's three component solves, then exit")
    ap.add_argument("--profile-io", action="store_true",
                    help="only time the polyMesh reader and a restart write/read on the workload mesh, then exit")
    ap.add_argument("--io-threads", type=int, default=16, help="reader threads of --profile-io")
    ap.add_argument("--io-dir", type=str, default=None, help="scratch directory of --profile-io (default: TMPDIR)")
    ap.add_argument("--calcp-cpu-n", type=int, default=128,
                    help="cells per direction of the CPU baseline sample of --profile-calcp and "
                         "--profile-momentum (0: none)")
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

    args.n = args.n or {1: 40, 2: 128, 3: 96, 4: 64, 5: 192}[args.case]
    if args.case in (3, 4):
        ctx = fcap.Context(0)
        if args.no_graphs:
            ctx.set_graphs(False)
        line = bench_case3(args, ctx) if args.case == 3 else bench_case4(args, ctx)
        ctx.close()
        print(json.dumps(line))
        return
    mesh = W.case_mesh(args.case, args.n)
    if args.profile_io:
        print(json.dumps(profile_io(mesh, args.io_threads, args.io_dir)))
        return
    ctx = fcap.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    if args.network_layout == "cells":
        ctx.set_network_layout(fcap.NET_CELL_ORDER)
    if args.profile_calcp or args.profile_momentum:
        cpu_n = 0 if args.no_cpu else args.calcp_cpu_n
        if args.profile_calcp:
            print(json.dumps(profile_calcp(mesh, ctx, args.profile_calcp, cpu_n)))
        else:
            print(json.dumps(profile_momentum(mesh, ctx, args.profile_momentum, cpu_n)))
        ctx.close()
        return
    s = fcap.System(mesh, ctx)
    n, nnz = mesh.n_cells, mesh.nnz
    xc, yc, zc, vol = s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol")
    nblobs = 32 if args.case == 5 else 16
    su0 = W.blob_rhs(xc, yc, zc, vol, nblobs, seed=2)
    mu = np.full(n, -1.0)
    s.assemble_laplacian(mu, None, su0)
    sizes = s.sizes()
    oracle_solve = None
    if not (args.no_cpu or args.profile_spmv or args.profile_gnn):
        oracle_solve = OracleSolve(mesh, s.get("a"), s.get("su"))
        oracle_solve.start()
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

    net = zero_net = None
    layer_ms = None
    if any(m in methods for m in ("fcg_learned", "fcg_learned_no_coarse", "fcg_learned_every_iter")):
        net = gnn.Network.load(args.weights or gnn.DEFAULT_WEIGHTS, ctx)
    code = {"ic_pcg_fp64": fcap.METHOD_ICCG, "ic_pcg_mixed": fcap.METHOD_ICCG_MIXED,
            "fcg_learned": fcap.METHOD_FCG_LEARNED, "fcg_jacobi_poly": fcap.METHOD_FCG_JACOBI,
            "fcg_learned_no_coarse": fcap.METHOD_FCG_LEARNED, "fcg_learned_every_iter": fcap.METHOD_FCG_LEARNED,
            "ic_pcg_fp64_levels": fcap.METHOD_ICCG}

    def step(method):
        s.set_solution(np.zeros(n))
        t0 = time.perf_counter()
        s.assemble_laplacian_dev()
        rc, info = s.solve(method, fcap.STOP_L2, 1e-8, 20000)
        return rc, info, (time.perf_counter() - t0) * 1e3, s.last_ms(0)

    results, sols = {}, {}
    for name in methods:
        ctx.set_sweep_mode(fcap.SWEEP_LEVELS if name == "ic_pcg_fp64_levels" else fcap.SWEEP_DEFAULT)
        ctx.set_fcg_coarse(0 if name == "fcg_learned_no_coarse" else 4096)
        ctx.set_fcg_network_period(1 if name == "fcg_learned_every_iter" else 0)
        if name in ("fcg_learned", "fcg_learned_no_coarse", "fcg_learned_every_iter"):
            net.attach(s)
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
        results[name]["rel_l2_host"] = host_rel_residual(s, sols[name])
        print(f"{name}: {results[name]}", file=sys.stderr, flush=True)
    ctx.set_sweep_mode(fcap.SWEEP_DEFAULT)
    ctx.set_fcg_coarse(4096)
    ctx.set_fcg_network_period(0)
    coarse = s.coarse_space() if ("fcg_learned" in methods or "fcg_jacobi_poly" in methods) else None
    if "ic_pcg_fp64" in sols and "ic_pcg_fp64_levels" in sols:
        results["ic_pcg_fp64_levels"]["bitwise_equal_to_ic_pcg_fp64"] = bool(
            np.array_equal(sols["ic_pcg_fp64"], sols["ic_pcg_fp64_levels"]))
    if net is not None:
        layer_ms = net.bench_layer(s, args.spmv_reps)
        layout = s.network_layout()
        net.close()
    if zero_net is not None:
        zero_net.close()
    if oracle_solve is not None:
        oracle_solve.join()
        if oracle_solve.x is not None:
            xo = oracle_solve.x
            for name in results:
                results[name]["err_vs_oracle_fp64"] = float(np.linalg.norm(sols[name] - xo) / np.linalg.norm(xo))
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
        "dtype": {"ic_pcg_fp64": "f64", "ic_pcg_mixed": "f32+f64", "fcg_learned": "bf16+f32+f64",
                  "fcg_learned_no_coarse": "bf16+f32+f64", "fcg_learned_every_iter": "bf16+f32+f64",
                  "fcg_jacobi_poly": "f32+f64",
                  "ic_pcg_fp64_levels": "f64"}[head],
        "data": "synthetic",
        "config": {"workload": workload, "cells": n, "nnz": nnz, "levels_fwd": sizes["nlev_fwd"],
                   "solver": head,
                   "preconditioner": ("BF16 message-passing network (4 layers, width 32) + 3 damped-Jacobi "
                                      "steps, omega 0.8 + coarse correction over aggregates of the mesh; "
                                      "the network on every second iteration"
                                      if head == "fcg_learned" else "IC(0)")},
        "solvers": results,
        "spmv": {"fp64_ms": spmv_ms, "fp64_GBps": gbs, "fp64_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                 "fp32_ms": spmv32_ms,
                 "fp32_GBps": b32 / (spmv32_ms * 1e-3) / 1e9, "bytes_fp64": b64, "bytes_fp32": b32},
    }
    if coarse is not None:
        line["coarse"] = {"aggregates": coarse[0], "cells_per_aggregate": coarse[1], "factored": coarse[2]}
    if oracle_solve is not None and oracle_solve.x is not None:
        line["oracle_fp64"] = dict(oracle_solve.info, rel_l2_host=host_rel_residual(s, oracle_solve.x))
    if "fcg_learned" in results and "ic_pcg_fp64" in results:
        ratio = results["ic_pcg_fp64"]["ms"] / results["fcg_learned"]["ms"]
        line["target"] = {"what": "learned mixed-precision FCG at least 2x faster than GPU FP64 IC(0)-PCG to 1e-8",
                          "required": 2.0, "measured": round(ratio, 3), "met": bool(ratio >= 2.0)}
    sweep_ms = s.sweep_bench(dilu=False, reps=10)
    line["sweeps"] = {"kernel": "FP64 DIC forward + backward sweep pair (src/iccg.f90:94-111)",
                      "pair_ms": sweep_ms, "bytes": sweep_pair_bytes(n, nnz),
                      "GBps": sweep_pair_bytes(n, nnz) / (sweep_ms * 1e-3) / 1e9}
    if layer_ms is not None:
        if layout[0] > 0:
            lb = gnn_layer_bytes_cl(n, layout[0], layout[2])
            kname = "k_gnn_layer_cl<6,false,false> (hidden message-passing layer, 256-cell clusters)"
            traffic = LAYER_TRAFFIC_CL.get((args.case, args.n))
        else:
            lb = gnn_layer_bytes(n)
            kname = "k_gnn_layer_quad<false,false,6> (hidden message-passing layer, cell order)"
            traffic = LAYER_TRAFFIC.get((args.case, args.n))
        lgbs = lb / (layer_ms * 1e-3) / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": round(lgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(lgbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                            "launch_ms": layer_ms, "bytes_per_launch": lb}
        line["network_layout"] = {"clusters": layout[0], "halo_max": layout[1], "halo_rows": layout[2],
                                  "halo_rows_per_cell": round(layout[2] / n, 4)}
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
