"""The benchmark configurations at their stated sizes on the GPU.

  1  hex box 40^3 (64,000 cells), Laplacian mu = -1, Dirichlet x = 0, RHS the discrete
     divergence of a seeded velocity with components U[-1, 1]; ICCG and DPCG to 1e-8.
  2  hex box 128^3 (2.10M cells), interior vertices +-20% (seeded), 16 Gaussian blobs: FP64
     against FP32-in-FP64 IC(0)-PCG.
  4  training the learned preconditioner on 64^3 perturbed systems, mini-batches of 8.
  5  hex box 192^3 (7.08M cells) perturbed and x-graded 1.05, 32 blobs: the learned flexible CG
     with the packaged weights against FP64 IC(0)-PCG.
(Configuration 3, the 96^3 permuted Kuhn-tet BiCGStab, is
tests/test_gpu_system.py::test_convdiff_tet_workload_bicgstab.)  Residuals are recomputed on
the host in FP64 from the assembled matrix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from freecappuccino_amd import fcap, gnn
from freecappuccino_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = fcap.Context(0)
    yield c
    c.close()


def _host_rel(s):
    n = s.n
    A = sp.csr_matrix((s.get("a"), s.get("ja") - 1, s.get("ioffset") - 1), shape=(n, n))
    su, x = s.get("su"), s.solution()
    return float(np.linalg.norm(su - A @ x) / np.linalg.norm(su)), x


def test_config1_box40_bitwise_vs_oracle(ctx):
    """Geometry, CSR, the Laplacian with its wall terms and the divergence right-hand side are
    the oracle's bits; with the sums in index order ICCG and DPCG to 1e-8 give the oracle's
    solution and iteration count bit for bit, and with the default sums the iteration count
    within one."""
    m = W.case_mesh(1, 40)
    n = m.n_cells
    geo = O.mesh_geometry(m)
    csr = O.csr_build(n, m.owner[:m.n_inner] + 1, m.neighbour + 1)
    u, v, w = W.cell_velocity(n, seed=1)
    su0 = O.rhs_divergence(m, O.face_flux_central(m, geo, u, v, w), np.zeros(n))
    mu, phi = np.full(n, -1.0), np.zeros(n + m.n_boundary)
    a_o, su_o = O.laplacian(m, geo, csr, mu, phi, su0)
    s = fcap.System(m, ctx)
    try:
        gflux = fcap.face_flux_central(m, fcap.Geometry(*(s.get(k) for k in (
            "arx", "ary", "arz", "xf", "yf", "zf", "vol", "xc", "yc", "zc", "facint"))), u, v, w, ctx)
        np.testing.assert_array_equal(fcap.rhs_divergence(m, gflux, np.zeros(n), ctx), su0)
        s.assemble_laplacian(mu, phi, su0)
        np.testing.assert_array_equal(s.get("a"), a_o)
        np.testing.assert_array_equal(s.get("su"), su_o)
        for method, ofn in ((fcap.METHOD_ICCG, O.iccg), (fcap.METHOD_DPCG, O.dpcg)):
            ref = ofn(csr, a_o, su_o, np.zeros(n), sor=1e-8, nsw=5000)
            s.set_solution(np.zeros(n))
            ctx.set_reduction_order(fcap.REDUCE_SEQUENTIAL)
            try:
                rc, info = s.solve(method, fcap.STOP_L1, 1e-8, 5000)
            finally:
                ctx.set_reduction_order(fcap.REDUCE_TREE)
            assert rc == fcap.FC_OK and info["iters"] == ref.iters, (method, info, ref.iters)
            np.testing.assert_array_equal(s.solution(), ref.fi)
            s.set_solution(np.zeros(n))
            rc, info = s.solve(method, fcap.STOP_L1, 1e-8, 5000)
            assert rc == fcap.FC_OK and abs(info["iters"] - ref.iters) <= 1, (method, info, ref.iters)
            print(f"config 1, method {method}: {info['iters']} iterations (oracle {ref.iters}), "
                  f"rel_l2 {info['rel_l2']:.2e}")
    finally:
        s.close()


def test_config2_box128_fp64_and_mixed_reach_1e8(ctx):
    """FP64 IC(0)-PCG and FP32 IC(0)-PCG inside FP64 refinement both reach 1e-8 in the FP64
    relative residual recomputed on the host; the mixed solution's error against the FP64 one is
    stated and below 1e-6."""
    m = W.case_mesh(2, 128)
    s = fcap.System(m, ctx)
    try:
        su0 = W.blob_rhs(s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol"), 16, seed=2)
        s.assemble_laplacian(np.full(m.n_cells, -1.0), None, su0)
        out = {}
        for name, method in (("fp64", fcap.METHOD_ICCG), ("mixed", fcap.METHOD_ICCG_MIXED)):
            s.set_solution(np.zeros(m.n_cells))
            rc, info = s.solve(method, fcap.STOP_L2, 1e-8, 20000)
            rel, x = _host_rel(s)
            assert rc == fcap.FC_OK and rel < 1e-8 and abs(rel - info["rel_l2"]) < 1e-12, (name, rel, info)
            out[name] = (x, info)
        err = np.linalg.norm(out["mixed"][0] - out["fp64"][0]) / np.linalg.norm(out["fp64"][0])
        print(f"config 2: FP64 {out['fp64'][1]['iters']} it {out['fp64'][1]['solve_ms']:.1f} ms, mixed "
              f"{out['mixed'][1]['iters']} it {out['mixed'][1]['solve_ms']:.1f} ms, error vs FP64 {err:.2e}")
        assert err < 1e-6
    finally:
        s.close()


def test_config4_training_at_64_cubed_batch_8(ctx):
    """Pairs of IC(0)-PCG iterates on two 64^3 perturbed systems (configuration 4's meshes),
    Adam on mini-batches of 8: the loss falls, and every step's loss is finite."""
    net = gnn.Network(gnn.init_params(0), ctx)
    systems = gnn.build_dataset(net, 64, 2, seed0=4000, ctx=ctx)
    try:
        assert net.num_samples == 2 * len(gnn.SAMPLE_ITERS)
        losses = net.train(200, batch=8, lr=3e-3, seed=1)
        assert np.all(np.isfinite(losses))
        first, last = losses[:30].mean(), losses[-30:].mean()
        print(f"config 4: loss {first:.4f} -> {last:.4f} over 200 steps of 8 at 64^3")
        assert last < 0.8 * first, (first, last)
    finally:
        net.clear_samples()
        for s in systems:
            s.close()
        net.close()


def test_config5_box192_learned_fcg_reaches_1e8(ctx):
    """The learned flexible CG with the packaged weights reaches 1e-8 in the FP64 residual
    recomputed on the host, with no IC(0)-PCG fallback, and its solution is FP64 IC(0)-PCG's
    within 1e-6; its coarse correction needs fewer than half the iterations of the same solve
    without it."""
    m = W.case_mesh(5, 192)
    s = fcap.System(m, ctx)
    net = gnn.Network.load(gnn.DEFAULT_WEIGHTS, ctx)
    try:
        su0 = W.blob_rhs(s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol"), 32, seed=2)
        s.assemble_laplacian(np.full(m.n_cells, -1.0), None, su0)
        s.set_solution(np.zeros(m.n_cells))
        rc, info64 = s.solve(fcap.METHOD_ICCG, fcap.STOP_L2, 1e-8, 20000)
        rel64, x64 = _host_rel(s)
        assert rc == fcap.FC_OK and rel64 < 1e-8
        net.attach(s)
        s.set_solution(np.zeros(m.n_cells))
        rc, info = s.solve(fcap.METHOD_FCG_LEARNED, fcap.STOP_L2, 1e-8, 20000)
        rel, x = _host_rel(s)
        err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
        print(f"config 5: learned FCG {info['iters']} it, {info['solve_ms']:.0f} ms, rel {rel:.2e}, error vs FP64 "
              f"IC(0)-PCG {err:.2e}; IC(0)-PCG {info64['iters']} it, {info64['solve_ms']:.0f} ms")
        assert rc == fcap.FC_OK and info["fallback_iters"] == 0 and rel < 1e-8, info
        assert err < 1e-6
        # the coarse correction over 4,096-cell aggregates (default) against none
        ctx.set_fcg_coarse(0)
        s.set_solution(np.zeros(m.n_cells))
        rc0, info0 = s.solve(fcap.METHOD_FCG_LEARNED, fcap.STOP_L2, 1e-8, 20000)
        print(f"config 5: learned FCG without the coarse correction {info0['iters']} it, "
              f"{info0['solve_ms']:.0f} ms")
        assert rc0 == fcap.FC_OK and info0["rel_l2"] < 1e-8, info0
        assert info["iters"] < 0.5 * info0["iters"], (info["iters"], info0["iters"])
    finally:
        ctx.set_fcg_coarse()
        net.close()
        s.close()
