"""Learned preconditioner on the GPU against the numpy oracle (oracle/gnn_oracle.py).

Edge features bitwise; the BF16/MFMA forward pass and the hand-written backward
pass within stated tolerances of the oracle evaluated at the same BF16
rounding points; training lowers the loss; flexible CG with the network
converges, falls back to IC(0)-PCG when the network output is not finite, and
a trained network needs fewer iterations than no network (Jacobi).
"""
import numpy as np
import pytest

from conftest import gnn_case
from freecappuccino_amd import fcap, gnn
from freecappuccino_amd import mesh as M
from freecappuccino_amd import workloads as W
from oracle import gnn_oracle as G

pytestmark = pytest.mark.gpu

# Tolerances.  The GPU evaluates in FP32 between the BF16 rounding points and
# rounds d_out to BF16 before its MFMA products; the oracle is float64 between
# the same points.  A value near a BF16 rounding boundary could round the other
# way, so comparisons are in relative L2 norm, not elementwise.  Measured on
# these cases (the tests print them): forward 8e-8..2e-7, loss 2e-9..4e-8,
# gradient 3e-4..4e-4 (worst parameter block 8e-4..9e-3); the bounds leave room
# for a few BF16 values rounding the other way, not for a wrong term.
FWD_RTOL = 1e-4       # ||y_gpu - y_oracle|| / ||y_oracle||, y = z/s - u (the network's part of z)
LOSS_RTOL = 1e-6
GRAD_COS = 0.9999     # cosine similarity of the full gradient vectors
GRAD_RTOL = 5e-3      # ||g_gpu - g_oracle|| / ||g_oracle||
BLOCK_RTOL = 5e-2     # the same, per parameter block


@pytest.fixture(scope="module")
def ctx():
    c = fcap.Context(0)
    yield c
    c.close()


def _system(mesh, ctx, seed=11):
    s = fcap.System(mesh, ctx)
    xc, yc, zc, vol = s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol")
    su0 = W.blob_rhs(xc, yc, zc, vol, 6, seed=seed) / vol.mean()
    s.assemble_laplacian(np.full(mesh.n_cells, -1.0), None, su0)
    return s


def _params(seed, dec_scale=20.0):
    th = gnn.init_params(seed).astype(np.float64)
    th[G.OFF_DEC_W:G.OFF_DEC_W + G.H] *= dec_scale     # make the network's share of z visible
    return th.astype(np.float32)


@pytest.fixture(scope="module")
def small(ctx):
    m = M.hex_box(6, perturb=0.2, seed=2)
    geo, csr, a, A, graph, feats = gnn_case(m)
    s = _system(m, ctx)
    np.testing.assert_array_equal(s.get("a"), a)
    yield m, A, graph, feats, s
    s.close()


def test_edge_features_bitwise(ctx, small):
    m, A, graph, (e1, e2, dg), s = small
    net = gnn.Network(None, ctx)
    g1, g2, gd = net.edge_features(s)
    np.testing.assert_array_equal(g1.astype(np.float64), e1)
    np.testing.assert_array_equal(g2.astype(np.float64), e2)
    np.testing.assert_array_equal(gd.astype(np.float64), dg)
    net.close()


def test_zero_network_is_jacobi(ctx, small):
    m, A, graph, feats, s = small
    net = gnn.Network(None, ctx)
    r = np.random.default_rng(3).standard_normal(m.n_cells).astype(np.float32)
    z = net.apply(s, r)
    np.testing.assert_allclose(z, r.astype(np.float64) / graph.diag, rtol=2e-6)
    net.close()


@pytest.mark.parametrize("pseed", [1, 7])
def test_forward_matches_oracle(ctx, small, pseed):
    m, A, graph, feats, s = small
    th = _params(pseed)
    net = gnn.Network(th, ctx)
    r = np.random.default_rng(10 + pseed).standard_normal(m.n_cells).astype(np.float32)
    z_gpu = net.apply(s, r).astype(np.float64)
    z_ref, cache, _ = G.forward(th, graph, r, bf=True, keep=True)
    sc, u = cache["s"], cache["u"]
    y_gpu, y_ref = z_gpu / sc - u, z_ref / sc - u
    assert np.linalg.norm(y_ref) > 0.1 * np.linalg.norm(u)        # the network part matters
    err = np.linalg.norm(y_gpu - y_ref) / np.linalg.norm(y_ref)
    print(f"forward pseed {pseed}: {err:.2e}")
    assert err < FWD_RTOL, err
    assert np.linalg.norm(z_gpu - z_ref) / np.linalg.norm(z_ref) < FWD_RTOL
    net.close()


@pytest.mark.parametrize("case", ["hex20_graded", "tet10_permuted"])
def test_forward_and_gradient_match_oracle_at_scale(ctx, case):
    """The forward pass (cluster layers, several clusters with halos) and the loss gradient on
    larger meshes than the 6^3 box: a 20^3 graded box (8,000 cells, 32 clusters) and 10^3 graded,
    permuted Kuhn tets (6,000 cells, ELL width 4), against the oracle at the same BF16 rounding
    points.  Measured (printed): forward 9e-8 / 2e-5, loss 9e-9 / 2e-8, gradient relative 7e-5 /
    6e-4 (hex / tets); the bounds leave about 10x."""
    m = (M.hex_box(20, perturb=0.2, seed=6, grade_axis=0, grade_ratio=1.1, grade_max_aspect=20.0)
         if case == "hex20_graded" else M.tet_box(10, grade_axis=1, grade_ratio=1.1, permute_seed=4))
    geo, csr, a, A, graph, feats = gnn_case(m)
    s = _system(m, ctx)
    th = _params(7)
    net = gnn.Network(th, ctx)
    try:
        clusters, _, _ = s.network_layout()
        assert clusters == (m.n_cells + 255) // 256
        rng = np.random.default_rng(31)
        r = rng.standard_normal(m.n_cells).astype(np.float32)
        z_gpu = net.apply(s, r).astype(np.float64)
        z_ref, cache, _ = G.forward(th, graph, r, bf=True, keep=True)
        sc, u = cache["s"], cache["u"]
        fwd = np.linalg.norm((z_gpu - z_ref) / sc) / np.linalg.norm(z_ref / sc - u)
        e = rng.standard_normal(m.n_cells).astype(np.float32)
        loss_gpu, g_gpu = net.loss_grad(s, r, e)
        loss_ref, g_ref = G.loss_and_grad(th, graph, A, [(r, e)], bf=True)
        g_gpu = g_gpu.astype(np.float64)
        cos = g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref))
        rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
        print(f"{case}: forward {fwd:.2e}, loss {abs(loss_gpu - loss_ref) / loss_ref:.2e}, "
              f"gradient cos {cos:.6f} rel {rel:.2e}")
        assert fwd < 1e-3
        assert abs(loss_gpu - loss_ref) <= 1e-6 * abs(loss_ref)
        assert cos > 0.9999 and rel < 5e-3, (cos, rel)
    finally:
        net.close()
        s.close()


@pytest.mark.parametrize("mesh", ["hex", "tet"])
def test_fused_first_layer_matches_stored_encoder_rows(ctx, mesh):
    """Inference computes the encoder rows inside the first layer from u (4 bytes gathered per
    neighbour); training's forward pass reads the rows the separate encoder pass stores.  They are
    the same BF16 values, so z is bitwise the same: with e = z of inference, the loss
    (z_train - e)^T A (z_train - e) / e^T A e that training evaluates is exactly 0, and one ulp
    off in one cell makes it positive.  ELL widths 6 (hex) and 4 (tet)."""
    m = M.hex_box(6, perturb=0.2, seed=2) if mesh == "hex" else M.tet_box(5, grade_axis=1, grade_ratio=1.1,
                                                                          permute_seed=3)
    th = _params(4)
    r = np.random.default_rng(21).standard_normal(m.n_cells).astype(np.float32)
    s = _system(m, ctx)
    net = gnn.Network(th, ctx)
    try:
        z = net.apply(s, r)
        assert np.abs(z).max() > 0
        loss, _ = net.loss_grad(s, r, z)
        assert loss == 0.0, loss
        e = z.copy()
        e[m.n_cells // 2] = np.nextafter(e[m.n_cells // 2], np.float32(np.inf))
        loss1, _ = net.loss_grad(s, r, e)
        assert loss1 > 0.0
    finally:
        net.close()
        s.close()


@pytest.mark.parametrize("case", ["hex6", "tet5", "hex23", "hex40_graded"])
def test_cluster_layout_matches_cell_order(ctx, case):
    """The inference layers over 256-cell clusters (rows renumbered, halos staged in LDS) give z
    bitwise equal to the cell-order layers: every message, MFMA product and epilogue value is
    formed with the same operations in the same order, so one wrong halo entry, local index or
    position anywhere in the four layers changes z.  Meshes: one partial cluster (6^3), permuted
    graded tets (ELL width 4, 3 clusters), 23^3 (12,167 cells, the last cluster partial) and a
    graded 40^3 box (250 clusters)."""
    m = {"hex6": lambda: M.hex_box(6, perturb=0.2, seed=2),
         "tet5": lambda: M.tet_box(5, grade_axis=1, grade_ratio=1.1, permute_seed=3),
         "hex23": lambda: M.hex_box(23, perturb=0.2, seed=4),
         "hex40_graded": lambda: M.hex_box(40, perturb=0.2, seed=5, grade_axis=0, grade_ratio=1.05,
                                           grade_max_aspect=100.0)}[case]()
    s = _system(m, ctx)
    net = gnn.Network(_params(4), ctx)
    try:
        clusters, halo_max, halo_total = s.network_layout()
        assert clusters == (m.n_cells + 255) // 256, (clusters, halo_max)
        assert 0 <= halo_max <= 384 and halo_max <= halo_total <= clusters * halo_max
        assert (halo_max > 0) == (clusters > 1)
        r = np.random.default_rng(5).standard_normal(m.n_cells).astype(np.float32)
        z_cl = net.apply(s, r)
        ctx.set_network_layout(fcap.NET_CELL_ORDER)
        z_cells = net.apply(s, r)
        assert np.abs(z_cells).max() > 0
        np.testing.assert_array_equal(z_cl, z_cells)
    finally:
        ctx.set_network_layout(fcap.NET_CLUSTERS)
        net.close()
        s.close()


@pytest.mark.parametrize("pseed,smooth", [(3, 0), (5, 0), (3, 3)])
def test_loss_and_gradient_match_oracle(ctx, small, pseed, smooth):
    """Loss and hand-written gradient against the oracle; with smooth = 3 the loss is taken after
    the solver's three damped-Jacobi steps (omega 0.8) and the gradient passes back through them."""
    m, A, graph, feats, s = small
    th = _params(pseed)
    net = gnn.Network(th, ctx)
    if smooth:
        net.set_training_smoothing(smooth, 0.8)
    rng = np.random.default_rng(20 + pseed)
    r = rng.standard_normal(m.n_cells).astype(np.float32)
    e = rng.standard_normal(m.n_cells).astype(np.float32)
    loss_gpu, g_gpu = net.loss_grad(s, r, e)
    loss_ref, g_ref = G.loss_and_grad(th, graph, A, [(r, e)], bf=True, smooth=smooth, omega=0.8)
    if smooth:
        loss_plain, _ = G.loss_and_grad(th, graph, A, [(r, e)], bf=True)
        assert abs(loss_ref - loss_plain) > 1e-2 * loss_plain   # the steps change the loss
    assert abs(loss_gpu - loss_ref) <= LOSS_RTOL * abs(loss_ref), (loss_gpu, loss_ref)
    g_gpu = g_gpu.astype(np.float64)
    cos = g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref))
    rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    blk = []
    print(f"gradient pseed {pseed} smooth {smooth}: loss {abs(loss_gpu - loss_ref) / abs(loss_ref):.2e} "
          f"cos {cos:.6f} rel {rel:.2e}")
    assert cos > GRAD_COS and rel < GRAD_RTOL, (cos, rel)
    # every parameter block receives its gradient (no block silently dropped)
    blocks = [(G.OFF_ENC_W, G.OFF_LAYER0), (G.OFF_DEC_W, G.N_PARAMS)]
    for l in range(G.L):
        o = G.OFF_LAYER0 + l * G.LAYER_STRIDE
        blocks += [(o, o + G.H * G.H), (o + G.H * G.H, o + 2 * G.H * G.H),
                   (o + 2 * G.H * G.H, o + 2 * G.H * G.H + G.H), (o + 2 * G.H * G.H + G.H, o + G.LAYER_STRIDE)]
    for lo, hi in blocks:
        nb = np.linalg.norm(g_ref[lo:hi])
        blk.append(np.linalg.norm(g_gpu[lo:hi] - g_ref[lo:hi]) / (nb + 1e-30))
        assert np.linalg.norm(g_gpu[lo:hi] - g_ref[lo:hi]) <= BLOCK_RTOL * nb + 1e-6 * np.linalg.norm(g_ref), (lo, hi)
    print(f"  worst parameter block {max(blk):.2e}")
    net.close()


def test_samples_are_ic0_pcg_iterates_of_the_homogeneous_system(ctx, small):
    """r = A e for every pair; e starts standard normal and gets smoother with the iteration count
    (the Rayleigh quotient e^T A^2 e / e^T A e falls as IC(0)-PCG removes the rough components)."""
    m, A, graph, feats, s = small
    net = gnn.Network(None, ctx)
    ks = (0, 1, 4, 8)
    net.add_samples(s, ks, seed=3)
    assert net.num_samples == len(ks)
    rq = []
    for i in range(len(ks)):
        r, e = net.sample(i, m.n_cells)
        e64 = e.astype(np.float64)
        Ae = A @ e64
        assert np.linalg.norm(r - Ae) <= 1e-4 * np.linalg.norm(Ae), i
        rq.append((Ae @ Ae) / (e64 @ Ae))
        if ks[i] == 0:
            assert abs(e64.mean()) < 0.25 and abs(e64.std() - 1.0) < 0.2
    assert rq[0] > rq[1] > rq[3], rq
    net.close()


def _train(ctx, n, n_systems, steps, seed0=0, lr=3e-3):
    net = gnn.Network(gnn.init_params(0), ctx)
    systems = gnn.build_dataset(net, n, n_systems, seed0=seed0, ctx=ctx)
    losses = net.train(steps, batch=8, lr=lr, seed=1)
    return net, systems, losses


@pytest.fixture(scope="module")
def trained(ctx):
    net, systems, losses = _train(ctx, 16, 4, 300)
    yield net, losses
    net.close()
    for s in systems:
        s.close()


def test_training_lowers_the_loss(trained):
    net, losses = trained
    assert np.all(np.isfinite(losses))
    first, last = losses[:20].mean(), losses[-20:].mean()
    assert last < 0.7 * first, (first, last)


def _fcg(ctx, s, params, sor=1e-8):
    net = gnn.Network(params, ctx)
    net.attach(s)
    s.set_solution(np.zeros(s.n))
    rc, info = s.solve(fcap.METHOD_FCG_LEARNED, fcap.STOP_L2, sor, 5000)
    net.detach(s)
    net.close()
    return rc, info


def test_fcg_learned_converges_and_beats_no_network(ctx, trained):
    net, _ = trained
    m = M.hex_box(16, perturb=0.2, seed=99)      # not a training mesh
    s = _system(m, ctx, seed=123)
    ctx.set_fcg_smoothing(0, 0.8)               # the network alone against Jacobi
    ctx.set_fcg_coarse(0)
    try:
        rc0, info0 = _fcg(ctx, s, np.zeros(gnn.N_PARAMS, np.float32))
        x0 = s.solution()
        rc1, info1 = _fcg(ctx, s, net.params)
        x1 = s.solution()
    finally:
        ctx.set_fcg_smoothing()
        ctx.set_fcg_coarse()
    assert rc0 == fcap.FC_OK and rc1 == fcap.FC_OK, (rc0, rc1, info0, info1)
    assert info0["rel_l2"] < 1e-8 and info1["rel_l2"] < 1e-8
    assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) < 1e-6
    assert info1["iters"] < 0.85 * info0["iters"], (info1["iters"], info0["iters"])
    s.close()


def test_fcg_with_zero_network_matches_jacobi_pcg(ctx):
    """A zero network is the fixed SPD preconditioner D^-1, and flexible CG with a fixed SPD
    preconditioner is PCG: to 1e-3, reached within the first inner FP32 solve (no restart of the
    refinement loop), the iteration count matches FP64 Jacobi-PCG up to FP32 rounding."""
    m = M.hex_box(12, perturb=0.2, seed=6)
    s = _system(m, ctx, seed=9)
    s.set_solution(np.zeros(s.n))
    rc_j, info_j = s.solve(fcap.METHOD_DPCG, fcap.STOP_L2, 1e-3, 5000)
    ctx.set_inner_tolerance(1e-3)       # the inner solve runs to the same 1e-3
    ctx.set_fcg_smoothing(0, 0.8)       # the network alone
    ctx.set_fcg_coarse(0)
    try:
        rc_f, info_f = _fcg(ctx, s, np.zeros(gnn.N_PARAMS, np.float32), sor=1e-3)
    finally:
        ctx.set_inner_tolerance(0.0)
        ctx.set_fcg_smoothing()
        ctx.set_fcg_coarse()
    assert rc_j == fcap.FC_OK and rc_f == fcap.FC_OK, (info_j, info_f)
    assert abs(info_f["iters"] - info_j["iters"]) <= 0.1 * info_j["iters"] + 5, (info_f["iters"], info_j["iters"])
    s.close()


def _poly_jacobi_pcg(A, b, steps, omega, tol, maxit=5000):
    """float64 PCG with M^-1 r = the Jacobi guess r/d followed by `steps` damped-Jacobi steps
    (a polynomial in D^-1 A times D^-1: symmetric, and positive for these omegas)."""
    d = A.diagonal()

    def prec(r):
        z = r / d
        for _ in range(steps):
            z = z + omega * (r - A @ z) / d
        return z

    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for it in range(1, maxit + 1):
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if np.linalg.norm(r) / nb < tol:
            return it
        z = prec(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return maxit


def test_fcg_smoothing_steps_match_polynomial_jacobi_pcg(ctx):
    """A zero network followed by 3 damped-Jacobi steps (omega 0.8) is a fixed SPD polynomial
    preconditioner, so flexible CG with it is PCG: the iteration count to 1e-3 matches a float64
    PCG with the same preconditioner in numpy up to FP32 rounding, and the steps cut the count
    against the network alone."""
    import scipy.sparse as sp
    m = M.hex_box(12, perturb=0.2, seed=6)
    s = _system(m, ctx, seed=9)
    n = s.n
    A = sp.csr_matrix((s.get("a"), s.get("ja") - 1, s.get("ioffset") - 1), shape=(n, n))
    # CG with a definite matrix and a preconditioner of the same sign takes the same steps
    # whether the pair is positive or negative definite
    ref = _poly_jacobi_pcg(A, s.get("su"), 3, 0.8, 1e-3)
    ctx.set_inner_tolerance(1e-3)
    ctx.set_fcg_coarse(0)
    try:
        ctx.set_fcg_smoothing(3, 0.8)
        rc3, info3 = _fcg(ctx, s, np.zeros(gnn.N_PARAMS, np.float32), sor=1e-3)
        ctx.set_fcg_smoothing(0, 0.8)
        rc0, info0 = _fcg(ctx, s, np.zeros(gnn.N_PARAMS, np.float32), sor=1e-3)
    finally:
        ctx.set_inner_tolerance(0.0)
        ctx.set_fcg_smoothing()
        ctx.set_fcg_coarse()
    assert rc3 == fcap.FC_OK and rc0 == fcap.FC_OK, (info3, info0)
    assert abs(info3["iters"] - ref) <= 0.1 * ref + 5, (info3["iters"], ref)
    assert info3["iters"] < 0.8 * info0["iters"], (info3["iters"], info0["iters"])
    s.close()


def test_learned_fcg_on_tets_reaches_fp64_tolerance(ctx, trained):
    """The learned flexible CG with the default Jacobi steps on a graded tetrahedral mesh the
    network was not trained on: 1e-8 in the recomputed FP64 residual, the solution that of FP64
    IC(0)-PCG."""
    net, _ = trained
    m = M.tet_box(8, grade_axis=0, grade_ratio=1.1, grade_max_aspect=4.0, permute_seed=2)
    s = _system(m, ctx, seed=77)
    try:
        s.set_solution(np.zeros(s.n))
        rc, info = s.solve(fcap.METHOD_ICCG, fcap.STOP_L2, 1e-12, 5000)
        x64 = s.solution()
        rc1, info1 = _fcg(ctx, s, net.params)
        assert rc == fcap.FC_OK and rc1 == fcap.FC_OK, (info, info1)
        assert info1["rel_l2"] < 1e-8 and info1["fallback_iters"] == 0, info1
        assert np.linalg.norm(s.solution() - x64) / np.linalg.norm(x64) < 1e-6
    finally:
        s.close()


def test_nonfinite_network_falls_back_to_ic0_pcg(ctx):
    m = M.hex_box(10, perturb=0.2, seed=5)
    s = _system(m, ctx, seed=8)
    rc, info = _fcg(ctx, s, np.full(gnn.N_PARAMS, np.nan, np.float32))
    assert rc == fcap.FC_FALLBACK_USED, (rc, info)
    assert info["fallback_iters"] > 0 and info["rel_l2"] < 1e-8
    s.close()


def test_stagnation_window_falls_back_to_ic0_pcg(ctx, trained):
    """Finite (trained) weights with the window rule of the inner solves made to fail (window 4,
    factor 0: no window after the first can beat 0 x the best before it): the first inner solve
    stops after 8 iterations and plain FP64 IC(0)-PCG finishes the solve from the refined
    iterate, with the caller's stopping rule (L2 relative to ||su||, or the reference's L1 rule
    relative to the first ||r0||_1) and what is left of nsw."""
    net, _ = trained
    m = M.hex_box(12, perturb=0.2, seed=5)
    s = _system(m, ctx, seed=8)
    n_net = gnn.Network(net.params, ctx)
    n_net.attach(s)
    ctx.set_fcg_stagnation(4, 0.0)
    ctx.set_inner_tolerance(1e-6)          # the first inner solve runs past 8 iterations
    try:
        for rule in (fcap.STOP_L2, fcap.STOP_L1):
            s.set_solution(np.zeros(s.n))
            rc, info = s.solve(fcap.METHOD_FCG_LEARNED, rule, 1e-8, 5000)
            assert rc == fcap.FC_FALLBACK_USED, (rule, info)
            assert info["iters"] == 8 and info["fallback_iters"] > 0, info
            if rule == fcap.STOP_L2:
                assert info["rel_l2"] < 1e-8, info
            else:
                assert info["rsm"] < 1e-8 and info["rel_l2"] < 1e-6, info
        # the fallback gets what is left of nsw: 8 FCG iterations + at most 4 IC(0)-PCG ones
        s.set_solution(np.zeros(s.n))
        rc, info = s.solve(fcap.METHOD_FCG_LEARNED, fcap.STOP_L2, 1e-8, 12)
        assert rc == fcap.FC_NOT_CONVERGED and info["iters"] == 8 and 1 <= info["fallback_iters"] <= 4, info
        # and with the test switched off the same weights converge without it
        ctx.set_fcg_stagnation(0, 0.81)
        s.set_solution(np.zeros(s.n))
        rc, info = s.solve(fcap.METHOD_FCG_LEARNED, fcap.STOP_L2, 1e-8, 5000)
        assert rc == fcap.FC_OK and info["fallback_iters"] == 0 and info["rel_l2"] < 1e-8, info
    finally:
        ctx.set_fcg_stagnation()
        ctx.set_inner_tolerance(0.0)
        n_net.detach(s)
        n_net.close()
        s.close()


def test_fcg_without_network_is_an_error(ctx):
    m = M.hex_box(4, seed=1)
    s = _system(m, ctx, seed=1)
    with pytest.raises(fcap.FcError):
        s.solve(fcap.METHOD_FCG_LEARNED, fcap.STOP_L2, 1e-8, 100)
    s.close()
