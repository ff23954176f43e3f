"""The learned-preconditioner oracle pinned on the CPU: finite differences, torch autograd,
closed forms, invariants.  (torch is imported inside the tests that need it, never at module
level: GPU test processes collect this module.)"""
import numpy as np
import scipy.sparse as sp

from freecappuccino_amd import gnn
from freecappuccino_amd import mesh as M
from oracle import gnn_oracle as G
from oracle import oracle as O


def _small_graph(n=4, seed=0):
    m = M.hex_box(n, perturb=0.2, seed=seed)
    geo = O.mesh_geometry(m)
    csr = O.csr_build(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1)
    a, _ = O.laplacian(m, geo, csr, np.full(m.n_cells, -1.0), np.zeros(m.n_cells + m.n_boundary),
                       np.zeros(m.n_cells))
    p0, j0, d0 = csr.ioffset - 1, csr.ja - 1, csr.diag - 1
    kface = np.full(j0.shape[0], -1, np.int64)
    kface[csr.ij_slot - 1] = np.arange(m.n_inner)
    kface[csr.ji_slot - 1] = np.arange(m.n_inner)
    e1, e2, dg = G.edge_features_np(p0, j0, d0, kface, geo, a)
    graph = G.Graph(p0, j0, d0, a[d0], e1, e2)
    A = sp.csr_matrix((a, j0, p0), shape=(m.n_cells, m.n_cells))
    return m, graph, A, (p0, j0, d0, a, e1, e2)


def test_layout_matches_the_package():
    assert G.N_PARAMS == gnn.N_PARAMS
    assert (G.OFF_LAYER0, G.LAYER_STRIDE, G.OFF_DEC_W) == (gnn.OFF_LAYER0, gnn.LAYER_STRIDE, gnn.OFF_DEC_W)


def test_bf16_rounding():
    assert G.bf16(1.0) == 1.0
    assert G.bf16(1.0 + 2.0 ** -8) == 1.0                    # tie -> even
    assert G.bf16(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7
    assert np.isnan(G.bf16(np.nan))


def test_zero_network_is_jacobi():
    """All-zero parameters: h = 0, y = 0, so z = s * u = r / a_ii exactly."""
    m, graph, A, _ = _small_graph()
    r = np.random.default_rng(1).standard_normal(m.n_cells)
    z = G.forward(np.zeros(G.N_PARAMS), graph, r, bf=False)
    np.testing.assert_allclose(z, r / graph.diag, rtol=1e-12)


def test_positive_homogeneity():
    m, graph, A, _ = _small_graph()
    th = gnn.init_params(3)
    r = np.random.default_rng(2).standard_normal(m.n_cells)
    z1 = G.forward(th, graph, r, bf=False)
    z2 = G.forward(th, graph, 7.5 * r, bf=False)
    np.testing.assert_allclose(z2, 7.5 * z1, rtol=1e-10)


def _probe_indices(rng):
    return np.concatenate([rng.choice(np.arange(G.OFF_ENC_W, G.OFF_LAYER0), 6, replace=False),
                           [G.OFF_LAYER0 + 5, G.OFF_LAYER0 + G.H * G.H + 7, G.OFF_LAYER0 + 2 * G.H * G.H + 3],
                           [G.OFF_LAYER0 + 2 * G.H * G.H + G.H + k for k in range(3)],
                           [G.OFF_LAYER0 + 3 * G.LAYER_STRIDE + 11, G.OFF_LAYER0 + 2 * G.LAYER_STRIDE + 1500,
                            G.OFF_DEC_W + 2, G.OFF_DEC_B]])


def test_backward_matches_finite_differences():
    """The explicit backward pass against central differences of the forward (float64, no BF16)."""
    m, graph, A, _ = _small_graph(n=3)
    rng = np.random.default_rng(4)
    th = gnn.init_params(5).astype(np.float64)
    th[G.OFF_DEC_W:G.OFF_DEC_W + G.H] *= 20.0          # make every block of parameters matter
    pairs = [(rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)) for _ in range(2)]
    loss, grad = G.loss_and_grad(th, graph, A, pairs, bf=False)
    eps = 1e-6
    for i in _probe_indices(rng):
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        lp, _ = G.loss_and_grad(tp, graph, A, pairs, bf=False)
        lm, _ = G.loss_and_grad(tm, graph, A, pairs, bf=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_smoothed_backward_matches_finite_differences():
    """With damped-Jacobi steps after the network in the loss, the backward pass (adjoint of the
    steps, D M^k D^-1) against central differences; the smoothed loss differs from the plain one,
    and zero steps give the plain loss exactly."""
    m, graph, A, _ = _small_graph(n=3)
    rng = np.random.default_rng(14)
    th = gnn.init_params(6).astype(np.float64)
    th[G.OFF_DEC_W:G.OFF_DEC_W + G.H] *= 20.0
    pairs = [(rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)) for _ in range(2)]
    loss, grad = G.loss_and_grad(th, graph, A, pairs, bf=False, smooth=3, omega=0.8)
    plain, _ = G.loss_and_grad(th, graph, A, pairs, bf=False)
    zero, _ = G.loss_and_grad(th, graph, A, pairs, bf=False, smooth=0, omega=0.8)
    assert zero == plain and abs(loss - plain) > 1e-3 * plain
    eps = 1e-6
    for i in _probe_indices(rng):
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        lp, _ = G.loss_and_grad(tp, graph, A, pairs, bf=False, smooth=3, omega=0.8)
        lm, _ = G.loss_and_grad(tm, graph, A, pairs, bf=False, smooth=3, omega=0.8)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_coarse_corrected_backward_matches_finite_differences():
    """With the Jacobi steps and the coarse correction in the loss, the backward pass (adjoint of
    the correction, dz - A P A_c^-1 P^T dz, then of the steps) against central differences; after
    the correction the residual of every aggregate is zero, whatever the network gives."""
    m, graph, A, _ = _small_graph(n=4)
    rng = np.random.default_rng(15)
    th = gnn.init_params(7).astype(np.float64)
    th[G.OFF_DEC_W:G.OFF_DEC_W + G.H] *= 20.0
    cell_agg = (np.arange(m.n_cells) * 4) // m.n_cells          # four aggregates of consecutive cells
    pairs = [(rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)) for _ in range(2)]
    loss, grad = G.loss_and_grad(th, graph, A, pairs, bf=False, smooth=2, omega=0.8, cell_agg=cell_agg)
    smoothed, _ = G.loss_and_grad(th, graph, A, pairs, bf=False, smooth=2, omega=0.8)
    assert abs(loss - smoothed) > 1e-3 * smoothed
    r = pairs[0][0]
    zc, P, _ = G.coarse_correct(rng.standard_normal(m.n_cells), r, A, cell_agg)
    assert np.abs(P.T @ (r - A @ zc)).max() < 1e-10 * np.abs(P.T @ r).max()
    eps = 1e-6
    for i in _probe_indices(rng):
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        lp, _ = G.loss_and_grad(tp, graph, A, pairs, bf=False, smooth=2, omega=0.8, cell_agg=cell_agg)
        lm, _ = G.loss_and_grad(tm, graph, A, pairs, bf=False, smooth=2, omega=0.8, cell_agg=cell_agg)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_smoothing_steps_are_jacobi_iterations():
    """smooth_steps from z = 0 is the Jacobi iteration for A z = r: its error contracts by
    (I - omega D^-1 A) each step, checked against the matrix power."""
    m, graph, A, _ = _small_graph(n=3)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(m.n_cells)
    x = np.linalg.solve(A.toarray() if hasattr(A, "toarray") else A, r)
    Mt = np.eye(m.n_cells) - 0.7 * (A.toarray() if hasattr(A, "toarray") else A) / graph.diag[:, None]
    z = G.smooth_steps(np.zeros(m.n_cells), r, A, graph.diag, 4, 0.7)
    np.testing.assert_allclose(z - x, np.linalg.matrix_power(Mt, 4) @ (-x), rtol=1e-10, atol=1e-12)


def test_backward_matches_torch_autograd():
    """The same gradient from torch autograd of an independent torch restatement (CPU)."""
    import torch
    m, graph, A, (p0, j0, d0, a, e1, e2) = _small_graph(n=4)
    rng = np.random.default_rng(8)
    th = gnn.init_params(9).astype(np.float64)
    r, e = rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)
    loss, grad = G.loss_and_grad(th, graph, A, [(r, e)], bf=False)
    t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    row, col = torch.as_tensor(graph.row), torch.as_tensor(graph.col)
    te1, te2 = torch.as_tensor(graph.e1), torch.as_tensor(graph.e2)
    dg = torch.as_tensor(graph.diag)
    rr, ee = torch.as_tensor(r), torch.as_tensor(e)
    v = rr / dg
    s = torch.sqrt(torch.mean(v * v)) + 1e-30
    u = v / s
    h = torch.relu(u[:, None] * t[0:32][None, :] + t[32:64][None, :])
    for l in range(G.L):
        o = G.OFF_LAYER0 + l * G.LAYER_STRIDE
        Ws, Wm = t[o:o + 1024].reshape(32, 32), t[o + 1024:o + 2048].reshape(32, 32)
        b, g = t[o + 2048:o + 2080], t[o + 2080:o + 2083]
        c = g[0] + g[1] * te1 + g[2] * te2
        agg = torch.zeros_like(h).index_add(0, row, c[:, None] * h[col])
        h = h + torch.relu(h @ Ws.T + agg @ Wm.T + b[None, :])
    z = s * (u + h @ t[G.OFF_DEC_W:G.OFF_DEC_W + 32] + t[G.OFF_DEC_B])
    At = torch.as_tensor(A.toarray())
    d = z - ee
    tl = (d @ (At @ d)) / (ee @ (At @ ee))
    tl.backward()
    assert abs(float(tl) - loss) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(grad, t.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_energy_loss_vanishes_for_the_exact_error():
    m, graph, A, _ = _small_graph()
    e = np.random.default_rng(6).standard_normal(m.n_cells)
    d = e - e
    assert float(d @ (A @ d)) == 0.0 and float(e @ (A @ e)) > 0.0
