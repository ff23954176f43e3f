"""CPU oracle of the learned preconditioner -- TEST INFRASTRUCTURE ONLY.

Defines, in numpy on the CPU, the message-passing network that
freecappuccino_amd/csrc/fcap_gnn.hip evaluates and trains on the GPU, with an
explicit backward pass, so the GPU forward pass, loss and hand-written
backward pass can be checked.  Only tests/ import this module.  (No torch
here: tests that run on the GPU import it, and a second HIP runtime must not
enter their process.)

Network (hidden width H = 32, L = 4 layers) on the cell graph of a CSR
matrix, for a residual r:
    s     = sqrt(mean((r_i / a_ii)^2)) + 1e-30
    u_i   = (r_i / a_ii) / s
    h0_i  = relu(w_enc * u_i + b_enc)
    for l in 0..L-1:
        c_k      = bf16(g0 + g1 * e1_k + g2 * e2_k)   (k: off-diagonal entry (i, j) of row i)
        agg_i    = sum_k c_k * h_j
        out_i    = Ws h_i + Wm agg_i + b
        h_i     <- h_i + relu(out_i)
    y_i   = w_dec . h_i + b_dec
    z_i   = s * (u_i + y_i)
Edge features of entry k in row i (face f between cells i and j):
    e1_k = (|S_f| / |x_j - x_i|) / a_ii,   e2_k = 1 - |S_f . d| / (|S_f| |d|)
BF16 points (as on the GPU): weights, the edge gates c_k, stored h, and agg
before the GEMM are rounded to bfloat16 (round to nearest even); the rest is
float64 here (FP32 on the GPU).  Rounding is identity in the backward pass.

Loss of a (residual, error) pair: L = (z - e)^T A (z - e) / (e^T A e).  With
`smooth` > 0 the z in the loss is the preconditioner's output after that many
damped-Jacobi steps z <- z + omega (r - A z) / a_ii, and the gradient passes
back through them: dz <- D M^k D^-1 dz with M = I - omega D^-1 A (A symmetric).

Pinning: the backward pass is checked against central finite differences of
the forward in float64 and against torch autograd (CPU) in tests/.
"""
from __future__ import annotations

import numpy as np

H = 32
L = 4
LAYER_STRIDE = 2 * H * H + H + 3
OFF_ENC_W, OFF_ENC_B = 0, H
OFF_LAYER0 = 2 * H
OFF_DEC_W = OFF_LAYER0 + L * LAYER_STRIDE
OFF_DEC_B = OFF_DEC_W + H
N_PARAMS = OFF_DEC_B + 1


def bf16(x):
    """Round to bfloat16 (nearest even) and return float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(nan, (u & 0xFFFF0000).astype(np.uint32) | np.where((u & 0xFFFF) != 0, 0x400000, 0).astype(np.uint32), r)
    return r.view(np.float32).astype(np.float64)


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    p = {"enc_w": t[OFF_ENC_W:OFF_ENC_W + H], "enc_b": t[OFF_ENC_B:OFF_ENC_B + H], "layers": []}
    for l in range(L):
        o = OFF_LAYER0 + l * LAYER_STRIDE
        p["layers"].append({"Ws": t[o:o + H * H].reshape(H, H), "Wm": t[o + H * H:o + 2 * H * H].reshape(H, H),
                            "b": t[o + 2 * H * H:o + 2 * H * H + H], "g": t[o + 2 * H * H + H:o + 2 * H * H + H + 3]})
    p["dec_w"] = t[OFF_DEC_W:OFF_DEC_W + H]
    p["dec_b"] = t[OFF_DEC_B]
    return p


class Graph:
    """Off-diagonal CSR entries as (row, col) pairs with edge features."""

    def __init__(self, ioffset0, ja0, diagpos0, a_diag, e1, e2):
        n = np.asarray(a_diag).shape[0]
        rows = np.repeat(np.arange(n), np.diff(ioffset0))
        off = np.ones(np.asarray(ja0).shape[0], bool)
        off[diagpos0] = False
        self.n = n
        self.row = rows[off]
        self.col = np.asarray(ja0)[off]
        self.e1 = np.asarray(e1, np.float64)[off]
        self.e2 = np.asarray(e2, np.float64)[off]
        self.diag = np.asarray(a_diag, np.float64)


def forward(theta, graph: Graph, r, bf=True, keep=False):
    rnd = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    p = unpack(rnd(theta))
    r = np.asarray(r, np.float64)
    v = r / graph.diag
    s = np.sqrt(np.mean(v * v)) + 1e-30
    u = v / s
    pre0 = u[:, None] * p["enc_w"][None, :] + p["enc_b"][None, :]
    h = rnd(np.maximum(pre0, 0.0))
    cache = {"u": u, "s": s, "pre0": pre0, "h": [h], "agg": [], "out": [], "c": []}
    for lp in p["layers"]:
        c = rnd(lp["g"][0] + lp["g"][1] * graph.e1 + lp["g"][2] * graph.e2)
        agg = np.zeros_like(h)
        np.add.at(agg, graph.row, c[:, None] * h[graph.col])
        agg = rnd(agg)
        out = h @ lp["Ws"].T + agg @ lp["Wm"].T + lp["b"][None, :]
        h = rnd(h + np.maximum(out, 0.0))
        cache["h"].append(h)
        cache["agg"].append(agg)
        cache["out"].append(out)
        cache["c"].append(c)
    y = h @ p["dec_w"] + p["dec_b"]
    z = s * (u + y)
    return (z, cache, p) if keep else z


def smooth_steps(z, r, A, diag, steps, omega):
    """`steps` damped-Jacobi steps z <- z + omega (r - A z) / diag."""
    z = np.asarray(z, np.float64)
    for _ in range(steps):
        z = z + omega * (np.asarray(r, np.float64) - A @ z) / diag
    return z


def coarse_correct(z, r, A, cell_agg):
    """z + P A_c^-1 P^T (r - A z), P the indicator of the aggregates cell_agg, A_c = P^T A P
    (dense FP64 solve)."""
    import scipy.sparse as sp
    n = z.shape[0]
    nc = int(cell_agg.max()) + 1
    P = sp.csr_matrix((np.ones(n), (np.arange(n), cell_agg)), shape=(n, nc))
    Ac = (P.T @ A @ P).toarray()
    return z + P @ np.linalg.solve(Ac, P.T @ (np.asarray(r, np.float64) - A @ z)), P, Ac


def loss_and_grad(theta, graph: Graph, A, pairs, bf=True, smooth=0, omega=0.8, cell_agg=None):
    """Mean loss over (r, e) pairs and its gradient (explicit backward pass); the loss is
    taken after `smooth` damped-Jacobi steps of the network's output and, with cell_agg (each
    cell's aggregate), the coarse correction after them."""
    grad = np.zeros(N_PARAMS)
    total = 0.0
    for r, e in pairs:
        z, cache, p = forward(theta, graph, r, bf, keep=True)
        e = np.asarray(e, np.float64)
        zs = smooth_steps(z, r, A, graph.diag, smooth, omega)
        if cell_agg is not None:
            zs, P, Ac = coarse_correct(zs, r, A, cell_agg)
        d = zs - e
        Ad = A @ d
        eAe = float(e @ (A @ e))
        total += float(d @ Ad) / eAe
        dz = 2.0 * Ad / eAe / len(pairs)
        if cell_agg is not None:
            # back through z_c = (I - P A_c^-1 P^T A) z + ...: dz <- dz - A P A_c^-1 P^T dz
            dz = dz - A @ (P @ np.linalg.solve(Ac, P.T @ dz))
        if smooth:
            # back through z_{k+1} = M z_k + omega D^-1 r: dz <- M^T dz = D M D^-1 dz
            t = dz / graph.diag
            for _ in range(smooth):
                t = t - omega * (A @ t) / graph.diag
            dz = graph.diag * t
        s, u = cache["s"], cache["u"]
        dy = s * dz
        h = cache["h"][-1]
        grad[OFF_DEC_W:OFF_DEC_W + H] += dy @ h
        grad[OFF_DEC_B] += dy.sum()
        dh = dy[:, None] * p["dec_w"][None, :]
        for l in range(L - 1, -1, -1):
            lp = p["layers"][l]
            h_in, agg, out, c = cache["h"][l], cache["agg"][l], cache["out"][l], cache["c"][l]
            dout = dh * (out > 0)
            o = OFF_LAYER0 + l * LAYER_STRIDE
            grad[o:o + H * H] += (dout.T @ h_in).ravel()
            grad[o + H * H:o + 2 * H * H] += (dout.T @ agg).ravel()
            grad[o + 2 * H * H:o + 2 * H * H + H] += dout.sum(0)
            dh_in = dh + dout @ lp["Ws"]
            dagg = dout @ lp["Wm"]
            np.add.at(dh_in, graph.col, c[:, None] * dagg[graph.row])
            dc = np.einsum("kc,kc->k", dagg[graph.row], h_in[graph.col])
            grad[o + 2 * H * H + H] += dc.sum()
            grad[o + 2 * H * H + H + 1] += (dc * graph.e1).sum()
            grad[o + 2 * H * H + H + 2] += (dc * graph.e2).sum()
            dh = dh_in
        dpre = dh * (cache["pre0"] > 0)
        grad[OFF_ENC_W:OFF_ENC_W + H] += dpre.T @ u
        grad[OFF_ENC_B:OFF_ENC_B + H] += dpre.sum(0)
    return total / len(pairs), grad


def edge_features_np(ioffset0, ja0, diagpos0, kface, geo, a):
    """Edge features (e1, e2) per CSR entry and the diagonal, from FP64 geometry and matrix.

    Evaluated in float64 in the kernel's order and rounded to bfloat16 (0 on the diagonal).
    """
    n = diagpos0.shape[0]
    rows = np.repeat(np.arange(n), np.diff(ioffset0))
    di = a[diagpos0][rows]
    off = np.ones(ja0.shape[0], bool)
    off[diagpos0] = False
    f = np.where(off, kface, 0)
    j = ja0
    sx, sy, sz = geo.arx[f], geo.ary[f], geo.arz[f]
    dx, dy, dz = geo.xc[j] - geo.xc[rows], geo.yc[j] - geo.yc[rows], geo.zc[j] - geo.zc[rows]
    sa = np.sqrt(sx * sx + sy * sy + sz * sz)
    dd = np.sqrt(dx * dx + dy * dy + dz * dz)
    with np.errstate(divide="ignore", invalid="ignore"):
        e1 = (sa / dd) / di
        e2 = 1.0 - np.abs(sx * dx + sy * dy + sz * dz) / (sa * dd)
    e1 = np.where(off, e1, 0.0)
    e2 = np.where(off, e2, 0.0)
    return bf16(e1), bf16(e2), a[diagpos0].astype(np.float32).astype(np.float64)
