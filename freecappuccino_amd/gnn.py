"""The learned preconditioner: network handle, parameter initialisation, training driver.

The network (message passing over the cell graph, hidden width 32, 4 layers,
BF16 weights/activations, MFMA node GEMMs) is evaluated and trained by the
library's kernels (csrc/fcap_gnn.hip); this module only drives them.  It is
the preconditioner of flexible CG (METHOD_FCG_LEARNED), with IC(0)-PCG as the
fallback on stagnation.

Training data (as for the 64^3 training case): (residual, error) pairs of
IC(0)-PCG iterates on assembled pressure systems of perturbed hex meshes, one
mesh seed per system.  The pairs come from IC(0)-PCG on the homogeneous
system A x = 0 started from seeded random vectors, recorded after 0, 1, 2, 4,
..., 32 iterations: the errors are exact (e = -x) and span rough to smooth,
which is what the preconditioner meets inside a Krylov method.  (Iterates
from zero on a smooth right-hand side only show smooth errors, which a
4-hop network cannot invert; networks trained on them help flexible CG far
less.)
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from typing import Iterable, List, Optional

import numpy as np

from . import fcap
from . import mesh as M
from . import workloads as W

H = 32
L = 4
LAYER_STRIDE = 2 * H * H + H + 3
OFF_ENC_W, OFF_ENC_B, OFF_LAYER0 = 0, H, 2 * H
OFF_DEC_W = OFF_LAYER0 + L * LAYER_STRIDE
OFF_DEC_B = OFF_DEC_W + H
N_PARAMS = OFF_DEC_B + 1

SAMPLE_ITERS = (0, 1, 2, 4, 8, 16, 32)

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "pressure_gnn.npy")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(seed: int) -> np.ndarray:
    """Scaled-uniform initialisation; the decoder starts small so z ~ Jacobi at step 0."""
    rng = np.random.default_rng(seed)
    th = np.zeros(N_PARAMS, np.float32)
    th[OFF_ENC_W:OFF_ENC_W + H] = rng.uniform(-1, 1, H)
    th[OFF_ENC_B:OFF_ENC_B + H] = rng.uniform(0, 0.1, H)
    sc = 0.5 / np.sqrt(H)
    for l in range(L):
        o = OFF_LAYER0 + l * LAYER_STRIDE
        th[o:o + H * H] = rng.uniform(-sc, sc, H * H)
        th[o + H * H:o + 2 * H * H] = rng.uniform(-sc, sc, H * H)
        th[o + 2 * H * H + H:o + 2 * H * H + H + 3] = [0.0, 0.5, 0.0]
    th[OFF_DEC_W:OFF_DEC_W + H] = rng.uniform(-0.1, 0.1, H) / np.sqrt(H)
    return th


class Network:
    def __init__(self, params: Optional[np.ndarray] = None, ctx: Optional[fcap.Context] = None):
        self.ctx = ctx or fcap.default_context()
        L_ = self.ctx.L
        assert L_.fc_gnn_num_params() == N_PARAMS
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        h = C.c_void_p()
        self.ctx.check(L_.fc_gnn_create(self.ctx.h, C.byref(h), None if p is None else _fp(p)))
        self.h = h
        self._attached = weakref.WeakSet()
        self.ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                for s in list(self._attached):
                    if getattr(s, "h", None):
                        self.ctx.L.fc_sys_set_network(s.h, None)
                self.ctx.L.fc_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        self.ctx.check(self.ctx.L.fc_gnn_get_params(self.h, _fp(out)))
        return out

    @params.setter
    def params(self, value):
        v = np.ascontiguousarray(value, np.float32)
        assert v.shape == (N_PARAMS,)
        self.ctx.check(self.ctx.L.fc_gnn_set_params(self.h, _fp(v)))

    def apply(self, system: fcap.System, r) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        self.ctx.check(self.ctx.L.fc_gnn_apply(self.h, system.h, _fp(r), _fp(z)))
        return z

    def bench(self, system: fcap.System, reps: int = 20) -> float:
        """Average ms of one application on `system` (device-resident input, HIP events)."""
        ms = self.ctx.L.fc_gnn_bench(self.h, system.h, int(reps))
        if ms < 0:
            raise fcap.FcError(-2, self.ctx.L.fc_last_error(self.ctx.h).decode())
        return ms

    def bench_layer(self, system: fcap.System, reps: int = 20) -> float:
        """Average ms of one launch of a hidden message-passing layer on `system` (HIP events)."""
        ms = self.ctx.L.fc_gnn_bench_layer(self.h, system.h, int(reps))
        if ms < 0:
            raise fcap.FcError(-2, self.ctx.L.fc_last_error(self.ctx.h).decode())
        return ms

    def edge_features(self, system: fcap.System):
        nnz = system.sizes()["nnz"]
        e1, e2, dg = np.zeros(nnz, np.float32), np.zeros(nnz, np.float32), np.zeros(system.n, np.float32)
        self.ctx.check(self.ctx.L.fc_gnn_edge_features(self.h, system.h, _fp(e1), _fp(e2), _fp(dg)))
        return e1, e2, dg

    def loss_grad(self, system: fcap.System, r, e):
        r = np.ascontiguousarray(r, np.float32)
        e = np.ascontiguousarray(e, np.float32)
        loss = C.c_double()
        g = np.zeros(N_PARAMS, np.float32)
        self.ctx.check(self.ctx.L.fc_gnn_loss_grad(self.h, system.h, _fp(r), _fp(e), C.byref(loss), _fp(g)))
        return loss.value, g

    def add_samples(self, system: fcap.System, iters: Iterable[int] = SAMPLE_ITERS, seed: int = 0):
        """(r, e) pairs of IC(0)-PCG on A x = 0 from seeded random starts, after each count in `iters`."""
        it = np.ascontiguousarray(list(iters), np.int32)
        self.ctx.check(self.ctx.L.fc_gnn_add_samples(self.h, system.h, it.shape[0],
                                                     it.ctypes.data_as(C.POINTER(C.c_int)), int(seed)))
        # the samples point at the system's matrix: closing the system drops them first
        system._sample_holders.add(self)

    @property
    def num_samples(self) -> int:
        return self.ctx.L.fc_gnn_num_samples(self.h)

    def sample(self, index: int, n: int):
        """(r, e) of training sample `index` (n = cells of its system)."""
        r, e = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self.ctx.check(self.ctx.L.fc_gnn_get_sample(self.h, index, _fp(r), _fp(e), n))
        return r, e

    def clear_samples(self):
        if getattr(self, "h", None):
            self.ctx.check(self.ctx.L.fc_gnn_clear_samples(self.h))

    def drop_system(self, system: fcap.System):
        """Forget the samples taken from `system` (it is about to be closed)."""
        if getattr(self, "h", None) and getattr(system, "h", None):
            self.ctx.check(self.ctx.L.fc_gnn_drop_system(self.h, system.h))

    def set_training_smoothing(self, steps: int = 3, omega: float = 0.8):
        """Take the solver's damped-Jacobi steps after the network into the training loss (and
        its gradient); 0 steps: the network's own output."""
        self.ctx.check(self.ctx.L.fc_gnn_set_training_smoothing(self.h, int(steps), float(omega)))

    def set_training_coarse(self, on: bool = True):
        """Take the solver's coarse correction (after the Jacobi steps) into the training loss
        and its gradient."""
        self.ctx.check(self.ctx.L.fc_gnn_set_training_coarse(self.h, int(bool(on))))

    def train(self, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(steps, np.float32)
        self.ctx.check(self.ctx.L.fc_gnn_train(self.h, steps, batch, float(lr), seed, _fp(losses)))
        return losses

    def attach(self, system: fcap.System):
        """Make this network the preconditioner of METHOD_FCG_LEARNED solves on `system`
        (the system keeps a borrowed pointer; closing the network detaches it)."""
        self.ctx.check(self.ctx.L.fc_sys_set_network(system.h, self.h))
        self._attached.add(system)

    def detach(self, system: fcap.System):
        self.ctx.check(self.ctx.L.fc_sys_set_network(system.h, None))
        self._attached.discard(system)

    def save(self, path: str = DEFAULT_WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.params)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, ctx: Optional[fcap.Context] = None) -> "Network":
        return cls(np.load(path, allow_pickle=False), ctx)


def graded_slab(nx: int = 192, nyz: int = 32, seed: int = 0):
    """The near-wall structure of the 192^3 graded mesh at 1/36 of its size: nx x nyz x nyz hex
    cells, interior vertices perturbed +-20%, x graded by 1.05 toward the x=0 wall (aspect up to
    100), and y, z scaled so that dy = dz = 1/nx as in the full mesh.  The network only sees
    4-hop neighbourhoods, which such a slab reproduces."""
    m = M.hex_box(nx, nyz, nyz, perturb=0.2, seed=seed, grade_axis=0, grade_ratio=1.05, grade_max_aspect=100.0)
    m.points = m.points.copy()
    m.points[:, 1:] *= nyz / nx
    return m


def training_system(n: int, seed: int, ctx: fcap.Context, nblobs: int = 8, kind: str = "perturbed") -> fcap.System:
    """A pressure system for training: kind "perturbed" is the n^3 perturbed hex mesh of the
    64^3 training case (seed); kind "graded" a graded slab (graded_slab, n cells along x)."""
    mesh = W.case_mesh(4, n, seed=seed) if kind == "perturbed" else graded_slab(n, max(8, n // 6), seed)
    s = fcap.System(mesh, ctx)
    xc, yc, zc, vol = s.get("xc"), s.get("yc"), s.get("zc"), s.get("vol")
    su0 = W.blob_rhs(xc, yc, zc, vol, nblobs, seed=10_000 + seed)
    s.assemble_laplacian(np.full(mesh.n_cells, -1.0), None, su0)
    return s


def build_dataset(net: Network, n: int, n_systems: int, iters=SAMPLE_ITERS, seed0: int = 0,
                  ctx: Optional[fcap.Context] = None, n_graded: int = 0, graded_n: int = 192) -> List[fcap.System]:
    """Assemble n_systems perturbed n^3 systems and n_graded graded slabs (graded_n cells along
    x) and add their IC(0)-PCG (r, e) pairs; returns the systems (closing one drops its samples)."""
    ctx = ctx or net.ctx
    systems = []
    for k in range(n_systems + n_graded):
        graded = k >= n_systems
        s = training_system(graded_n if graded else n, seed0 + k, ctx, kind="graded" if graded else "perturbed")
        net.add_samples(s, iters, seed=seed0 + k)
        systems.append(s)
    return systems
