import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")

MESH_CASES = ["hex4_uniform", "hex8_perturbed", "hex6x5x4_graded", "tet3_permuted", "hex12_perturbed"]
CONVDIFF_CASES = ["tet4_convdiff", "hex8_convdiff"]
CALCP_CASES = ["hex6_calcp", "tet3_calcp"]
GRAD_CASES = ["hex5_grad", "tet3_grad"]
ADJUST_CASES = ["hex5_adjust", "tet3_adjust"]
BC_CASES = ["hex5_bc", "tet3_bc"]
PISO_FIELDS = ("den", "apu", "u", "v", "w", "dudx", "dvdx", "dwdx")
CALCP_FIELDS = ("den", "apu", "apv", "apw", "u", "v", "w", "p", "dudx", "dvdx", "dwdx", "dpdx")


def calcp_uniform_case(n=4, rho=1.3, vel=(0.7, -0.4, 0.25), grad=(0.3, -1.1, 0.6)):
    """A box of n^3 unit cubes with constant density and velocity and the linear pressure
    p = grad . x with its exact gradient: there the Rhie-Chow terms of facefluxmass cancel
    (dpxi = Dpu g_x |d| n_x and dpex = Dpu g_x for |d| = 1), so the mass flux is rho u . S
    and cap = -rho (Vol/a_P)_f |S| / |d|."""
    from freecappuccino_amd import mesh as M

    m = M.hex_box(n)
    m.points = m.points * n
    rng = np.random.default_rng(3)
    fields = {"den": np.full(m.n_cells, rho), "apu": 0.5 + rng.random(m.n_cells),
              "apv": 0.5 + rng.random(m.n_cells), "apw": 0.5 + rng.random(m.n_cells)}
    return m, fields, np.array(vel), np.array(grad)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


class Golden:
    """One golden fixture: the mesh, the inputs and the reference's outputs."""

    def __init__(self, name):
        from freecappuccino_amd.mesh import PolyMesh

        self.name = name
        z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
        self.z = z
        self.mesh = PolyMesh(points=z["mesh_points"], face_ptr=z["mesh_face_ptr"],
                             face_nodes=z["mesh_face_nodes"], owner=z["mesh_owner"],
                             neighbour=z["mesh_neighbour"], n_cells=int(z["mesh_n_cells"][0]),
                             patches=[tuple(p) for p in json.loads(str(z["mesh_patches"]))])
        self.solves = json.loads(str(z["solves"]))

    def inp(self, k):
        return self.z[f"in_{k}"]

    def out(self, solve, k):
        return self.z[f"{solve}_{k}"]

    @property
    def first(self):
        return self.solves[0]


@pytest.fixture(scope="session")
def golden_loader():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = Golden(name)
        return cache[name]

    return get


def load_transcript():
    with open(os.path.join(GOLDEN, "solver5x5.json")) as f:
        return json.load(f)


def gnn_case(mesh, mu=-1.0):
    """Oracle-side graph of the learned preconditioner for the Laplacian of `mesh`:
    (geometry, CSR, matrix values, scipy matrix, oracle Graph)."""
    import scipy.sparse as sp
    from oracle import gnn_oracle as G
    from oracle import oracle as O

    geo = O.mesh_geometry(mesh)
    csr = O.csr_build(mesh.n_cells, mesh.owner[:mesh.n_inner] + 1, mesh.neighbour + 1)
    a, _ = O.laplacian(mesh, geo, csr, np.full(mesh.n_cells, mu), np.zeros(mesh.n_cells + mesh.n_boundary),
                       np.zeros(mesh.n_cells))
    p0, j0, d0 = csr.ioffset - 1, csr.ja - 1, csr.diag - 1
    kface = np.full(j0.shape[0], -1, np.int64)
    kface[csr.ij_slot - 1] = np.arange(mesh.n_inner)
    kface[csr.ji_slot - 1] = np.arange(mesh.n_inner)
    e1, e2, dg = G.edge_features_np(p0, j0, d0, kface, geo, a)
    graph = G.Graph(p0, j0, d0, a[d0], e1, e2)
    A = sp.csr_matrix((a, j0, p0), shape=(mesh.n_cells, mesh.n_cells))
    return geo, csr, a, A, graph, (e1, e2, dg)
