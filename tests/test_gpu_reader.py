"""The polyMesh reader feeding the GPU path: the reference's example meshes (cavity, 400
cells; pitzDaily, 12,225 cells) read by csrc/fcap_polymesh.cpp, handed to fc_sys_create, and
their geometry, CSR structure, Laplacian and IC(0)-PCG solve computed on the GPU.

Anchors: the reference's own mesh_geometry, create_CSR_matrix_from_mesh_data, laplacian and
iccg run on the same meshes (tests/golden/cavity_reader.npz holds its arrays,
pitzdaily_reader_digest.npz the SHA-256 of each of its arrays).  With the solver's sums in
index order every array is the reference's bit for bit.
"""
import hashlib
import os
import tarfile

import numpy as np
import pytest

from freecappuccino_amd import fcap
from freecappuccino_amd.polymesh import read_polymesh

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GEO = ("arx", "ary", "arz", "xf", "yf", "zf", "vol", "xc", "yc", "zc", "facint")
CSR = ("ioffset", "ja", "diag", "ij_slot", "ji_slot")


@pytest.fixture(scope="module")
def ctx():
    c = fcap.Context(0)
    yield c
    c.close()


def _sha(x):
    return hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest()


def _gpu_chain(m, z, ctx):
    s = fcap.System(m, ctx)
    try:
        out = {k: s.get(k) for k in GEO + CSR}
        su0 = z["in_su0"] if "in_su0" in z.files else np.random.default_rng(int(z["in_seed"][0])).standard_normal(
            m.n_cells)
        s.assemble_laplacian(z["in_mu"], z["in_phi"], su0)
        out["a"], out["su"] = s.get("a"), s.get("su")
        s.set_solution(z["in_fi0"])
        ctx.set_reduction_order(fcap.REDUCE_SEQUENTIAL)
        try:
            rc, info = s.solve(fcap.METHOD_ICCG, fcap.STOP_L1, float(z["in_sor"][0]), int(z["in_nsw"][0]))
        finally:
            ctx.set_reduction_order(fcap.REDUCE_TREE)
        out["fi"], out["res"] = s.solution(), s.get("res")
        return out, rc, info
    finally:
        s.close()


def test_cavity_reader_to_gpu_bitwise_vs_reference(ctx):
    m = read_polymesh(os.path.join(GOLDEN, "polymesh_cavity"))
    z = np.load(os.path.join(GOLDEN, "cavity_reader.npz"), allow_pickle=False)
    out, rc, info = _gpu_chain(m, z, ctx)
    for k, v in out.items():
        ref = z[f"iccg_{k}"]
        np.testing.assert_array_equal(v[:ref.size], ref, err_msg=k)
    assert rc == fcap.FC_OK and info["iters"] == int(z["iccg_iters"][0])
    assert info["res0"] == float(z["iccg_resor"][0])


def test_pitzdaily_reader_to_gpu_bitwise_vs_reference(ctx, tmp_path):
    with tarfile.open(os.path.join(GOLDEN, "polymesh_pitzDaily.tar.gz")) as t:
        t.extractall(tmp_path, filter="data")
    m = read_polymesh(str(tmp_path / "polyMesh"))
    z = np.load(os.path.join(GOLDEN, "pitzdaily_reader_digest.npz"), allow_pickle=False)
    out, rc, info = _gpu_chain(m, z, ctx)
    sizes = {"facint": m.n_inner}
    for k, v in out.items():
        assert _sha(v[:sizes[k]] if k in sizes else v) == str(z[f"iccg_{k}"]), k
    assert rc == fcap.FC_OK and info["iters"] == int(z["iccg_iters"][0])
    assert info["res0"] == float(z["iccg_resor"][0])
