"""The polyMesh reader (csrc/fcap_polymesh.cpp) on the reference's own example meshes and on
round trips.  Host code only: these run without a GPU.

Anchors: the header notes and boundary files of the cavity and pitzDaily meshes shipped with
the reference (examples/*/…-setup.tar.gz, copied as fixtures by
tests/golden/make_polymesh_fixtures.py), and what any valid mesh satisfies: closed cells
(sum of outward face areas is zero), positive volumes summing to the domain's, inner faces
with owner < neighbour in owner order.  (The reference's own OpenFOAM reader aborts under the
flang available here, so its output cannot be compared directly.)
"""
import os
import tarfile

import numpy as np
import pytest

from freecappuccino_amd import fcap
from freecappuccino_amd import mesh as M
from freecappuccino_amd.polymesh import read_polymesh
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CAVITY = os.path.join(GOLDEN, "polymesh_cavity")


@pytest.fixture(scope="module")
def pitzdaily(tmp_path_factory):
    d = tmp_path_factory.mktemp("pitz")
    with tarfile.open(os.path.join(GOLDEN, "polymesh_pitzDaily.tar.gz")) as t:
        t.extractall(d, filter="data")
    return str(d / "polyMesh")


def _closed_cells(m, geo):
    """Per cell: sum over its faces of the outward area vector (owner +S, neighbour -S)."""
    s = np.zeros((m.n_cells, 3))
    S = np.stack([geo.arx, geo.ary, geo.arz], axis=1)
    np.add.at(s, m.owner, S)
    np.add.at(s, m.neighbour, -S[:m.n_inner])
    return np.abs(s).max() / np.abs(S).max()


def _check_structure(m):
    m.validate()
    own_in = m.owner[:m.n_inner]
    assert np.all(own_in < m.neighbour)
    assert np.all(np.diff(own_in) >= 0)                        # inner faces in owner order


def test_cavity_counts_and_boundary():
    m = read_polymesh(CAVITY)
    # nPoints: 882 nCells: 400 nFaces: 1640 nInternalFaces: 760 (owner header note)
    assert (m.points.shape[0], m.n_cells, m.n_faces, m.n_inner) == (882, 400, 1640, 760)
    assert m.patches == [("wall", 20, 760), ("wall", 60, 780), ("symmetry", 800, 840)]
    assert np.all(np.diff(m.face_ptr) == 4)
    _check_structure(m)
    nwal, wstart, iwall = fcap.wall_info(m)
    assert (nwal, wstart) == (80, 760)                         # consecutive wall patches merge


def test_cavity_geometry_invariants():
    m = read_polymesh(CAVITY)
    geo = O.mesh_geometry(m)
    lo, hi = m.points.min(0), m.points.max(0)
    assert np.all(geo.vol > 0)
    assert abs(geo.vol.sum() - np.prod(hi - lo)) <= 1e-12 * np.prod(hi - lo)
    assert _closed_cells(m, geo) < 1e-12
    # a Laplacian solve on it converges (the mesh is usable end to end)
    csr = O.csr_build(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1)
    a, su = O.laplacian(m, geo, csr, np.full(m.n_cells, -1.0), np.zeros(m.n_cells + m.n_boundary),
                        geo.vol.copy())
    r = O.iccg(csr, a, su, np.zeros(m.n_cells), sor=1e-10, nsw=1000)
    assert r.iters < 1000 and np.all(np.isfinite(r.fi))


def test_pitzdaily_counts_and_geometry(pitzdaily):
    m = read_polymesh(pitzdaily)
    # nPoints: 25012 nCells: 12225 nFaces: 49180 nInternalFaces: 24170
    assert (m.points.shape[0], m.n_cells, m.n_faces, m.n_inner) == (25012, 12225, 49180, 24170)
    assert [p[0] for p in m.patches] == ["inlet", "outlet", "wall", "wall", "symmetry"]
    _check_structure(m)
    geo = O.mesh_geometry(m)
    assert np.all(geo.vol > 0)
    assert _closed_cells(m, geo) < 1e-10


def test_thread_count_does_not_change_the_result(pitzdaily):
    a = read_polymesh(pitzdaily, nthreads=1)
    for nt in (3, 8, 0):
        b = read_polymesh(pitzdaily, nthreads=nt)
        for k in ("points", "face_ptr", "face_nodes", "owner", "neighbour"):
            np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=f"{k} nthreads={nt}")


@pytest.mark.parametrize("fmt", ["foam", "native"])
@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_round_trip(tmp_path, fmt, kind):
    g = M.hex_box(5, perturb=0.2, seed=4) if kind == "hex" else M.tet_box(3, permute_seed=2)
    M.write_polymesh(g, str(tmp_path), fmt=fmt)
    r = read_polymesh(str(tmp_path), nthreads=2)
    for k in ("points", "face_ptr", "face_nodes", "owner", "neighbour"):
        np.testing.assert_array_equal(getattr(g, k), getattr(r, k), err_msg=k)
    assert r.n_cells == g.n_cells and r.patches == [tuple(p) for p in g.patches]


def _copy_cavity(dst):
    for f in os.listdir(CAVITY):
        with open(os.path.join(CAVITY, f), "rb") as a, open(os.path.join(dst, f), "wb") as b:
            b.write(a.read())


@pytest.mark.parametrize("case", ["missing", "bctype", "count", "vertex", "coverage"])
def test_errors_are_reported(tmp_path, case):
    d = str(tmp_path)
    if case != "missing":
        _copy_cavity(d)
    if case == "bctype":
        with open(os.path.join(d, "boundary"), "w") as f:
            f.write("#type nFaces startFace\nwall 80 760\nfrontAndBack 800 840\n")
    elif case == "count":
        t = open(os.path.join(d, "owner")).read().replace("\n1640\n(", "\n1641\n(")
        open(os.path.join(d, "owner"), "w").write(t)
    elif case == "vertex":
        t = open(os.path.join(d, "faces")).read().replace("4(1 22 463 442)", "4(1 22 463 9999)")
        open(os.path.join(d, "faces"), "w").write(t)
    elif case == "coverage":
        with open(os.path.join(d, "boundary"), "w") as f:
            f.write("#type nFaces startFace\nwall 80 760\nsymmetry 700 840\n")
    with pytest.raises(fcap.FcError) as ei:
        read_polymesh(d if case != "missing" else os.path.join(d, "nope"))
    assert str(ei.value)


# ------------------------------------------- the reference on meshes this reader reads

def _sha(x):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest()


def _oracle_chain(m, z):
    """Oracle geometry, CSR, Laplacian (mu, zero wall values, su0 seeded as the fixture's) and
    ICCG on a reader mesh."""
    geo = O.mesh_geometry(m)
    csr = O.csr_build(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1)
    su0 = z["in_su0"] if "in_su0" in z.files else np.random.default_rng(int(z["in_seed"][0])).standard_normal(m.n_cells)
    a, su = O.laplacian(m, geo, csr, z["in_mu"], z["in_phi"], su0)
    r = O.iccg(csr, a, su, z["in_fi0"], sor=float(z["in_sor"][0]), nsw=int(z["in_nsw"][0]))
    out = {k: getattr(geo, k) for k in ("arx", "ary", "arz", "xf", "yf", "zf", "vol", "xc", "yc", "zc", "facint")}
    out.update({k: getattr(csr, k) for k in ("ioffset", "ja", "diag", "ij_slot", "ji_slot")})
    out.update({"a": a, "su": su, "fi": r.fi, "res": r.res})
    return out, r


def test_oracle_matches_the_reference_on_the_cavity_mesh():
    """The reference's mesh_geometry, create_CSR_matrix_from_mesh_data, laplacian and iccg
    (oracle/_ref/ref_driver) on the cavity mesh as this reader reads it: the oracle's arrays
    bit for bit, and its iteration count and initial residual."""
    m = read_polymesh(CAVITY)
    z = np.load(os.path.join(GOLDEN, "cavity_reader.npz"), allow_pickle=False)
    out, r = _oracle_chain(m, z)
    for k, v in out.items():
        np.testing.assert_array_equal(v, z[f"iccg_{k}"], err_msg=k)
    assert r.iters == int(z["iccg_iters"][0]) and r.res0 == float(z["iccg_resor"][0])


def test_oracle_matches_the_reference_on_the_pitzdaily_mesh(pitzdaily):
    """The same chain on pitzDaily (12,225 cells): every array's SHA-256 equals the digest of
    the reference's array, and the iteration count and initial residual are the reference's."""
    m = read_polymesh(pitzdaily)
    z = np.load(os.path.join(GOLDEN, "pitzdaily_reader_digest.npz"), allow_pickle=False)
    out, r = _oracle_chain(m, z)
    for k, v in out.items():
        assert _sha(v) == str(z[f"iccg_{k}"]), k
    assert r.iters == int(z["iccg_iters"][0]) and r.res0 == float(z["iccg_resor"][0])
