"""GPU parity of every hot-path function against the oracle and the reference's golden vectors.

Bit-exact: geometry, CSR structure and slot maps, Laplacian and
convection-diffusion assembly, face fluxes and divergence, SpMV, DIC/DILU
pivots and the triangular sweeps (all run the reference's per-row order).
Solvers: the reference's iteration counts (+-1 for the reduction order) and
a solution within 1e-10 relative FP64 residual of the reference's.
"""
import numpy as np
import pytest

from conftest import CONVDIFF_CASES, MESH_CASES, load_transcript
from freecappuccino_amd import fcap
from freecappuccino_amd import mesh as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = fcap.Context(0)
    yield c
    c.close()


def _csr_of(g):
    m = g.mesh
    return fcap.CSR(*(g.out(g.first, k) for k in ("ioffset", "ja", "diag", "ij_slot", "ji_slot")))


def _geo_of(g):
    return fcap.Geometry(*(g.out(g.first, k) for k in ("arx", "ary", "arz", "xf", "yf", "zf", "vol", "xc", "yc",
                                                       "zc", "facint")))


@pytest.mark.parametrize("case", MESH_CASES + CONVDIFF_CASES)
def test_geometry_bitwise(ctx, golden_loader, case):
    g = golden_loader(case)
    geo = fcap.mesh_geometry(g.mesh, ctx)
    for k in ("arx", "ary", "arz", "xf", "yf", "zf", "vol", "xc", "yc", "zc", "facint"):
        np.testing.assert_array_equal(getattr(geo, k), g.out(g.first, k), err_msg=k)


@pytest.mark.parametrize("case", MESH_CASES + CONVDIFF_CASES)
def test_csr_bitwise(ctx, golden_loader, case):
    g = golden_loader(case)
    m = g.mesh
    c = fcap.create_csr_matrix_from_mesh_data(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1, ctx)
    for k in ("ioffset", "ja", "diag", "ij_slot", "ji_slot"):
        np.testing.assert_array_equal(getattr(c, k), g.out(g.first, k), err_msg=k)


@pytest.mark.parametrize("case", MESH_CASES)
def test_laplacian_bitwise(ctx, golden_loader, case):
    g = golden_loader(case)
    a, su = fcap.laplacian(g.mesh, _geo_of(g), _csr_of(g), g.inp("mu"), g.inp("phi"), g.inp("su0"), ctx)
    np.testing.assert_array_equal(a, g.out(g.first, "a"))
    np.testing.assert_array_equal(su, g.out(g.first, "su"))


def test_laplacian_wall_values_bitwise(ctx, golden_loader):
    """Non-zero Dirichlet values and a non-constant mu exercise su += a(diag)*phi_b (fvm_laplacian.f90:152)."""
    g = golden_loader("hex6x5x4_graded")
    m = g.mesh
    rng = np.random.default_rng(5)
    mu = -rng.uniform(0.5, 2.0, m.n_cells)
    phi = rng.standard_normal(m.n_cells + m.n_boundary)
    su0 = rng.standard_normal(m.n_cells)
    geo, csr = _geo_of(g), _csr_of(g)
    a, su = fcap.laplacian(m, geo, csr, mu, phi, su0, ctx)
    og = O.Geometry(*(getattr(geo, k) for k in ("arx", "ary", "arz", "xf", "yf", "zf", "vol", "xc", "yc", "zc",
                                                "facint")))
    oc = O.CSR(csr.ioffset, csr.ja, csr.diag, csr.ij_slot, csr.ji_slot)
    a_o, su_o = O.laplacian(m, og, oc, mu, phi, su0)
    np.testing.assert_array_equal(a, a_o)
    np.testing.assert_array_equal(su, su_o)


@pytest.mark.parametrize("case", CONVDIFF_CASES)
def test_convdiff_and_fluxes_bitwise(ctx, golden_loader, case):
    g = golden_loader(case)
    m = g.mesh
    geo, csr = _geo_of(g), _csr_of(g)
    flux = fcap.face_flux_central(m, geo, g.inp("u"), g.inp("v"), g.inp("w"), ctx)
    np.testing.assert_array_equal(flux, g.inp("flux"))
    a = fcap.convdiff(m, geo, csr, g.inp("vis"), flux, 0.7, ctx)
    np.testing.assert_array_equal(a, g.inp("a_convdiff"))
    su = fcap.rhs_divergence(m, flux, np.zeros(m.n_cells), ctx)
    np.testing.assert_array_equal(su, O.rhs_divergence(m, flux, np.zeros(m.n_cells)))


@pytest.mark.parametrize("case", MESH_CASES)
def test_spmv_pivots_sweeps_bitwise(ctx, golden_loader, case):
    g = golden_loader(case)
    csr = _csr_of(g)
    oc = O.CSR(csr.ioffset, csr.ja, csr.diag, csr.ij_slot, csr.ji_slot)
    a = g.out(g.first, "a")
    x = np.random.default_rng(0).standard_normal(csr.diag.shape[0])
    np.testing.assert_array_equal(fcap.spmv(csr, a, x, ctx), O.spmv(oc, a, x))
    d = fcap.dic_pivots(csr, a, ctx)
    np.testing.assert_array_equal(d, O.dic_pivots(oc, a))
    np.testing.assert_array_equal(fcap.lu_apply(csr, a, d, x, ctx), O.lu_apply(oc, a, d, x))


@pytest.mark.parametrize("case", CONVDIFF_CASES)
def test_dilu_pivots_bitwise(ctx, golden_loader, case):
    g = golden_loader(case)
    csr = _csr_of(g)
    oc = O.CSR(csr.ioffset, csr.ja, csr.diag, csr.ij_slot, csr.ji_slot)
    a = g.inp("a_convdiff")
    d = fcap.dilu_pivots(csr, a, ctx)
    np.testing.assert_array_equal(d, O.dilu_pivots(oc, a))
    r = np.random.default_rng(1).standard_normal(csr.diag.shape[0])
    np.testing.assert_array_equal(fcap.lu_apply(csr, a, d, r, ctx), O.lu_apply(oc, a, d, r))


def _dense(csr, a):
    import scipy.sparse as sp
    n = csr.diag.shape[0]
    return sp.csr_matrix((a, csr.ja - 1, csr.ioffset - 1), shape=(n, n))


@pytest.mark.parametrize("case", MESH_CASES)
@pytest.mark.parametrize("solver", ["iccg", "dpcg", "bicgstab"])
def test_solvers_match_reference(ctx, golden_loader, case, solver):
    """The GPU solve reaches the reference's solution: relative FP64 difference of the residuals <= 1e-10."""
    g = golden_loader(case)
    csr = _csr_of(g)
    a, su = g.out(g.first, "a"), g.out(g.first, "su")
    r = getattr(fcap, solver)(csr, a, su, g.inp("fi0"), sor=float(g.inp("sor")[0]), nsw=int(g.inp("nsw")[0]),
                              ctx=ctx)
    ref_iters = int(g.out(solver, "iters")[0])
    assert r.status == 0
    assert abs(r.info["iters"] - ref_iters) <= 1, (r.info["iters"], ref_iters)
    A = _dense(csr, a)
    fi_ref = g.out(solver, "fi")
    bn = np.linalg.norm(su)
    # both solutions are solutions of the same system to the same tolerance
    assert np.linalg.norm(A @ (r.fi - fi_ref)) / bn <= 1e-6
    res_gpu = np.linalg.norm(su - A @ r.fi) / bn
    res_ref = np.linalg.norm(su - A @ fi_ref) / bn
    assert abs(res_gpu - res_ref) <= 1e-10 + 0.5 * res_ref
    assert abs(r.info["rel_l2"] - res_gpu) <= 1e-12 + 1e-6 * res_gpu


@pytest.mark.parametrize("case", CONVDIFF_CASES)
def test_bicgstab_convdiff_matches_reference(ctx, golden_loader, case):
    g = golden_loader(case)
    csr = _csr_of(g)
    a, su = g.inp("a_convdiff"), g.inp("su_convdiff")
    r = fcap.bicgstab(csr, a, su, g.inp("fi0"), sor=float(g.inp("sor")[0]), nsw=int(g.inp("nsw")[0]), ctx=ctx)
    assert r.status == 0
    assert abs(r.info["iters"] - int(g.out("bicgstab", "iters")[0])) <= 1
    A = _dense(csr, a)
    bn = np.linalg.norm(su)
    assert np.linalg.norm(A @ (r.fi - g.out("bicgstab", "fi"))) / bn <= 1e-6


def test_solver_transcript_5x5(ctx):
    """tests/output.txt: iteration counts 6/5/5, res0 lines and the solutions (warm starts as in the reference)."""
    t = load_transcript()
    csr = fcap.CSR(np.array(t["ioffset"], np.int32), np.array(t["ja"], np.int32), np.array(t["diag"], np.int32),
                   np.zeros(0, np.int32), np.zeros(0, np.int32))
    r1 = fcap.bicgstab(csr, t["a1"], t["b1"], np.zeros(5), ctx=ctx, variant=fcap.VARIANT_TESTS)
    r2 = fcap.iccg(csr, t["a2"], t["b2"], r1.fi, ctx=ctx, variant=fcap.VARIANT_TESTS)
    r3 = fcap.dpcg(csr, t["a2"], t["b2"], r2.fi, ctx=ctx, variant=fcap.VARIANT_TESTS)
    assert (r1.info["iters"], r2.info["iters"], r3.info["iters"]) == (6, 5, 5)
    assert f"{r1.info['res0']:10.3E}" == " 2.903E+01"
    assert f"{r2.info['res0']:10.3E}" == " 3.725E+01"
    np.testing.assert_allclose(r1.fi, np.linalg.solve(np.array(t["a1"]).reshape(5, 5), t["b1"]), rtol=1e-12)
    np.testing.assert_allclose(r3.fi, np.linalg.solve(np.array(t["a2"]).reshape(5, 5), t["b2"]), rtol=1e-12)
    assert [f"{v:5.2f}" for v in r1.fi] == [f"{v:5.2f}" for v in t["xa1"]]


def test_early_exit_and_status_codes(ctx, golden_loader):
    g = golden_loader("hex4_uniform")
    csr = _csr_of(g)
    a = g.out(g.first, "a")
    x = np.random.default_rng(2).standard_normal(csr.diag.shape[0])
    su = O.spmv(O.CSR(csr.ioffset, csr.ja, csr.diag, csr.ij_slot, csr.ji_slot), a, x)
    r = fcap.iccg(csr, a, su, x, ctx=ctx)          # exact initial guess: res0 < 1e-13
    assert r.status == fcap.FC_EARLY_EXIT and r.info["iters"] == 0
    r = fcap.dpcg(csr, a, g.out(g.first, "su"), np.zeros_like(x), sor=1e-30, nsw=3, ctx=ctx)
    assert r.status == fcap.FC_NOT_CONVERGED and r.info["iters"] == 3
    bad = a.copy()
    bad[csr.diag[0] - 1] = np.nan
    with pytest.raises(fcap.FcError):
        fcap.iccg(csr, bad, g.out(g.first, "su"), np.zeros_like(x), ctx=ctx)


def test_structure_errors_are_reported(ctx):
    own1 = np.array([1, 1, 2], np.int32)
    nei1 = np.array([2, 2, 3], np.int32)            # duplicate pair (1,2)
    with pytest.raises(fcap.FcError) as e:
        fcap.create_csr_matrix_from_mesh_data(3, own1, nei1, ctx)
    assert e.value.status == -5
    with pytest.raises(fcap.FcError):
        fcap.create_csr_matrix_from_mesh_data(3, np.array([1], np.int32), np.array([7], np.int32), ctx)
