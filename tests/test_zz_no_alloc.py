"""fc_dpcg_noalloc (no device allocation: its buffers live in a static device
array) against fc_dpcg on the same inputs, through the C ABI.

Two calls in a row, the second on a smaller system, so the second call runs on
the scratch the first one left behind: both give fc_dpcg's solution, residual
vector and iteration count bit for bit (the same kernels and the same sums).
Sizes beyond the scratch and host pointers are refused.
"""
import ctypes as C

import numpy as np
import pytest

from freecappuccino_amd import fcap
from freecappuccino_amd import mesh as M
from freecappuccino_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 20


class _Dev:
    """hipMalloc'd copies of host arrays, freed on close."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, x):
        x = np.ascontiguousarray(x)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(x.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, x.ctypes.data, x.nbytes, 1) == 0
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _system(nside, seed):
    m = M.hex_box(nside, perturb=0.2, seed=seed)
    geo = O.mesh_geometry(m)
    csr = O.csr_build(m.n_cells, m.owner[:m.n_inner] + 1, m.neighbour + 1)
    su0 = W.blob_rhs(geo.xc, geo.yc, geo.zc, geo.vol, 6, seed=seed)
    a, su = O.laplacian(m, geo, csr, np.full(m.n_cells, -1.0), np.zeros(m.n_cells + m.n_boundary), su0)
    return csr, a, su


def _noalloc(ctx, dev, csr, a, su, fi0, sor, nsw):
    n, nnz = csr.diag.shape[0], csr.ja.shape[0]
    p = [dev.put(csr.ioffset.astype(np.int32)), dev.put(csr.ja.astype(np.int32)),
         dev.put(csr.diag.astype(np.int32)), dev.put(a), dev.put(su), dev.put(fi0), dev.put(np.zeros(n))]
    info = fcap.SolveInfo()
    rc = ctx.L.fc_dpcg_noalloc(ctx.h, n, nnz, *p, float(sor), int(nsw), C.byref(info))
    return rc, dev.get(p[5], fi0), dev.get(p[6], fi0), info.as_dict()


def test_noalloc_dpcg_matches_dpcg_on_two_calls():
    ctx = fcap.Context(0)
    dev = _Dev()
    try:
        for nside, seed in ((20, 3), (11, 4)):     # the second call is smaller
            csr, a, su = _system(nside, seed)
            n = csr.diag.shape[0]
            print(f"noalloc dpcg: {n} cells", flush=True)
            fi0 = np.zeros(n)
            ref = fcap.dpcg(fcap.CSR(csr.ioffset, csr.ja, csr.diag, csr.ij_slot, csr.ji_slot), a, su, fi0,
                            sor=1e-8, nsw=3000, ctx=ctx)
            rc, fi, res, info = _noalloc(ctx, dev, csr, a, su, fi0, 1e-8, 3000)
            assert rc == ref.status == fcap.FC_OK, (rc, ref.status)
            assert info["iters"] == ref.info["iters"]
            np.testing.assert_array_equal(fi, ref.fi)
            np.testing.assert_array_equal(res, ref.res)
            assert info["res0"] == ref.info["res0"] and info["resl"] == ref.info["resl"]
            assert abs(info["rel_l2"] - ref.info["rel_l2"]) <= 1e-6 * ref.info["rel_l2"]
    finally:
        dev.close()
        ctx.close()


def test_noalloc_dpcg_refuses_oversize_and_host_arrays():
    ctx = fcap.Context(0)
    dev = _Dev()
    try:
        csr, a, su = _system(4, 1)
        n, nnz = csr.diag.shape[0], csr.ja.shape[0]
        p = [dev.put(csr.ioffset.astype(np.int32)), dev.put(csr.ja.astype(np.int32)),
             dev.put(csr.diag.astype(np.int32)), dev.put(a), dev.put(su), dev.put(np.zeros(n)), None]
        info = fcap.SolveInfo()
        # more cells than the scratch holds: refused before any array is read
        rc = ctx.L.fc_dpcg_noalloc(ctx.h, MAX_CELLS + 1, nnz, *p, 1e-8, 10, C.byref(info))
        assert rc == fcap.FC_ERR_ARGUMENT
        assert "1048576" in ctx.L.fc_last_error(ctx.h).decode()
        # a host array
        host_su = np.ascontiguousarray(su)
        q = list(p)
        q[4] = host_su.ctypes.data
        rc = ctx.L.fc_dpcg_noalloc(ctx.h, n, nnz, *q, 1e-8, 10, C.byref(info))
        assert rc == fcap.FC_ERR_ARGUMENT
        # the refused calls left nothing behind: a valid call still matches fc_dpcg
        ref = fcap.dpcg(fcap.CSR(csr.ioffset, csr.ja, csr.diag, csr.ij_slot, csr.ji_slot), a, su, np.zeros(n),
                        sor=1e-8, nsw=500, ctx=ctx)
        rc, fi, _, info2 = _noalloc(ctx, dev, csr, a, su, np.zeros(n), 1e-8, 500)
        assert rc == fcap.FC_OK and info2["iters"] == ref.info["iters"]
        np.testing.assert_array_equal(fi, ref.fi)
    finally:
        dev.close()
        ctx.close()
