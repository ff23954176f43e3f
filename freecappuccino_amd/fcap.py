"""ctypes front end of libfcap.so and the host-side mirror of the reference's solver interface.

The library is the product: every computation below runs in its HIP kernels
on the GPU.  If libfcap.so is missing or no GPU is present, the functions
raise -- there is no CPU fallback.

Mirrors of the reference entry points (explicit-argument forms, 1-based index
arrays, as src/ and tests/test_sparse_solvers.f90 use them):

    iccg(...)      src/iccg.f90:3        DIC-preconditioned CG
    dpcg(...)      src/dpcg.f90:3        Jacobi-preconditioned CG
    bicgstab(...)  src/bicgstab.f90:1    DILU-preconditioned BiCGStab
    laplacian(...) src/fvm_laplacian.f90:1
    create_csr_matrix_from_mesh_data(...) src/sparse_matrix.f90:42
    mesh_geometry(...) src/mesh_geometry_and_topology.f90:310 (geometry part)
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfcap.so")

FC_OK, FC_NOT_CONVERGED, FC_EARLY_EXIT, FC_FALLBACK_USED = 0, 1, 2, 3
FC_ERR_ARGUMENT, FC_ERR_HIP, FC_ERR_NONFINITE, FC_ERR_ZERO_PIVOT = -1, -2, -3, -4
FC_ERR_STRUCTURE, FC_ERR_ALLOC, FC_ERR_NO_DEVICE = -5, -6, -7
STATUS_NAMES = {0: "FC_OK", 1: "FC_NOT_CONVERGED", 2: "FC_EARLY_EXIT", 3: "FC_FALLBACK_USED",
                -1: "FC_ERR_ARGUMENT", -2: "FC_ERR_HIP", -3: "FC_ERR_NONFINITE", -4: "FC_ERR_ZERO_PIVOT",
                -5: "FC_ERR_STRUCTURE", -6: "FC_ERR_ALLOC", -7: "FC_ERR_NO_DEVICE"}
VARIANT_SRC, VARIANT_TESTS = 0, 1
METHOD_ICCG, METHOD_DPCG, METHOD_BICGSTAB = 1, 2, 3
METHOD_ICCG_MIXED, METHOD_BICGSTAB_MIXED, METHOD_FCG_LEARNED, METHOD_FCG_JACOBI = 11, 13, 21, 22
STOP_L1, STOP_L2 = 0, 1
SWEEP_LEVELS, SWEEP_DEPENDENT, SWEEP_PARTS = 0, 2, 3
REDUCE_TREE, REDUCE_SEQUENTIAL = 0, 1
NET_CLUSTERS, NET_CELL_ORDER = 0, 1
SWEEP_DEFAULT = SWEEP_PARTS
ARR = dict(a=1, su=2, ioffset=3, ja=4, diag=5, ij_slot=6, ji_slot=7, vol=8, xc=9, yc=10, zc=11, facint=12,
           arx=13, ary=14, arz=15, xf=16, yf=17, zf=18, res=19, pivots=20)


class FcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SolveInfo(C.Structure):
    _fields_ = [("res0", C.c_double), ("resl", C.c_double), ("rsm", C.c_double), ("rel_l2", C.c_double),
                ("iters", C.c_int), ("outer", C.c_int), ("status", C.c_int), ("fallback_iters", C.c_int),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def library() -> C.CDLL:
    """Load libfcap.so (built in-tree).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FcError(-7, f"{LIB_PATH} not built: run __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, ip, dp = C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    si = C.POINTER(SolveInfo)
    sig = {
        "fc_init": ([C.POINTER(vp), i], i), "fc_finalize": ([vp], i), "fc_set_stream": ([vp, vp], i),
        "fc_last_error": ([vp], C.c_char_p), "fc_set_report": ([vp, i, C.c_char_p], i),
        "fc_set_variant": ([vp, i], i),
        "fc_set_inner_tolerance": ([vp, d], i), "fc_set_fcg_window": ([vp, i], i),
        "fc_set_sweep_mode": ([vp, i], i), "fc_set_fcg_smoothing": ([vp, i, d], i),
        "fc_set_fcg_stagnation": ([vp, i, d], i),
        "fc_set_sweep_wait_limit": ([vp, i], i), "fc_set_graphs": ([vp, i], i),
        "fc_set_sweep_waves": ([vp, i], i),
        "fc_set_reduction_order": ([vp, i], i), "fc_set_network_layout": ([vp, i], i),
        "fc_sys_network_layout": ([vp, C.POINTER(i), C.POINTER(i), C.POINTER(C.c_longlong)], i),
        "fc_set_fcg_coarse": ([vp, i], i), "fc_set_fcg_network_period": ([vp, i], i),
        "fc_sys_coarse_space": ([vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), vp, vp], i),
        "fc_sys_precondition": ([vp, vp, vp], i),
        "fc_dpcg_noalloc": ([vp, i, i, vp, vp, vp, vp, vp, vp, vp, d, i, si], i),
        "fc_mesh_geometry": ([vp, i, dp, i, i, ip, ip, ip, ip, i] + [dp] * 11, i),
        "fc_csr_build": ([vp, i, i, ip, ip, ip, ip, ip, ip, ip], i),
        "fc_laplacian": ([vp, i, i, ip, ip] + [dp] * 12 + [i, i, i, i, i, i, ip, ip, dp, ip, ip, ip, ip, ip, i,
                                                          dp, dp, dp, dp], i),
        "fc_convdiff": ([vp, i, i, ip, ip] + [dp] * 12 + [i, i, ip, ip, ip, ip, ip, i, d, dp], i),
        "fc_face_flux_central": ([vp, i, ip, ip, dp, dp, dp, dp, dp, dp, dp, i, dp], i),
        "fc_grad_gauss": ([vp, i, i, i, ip, ip] + [dp] * 12 + [i, ip, i, ip, ip, ip, dp, i, dp, dp, dp], i),
        "fc_calcp_assemble": ([vp, i, i, ip, ip] + [dp] * 23 + [i, i, ip, ip, dp, ip, ip, ip, ip, ip, i]
                              + [dp] * 6, i),
        "fc_calcp_nonorth": ([vp, i, i, ip, ip] + [dp] * 14 + [i, i, ip, ip, dp, ip, ip, ip, ip, ip, i]
                             + [dp] * 3, i),
        "fc_adjust_mass_flow": ([vp, i, i, i, ip, dp, dp, dp, i, i, dp, i, i, i, d] + [dp] * 6, i),
        "fc_piso_assemble": ([vp, i, i, ip, ip] + [dp] * 19 + [i, i, ip, ip, dp, ip, ip, ip, ip, ip, i]
                             + [dp] * 6, i),
        "fc_bpres": ([vp, i, i, i, i, ip, dp, dp, dp, dp, dp, dp, ip, i, i, dp, i, dp], i),
        "fc_correct_velocity_bc": ([vp, i, i, i, i, ip, dp, dp, dp, ip, d, dp, dp, dp, dp, dp], i),
        "fc_sys_piso_assemble": ([vp] + [vp] * 9, i),
        "fc_sys_pin_reference": ([vp, i, vp], i),
        "fc_sys_bpres": ([vp, ip, i, vp, vp], i),
        "fc_sys_grad_pp": ([vp, ip, i, i, vp], i),
        "fc_sys_correct_velocity_bc": ([vp, ip, d, vp, vp, vp, vp, vp], i),
        "fc_calcp_correct": ([vp, i, i, ip, ip, ip, ip, ip, ip, ip, i] + [dp] * 7 + [d, i, i, ip, ip]
                             + [dp] * 7, i),
        "fc_rhs_divergence": ([vp, i, i, ip, ip, dp, dp], i),
        "fc_spmv": ([vp, i, i, ip, ip, dp, dp, dp], i),
        "fc_dic_pivots": ([vp, i, i, ip, ip, ip, dp, dp], i),
        "fc_dilu_pivots": ([vp, i, i, ip, ip, ip, dp, dp], i),
        "fc_lu_apply": ([vp, i, i, ip, ip, ip, dp, dp, dp, dp], i),
        "fc_sys_create": ([vp, C.POINTER(vp), i, dp, i, i, ip, ip, ip, ip, i, i, i, i], i),
        "fc_sys_destroy": ([vp], i),
        "fc_sys_assemble_laplacian": ([vp, dp, dp, dp], i),
        "fc_sys_assemble_laplacian_dev": ([vp], i),
        "fc_sys_assemble_convdiff": ([vp, dp, dp, d, dp], i),
        "fc_sys_calcp_assemble": ([vp] + [vp] * 13, i),
        "fc_sys_grad_gauss": ([vp, ip, vp, i, vp], i),
        "fc_sys_calcp_correct": ([vp, vp, vp, vp, vp, vp, d, i, vp, vp, vp, vp, vp], i),
        "fc_sys_calcp_nonorth": ([vp, vp, vp, vp, vp], i),
        "fc_sys_set_rhs": ([vp, dp], i), "fc_sys_set_solution": ([vp, dp], i),
        "fc_sys_solve_momentum": ([vp, C.POINTER(dp), C.POINTER(dp), dp, C.POINTER(dp), C.POINTER(dp), i, d, i,
                                   si], i),
        "fc_sys_get_solution": ([vp, dp], i),
        "fc_sys_solve": ([vp, i, i, d, i, si], i),
        "fc_sys_get": ([vp, i, vp, C.c_size_t], i),
        "fc_sys_device_ptr": ([vp, i], vp),
        "fc_sys_sizes": ([vp, ip, ip, ip, ip, ip], i),
        "fc_sys_last_ms": ([vp, i], d),
        "fc_sys_spmv_bench": ([vp, i, i], d),
        "fc_sys_sweep_bench": ([vp, i, i], d),
        "fc_copy_bench": ([vp, C.c_size_t, i], d),
        "fc_gnn_num_params": ([], i),
        "fc_gnn_create": ([vp, C.POINTER(vp), fp], i),
        "fc_gnn_destroy": ([vp], i),
        "fc_gnn_set_params": ([vp, fp], i),
        "fc_gnn_get_params": ([vp, fp], i),
        "fc_gnn_apply": ([vp, vp, fp, fp], i),
        "fc_gnn_bench": ([vp, vp, i], d),
        "fc_gnn_bench_layer": ([vp, vp, i], d),
        "fc_gnn_edge_features": ([vp, vp, fp, fp, fp], i),
        "fc_gnn_loss_grad": ([vp, vp, fp, fp, dp, fp], i),
        "fc_gnn_add_samples": ([vp, vp, i, ip, C.c_uint], i),
        "fc_gnn_num_samples": ([vp], i),
        "fc_gnn_get_sample": ([vp, i, fp, fp, i], i),
        "fc_gnn_clear_samples": ([vp], i),
        "fc_gnn_drop_system": ([vp, vp], i),
        "fc_gnn_train": ([vp, i, i, C.c_float, C.c_uint, fp], i),
        "fc_gnn_set_training_smoothing": ([vp, i, d], i), "fc_gnn_set_training_coarse": ([vp, i], i),
        "fc_sys_set_network": ([vp, vp], i),
    }
    for name in ("fc_iccg", "fc_dpcg", "fc_bicgstab"):
        sig[name] = ([vp, i, i, ip, ip, ip, dp, dp, dp, dp, i, ip, ip, dp, dp, d, i, si], i)
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """A device context (stream, workspace, structure cache) on one GPU."""

    def __init__(self, device: int = 0):
        self.L = library()
        h = C.c_void_p()
        rc = self.L.fc_init(C.byref(h), device)
        if rc != 0:
            raise FcError(rc, "fc_init failed (is a GPU visible?)")
        self.h = h
        # objects whose library handles live on this context (systems, networks):
        # they are closed before the context, whatever order Python collects them in
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if getattr(self, "h", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            self.L.fc_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, allow=(FC_OK,)):
        if rc not in allow:
            raise FcError(rc, self.L.fc_last_error(self.h).decode())
        return rc

    def set_variant(self, variant):
        self.check(self.L.fc_set_variant(self.h, variant))

    def set_inner_tolerance(self, eta: float):
        """Inner FP32 tolerance of the refinement solvers (0: the method's default)."""
        self.check(self.L.fc_set_inner_tolerance(self.h, float(eta)))

    def set_fcg_window(self, m: int = 0):
        """Directions kept by flexible CG (1..16, rounded up to a multiple of 4; 0: 4 with the
        coarse correction, 12 without)."""
        self.check(self.L.fc_set_fcg_window(self.h, int(m)))

    def set_fcg_smoothing(self, steps: int = 3, omega: float = 0.8):
        """Damped-Jacobi steps after the network in the learned FCG's preconditioner."""
        self.check(self.L.fc_set_fcg_smoothing(self.h, int(steps), float(omega)))

    def set_fcg_coarse(self, cells_per_aggregate: int = 4096):
        """Coarse correction of the learned FCG's preconditioner over aggregates of this many
        cells of the network's Morton order (a multiple of 256; 0 switches it off)."""
        self.check(self.L.fc_set_fcg_coarse(self.h, int(cells_per_aggregate)))

    def set_fcg_network_period(self, period: int = 0):
        """The learned FCG runs its network on the first of every `period` iterations (1..4; the
        others start from the Jacobi guess); 0: 2 with the coarse correction, 1 without."""
        self.check(self.L.fc_set_fcg_network_period(self.h, int(period)))

    def set_network_layout(self, layout: int = NET_CLUSTERS):
        """Layout of the network's inference layers: NET_CLUSTERS (compact 256-cell clusters
        staged in LDS, where every halo fits) or NET_CELL_ORDER; bitwise the same z."""
        self.check(self.L.fc_set_network_layout(self.h, int(layout)))

    def set_fcg_stagnation(self, window: int = 64, factor: float = 0.81):
        """Stagnation test of the learned FCG's inner solves (window 0: off; factor 0 forces the
        IC(0)-PCG fallback after the second window)."""
        self.check(self.L.fc_set_fcg_stagnation(self.h, int(window), float(factor)))

    def set_sweep_mode(self, mode: int):
        """Schedule of the triangular sweeps: SWEEP_LEVELS, SWEEP_DEPENDENT or SWEEP_PARTS
        (bit-identical results)."""
        self.check(self.L.fc_set_sweep_mode(self.h, int(mode)))

    def set_sweep_waves(self, waves: int = 0):
        """Persistent waves of the dependency-driven sweeps and pivots (0: automatic)."""
        self.check(self.L.fc_set_sweep_waves(self.h, int(waves)))

    def set_sweep_wait_limit(self, microseconds: int = 0):
        """Wall-clock limit of one wait of the dependency-driven sweeps (0: the built-in 2 s)."""
        self.check(self.L.fc_set_sweep_wait_limit(self.h, int(microseconds)))

    def set_reduction_order(self, order: int = 0):
        """REDUCE_TREE (default) or REDUCE_SEQUENTIAL: the FP64 solvers' sums in index order, which
        makes their iterates bit-identical to the reference's serial solvers (verification mode)."""
        self.check(self.L.fc_set_reduction_order(self.h, int(order)))

    def set_graphs(self, on: bool = True):
        """Replay solver iterations as HIP graphs (default) or launch kernels one by one."""
        self.check(self.L.fc_set_graphs(self.h, 1 if on else 0))

    def set_report(self, on: bool, name: str = "p"):
        self.check(self.L.fc_set_report(self.h, 1 if on else 0, name.encode()))

    def copy_bench(self, nbytes: int, reps: int = 10) -> float:
        """Average ms of a plain device copy of nbytes (HBM counter calibration)."""
        ms = self.L.fc_copy_bench(self.h, nbytes, reps)
        if ms < 0:
            raise FcError(-2, self.L.fc_last_error(self.h).decode())
        return ms

    def set_stream(self, stream_ptr: int):
        self.check(self.L.fc_set_stream(self.h, C.c_void_p(stream_ptr)))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---------------------------------------------------------------- geometry

@dataclass
class Geometry:
    arx: np.ndarray
    ary: np.ndarray
    arz: np.ndarray
    xf: np.ndarray
    yf: np.ndarray
    zf: np.ndarray
    vol: np.ndarray
    xc: np.ndarray
    yc: np.ndarray
    zc: np.ndarray
    facint: np.ndarray


def mesh_geometry(mesh, ctx: Optional[Context] = None) -> Geometry:
    """Geometry of a PolyMesh, computed on the GPU (src/mesh_geometry_and_topology.f90:859-1081)."""
    ctx = ctx or default_context()
    nf, nif, n = mesh.n_faces, mesh.n_inner, mesh.n_cells
    g = Geometry(*(np.zeros(nf) for _ in range(6)), *(np.zeros(n) for _ in range(4)), np.zeros(max(nif, 1)))
    xyz = _f64(mesh.points)
    fn1 = _i32(mesh.face_nodes.astype(np.int64) + 1)
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1) if nif else np.zeros(1, np.int32)
    ctx.check(ctx.L.fc_mesh_geometry(ctx.h, xyz.shape[0], _dp(xyz), nf, nif, _ip(_i32(mesh.face_ptr)), _ip(fn1),
                                     _ip(own1), _ip(nei1), n,
                                     *(_dp(getattr(g, k)) for k in ("arx", "ary", "arz", "xf", "yf", "zf", "vol",
                                                                    "xc", "yc", "zc", "facint"))))
    g.facint = g.facint[:nif]
    return g


@dataclass
class CSR:
    ioffset: np.ndarray
    ja: np.ndarray
    diag: np.ndarray
    ij_slot: np.ndarray
    ji_slot: np.ndarray


def create_csr_matrix_from_mesh_data(n_cells: int, owner1, neighbour1, ctx: Optional[Context] = None) -> CSR:
    """CSR structure and face slot maps on the GPU (src/sparse_matrix.f90:42-172); 1-based in and out."""
    ctx = ctx or default_context()
    owner1, neighbour1 = _i32(owner1), _i32(neighbour1)
    nif = neighbour1.shape[0]
    nnz = 2 * nif + n_cells
    c = CSR(np.zeros(n_cells + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(n_cells, np.int32),
            np.zeros(max(nif, 1), np.int32), np.zeros(max(nif, 1), np.int32))
    ctx.check(ctx.L.fc_csr_build(ctx.h, n_cells, nif, _ip(owner1), _ip(neighbour1), _ip(c.ioffset), _ip(c.ja),
                                 _ip(c.diag), _ip(c.ij_slot), _ip(c.ji_slot)))
    c.ij_slot, c.ji_slot = c.ij_slot[:nif], c.ji_slot[:nif]
    return c


def wall_info(mesh):
    """(nwal, iWallFacesStart, iWallStart) as src/mesh_geometry_and_topology.f90:409-605 derives them."""
    ws, wc = mesh.wall_range()
    cnt = {"inlet": 0, "outlet": 0, "symmetry": 0}
    for t, nf, _ in mesh.patches:
        if t in cnt:
            cnt[t] += nf
    return wc, ws, mesh.n_cells + cnt["inlet"] + cnt["outlet"] + cnt["symmetry"]


def laplacian(mesh, geo: Geometry, csr: CSR, mu, phi, su0, ctx: Optional[Context] = None):
    """laplacian(mu, phi) on the GPU (src/fvm_laplacian.f90:1-163); returns (a, su)."""
    ctx = ctx or default_context()
    n, nif = mesh.n_cells, mesh.n_inner
    nwal, wstart, iwall = wall_info(mesh)
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    a = np.zeros(mesh.nnz)
    su = _f64(su0).copy()
    phi = _f64(phi)
    zi, zd = np.zeros(1, np.int32), np.zeros(1)
    ctx.check(ctx.L.fc_laplacian(ctx.h, n, nif, _ip(own1), _ip(nei1), _dp(geo.arx), _dp(geo.ary), _dp(geo.arz),
                                 _dp(_f64(geo.facint)), _dp(geo.xc), _dp(geo.yc), _dp(geo.zc), _dp(geo.xf),
                                 _dp(geo.yf), _dp(geo.zf), _dp(_f64(mu)), _dp(phi), phi.shape[0], nwal, wstart,
                                 iwall, 0, 0, _ip(zi), _ip(zi), _dp(zd), _ip(csr.ioffset), _ip(csr.ja),
                                 _ip(csr.diag), _ip(_i32(csr.ij_slot)), _ip(_i32(csr.ji_slot)), mesh.nnz, _dp(a),
                                 _dp(su), _dp(zd), _dp(zd)))
    return a, su


def convdiff(mesh, geo: Geometry, csr: CSR, vis, flux, urf=0.7, ctx: Optional[Context] = None):
    """Momentum-like upwind convection-diffusion matrix on the GPU."""
    ctx = ctx or default_context()
    n, nif = mesh.n_cells, mesh.n_inner
    nwal, wstart, _ = wall_info(mesh)
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    a = np.zeros(mesh.nnz)
    ctx.check(ctx.L.fc_convdiff(ctx.h, n, nif, _ip(own1), _ip(nei1), _dp(geo.arx), _dp(geo.ary), _dp(geo.arz),
                                _dp(_f64(geo.facint)), _dp(geo.xc), _dp(geo.yc), _dp(geo.zc), _dp(geo.xf),
                                _dp(geo.yf), _dp(geo.zf), _dp(_f64(vis)), _dp(_f64(flux)), nwal, wstart,
                                _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag), _ip(_i32(csr.ij_slot)),
                                _ip(_i32(csr.ji_slot)), mesh.nnz, float(urf), _dp(a)))
    return a


def boundary_groups(mesh):
    """(count, faces before the group) of the inlet, outlet, symmetry, wall and pressure-outlet
    groups, as src/mesh_geometry_and_topology.f90:409-470 derives them from the boundary file."""
    groups = ("inlet", "outlet", "symmetry", "wall", "prOutlet")
    walls = ("wall", "wallIsoth", "wallAdiab", "wallQFlux")
    bnd = np.zeros(10, np.int32)
    for t, nf, st in mesh.patches:
        g = 3 if t in walls else groups.index(t) if t in groups else -1
        if g < 0:
            continue
        if bnd[2 * g] == 0:
            bnd[2 * g + 1] = st
        bnd[2 * g] += nf
    return bnd


def grad_gauss(mesh, geo: Geometry, u, nigrad=1, oc=None, ctx: Optional[Context] = None):
    """Gauss gradient on the GPU (src/grad_gauss.f90) of u (cells, then boundary values in group
    order); returns (dudx, dudy, dudz).  oc: optional O-C cuts {ijl, ijr (0-based cells), foc,
    faces (0-based face of the first cut, the others following)}."""
    ctx = ctx or default_context()
    n, nif, nf = mesh.n_cells, mesh.n_inner, mesh.n_faces
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    u = _f64(u)
    out = [np.zeros(n) for _ in range(3)]
    zi, zd = np.zeros(1, np.int32), np.zeros(1)
    noc = 0 if oc is None else len(oc["ijl"])
    if noc:
        ijlf = _i32(int(oc["faces"]) + np.arange(noc) + 1)
        ijl, ijr, foc = _i32(np.asarray(oc["ijl"]) + 1), _i32(np.asarray(oc["ijr"]) + 1), _f64(oc["foc"])
    else:
        ijlf, ijl, ijr, foc = zi, zi, zi, zd
    g = [_f64(getattr(geo, k)) for k in ("arx", "ary", "arz", "facint", "xc", "yc", "zc", "xf", "yf", "zf", "vol")]
    ctx.check(ctx.L.fc_grad_gauss(ctx.h, n, nf, nif, _ip(own1), _ip(nei1), *[_dp(x) for x in g], _dp(u), u.shape[0],
                                  _ip(boundary_groups(mesh)), noc, _ip(ijlf), _ip(ijl), _ip(ijr), _dp(foc),
                                  int(nigrad),
                                  *[_dp(x) for x in out]))
    return tuple(out)


CALCP_FIELDS = ("den", "apu", "apv", "apw", "u", "v", "w", "p")
CALCP_GRADS = ("dudx", "dvdx", "dwdx", "dpdx")


def calcp_assemble(mesh, geo: Geometry, csr: CSR, fields, oc=None, ctx: Optional[Context] = None):
    """Pressure-correction assembly on the GPU: calcp's face loops around facefluxmass
    (src/calcp-multiple_correction_SIMPLE.f90:34-107, src/facefluxmass.f90:34-199).
    fields: den, apu, apv, apw, u, v, w, p (n each) and dudx, dvdx, dwdx, dpdx (n x 3).
    oc: optional O-C cuts {ijl, ijr (0-based cells), foc, faces (0-based first cut face)}.
    Returns a dict with a, su, flmass (and al, ar, fmoc with cuts)."""
    ctx = ctx or default_context()
    n, nif = mesh.n_cells, mesh.n_inner
    own1 = _i32(mesh.owner[:nif].astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    fl = [_f64(fields[k]) for k in CALCP_FIELDS]
    gr = [_f64(np.asarray(fields[k]).reshape(n, 3)) for k in CALCP_GRADS]
    noc = 0 if oc is None else len(oc["ijl"])
    ijl = _i32(np.asarray(oc["ijl"]) + 1) if noc else np.zeros(1, np.int32)
    ijr = _i32(np.asarray(oc["ijr"]) + 1) if noc else np.zeros(1, np.int32)
    foc = _f64(oc["foc"]) if noc else np.zeros(1)
    oc0 = int(oc["faces"]) if noc else 0
    out = {"a": np.zeros(mesh.nnz), "su": np.zeros(n), "flmass": np.zeros(max(nif, 1)),
           "al": np.zeros(max(noc, 1)), "ar": np.zeros(max(noc, 1)), "fmoc": np.zeros(max(noc, 1))}
    g = [_f64(getattr(geo, k)) for k in ("xc", "yc", "zc", "vol", "xf", "yf", "zf", "arx", "ary", "arz", "facint")]
    ctx.check(ctx.L.fc_calcp_assemble(ctx.h, n, nif, _ip(own1), _ip(nei1), *[_dp(x) for x in g],
                                      *[_dp(x) for x in fl], *[_dp(x) for x in gr], noc, oc0, _ip(ijl), _ip(ijr),
                                      _dp(foc), _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag),
                                      _ip(_i32(csr.ij_slot)), _ip(_i32(csr.ji_slot)), mesh.nnz,
                                      *[_dp(out[k]) for k in ("a", "su", "flmass", "al", "ar", "fmoc")]))
    out["flmass"] = out["flmass"][:nif]
    for k in ("al", "ar", "fmoc"):
        out[k] = out[k][:noc]
    return out


def _oc_host(oc):
    noc = 0 if oc is None else len(oc["ijl"])
    ijl = _i32(np.asarray(oc["ijl"]) + 1) if noc else np.zeros(1, np.int32)
    ijr = _i32(np.asarray(oc["ijr"]) + 1) if noc else np.zeros(1, np.int32)
    foc = _f64(oc["foc"]) if noc else np.zeros(1)
    return noc, (int(oc["faces"]) if noc else 0), ijl, ijr, foc


PISO_FIELDS = ("den", "apu", "u", "v", "w")
PISO_GRADS = ("dudx", "dvdx", "dwdx")


def piso_assemble(mesh, geo: Geometry, csr: CSR, fields, oc=None, ctx: Optional[Context] = None):
    """PISO's pressure assembly on the GPU: the face loop of src/PISO_multiple_correction.f90:
    107-148 around facefluxmass_piso (src/facefluxmass.f90:294-358).  fields: den, apu, u, v, w
    (n each) and dudx, dvdx, dwdx (n x 3).  Returns a dict with a, su, flmass (and al, ar, fmoc
    with cuts)."""
    ctx = ctx or default_context()
    n, nif = mesh.n_cells, mesh.n_inner
    own1 = _i32(mesh.owner[:nif].astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    fl = [_f64(fields[k]) for k in PISO_FIELDS]
    gr = [_f64(np.asarray(fields[k]).reshape(n, 3)) for k in PISO_GRADS]
    noc, oc0, ijl, ijr, foc = _oc_host(oc)
    out = {"a": np.zeros(mesh.nnz), "su": np.zeros(n), "flmass": np.zeros(max(nif, 1)),
           "al": np.zeros(max(noc, 1)), "ar": np.zeros(max(noc, 1)), "fmoc": np.zeros(max(noc, 1))}
    g = [_f64(getattr(geo, k)) for k in ("xc", "yc", "zc", "vol", "xf", "yf", "zf", "arx", "ary", "arz", "facint")]
    ctx.check(ctx.L.fc_piso_assemble(ctx.h, n, nif, _ip(own1), _ip(nei1), *[_dp(x) for x in g],
                                     *[_dp(x) for x in fl], *[_dp(x) for x in gr], noc, oc0, _ip(ijl), _ip(ijr),
                                     _dp(foc), _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag),
                                     _ip(_i32(csr.ij_slot)), _ip(_i32(csr.ji_slot)), mesh.nnz,
                                     *[_dp(out[k]) for k in ("a", "su", "flmass", "al", "ar", "fmoc")]))
    out["flmass"] = out["flmass"][:nif]
    for k in ("al", "ar", "fmoc"):
        out[k] = out[k][:noc]
    return out


def bpres(mesh, geo: Geometry, p, dpdx, istage, noc=0, oc_faces_start=0, ctx: Optional[Context] = None):
    """bpres(p, istage) on the GPU (src/bpres.f90:3-152): returns p (numTotal) with its boundary
    values set; dpdx (n x 3) is read from istage 2 on."""
    ctx = ctx or default_context()
    n = mesh.n_cells
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    out = _f64(p).copy()
    dg = _f64(np.asarray(dpdx).reshape(n, 3)) if dpdx is not None else np.zeros(1)
    g = [_f64(getattr(geo, k)) for k in ("xc", "yc", "zc", "xf", "yf", "zf")]
    ctx.check(ctx.L.fc_bpres(ctx.h, n, out.shape[0], mesh.n_faces, mesh.n_inner, _ip(own1), *[_dp(x) for x in g],
                             _ip(boundary_groups(mesh)), int(noc), int(oc_faces_start), _dp(dg), int(istage),
                             _dp(out)))
    return out


def correct_velocity_bc(mesh, geo: Geometry, flomas, den, u, v, w, ctx: Optional[Context] = None):
    """correctBoundaryConditionsVelocity on the GPU (src/correctBoundaryConditionsVelocity.f90:1-79):
    returns (u, v, w, fmo); u, v, w hold the cells then the boundary values in group order."""
    ctx = ctx or default_context()
    n = mesh.n_cells
    bnd = boundary_groups(mesh)
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    vel = [_f64(x).copy() for x in (u, v, w)]
    nout = int(bnd[2])
    fmo = np.zeros(max(nout, 1))
    ctx.check(ctx.L.fc_correct_velocity_bc(ctx.h, n, vel[0].shape[0], mesh.n_faces, mesh.n_inner, _ip(own1),
                                           _dp(_f64(geo.arx)), _dp(_f64(geo.ary)), _dp(_f64(geo.arz)), _ip(bnd),
                                           float(flomas), _dp(_f64(den)), *[_dp(x) for x in vel], _dp(fmo)))
    return (*vel, fmo[:nout])


def calcp_nonorth(mesh, geo: Geometry, csr: CSR, den, apu, dpdx, flmass, oc=None, fmoc=None,
                  ctx: Optional[Context] = None):
    """Non-orthogonal corrector source of calcp on the GPU
    (src/calcp-multiple_correction_SIMPLE.f90:185-229, fluxmc src/facefluxmass.f90:520-607).
    dpdx: gradient of the pressure correction (n x 3).  Returns (su, flmass + fmcor,
    fmoc + fmcor); su starts from zero."""
    ctx = ctx or default_context()
    n, nif = mesh.n_cells, mesh.n_inner
    own1 = _i32(mesh.owner[:nif].astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    noc, oc0, ijl, ijr, foc = _oc_host(oc)
    su = np.zeros(n)
    fl = np.zeros(max(nif, 1))
    fl[:nif] = _f64(flmass)[:nif]
    fm = _f64(fmoc).copy() if noc else np.zeros(1)
    g = [_f64(getattr(geo, k)) for k in ("xc", "yc", "zc", "vol", "xf", "yf", "zf", "arx", "ary", "arz", "facint")]
    ctx.check(ctx.L.fc_calcp_nonorth(ctx.h, n, nif, _ip(own1), _ip(nei1), *[_dp(x) for x in g], _dp(_f64(den)),
                                     _dp(_f64(apu)), _dp(_f64(np.asarray(dpdx).reshape(n, 3))), noc, oc0, _ip(ijl),
                                     _ip(ijr), _dp(foc), _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag),
                                     _ip(_i32(csr.ij_slot)), _ip(_i32(csr.ji_slot)), mesh.nnz, _dp(su), _dp(fl),
                                     _dp(fm)))
    return su, fl[:nif], fm[:noc]


def adjust_mass_flow(mesh, geo: Geometry, flomas, den, u, v, w, su, fmi, ctx: Optional[Context] = None):
    """adjustMassFlow on the GPU (src/adjustMassFlow.f90:1-71): inlet fluxes fmi into su, outlet
    fluxes from the owner cells scaled so that they carry flomas.  u, v, w hold the cells then the
    boundary values in the reference's group order.  Returns (su, u, v, w, fmo)."""
    ctx = ctx or default_context()
    n = mesh.n_cells
    bnd = boundary_groups(mesh)
    ninl, inl0, nout, out0 = (int(x) for x in bnd[:4])
    own1 = _i32(mesh.owner.astype(np.int64) + 1)
    vel = [_f64(x).copy() for x in (u, v, w)]
    ntot = vel[0].shape[0]
    s = _f64(su).copy()
    fmo = np.zeros(max(nout, 1))
    fmi_ = _f64(fmi) if ninl else np.zeros(1)
    ctx.check(ctx.L.fc_adjust_mass_flow(ctx.h, n, ntot, mesh.n_faces, _ip(own1), _dp(_f64(geo.arx)),
                                        _dp(_f64(geo.ary)), _dp(_f64(geo.arz)), ninl, inl0, _dp(fmi_), nout, out0,
                                        n + ninl, float(flomas), _dp(_f64(den)), *[_dp(x) for x in vel], _dp(s),
                                        _dp(fmo)))
    return (s, *vel, fmo[:nout])


def calcp_correct(mesh, csr: CSR, a, pp, dpdx, vol, apu, apv, apw, urf, pref_cell, flmass, u, v, w, p,
                  oc=None, ar=None, fmoc=None, ctx: Optional[Context] = None):
    """Corrections of calcp on the GPU (src/calcp-multiple_correction_SIMPLE.f90:145-183): face mass
    fluxes by a(ij) (pp_N - pp_P), O-C fluxes by ar (pp_r - pp_l), u, v, w by -grad(pp) Vol/a_P and
    p by urf (pp - pp[pref_cell]); pref_cell is 0-based.  Returns (flmass, fmoc, u, v, w, p)."""
    ctx = ctx or default_context()
    n, nif = mesh.n_cells, mesh.n_inner
    own1 = _i32(mesh.owner[:nif].astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    noc, _, ijl, ijr, _ = _oc_host(oc)
    ar_ = _f64(ar) if noc else np.zeros(1)
    fm = _f64(fmoc).copy() if noc else np.zeros(1)
    fl = np.zeros(max(nif, 1))
    fl[:nif] = _f64(flmass)[:nif]
    cells = [_f64(x).copy() for x in (u, v, w, p)]
    ctx.check(ctx.L.fc_calcp_correct(ctx.h, n, nif, _ip(own1), _ip(nei1), _ip(csr.ioffset), _ip(csr.ja),
                                     _ip(csr.diag), _ip(_i32(csr.ij_slot)), _ip(_i32(csr.ji_slot)), mesh.nnz,
                                     _dp(_f64(a)), _dp(_f64(pp)), _dp(_f64(np.asarray(dpdx).reshape(n, 3))),
                                     _dp(_f64(vol)), _dp(_f64(apu)), _dp(_f64(apv)), _dp(_f64(apw)), float(urf),
                                     int(pref_cell) + 1, noc, _ip(ijl), _ip(ijr), _dp(ar_), _dp(fm), _dp(fl),
                                     *[_dp(x) for x in cells]))
    return (fl[:nif], fm[:noc], *cells)


def face_flux_central(mesh, geo: Geometry, u, v, w, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    nif = mesh.n_inner
    own1 = _i32(mesh.owner[:nif].astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    flux = np.zeros(max(nif, 1))
    ctx.check(ctx.L.fc_face_flux_central(ctx.h, nif, _ip(own1), _ip(nei1), _dp(_f64(geo.facint)),
                                         _dp(_f64(geo.arx[:nif])), _dp(_f64(geo.ary[:nif])),
                                         _dp(_f64(geo.arz[:nif])), _dp(_f64(u)), _dp(_f64(v)), _dp(_f64(w)),
                                         mesh.n_cells, _dp(flux)))
    return flux[:nif]


def rhs_divergence(mesh, flux, su0, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    nif = mesh.n_inner
    own1 = _i32(mesh.owner[:nif].astype(np.int64) + 1)
    nei1 = _i32(mesh.neighbour.astype(np.int64) + 1)
    su = _f64(su0).copy()
    ctx.check(ctx.L.fc_rhs_divergence(ctx.h, mesh.n_cells, nif, _ip(own1), _ip(nei1), _dp(_f64(flux)), _dp(su)))
    return su


# ----------------------------------------------------------------- kernels

def spmv(csr: CSR, a, x, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    n = csr.diag.shape[0]
    y = np.zeros(n)
    ctx.check(ctx.L.fc_spmv(ctx.h, n, csr.ja.shape[0], _ip(csr.ioffset), _ip(csr.ja), _dp(_f64(a)), _dp(_f64(x)),
                            _dp(y)))
    return y


def dic_pivots(csr: CSR, a, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    n = csr.diag.shape[0]
    d = np.zeros(n)
    ctx.check(ctx.L.fc_dic_pivots(ctx.h, n, csr.ja.shape[0], _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag),
                                  _dp(_f64(a)), _dp(d)))
    return d


def dilu_pivots(csr: CSR, a, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    n = csr.diag.shape[0]
    d = np.zeros(n)
    ctx.check(ctx.L.fc_dilu_pivots(ctx.h, n, csr.ja.shape[0], _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag),
                                   _dp(_f64(a)), _dp(d)))
    return d


def lu_apply(csr: CSR, a, d, r, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    n = csr.diag.shape[0]
    z = np.zeros(n)
    ctx.check(ctx.L.fc_lu_apply(ctx.h, n, csr.ja.shape[0], _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag),
                                _dp(_f64(a)), _dp(_f64(d)), _dp(_f64(r)), _dp(z)))
    return z


# ----------------------------------------------------------------- solvers

@dataclass
class SolveResult:
    status: int
    fi: np.ndarray
    res: np.ndarray
    info: dict


def _solve(name, csr: CSR, a, su, fi0, sor, nsw, ctx, noc_arrays, variant):
    ctx = ctx or default_context()
    ctx.set_variant(variant)
    n = csr.diag.shape[0]
    fi = _f64(fi0).copy()
    res = np.zeros(n)
    info = SolveInfo()
    if noc_arrays is None:
        noc, ijl, ijr, al, ar = 0, np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1), np.zeros(1)
    else:
        ijl, ijr, al, ar = (_i32(noc_arrays[0]), _i32(noc_arrays[1]), _f64(noc_arrays[2]), _f64(noc_arrays[3]))
        noc = ijl.shape[0]
    fn = getattr(ctx.L, name)
    rc = fn(ctx.h, n, csr.ja.shape[0], _ip(csr.ioffset), _ip(csr.ja), _ip(csr.diag), _dp(_f64(a)), _dp(_f64(su)),
            _dp(fi), _dp(res), noc, _ip(ijl), _ip(ijr), _dp(al), _dp(ar), float(sor), int(nsw), C.byref(info))
    ctx.set_variant(VARIANT_SRC)
    ctx.check(rc, allow=(FC_OK, FC_NOT_CONVERGED, FC_EARLY_EXIT))
    return SolveResult(rc, fi, res, info.as_dict())


def iccg(csr, a, su, fi0, sor=1e-8, nsw=1000, ctx=None, noc_arrays=None, variant=VARIANT_SRC):
    """iccg(fi, ifi): DIC-preconditioned CG on the GPU (src/iccg.f90)."""
    return _solve("fc_iccg", csr, a, su, fi0, sor, nsw, ctx, noc_arrays, variant)


def dpcg(csr, a, su, fi0, sor=1e-8, nsw=1000, ctx=None, noc_arrays=None, variant=VARIANT_SRC):
    """dpcg(fi, ifi): Jacobi-preconditioned CG on the GPU (src/dpcg.f90)."""
    return _solve("fc_dpcg", csr, a, su, fi0, sor, nsw, ctx, noc_arrays, variant)


def bicgstab(csr, a, su, fi0, sor=1e-8, nsw=1000, ctx=None, noc_arrays=None, variant=VARIANT_SRC):
    """bicgstab(fi, ifi): DILU-preconditioned BiCGStab on the GPU (src/bicgstab.f90)."""
    return _solve("fc_bicgstab", csr, a, su, fi0, sor, nsw, ctx, noc_arrays, variant)


# ---------------------------------------------------------- resident system

class System:
    """A mesh, its CSR structure and an assembled system resident in HBM."""

    def __init__(self, mesh, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.mesh = mesh
        L = self.ctx.L
        nwal, wstart, iwall = wall_info(mesh)
        xyz = _f64(mesh.points)
        fn1 = _i32(mesh.face_nodes.astype(np.int64) + 1)
        own1 = _i32(mesh.owner.astype(np.int64) + 1)
        nei1 = _i32(mesh.neighbour.astype(np.int64) + 1) if mesh.n_inner else np.zeros(1, np.int32)
        h = C.c_void_p()
        self.ctx.check(L.fc_sys_create(self.ctx.h, C.byref(h), xyz.shape[0], _dp(xyz), mesh.n_faces, mesh.n_inner,
                                       _ip(_i32(mesh.face_ptr)), _ip(fn1), _ip(own1), _ip(nei1), mesh.n_cells,
                                       nwal, wstart, iwall))
        self.h = h
        self.n = mesh.n_cells
        self.ntotal = mesh.n_cells + mesh.n_boundary
        self._sample_holders = weakref.WeakSet()   # networks holding training samples of this system
        self.ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None):
            for net in list(getattr(self, "_sample_holders", ())):
                net.drop_system(self)
            if getattr(self.ctx, "h", None):
                self.ctx.L.fc_sys_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble_laplacian(self, mu, phi=None, su0=None):
        phi = np.zeros(self.ntotal) if phi is None else _f64(phi)
        self.ctx.check(self.ctx.L.fc_sys_assemble_laplacian(self.h, _dp(_f64(mu)), _dp(phi),
                                                            _dp(None if su0 is None else _f64(su0))))

    def assemble_laplacian_dev(self):
        self.ctx.check(self.ctx.L.fc_sys_assemble_laplacian_dev(self.h))

    def assemble_convdiff(self, vis, flux, urf=0.7, su0=None):
        self.ctx.check(self.ctx.L.fc_sys_assemble_convdiff(self.h, _dp(_f64(vis)), _dp(_f64(flux)), float(urf),
                                                           _dp(None if su0 is None else _f64(su0))))

    def calcp_assemble(self, den, apu, apv, apw, u, v, w, p, dudx, dvdx, dwdx, dpdx, flmass=None):
        """calcp's pressure-correction assembly on the system's own mesh
        (src/calcp-multiple_correction_SIMPLE.f90:34-107): a and su go into the system.
        Each argument is a numpy array (copied) or an int device address (used in
        place); gradients are (numCells, 3) arrays.  Returns flmass (numInnerFaces)
        when it was not given as a device address."""
        keep = []

        def arg(x):
            if isinstance(x, int):
                return C.c_void_p(x)
            arr = np.ascontiguousarray(x, dtype=np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.c_void_p)

        out = None
        if flmass is None:
            out = np.zeros(max(self.mesh.n_inner, 1))
            flmass = out
        args = [arg(x) for x in (den, apu, apv, apw, u, v, w, p, dudx, dvdx, dwdx, dpdx, flmass)]
        self.ctx.check(self.ctx.L.fc_sys_calcp_assemble(self.h, *args))
        return None if out is None else out[:self.mesh.n_inner]

    def piso_assemble(self, den, apu, u, v, w, dudx, dvdx, dwdx, flmass=None):
        """PISO's pressure assembly (facefluxmass_piso) on the system's own mesh: a and su go
        into the system, its solution vector is left as it is.  Arguments as for
        calcp_assemble."""
        keep = []

        def arg(x):
            if isinstance(x, int):
                return C.c_void_p(x)
            arr = np.ascontiguousarray(x, dtype=np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.c_void_p)

        out = None
        if flmass is None:
            out = np.zeros(max(self.mesh.n_inner, 1))
            flmass = out
        args = [arg(x) for x in (den, apu, u, v, w, dudx, dvdx, dwdx, flmass)]
        self.ctx.check(self.ctx.L.fc_sys_piso_assemble(self.h, *args))
        return None if out is None else out[:self.mesh.n_inner]

    def pin_reference(self, pref_cell, p):
        """PISO's reference pressure in the system's matrix (row pref_cell, 0-based, zeroed, its
        diagonal 1, su = p[pref_cell]); p: numpy array or int device address."""
        arr = None if isinstance(p, int) else np.ascontiguousarray(p, np.float64)
        ptr = C.c_void_p(p) if isinstance(p, int) else arr.ctypes.data_as(C.c_void_p)
        self.ctx.check(self.ctx.L.fc_sys_pin_reference(self.h, int(pref_cell) + 1, ptr))

    def bpres(self, p, istage, dpdx=None):
        """bpres(p, istage) on the system's mesh; p (numTotal) numpy (a new array is returned) or
        int device address (updated in place)."""
        bnd = _i32(boundary_groups(self.mesh))
        keep = []

        def arg(x, copy=False):
            if x is None:
                return C.c_void_p(None)
            if isinstance(x, int):
                return C.c_void_p(x)
            arr = np.array(x, dtype=np.float64, copy=True) if copy else np.ascontiguousarray(x, np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.c_void_p)

        pa = arg(p, copy=True)
        da = arg(dpdx)
        self.ctx.check(self.ctx.L.fc_sys_bpres(self.h, _ip(bnd), int(istage), da, pa))
        return None if isinstance(p, int) else keep[0]

    def grad_pp(self, nipgrad=2, nigrad=1, out=None):
        """The corrector's gradient of the system's solution pp (calcp:131-136): nipgrad stages
        of bpres + Gauss gradient; returns (numCells, 3) or None when out is a device address."""
        bnd = _i32(boundary_groups(self.mesh))
        res = None
        if out is None:
            res = np.zeros((self.n, 3))
            out = res
        o_arg = C.c_void_p(out) if isinstance(out, int) else out.ctypes.data_as(C.c_void_p)
        self.ctx.check(self.ctx.L.fc_sys_grad_pp(self.h, _ip(bnd), int(nipgrad), int(nigrad), o_arg))
        return res

    def correct_velocity_bc(self, flomas, den, u, v, w):
        """correctBoundaryConditionsVelocity on the system's mesh; numpy u, v, w are copied and
        (u, v, w, fmo) returned; int device addresses are updated in place (fmo then returned)."""
        bnd = _i32(boundary_groups(self.mesh))
        keep = []

        def arg(x, copy=False):
            if isinstance(x, int):
                return C.c_void_p(x)
            arr = np.array(x, dtype=np.float64, copy=True) if copy else np.ascontiguousarray(x, np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.c_void_p)

        fmo = np.zeros(max(int(bnd[2]), 1))
        dn = arg(den)
        vel = [arg(x, copy=True) for x in (u, v, w)]
        self.ctx.check(self.ctx.L.fc_sys_correct_velocity_bc(self.h, _ip(bnd), float(flomas), dn, *vel,
                                                             fmo.ctypes.data_as(C.c_void_p)))
        if isinstance(u, int):
            return fmo[:int(bnd[2])]
        return (*keep[1:4], fmo[:int(bnd[2])])

    def grad_gauss(self, u, nigrad=1, out=None):
        """Gauss gradient of u (cells, then boundary values in group order) on the system's own
        mesh (src/grad_gauss.f90:1-124), in the reference's (3, numCells) layout: returns an
        (numCells, 3) array, or None when out is an int device address (written in place).
        u may be a numpy array or an int device address."""
        bnd = _i32(boundary_groups(self.mesh))
        u_arg = C.c_void_p(u) if isinstance(u, int) else _f64(u)
        if not isinstance(u, int) and u_arg.shape[0] < self.ntotal:
            raise ValueError(f"u needs {self.ntotal} entries (cells, then boundary values)")
        res = None
        if out is None:
            res = np.zeros((self.n, 3))
            out = res
        o_arg = C.c_void_p(out) if isinstance(out, int) else out.ctypes.data_as(C.c_void_p)
        u_ptr = u_arg if isinstance(u, int) else u_arg.ctypes.data_as(C.c_void_p)
        self.ctx.check(self.ctx.L.fc_sys_grad_gauss(self.h, _ip(bnd), u_ptr, int(nigrad), o_arg))
        return res

    def calcp_correct(self, dpdx, apu, apv, apw, urf, pref_cell, flmass, u, v, w, p, pp=None):
        """calcp's corrections (src/calcp-multiple_correction_SIMPLE.f90:145-183) with the
        system's matrix and, when pp is None, its solution; pref_cell is 0-based.  numpy
        arguments are copied (the updated flmass, u, v, w, p are returned as new arrays);
        int device addresses are read and updated in place (then None is returned)."""
        keep = []

        def arg(x, copy=False):
            if isinstance(x, int):
                return C.c_void_p(x)
            arr = np.array(x, dtype=np.float64, copy=True) if copy else np.ascontiguousarray(x, np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.c_void_p)

        on_device = [isinstance(x, int) for x in (flmass, u, v, w, p)]
        if any(on_device) and not all(on_device):
            raise ValueError("flmass, u, v, w, p: all numpy arrays or all device addresses")
        ins = [C.c_void_p(None) if pp is None else arg(pp)] + [arg(x) for x in (dpdx, apu, apv, apw)]
        nin = len(keep)
        outs = [arg(x, copy=True) for x in (flmass, u, v, w, p)]
        self.ctx.check(self.ctx.L.fc_sys_calcp_correct(self.h, *ins, float(urf), int(pref_cell) + 1, *outs))
        return None if all(on_device) else tuple(keep[nin:])

    def calcp_nonorth(self, den, apu, dpdx, flmass):
        """Source of calcp's non-orthogonal corrector (src/calcp-multiple_correction_SIMPLE.f90:
        185-229) into the system's su, for the corrector's next solve.  numpy arguments are
        copied and the updated flmass is returned; an int device address for flmass is
        updated in place (then None is returned)."""
        keep = []

        def arg(x, copy=False):
            if isinstance(x, int):
                return C.c_void_p(x)
            arr = np.array(x, dtype=np.float64, copy=True) if copy else np.ascontiguousarray(x, np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.c_void_p)

        args = [arg(den), arg(apu), arg(dpdx), arg(flmass, copy=True)]
        self.ctx.check(self.ctx.L.fc_sys_calcp_nonorth(self.h, *args))
        return None if isinstance(flmass, int) else keep[-1]

    def set_rhs(self, su):
        self.ctx.check(self.ctx.L.fc_sys_set_rhs(self.h, _dp(_f64(su))))

    def solve_momentum(self, sp, src, urf, phi, method=METHOD_BICGSTAB, sor=1e-8, nsw=1000):
        """calcuvw's three component solves on the last convection-diffusion assembly
        (src/calcuvw.f90:414-556).  sp, src, phi: three arrays each (u, v, w); urf: three
        factors.  Returns (status, [phi_u, phi_v, phi_w], [apu, apv, apw], [info x 3])."""
        n = self.n
        sp_ = [_f64(x) for x in sp]
        src_ = [_f64(x) for x in src]
        phi_ = [_f64(x).copy() for x in phi]
        ap_ = [np.zeros(n) for _ in range(3)]
        urf_ = _f64(urf)
        arr = lambda xs: (C.POINTER(C.c_double) * 3)(*[_dp(x) for x in xs])  # noqa: E731
        infos = (SolveInfo * 3)()
        rc = self.ctx.L.fc_sys_solve_momentum(self.h, arr(sp_), arr(src_), _dp(urf_), arr(phi_), arr(ap_),
                                              int(method), float(sor), int(nsw), infos)
        self.ctx.check(rc, allow=(FC_OK, FC_NOT_CONVERGED, FC_EARLY_EXIT))
        return rc, phi_, ap_, [x.as_dict() for x in infos]

    def set_solution(self, fi):
        self.ctx.check(self.ctx.L.fc_sys_set_solution(self.h, _dp(_f64(fi))))

    def solution(self):
        fi = np.zeros(self.n)
        self.ctx.check(self.ctx.L.fc_sys_get_solution(self.h, _dp(fi)))
        return fi

    def network_layout(self):
        """(clusters, largest halo, halo rows in all) of the network's cluster layout on this
        system (0 clusters: none); builds the network's graph data if needed."""
        c, h, t = C.c_int(), C.c_int(), C.c_longlong()
        self.ctx.check(self.ctx.L.fc_sys_network_layout(self.h, C.byref(c), C.byref(h), C.byref(t)))
        return c.value, h.value, t.value

    def coarse_space(self, with_arrays: bool = False):
        """(aggregates, cells per aggregate, factored) of the learned FCG's coarse space on this
        system's current matrix; with_arrays adds each cell's aggregate and A_c = P^T A P."""
        nc, agg, ok = C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(self.ctx.L.fc_sys_coarse_space(self.h, C.byref(nc), C.byref(agg), C.byref(ok), None, None))
        if not with_arrays or nc.value == 0:
            return nc.value, agg.value, bool(ok.value)
        cell_agg = np.zeros(self.n, np.int32)
        ac = np.zeros((nc.value, nc.value))
        self.ctx.check(self.ctx.L.fc_sys_coarse_space(self.h, C.byref(nc), C.byref(agg), C.byref(ok),
                                                      cell_agg.ctypes.data_as(C.c_void_p),
                                                      ac.ctypes.data_as(C.c_void_p)))
        return nc.value, agg.value, bool(ok.value), cell_agg, ac.T.copy()

    def precondition(self, r):
        """One application of the learned FCG's preconditioner (attached network, Jacobi steps,
        coarse correction) to r, in FP32."""
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        self.ctx.check(self.ctx.L.fc_sys_precondition(self.h, r.ctypes.data_as(C.c_void_p),
                                                      z.ctypes.data_as(C.c_void_p)))
        return z

    def solve(self, method=METHOD_ICCG, stop_rule=STOP_L1, sor=1e-8, nsw=1000):
        info = SolveInfo()
        rc = self.ctx.L.fc_sys_solve(self.h, method, stop_rule, float(sor), int(nsw), C.byref(info))
        self.ctx.check(rc, allow=(FC_OK, FC_NOT_CONVERGED, FC_EARLY_EXIT, FC_FALLBACK_USED))
        return rc, info.as_dict()

    def get(self, name):
        sizes = self.sizes()
        n, nif, nnz = sizes["n"], sizes["nif"], sizes["nnz"]
        nf = self.mesh.n_faces
        count = {"a": nnz, "su": n, "ioffset": n + 1, "ja": nnz, "diag": n, "ij_slot": nif, "ji_slot": nif,
                 "vol": n, "xc": n, "yc": n, "zc": n, "facint": nif, "arx": nf, "ary": nf, "arz": nf,
                 "xf": nf, "yf": nf, "zf": nf, "res": n, "pivots": n}[name]
        dtype = np.int32 if name in ("ioffset", "ja", "diag", "ij_slot", "ji_slot") else np.float64
        out = np.zeros(max(count, 1), dtype)
        self.ctx.check(self.ctx.L.fc_sys_get(self.h, ARR[name], out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out[:count]

    def device_ptr(self, name) -> int:
        return self.ctx.L.fc_sys_device_ptr(self.h, ARR[name]) or 0

    def sizes(self):
        v = [C.c_int() for _ in range(5)]
        self.ctx.check(self.ctx.L.fc_sys_sizes(self.h, *(C.byref(x) for x in v)))
        return dict(zip(("n", "nif", "nnz", "nlev_fwd", "nlev_bwd"), (x.value for x in v)))

    def last_ms(self, stage):
        return self.ctx.L.fc_sys_last_ms(self.h, stage)

    def sweep_bench(self, dilu=False, reps=10):
        """Average device ms of one FP64 IC(0) (or DILU) forward + backward sweep pair."""
        ms = self.ctx.L.fc_sys_sweep_bench(self.h, 1 if dilu else 0, int(reps))
        if ms < 0:
            raise FcError(-2, self.ctx.L.fc_last_error(self.ctx.h).decode())
        return ms

    def spmv_bench(self, precision=64, reps=20):
        ms = self.ctx.L.fc_sys_spmv_bench(self.h, precision, reps)
        if ms < 0:
            raise FcError(-2, self.ctx.L.fc_last_error(self.ctx.h).decode())
        return ms
