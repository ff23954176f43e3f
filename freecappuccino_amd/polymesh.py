"""polyMesh reader: the library's parallel host parser behind a PolyMesh.

    from freecappuccino_amd.polymesh import read_polymesh
    mesh = read_polymesh("case/constant/polyMesh")      # OpenFOAM ASCII or the native layout

Mirrors the reading part of the reference's mesh set-up
(src/mesh_geometry_and_topology.f90:409-605): the same five files, the same
two layouts, nCells from the owner note.  The geometry then comes from the
GPU (``fcap.System(mesh)`` or ``fcap.mesh_geometry``).  No GPU is needed to
read a mesh.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import fcap
from .mesh import PolyMesh

_BOUND = set()


def _bind(L: C.CDLL) -> C.CDLL:
    if id(L) in _BOUND:
        return L
    vp, i, ip = C.c_void_p, C.c_int, C.POINTER(C.c_int)
    sig = {
        "fc_polymesh_read": ([C.c_char_p, i, C.POINTER(vp), C.c_char_p, i], i),
        "fc_polymesh_sizes": ([vp, ip, ip, ip, ip, ip, ip], i),
        "fc_polymesh_arrays": ([vp, C.POINTER(C.c_double), ip, ip, ip, ip], i),
        "fc_polymesh_patch": ([vp, i, C.c_char_p, i, ip, ip], i),
        "fc_polymesh_free": ([vp], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _BOUND.add(id(L))
    return L


def read_polymesh(directory: str, nthreads: int = 0, lib: Optional[C.CDLL] = None) -> PolyMesh:
    """Read points/faces/owner/neighbour/boundary from `directory` (0-based PolyMesh)."""
    L = _bind(lib or fcap.library())
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.fc_polymesh_read(directory.encode(), int(nthreads), C.byref(h), err, len(err))
    if rc != 0:
        raise fcap.FcError(rc, err.value.decode() or "fc_polymesh_read failed")
    try:
        v = [C.c_int() for _ in range(6)]
        L.fc_polymesh_sizes(h, *(C.byref(x) for x in v))
        nn, nf, nif, n, nfn, npat = (x.value for x in v)
        xyz = np.zeros(3 * nn)
        ptr = np.zeros(nf + 1, np.int32)
        fn = np.zeros(max(nfn, 1), np.int32)
        own = np.zeros(max(nf, 1), np.int32)
        nei = np.zeros(max(nif, 1), np.int32)
        ipt = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        L.fc_polymesh_arrays(h, xyz.ctypes.data_as(C.POINTER(C.c_double)), ipt(ptr), ipt(fn), ipt(own), ipt(nei))
        patches = []
        name = C.create_string_buffer(64)
        for k in range(npat):
            a, b = C.c_int(), C.c_int()
            L.fc_polymesh_patch(h, k, name, len(name), C.byref(a), C.byref(b))
            patches.append((name.value.decode(), a.value, b.value))
    finally:
        L.fc_polymesh_free(h)
    return PolyMesh(points=xyz.reshape(nn, 3), face_ptr=ptr, face_nodes=fn[:nfn] - 1, owner=own[:nf] - 1,
                    neighbour=nei[:nif] - 1, n_cells=n, patches=patches)
