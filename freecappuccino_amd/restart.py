"""Restart records of the reference (src/write_restart_files.f90:22-47, src/readfiles.f90:19-43).

One Fortran unformatted sequential record per `write`, in this order:
(itime, time), gradpcmf (only with constant mass flow), then the float64
fields of FIELDS.  The library's C-ABI (fc_restart_write / _probe / _read)
does the I/O; this module names the records.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import fcap

# record order after the header; the sizes the reference allocates them with
FIELDS = ("flmass", "u", "v", "w", "p", "te", "ed", "t", "vis", "uu", "vv", "ww", "uv", "uw", "vw",
          "uo", "vo", "wo", "teo", "edo")
SIZES = {"flmass": "numInnerFaces", "uu": "numCells", "vv": "numCells", "ww": "numCells", "uv": "numCells",
         "uw": "numCells", "vw": "numCells"}   # the rest: numTotal


def _bind(L):
    if getattr(L, "_restart_bound", False):
        return L
    ll, dp, vp = C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.c_char_p
    L.fc_restart_write.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, ll,
                                   C.POINTER(dp), C.c_char_p, C.c_int]
    L.fc_restart_probe.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, dp, C.c_int, C.POINTER(C.c_int), ll,
                                   C.c_char_p, C.c_int]
    L.fc_restart_read.argtypes = [vp, C.c_int, C.c_int, ll, C.POINTER(dp), C.c_char_p, C.c_int]
    for f in (L.fc_restart_write, L.fc_restart_probe, L.fc_restart_read):
        f.restype = C.c_int
    L._restart_bound = True
    return L


def write(path: str, fields: Dict[str, np.ndarray], itime: int = 0, time: float = 0.0,
          gradpcmf: Optional[float] = None, lib: Optional[C.CDLL] = None):
    """Write every field of FIELDS (all must be given, as the reference writes them all)."""
    L = _bind(lib or fcap.library())
    arrs = [np.ascontiguousarray(fields[k], np.float64) for k in FIELDS]
    lens = np.array([a.size for a in arrs], np.int64)
    ptrs = (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    err = C.create_string_buffer(256)
    rc = L.fc_restart_write(path.encode(), int(itime), float(time), 0 if gradpcmf is None else 1,
                            0.0 if gradpcmf is None else float(gradpcmf), len(arrs),
                            lens.ctypes.data_as(C.POINTER(C.c_longlong)), ptrs, err, len(err))
    if rc != 0:
        raise fcap.FcError(rc, err.value.decode())


def read(path: str, const_mflux: bool = False, lib: Optional[C.CDLL] = None):
    """Returns (itime, time, gradpcmf or None, {name: array}) for the records in the file."""
    L = _bind(lib or fcap.library())
    it, t, g, nrec = C.c_int(), C.c_double(), C.c_double(), C.c_int()
    lens = np.zeros(64, np.int64)
    err = C.create_string_buffer(256)
    rc = L.fc_restart_probe(path.encode(), int(const_mflux), C.byref(it), C.byref(t), C.byref(g), lens.size,
                            C.byref(nrec), lens.ctypes.data_as(C.POINTER(C.c_longlong)), err, len(err))
    if rc != 0:
        raise fcap.FcError(rc, err.value.decode())
    n = min(nrec.value, lens.size)
    arrs = [np.zeros(int(lens[k])) for k in range(n)]
    ptrs = (C.POINTER(C.c_double) * max(n, 1))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    rc = L.fc_restart_read(path.encode(), int(const_mflux), n, lens.ctypes.data_as(C.POINTER(C.c_longlong)), ptrs,
                           err, len(err))
    if rc != 0:
        raise fcap.FcError(rc, err.value.decode())
    names = list(FIELDS) + [f"record{k}" for k in range(len(FIELDS), n)]
    return it.value, t.value, (g.value if const_mflux else None), dict(zip(names, arrs))
