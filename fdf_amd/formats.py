"""The reference's on-disk formats, through the library's readers and writers.

    read_grid_file / write_grid_file   FDFGRID grid files (io.c:101-176, grid.c:356-383)
    read_petsc_vec / write_petsc_vec   PETSc binary Vec files (io.c:253-436, 657-745)
    field_file_name                    results/<field>%05d_0.<ext> (io.c:680-682)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L


def _b(s) -> bytes:
    return os.fsencode(s)


def read_grid_file(path, coordinates: bool = True):
    """(IM, JM, KM, coor [(KM+1)(JM+1)(IM+1), 3] or None, nblk) of block 0."""
    lib = L.load()
    IM, JM, KM, nb = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    L.check(lib.picgpu_read_grid_file(_b(path), C.byref(IM), C.byref(JM), C.byref(KM), C.byref(nb), None))
    coor = None
    if coordinates:
        coor = np.empty(((IM.value + 1) * (JM.value + 1) * (KM.value + 1), 3), np.float64)
        L.check(lib.picgpu_read_grid_file(_b(path), C.byref(IM), C.byref(JM), C.byref(KM), C.byref(nb),
                                          coor.ctypes.data_as(C.c_void_p)))
    return IM.value, JM.value, KM.value, coor, nb.value


def write_grid_file(path, IM: int, JM: int, KM: int, coor) -> None:
    coor = np.ascontiguousarray(coor, np.float64).reshape(-1, 3)
    assert coor.shape[0] == (IM + 1) * (JM + 1) * (KM + 1)
    L.check(L.load().picgpu_write_grid_file(_b(path), IM, JM, KM, coor.ctypes.data_as(C.c_void_p)))


def field_file_name(field: str, step: int, directory: str = "results", ext: str = "dat") -> str:
    buf = C.create_string_buffer(4096)
    L.check(L.load().picgpu_field_file_name(buf, len(buf), _b(directory), field.encode(), int(step), ext.encode()))
    return os.fsdecode(buf.value)


def petsc_vec_length(path) -> int:
    n = C.c_int64()
    L.check(L.load().picgpu_petsc_vec_length(_b(path), C.byref(n)))
    return n.value


def read_petsc_vec(path) -> np.ndarray:
    n = petsc_vec_length(path)
    out = np.empty(n, np.float64)
    L.check(L.load().picgpu_read_petsc_vec(_b(path), n, out.ctypes.data_as(C.c_void_p)))
    return out


def write_petsc_vec(path, data, block_size: int = 1) -> None:
    data = np.ascontiguousarray(data, np.float64).ravel()
    L.check(L.load().picgpu_write_petsc_vec(_b(path), data.size, int(block_size), data.ctypes.data_as(C.c_void_p)))


def write_vts(path, IM: int, JM: int, KM: int, coor, name: str, field) -> None:
    """Structured-grid VTK file of one node field (1 or 3 components)."""
    coor = np.ascontiguousarray(coor, np.float64).reshape(-1, 3)
    field = np.ascontiguousarray(field, np.float64)
    ncomp = field.size // coor.shape[0]
    L.check(L.load().picgpu_write_vts(_b(path), IM, JM, KM, coor.ctypes.data_as(C.c_void_p), name.encode(), ncomp,
                                      field.ctypes.data_as(C.c_void_p)))


def write_vtp(path, pos, name: str, field) -> None:
    """Poly-data VTK file of points with one point field (1 or 3 components)."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    field = np.ascontiguousarray(field, np.float64)
    n = pos.shape[0]
    ncomp = field.size // max(n, 1) if n else 3
    L.check(L.load().picgpu_write_vtp(_b(path), n, pos.ctypes.data_as(C.c_void_p), name.encode(), ncomp,
                                      field.ctypes.data_as(C.c_void_p)))
