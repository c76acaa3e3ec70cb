"""CPU restatement of the reference's on-disk formats, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
readers and writers (fdf_amd/csrc/io.hip); the product path never uses it.

PETSc binary Vec files.  The reference writes and reads them through PETSc
3.19.2 (a 64-bit-index build, compile_log:11) with PetscViewerBinaryOpen +
VecView / VecLoad (io.c:253-436, 657-745).  PETSc is not in this image, so
the format is restated from PETSc's published binary layout: the Vec class
id VEC_FILE_CLASSID = 1211214 and the global length, each a PetscInt (8
bytes here, 4 in 32-bit-index builds), then the values as doubles, all
big-endian; the viewer also writes "<file>.info", which holds
"-vecload_block_size <bs>" for vectors with block size > 1.  File names
follow io.c:680-682: results/<field>%05d_0.<ext>.

FDFGRID grid files (README.md:134-146): "FDFGRID", the block count, one
"IM JM KM" line per block, then "x y z" per node, k slowest, i fastest.
ReadGridFile (io.c:101-176) takes the first token as the block count when it
is not "FDFGRID"; AssignGridCoordinates (grid.c:356-383) then skips nblk + 2
lines before the coordinates whatever the first token was.
"""
from __future__ import annotations

import os

import numpy as np

VEC_FILE_CLASSID = 1211214


def write_petsc_vec(path, data, block_size: int = 1, int_bytes: int = 8) -> None:
    data = np.asarray(data, np.float64).ravel()
    it = ">i8" if int_bytes == 8 else ">i4"
    with open(path, "wb") as f:
        f.write(np.array([VEC_FILE_CLASSID, data.size], dtype=it).tobytes())
        f.write(data.astype(">f8").tobytes())
    with open(str(path) + ".info", "w") as f:
        if block_size > 1:
            f.write(f"-vecload_block_size {block_size}\n")


def read_petsc_vec(path) -> np.ndarray:
    raw = open(path, "rb").read()
    if int(np.frombuffer(raw[:4], ">i4")[0]) == VEC_FILE_CLASSID:
        n, off = int(np.frombuffer(raw[4:8], ">i4")[0]), 8
    else:
        cid, n = (int(v) for v in np.frombuffer(raw[:16], ">i8"))
        if cid != VEC_FILE_CLASSID:
            raise ValueError("not a PETSc Vec file")
        off = 16
    vals = np.frombuffer(raw[off:off + 8 * n], ">f8")
    if vals.size != n:
        raise ValueError("short file")
    return vals.astype(np.float64)


def field_file_name(field: str, step: int, directory: str = "results", ext: str = "dat") -> str:
    return os.path.join(directory, f"{field}{step:05d}_0.{ext}")


def read_grid_file(path):
    """(IM, JM, KM, coor, nblk) with the reference's header handling."""
    lines = open(path).read().splitlines()
    toks = " ".join(lines).split()
    if toks[0] == "FDFGRID":
        nblk, p = int(toks[1]), 2
    else:
        nblk, p = int(toks[0]), 1
    nblk = nblk or 1
    IM, JM, KM = (int(t) for t in toks[p:p + 3])
    nn = (IM + 1) * (JM + 1) * (KM + 1)
    body = lines[nblk + 2:nblk + 2 + nn]
    if len(body) < nn:
        raise ValueError("Error reading grid coordinates from file")
    coor = np.array([[float(v) for v in ln.split()[:3]] for ln in body], np.float64)
    return IM, JM, KM, coor, nblk


def write_grid_file(path, IM, JM, KM, coor, header: bool = True) -> None:
    coor = np.asarray(coor, np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        if header:
            f.write("FDFGRID\n")
        f.write(f"1\n{IM} {JM} {KM}\n")
        for x, y, z in coor.tolist():
            f.write(f"{x!r} {y!r} {z!r}\n")


def read_vtk_xml(path):
    """Arrays of a VTK XML file with appended raw data: {Name: ndarray}, plus the
    header text.  Counts are read with the width header_type names (UInt32 if
    absent), as VTK readers do; offsets count from the byte after "_"."""
    import re

    raw = open(path, "rb").read()
    head, rest = raw.split(b'<AppendedData encoding="raw">', 1)
    start = rest.index(b"_") + 1
    blob = rest[start:]
    text = head.decode()
    hdr = re.search(r'header_type="(\w+)"', text)
    cw = 8 if hdr and hdr.group(1) == "UInt64" else 4
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)"(?: NumberOfComponents="(\d+)")? '
                         r'format="appended" offset="(\d+)"', text):
        typ, name, nc, off = m.group(1), m.group(2), int(m.group(3) or 1), int(m.group(4))
        nbytes = int.from_bytes(blob[off:off + cw], "little")
        dt = {"Float64": "<f8", "Int64": "<i8", "Int32": "<i4"}[typ]
        a = np.frombuffer(blob[off + cw:off + cw + nbytes], dt)
        arrays[name] = a.reshape(-1, nc) if nc > 1 else a
    return arrays, text
