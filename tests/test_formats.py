"""The library's readers and writers of the reference's on-disk formats
(host code, no GPU) against the numpy restatement in oracle/formats.py and
against bytes fixed by the format itself (PETSc's class id, big-endian
IEEE-754 doubles)."""
import os

import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import formats as F
from fdf_amd import workloads as W
from oracle import formats as O


def test_known_bytes_of_a_vec_file(tmp_path):
    """[1.0, -2.0, 0.5]: 64-bit PetscInt class id 1211214 = 0x127B4E and length 3,
    then the doubles big-endian."""
    p = tmp_path / "v.dat"
    F.write_petsc_vec(p, [1.0, -2.0, 0.5], block_size=3)
    raw = p.read_bytes()
    assert raw[:16] == bytes.fromhex("0000000000127b4e" "0000000000000003")
    assert raw[16:] == bytes.fromhex("3ff0000000000000" "c000000000000000" "3fe0000000000000")
    assert (tmp_path / "v.dat.info").read_text() == "-vecload_block_size 3\n"


def test_writer_matches_restatement_byte_for_byte(tmp_path):
    x = np.random.default_rng(1).normal(size=3 * 1000 + 7)
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    F.write_petsc_vec(a, x, block_size=1)
    O.write_petsc_vec(b, x, block_size=1)
    assert a.read_bytes() == b.read_bytes()
    assert (tmp_path / "a.dat.info").read_text() == ""


@pytest.mark.parametrize("int_bytes", [8, 4])
def test_reader_reads_both_header_widths(tmp_path, int_bytes):
    x = np.random.default_rng(2).normal(size=(2_500_000,))  # more than one I/O chunk
    p = tmp_path / f"v{int_bytes}.dat"
    O.write_petsc_vec(p, x, int_bytes=int_bytes)
    assert F.petsc_vec_length(p) == x.size
    assert np.array_equal(F.read_petsc_vec(p), x)


def test_round_trip_special_values(tmp_path):
    x = np.array([0.0, -0.0, np.inf, -np.inf, 5e-324, 1.7976931348623157e308, np.pi])
    p = tmp_path / "s.dat"
    F.write_petsc_vec(p, x)
    y = F.read_petsc_vec(p)
    assert y.tobytes() == x.tobytes()
    assert np.array_equal(O.read_petsc_vec(p).view(np.int64), x.view(np.int64))


def test_vec_errors(tmp_path):
    lib = L.load()
    p = tmp_path / "v.dat"
    O.write_petsc_vec(p, np.arange(6.0))
    out = np.empty(5)
    with pytest.raises(L.PicgpuError):  # length mismatch (io.c:301-305)
        L.check(lib.picgpu_read_petsc_vec(os.fsencode(p), 5, out.ctypes.data))
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"\x00" * 32)
    with pytest.raises(L.PicgpuError):
        F.read_petsc_vec(bad)
    with pytest.raises(L.PicgpuError):
        F.read_petsc_vec(tmp_path / "missing.dat")


def test_field_file_names():
    assert F.field_file_name("ufield", 12) == "results/ufield00012_0.dat"
    assert F.field_file_name("position", 123456, "out", "dat") == "out/position123456_0.dat"
    assert F.field_file_name("vfield", 3) == O.field_file_name("vfield", 3)


def test_grid_file_round_trip(tmp_path):
    IM, JM, KM = 6, 5, 7
    coor = W.perturbed_cube(IM, JM, KM, seed=9)
    p = tmp_path / "grid.dat"
    F.write_grid_file(p, IM, JM, KM, coor)
    im, jm, km, c, nb = F.read_grid_file(p)
    assert (im, jm, km, nb) == (IM, JM, KM, 1) and np.array_equal(c, coor)
    im, jm, km, c2, nb = O.read_grid_file(p)
    assert (im, jm, km) == (IM, JM, KM) and np.array_equal(c2, coor)
    # a file written by the restatement (shortest repr) reads to the same doubles
    q = tmp_path / "grid2.dat"
    O.write_grid_file(q, IM, JM, KM, coor)
    assert np.array_equal(F.read_grid_file(q)[3], coor)


def test_grid_file_header_rules(tmp_path):
    IM, JM, KM = 2, 2, 2
    coor = W.uniform_grid(IM, JM, KM)
    # sizes are read with or without the FDFGRID tag ...
    q = tmp_path / "nohdr.dat"
    O.write_grid_file(q, IM, JM, KM, coor, header=False)
    assert F.read_grid_file(q, coordinates=False)[:3] == (IM, JM, KM)
    # ... but the coordinate read skips nblk + 2 lines as the reference does,
    # so a file without the tag loses its first node and runs out of lines
    with pytest.raises(L.PicgpuError):
        F.read_grid_file(q)
    with pytest.raises(ValueError):
        O.read_grid_file(q)
    # block count 0 means one block; a second block's sizes are read and skipped
    r = tmp_path / "two.dat"
    body = "".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in coor.tolist())
    r.write_text("FDFGRID\n2\n2 2 2\n4 4 4\n" + body)
    im, jm, km, c, nb = F.read_grid_file(r)
    assert (im, jm, km, nb) == (2, 2, 2, 2) and np.array_equal(c, coor)
    with pytest.raises(L.PicgpuError):
        F.read_grid_file(tmp_path / "missing.dat")


VTS_HEAD = """<?xml version="1.0"?>
<VTKFile type="StructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt64">
  <StructuredGrid WholeExtent="0 {mx} 0 {my} 0 {mz}">
    <Piece Extent="0 {mx} 0 {my} 0 {mz}">
      <Points>
        <DataArray type="Float64" Name="Position" NumberOfComponents="3" format="appended" offset="0" />
      </Points>
      <PointData Scalars="{name}">
        <DataArray type="Float64" Name="{name}" NumberOfComponents="{nc}" format="appended" offset="{off}" />
      </PointData>
    </Piece>
  </StructuredGrid>
  """


@pytest.mark.parametrize("ncomp", [1, 3])
def test_vts_layout_and_values(tmp_path, ncomp):
    IM, JM, KM = 4, 3, 5
    coor = W.perturbed_cube(IM, JM, KM, seed=2)
    nn = (IM + 1) * (JM + 1) * (KM + 1)
    field = np.random.default_rng(3).normal(size=(nn, ncomp)).squeeze()
    p = tmp_path / "Ucat_00007.vts"
    F.write_vts(p, IM, JM, KM, coor, "Ucat", field)
    arrays, text = O.read_vtk_xml(p)
    assert text == VTS_HEAD.format(mx=IM, my=JM, mz=KM, name="Ucat", nc=ncomp, off=8 + 3 * nn * 8)
    assert np.array_equal(arrays["Position"], coor) and np.array_equal(arrays["Ucat"], field)
    assert p.read_bytes().endswith(b"\n </AppendedData>\n</VTKFile>\n")


def test_vtp_layout_and_values(tmp_path):
    n = 1000
    rng = np.random.default_rng(4)
    pos, vel = rng.random((n, 3)), rng.normal(size=(n, 3))
    p = tmp_path / "velocity_00003.vtp"
    F.write_vtp(p, pos, "velocity", vel)
    arrays, text = O.read_vtk_xml(p)
    assert f'<Piece NumberOfPoints="{n}" NumberOfVerts="{n}" NumberOfLines="0"' in text
    offs = [int(v) for v in __import__("re").findall(r'offset="(\d+)"', text)]
    assert offs == [0, 8 + 24 * n, 2 * (8 + 24 * n), 2 * (8 + 24 * n) + 8 + 8 * n]
    assert np.array_equal(arrays["Position"], pos) and np.array_equal(arrays["velocity"], vel)
    assert np.array_equal(arrays["connectivity"], np.arange(n)) and np.array_equal(arrays["offsets"], np.arange(1, n + 1))
