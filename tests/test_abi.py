"""The C-ABI library loads and exports every entry point include/picgpu.h declares (no GPU calls)."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np

from fdf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ensure_built():
    if not os.path.exists(_lib.LIB_PATH):
        from fdf_amd import build

        build.build()


def _declared():
    src = open(os.path.join(ROOT, "include", "picgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(picgpu_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_reference_stages():
    names = _declared()
    for n in ("picgpu_advect", "picgpu_locate", "picgpu_interpolate", "picgpu_scatter", "picgpu_step",
              "picgpu_remove_lost", "picgpu_remove_out_of_bounds", "picgpu_reset_statuses",
              "picgpu_deposit_faces", "picgpu_mlp_train"):
        assert n in names


def test_library_exports_every_declared_symbol():
    _ensure_built()
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.exported_symbols()) == _declared()


def test_options_struct_matches_header_defaults():
    _ensure_built()
    lib = _lib.load()
    opt = _lib.PicgpuOptions()
    lib.picgpu_default_options(ctypes.byref(opt))
    assert opt.max_traversal == 1000
    assert opt.distance_threshold == 1e-11
    assert opt.weight_clamp == 1e-14
    assert opt.bbox_tolerance == 1e-6
    assert opt.precision == _lib.FP64


def test_c_consumer_sees_the_binding_layout(tmp_path):
    """A C program built with gcc against include/picgpu.h (what a PETSc-side
    shim is) agrees with the Python binding on every struct's size and field
    offsets, and calls the library through dlopen: the default options and
    a PETSc binary Vec file whose bytes the format fixes."""
    _ensure_built()
    gcc = shutil.which("gcc")
    assert gcc, "gcc is needed to build the C consumer"
    exe = tmp_path / "abi_consumer"
    subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "abi_consumer.c"), "-o", str(exe), "-ldl"], check=True)
    vec = tmp_path / "v.dat"
    out = subprocess.run([str(exe), _lib.LIB_PATH, str(vec)], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    structs = {"picgpu_options": _lib.PicgpuOptions, "picgpu_stats": _lib.PicgpuStats,
               "picgpu_boundary_config": _lib.PicgpuBoundaryConfig}
    seen = {k: set() for k in structs}
    for line in out:
        w = line.split()
        if w[0] == "sizeof":
            assert ctypes.sizeof(structs[w[1]]) == int(w[2]), line
        elif "." in w[0] and w[0].split(".")[0] in structs:
            name, field = w[0].split(".")
            desc = getattr(structs[name], field)
            assert (desc.offset, desc.size) == (int(w[1]), int(w[2])), line
            seen[name].add(field)
        elif w[0] == "default":
            kv = dict(zip(w[1::2], w[2::2]))
            opt = _lib.PicgpuOptions()
            _lib.load().picgpu_default_options(ctypes.byref(opt))
            for k, v in kv.items():
                assert float(v) == float(getattr(opt, k)), (k, v)
        elif w[0] == "write_petsc_vec":
            assert w[1] == "0"
    for name, cls in structs.items():
        assert seen[name] == {f[0] for f in cls._fields_}, name
    data = vec.read_bytes()
    assert int.from_bytes(data[:8], "big") == 1211214 and int.from_bytes(data[8:16], "big") == 6
    assert np.array_equal(np.frombuffer(data[16:], ">f8"), [1.0, -2.5, 3.25, 0.0, 1e300, -1e-300])
    assert (tmp_path / "v.dat.info").read_text().strip() == "-vecload_block_size 3"
