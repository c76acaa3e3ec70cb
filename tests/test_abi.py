"""The C-ABI library loads and exports every entry point include/picgpu.h declares (no GPU calls)."""
import ctypes
import os
import re

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
