"""The boundary-condition file parser of the library (host code, no GPU):
ParseAllBoundaryConditions semantics (io.c:2515-2680) and the handler
parameters the inlet handlers read (BC_Handlers.c:246-251, 658-660)."""
import pytest

from fdf_amd import _lib as L
from fdf_amd.boundary import BoundaryConfig


def _write(tmp_path, text):
    p = tmp_path / "bcs.dat"
    p.write_text(text)
    return p


def test_defaults_are_noslip_walls():
    cfg = BoundaryConfig.default()
    assert cfg.type == [L.BCTYPE_WALL] * 6 and cfg.handler == [L.BC_WALL_NOSLIP] * 6
    assert cfg.vmax == [1.0] * 6 and all(v == [0.0, 0.0, 0.0] for v in cfg.vel)


def test_parse_faces_types_handlers_and_parameters(tmp_path):
    p = _write(tmp_path, "# comment\n\n   -Zeta  INLET constant_velocity vz=2.5 junk vx=0.1\n"
                         "+zeta nograd ALLCOPY\n+Eta inlet Parabolic vmax=3 vmax=4\n-Xi WALL noslip\n"
                         "-Eta WALL\n")
    cfg = BoundaryConfig.from_file(p)
    assert cfg.type == [L.BCTYPE_WALL, L.BCTYPE_WALL, L.BCTYPE_WALL, L.BCTYPE_INLET, L.BCTYPE_INLET,
                        L.BCTYPE_NOGRAD]
    assert cfg.handler == [L.BC_WALL_NOSLIP, L.BC_WALL_NOSLIP, L.BC_WALL_NOSLIP, L.BC_INLET_PARABOLIC,
                           L.BC_INLET_CONSTANT_VELOCITY, L.BC_NOGRAD_COPY_GHOST]
    assert cfg.vel[4] == [0.1, 0.0, 2.5]
    assert cfg.vmax[3] == 4.0  # the last value of a repeated key wins


def test_a_later_line_replaces_the_face(tmp_path):
    p = _write(tmp_path, "-Zeta INLET constant_velocity vz=2\n-Zeta INLET parabolic\n")
    cfg = BoundaryConfig.from_file(p)
    assert cfg.handler[4] == L.BC_INLET_PARABOLIC and cfg.vel[4] == [0.0, 0.0, 0.0] and cfg.vmax[4] == 1.0


@pytest.mark.parametrize("text", ["-Q WALL noslip\n", "-Xi SOLID noslip\n", "-Xi WALL sticky\n",
                                  "-Xi WALL allcopy\n", "-Xi INLET noslip\n", "-Xi NOGRAD noslip\n"])
def test_parse_errors(tmp_path, text):
    with pytest.raises(L.PicgpuError) as e:
        BoundaryConfig.from_file(_write(tmp_path, text))
    assert e.value.code == -1


def test_missing_file(tmp_path):
    with pytest.raises(L.PicgpuError):
        BoundaryConfig.from_file(tmp_path / "nope.dat")
