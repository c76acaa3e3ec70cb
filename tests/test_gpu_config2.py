"""Config 2 at full size: the 64x64x512 bent tube with its Poiseuille field,
4,194,304 particles placed in random cells (each uploaded with that cell as
the walk's start), 50 RK2 steps with face deposition, and the face volume
fractions that are config 2's output.  Reduced-size runs of the same
workload are checked against the oracle particle by particle
(test_gpu_parity, case tube_rk2); here, at full size, through properties
that hold at any size:

- every particle is located from its start cell;
- the count deposited on each direction's faces sums to the located
  particles (a particle adds 1 - a and a to its two faces);
- the volume fractions are the oracle's formula (count * V_p over half the
  two adjacent cells' volumes) applied to the device's own face counts, bit
  for bit;
- the walks stay short (a particle moves at most half a cell per step).
"""
import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


def _particles_in_cells(coor, IM, JM, KM, n, seed):
    """n points at trilinear coordinates in [0.05, 0.95]^3 of uniformly drawn
    cells, and those cells."""
    rng = np.random.default_rng(seed)
    cell = np.stack([rng.integers(0, IM, n), rng.integers(0, JM, n), rng.integers(0, KM, n)], 1)
    xi = rng.uniform(0.05, 0.95, (n, 3))
    c3 = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    pos = np.zeros((n, 3))
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                wgt = ((xi[:, 0] if di else 1 - xi[:, 0]) * (xi[:, 1] if dj else 1 - xi[:, 1]) *
                       (xi[:, 2] if dk else 1 - xi[:, 2]))
                pos += wgt[:, None] * c3[cell[:, 2] + dk, cell[:, 1] + dj, cell[:, 0] + di]
    return pos, cell


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config2_full_size_properties(swarm_cls, oracle, precision):
    d = W.config(2, with_particles=False)
    IM, JM, KM, n, steps = d["IM"], d["JM"], d["KM"], d["n"], d["steps"]
    assert (IM, JM, KM, n, steps) == (64, 64, 512, 4_194_304, 50)
    vp = 1.0 / n
    _say("config2 full size", precision, IM, JM, KM, n, steps)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, d["coor"])
        g.upload_ucat(d["ucat"])
        pos, cell = _particles_in_cells(d["coor"], IM, JM, KM, n, 1002)
        g.upload_particles(pos, cell=cell)
        g.LocateAllParticlesInGrid()
        st0 = g.stats()
        g.InterpolateAllFieldsToSwarm()
        g.set_particle_volume(vp)
        g.AdvanceSimulation(d["dt"], steps, rk2=True, deposit_cells=False, deposit_faces=True)
        st = g.stats()
        faces = [g.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)]
        vfs = [g.field(f) for f in (L.FIELD_IFACE_VOLFRAC, L.FIELD_JFACE_VOLFRAC, L.FIELD_KFACE_VOLFRAC)]
    located = st["located"]
    _say("located", st0["located"], "->", located, "walk iterations per particle-step",
         st["walk_iterations"] / max(1, located))
    assert st0["located"] == n and st0["lost"] == 0
    # half a cell per step at most: after 50 steps the particles of the last
    # 25 of 512 k-planes may have left through the outlet, no others
    assert 0.9 * n < located <= n and st["n_particles"] == located
    assert st["walk_iterations"] <= 3 * located
    tol = 1e-9 if precision == "fp64" else 1e-5
    for dd, (f, vf) in enumerate(zip(faces, vfs)):
        f4 = f.reshape(-1, 4)
        assert (f4[:, 0] >= 0).all()
        assert abs(f4[:, 0].sum() - located) <= tol * located
        assert np.array_equal(vf, oracle.face_volfrac(IM, JM, KM, d["coor"], dd, f, vp))
