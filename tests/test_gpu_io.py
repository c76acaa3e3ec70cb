"""Output and restart files of a GPU context (PETSc binary Vec files),
checked with the numpy restatement of the format (oracle/formats.py)."""
import os

import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W
from fdf_amd.boundary import channel
from oracle import formats as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


def _f(d, name, step):
    return O.read_petsc_vec(O.field_file_name(name, step, str(d)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_write_output_then_restart(swarm_cls, oracle, tmp_path, precision):
    _say("output + restart", precision)
    IM, JM, KM = 16, 16, 40
    coor = W.tube_grid(IM, JM, KM)
    pos = W.particles_uniform_tube(20000, 91)
    step = 7
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.BoundarySystem_Create(channel(4, "constant_velocity", 1.0))
        g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        g.AdvanceSimulation(0.05, 3, rk2=True, deposit_cells=True)
        g.WriteOutput(step, str(tmp_path))
        d = g.download_particles()
        ucat, ucont = g.field(L.FIELD_UCAT), g.field(L.FIELD_UCONT)
        count, pavg = g.field(L.FIELD_PARTICLE_COUNT), g.field(L.FIELD_P_AVERAGE)
    n = len(d["pid"])
    assert np.array_equal(_f(tmp_path, "position", step), d["pos"].ravel())
    assert np.array_equal(_f(tmp_path, "velocity", step), d["vel"].ravel())
    assert np.array_equal(_f(tmp_path, "ufield", step), ucat)
    assert np.array_equal(_f(tmp_path, "vfield", step), ucont)
    assert np.array_equal(_f(tmp_path, "pfield", step), pavg)
    assert np.array_equal(_f(tmp_path, "ParticleCount", step), count)
    assert not _f(tmp_path, "nvfield", step).any()
    info = open(O.field_file_name("position", step, str(tmp_path)) + ".info").read()
    assert info == "-vecload_block_size 3\n"

    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.ReadRestart(step, str(tmp_path))
        assert g.n == n
        r = g.download_particles()
        assert np.array_equal(g.field(L.FIELD_UCAT), ucat)
        assert np.array_equal(g.field(L.FIELD_UCONT), ucont)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        r2 = g.download_particles()
    # PrepareLoadedSwarmForRelocation: no cell, UNINITIALIZED, PIDs 0..n-1, file order kept
    assert np.array_equal(r["pos"], d["pos"]) and np.array_equal(r["vel"], d["vel"])
    assert (r["cell"] == -1).all() and (r["status"] == L.UNINITIALIZED).all()
    assert np.array_equal(np.sort(r["pid"]), np.arange(n))
    o = np.argsort(r2["pid"])
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, r["pos"])
    np.testing.assert_array_equal(r2["cell"][o], cell)
    np.testing.assert_array_equal(r2["status"][o], st)


def test_entry_points_after_single_step_calls(swarm_cls, tmp_path):
    """Step calls that end between compactions leave the removed particles in
    the arrays; output files, stage-wise calls and a grid upload must see only
    the live particles."""
    _say("entry points after single-step calls")
    IM, JM, KM = 16, 16, 40
    coor = W.tube_grid(IM, JM, KM)
    pos = W.particles_uniform_tube(20000, 92)
    with swarm_cls(precision="fp64") as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.BoundarySystem_Create(channel(4, "constant_velocity", 1.0))
        g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        for _ in range(3):  # one call per step, no compaction yet
            g.AdvanceSimulation(0.08, 1, rk2=True, deposit_cells=True)
        n = g.n
        assert n == g.stats()["n_particles"] < 20000  # some left through the open end
        g.WriteOutput(3, str(tmp_path))
        d = g.download_particles()
        assert len(d["pid"]) == n and (d["status"] == L.ACTIVE_AND_LOCATED).all()
        assert np.array_equal(_f(tmp_path, "position", 3), d["pos"].ravel())
        # the step interpolated at the final positions: the stage-wise call
        # (which compacts first) gives the same velocities
        g.AdvanceSimulation(0.08, 2, rk2=True, deposit_cells=True)
        n2 = g.n
        g.InterpolateAllFieldsToSwarm()
        d2 = g.download_particles()
        assert len(d2["pid"]) == n2
        g.AdvanceSimulation(0.08, 1, rk2=True, deposit_cells=True)
        d3 = g.download_particles()
        g.InterpolateAllFieldsToSwarm()
        assert np.array_equal(g.download_particles()["vel"], d3["vel"])
        # a grid upload keeps the live particles only
        g.AdvanceSimulation(0.08, 1, rk2=True, deposit_cells=True)
        n3 = g.n
        g.upload_grid(IM, JM, KM, coor)
        assert g.n == n3
        g.compute_metrics()
        g.upload_ucat(np.zeros(((IM + 1) * (JM + 1) * (KM + 1), 3)))
        g.LocateAllParticlesInGrid()
        assert (g.download_particles()["status"] == L.ACTIVE_AND_LOCATED).all()


def test_restart_errors(swarm_cls, tmp_path):
    _say("restart errors")
    IM, JM, KM = 4, 4, 4
    with swarm_cls(precision="fp64") as g:
        g.upload_grid(IM, JM, KM, W.uniform_grid(IM, JM, KM))
        with pytest.raises(L.PicgpuError):  # nothing written yet
            g.ReadRestart(1, str(tmp_path))
        g.upload_ucat(np.zeros((5 * 5 * 5, 3)))
        g.upload_particles(np.full((10, 3), 0.5))
        g.WriteOutput(1, str(tmp_path))
        # a grid of another size does not take these files
        g.upload_grid(IM + 1, JM, KM, W.uniform_grid(IM + 1, JM, KM))
        with pytest.raises(L.PicgpuError):
            g.ReadRestart(1, str(tmp_path))
        g.upload_grid(IM, JM, KM, W.uniform_grid(IM, JM, KM))
        os.remove(O.field_file_name("velocity", 1, str(tmp_path)))  # optional file
        g.ReadRestart(1, str(tmp_path))
        assert g.n == 10
        assert (g.download_particles()["vel"] == 0).all()


def test_write_vtk(swarm_cls, tmp_path):
    _say("vtk output")
    IM, JM, KM = 12, 10, 14
    coor = W.perturbed_cube(IM, JM, KM, seed=23)
    with swarm_cls(precision="fp32") as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.BoundarySystem_Create(channel(4, "constant_velocity", 1.0))
        g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
        g.InitializeParticleBasicProperties(5000, volumetric=True)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        g.AdvanceSimulation(0.01, 2, deposit_cells=True)
        g.WriteVTK(2, str(tmp_path))
        d = g.download_particles()
        ucat, ucont = g.field(L.FIELD_UCAT), g.field(L.FIELD_UCONT)
        pavg, count = g.field(L.FIELD_P_AVERAGE), g.field(L.FIELD_PARTICLE_COUNT)
    for name, exp in (("Ucat", ucat.reshape(-1, 3)), ("Ucont", ucont.reshape(-1, 3)), ("P", pavg),
                      ("ParticleCount", count)):
        a, _ = O.read_vtk_xml(os.path.join(str(tmp_path), f"{name}_00002.vts"))
        assert np.array_equal(a["Position"], coor) and np.array_equal(a[name], exp), name
    a, _ = O.read_vtk_xml(os.path.join(str(tmp_path), "velocity_00002.vtp"))
    assert np.array_equal(a["Position"], d["pos"]) and np.array_equal(a["velocity"], d["vel"])
    assert np.array_equal(a["connectivity"], np.arange(len(d["pid"])))
