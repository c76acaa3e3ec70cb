"""The gfx950 path through the C-ABI against the committed golden fixtures
(tests/golden/, written by the CPU oracle).

FP64 mode: cells, statuses, walk iterations, weights, velocities, positions
and deposits bit-exact.  FP32 mode: cells, statuses and walk iterations
exact; weights, velocities and deposits within 1e-5 of the FP64 values
(relative to the field scale); positions after steps within 1e-4 (the FP32
velocities drift the trajectories)."""
import os

import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FP32_TOL = 1e-5


def _load(name):
    return np.load(os.path.join(GOLDEN, name))


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _by_pid(d):
    o = np.argsort(d["pid"])
    return {k: v[o] for k, v in d.items()}


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


def _check_locate(out, cell, st, steps, w, walk_steps, precision):
    np.testing.assert_array_equal(out["cell"], cell)
    np.testing.assert_array_equal(out["status"], st)
    np.testing.assert_array_equal(walk_steps, steps)
    ok = st == 1
    if precision == "fp64":
        assert np.array_equal(out["weight"][ok], w[ok])
    else:
        assert np.abs(out["weight"][ok] - w[ok]).max() < FP32_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_golden_walk_uniform_cfg1(swarm_cls, precision):
    _say("golden_walk_uniform_cfg1", precision)
    g = _load("walk_uniform_cfg1.npz")
    IM, JM, KM = (int(v) for v in g["dims"])
    with swarm_cls(precision=precision, collect_walk_stats=True) as s:
        s.upload_grid(IM, JM, KM, W.uniform_grid(IM, JM, KM))
        s.upload_particles(g["pos"])
        s.LocateAllParticlesInGrid()
        out = s.download_particles()
        steps = s.walk_steps()
        hist = s.stats()["walk_hist"]
    _check_locate(out, g["cell"], g["status"], g["steps"], g["weight"], steps, precision)
    np.testing.assert_array_equal(hist, np.bincount(np.minimum(g["steps"], 63), minlength=64))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_golden_walk_warped_small(swarm_cls, precision):
    _say("golden_walk_warped_small", precision)
    g = _load("walk_warped_small.npz")
    IM, JM, KM = (int(v) for v in g["dims"])
    n = len(g["pos"])
    with swarm_cls(precision=precision, collect_walk_stats=True) as s:
        s.upload_grid(IM, JM, KM, g["coor"])
        s.upload_particles(g["pos"])
        s.LocateAllParticlesInGrid()
        out, steps = s.download_particles(), s.walk_steps()
        _check_locate(out, g["cell"], g["status"], g["steps"], g["weight"], steps, precision)
        s.upload_particles(g["pos"], cell=g["start"], status=np.zeros(n, np.int64))
        s.LocateAllParticlesInGrid()
        out, steps = s.download_particles(), s.walk_steps()
        _check_locate(out, g["cell2"], g["status2"], g["steps2"], g["weight2"], steps, precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_golden_interp_and_deposit(swarm_cls, precision):
    _say("golden_interp_and_deposit", precision)
    g = _load("walk_warped_small.npz")
    f = _load("interp_deposit_warped_small.npz")
    IM, JM, KM = (int(v) for v in g["dims"])
    with swarm_cls(precision=precision) as s:
        s.upload_grid(IM, JM, KM, g["coor"])
        s.upload_ucat(f["ucat"])
        s.upload_particles(g["pos"], cell=g["cell"], weight=g["weight"], status=g["status"], pid=f["pid"])
        s.InterpolateAllFieldsToSwarm()
        out = _by_pid(s.download_particles())
        corner = s.field(L.FIELD_CORNER_VELOCITY)
        s.ScatterAllParticleFieldsToEulerFields()
        count, pavg = s.field(L.FIELD_PARTICLE_COUNT), s.field(L.FIELD_P_AVERAGE)
        s.DepositParticlesToFaces()
        faces = [s.field(L.FIELD_IFACE), s.field(L.FIELD_JFACE), s.field(L.FIELD_KFACE)]
    o = np.argsort(f["pid"])
    vel = f["vel"][o]
    assert np.array_equal(count, f["count"])
    assert np.array_equal(pavg, f["pavg"])
    scale = np.abs(f["ucat"]).max()
    if precision == "fp64":
        assert np.array_equal(corner, f["corner"].ravel())
        assert np.array_equal(out["vel"], vel)
        for a, b in zip(faces, (f["fi"], f["fj"], f["fk"])):
            assert np.array_equal(a, b)
    else:
        assert np.abs(corner - f["corner"].ravel()).max() <= FP32_TOL * scale
        assert np.abs(out["vel"] - vel).max() <= FP32_TOL * scale
        for a, b in zip(faces, (f["fi"], f["fj"], f["fk"])):
            a4, b4 = a.reshape(-1, 4), b.reshape(-1, 4)
            assert np.abs(a4[:, 0] - b4[:, 0]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 0]).max())
            assert np.abs(a4[:, 1:] - b4[:, 1:]).max() <= FP32_TOL * scale * max(1.0, np.abs(b4[:, 0]).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("path", ["fused", "staged"])
def test_golden_rk2_steps(swarm_cls, precision, path):
    _say("golden_rk2_steps", precision, path)
    g = _load("walk_warped_small.npz")
    f = _load("interp_deposit_warped_small.npz")
    s0 = _load("steps_rk2_warped_small.npz")
    IM, JM, KM = (int(v) for v in g["dims"])
    with swarm_cls(precision=precision) as s:
        s.upload_grid(IM, JM, KM, g["coor"])
        s.upload_ucat(f["ucat"])
        s.upload_particles(s0["in_pos"], vel=s0["in_vel"], cell=s0["in_cell"], weight=s0["in_weight"],
                           status=s0["in_status"], pid=s0["in_pid"], P=s0["in_P"])
        s.AdvanceSimulation(float(s0["dt"]), 2, rk2=True, deposit_cells=True, deposit_faces=True,
                            staged=path == "staged")
        s.walk_stats(True)
        n_before = s.n
        s.AdvanceSimulation(float(s0["dt"]), 1, rk2=True, deposit_cells=True, deposit_faces=True,
                            staged=path == "staged")
        st = s.stats()
        out = _by_pid(s.download_particles())
    if path == "fused":
        # the histogram of the last step covers every particle it stepped
        h = st["walk_hist"]
        assert h.sum() == n_before
        # the last bin holds the walks of 63 or more iterations (cycling walks
        # count MAX_TRAVERSAL = 1000, at most two walks per RK2 step)
        below = (h[:63] * np.arange(63)).sum()
        assert below + 63 * h[63] <= st["walk_iterations"] <= below + 2000 * h[63]
    o = np.argsort(s0["out_pid"])
    exp = {k[4:]: s0[k][o] for k in s0.files if k.startswith("out_")}
    if precision == "fp64":
        for k in ("pid", "cell", "status", "pos", "vel", "weight"):
            assert np.array_equal(out[k], exp[k]), k
    else:
        common, ia, ib = np.intersect1d(out["pid"], exp["pid"], return_indices=True)
        assert len(common) >= 0.99 * len(exp["pid"])
        assert np.abs(out["pos"][ia] - exp["pos"][ib]).max() < 1e-4
        same = (out["cell"][ia] == exp["cell"][ib]).all(1)
        assert same.mean() > 0.99


def test_golden_mlp(swarm_cls):
    _say("golden_mlp")
    m = _load("mlp_small.npz")
    with swarm_cls(precision="fp32") as s:
        s.upload_grid(4, 4, 4, W.uniform_grid(4, 4, 4))
        s.mlp_set_params(m["params"].astype(np.float32))
        y = s.mlp_forward(m["x"].astype(np.float32))
        loss, grad = s.mlp_loss_grad(m["x"].astype(np.float32), m["t"].astype(np.float32))
    err = np.abs(y - m["y"])
    assert np.median(err) < 1e-5 and err.max() < 5e-3
    assert abs(loss - float(m["loss"])) <= 1e-4 * abs(float(m["loss"]))
    assert np.linalg.norm(grad - m["grad"]) / np.linalg.norm(m["grad"]) < 2e-3
