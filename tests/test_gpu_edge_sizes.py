"""Edge sizes on the HIP path: an empty swarm, a single particle, and a grid
smaller than one 8x8x4 tile, through the fused and the staged step, against
the CPU oracle (FP64 bit-exact; FP32 cells exact and fields within 1e-5).
A swarm that leaves the grid entirely is config 1's case in
test_gpu_parity.test_full_steps_against_oracle."""
import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
FACE_FIELDS = (L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


def _by_pid(d):
    o = np.argsort(d["pid"])
    return {k: v[o] for k, v in d.items()}


def _oracle_steps(oracle, IM, JM, KM, coor, ucat, pos, dt, nsteps):
    """The reference's setup (locate from (0,0,0), interpolate) and nsteps
    RK2 steps on the oracle; returns the state and the corner velocities."""
    corner = oracle.center_to_corner(IM, JM, KM, ucat)
    n = pos.shape[0]
    s = {"pos": pos.copy(), "vel": np.zeros((n, 3)), "cell": np.full((n, 3), -1, np.int64),
         "weight": np.zeros((n, 3)), "status": np.full(n, 4, np.int64), "pid": np.arange(n, dtype=np.int64),
         "P": np.full(n, 101325.0)}
    c, w, st, _ = oracle.locate(IM, JM, KM, coor, s["pos"], status=s["status"])
    s.update(cell=c, weight=w, status=st)
    s["vel"] = oracle.interp(IM, JM, KM, corner, c, w)
    for _ in range(nsteps):
        oracle.step(IM, JM, KM, coor, corner, s, dt, rk2=True)
    return s, corner


def _gpu_steps(swarm_cls, precision, IM, JM, KM, coor, ucat, pos, dt, nsteps, staged, faces_of_start=False):
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_ucat(ucat)
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        g.AdvanceSimulation(dt, nsteps, rk2=True, deposit_cells=True, deposit_faces=True, staged=staged,
                            faces_of_start=faces_of_start)
        out = _by_pid(g.download_particles())
        return (out, g.n, g.stats(), g.field(L.FIELD_PARTICLE_COUNT), g.field(L.FIELD_P_AVERAGE),
                [g.field(f) for f in FACE_FIELDS])


def _grid(IM, JM, KM, seed):
    coor = W.perturbed_cube(IM, JM, KM, seed=seed, amp=0.15)
    ucat = W.ucat_random_fourier(coor, IM, JM, KM, seed=seed + 1)
    dt = 0.5 / (np.abs(ucat).max() * max(IM, JM, KM))
    return coor, ucat, dt


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("path,faces_of_start", [("fused", False), ("fused", True), ("staged", False)])
def test_empty_swarm(swarm_cls, oracle, precision, path, faces_of_start):
    """No particles: locate, interpolation and three RK2 steps with cell and
    face deposition succeed and deposit nothing (zero counts, P averages and
    faces, as the oracle gives for an empty swarm).  faces_of_start (the
    in-step deposition's faces) is an option of the fused step only."""
    _say("empty swarm", precision, path, faces_of_start)
    IM, JM, KM = 12, 10, 9
    coor, ucat, dt = _grid(IM, JM, KM, 41)
    out, n, st, cnt, pav, faces = _gpu_steps(swarm_cls, precision, IM, JM, KM, coor, ucat, np.zeros((0, 3)), dt,
                                             3, path == "staged", faces_of_start)
    assert n == 0 and st["particle_steps"] == 0 and st["located"] == 0
    assert all(v.shape[0] == 0 for v in out.values())
    e = np.zeros((0, 3), np.int64)
    ref_cnt = oracle.count_cells(IM, JM, KM, e)
    assert np.array_equal(cnt, ref_cnt) and not cnt.any()
    assert np.array_equal(pav, oracle.scatter_P(IM, JM, KM, e, np.zeros(0), np.zeros(0, np.int64), ref_cnt))
    ref = oracle.deposit_faces(IM, JM, KM, e, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int64))
    for a, b in zip(faces, ref):
        assert np.array_equal(a, b) and not a.any()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("path", ["fused", "staged"])
@pytest.mark.parametrize("case", ["one_particle", "grid_below_one_tile"])
def test_small_sizes_against_oracle(swarm_cls, oracle, precision, path, case):
    """One particle in a 12x10x9 grid (one lane of one tile busy), and 500
    particles in a 3x2x5 grid (fewer cells than one 8x8x4 tile in every
    direction but k, which has a seam between two tiles; about a tenth of
    the particles leave through the walls): four RK2 steps."""
    _say("small sizes", case, precision, path)
    if case == "one_particle":
        IM, JM, KM = 12, 10, 9
        coor, ucat, dt = _grid(IM, JM, KM, 43)
        pos = np.array([[0.43, 0.57, 0.31]])
    else:
        IM, JM, KM = 3, 2, 5
        coor, ucat, dt = _grid(IM, JM, KM, 45)
        pos = 0.02 + 0.96 * np.random.default_rng(46).random((500, 3))
    nsteps = 4
    s, corner = _oracle_steps(oracle, IM, JM, KM, coor, ucat, pos, dt, nsteps)
    assert len(s["pid"]) == (1 if case == "one_particle" else 444)
    out, n, st, cnt, pav, faces = _gpu_steps(swarm_cls, precision, IM, JM, KM, coor, ucat, pos, dt, nsteps,
                                             path == "staged")
    assert n == len(s["pid"])
    np.testing.assert_array_equal(out["pid"], s["pid"])
    np.testing.assert_array_equal(out["status"], s["status"])
    assert np.array_equal(cnt, oracle.count_cells(IM, JM, KM, out["cell"]))
    assert set(np.unique(pav)) <= {0.0, 101325.0} and ((pav > 0) == (cnt > 0)).all()
    ref = oracle.deposit_faces(IM, JM, KM, out["cell"], out["weight"], out["vel"], out["pid"])
    if precision == "fp64":
        for k in ("cell", "pos", "vel", "weight"):
            assert np.array_equal(out[k], s[k]), k
        for a, b in zip(faces, ref):
            assert np.array_equal(a, b)
    else:
        # ownership exact on the variant's own positions, fields within the tolerance
        np.testing.assert_array_equal(out["cell"], s["cell"])
        assert np.abs(out["pos"] - s["pos"]).max() < 1e-6
        c2, w2, _, _ = oracle.locate(IM, JM, KM, coor, out["pos"], cell=out["cell"],
                                     status=np.zeros(len(out["pid"]), np.int64))
        np.testing.assert_array_equal(c2, out["cell"])
        assert np.abs(w2 - out["weight"]).max() < FP32_TOL
        v2 = oracle.interp(IM, JM, KM, corner, c2, w2)
        assert np.abs(v2 - out["vel"]).max() <= FP32_TOL * np.abs(ucat).max()
        for a, b in zip(faces, ref):
            a4, b4 = a.reshape(-1, 4), b.reshape(-1, 4)
            assert np.abs(a4[:, 0] - b4[:, 0]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 0]).max())
            assert np.abs(a4[:, 1:] - b4[:, 1:]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 1:]).max())


def test_sizes_beyond_32_bit_indices_are_rejected(swarm_cls):
    """Cell numbers and particle numbers are 32-bit on the device: a grid of
    2^31 - 2 cells or more, and 2^31 - 1 particles or more, are refused with
    PICGPU_E_UNSUPPORTED before anything is read from the caller's arrays
    (here one-element stand-ins)."""
    _say("size limits")
    one = np.zeros(3)
    with swarm_cls(precision="fp64") as g:
        ptr = one.ctypes.data_as(L.C.c_void_p)
        assert g.lib.picgpu_upload_grid(g.h, 2048, 1024, 1024, ptr) == -6
        assert g.lib.picgpu_upload_grid(g.h, 1291, 1291, 1291, ptr) == -6  # 2,151,685,171 cells
        g.upload_grid(4, 4, 4, W.uniform_grid(4, 4, 4))
        cells = np.zeros(3, np.int64)
        st = np.zeros(1, np.int64)
        assert g.lib.picgpu_upload_particles(g.h, (1 << 31) - 1, ptr, None, cells.ctypes.data_as(L.C.c_void_p), None,
                                             st.ctypes.data_as(L.C.c_void_p), st.ctypes.data_as(L.C.c_void_p),
                                             None) == -6
        # the context stays usable
        g.upload_particles(np.array([[0.3, 0.4, 0.55]]))
        g.LocateAllParticlesInGrid()
        assert g.n == 1 and (g.download_particles()["cell"] == [[1, 1, 2]]).all()
