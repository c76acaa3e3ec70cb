"""GPU parity of the Eulerian field preparation against the CPU oracle.

Bar: metrics (Csi, Eta, Zet, Aj), the orientation, Ucont and Ucat after the
boundary handlers, SetInitialInteriorField and Contra2Cart bit-exact in both
precision modes (the preparation is FP64 in both); the particle loop's corner
velocities derived from the prepared Ucat bit-exact in FP64 mode and within
1e-5 of the field scale in FP32 mode; particle steps driven by the prepared
field bit-exact in FP64 mode.  At the 256 x 256 x 1024 grid: the cell systems
M u = q hold, the discrete closure holds and a second preparation step
reproduces the first bit for bit.
"""
import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W
from fdf_amd.boundary import BoundaryConfig, channel

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


GRIDS = {
    # 69 x 8 x 36 interior cells: two i tiles, two j tiles, three k chunks of the Contra2Cart kernel
    "perturbed": lambda: (70, 9, 37, W.perturbed_cube(70, 9, 37, seed=3001)),
    "tube": lambda: (16, 16, 40, W.tube_grid(16, 16, 40)),
    "mirrored": lambda: (12, 10, 14, W.perturbed_cube(12, 10, 14, seed=3002) * np.array([-1.0, 1.0, 1.0])),
}


def _parabolic_negz():
    cfg = BoundaryConfig()
    cfg.set_face("+Zeta", "INLET", "parabolic", vmax=2.0)
    cfg.set_face("-Zeta", "NOGRAD", "allcopy")
    return cfg


def _inlet_x():
    cfg = BoundaryConfig()
    cfg.set_face("-Xi", "INLET", "constant_velocity", vx=0.5, vy=0.05)
    cfg.set_face("+Xi", "NOGRAD", "allcopy")
    cfg.set_face("-Eta", "NOGRAD", "allcopy")
    return cfg


CONFIGS = {
    "channel_z": (lambda: channel(4, "constant_velocity", 1.0), 1, (0.0, 0.0, 1.0)),
    "parabolic_negz": (_parabolic_negz, 2, (0.0, 0.0, 2.0)),
    "inlet_x": (_inlet_x, 1, (0.5, 0.0, 0.0)),
}


def _oracle_prepare(oracle, IM, JM, KM, coor, cfg, finit, uin, steps):
    csi, eta, zet, aj, ori = oracle.metrics(IM, JM, KM, coor)
    n = len(csi)
    ucat, ucont = np.zeros((n, 3)), np.zeros((n, 3))
    oracle.bc_initialize(IM, JM, KM, cfg, csi, eta, zet, ucat, ucont)
    hist = [(ucat.copy(), ucont.copy())]
    for step in range(steps):
        oracle.set_eulerian_fields(IM, JM, KM, cfg, csi, eta, zet, step, finit, uin, ori, ucat, ucont)
        hist.append((ucat.copy(), ucont.copy()))
    return (csi, eta, zet, aj, ori), hist


def _eq(a, b):
    return np.array_equal(np.asarray(a).ravel(), np.asarray(b).ravel())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("grid", list(GRIDS))
def test_metrics_bit_exact(swarm_cls, oracle, grid, precision):
    _say("eulerian metrics", grid, precision)
    IM, JM, KM, coor = GRIDS[grid]()
    csi, eta, zet, aj, ori = oracle.metrics(IM, JM, KM, coor)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        o = g.compute_metrics()
        got = [g.field(f) for f in (L.FIELD_CSI, L.FIELD_ETA, L.FIELD_ZET, L.FIELD_AJ)]
        zero = [g.field(L.FIELD_UCAT), g.field(L.FIELD_UCONT)]
    assert o == ori == 1
    for a, b, name in zip(got, (csi, eta, zet, aj), ("csi", "eta", "zet", "aj")):
        assert _eq(a, b), name
    assert not any(z.any() for z in zero)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("config", list(CONFIGS))
@pytest.mark.parametrize("grid", list(GRIDS))
def test_set_eulerian_fields_bit_exact(swarm_cls, oracle, grid, config, precision):
    _say("eulerian fields", grid, config, precision)
    IM, JM, KM, coor = GRIDS[grid]()
    make, finit, uin = CONFIGS[config]
    cfg = make()
    _, hist = _oracle_prepare(oracle, IM, JM, KM, coor, cfg, finit, uin, 2)
    pos = np.random.default_rng(3).random((500, 3))  # a few particles so that the interpolation runs
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.BoundarySystem_Create(cfg)
        stages = [(g.field(L.FIELD_UCAT), g.field(L.FIELD_UCONT))]
        for step in range(2):
            g.SetEulerianFields(step, 0, finit, uin)
            stages.append((g.field(L.FIELD_UCAT), g.field(L.FIELD_UCONT)))
        g.upload_particles(pos)
        g.InterpolateAllFieldsToSwarm()
        corner = g.field(L.FIELD_CORNER_VELOCITY)
    for s, ((gu, gc), (ou, oc)) in enumerate(zip(stages, hist)):
        assert _eq(gu, ou), f"ucat after stage {s}"
        assert _eq(gc, oc), f"ucont after stage {s}"
    assert np.abs(hist[-1][0]).max() > 0
    ref = oracle.center_to_corner(IM, JM, KM, hist[-1][0])
    if precision == "fp64":
        assert _eq(corner, ref)
    else:
        assert np.abs(corner - ref.ravel()).max() <= FP32_TOL * np.abs(ref).max()


@pytest.mark.parametrize("grid", ["perturbed", "tube"])
def test_contra2cart_random_ucont(swarm_cls, oracle, grid):
    _say("contra2cart random", grid)
    IM, JM, KM, coor = GRIDS[grid]()
    csi, eta, zet, _, _ = oracle.metrics(IM, JM, KM, coor)
    ucont = np.random.default_rng(21).normal(size=csi.shape) * np.abs(csi).max()
    ref = np.zeros_like(ucont)
    oracle.contra2cart(IM, JM, KM, ucont, csi, eta, zet, ref)
    with swarm_cls(precision="fp64") as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.upload_ucont(ucont)
        g.Contra2Cart()
        got = g.field(L.FIELD_UCAT)
    assert _eq(got, ref)


def test_errors(swarm_cls, oracle):
    _say("eulerian errors")
    IM, JM, KM = 4, 4, 4
    base = W.uniform_grid(IM, JM, KM).reshape(KM + 1, JM + 1, IM + 1, 3)
    with swarm_cls(precision="fp64") as g:
        g.upload_grid(IM, JM, KM, base.reshape(-1, 3))
        with pytest.raises(L.PicgpuError) as e:
            g.Contra2Cart()  # before compute_metrics
        assert e.value.code == -2
        g.compute_metrics()
        with pytest.raises(L.PicgpuError) as e:
            g.SetEulerianFields(0)  # before BoundarySystem_Create
        assert e.value.code == -2
        bad = BoundaryConfig()
        bad.set_face("+Zeta", "OUTLET", "conservation")
        with pytest.raises(L.PicgpuError) as e:
            g.BoundarySystem_Create(bad)
        assert e.value.code == -6
        # zero-volume cells: the i = 2 node plane collapsed onto i = 1
        flat = base.copy()
        flat[:, :, 2, 0] = flat[:, :, 1, 0]
        g.upload_grid(IM, JM, KM, flat.reshape(-1, 3))
        with pytest.raises(L.PicgpuError) as e:
            g.compute_metrics()
        assert e.value.code == -7
        with pytest.raises(RuntimeError):
            oracle.metrics(IM, JM, KM, flat.reshape(-1, 3))
        # folded grid: Jacobians of both signs
        fold = base.copy()
        fold[:, :, 2:, 0] = 1.75 - fold[:, :, 2:, 0]
        g.upload_grid(IM, JM, KM, fold.reshape(-1, 3))
        with pytest.raises(L.PicgpuError) as e:
            g.compute_metrics()
        assert e.value.code == -8
        # a tiny grid: the Jacobian passes (det 8e-12), the Contra2Cart matrices do not (det 6.4e-23)
        tiny = base * 2e-4 * 4
        g.upload_grid(IM, JM, KM, tiny.reshape(-1, 3))
        g.compute_metrics()
        g.BoundarySystem_Create(channel())
        with pytest.raises(L.PicgpuError) as e:
            g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
        assert e.value.code == -7


def _oracle_steps(oracle, IM, JM, KM, coor, ucat, pos, dt, nsteps):
    corner = oracle.center_to_corner(IM, JM, KM, ucat)
    n = len(pos)
    s = {"pos": pos.copy(), "vel": np.zeros((n, 3)), "cell": np.full((n, 3), -1, np.int64),
         "weight": np.zeros((n, 3)), "status": np.full(n, 4, np.int64), "pid": np.arange(n, dtype=np.int64),
         "P": np.full(n, 101325.0)}
    c, w, st, _ = oracle.locate(IM, JM, KM, coor, s["pos"], status=s["status"])
    s.update(cell=c, weight=w, status=st)
    s["vel"] = oracle.interp(IM, JM, KM, corner, c, w)
    for _ in range(nsteps):
        oracle.step(IM, JM, KM, coor, corner, s, dt)
    return s


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_prepared_field_drives_particle_steps(swarm_cls, oracle, precision):
    """The tube with a constant-velocity inlet: the particle loop advects on
    the Ucat the preparation derived, exactly as the oracle does."""
    _say("eulerian -> particle steps", precision)
    IM, JM, KM = 16, 16, 48
    coor = W.tube_grid(IM, JM, KM)
    cfg = channel(4, "constant_velocity", 1.0)
    _, hist = _oracle_prepare(oracle, IM, JM, KM, coor, cfg, 1, (0.0, 0.0, 1.0), 1)
    ucat = hist[-1][0]
    pos = W.particles_uniform_tube(20000, 77)
    dt, nsteps = 0.05, 6
    s = _oracle_steps(oracle, IM, JM, KM, coor, ucat, pos, dt, nsteps)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.compute_metrics()
        g.BoundarySystem_Create(cfg)
        g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        g.AdvanceSimulation(dt, nsteps, deposit_cells=True)
        d = g.download_particles()
    o = np.argsort(d["pid"])
    out = {k: v[o] for k, v in d.items()}
    # about two thirds of the walks from cell (0,0,0) cycle on the O-grid and
    # end LOST in the first locate (the reference's start rule); the rest stay
    assert len(s["pid"]) > 5000
    if precision == "fp64":
        for k in ("pid", "cell", "status", "pos", "vel", "weight"):
            assert np.array_equal(out[k], s[k]), k
    else:
        common, ia, ib = np.intersect1d(out["pid"], s["pid"], return_indices=True)
        assert len(common) >= 0.99 * len(s["pid"])
        assert np.abs(out["pos"][ia] - s["pos"][ib]).max() < 1e-4
        assert (out["cell"][ia] == s["cell"][ib]).all(1).mean() > 0.99


def test_full_size_properties(swarm_cls):
    """256 x 256 x 1024 perturbed cube: sampled cell systems, closure, idempotence."""
    IM, JM, KM = 256, 256, 1024
    _say("eulerian full size", IM, JM, KM)
    coor = W.perturbed_cube(IM, JM, KM, seed=3005)
    with swarm_cls(precision="fp32") as g:
        g.upload_grid(IM, JM, KM, coor)
        del coor
        g.compute_metrics()
        g.BoundarySystem_Create(channel(4, "constant_velocity", 1.0))
        g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
        u0 = g.field(L.FIELD_UCAT)
        g.SetEulerianFields(1, 0, 1, (0.0, 0.0, 1.0))
        u1 = g.field(L.FIELD_UCAT)
        assert np.array_equal(u0, u1)
        del u1
        shp = (KM + 1, JM + 1, IM + 1, 3)
        u = u0.reshape(shp)
        c, e, z, q = (g.field(f).reshape(shp) for f in (L.FIELD_CSI, L.FIELD_ETA, L.FIELD_ZET, L.FIELD_UCONT))
    rng = np.random.default_rng(8)
    m = 200_000
    i, j, k = (rng.integers(1, n, m) for n in (IM, JM, KM))  # interior cells 1..IM-1
    M = np.stack([0.5 * (c[k, j, i - 1] + c[k, j, i]), 0.5 * (e[k, j - 1, i] + e[k, j, i]),
                  0.5 * (z[k - 1, j, i] + z[k, j, i])], -2)
    qq = np.stack([0.5 * (q[k, j, i - 1, 0] + q[k, j, i, 0]), 0.5 * (q[k, j - 1, i, 1] + q[k, j, i, 1]),
                   0.5 * (q[k - 1, j, i, 2] + q[k, j, i, 2])], -1)
    res = float(np.abs(np.einsum("nab,nb->na", M, u[k, j, i]) - qq).max())
    qmax = float(np.abs(qq).max())
    _say("max |M u - q|", res, "max |q|", qmax)
    assert res <= 1e-12 * qmax
    # closure on sampled cells away from the copied boundary planes
    ci, cj, ck = (rng.integers(0, n - 1, m) for n in (IM, JM, KM))
    clo = (c[ck + 1, cj + 1, ci + 1] - c[ck + 1, cj + 1, ci]) + (e[ck + 1, cj + 1, ci + 1] - e[ck + 1, cj, ci + 1]) + \
        (z[ck + 1, cj + 1, ci + 1] - z[ck, cj + 1, ci + 1])
    # rounding of the O(1) coordinates in the O(h) edge differences bounds the residual
    clo_max, cmax = float(np.abs(clo).max()), float(np.abs(c).max())
    _say("max closure residual", clo_max, "max |csi|", cmax)
    assert clo_max <= 1e-12 * cmax
    # interior flow is along +z, mostly
    assert np.median(u[1:-1, 1:-1, 1:-1, 2]) > 0.5


def test_step_flag_prepares_fields_every_step(swarm_cls):
    """PICGPU_STEP_FIELDS runs the preparation before each step; the result
    equals preparing by hand before each single step (FP64, per PID)."""
    _say("eulerian step flag")
    IM, JM, KM = 16, 16, 40
    coor = W.tube_grid(IM, JM, KM)
    cfg = channel(4, "constant_velocity", 1.0)
    pos = W.particles_uniform_tube(20000, 78)
    outs = []
    for by_hand in (False, True):
        with swarm_cls(precision="fp64") as g:
            g.upload_grid(IM, JM, KM, coor)
            g.compute_metrics()
            g.BoundarySystem_Create(cfg)
            g.SetEulerianFields(0, 0, 1, (0.0, 0.0, 1.0))
            g.upload_particles(pos)
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            if by_hand:
                for s in range(3):
                    g.SetEulerianFields(s + 1, 0)
                    g.AdvanceSimulation(0.05, 1, rk2=True, deposit_faces=True)
            else:
                g.AdvanceSimulation(0.05, 3, rk2=True, deposit_faces=True, fields=True)
            d = g.download_particles()
            o = np.argsort(d["pid"])
            outs.append(({k: v[o] for k, v in d.items()}, g.field(L.FIELD_UCAT), g.field(L.FIELD_KFACE)))
    (a, ua, fa), (b, ub, fb) = outs
    assert len(a["pid"]) > 5000
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(ua, ub) and np.array_equal(fa, fb)
