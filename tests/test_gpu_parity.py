"""GPU parity: the gfx950 path through the C-ABI against the CPU oracle.

Bar: owning cells, statuses and walk iterations bit-exact in both precision
modes; FP64 mode weights, velocities, positions and deposits bit-exact; FP32
mode weights/velocities/deposits within 1e-5 (relative to the field scale)
of the FP64 oracle.  Particles are compared per PID (the device keeps them
sorted by cell).
"""
import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


def _by_pid(d):
    o = np.argsort(d["pid"])
    return {k: v[o] for k, v in d.items()}


GRIDS = {
    "cube": lambda: (20, 20, 280, W.uniform_grid(20, 20, 280)),
    "warped": lambda: (16, 16, 32, W.perturbed_cube(16, 16, 32, seed=11)),
    "warped_strong": lambda: (8, 8, 12, W.perturbed_cube(8, 8, 12, seed=12, amp=0.35)),
    "tube": lambda: (16, 16, 48, W.tube_grid(16, 16, 48)),
}


def _points(name, n, seed):
    if name == "tube":
        return W.particles_uniform_tube(n, seed)
    p = np.random.default_rng(seed).random((n, 3))
    p[: n // 50] = p[: n // 50] * 1.4 - 0.2  # some outside the domain
    return p


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("grid", list(GRIDS))
def test_locate_from_scratch(swarm_cls, oracle, grid, precision):
    _say("locate_from_scratch", grid, precision)
    IM, JM, KM, coor = GRIDS[grid]()
    pos = _points(grid, 20000, 7)
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos)
    with swarm_cls(precision=precision, collect_walk_stats=True) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        gsteps = g.walk_steps()
        d = g.download_particles()
        s = g.stats()
    o = np.argsort(d["pid"])
    np.testing.assert_array_equal(d["status"][o], st)
    np.testing.assert_array_equal(d["cell"][o], cell)
    np.testing.assert_array_equal(gsteps[o], steps)
    ok = st == 1
    if precision == "fp64":
        assert np.array_equal(d["weight"][o][ok].view(np.int64), w[ok].view(np.int64))
    else:
        assert np.abs(d["weight"][o][ok] - w[ok]).max() < FP32_TOL
        _say("fp64 fallbacks", s["fp64_fallbacks"], "of", 6 * s["walk_iterations"], "faces")
    assert s["located"] == ok.sum() and s["lost"] == (~ok).sum()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_locate_from_previous_cell_and_max_traversal(swarm_cls, oracle, precision):
    _say("locate_prev_cell", precision)
    IM, JM, KM = 2, 3, 1100
    coor = W.uniform_grid(IM, JM, KM)
    rng = np.random.default_rng(1)
    pos = rng.random((5000, 3)) * [0.98, 0.98, 1.0] + [0.01, 0.01, 0.0]
    prev = np.floor(pos * [IM, JM, KM]).astype(np.int64) + rng.integers(-3, 4, (5000, 3))
    prev[:2500] = -1  # half start from (0,0,0): k >= 1000 is LOST by the cap
    status = np.zeros(5000, np.int64)
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos, cell=prev, status=status)
    assert (st == 3).sum() > 100
    # prev cells outside the grid fall back to (0,0,0) in both
    with swarm_cls(precision=precision, collect_walk_stats=True) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_particles(pos, cell=prev, status=status)
        g.LocateAllParticlesInGrid()
        gsteps = g.walk_steps()
        raw = g.download_particles()
    o = np.argsort(raw["pid"])
    d = {k: v[o] for k, v in raw.items()}
    np.testing.assert_array_equal(d["status"], st)
    np.testing.assert_array_equal(d["cell"], cell)
    np.testing.assert_array_equal(gsteps[o], steps)
    ok = st == 1
    if precision == "fp64":
        assert np.array_equal(d["weight"][ok], w[ok])


def test_degenerate_grid_is_an_error(swarm_cls):
    _say("degenerate")
    coor = W.uniform_grid(1, 1, 1)
    coor[:] = 0.0
    with swarm_cls() as g:
        g.upload_grid(1, 1, 1, coor)
        g.upload_particles(np.zeros((1, 3)))
        with pytest.raises(L.PicgpuError) as e:
            g.LocateAllParticlesInGrid()
        assert e.value.code == -3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("grid", ["warped", "tube"])
def test_interpolate(swarm_cls, oracle, grid, precision):
    _say("interpolate", grid, precision)
    IM, JM, KM, coor = GRIDS[grid]()
    ucat = (W.ucat_poiseuille(coor, IM, JM, KM, seed=5) if grid == "tube"
            else W.ucat_taylor_green(coor, IM, JM, KM))
    pos = _points(grid, 20000, 3)
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    corner = oracle.center_to_corner(IM, JM, KM, ucat)
    vel = oracle.interp(IM, JM, KM, corner, cell, w)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_ucat(ucat)
        g.upload_particles(pos, cell=cell, weight=w, status=st)
        g.InterpolateAllFieldsToSwarm()
        gc = g.field(L.FIELD_CORNER_VELOCITY)
        d = _by_pid(g.download_particles())
    scale = np.abs(ucat).max()
    if precision == "fp64":
        assert np.array_equal(gc, corner.ravel())
        assert np.array_equal(d["vel"], vel)
    else:
        assert np.abs(gc - corner.ravel()).max() <= FP32_TOL * scale
        assert np.abs(d["vel"] - vel).max() <= FP32_TOL * scale


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_interpolation_error_diagnostic(swarm_cls, oracle, precision):
    """The reference's known-answer diagnostic LOG_INTERPOLATION_ERROR
    (logging.c:533-580 with AnalyticalSolution.c:932-962) on the HIP path:
    the L2 error, in percent, of the interpolated particle velocities against
    (sin x, sin y, sin z) when the cell field is that function at the
    centroids (finit 1), and against (0, 0, 1) for the uniform field
    (finit 0).  It equals the oracle's (FP64 bit for bit, FP32 within 1e-4
    percent); the uniform field's error is zero; and on uniform grids it
    falls by 2^1.5 per halving of h, the order of centre->corner + trilinear
    with the boundary's one-sided corner averages."""
    from oracle.oracle import interpolation_error_percent

    _say("interpolation error diagnostic", precision)
    side = 2.0
    errs = {}
    for finit, m, grid in ((1, 16, "uniform"), (1, 32, "uniform"), (1, 16, "perturbed"), (0, 16, "perturbed")):
        coor = (W.perturbed_cube(m, m, m, seed=7) if grid == "perturbed" else W.uniform_grid(m, m, m)) * side
        ucat = W.ucat_sine(coor, m, m, m) if finit == 1 else W.ucat_uniform(m, m, m)
        pos = side * np.random.default_rng(m + finit).random((20000, 3))
        cell, w, st, _ = oracle.locate(m, m, m, coor, pos)
        vo = oracle.interp(m, m, m, oracle.center_to_corner(m, m, m, ucat), cell, w)
        with swarm_cls(precision=precision) as g:
            g.upload_grid(m, m, m, coor)
            g.upload_ucat(ucat)
            g.upload_particles(pos)
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            ed = g.LOG_INTERPOLATION_ERROR(finit)  # the library's own, over the whole swarm
            out = _by_pid(g.download_particles())
        # the device reduction equals the formula over every particle (LOST
        # ones included, with the zero velocity interpolation gives them)
        ea = interpolation_error_percent(out["pos"], out["vel"], finit)
        assert abs(ed - ea) <= 1e-12 + 1e-10 * ea
        # walks from cell (0,0,0) that cycle end LOST (the reference's start
        # rule); the comparisons below are taken over the located particles
        ok = st == 1
        assert ok.mean() > 0.95
        np.testing.assert_array_equal(out["status"], st)
        np.testing.assert_array_equal(out["cell"][ok], cell[ok])
        eg = interpolation_error_percent(out["pos"][ok], out["vel"][ok], finit)
        eo = interpolation_error_percent(pos[ok], vo[ok], finit)
        _say("finit", finit, grid, m, "error %", eg, "oracle", eo)
        if precision == "fp64":
            assert np.array_equal(out["vel"][ok], vo[ok]) and eg == eo
        else:
            assert abs(eg - eo) <= 1e-4
        errs[(finit, m, grid)] = eg
    assert errs[(0, 16, "perturbed")] <= (0.0 if precision == "fp64" else 1e-4)
    assert 2.6 < errs[(1, 16, "uniform")] / errs[(1, 32, "uniform")] < 3.0
    with swarm_cls(precision=precision) as g:
        g.upload_grid(4, 4, 4, W.uniform_grid(4, 4, 4))
        g.upload_particles(np.zeros((0, 3)))
        assert np.isnan(g.LOG_INTERPOLATION_ERROR(1))  # 0/0, as the reference's norms give
        with pytest.raises(L.PicgpuError) as e:
            g.LOG_INTERPOLATION_ERROR(2)  # the reference defines finit 0 and 1 only
        assert e.value.code == -1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_advect_and_removal(swarm_cls, oracle, precision):
    _say("advect_remove", precision)
    IM, JM, KM, coor = GRIDS["warped"]()
    rng = np.random.default_rng(4)
    pos = rng.random((10000, 3))
    vel = rng.normal(0, 0.1, (10000, 3))
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_particles(pos, vel=vel, cell=cell, weight=w, status=st)
        g.ResetAllParticleStatuses()
        g.UpdateAllParticlePositions(0.37)
        d = _by_pid(g.download_particles())
        vv = vel.astype(np.float32).astype(np.float64) if precision == "fp32" else vel
        exp = pos.copy()
        exp += vv * 0.37
        assert np.array_equal(d["pos"], exp)
        assert (d["status"] == np.where(st == 1, 0, st)).all()
        g.LocateAllParticlesInGrid()
        n0 = g.n
        rl = g.CheckAndRemoveLostParticles()
        ro = g.CheckAndRemoveOutOfBoundsParticles()
        d2 = _by_pid(g.download_particles())
    c2, w2, st2, _ = oracle.locate(IM, JM, KM, coor, exp, cell=cell, status=np.zeros(10000, np.int64))
    lo, hi = oracle.bbox(IM, JM, KM, coor)
    inbox = ((exp >= lo) & (exp <= hi)).all(1)
    keep = (st2 != 3) & inbox
    assert rl == (st2 == 3).sum() and ro == ((st2 != 3) & ~inbox).sum()
    assert n0 - rl - ro == keep.sum() == len(d2["pid"])
    np.testing.assert_array_equal(d2["pid"], np.arange(10000)[keep])
    np.testing.assert_array_equal(d2["cell"], c2[keep])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_scatter_and_faces(swarm_cls, oracle, precision):
    _say("scatter_faces", precision)
    IM, JM, KM, coor = GRIDS["warped"]()
    rng = np.random.default_rng(6)
    pos = rng.random((30000, 3))
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    vel = rng.normal(0, 1.0, (30000, 3))
    P = 101325.0 + rng.normal(0, 10.0, 30000)
    pid = rng.permutation(30000).astype(np.int64) * 3 + 11
    cnt = oracle.count_cells(IM, JM, KM, cell)
    pav = oracle.scatter_P(IM, JM, KM, cell, P, pid, cnt)
    fi, fj, fk = oracle.deposit_faces(IM, JM, KM, cell, w, vel, pid)
    vp = 2.0e-3
    vf = [oracle.face_volfrac(IM, JM, KM, coor, d, f, vp) for d, f in enumerate((fi, fj, fk))]
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_particles(pos, vel=vel, cell=cell, weight=w, status=st, pid=pid, P=P)
        g.ScatterAllParticleFieldsToEulerFields()
        g.DepositParticlesToFaces(particle_volume=vp)
        gcnt = g.field(L.FIELD_PARTICLE_COUNT)
        gpav = g.field(L.FIELD_P_AVERAGE)
        gf = [g.field(L.FIELD_IFACE), g.field(L.FIELD_JFACE), g.field(L.FIELD_KFACE)]
        gf2 = [g.field(L.FIELD_IFACE), g.field(L.FIELD_JFACE), g.field(L.FIELD_KFACE)]
        gvf = [g.field(f) for f in (L.FIELD_IFACE_VOLFRAC, L.FIELD_JFACE_VOLFRAC, L.FIELD_KFACE_VOLFRAC)]
        n = g.n
    assert n == 30000
    assert np.array_equal(gcnt, cnt)
    if precision == "fp64":
        assert np.array_equal(gpav, pav)
        for a, b in zip(gf, (fi, fj, fk)):
            assert np.array_equal(a, b)
        for a, b in zip(gvf, vf):
            assert np.array_equal(a, b)
    else:
        assert np.abs(gpav - pav).max() <= FP32_TOL * 101325.0
        for a, b in zip(gf, (fi, fj, fk)):
            a4, b4 = a.reshape(-1, 4), b.reshape(-1, 4)
            assert np.abs(a4[:, 0] - b4[:, 0]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 0]).max())
            assert np.abs(a4[:, 1:] - b4[:, 1:]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 1:]).max())
        for a, b in zip(gvf, vf):
            assert np.abs(a - b).max() <= FP32_TOL * np.abs(b).max()
        # the volume fraction of the FP32 counts themselves is the oracle's arithmetic, bit for bit
        for d, (a, f) in enumerate(zip(gvf, gf)):
            assert np.array_equal(a, oracle.face_volfrac(IM, JM, KM, coor, d, f, vp))
    for a, b in zip(gf, gf2):
        assert np.array_equal(a, b)


def _run_oracle_steps(oracle, d, nsteps, rk2):
    IM, JM, KM, coor = d["IM"], d["JM"], d["KM"], d["coor"]
    corner = oracle.center_to_corner(IM, JM, KM, d["ucat"])
    n = d["pos"].shape[0]
    s = {"pos": d["pos"].copy(), "vel": np.zeros((n, 3)), "cell": np.full((n, 3), -1, np.int64),
         "weight": np.zeros((n, 3)), "status": np.full(n, 4, np.int64), "pid": np.arange(n, dtype=np.int64),
         "P": np.full(n, 101325.0)}
    # initial setup (simulation.c:129-175): locate, interpolate
    c, w, st, _ = oracle.locate(IM, JM, KM, coor, s["pos"], status=s["status"])
    s.update(cell=c, weight=w, status=st)
    s["vel"] = oracle.interp(IM, JM, KM, corner, c, w)
    for _ in range(nsteps):
        oracle.step(IM, JM, KM, coor, corner, s, d["dt"], rk2=rk2)
    return s, corner


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("path", ["fused", "staged"])
@pytest.mark.parametrize("case", ["config1", "tube_rk2", "warped_rk2"])
def test_full_steps_against_oracle(swarm_cls, oracle, case, precision, path):
    _say("full_steps", case, precision, path)
    if case == "config1":
        d = W.config(1)
        nsteps, rk2 = (10 if precision == "fp64" else 5), False
    elif case == "tube_rk2":
        d = W.config(2, scale=0.25)
        d["pos"] = d["pos"][:40000]
        nsteps, rk2 = 6, True
    else:
        d = W.config(5, scale=0.0625)
        d["pos"] = d["pos"][:40000]
        nsteps, rk2 = 5, True
    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    s, corner = _run_oracle_steps(oracle, d, nsteps, rk2)
    vp = 1.0 / (IM * JM * KM)  # particle volume of the volume fractions (config 2's output)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, d["coor"])
        g.upload_ucat(d["ucat"])
        g.upload_particles(d["pos"])
        g.set_particle_volume(vp)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        # split the steps over two calls: the fused path compacts between calls
        g.AdvanceSimulation(d["dt"], nsteps // 2, rk2=rk2, deposit_cells=True, deposit_faces=True,
                            staged=path == "staged")
        g.AdvanceSimulation(d["dt"], nsteps - nsteps // 2, rk2=rk2, deposit_cells=True, deposit_faces=True,
                            staged=path == "staged")
        out = _by_pid(g.download_particles())
        cnt = g.field(L.FIELD_PARTICLE_COUNT)
        pav = g.field(L.FIELD_P_AVERAGE)
        faces = [g.field(L.FIELD_IFACE), g.field(L.FIELD_JFACE), g.field(L.FIELD_KFACE)]
        vfs = [g.field(f) for f in (L.FIELD_IFACE_VOLFRAC, L.FIELD_JFACE_VOLFRAC, L.FIELD_KFACE_VOLFRAC)]
    for dd, (a, f) in enumerate(zip(vfs, faces)):
        # the volume fraction of the deposited counts, bit for bit the oracle's arithmetic
        assert np.array_equal(a, oracle.face_volfrac(IM, JM, KM, d["coor"], dd, f, vp))
    assert set(np.unique(pav)) <= {0.0, 101325.0} and ((pav > 0) == (cnt > 0)).all()
    if case == "config1" and nsteps == 10:
        # z moves by 10 x 0.1: every particle leaves the unit cube (simulation.c:263-265 removes them)
        assert len(s["pid"]) <= 1
    if precision == "fp64":
        np.testing.assert_array_equal(out["pid"], s["pid"])
        np.testing.assert_array_equal(out["cell"], s["cell"])
        np.testing.assert_array_equal(out["status"], s["status"])
        assert np.array_equal(out["pos"], s["pos"])
        assert np.array_equal(out["vel"], s["vel"])
        assert np.array_equal(out["weight"], s["weight"])
        assert np.array_equal(cnt, oracle.count_cells(IM, JM, KM, s["cell"]))
        ref = oracle.deposit_faces(IM, JM, KM, s["cell"], s["weight"], s["vel"], s["pid"])
        for a, b in zip(faces, ref):
            assert np.array_equal(a, b)
        for dd, (a, f) in enumerate(zip(vfs, ref)):
            assert np.array_equal(a, oracle.face_volfrac(IM, JM, KM, d["coor"], dd, f, vp))
    else:
        # FP32 trajectories drift from FP64 ones; ownership is exact on the
        # variant's own positions, and the drift is small
        common, ia, ib = np.intersect1d(out["pid"], s["pid"], return_indices=True)
        assert len(s["pid"]) > 1000 and len(common) >= 0.99 * len(s["pid"])
        dpos = np.abs(out["pos"][ia] - s["pos"][ib]).max()
        _say("fp32 max position drift vs fp64", dpos)
        assert dpos < 1e-4
        c2, w2, st2, _ = oracle.locate(IM, JM, KM, d["coor"], out["pos"], cell=out["cell"],
                                       status=np.zeros(len(out["pid"]), np.int64))
        # re-locating the FP32 positions from their own cells keeps the cells
        np.testing.assert_array_equal(c2, out["cell"])
        assert np.abs(w2 - out["weight"]).max() < FP32_TOL
        v2 = oracle.interp(IM, JM, KM, corner, c2, w2)
        assert np.abs(v2 - out["vel"]).max() <= FP32_TOL * np.abs(d["ucat"]).max()
        assert np.array_equal(cnt, oracle.count_cells(IM, JM, KM, out["cell"]))
        # faces of the FP32 run's own state against the FP64 oracle's sums of it
        ref = oracle.deposit_faces(IM, JM, KM, out["cell"], out["weight"], out["vel"], out["pid"])
        for a, b in zip(faces, ref):
            a4, b4 = a.reshape(-1, 4), b.reshape(-1, 4)
            assert np.abs(a4[:, 0] - b4[:, 0]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 0]).max())
            assert np.abs(a4[:, 1:] - b4[:, 1:]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 1:]).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_crowded_tiles_with_fast_particles(swarm_cls, oracle, precision):
    """Tiles of 8x8x4 cells holding thousands of particles, more than a
    thousand of which change cell in a step: the fused step's lanes run many
    particles each and most walks leave their cell, and its staged results
    overflow to direct writes.  FP64: bit-exact with the oracle; both
    precisions: identical to the stage-by-stage path."""
    _say("crowded_tiles", precision)
    IM, JM, KM = 16, 16, 8
    coor = W.perturbed_cube(IM, JM, KM, seed=31, amp=0.15)
    ucat = W.ucat_random_fourier(coor, IM, JM, KM, seed=32)
    n = 40000
    pos = 0.05 + 0.9 * np.random.default_rng(33).random((n, 3))
    d = {"IM": IM, "JM": JM, "KM": KM, "coor": coor, "ucat": ucat, "pos": pos,
         "dt": 1.5 / (np.abs(ucat).max() * max(IM, JM, KM))}  # up to ~1.5 cells a step
    nsteps = 3
    # by construction: per starting tile, more than 600 particles change cell
    # in the first step (each with a walk that leaves its cell)
    s1, _ = _run_oracle_steps(oracle, d, 1, True)
    c0, _, _, _ = oracle.locate(IM, JM, KM, coor, pos, status=np.full(n, 4, np.int64))
    start = c0[s1["pid"]]  # (i, j, k)
    moved = (start != s1["cell"]).any(axis=1)
    tile = (start[:, 2] // 4) * 4 + (start[:, 1] // 8) * 2 + start[:, 0] // 8  # 2 x 2 x 2 tiles
    _say("moved per tile", np.bincount(tile[moved], minlength=8))
    assert np.bincount(tile[moved], minlength=8).max() > 600
    s, _ = _run_oracle_steps(oracle, d, nsteps, True)
    outs = []
    for staged in (False, True):
        with swarm_cls(precision=precision) as g:
            g.upload_grid(IM, JM, KM, coor)
            g.upload_ucat(ucat)
            g.upload_particles(pos)
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            g.AdvanceSimulation(d["dt"], nsteps, rk2=True, deposit_cells=True, deposit_faces=True, staged=staged)
            outs.append((_by_pid(g.download_particles()), [g.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE,
                                                                                 L.FIELD_KFACE)]))
    (a, fa), (b, fb) = outs
    for k in ("pid", "cell", "pos", "vel", "weight", "status"):
        assert np.array_equal(a[k], b[k]), k
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
    if precision == "fp64":
        for k in ("pid", "cell", "pos", "vel", "weight", "status"):
            assert np.array_equal(a[k], s[k]), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_single_step_calls_match_one_call(swarm_cls, precision):
    """Twenty one-step calls (each continuing from the key index the previous
    call left; the compaction every 16 steps falls inside the sequence) give
    exactly what one twenty-step call gives, with and without deposition in
    between."""
    _say("single_step_calls", precision)
    d = W.config(5, scale=0.02)
    d["pos"] = d["pos"][:30000]
    outs = []
    for split in (False, True):
        with swarm_cls(precision=precision) as g:
            g.upload_grid(d["IM"], d["JM"], d["KM"], d["coor"])
            g.upload_ucat(d["ucat"])
            g.upload_particles(d["pos"])
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            if split:
                for s in range(20):
                    g.AdvanceSimulation(d["dt"] * 3, 1, rk2=True, deposit_cells=True,
                                        deposit_faces=s % 3 == 2 or s == 19)
            else:
                g.AdvanceSimulation(d["dt"] * 3, 20, rk2=True, deposit_cells=True, deposit_faces=True)
            st = g.stats()
            n_live = g.n  # after the calls a compaction is still pending
            p = g.download_particles()
            assert n_live == st["n_particles"] == len(p["pid"]) == st["located"]
            outs.append((_by_pid(p), g.field(L.FIELD_KFACE), g.field(L.FIELD_P_AVERAGE), st["located"]))
    (a, fa, pa, la), (b, fb, pb, lb) = outs
    for k in ("pid", "cell", "pos", "vel", "weight", "status"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(fa, fb) and np.array_equal(pa, pb) and la == lb


def test_timing_select(swarm_cls):
    """With a kernel-name prefix selected, only the matching launches are timed."""
    _say("timing select")
    d = W.config(5, scale=0.02)
    d["pos"] = d["pos"][:20000]
    with swarm_cls(precision="fp32") as g:
        g.upload_grid(d["IM"], d["JM"], d["KM"], d["coor"])
        g.upload_ucat(d["ucat"])
        g.upload_particles(d["pos"])
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        g.timing(True)
        g.timing_select("step_fused")
        g.timing_reset()
        g.AdvanceSimulation(d["dt"], 3, rk2=True, deposit_faces=True)
        g.synchronize()
        ms, n = g.timing_query("step_fused_fp32")
        assert n == 3 and ms > 0
        assert g.timing_query("step_finish_fp32")[1] == 0 and g.timing_query("center_to_corner")[1] == 0
        g.timing_select("")
        g.timing_reset()
        g.AdvanceSimulation(d["dt"], 2, rk2=True, deposit_faces=True)
        g.synchronize()
        # the second step deposits the first step's state, the call's finishing pass the second's
        assert g.timing_query("step_fused_fp32")[1] == 2 and g.timing_query("step_finish_fp32")[1] == 1
        assert g.timing_query("face_seams")[1] == 2
        g.timing(False)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_long_fused_run_matches_staged(swarm_cls, precision):
    """20 steps in one call (the fused path compacts the swarm mid-call) give
    the same particles, and the same walk iterations, as the stage-by-stage
    path."""
    _say("long_fused_vs_staged", precision)
    d = W.config(5, scale=0.02)
    d["pos"] = d["pos"][:60000]
    outs = []
    for staged in (False, True):
        with swarm_cls(precision=precision) as g:
            g.upload_grid(d["IM"], d["JM"], d["KM"], d["coor"])
            g.upload_ucat(d["ucat"])
            g.upload_particles(d["pos"])
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            g.AdvanceSimulation(d["dt"] * 4, 20, rk2=True, deposit_cells=True, deposit_faces=True, staged=staged)
            outs.append((_by_pid(g.download_particles()), g.field(L.FIELD_KFACE), g.stats()))
    (a, fa, sa), (b, fb, sb) = outs
    for k in ("pid", "cell", "pos", "vel", "weight", "status"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(fa, fb)
    assert sa["particle_steps"] == sb["particle_steps"]
    # the fused walks (some continued from where the straight-line pass left
    # them) take as many iterations as the stage kernels' walks
    assert sa["walk_iterations"] == sb["walk_iterations"]
    assert len(a["pid"]) < 60000  # some particles left the domain


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["warped", "crowded", "crowded_wide_pid"])
def test_steps_deposit_the_state_they_start_from(swarm_cls, oracle, case, precision):
    """Inside a call every step deposits the state it starts from (the
    previous step's deposition, made while the step walks its particles);
    with faces_of_start the faces keep the one the call's last step made.
    It equals the oracle's deposition of the state one step before the end:
    bit for bit in FP64, within 1e-5 in FP32 (of the FP32 run's own state).
    The crowded case has thousands of particles per tile, more than the
    kernel stages in LDS.  Its wide-PID form gives PIDs above 2^32 whose low
    32 bits are in another order than the PIDs, so the buckets must be put
    in the order of the 64-bit PIDs (the 32-bit copy serves only PIDs in
    [0, 2^31))."""
    _say("deposit_start_state", case, precision)
    if case == "warped":
        d = W.config(5, scale=0.02)
        d["pos"] = d["pos"][:30000]
        d["dt"] *= 3
    else:
        IM, JM, KM = 16, 16, 8
        coor = W.perturbed_cube(IM, JM, KM, seed=31, amp=0.15)
        ucat = W.ucat_random_fourier(coor, IM, JM, KM, seed=32)
        d = {"IM": IM, "JM": JM, "KM": KM, "coor": coor, "ucat": ucat,
             "pos": 0.05 + 0.9 * np.random.default_rng(33).random((40000, 3)),
             "dt": 1.5 / (np.abs(ucat).max() * max(IM, JM, KM))}
    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    nsteps = 4
    npart = d["pos"].shape[0]
    pid = None
    if case == "crowded_wide_pid":
        rng = np.random.default_rng(34)
        pid = (rng.permutation(npart).astype(np.int64) << 32) | rng.integers(0, 1 << 31, npart)
        lo = pid & 0xFFFFFFFF
        assert (np.argsort(lo) != np.argsort(pid)).any()

    def run(n, faces_of_start):
        with swarm_cls(precision=precision) as g:
            g.upload_grid(IM, JM, KM, d["coor"])
            g.upload_ucat(d["ucat"])
            if pid is None:
                g.upload_particles(d["pos"])
            else:
                g.upload_particles(d["pos"], pid=pid)
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            g.AdvanceSimulation(d["dt"], n, rk2=True, deposit_cells=False, deposit_faces=True,
                                faces_of_start=faces_of_start)
            return _by_pid(g.download_particles()), [g.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE,
                                                                          L.FIELD_KFACE)]

    before, _ = run(nsteps - 1, False)
    after, faces = run(nsteps, True)
    assert len(after["pid"]) < len(before["pid"]) or case == "warped"
    ref = oracle.deposit_faces(IM, JM, KM, before["cell"], before["weight"], before["vel"], before["pid"])
    if precision == "fp64":
        if pid is None:
            s, _ = _run_oracle_steps(oracle, d, nsteps - 1, True)
            for k in ("pid", "cell", "pos", "vel", "weight"):
                assert np.array_equal(before[k], s[k]), k
        for a, b in zip(faces, ref):
            assert np.array_equal(a, b)
    else:
        for a, b in zip(faces, ref):
            a4, b4 = a.reshape(-1, 4), b.reshape(-1, 4)
            assert np.abs(a4[:, 0] - b4[:, 0]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 0]).max())
            assert np.abs(a4[:, 1:] - b4[:, 1:]).max() <= FP32_TOL * max(1.0, np.abs(b4[:, 1:]).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_equal_pids_are_each_deposited_once(swarm_cls, oracle, precision):
    """A caller may upload equal PIDs (here a tenth of the PIDs distinct,
    each shared by about ten particles).  The in-step deposition, which picks
    each cell's particles by PID, sums every particle once: the faces equal
    the oracle's deposition of the same state up to summation order (equal
    PIDs leave the order among them open), and each direction's face counts
    add up to the particles."""
    _say("equal pids", precision)
    IM, JM, KM = 16, 16, 8
    coor = W.perturbed_cube(IM, JM, KM, seed=51, amp=0.15)
    ucat = W.ucat_random_fourier(coor, IM, JM, KM, seed=52)
    rng = np.random.default_rng(53)
    npart = 20000
    pos = 0.05 + 0.9 * rng.random((npart, 3))
    pid = rng.integers(0, npart // 10, npart).astype(np.int64)
    dt = 1.5 / (np.abs(ucat).max() * max(IM, JM, KM))

    def run(n, faces_of_start):
        with swarm_cls(precision=precision) as g:
            g.upload_grid(IM, JM, KM, coor)
            g.upload_ucat(ucat)
            g.upload_particles(pos, pid=pid)
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            g.AdvanceSimulation(dt, n, rk2=True, deposit_cells=False, deposit_faces=True,
                                faces_of_start=faces_of_start)
            return g.download_particles(), [g.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)]

    before, _ = run(2, False)
    _, faces = run(3, True)
    nb = len(before["pid"])
    ref = oracle.deposit_faces(IM, JM, KM, before["cell"], before["weight"], before["vel"], before["pid"])
    tol = 1e-12 if precision == "fp64" else FP32_TOL
    for a, b in zip(faces, ref):
        a4, b4 = a.reshape(-1, 4), b.reshape(-1, 4)
        assert abs(a4[:, 0].sum() - nb) <= tol * nb
        assert np.abs(a4[:, 0] - b4[:, 0]).max() <= tol * max(1.0, np.abs(b4[:, 0]).max())
        assert np.abs(a4[:, 1:] - b4[:, 1:]).max() <= tol * max(1.0, np.abs(b4[:, 1:]).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["fused", "staged", "scatter", "empty"])
def test_particle_count_summary(swarm_cls, precision, case):
    """CalculateParticleCountPerCell's logged diagnostics (ParticleMotion.c:1021-1026:
    VecSum and VecMax of ParticleCount) reduced on the device equal numpy's on
    the downloaded field: the total, the largest count and the first flat
    node slot holding it.  Two cells hold the largest count (40 particles
    each), so the first of them must be reported; 500 other cells hold one
    particle each.  Counts after a fused step, a staged step, the stage-wise
    scatter, and an empty swarm (0, 0, slot 0)."""
    _say("count summary", case, precision)
    IM, JM, KM = 16, 12, 10
    coor = W.uniform_grid(IM, JM, KM)
    ucat = W.ucat_uniform(IM, JM, KM, (0.0, 0.0, 1.0))
    rng = np.random.default_rng(61)
    h = np.array([1.0 / IM, 1.0 / JM, 1.0 / KM])
    cells = [(3, 4, 2)] * 40 + [(11, 9, 7)] * 40
    flat = rng.choice(IM * JM * KM, 520, replace=False)
    others = [(int(f % IM), int(f // IM % JM), int(f // (IM * JM))) for f in flat]
    cells += [c for c in others if c not in ((3, 4, 2), (11, 9, 7))][:500]
    cells = np.array(cells, np.float64)
    pos = (cells + 0.3 + 0.4 * rng.random(cells.shape)) * h
    if case == "empty":
        pos = np.zeros((0, 3))
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_ucat(ucat)
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        if case == "scatter":
            g.ScatterAllParticleFieldsToEulerFields()
        else:
            g.AdvanceSimulation(1e-9, 1, rk2=True, deposit_cells=True, staged=case == "staged")
        total, mx, at = g.particle_count_summary()
        cnt = g.field(L.FIELD_PARTICLE_COUNT)
    assert cnt.shape[0] == (IM + 1) * (JM + 1) * (KM + 1)
    assert (total, mx, at) == (cnt.sum(), cnt.max(), int(np.argmax(cnt)))
    if case == "empty":
        assert (total, mx, at) == (0.0, 0.0, 0)
    else:
        assert (total, mx) == (580.0, 40.0)
        assert at == (2 * (JM + 1) + 4) * (IM + 1) + 3  # cell (3,4,2), before (11,9,7)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reference_unit_test_point_near_the_upper_face(swarm_cls, oracle, precision):
    """The expectation of the reference's one unit test (old/test_walkingsearch.c:103-128):
    in a one-cell unit grid the point (0.9999999999999, 0.5, 0.5), 1e-13 from
    the upper i-face, is found in cell (0,0,0); weights as the oracle's (FP64
    bit for bit)."""
    _say("reference unit test", precision)
    coor = W.uniform_grid(1, 1, 1)
    pos = np.array([[0.9999999999999, 0.5, 0.5]])
    c, w, st, _ = oracle.locate(1, 1, 1, coor, pos)
    assert st[0] == L.ACTIVE_AND_LOCATED and (c[0] == 0).all()
    with swarm_cls(precision=precision) as g:
        g.upload_grid(1, 1, 1, coor)
        g.upload_particles(pos)
        g.LocateAllParticlesInGrid()
        d = g.download_particles()
    assert d["status"][0] == L.ACTIVE_AND_LOCATED
    np.testing.assert_array_equal(d["cell"], c)
    if precision == "fp64":
        assert np.array_equal(d["weight"], w)
    else:
        assert np.abs(d["weight"] - w).max() < FP32_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_runs_are_bit_identical(swarm_cls, precision):
    """Two runs of one crowded workload (about 29 particles per cell, 7,500
    per 8x8x4 tile, more than the step kernel stages in LDS; slots are taken
    by atomics in whatever order the hardware runs them, and the swarm is
    compacted once in the 20 steps) give the same
    particles, per PID, and the same counts, P averages and faces, bit for
    bit: deposits are sums in PID order, with no float atomics."""
    _say("bit-identical runs", precision)
    IM, JM, KM = 16, 16, 8
    coor = W.perturbed_cube(IM, JM, KM, seed=71, amp=0.15)
    ucat = W.ucat_random_fourier(coor, IM, JM, KM, seed=72)
    pos = 0.05 + 0.9 * np.random.default_rng(73).random((60000, 3))
    dt = 1.0 / (np.abs(ucat).max() * max(IM, JM, KM))
    fields = (L.FIELD_PARTICLE_COUNT, L.FIELD_P_AVERAGE, L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)

    def run():
        with swarm_cls(precision=precision) as g:
            g.upload_grid(IM, JM, KM, coor)
            g.upload_ucat(ucat)
            g.upload_particles(pos)
            g.LocateAllParticlesInGrid()
            g.InterpolateAllFieldsToSwarm()
            g.AdvanceSimulation(dt, 20, rk2=True, deposit_cells=True, deposit_faces=True)
            return _by_pid(g.download_particles()), [g.field(f) for f in fields]

    a, fa = run()
    b, fb = run()
    assert 0 < len(a["pid"]) < pos.shape[0]
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
