"""The walk's edge cases through the gfx950 path, both precision modes,
against the oracle: cells, statuses and walk iterations exact (weights
bit-exact in FP64 mode).  Points are placed where random points almost never
land:

  * exactly on interior face planes, each walked to from both sides (the
    first cell reached takes a point in its zero band, walkingsearch.c:219-222);
  * outside one face but inside another face's zero band, which the
    reference accepts (position >= 0, walkingsearch.c:407-462, :1197);
  * at distances thr * (1 +- 2^-20), thr * (1 +- 1e-3), thr / 2 and 0 from
    warped faces, thr = 1e-11 h the zero band (walkingsearch.c:10-12, :827);
  * in the wedge gaps and overlaps between the two k-face planes that
    neighbouring cells of a warped grid build from different corner vertices;
  * in cells so small that the absolute 1e-9 orientation test
    (walkingsearch.c:181) never flips a normal.

In FP32 mode these points defeat the certified FP32 filter, so the test
also asserts that the FP64 recomputation ran, in the locate and in a fused
step (dt = 0, every particle re-located from its own cell).  A degenerate
face (walkingsearch.c:147-155) is an error in both modes and on both paths;
a fused step that meets one leaves the particles unusable until new ones
are uploaded."""
import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu

OFF = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 7), (1, 6), (2, 5), (3, 4)]
FACES = {"BACK": (0, 3, 2, 1), "FRONT": (4, 7, 6, 5), "BOTTOM": (0, 1, 6, 7), "TOP": (3, 4, 5, 2),
         "LEFT": (0, 7, 4, 3), "RIGHT": (1, 2, 5, 6)}


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def swarm_cls():
    from fdf_amd.swarm import PicGpu

    return PicGpu


def _vertices(coor, IM, JM, KM, i, j, k):
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    return np.array([c[k + dk, j + dj, i + di] for di, dj, dk in OFF])


def _plane(V, face):
    """Unit normal (oriented away from the centroid) and face centre, as the
    reference builds them (walkingsearch.c:102-227), and the zero band."""
    v1, v2, v3, v4 = (V[q] for q in FACES[face])
    n = np.cross(v2 - v1, v4 - v1)
    fc = 0.25 * (v1 + v2 + v3 + v4)
    cc = V.mean(0)
    if n @ (cc - fc) > 1e-9:
        n = -n
    h = np.mean([np.linalg.norm(V[b] - V[a]) for a, b in EDGES])
    return n / np.linalg.norm(n), fc, 1e-11 * h


def _compare(swarm_cls, oracle, IM, JM, KM, coor, pos, start, precision, expect_fallbacks):
    n = len(pos)
    status = np.zeros(n, np.int64)
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos, cell=start, status=status)
    with swarm_cls(precision=precision, collect_walk_stats=True) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_ucat(W.ucat_uniform(IM, JM, KM, (0.0, 0.0, 0.0)))
        g.upload_particles(pos, cell=start, status=status)
        g.LocateAllParticlesInGrid()
        s_loc = g.stats()
        gsteps = g.walk_steps()
        d = g.download_particles()
        o = np.argsort(d["pid"])
        np.testing.assert_array_equal(d["status"][o], st)
        np.testing.assert_array_equal(d["cell"][o], cell)
        np.testing.assert_array_equal(gsteps[o], steps)
        ok = st == 1
        if precision == "fp64":
            assert np.array_equal(d["weight"][o][ok], w[ok])
        else:
            assert np.abs(d["weight"][o][ok] - w[ok]).max() < 1e-5
        # a fused step with dt = 0: every located particle is walked again
        # from its own cell at the same position (zero field, Euler), after
        # the LOST ones are removed as the reference does
        g.CheckAndRemoveLostParticles()
        g.CheckAndRemoveOutOfBoundsParticles()
        g.InterpolateAllFieldsToSwarm()
        g.AdvanceSimulation(0.0, 1, rk2=False, deposit_cells=True, deposit_faces=True)
        s_step = g.stats()
        d2 = g.download_particles()
    o2 = np.argsort(d2["pid"])
    c2, w2, st2, _ = oracle.locate(IM, JM, KM, coor, pos[ok], cell=cell[ok], status=np.zeros(ok.sum(), np.int64))
    lo, hi = oracle.bbox(IM, JM, KM, coor)
    keep = (st2 == 1) & ((pos[ok] >= lo) & (pos[ok] <= hi)).all(1)
    np.testing.assert_array_equal(d2["pid"][o2], np.arange(n)[ok][keep])
    np.testing.assert_array_equal(d2["cell"][o2], c2[keep])
    if precision == "fp64":
        assert np.array_equal(d2["weight"][o2], w2[keep])
    _say(precision, "fp64 fallback faces: locate", s_loc["fp64_fallbacks"], "fused step", s_step["fp64_fallbacks"])
    if precision == "fp32" and expect_fallbacks:
        assert s_loc["fp64_fallbacks"] > 0 and s_step["fp64_fallbacks"] > 0
    return st, cell


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_points_on_face_planes_from_both_sides(swarm_cls, oracle, precision):
    _say("on_face_planes", precision)
    IM = JM = KM = 8  # nodes at multiples of 1/8: the planes are exact
    coor = W.uniform_grid(IM, JM, KM)
    rng = np.random.default_rng(21)
    m = 3000
    pos = rng.random((m, 3)) * 0.9 + 0.05
    axis = rng.integers(0, 3, m)
    plane = rng.integers(1, 8, m)
    pos[np.arange(m), axis] = plane / 8.0
    lower = np.floor(pos * 8).astype(np.int64)
    lower[np.arange(m), axis] = plane - 1
    upper = lower.copy()
    upper[np.arange(m), axis] = plane
    pos2 = np.concatenate([pos, pos])
    start = np.concatenate([lower, upper])
    st, cell = _compare(swarm_cls, oracle, IM, JM, KM, coor, pos2, start, precision, True)
    # the first cell reached keeps the point: the walk from either side stops at once
    assert (st == 1).all()
    np.testing.assert_array_equal(cell, start)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_outside_but_in_another_faces_zero_band(swarm_cls, oracle, precision):
    _say("outside_but_on_face", precision)
    IM = JM = KM = 8
    coor = W.uniform_grid(IM, JM, KM)
    rng = np.random.default_rng(22)
    m = 2000
    ijk = rng.integers(1, 6, (m, 3))
    pos = (ijk + rng.random((m, 3)) * 0.8 + 0.1) / 8.0
    pos[:, 0] = ijk[:, 0] / 8.0  # on the LEFT face plane of cell ijk
    pos[:, 1] = (ijk[:, 1] + 1) / 8.0 + rng.random(m) * 0.02  # beyond its TOP face
    st, cell = _compare(swarm_cls, oracle, IM, JM, KM, coor, pos, ijk.astype(np.int64), precision, True)
    # accepted in the start cell although the point is outside its TOP face
    assert (st == 1).all() and (cell == ijk).all()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_points_at_the_zero_band_of_warped_faces(swarm_cls, oracle, precision):
    _say("zero_band", precision)
    IM, JM, KM = 8, 8, 12
    coor = W.perturbed_cube(IM, JM, KM, seed=12, amp=0.3)
    rng = np.random.default_rng(23)
    factors = [0.0, 0.5, 1 - 2.0 ** -20, 1 + 2.0 ** -20, 1 - 1e-3, 1 + 1e-3, 2.0]
    pos, start = [], []
    for _ in range(250):
        i, j, k = rng.integers(1, IM - 1), rng.integers(1, JM - 1), rng.integers(1, KM - 1)
        V = _vertices(coor, IM, JM, KM, i, j, k)
        face = list(FACES)[rng.integers(0, 6)]
        nhat, fc, thr = _plane(V, face)
        # a tangential offset keeps the point inside the face's other bounds
        t = np.cross(nhat, rng.normal(size=3))
        t *= 0.05 * np.linalg.norm(V[5] - V[0]) / np.linalg.norm(t)
        for f in factors:
            for sgn in (1.0, -1.0):
                pos.append(fc + t - sgn * f * thr * nhat)
                start.append((i, j, k))
    pos, start = np.array(pos), np.array(start, np.int64)
    _compare(swarm_cls, oracle, IM, JM, KM, coor, pos, start, precision, True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_k_face_wedge_gaps_and_overlaps(swarm_cls, oracle, precision):
    _say("k_face_wedges", precision)
    IM, JM, KM = 8, 8, 12
    coor = W.perturbed_cube(IM, JM, KM, seed=12, amp=0.35)
    rng = np.random.default_rng(24)
    gap, over = [], []
    for _ in range(400):
        i, j, k = rng.integers(0, IM), rng.integers(0, JM), rng.integers(0, KM - 1)
        Vl, Vu = _vertices(coor, IM, JM, KM, i, j, k), _vertices(coor, IM, JM, KM, i, j, k + 1)
        nF, fc, _ = _plane(Vl, "FRONT")
        nB, _, _ = _plane(Vu, "BACK")
        axis = nF - nB
        axis /= np.linalg.norm(axis)
        for _ in range(20):
            # near the shared face: inside it tangentially, a small step along the axis
            a, b = rng.random(2) * 0.8 + 0.1
            q = (1 - a) * (1 - b) * Vu[0] + a * (1 - b) * Vu[1] + a * b * Vu[2] + (1 - a) * b * Vu[3]
            p = q + axis * rng.normal() * 0.02 * np.linalg.norm(Vl[5] - Vl[0])
            dF, dB = (fc - p) @ nF, (fc - p) @ nB
            if dF < -1e-12 and dB < -1e-12:
                gap.append((p, (i, j, k)))
            elif dF > 1e-12 and dB > 1e-12:
                over.append((p, (i, j, k)))
    _say("wedge points: gap", len(gap), "overlap", len(over))
    assert len(gap) > 20 and len(over) > 20
    pts = [p for p, _ in gap + over]
    lower = [c for _, c in gap + over]
    upper = [(c[0], c[1], c[2] + 1) for c in lower]
    pos = np.array(pts * 3)
    start = np.array(lower + upper + [(-1, -1, -1)] * len(pts), np.int64)
    _compare(swarm_cls, oracle, IM, JM, KM, coor, pos, start, precision, False)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cells_too_small_for_the_orientation_test(swarm_cls, oracle, precision):
    _say("tiny_cells", precision)
    IM, JM, KM = 10, 10, 10
    coor = W.perturbed_cube(IM, JM, KM, seed=25, amp=0.2) * 1e-3  # h = 1e-4: |n (cc - fc)| ~ 1e-12
    pos = np.random.default_rng(26).random((4000, 3)) * 1e-3
    start = np.full((4000, 3), -1, np.int64)
    _compare(swarm_cls, oracle, IM, JM, KM, coor, pos, start, precision, False)


def _degenerate_corner_grid():
    IM, JM, KM = 4, 4, 4
    coor = W.uniform_grid(IM, JM, KM).reshape(KM + 1, JM + 1, IM + 1, 3)
    coor[KM, JM, IM] = coor[KM, JM, IM - 1]  # cell (3,3,3): its FRONT face has |n| = 0
    return IM, JM, KM, coor.reshape(-1, 3)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_degenerate_face_is_an_error_in_locate(swarm_cls, precision):
    _say("degenerate_locate", precision)
    IM, JM, KM, coor = _degenerate_corner_grid()
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_particles(np.array([[0.9, 0.9, 0.9]]), cell=np.array([[3, 3, 3]]), status=np.zeros(1, np.int64))
        with pytest.raises(L.PicgpuError) as e:
            g.LocateAllParticlesInGrid()
        assert e.value.code == -3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_degenerate_face_in_a_fused_step(swarm_cls, precision):
    _say("degenerate_fused", precision)
    IM, JM, KM, coor = _degenerate_corner_grid()
    pos = np.array([[0.9, 0.9, 0.9], [0.1, 0.1, 0.1], [0.6, 0.3, 0.4]])
    cell = np.array([[3, 3, 3], [0, 0, 0], [2, 1, 1]], np.int64)
    status = np.ones(3, np.int64)
    with swarm_cls(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_ucat(W.ucat_uniform(IM, JM, KM, (0.0, 0.0, 0.0)))
        g.upload_particles(pos, cell=cell, status=status)
        with pytest.raises(L.PicgpuError) as e:
            g.AdvanceSimulation(0.01, 1, rk2=True, deposit_faces=True)
        assert e.value.code == -3
        # the particles are unusable until new ones are uploaded
        for call in (g.download_particles, lambda: g.AdvanceSimulation(0.01, 1), g.LocateAllParticlesInGrid):
            with pytest.raises(L.PicgpuError) as e:
                call()
            assert e.value.code == -2
        g.upload_particles(pos[1:], cell=cell[1:], status=status[1:])
        g.AdvanceSimulation(0.0, 2, rk2=True, deposit_faces=True)
        d = g.download_particles()
    assert sorted(d["pid"]) == [0, 1]
    np.testing.assert_array_equal(d["cell"][np.argsort(d["pid"])], cell[1:])
