"""Pins the seeding oracle (oracle/seedref.c).

The random stream is checked against the C library's own srand48/drand48
(the functions PETSc's PETSCRAND48 type calls); the placement against a
Python restatement driven by that same libc stream, and against closed forms
on a uniform grid.  A wrong draw order, a wrong cell pick or a wrong corner
order in the blend fails one of these."""
import ctypes as C

import numpy as np
import pytest

from fdf_amd import workloads as W

libc = C.CDLL(None)
libc.drand48.restype = C.c_double
libc.srand48.argtypes = [C.c_long]


def libc_stream(seed, count):
    libc.srand48(seed)
    return [libc.drand48() for _ in range(count)]


@pytest.mark.parametrize("seed", [202403, 202404, 0, 12345, -3, 2**31 + 17])
def test_drand48_is_the_libc_stream(oracle, seed):
    assert np.array_equal(oracle.drand48_stream(seed, 20000), np.array(libc_stream(seed, 20000)))


def _blend(coor, IM, JM, i, j, k, xi, eta, zta):
    N4 = coor.reshape(-1, JM + 1, IM + 1, 3)
    x = y = z = 0.0
    for c in range(8):
        v = N4[k + (c >> 2 & 1), j + (c >> 1 & 1), i + (c & 1)]
        n = ((xi if c & 1 else 1.0 - xi) * (eta if c & 2 else 1.0 - eta)) * (zta if c & 4 else 1.0 - zta)
        x += n * float(v[0])
        y += n * float(v[1])
        z += n * float(v[2])
    return x, y, z


def _pick(r, n):
    return min(max(0, int(r * n)), n - 1)


def _clamp(v):
    return min(max(0.0, v), 1.0 - 1.0e-7)


def python_seed(coor, IM, JM, KM, n, volumetric, face, seed):
    """InitializeParticleBasicProperties restated in Python on libc's drand48."""
    libc.srand48(seed)
    d = libc.drand48
    out = []
    for _ in range(n):
        if volumetric:
            i, j, k = _pick(d(), IM), _pick(d(), JM), _pick(d(), KM)
            xi, eta, zta = d(), d(), d()
        else:
            axis, hi = face // 2, face % 2
            edge = 1.0 - 1.0e-6 if hi else 1.0e-6
            if axis == 0:
                i, xi = (IM - 1 if hi else 0), edge
                j, k = _pick(d(), JM), _pick(d(), KM)
                eta, zta = _clamp(d()), _clamp(d())
            elif axis == 1:
                j, eta = (JM - 1 if hi else 0), edge
                i, k = _pick(d(), IM), _pick(d(), KM)
                xi, zta = _clamp(d()), _clamp(d())
            else:
                k, zta = (KM - 1 if hi else 0), edge
                i, j = _pick(d(), IM), _pick(d(), JM)
                xi, eta = _clamp(d()), _clamp(d())
        out.append(_blend(coor, IM, JM, i, j, k, xi, eta, zta))
    return np.array(out)


@pytest.mark.parametrize("mode", ["volume", 0, 1, 2, 3, 4, 5])
def test_seeding_matches_libc_driven_restatement(oracle, mode):
    IM, JM, KM = 7, 6, 9
    coor = W.perturbed_cube(IM, JM, KM, seed=5)
    vol = mode == "volume"
    face = 4 if vol else mode
    got = oracle.seed_particles(IM, JM, KM, coor, 3000, volumetric=vol, face=face)
    exp = python_seed(coor, IM, JM, KM, 3000, vol, face, 202403)
    assert np.array_equal(got, exp)


def test_seeding_closed_forms_on_a_uniform_grid(oracle):
    IM, JM, KM = 10, 8, 12
    coor = W.uniform_grid(IM, JM, KM, lo=(0.0, -1.0, 2.0), hi=(1.0, 1.0, 5.0))
    h = np.array([1.0 / IM, 2.0 / JM, 3.0 / KM])
    lo = np.array([0.0, -1.0, 2.0])
    r = oracle.drand48_stream(202403, 6 * 5000).reshape(-1, 6)
    cell = np.minimum((r[:, :3] * [IM, JM, KM]).astype(np.int64), [IM - 1, JM - 1, KM - 1])
    exp = lo + (cell + r[:, 3:]) * h
    got = oracle.seed_particles(IM, JM, KM, coor, 5000, volumetric=True)
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-14)
    # inlet -Zeta: 1e-6 of a cell above the face, spread over the whole face
    s = oracle.seed_particles(IM, JM, KM, coor, 5000, volumetric=False, face=4)
    np.testing.assert_allclose(s[:, 2], lo[2] + 1e-6 * h[2], rtol=0, atol=1e-14)
    assert s[:, 0].min() < 0.05 and s[:, 0].max() > 0.95 and s[:, 1].min() < -0.9 and s[:, 1].max() > 0.9
    # +Xi: 1e-6 of a cell inside the last cell's +x face
    s = oracle.seed_particles(IM, JM, KM, coor, 100, volumetric=False, face=1)
    np.testing.assert_allclose(s[:, 0], 1.0 - 1e-6 * h[0], rtol=0, atol=1e-14)


def test_volumetric_cells_are_uniform(oracle):
    IM, JM, KM = 4, 4, 4
    coor = W.uniform_grid(IM, JM, KM)
    p = oracle.seed_particles(IM, JM, KM, coor, 64000, volumetric=True)
    cells = np.floor(p * 4).astype(int)
    counts = np.bincount((cells[:, 2] * 4 + cells[:, 1]) * 4 + cells[:, 0], minlength=64)
    assert counts.min() > 850 and counts.max() < 1150  # 1000 expected per cell
