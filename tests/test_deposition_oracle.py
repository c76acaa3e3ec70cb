"""The oracle's face deposition and face volume fraction (definitions with no
reference counterpart), pinned by what the mathematics fixes: exactly
representable grids where every volume is known, affine cells whose volume
is |det| of the edge matrix, a tiling identity on a warped grid, single
particles, and a summation order chosen so that the other plausible order
rounds differently."""
import numpy as np

from fdf_amd import workloads as W

OFF = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]


def _vertices(coor, IM, JM, KM, i, j, k):
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    return np.array([c[k + dk, j + dj, i + di] for di, dj, dk in OFF])


def _dyadic_grid(IM, JM, KM, h=(0.25, 0.5, 0.125)):
    """Uniform grid whose coordinates and cell volume are exact binary fractions."""
    return W.uniform_grid(IM, JM, KM, hi=(h[0] * IM, h[1] * JM, h[2] * KM)), h[0] * h[1] * h[2]


def test_face_sum_is_lower_sum_plus_upper_sum(oracle):
    # lower cell (0,0,0): one particle with a1 * u = 0.5 * 2 = 1; upper cell
    # (1,0,0): two particles with (1 - a1) * u = 2^-53 each.  Lower + (U1 +
    # U2) = 1 + 2^-52; one running sum 1 + U1 + U2 would round to 1.
    IM, JM, KM = 2, 1, 1
    cell = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 0]], np.int64)
    w = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    vel = np.array([[2.0, 0, 0], [2.0 ** -52, 0, 0], [2.0 ** -52, 0, 0]])
    fi, _, _ = oracle.deposit_faces(IM, JM, KM, cell, w, vel, np.array([5, 9, 7]))
    f = fi.reshape(KM, JM, IM + 1, 4)
    assert f[0, 0, 1, 1] == 1.0 + 2.0 ** -52
    assert f[0, 0, 1, 0] == 1.5
    assert f[0, 0, 0, 1] == 0.5 * 2.0 and f[0, 0, 2, 1] == 2.0 ** -52  # boundary faces: one cell each


def test_face_sums_follow_pid_order(oracle):
    # one cell, three particles whose momenta 1, 2^-53, 2^-53 round differently
    # when the small ones are summed first (PID order 3 < 5 < 8 puts them first)
    IM, JM, KM = 1, 1, 1
    cell = np.zeros((3, 3), np.int64)
    w = np.zeros((3, 3))  # weight a = 0: everything on the lower face, 1 - a = 1
    vel = np.array([[1.0, 0, 0], [2.0 ** -53, 0, 0], [2.0 ** -53, 0, 0]])
    fi, _, _ = oracle.deposit_faces(IM, JM, KM, cell, w, vel, np.array([8, 3, 5]))
    assert fi.reshape(-1, 4)[0, 1] == 1.0 + 2.0 ** -52
    fi2, _, _ = oracle.deposit_faces(IM, JM, KM, cell, w, vel, np.array([1, 3, 5]))
    assert fi2.reshape(-1, 4)[0, 1] == 1.0


def test_cell_volume_exact_on_a_dyadic_grid(oracle):
    IM, JM, KM = 3, 2, 4
    coor, vol = _dyadic_grid(IM, JM, KM)
    for (i, j, k) in [(0, 0, 0), (2, 1, 3), (1, 0, 2)]:
        assert oracle.cell_volume(_vertices(coor, IM, JM, KM, i, j, k)) == vol


def test_cell_volume_of_affine_cells_is_abs_det(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rng.normal(size=(3, 3)) + 2.0 * np.eye(3) * rng.choice([-1, 1])
        b = rng.normal(size=3)
        V = np.array([A @ np.array(o, float) + b for o in OFF])
        assert abs(oracle.cell_volume(V) - abs(np.linalg.det(A))) <= 1e-12 * abs(np.linalg.det(A))


def test_cell_volumes_tile_a_warped_cube(oracle):
    # the six tetrahedra of neighbouring cells split their shared faces along
    # the same diagonal, so the cell volumes of a grid whose boundary nodes
    # stay on the unit cube's faces add up to 1 (at this amplitude no
    # tetrahedron is inverted)
    IM, JM, KM = 6, 5, 7
    coor = W.perturbed_cube(IM, JM, KM, seed=9, amp=0.2)
    tot = sum(oracle.cell_volume(_vertices(coor, IM, JM, KM, i, j, k))
              for k in range(KM) for j in range(JM) for i in range(IM))
    assert abs(tot - 1.0) < 1e-13
    # a wrong decomposition (last tetrahedron through V6 instead of V4) does not tile
    V = _vertices(coor, IM, JM, KM, 2, 2, 3)
    V2 = V.copy()
    V2[4] = V[6]
    assert oracle.cell_volume(V2) != oracle.cell_volume(V)


def test_face_volfrac_closed_forms(oracle):
    IM, JM, KM = 3, 2, 4
    coor, vol = _dyadic_grid(IM, JM, KM)
    rng = np.random.default_rng(5)
    vp = 0.0078125  # 2^-7: every product below is exact
    for d in range(3):
        nf = (IM + (d == 0)) * (JM + (d == 1)) * (KM + (d == 2))
        face = np.zeros((nf, 4))
        face[:, 0] = rng.integers(0, 9, nf) * 0.25
        face[:, 1:] = rng.normal(size=(nf, 3))
        vf = oracle.face_volfrac(IM, JM, KM, coor, d, face, vp)
        shape = (KM + (d == 2), JM + (d == 1), IM + (d == 0))
        idx = np.indices(shape)  # k, j, i
        ax = [2, 1, 0][d]
        boundary = (idx[ax] == 0) | (idx[ax] == shape[ax] - 1)
        vdual = np.where(boundary.ravel(), 0.5 * vol, vol)
        assert np.array_equal(vf, face[:, 0] * vp / vdual)


def test_face_volfrac_of_one_particle(oracle):
    IM, JM, KM = 4, 3, 5
    coor = W.perturbed_cube(IM, JM, KM, seed=2, amp=0.25)
    rng = np.random.default_rng(8)
    pos = rng.random((200, 3)) * 0.6 + 0.2
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    q = int(np.where(st == 1)[0][0])
    c, a = cell[q], w[q]
    faces = oracle.deposit_faces(IM, JM, KM, cell[q:q + 1], w[q:q + 1], np.ones((1, 3)), np.arange(1))
    vp = 1e-3
    vols = {}

    def vol(i, j, k):
        if (i, j, k) not in vols:
            vols[(i, j, k)] = oracle.cell_volume(_vertices(coor, IM, JM, KM, i, j, k))
        return vols[(i, j, k)]

    for d in range(3):
        vf = oracle.face_volfrac(IM, JM, KM, coor, d, faces[d], vp)
        shape = (KM + (d == 2), JM + (d == 1), IM + (d == 0))
        vf = vf.reshape(shape)
        e = np.eye(3, dtype=np.int64)[d]
        lo_face, hi_face = c.copy(), c + e  # face numbers (i, j, k) of the cell's two faces
        lower_nb = c - e
        v_lo = vol(*c) + (vol(*lower_nb) if lower_nb[d] >= 0 else 0.0)
        upper_nb = c + e
        v_hi = vol(*c) + (vol(*upper_nb) if upper_nb[d] < (IM, JM, KM)[d] else 0.0)
        got_lo = vf[lo_face[2], lo_face[1], lo_face[0]]
        got_hi = vf[hi_face[2], hi_face[1], hi_face[0]]
        assert abs(got_lo - (1 - a[d]) * vp / (0.5 * v_lo)) <= 1e-14 * got_lo
        assert abs(got_hi - a[d] * vp / (0.5 * v_hi)) <= 1e-14 * max(got_hi, 1e-300)
        assert np.count_nonzero(vf) <= 2
