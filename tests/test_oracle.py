"""Pins the CPU oracle (oracle/picref.c) without the reference.

The reference cannot be built here (PETSc/MPI absent) and ships no tests or
golden data, so the oracle is pinned by what the mathematics fixes: closed
forms on uniform grids, brute force on tiny warped grids, invariants and the
two expectations of the reference's only unit test (old/test_walkingsearch.c).
Each check is chosen so that a plausible slip (wrong vertex order, sign,
index, dropped term) fails it.
"""
import math

import numpy as np
import pytest

from fdf_amd import workloads as W


def _uniform(IM, JM, KM):
    return W.uniform_grid(IM, JM, KM)


# --------------------------------------------------------------------- walk
def test_walk_uniform_closed_form(oracle):
    """Cell = floor, weight = fractional coordinate, walk length from (0,0,0) = max(i,j,k)+1."""
    IM, JM, KM = 20, 20, 280
    coor = _uniform(IM, JM, KM)
    pos = W.particles_uniform_cube(10_000, 1001)
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos)
    exp = np.floor(pos * [IM, JM, KM]).astype(np.int64)
    assert (st == 1).all()
    np.testing.assert_array_equal(cell, exp)
    np.testing.assert_allclose(w, pos * [IM, JM, KM] - exp, atol=1e-12)
    np.testing.assert_array_equal(steps, exp.max(axis=1) + 1)


def test_walk_from_previous_cell(oracle):
    """Start rule walkingsearch.c:571-629: a valid previous cell is the start."""
    IM, JM, KM = 10, 12, 14
    coor = _uniform(IM, JM, KM)
    rng = np.random.default_rng(5)
    pos = rng.random((2000, 3)) * 0.98 + 0.01
    exp = np.floor(pos * [IM, JM, KM]).astype(np.int64)
    prev = np.clip(exp + rng.integers(-2, 3, size=exp.shape), 0, [IM - 1, JM - 1, KM - 1])
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos, cell=prev, status=np.zeros(2000, np.int64))
    np.testing.assert_array_equal(cell, exp)
    np.testing.assert_array_equal(steps, np.abs(exp - prev).max(axis=1) + 1)


def test_walk_outside_domain_is_lost(oracle):
    IM, JM, KM = 8, 8, 8
    coor = _uniform(IM, JM, KM)
    pos = np.array([[1.5, 0.53, 0.47], [0.53, -0.2, 0.47], [0.53, 0.47, 2.0]])
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos)
    assert (st == 3).all() and (cell == -1).all()


def test_walk_far_start_hits_max_traversal(oracle):
    """From (0,0,0) a point in cell k >= 1000 needs > 1000 iterations: LOST (walkingsearch.c:1123, 1206-1212)."""
    IM, JM, KM = 2, 2, 1100
    coor = _uniform(IM, JM, KM)
    pos = np.array([[0.3, 0.7, 998.5 / 1100], [0.3, 0.7, 999.5 / 1100], [0.3, 0.7, 1050.5 / 1100]])
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos)
    assert cell[0, 2] == 998 and st[0] == 1 and steps[0] == 999
    assert cell[1, 2] == 999 and st[1] == 1 and steps[1] == 1000
    assert st[2] == 3 and steps[2] == 1000


def test_walk_old_unit_test_expectations(oracle):
    """old/test_walkingsearch.c:103-128 (point near x=1 found in a 1-cell grid)
    and :72-100 (degenerate cell is an error)."""
    coor = _uniform(1, 1, 1)
    cell, w, st, _ = oracle.locate(1, 1, 1, coor, np.array([[0.9999999999999, 0.5, 0.5]]))
    assert st[0] == 1 and tuple(cell[0]) == (0, 0, 0)
    bad = coor.copy()
    bad[:] = 0.0  # every vertex coincident
    with pytest.raises(RuntimeError):
        oracle.locate(1, 1, 1, bad, np.array([[0.0, 0.0, 0.0]]))


def test_on_face_points_taken_by_first_cell_reached(oracle):
    """A point on an interior face plane has d == 0 there and is accepted by
    the cell the walk reaches first.  From (0,0,0) the walk moves diagonally;
    x = 0.25 is the RIGHT plane of cell i=4, reached at iteration 5 when
    j = k = 4 too: that cell is accepted although y and z lie beyond its TOP
    and FRONT planes (position 1: one zero, other faces negative), and the
    clamped distances give weights ~ (1, 1, 1)."""
    IM, JM, KM = 20, 20, 20
    coor = _uniform(IM, JM, KM)
    pos = np.array([[0.25, 0.33, 0.41], [0.33, 0.5, 0.41], [0.5, 0.5, 0.5]])
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos)
    np.testing.assert_array_equal(cell[0], [4, 4, 4])
    assert steps[0] == 5 and st[0] == 1
    np.testing.assert_allclose(w[0], [1.0, 1.0, 1.0], atol=1e-12)
    np.testing.assert_array_equal(cell[1], [6, 9, 8])
    np.testing.assert_array_equal(cell[2], [9, 9, 9])


def test_outside_but_on_another_face_plane_is_accepted(oracle):
    """position >= 0 accepts a cell when some d == 0 even if another d < 0
    (walkingsearch.c:438-455 with :1197-1199): y = 0 exactly lies on the
    BOTTOM plane of every j=0 cell, so the walk stops at its first cell."""
    IM, JM, KM = 10, 10, 10
    coor = _uniform(IM, JM, KM)
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, np.array([[0.35, 0.0, 0.05]]))
    assert st[0] == 1 and tuple(cell[0]) == (0, 0, 0) and steps[0] == 1


def test_face_distance_signs_and_threshold(oracle):
    """Unit cube cell: d = distances to the six faces, enum order LEFT..BACK."""
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 1, 1], [1, 1, 1], [1, 0, 1], [0, 0, 1]], float)
    r, d, pos = oracle.evaluate(V, np.array([0.2, 0.3, 0.4]))
    assert r == 0 and pos == 0
    np.testing.assert_allclose(d, [0.2, 0.8, 0.3, 0.7, 0.6, 0.4], atol=1e-15)
    r, d, pos = oracle.evaluate(V, np.array([1.3, 0.3, -0.4]))
    assert pos == -1 and d[1] < 0 and d[5] < 0
    # inside the 1e-11*h band -> zero
    r, d, pos = oracle.evaluate(V, np.array([0.5e-11, 0.5, 0.5]))
    assert d[0] == 0.0 and pos == 0
    r, d, pos = oracle.evaluate(V, np.array([2e-11, 0.5, 0.5]))
    assert d[0] > 0.0
    assert abs(oracle.char_size(V) - 1.0) < 1e-15


def test_weights_clamp(oracle):
    w = oracle.update_weights([0.0, 0.5, -1.0, 0.25, 0.3, 0.3])
    np.testing.assert_allclose(w, [1e-14 / 0.5, 1e-14 / (1e-14 + 0.25), 0.5], rtol=1e-12)


def _py_walk(oracle, V_of, IM, JM, KM, p, start=(0, 0, 0), cap=1000):
    """Second statement of the walk driver (walkingsearch.c:1123-1212) over the
    oracle's cell evaluation: returns (cell, status, steps, cycled)."""
    i, j, k = start
    seen = {}
    for it in range(1, cap + 1):
        if not (0 <= i < IM and 0 <= j < JM and 0 <= k < KM):
            return None, 3, it, False
        r, d, ps = oracle.evaluate(V_of(i, j, k), p)
        assert r == 0
        if ps >= 0:
            return (i, j, k), 1, it, False
        if (i, j, k) in seen:
            # deterministic map on a finite set: a revisit means a cycle, the
            # cap will be reached
            return None, 3, cap, True
        seen[(i, j, k)] = it
        if d[4] < 0:
            k += 1
        elif d[5] < 0:
            k -= 1
        if d[0] < 0:
            i -= 1
        elif d[1] < 0:
            i += 1
        if d[2] < 0:
            j -= 1
        elif d[3] < 0:
            j += 1
    return None, 3, cap, False


@pytest.mark.parametrize("amp,min_found,max_cycled", [(0.2, 480, 0.02), (0.35, 400, 0.1)])
def test_walk_bruteforce_on_warped_grid(oracle, amp, min_found, max_cycled):
    """Tiny perturbed cube.  (1) An independent Python walk over the oracle's
    cell evaluation reproduces the C walk (cell, status, iterations).  (2) The
    accepted cell always accepts the point; with no accepting cell the walk is
    LOST.  (3) Where exactly one cell accepts the point the walk finds it,
    unless the reference's diagonal step rule cycles (warped faces), which the
    Python walk detects."""
    IM, JM, KM = 4, 4, 6
    coor = W.perturbed_cube(IM, JM, KM, seed=77, amp=amp)
    pos = np.random.default_rng(3).random((600, 3))
    cell, w, st, steps = oracle.locate(IM, JM, KM, coor, pos)
    c3 = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]

    def V_of(i, j, k):
        return np.array([c3[k + dk, j + dj, i + di] for di, dj, dk in off])

    found = cycled = 0
    for p in range(pos.shape[0]):
        pc, pst, pit, cyc = _py_walk(oracle, V_of, IM, JM, KM, pos[p])
        assert pst == st[p] and pit == steps[p]
        if pst == 1:
            assert tuple(cell[p]) == pc
        owners = [(i, j, k) for k in range(KM) for j in range(JM) for i in range(IM)
                  if oracle.evaluate(V_of(i, j, k), pos[p])[2] >= 0]
        if st[p] == 1:
            assert tuple(cell[p]) in owners
        if len(owners) == 0:
            assert st[p] == 3
        if len(owners) == 1:
            if st[p] == 1:
                assert tuple(cell[p]) == owners[0]
                found += 1
            else:
                assert cyc
                cycled += 1
    assert found > min_found and cycled <= max_cycled * (found + cycled)


# --------------------------------------------------------------- interpolation
def test_center_to_corner_linear_field(oracle):
    IM, JM, KM = 6, 5, 7
    coor = _uniform(IM, JM, KM)
    A = np.array([[1.0, 2.0, -1.0], [0.5, -0.25, 3.0], [2.0, 1.0, 0.0]])
    b = np.array([0.1, -0.2, 0.3])
    cc = W.cell_centroids(coor, IM, JM, KM)
    ucat = W.to_node_slots(cc @ A.T + b, IM, JM, KM)
    corner = oracle.center_to_corner(IM, JM, KM, ucat).reshape(KM + 1, JM + 1, IM + 1, 3)
    nodes = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    inner = nodes[1:-1, 1:-1, 1:-1] @ A.T + b
    np.testing.assert_allclose(corner[1:-1, 1:-1, 1:-1], inner, atol=1e-13)
    # a corner node has one cell: its value is that cell's
    np.testing.assert_allclose(corner[0, 0, 0], cc[0, 0, 0] @ A.T + b, atol=1e-14)
    # an edge node between two cells along i: mean of the two
    np.testing.assert_allclose(corner[0, 0, 3], 0.5 * ((cc[0, 0, 2] + cc[0, 0, 3]) @ A.T) + b, atol=1e-13)


def test_trilinear_reproduces_trilinear_field(oracle):
    IM, JM, KM = 5, 6, 4
    coor = _uniform(IM, JM, KM)
    nodes = coor.reshape(-1, 3)
    f = lambda x: np.stack([1 + x[:, 0] + 2 * x[:, 1] * x[:, 2], x[:, 0] * x[:, 1] * x[:, 2] - 3 * x[:, 2],
                            np.ones(len(x)) * 7.0], axis=-1)
    corner = f(nodes)
    pos = np.random.default_rng(9).random((500, 3))
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    vel = oracle.interp(IM, JM, KM, corner, cell, w)
    np.testing.assert_allclose(vel, f(pos), atol=1e-12)
    # LOST particle (cell -1): zero velocity (interpolation.c:1356-1372)
    v0 = oracle.interp(IM, JM, KM, corner, np.array([[-1, -1, -1]]), np.array([[0.5, 0.5, 0.5]]))
    assert (v0 == 0).all()


def test_uniform_field_known_answer(oracle):
    """LOG_INTERPOLATION_ERROR's finit=0 case: (0,0,1) interpolates exactly."""
    d = W.config(1)
    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    cell, w, st, _ = oracle.locate(IM, JM, KM, d["coor"], d["pos"])
    corner = oracle.center_to_corner(IM, JM, KM, d["ucat"])
    vel = oracle.interp(IM, JM, KM, corner, cell, w)
    assert (vel == np.array([0.0, 0.0, 1.0])).all()
    from oracle.oracle import interpolation_error_percent

    assert interpolation_error_percent(d["pos"][st == 1], vel[st == 1], 0) == 0.0


def _sine_error(oracle, m, interior, L=2.0):
    """LOG_INTERPOLATION_ERROR's finit = 1 diagnostic on an m^3 cube of side L."""
    from oracle.oracle import interpolation_error_percent

    coor = W.uniform_grid(m, m, m) * L
    ucat = W.ucat_sine(coor, m, m, m)
    rng = np.random.default_rng(1)
    pos = L * (0.25 + 0.5 * rng.random((20000, 3))) if interior else L * rng.random((20000, 3))
    cell, w, st, _ = oracle.locate(m, m, m, coor, pos)
    vel = oracle.interp(m, m, m, oracle.center_to_corner(m, m, m, ucat), cell, w)
    ok = st == 1
    assert ok.all()
    return interpolation_error_percent(pos[ok], vel[ok], 1)


def test_sine_field_interpolation_error_converges(oracle):
    """LOG_INTERPOLATION_ERROR's finit = 1 case, (sin x, sin y, sin z) at the
    cell centroids: centre->corner averaging plus trilinear interpolation is
    second order inside the domain (the error falls 4x per halving of h), and
    the one-sided corner averages of the boundary layer bring the whole
    domain's L2 error to order 3/2 (2^1.5 = 2.83x).  A corner averaged over
    the wrong cells, or weights taken from the wrong faces, would be first
    order or worse."""
    inner = [_sine_error(oracle, m, True) for m in (8, 16, 32)]
    whole = [_sine_error(oracle, m, False) for m in (8, 16, 32)]
    for e in (inner, whole):
        assert all(a > b > 0 for a, b in zip(e, e[1:]))
    for a, b in zip(inner, inner[1:]):
        assert 3.8 < a / b < 4.2
    for a, b in zip(whole, whole[1:]):
        assert 2.6 < a / b < 3.0
    assert inner[-1] < 0.1  # percent


# ------------------------------------------------------------------ deposition
def test_count_and_P_average(oracle):
    IM, JM, KM = 7, 6, 9
    coor = _uniform(IM, JM, KM)
    pos = np.random.default_rng(2).random((3000, 3))
    pos[:10] += 2.0  # lost
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    cnt = oracle.count_cells(IM, JM, KM, cell)
    assert cnt.sum() == (st == 1).sum() == 2990
    exp = np.zeros((KM + 1, JM + 1, IM + 1))
    for c in cell[st == 1]:
        exp[c[2], c[1], c[0]] += 1
    np.testing.assert_array_equal(cnt, exp.ravel())
    P = np.full(3000, 101325.0)
    avg = oracle.scatter_P(IM, JM, KM, cell, P, np.arange(3000), cnt)
    assert set(np.unique(avg)) <= {0.0, 101325.0}
    assert ((avg == 101325.0) == (cnt > 0)).all()


def test_face_deposition_conservation(oracle):
    IM, JM, KM = 6, 5, 8
    coor = W.perturbed_cube(IM, JM, KM, seed=5)
    pos = np.random.default_rng(4).random((4000, 3))
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, pos)
    vel = np.random.default_rng(8).normal(size=(4000, 3))
    ok = st == 1
    fi, fj, fk = oracle.deposit_faces(IM, JM, KM, cell, w, vel, np.arange(4000))
    for f in (fi, fj, fk):
        f = f.reshape(-1, 4)
        assert abs(f[:, 0].sum() - ok.sum()) < 1e-9
        np.testing.assert_allclose(f[:, 1:].sum(0), vel[ok].sum(0), atol=1e-9)
    # per-row: every particle of row (j,k) deposits weight 1 onto that row's i-faces
    fi4 = fi.reshape(KM, JM, IM + 1, 4)
    rows = np.zeros((KM, JM))
    for c in cell[ok]:
        rows[c[2], c[1]] += 1
    np.testing.assert_allclose(fi4[..., 0].sum(-1), rows, atol=1e-12)
    # single particle: faces i and i+1 get 1-a and a
    fi1, _, _ = oracle.deposit_faces(IM, JM, KM, cell[ok][:1], w[ok][:1], vel[ok][:1], np.arange(1))
    c = cell[ok][0]
    g = fi1.reshape(KM, JM, IM + 1, 4)
    assert abs(g[c[2], c[1], c[0], 0] - (1 - w[ok][0, 0])) < 1e-15
    assert abs(g[c[2], c[1], c[0] + 1, 0] - w[ok][0, 0]) < 1e-15


# ------------------------------------------------------------------ time steps
def _state(pos, vel=None):
    n = pos.shape[0]
    return {"pos": pos.copy(), "vel": np.zeros((n, 3)) if vel is None else vel.copy(),
            "cell": np.full((n, 3), -1, np.int64), "weight": np.zeros((n, 3)),
            "status": np.full(n, 4, np.int64), "pid": np.arange(n, dtype=np.int64), "P": np.full(n, 101325.0)}


def test_euler_step_uniform_flow_config1(oracle):
    d = W.config(1)
    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    coor, ucat = d["coor"], d["ucat"]
    corner = oracle.center_to_corner(IM, JM, KM, ucat)
    s = _state(d["pos"][:2000])
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, s["pos"])
    s.update(cell=cell, weight=w, status=st)
    s["vel"] = oracle.interp(IM, JM, KM, corner, cell, w)
    z0 = dict(zip(s["pid"], s["pos"][:, 2]))
    for step in range(3):
        oracle.step(IM, JM, KM, coor, corner, s, 0.1)
    # survivors moved exactly 0.3 in z; the rest left the unit cube
    expect_alive = sorted(p for p, z in z0.items() if z + 0.1 + 0.1 + 0.1 < 1.0 + 1e-9)
    assert len(s["pid"]) in (len(expect_alive), len(expect_alive) - 1, len(expect_alive) + 1)
    for p, z in zip(s["pid"], s["pos"][:, 2]):
        assert abs(z - (((z0[p] + 0.1) + 0.1) + 0.1)) == 0.0


def test_rk2_order_on_linear_field(oracle):
    """u = (0, 0, alpha z): Heun's one-step error is ~(alpha dt)^3 z / 6, third order."""
    IM, JM, KM = 4, 4, 64
    coor = _uniform(IM, JM, KM, )
    alpha = 0.8
    nodes = coor.reshape(-1, 3)
    corner = np.stack([0 * nodes[:, 0], 0 * nodes[:, 0], alpha * nodes[:, 2]], -1)
    errs = []
    for dt in (0.2, 0.1):
        s = _state(np.array([[0.37, 0.61, 0.3]]))
        cell, w, st, _ = oracle.locate(IM, JM, KM, coor, s["pos"])
        s.update(cell=cell, weight=w, status=st)
        s["vel"] = oracle.interp(IM, JM, KM, corner, cell, w)
        oracle.step(IM, JM, KM, coor, corner, s, dt, rk2=True)
        exact = 0.3 * math.exp(alpha * dt)
        errs.append(abs(s["pos"][0, 2] - exact))
    assert errs[0] < 1e-3
    assert 6.0 < errs[0] / errs[1] < 10.0


def test_rk2_uniform_flow_exact(oracle):
    IM, JM, KM = 4, 4, 40
    coor = _uniform(IM, JM, KM)
    nodes = coor.reshape(-1, 3)
    corner = np.tile([0.1, -0.05, 0.5], (nodes.shape[0], 1))
    s = _state(np.array([[0.3, 0.6, 0.1], [0.7, 0.2, 0.2]]))
    cell, w, st, _ = oracle.locate(IM, JM, KM, coor, s["pos"])
    s.update(cell=cell, weight=w, status=st, vel=oracle.interp(IM, JM, KM, corner, cell, w))
    x0 = s["pos"].copy()
    oracle.step(IM, JM, KM, coor, corner, s, 0.1, rk2=True)
    np.testing.assert_allclose(s["pos"], x0 + 0.1 * np.array([0.1, -0.05, 0.5]), atol=1e-15)


# ------------------------------------------------------------------------- MLP
def _mlp_params(seed, H=64, w1=0.7):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, w1, 3 * H), rng.normal(0, 0.1, H), rng.normal(0, 1 / 8, H * H),
                           rng.normal(0, 0.1, H), rng.normal(0, 1 / 8, 3 * H), rng.normal(0, 0.1, 3)])


@pytest.mark.parametrize("w1", [0.7, 4.0])
def test_mlp_gradient_matches_finite_differences(oracle, w1):
    """Central differences; with first-layer weights of scale 4 about half of
    the first-layer gates are clamped, so both branches of the activation's
    derivative are exercised."""
    p = _mlp_params(1, w1=w1)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (64, 3))
    t = rng.uniform(0, 1, (64, 3))
    loss, g = oracle.mlp_loss_grad(p, x, t, round_bf16=False)
    idx = rng.choice(p.size, 60, replace=False)
    eps = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (oracle.mlp_loss_grad(pp, x, t, False)[0] - oracle.mlp_loss_grad(pm, x, t, False)[0]) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-7 + 1e-5 * abs(fd)
    y = oracle.mlp_forward(p, x, round_bf16=False)
    assert abs(loss - ((y - t) ** 2).mean()) < 1e-14


def test_mlp_forward_by_hand(oracle):
    """One hidden unit through each layer, written out, with the hard GELU's
    gate inside (0, 1) on the way in and saturated at 1 and at 0 on two
    more units."""
    H = 64
    p = np.zeros(oracle.lib.ref_mlp_param_count(H))
    W1, b1 = p[:3 * H].reshape(H, 3), p[3 * H:4 * H]
    off = 4 * H
    W2, b2 = p[off:off + H * H].reshape(H, H), p[off + H * H:off + H * H + H]
    off += H * H + H
    W3, b3 = p[off:off + 3 * H].reshape(3, H), p[off + 3 * H:]
    W1[0] = [1.0, -2.0, 0.5]
    b1[0] = 0.25
    W2[3, 0] = 1.0
    b2[3] = -0.5
    W3[1, 3] = 3.0
    b3[:] = [0.1, 0.2, 0.3]
    b2[5] = 7.0    # z = 7: gate clamped at 1, act = 7
    W3[0, 5] = 0.5
    b2[6] = -7.0   # z = -7: gate clamped at 0, act = 0
    W3[2, 6] = 4.0
    c = float(np.float32(1.0 / math.sqrt(2.0 * math.pi)))  # binary32 value of Phi'(0)
    act = lambda z: z * min(max(c * z + 0.5, 0.0), 1.0)  # GELU, hard form
    x = np.array([0.3, -0.1, 0.7])
    z1 = 0.3 + 0.2 + 0.35 + 0.25
    z2 = act(z1) - 0.5
    assert 0 < c * z1 + 0.5 < 1 and 0 < c * z2 + 0.5 < 1
    y = oracle.mlp_forward(p, x[None], round_bf16=False)[0]
    np.testing.assert_allclose(y, [0.1 + 0.5 * 7.0, 0.2 + 3.0 * act(z2), 0.3], rtol=1e-14)
    assert act(0.0) == 0.0 and act(3.0 / c) == 3.0 / c and act(-3.0 / c) == 0.0
