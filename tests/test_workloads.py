"""The synthetic configurations are well-formed grids (right-handed, no
degenerate faces) and the particle generators stay inside them."""
import math

import numpy as np
import pytest

from fdf_amd import workloads as W


def _cell_checks(coor, IM, JM, KM):
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    v0 = c[:-1, :-1, :-1]
    ex = c[:-1, :-1, 1:] - v0
    ey = c[:-1, 1:, :-1] - v0
    ez = c[1:, :-1, :-1] - v0
    jac = np.einsum("...i,...i->...", np.cross(ex, ey), ez)
    # smallest face-normal magnitude |(v2-v1)x(v4-v1)| over the k-faces (walkingsearch.c:147 threshold 1e-12)
    nk = np.linalg.norm(np.cross(c[:, :-1, 1:] - c[:, :-1, :-1], c[:, 1:, :-1] - c[:, :-1, :-1]), axis=-1)
    ni = np.linalg.norm(np.cross(c[:-1, 1:, :] - c[:-1, :-1, :], c[1:, :-1, :] - c[:-1, :-1, :]), axis=-1)
    nj = np.linalg.norm(np.cross(c[:-1, :, 1:] - c[:-1, :, :-1], c[1:, :, :-1] - c[:-1, :, :-1]), axis=-1)
    return jac, min(nk.min(), ni.min(), nj.min())


@pytest.mark.parametrize("maker", ["tube", "cube"])
def test_grids_right_handed_and_nondegenerate(maker):
    if maker == "tube":
        IM, JM, KM = 32, 32, 64
        coor = W.tube_grid(IM, JM, KM)
    else:
        IM, JM, KM = 16, 16, 64
        coor = W.perturbed_cube(IM, JM, KM, seed=3005)
    jac, nmin = _cell_checks(coor, IM, JM, KM)
    assert (jac > 0).all()
    assert nmin > 1e-9


def test_config2_tube_full_size_nondegenerate():
    d = W.config(2, with_particles=False)
    jac, nmin = _cell_checks(d["coor"], d["IM"], d["JM"], d["KM"])
    assert (jac > 0).all() and nmin > 1e-9


def test_tube_particles_inside_analytic_tube():
    p = W.particles_uniform_tube(20000, 7)
    r = np.hypot(p[:, 0], p[:, 1])
    rho = np.hypot(r - W.TUBE_R, p[:, 2])
    th = np.arctan2(p[:, 1], p[:, 0])
    assert rho.max() <= W.RHO_MAX * W.TUBE_A + 1e-12
    assert th.min() >= -W.TUBE_SWEEP - 1e-12 and th.max() <= 1e-12
    # volume element: more particles on the outer side (R + rho cos phi)
    assert (r > W.TUBE_R).mean() > 0.5


def test_random_fourier_field_is_divergence_free_modes():
    ks, amps, _ = W.fourier_modes(2005)
    assert ks.shape == (64, 3)
    assert np.abs((ks * amps).sum(1)).max() < 1e-12
    nk = np.linalg.norm(ks, axis=1)
    assert nk.min() >= 1 and nk.max() <= 8


def test_config_shapes_small():
    for c in (1, 2, 3, 4, 5):
        d = W.config(c, scale=0.05 if c != 1 else 1.0)
        IM, JM, KM = d["IM"], d["JM"], d["KM"]
        assert d["coor"].shape == ((IM + 1) * (JM + 1) * (KM + 1), 3)
        assert d["ucat"].shape == d["coor"].shape
        assert d["pos"].shape == (d["n"], 3)
        assert d["dt"] > 0
