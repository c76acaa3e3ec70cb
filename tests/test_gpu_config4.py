"""Config 4 at full size: the 128x128x1024 bent tube and 16,777,216 fresh
positions (uniform in the tube, no cells), located once with the
reference's start rule (cell (0,0,0), walkingsearch.c:603-619) and once
from the learned guess, trained as the bench trains it.  The guess is
scored by the walk iterations it saves; the owners are the walk's:

- from the guess every position is located, and the mean walk is short
  (under 5 iterations) and far shorter than from (0,0,0), which loses many
  walks to MAX_TRAVERSAL on the bent tube;
- on a random sample of 131,072 positions the oracle's walk, started from
  the device's cell, keeps that cell (the device's owner is an owner by the
  reference's classification), with the same weights: bit for bit in FP64,
  within 1e-5 in FP32.
"""
import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu

NSAMPLE = 131_072
FP32_TOL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def cfg4():
    return W.config(4)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config4_full_size_relocation(cfg4, oracle, precision):
    from fdf_amd import mlp as M
    from fdf_amd.swarm import PicGpu

    d = cfg4
    IM, JM, KM, n = d["IM"], d["JM"], d["KM"], d["n"]
    assert (IM, JM, KM, n) == (128, 128, 1024, 16_777_216)
    _say("config4 full size", precision, IM, JM, KM, n)
    with PicGpu(precision=precision) as g:
        g.upload_grid(IM, JM, KM, d["coor"])
        g.upload_ucat(d["ucat"])
        g.upload_particles(d["pos"])
        g.LocateAllParticlesInGrid()
        plain = g.stats()
        M.train_from_walk(g, d, iters=12000, batch=16384, lr=2e-3)
        g.upload_particles(d["pos"])
        g.LocateAllParticlesInGrid()
        st = g.stats()
        out = g.download_particles()
    plain_mean = plain["walk_iterations"] / n
    mean = st["walk_iterations"] / n
    _say("from (0,0,0): lost", plain["lost"], "mean walk", plain_mean, "; from the guess: lost", st["lost"],
         "mean walk", mean)
    assert plain["located"] + plain["lost"] == n and plain["lost"] > 0
    assert st["located"] == n and st["lost"] == 0
    assert mean < 5.0 and plain_mean > 8.0 * mean

    pick = np.random.default_rng(404).choice(n, NSAMPLE, replace=False)
    pos, cell, w = out["pos"][pick], out["cell"][pick], out["weight"][pick]
    assert (out["status"][pick] == L.ACTIVE_AND_LOCATED).all()
    assert np.array_equal(pos, d["pos"][out["pid"][pick]])  # a locate moves nothing
    c2, w2, st2, _ = oracle.locate(IM, JM, KM, d["coor"], pos, cell=cell, status=np.zeros(NSAMPLE, np.int64))
    np.testing.assert_array_equal(st2, L.ACTIVE_AND_LOCATED)
    np.testing.assert_array_equal(c2, cell)
    if precision == "fp64":
        assert np.array_equal(w2, w)
    else:
        assert np.abs(w2 - w).max() < FP32_TOL
