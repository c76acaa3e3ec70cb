"""GPU MLP (learned initial guess) against the CPU FP64 oracle.

The oracle applies the device's rounding rule (BF16 weights and BF16 matmul
inputs a1, a2, g2; FP64 elsewhere) and the same hard-GELU activation, so
the remaining differences come from FP32 accumulation: a rare one-ulp flip
of a BF16-rounded activation moves one output by ~1e-3.  Tolerances are
written per test.
"""
import numpy as np
import pytest

from fdf_amd import mlp as M
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    from fdf_amd.swarm import PicGpu

    g = PicGpu(precision="fp32")
    g.upload_grid(4, 4, 4, W.uniform_grid(4, 4, 4))
    yield g
    g.close()


def _bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_mlp_forward_matches_oracle(ctx, oracle):
    _say("mlp_forward")
    p = M.initial_params(1)
    ctx.mlp_set_params(p)
    x = np.random.default_rng(2).uniform(-1, 1, (4096 + 17, 3)).astype(np.float32)
    y = ctx.mlp_forward(x)
    yo = oracle.mlp_forward(p.astype(np.float64), x.astype(np.float64), round_bf16=True)
    err = np.abs(y - yo)
    _say("forward err median", np.median(err), "max", err.max())
    assert np.median(err) < 1e-5
    assert err.max() < 5e-3
    # the same samples in another order give the same outputs (no cross-sample coupling)
    perm = np.random.default_rng(3).permutation(x.shape[0])
    assert np.array_equal(ctx.mlp_forward(x[perm]), y[perm])


def test_mlp_forward_many_tiles_per_wavefront(ctx, oracle):
    """Above 1024 blocks x 4 wavefronts x 32 = 131,072 samples each wavefront
    of the inference kernel loops over several 32-sample tiles; this size
    gives four or five tiles per wavefront and a ragged last tile."""
    _say("mlp_forward many tiles")
    p = M.initial_params(8)
    ctx.mlp_set_params(p)
    x = np.random.default_rng(9).uniform(-1, 1, (3 * 196_608 + 29, 3)).astype(np.float32)
    y = ctx.mlp_forward(x)
    yo = oracle.mlp_forward(p.astype(np.float64), x.astype(np.float64), round_bf16=True)
    err = np.abs(y - yo)
    _say("forward err median", np.median(err), "max", err.max())
    assert np.median(err) < 1e-5
    assert err.max() < 5e-3
    # samples from late in the waves' loops, shifted to other lanes and tiles
    # in a small launch, give the same outputs
    tail = 4096 + 7
    assert np.array_equal(ctx.mlp_forward(x[-tail:]), y[-tail:])


def test_mlp_gradient_matches_oracle(ctx, oracle):
    _say("mlp_grad")
    p = M.initial_params(4)
    ctx.mlp_set_params(p)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (2048, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (2048, 3)).astype(np.float32)
    loss, g = ctx.mlp_loss_grad(x, t)
    lo, go = oracle.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                  round_bf16=True)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    _say("loss", loss, lo, "grad rel err", rel)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert rel < 2e-3
    # and against the unrounded FP64 gradient the BF16 effect stays small
    lu, gu = oracle.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                  round_bf16=False)
    assert np.linalg.norm(g - gu) / np.linalg.norm(gu) < 5e-2
    # deterministic
    loss2, g2 = ctx.mlp_loss_grad(x, t)
    assert loss2 == loss and np.array_equal(g, g2)


def test_mlp_guess_cells_match_forward(ctx, oracle):
    _say("mlp_guess")
    p = M.initial_params(6)
    ctx.mlp_set_params(p)
    pos = np.random.default_rng(7).random((1000, 3))
    cells = ctx.mlp_guess(pos)
    # the context normalises with the padded grid bounding box
    lo, hi = oracle.bbox(4, 4, 4, W.uniform_grid(4, 4, 4))
    x = ((pos - lo) * (2.0 / (hi - lo)) - 1.0).astype(np.float32)
    y = ctx.mlp_forward(x)
    exp = np.clip(np.floor(y.astype(np.float32) * np.float32(4)), 0, 3).astype(np.int64)
    assert (cells == exp).mean() > 0.999


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_training_saves_walk_steps_and_keeps_ownership(oracle, precision):
    """Bent tube, fresh uniform positions located from scratch: the trained
    guess cuts walk iterations, locates points the (0,0,0) start loses to
    MAX_TRAVERSAL, and owning cells agree with the reference start wherever
    the point has a unique owner."""
    _say("mlp_training", precision)
    from fdf_amd.swarm import PicGpu

    d = W.config(4, scale=0.02, with_particles=False)
    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    train = W.particles_uniform_tube(200_000, 11)
    fresh = W.particles_uniform_tube(50_000, 12)
    res = {}
    with PicGpu(precision=precision, collect_walk_stats=True) as g:
        g.upload_grid(IM, JM, KM, d["coor"])
        info = M.train_from_walk(g, {"IM": IM, "JM": JM, "KM": KM, "coor": d["coor"], "pos": train, "n": 200_000},
                                 iters=1500)
        _say("train", info)
        for use in (False, True):
            g.mlp_enable(use)
            g.upload_particles(fresh)
            g.LocateAllParticlesInGrid()
            out = g.download_particles()
            res[use] = (out, g.walk_steps().astype(np.int64))
    (a, sa), (b, sb) = res[False], res[True]
    oa, ob = np.argsort(a["pid"]), np.argsort(b["pid"])
    ca, cb = a["cell"][oa], b["cell"][ob]
    sta, stb = a["status"][oa], b["status"][ob]
    both = (sta == 1) & (stb == 1)
    _say("mean walk iterations (located by both) without/with guess", sa[both].mean(), sb[both].mean(),
         "lost", (sta == 3).sum(), (stb == 3).sum())
    assert sb[both].mean() * 3 < sa[both].mean()
    assert (stb == 3).sum() <= min((sta == 3).sum(), 0.01 * len(stb))
    # ownership: identical where both located, except points with more than one owner
    diff = np.where(both & (ca != cb).any(1))[0]
    assert len(diff) <= 1e-3 * both.sum()
    c3 = d["coor"].reshape(KM + 1, JM + 1, IM + 1, 3)
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]
    for q in diff:
        p = fresh[q]
        for c in (ca[q], cb[q]):
            V = np.array([c3[c[2] + dk, c[1] + dj, c[0] + di] for di, dj, dk in off])
            assert oracle.evaluate(V, p)[2] >= 0  # both cells accept the point: an overlap


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_guess_locate_grouped_by_start_cell_matches_plain_locate(precision):
    """An FP64 locate of at least 2**20 particles from guessed start cells
    visits them grouped by start cell (FP32 in index order).  Each walk is
    unchanged by the order: cells, weights and statuses equal those of the
    walks from (0,0,0), and the cells are the closed form floor(32 x) of the
    uniform grid."""
    _say("grouped locate", precision)
    from fdf_amd.swarm import PicGpu

    n = (1 << 20) + 4099
    pos = np.random.default_rng(21).random((n, 3))
    res, grouped = {}, {}
    for use in (False, True):
        with PicGpu(precision=precision) as g:
            g.upload_grid(32, 32, 32, W.uniform_grid(32, 32, 32))
            if use:
                g.mlp_set_params(M.initial_params(13))  # an untrained guess: any start cell
                g.mlp_enable(True)
            g.upload_particles(pos)
            g.timing(True)
            g.timing_select("key_histogram")  # the first kernel of the grouping index
            g.LocateAllParticlesInGrid()
            g.synchronize()
            grouped[use] = g.timing_query("key_histogram")[1]
            res[use] = g.download_particles()
    # grouping ran exactly where it is meant to: FP64 with a guess
    assert grouped[False] == 0
    assert grouped[True] == (1 if precision == "fp64" else 0)
    a, b = res[False], res[True]
    oa, ob = np.argsort(a["pid"]), np.argsort(b["pid"])
    for k in ("cell", "weight", "status"):
        assert np.array_equal(a[k][oa], b[k][ob]), k
    assert (b["status"] == 1).all()
    assert np.array_equal(b["pos"][ob], pos)
    assert (b["cell"][ob] == np.floor(pos * 32).astype(np.int64)).all(axis=1).mean() > 1 - 1e-5
