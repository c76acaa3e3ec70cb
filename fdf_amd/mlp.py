"""Training driver of the learned initial guess (3->64->64->3 MLP on the GPU).

The network is trained on (position, cell) pairs that the walking search
itself produces: a subset of the swarm is located from scratch with the
reference start rule, the located pairs become the training set, and the
trained weights are installed in the target context, whose locate then
starts unlocated particles at the MLP's cell.  The exact walk always
finishes from the guess, so the owning cell is decided by the walk.
"""
from __future__ import annotations

import time

import numpy as np

from .swarm import PicGpu


def train_from_walk(ctx: PicGpu, d: dict, n_pairs: int | None = None, iters: int = 1500, batch: int = 4096,
                    lr: float = 3e-3, seed: int = 3004, rounds: int = 2, kernel_timing: bool = False) -> dict:
    """Train ctx's MLP on pairs located by the walk for the first n_pairs
    particles of workload d (device positions d["pos_dev"] or host d["pos"]).

    Round 1 locates with the reference start rule; on curvilinear grids that
    start (cell (0,0,0)) loses many walks, so each later round re-locates the
    still-LOST particles from the current network's guess and trains again
    on the enlarged set of walk-produced pairs.  With kernel_timing the
    training kernels are timed with HIP events and the result holds their
    totals under "kernels" ({name: (total ms, launches)})."""
    n_pairs = int(n_pairs or d.get("train_pairs") or min(2_097_152, d["n"]))
    tmp = PicGpu(precision=ctx.precision)
    hist = []
    kernels = {}
    try:
        if kernel_timing:
            tmp.timing(True)
            tmp.timing_select("mlp_")
        if "pos_dev" in d:
            tmp.upload_grid_device(d["IM"], d["JM"], d["KM"], d["coor"].data_ptr())
            n_pairs = min(n_pairs, int(d["pos_dev"].shape[0]))
            tmp.init_particles_device(n_pairs, d["pos_dev"].data_ptr())
        else:
            tmp.upload_grid(d["IM"], d["JM"], d["KM"], d["coor"])
            n_pairs = min(n_pairs, d["pos"].shape[0])
            tmp.upload_particles(d["pos"][:n_pairs])
        t0 = time.perf_counter()
        walk_s = train_s = 0.0
        loss = float("nan")
        for r in range(rounds):
            t0 = time.perf_counter()
            tmp.LocateAllParticlesInGrid()  # LOST particles are walked again, from the guess after round 1
            tmp.synchronize()
            st = tmp.stats()
            t1 = time.perf_counter()
            walk_s += t1 - t0
            located = n_pairs - st["lost"] if r > 0 else st["located"]
            loss = tmp.mlp_train(max_pairs=n_pairs, batch=batch, iters=iters, lr=lr, seed=seed + r)
            tmp.synchronize()
            train_s += time.perf_counter() - t1
            tmp.mlp_enable(True)
            hist.append({"round": r, "lost": int(st["lost"]), "final_mse": float(loss)})
            if st["lost"] == 0:
                break
        params = tmp.mlp_get_params()
        if kernel_timing:
            tmp.synchronize()
            for k in ("mlp_fwd_bwd", "mlp_reduce", "mlp_adam", "mlp_guess"):
                kernels[k] = tmp.timing_query(k)
    finally:
        tmp.close()
    ctx.mlp_set_params(params)
    ctx.mlp_enable(True)
    out = {"pairs": int(located), "walk_s": walk_s, "train_s": train_s, "iters_per_round": iters,
           "batch": batch, "final_mse": float(loss), "rounds": hist}
    if kernel_timing:
        out["kernels"] = kernels
    return out


def forward_flops(H: int = 64) -> int:
    """Dense-product FLOPs of one sample's forward pass (3->H->H->3)."""
    return 2 * (3 * H + H * H + H * 3)


def train_flops(H: int = 64) -> int:
    """Dense-product FLOPs of one sample's forward and backward pass: the
    forward, then dW3, dA2, dW2, dA1 and dW1 (the input gradient is not
    formed)."""
    return forward_flops(H) + 2 * (H * 3 + H * 3 + H * H + H * H + 3 * H)


def train_on_swarm(ctx: PicGpu, d: dict, **kw) -> dict:
    return train_from_walk(ctx, d, **kw)


def initial_params(seed: int = 0, H: int = 64) -> np.ndarray:
    """Random parameters in the C-ABI layout (tests)."""
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, 1.0, 3 * H), rng.normal(0, 0.1, H), rng.normal(0, 1 / 8, H * H),
                           rng.normal(0, 0.1, H), rng.normal(0, 1 / 8, 3 * H), rng.normal(0, 0.1, 3)]).astype(
        np.float32)
