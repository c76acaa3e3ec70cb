"""Writes the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py [fixture ...]

The fixtures are small seeded cases (inputs and the oracle's outputs) that
pin the oracle against regressions and give the GPU tests fixed expected
values: the reference ships no usable golden vectors for this path, so the
expected outputs are the oracle's.  Every array is stored in an .npz file
(no pickled objects).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fdf_amd import workloads as W  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def _save(name: str, **arrays) -> None:
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, sum(a.nbytes for a in arrays.values()), "bytes uncompressed")


def walk_uniform_cfg1(o: Oracle) -> None:
    """Config 1 grid (20x20x280 unit cube), 2000 positions located from
    scratch (start (0,0,0)): cells, statuses, weights, walk iterations."""
    IM, JM, KM = 20, 20, 280
    coor = W.uniform_grid(IM, JM, KM)
    pos = np.random.default_rng(1001).random((2000, 3))
    pos[:40] = pos[:40] * 1.2 - 0.1  # a few outside the domain
    cell, w, st, steps = o.locate(IM, JM, KM, coor, pos)
    _save("walk_uniform_cfg1.npz", dims=np.array([IM, JM, KM]), pos=pos, cell=cell, weight=w, status=st,
          steps=steps)


def walk_warped_small(o: Oracle) -> None:
    """8x8x16 cube with interior nodes displaced by 0.35 cell widths: wedge
    gaps, overlaps and cycling walks.  Positions located from scratch and
    again from a neighbouring start cell."""
    IM, JM, KM = 8, 8, 16
    coor = W.perturbed_cube(IM, JM, KM, seed=12, amp=0.35)
    pos = np.random.default_rng(1002).random((2000, 3))
    cell, w, st, steps = o.locate(IM, JM, KM, coor, pos)
    start = np.clip(cell + np.array([1, -1, 2]), 0, [IM - 1, JM - 1, KM - 1])
    start[st != 1] = -1
    cell2, w2, st2, steps2 = o.locate(IM, JM, KM, coor, pos, cell=start, status=np.zeros(len(pos), np.int64))
    _save("walk_warped_small.npz", dims=np.array([IM, JM, KM]), coor=coor, pos=pos, cell=cell, weight=w,
          status=st, steps=steps, start=start, cell2=cell2, weight2=w2, status2=st2, steps2=steps2)


VP = 1.0e-3  # particle volume of the face volume fractions


def interp_and_deposit_small(o: Oracle) -> None:
    """On the same warped grid: a random ucat -> corner velocities, trilinear
    interpolation at the located particles, cell counts, P average, the face
    deposits (count + momentum) and the face volume fractions (particle
    volume VP)."""
    g = np.load(os.path.join(HERE, "walk_warped_small.npz"))
    IM, JM, KM = (int(v) for v in g["dims"])
    ucat = np.zeros(((KM + 1) * (JM + 1) * (IM + 1), 3))
    rng = np.random.default_rng(1003)
    u = ucat.reshape(KM + 1, JM + 1, IM + 1, 3)
    u[:KM, :JM, :IM] = rng.normal(0.0, 1.0, (KM, JM, IM, 3))
    corner = o.center_to_corner(IM, JM, KM, ucat)
    cell, w = g["cell"], g["weight"]
    vel = o.interp(IM, JM, KM, corner, cell, w)
    pid = rng.permutation(len(cell)).astype(np.int64)
    count = o.count_cells(IM, JM, KM, cell)
    P = np.full(len(cell), 101325.0)
    pavg = o.scatter_P(IM, JM, KM, cell, P, pid, count)
    fi, fj, fk = o.deposit_faces(IM, JM, KM, cell, w, vel, pid)
    vf = [o.face_volfrac(IM, JM, KM, g["coor"], d, f, VP) for d, f in enumerate((fi, fj, fk))]
    _save("interp_deposit_warped_small.npz", ucat=ucat, corner=corner, vel=vel, pid=pid, count=count, pavg=pavg,
          fi=fi, fj=fj, fk=fk, vp=np.array(VP), vfi=vf[0], vfj=vf[1], vfk=vf[2])


def steps_rk2_small(o: Oracle) -> None:
    """Three RK2 steps (Heun, locate after each position update, removal of
    LOST and out-of-box particles) on the warped grid with the same ucat."""
    g = np.load(os.path.join(HERE, "walk_warped_small.npz"))
    f = np.load(os.path.join(HERE, "interp_deposit_warped_small.npz"))
    IM, JM, KM = (int(v) for v in g["dims"])
    coor, corner = g["coor"], f["corner"]
    ok = g["status"] == 1
    s = {"pos": g["pos"][ok], "vel": f["vel"][ok], "cell": g["cell"][ok], "weight": g["weight"][ok],
         "status": g["status"][ok], "pid": f["pid"][ok], "P": np.full(int(ok.sum()), 101325.0)}
    init = {k: v.copy() for k, v in s.items()}
    dt = 0.02
    for _ in range(3):
        o.step(IM, JM, KM, coor, corner, s, dt, rk2=True)
    _save("steps_rk2_warped_small.npz", dt=np.array(dt), **{"in_" + k: v for k, v in init.items()},
          **{"out_" + k: v for k, v in s.items()})


def mlp_small(o: Oracle) -> None:
    """MLP forward and loss gradient (BF16 rounding of the matmul operands as
    on the device, and without) for fixed parameters and inputs."""
    from fdf_amd.mlp import initial_params

    p = initial_params(7).astype(np.float64)
    rng = np.random.default_rng(1004)
    x = rng.uniform(-1, 1, (64, 3))
    t = rng.uniform(0, 1, (64, 3))
    y = o.mlp_forward(p, x, round_bf16=True)
    loss, grad = o.mlp_loss_grad(p, x, t, round_bf16=True)
    loss0, grad0 = o.mlp_loss_grad(p, x, t, round_bf16=False)
    _save("mlp_small.npz", params=p, x=x, t=t, y=y, loss=np.array(loss), grad=grad, loss_fp64=np.array(loss0),
          grad_fp64=grad0)


FIXTURES = {"walk_uniform_cfg1": walk_uniform_cfg1, "walk_warped_small": walk_warped_small,
            "interp_deposit_warped_small": interp_and_deposit_small, "steps_rk2_warped_small": steps_rk2_small,
            "mlp_small": mlp_small}


def main() -> None:
    """All fixtures, or the ones named on the command line (later fixtures
    read the walk fixture, so it is written first when named)."""
    o = Oracle()
    names = sys.argv[1:] or list(FIXTURES)
    for name in FIXTURES:
        if name in names:
            FIXTURES[name](o)


if __name__ == "__main__":
    main()
