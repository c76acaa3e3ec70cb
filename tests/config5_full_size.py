"""Run as a program by tests/test_gpu_config5.py (one precision mode per run).

Config 5 at full size: the 256x256x1024 perturbed cube with its random
Fourier field and 134,217,728 particles, three RK2 steps with face
deposition, in both precision modes (the bench's workload, and its step
kernel at the bench's launch shapes).  The whole swarm is checked through
properties that hold at any size, and a random sample of 131,072 particles
against the oracle:

- the oracle steps the sampled particles on the full grid from the state
  the device held before the steps (each particle's step depends only on
  its own state, the grid and the field).  FP64: the same survivors, and
  their positions, cells, weights and velocities bit for bit.  FP32: the
  same survivors, positions within 1e-6, and at the device's own positions
  the oracle's walk, started from the device's cell, keeps that cell, with
  weights and velocities there within the FP32 tolerance;
- every survivor is ACTIVE_AND_LOCATED, and each direction's face counts
  add up to the located particles (a particle adds 1 - a and a to its two
  faces).

The initial locate starts from the learned guess, as the bench's FP32 run
does; the guess picks only where each walk starts.
"""
import os
import sys

import numpy as np
import torch  # before the library uses the GPU, as in bench.py (imported after, it finds no GPU)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from fdf_amd import _lib as L  # noqa: E402

FP32_TOL = 1e-5
NSAMPLE = 131_072
NSTEPS = 3


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _config5_on_device():
    from fdf_amd import workloads as W

    d = W.config(5, xp=W.namespace("torch", "cuda"), with_particles=False)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1005)
    d["pos_dev"] = torch.rand((d["n"], 3), dtype=torch.float64, device="cuda", generator=gen)
    d["coor"] = d["coor"].contiguous()
    d["ucat"] = d["ucat"].contiguous()
    return d


def _rows(n, pid):
    """row[p] = the row of PID p in pid (-1 where absent); PIDs lie in [0, n)."""
    assert pid.min() >= 0 and pid.max() < n
    row = np.full(n, -1, np.int64)
    row[pid] = np.arange(pid.shape[0])
    return row


def run(precision):
    from fdf_amd import mlp as M
    from fdf_amd.swarm import PicGpu
    from oracle.oracle import Oracle

    oracle = Oracle()

    d = _config5_on_device()
    IM, JM, KM, n, dt = d["IM"], d["JM"], d["KM"], d["n"], d["dt"]
    assert (IM, JM, KM, n) == (256, 256, 1024, 134_217_728)
    _say("config5 full size", precision, IM, JM, KM, n)
    with PicGpu(precision=precision) as g:
        g.upload_grid_device(IM, JM, KM, d["coor"].data_ptr())
        g.upload_ucat_device(d["ucat"].data_ptr())
        g.init_particles_device(n, d["pos_dev"].data_ptr())
        M.train_on_swarm(g, d)
        g.LocateAllParticlesInGrid()
        g.InterpolateAllFieldsToSwarm()
        st0 = g.stats()
        _say("initial locate: located", st0["located"], "lost", st0["lost"])
        s0 = g.download_particles()
        pick = np.sort(np.random.default_rng(505).choice(s0["pid"].shape[0], NSAMPLE, replace=False))
        s = {k: np.ascontiguousarray(s0[k][pick]) for k in ("pos", "vel", "cell", "weight", "status", "pid")}
        s["P"] = np.full(NSAMPLE, 101325.0)  # the reference's constant P (ParticleSwarm.c:514-518)
        del s0
        g.AdvanceSimulation(dt, NSTEPS, rk2=True, deposit_cells=False, deposit_faces=True)
        st = g.stats()
        faces = [g.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)]
        out = g.download_particles()
    coor = d["coor"].cpu().numpy()
    ucat = d["ucat"].cpu().numpy()
    del d
    torch.cuda.empty_cache()
    located = st["located"]
    _say("after", NSTEPS, "steps: particles", located, "walk iterations per particle-step",
         st["walk_iterations"] / max(1, located))

    # the whole swarm
    assert st["n_particles"] == located == out["pid"].shape[0]
    assert 0.99 * n < located <= n
    assert (out["status"] == L.ACTIVE_AND_LOCATED).all()
    tol = 1e-9 if precision == "fp64" else FP32_TOL
    for f in faces:
        f4 = f.reshape(-1, 4)
        assert (f4[:, 0] >= 0).all()
        assert abs(f4[:, 0].sum() - located) <= tol * located
    del faces

    # the sample, stepped by the oracle from the same start
    sample_pid = s["pid"].copy()
    corner = oracle.center_to_corner(IM, JM, KM, ucat)
    umax = float(np.abs(ucat).max())
    del ucat
    for _ in range(NSTEPS):
        oracle.step(IM, JM, KM, coor, corner, s, dt, rk2=True)
    row = _rows(n, out["pid"])
    dev_surv = np.sort(sample_pid[row[sample_pid] >= 0])
    ref_surv = np.sort(s["pid"])
    _say("sample survivors: device", dev_surv.shape[0], "oracle", ref_surv.shape[0])
    if precision == "fp64":
        np.testing.assert_array_equal(dev_surv, ref_surv)
        r = row[s["pid"]]
        for k in ("pos", "cell", "weight", "vel"):
            assert np.array_equal(out[k][r], s[k]), k
        _say("sample checks passed")
        return
    # FP32: the same survivors but at most a few, near the domain's walls
    assert np.setxor1d(dev_surv, ref_surv).shape[0] <= 13
    common = np.intersect1d(dev_surv, ref_surv)
    r = row[common]
    ro = _rows(n, s["pid"])[common]
    assert np.abs(out["pos"][r] - s["pos"][ro]).max() < 1e-6
    c2, w2, st2, _ = oracle.locate(IM, JM, KM, coor, out["pos"][r], cell=out["cell"][r],
                                   status=np.zeros(r.shape[0], np.int64))
    np.testing.assert_array_equal(st2, L.ACTIVE_AND_LOCATED)
    np.testing.assert_array_equal(c2, out["cell"][r])
    assert np.abs(w2 - out["weight"][r]).max() < FP32_TOL
    v2 = oracle.interp(IM, JM, KM, corner, c2, w2)
    assert np.abs(v2 - out["vel"][r]).max() <= FP32_TOL * umax
    _say("sample checks passed")


if __name__ == "__main__":
    run(sys.argv[1])
    print("ok", flush=True)
