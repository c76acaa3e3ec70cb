#!/usr/bin/env python3
"""Particle-steps/s of PICurv's per-step particle loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 5] [--scale 1.0]

A step is one pass of the loop over the whole swarm: reset, advect (RK2 or
Euler) with the walking-search locate, removal, centre->corner +
trilinear interpolation, sort by cell and deposition.  Inputs (grid, field,
particles) are synthesised on the GPU and resident in HBM before timing.
Two variants run back to back on the same inputs:
  fp64      plain FP64, the reference arithmetic (bit-exact cells/weights)
  fp32_mlp  FP32 weights/interpolation/advection/deposition, FP64 positions
            and near-face decisions, MLP start cell for unlocated particles
`value` is the fp32_mlp throughput.  One JSON line is printed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/sec (locate+interp+advect+deposit), FP32+MLP vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
BF16_DENSE_PEAK_TFLOPS = 2500.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel: str, n: int, g: dict, rsize: int, deposit_share: float = 0.0) -> float:
    """Compulsory HBM bytes of one launch of `kernel` over n particles.

    The fused step moves the particle state once in and once out (S bytes
    each way: pos f64 24 + packed cell 4 + status 4, plus the FP64 velocity
    24 in the FP64 variant) and reads each node's coordinates (24 B) and
    corner velocity (3 reals) once.  Every launch of a call but the first
    also deposits its particles' start state onto the faces (count and
    momentum, 4 reals per face): deposit_share is the fraction of launches
    that do.  Grid-record caches, sort scratch and re-reads are
    implementation choices and are not counted."""
    Nn, Nc = g["nnode"], g["ncell"]
    S = 32 if rsize == 4 else 56
    if kernel.startswith("step_fused"):
        return n * 2 * S + Nn * (24 + 3 * rsize) + deposit_share * g["nface"] * 4 * rsize
    if kernel.startswith("step_finish"):
        # position and cell read, weights and velocity written, faces written
        return n * (24 + 4 + 6 * rsize) + Nn * (24 + 3 * rsize) + g["nface"] * 4 * rsize
    if kernel == "deposit_faces":
        # weights and velocity of each particle once; count + 3 momentum per face
        return n * 6 * rsize + g["nface"] * 4 * rsize
    if kernel == "gather_swarm":
        rec = 24 + 3 * rsize + 3 * rsize + 4 + 4 + 8 + rsize
        return n * (2 * rec + 4)
    if kernel == "key_reindex":
        return n * (4 + 4 + 4 + 4)
    if kernel.startswith("locate"):
        return n * (24 + 4 + 4 + 4 + 4 + 3 * rsize) + Nn * 24
    if kernel == "interp":
        return n * (4 + 3 * rsize + 3 * rsize) + Nn * 3 * rsize
    if kernel == "center_to_corner":
        return Nc * 3 * rsize + Nn * 3 * rsize
    return float("nan")


def design_bytes(kernel: str, n: int, g: dict, rsize: int, deposit_share: float = 0.0) -> float:
    """Bytes the implementation itself moves per launch of the fused step:
    particle fields (in: index entry, cell, position; out: position, cell,
    slot), one face-plane record per cell (128 B FP32, 304 B FP64), the
    staged corner velocities and, on the launches that deposit, the faces
    and the seam halves between tiles (about half a face per cell)."""
    if not kernel.startswith("step_fused"):
        return float("nan")
    rec = 128 if rsize == 4 else 304
    seams = g["ncell"] * 0.5 * 4 * rsize
    return n * ((4 + 4 + 24) + (24 + 4 + 4)) + g["ncell"] * rec + g["nnode"] * 3 * rsize + \
        deposit_share * (g["nface"] * 4 * rsize + seams)


def step_bytes(n: int, g: dict, rsize: int, faces: bool) -> float:
    """Per-step algorithmic bytes B = N*2S + Nn*(24+c) + Ncell*c + Nn*c + D."""
    S = 32 if rsize == 4 else 56
    c = 3 * rsize
    D = g["nface"] * 4 * rsize if faces else g["ncell"] * 2 * rsize
    return n * 2 * S + g["nnode"] * (24 + c) + g["ncell"] * c + g["nnode"] * c + D


def measured_traffic(workload: str, variant: str, kernel: str, later_share: float) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summaries (the L2's memory requests counted by size: 32, 64 and 128 B),
    when one matches this workload and variant; otherwise None.  The
    summaries give the first launch of a call and the later ones (which also
    deposit) separately; later_share is the fraction of later launches."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.txt"))):
        kv = {}
        with open(path) as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        if kv.get("workload") == workload and kv.get("variant") == variant and kv.get("kernel") == kernel:
            try:
                first = float(kv["traffic_bytes_first_launch"])
                later = float(kv["traffic_bytes_later_launch"])
            except (KeyError, ValueError):
                return None
            return (1.0 - later_share) * first + later_share * later
    return None


def calibration_copy(nbytes: int = 1 << 30) -> float:
    """A plain device copy of known size, so that counter runs of this command
    can check their byte counts against it; then five more, timed with
    events, for the measured copy bandwidth (bytes read + written per second,
    GB/s), the roofline's second denominator."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    bw = 2 * nbytes / (e0.elapsed_time(e1) / 5 * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    log(f"calibration copy: {nbytes} bytes read, {nbytes} bytes written; measured copy bandwidth {bw:.0f} GB/s")
    return bw


def make_inputs(cfg: int, scale: float, device: str):
    import torch

    from fdf_amd import workloads as W

    xp = W.namespace("torch", device)
    d = W.config(cfg, scale=scale, xp=xp, with_particles=False)
    n = d["n"]
    gen = torch.Generator(device=device)
    gen.manual_seed(1000 + cfg)
    if d["particles"] == "cube":
        pos = torch.rand((n, 3), dtype=torch.float64, device=device, generator=gen)
    else:
        pos = torch.as_tensor(W.particles_uniform_tube(n, 1000 + cfg), device=device)
    d["pos_dev"] = pos.contiguous()
    d["coor"] = d["coor"].contiguous()
    d["ucat"] = d["ucat"].contiguous()
    return d


def run_variant(name: str, d: dict, steps: int, warmup: int, mlp: dict | None, call_steps: int = 0):
    import torch

    from fdf_amd.swarm import PicGpu

    precision = "fp64" if name == "fp64" else "fp32"
    use_mlp = name.endswith("mlp") and mlp is not None
    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    g = PicGpu(precision=precision, use_mlp_guess=False)
    g.upload_grid_device(IM, JM, KM, d["coor"].data_ptr())
    g.upload_ucat_device(d["ucat"].data_ptr())
    fields = "ucont" in d
    if fields:
        # Ucat comes from the field preparation: metrics, zero-gradient faces,
        # the synthetic field's Ucont, then one preparation (boundaries,
        # Contra2Cart, boundaries); every step repeats the preparation
        from fdf_amd.boundary import BoundaryConfig

        g.compute_metrics()
        bc = BoundaryConfig()
        for f in range(6):
            bc.set_face(f, "NOGRAD", "allcopy")
        g.BoundarySystem_Create(bc)
        g.upload_ucont_device(d["ucont"].data_ptr())
        g.SetEulerianFields(1, 0)
    n0 = int(d["pos_dev"].shape[0])
    g.init_particles_device(n0, d["pos_dev"].data_ptr())
    torch.cuda.synchronize()
    res = {"variant": name, "precision": precision, "mlp": bool(use_mlp)}
    if use_mlp:
        res["mlp_train"] = mlp["train"](g)
        # the guess kernel of the initial locate is timed with events (one launch)
        g.timing(True)
        g.timing_select("mlp_guess")
        g.timing_reset()
    t0 = time.perf_counter()
    g.LocateAllParticlesInGrid()
    st0 = g.stats()
    g.InterpolateAllFieldsToSwarm()
    torch.cuda.synchronize()
    res["initial_locate_s"] = time.perf_counter() - t0
    if use_mlp:
        res["mlp_mfma"] = mlp_mfma(res["mlp_train"], g.timing_query("mlp_guess"), n0)
        g.timing(False)
        g.timing_select("")
    res["initial_located"], res["initial_lost"] = st0["located"], st0["lost"]
    res["initial_walk_iterations_per_particle"] = st0["walk_iterations"] / max(1, n0)
    log(f"{name}: initial locate {res['initial_locate_s']:.2f}s located={st0['located']} lost={st0['lost']}")
    flags = dict(rk2=d["rk2"], deposit_cells=("cells" in d["deposit"]), deposit_faces=("faces" in d["deposit"]),
                 fields=fields)
    for _ in range(warmup):
        g.AdvanceSimulation(d["dt"], 1, **flags)
    torch.cuda.synchronize()
    # HIP events around the step kernel only: each timed launch records two
    # events on the stream, and events around every launch would lengthen the
    # step they measure
    g.timing(True)
    g.timing_select("step_fused")
    g.timing_reset()
    ps0 = g.stats()["particle_steps"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if call_steps > 0:  # the timed steps as calls of call_steps steps each
        for s0 in range(0, steps, call_steps):
            g.AdvanceSimulation(d["dt"], min(call_steps, steps - s0), **flags)
    else:
        g.AdvanceSimulation(d["dt"], steps, **flags)
    g.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = g.stats()
    res["seconds"] = t1 - t0
    res["steps_per_call"] = call_steps if call_steps > 0 else steps
    res["particle_steps"] = st["particle_steps"] - ps0
    res["value"] = res["particle_steps"] / res["seconds"]
    res["ms_per_step"] = 1e3 * res["seconds"] / max(1, steps)
    res["n_end"] = st["n_particles"]
    res["walk_iterations_last_locate_per_particle"] = st["walk_iterations"] / max(1, st["n_particles"])
    res["fp64_fallback_faces_last_locate"] = st["fp64_fallbacks"]
    timed = {}
    for k in ("step_fused_fp64", "step_fused_fp32"):
        ms, cnt = g.timing_query(k)
        if cnt:
            timed[k] = {"total_ms": ms, "launches": cnt, "avg_ms": ms / cnt}
    res["dominant_kernel_timed"] = timed
    # per-kernel breakdown: more steps with every launch timed (not part of
    # the measured rate)
    g.timing_select("")
    g.timing_reset()
    res["breakdown_steps"] = min(steps, 16)
    g.AdvanceSimulation(d["dt"], res["breakdown_steps"], **flags)
    g.synchronize()
    kernels = {}
    for k in ("step_fused_fp64", "step_fused_fp32", "step_finish_fp64", "step_finish_fp32", "face_seams",
              "step_list_fp64", "step_list_fp32", "key_histogram",
              "key_scatter", "key_reindex", "step_deferred", "cell_planes", "bucket_pid_order", "iota", "gather_swarm", "locate_fp64", "locate_fp32",
              "interp", "center_to_corner", "rk2_predict", "rk2_correct", "keep_flags", "compact", "cell_histogram",
              "sort_scatter", "scan_reduce", "scan_tile", "scatter_P", "deposit_faces", "reset_advect", "compact_tail",
              "keep_unlocated", "contra2cart", "bc_face", "flag_fill"):
        ms, cnt = g.timing_query(k)
        if cnt:
            kernels[k] = {"total_ms": ms, "launches": cnt, "avg_ms": ms / cnt}
    res["kernels"] = kernels
    if fields and "contra2cart" in kernels:
        # Contra2Cart per interior cell: node coordinates and Ucont read
        # (2 x 24 B; the face vectors are functions of the coordinates), Ucat
        # written (24 B) and its cell-linear copy (3 reals)
        rs = 8 if precision == "fp64" else 4
        ncell_in = (IM - 1) * (JM - 1) * (KM - 1)
        byts = ncell_in * (2 * 24 + 24 + 3 * rs)
        avg = kernels["contra2cart"]["avg_ms"]
        res["fields"] = {"contra2cart_avg_ms": avg, "algorithmic_bytes": byts,
                         "achieved_GBs": byts / (avg * 1e-3) / 1e9,
                         "frac": byts / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "cells_per_s": ncell_in / (avg * 1e-3),
                         "bc_face_ms_per_step": kernels.get("bc_face", {}).get("total_ms", 0.0) /
                         max(1, res["breakdown_steps"])}
    g.timing(False)
    # walk iterations per particle-step, from one more (untimed) step with statistics on
    g.walk_stats(True)
    g.AdvanceSimulation(d["dt"], 1, **flags)
    st = g.stats()
    res["walk_iterations_per_particle_step"] = walk_percentiles(st["walk_hist"])
    res["walk_iterations_per_particle_step"]["mean"] = st["walk_iterations"] / max(1, int(st["walk_hist"].sum()))
    g.walk_stats(False)
    if d.get("io_dir"):
        # output files of one step (PETSc binary Vec files), then the restart
        # read of the same files; bytes = file payloads
        n_now = g.n
        nn = (IM + 1) * (JM + 1) * (KM + 1)
        byts = 2 * 24 * n_now + (2 * 24 + 3 * 8) * nn
        t0 = time.perf_counter()
        g.WriteOutput(1, d["io_dir"])
        t1 = time.perf_counter()
        g.ReadRestart(1, d["io_dir"])
        t2 = time.perf_counter()
        res["io"] = {"bytes": byts, "write_s": t1 - t0, "read_s": t2 - t1,
                     "write_GBs": byts / (t1 - t0) / 1e9, "read_GBs": byts / (t2 - t1) / 1e9,
                     "directory": d["io_dir"]}
        log(f"{name}: output {byts / 1e9:.2f} GB written in {t1 - t0:.2f} s, read back in {t2 - t1:.2f} s")
    g.close()
    log(f"{name}: {res['value']:.4g} particle-steps/s, {res['ms_per_step']:.3f} ms/step")
    return res


def walk_percentiles(hist) -> dict:
    """Median and 99th percentile of a walk-iteration histogram (last bin: 63 or
    more, reported as 63), with the histogram's mean."""
    import numpy as np

    h = np.asarray(hist, np.float64)
    n = h.sum()
    if n <= 0:
        return {"mean": None, "p50": None, "p99": None, "particles": 0}
    c = np.cumsum(h) / n
    b = np.arange(len(h))
    return {"mean": float((h * b).sum() / n), "p50": int(np.searchsorted(c, 0.5)),
            "p99": int(np.searchsorted(c, 0.99)), "max": int(b[h > 0].max()), "particles": int(n)}


def precision_errors(d: dict, sample: int) -> dict:
    """FP32-mode error against FP64 for one step from the same state.

    The FP64 mode is bit-exact with the FP64 oracle (tests/), so it stands in
    for the oracle at this size.  Both contexts start from the FP64 located
    state of the first `sample` particles on the full grid and take one step
    of the workload; particles are matched by PID."""
    import numpy as np

    from fdf_amd import _lib as L
    from fdf_amd.swarm import PicGpu

    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    pos = d["pos_dev"][:sample].contiguous()
    flags = dict(rk2=d["rk2"], deposit_cells=False, deposit_faces=True)
    out = {}
    with PicGpu(precision="fp64") as g64:
        g64.upload_grid_device(IM, JM, KM, d["coor"].data_ptr())
        g64.upload_ucat_device(d["ucat"].data_ptr())
        g64.init_particles_device(sample, pos.data_ptr())
        g64.LocateAllParticlesInGrid()
        g64.InterpolateAllFieldsToSwarm()
        s0 = g64.download_particles()
        g64.AdvanceSimulation(d["dt"], 1, **flags)
        out["fp64"] = (g64.download_particles(), [g64.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)])
    with PicGpu(precision="fp32") as g32:
        g32.upload_grid_device(IM, JM, KM, d["coor"].data_ptr())
        g32.upload_ucat_device(d["ucat"].data_ptr())
        g32.upload_particles(s0["pos"], vel=s0["vel"], cell=s0["cell"], weight=s0["weight"], status=s0["status"],
                             pid=s0["pid"])
        g32.AdvanceSimulation(d["dt"], 1, **flags)
        out["fp32"] = (g32.download_particles(), [g32.field(f) for f in (L.FIELD_IFACE, L.FIELD_JFACE, L.FIELD_KFACE)])
    (a, fa), (b, fb) = out["fp32"], out["fp64"]
    common, ia, ib = np.intersect1d(a["pid"], b["pid"], return_indices=True)
    scale = float(d["ucat"].abs().max().item())
    same = (a["cell"][ia] == b["cell"][ib]).all(1)
    ok = same
    cnt = np.concatenate([f.reshape(-1, 4)[:, 0] for f in fb])
    mom = np.concatenate([f.reshape(-1, 4)[:, 1:].ravel() for f in fb])
    cnt32 = np.concatenate([f.reshape(-1, 4)[:, 0] for f in fa])
    mom32 = np.concatenate([f.reshape(-1, 4)[:, 1:].ravel() for f in fa])
    return {
        "sample": f"first {sample} particles of the workload on the full grid, one step from the FP64 located state",
        "reference": "FP64 mode (bit-exact with the FP64 oracle)",
        "particles_fp64": int(len(b["pid"])), "particles_fp32": int(len(a["pid"])),
        "cell_mismatch": int((~same).sum()),
        "weight_max_abs": float(np.abs(a["weight"][ia][ok] - b["weight"][ib][ok]).max()),
        "velocity_max_rel": float(np.abs(a["vel"][ia][ok] - b["vel"][ib][ok]).max() / scale),
        "position_max_abs": float(np.abs(a["pos"][ia] - b["pos"][ib]).max()),
        "face_count_max_rel": float(np.abs(cnt32 - cnt).max() / max(1.0, np.abs(cnt).max())),
        "face_momentum_max_rel": float(np.abs(mom32 - mom).max() / (scale * max(1.0, np.abs(cnt).max()))),
    }


def mlp_mfma(train: dict, guess: tuple, n: int) -> dict:
    """FLOP rates of the MLP kernels against the BF16 dense MFMA peak: the
    training forward+backward (batch samples per launch) and the start-cell
    guess of the initial locate (n samples, one launch).  FLOPs are the
    network's dense products, whether they run on MFMA or the VALU."""
    from fdf_amd import mlp as M

    out = {"peak": BF16_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s"}
    rows = []
    fb = train.get("kernels", {}).get("mlp_fwd_bwd")
    if fb and fb[1]:
        rows.append(("train_fwd_bwd", "k_mlp_fwd_bwd", train["batch"] * M.train_flops(), fb[0] / fb[1], fb[1]))
    if guess[1]:
        rows.append(("guess", "k_mlp_infer<true>", n * M.forward_flops(), guess[0] / guess[1], guess[1]))
    for name, kern, flops, avg_ms, cnt in rows:
        tf = flops / (avg_ms * 1e-3) / 1e12
        out[name] = {"kernel": kern, "flops_per_launch": flops, "avg_ms": avg_ms, "launches": cnt,
                     "achieved": tf, "frac": tf / BF16_DENSE_PEAK_TFLOPS}
    return out


def roofline_for(res: dict, d: dict, g: dict) -> dict:
    rsize = 8 if res["precision"] == "fp64" else 4
    kern = res["dominant_kernel_timed"]  # the step kernel, timed over the timed steps
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    n_avg = res["particle_steps"] / max(1, res["steps"])
    faces = "faces" in d["deposit"] and "ucont" not in d  # per-step field preparation: a pass of its own
    # every launch but the first of a call deposits the state it starts from
    spc = min(res.get("steps_per_call", res["steps"]), res["steps"])
    share = (spc - 1) / spc if faces and spc > 0 else 0.0
    b = algorithmic_bytes(dom, int(n_avg), g, rsize, share)
    achieved = b / (kern[dom]["avg_ms"] * 1e-3) / 1e9
    return {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(d["workload_name"], res["variant"], dom, share),
            "algorithmic_bytes_per_launch": b,
            "design_bytes_per_launch": design_bytes(dom, int(n_avg), g, rsize, share),
            "avg_launch_ms": kern[dom]["avg_ms"], "particles_per_launch": n_avg,
            "share_of_step_time": kern[dom]["total_ms"] / (1e3 * res["seconds"])}


def _host_cpus() -> dict:
    """CPUs this process may run on, the machine's count and the CPU model
    (what lscpu prints as "Model name", read from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"nproc": avail, "cpu_count": os.cpu_count(), "model": model}


def _sub_block(d: dict, dims) -> dict:
    """A block of ni x nj x nk cells from the middle of the workload's grid, its
    field and the workload's particles inside the block's lattice box (the
    same density as the whole workload), all on the host."""
    import torch

    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    ni, nj, nk = (min(a, b) for a, b in zip(dims, (IM, JM, KM)))
    i0, j0, k0 = (IM - ni) // 2, (JM - nj) // 2, (KM - nk) // 2
    if (ni, nj, nk) == (IM, JM, KM):  # the whole workload (any grid shape, every particle)
        return {"IM": IM, "JM": JM, "KM": KM, "origin": (0, 0, 0), "coor": d["coor"].cpu().numpy().reshape(-1, 3),
                "ucat": d["ucat"].cpu().numpy().reshape(-1, 3), "pos": d["pos_dev"].cpu().numpy()}
    c = d["coor"].reshape(KM + 1, JM + 1, IM + 1, 3)[k0:k0 + nk + 1, j0:j0 + nj + 1, i0:i0 + ni + 1]
    u = torch.zeros((nk + 1, nj + 1, ni + 1, 3), dtype=torch.float64, device=d["ucat"].device)
    u[:nk, :nj, :ni] = d["ucat"].reshape(KM + 1, JM + 1, IM + 1, 3)[k0:k0 + nk, j0:j0 + nj, i0:i0 + ni]
    lo = torch.tensor([i0 / IM, j0 / JM, k0 / KM], dtype=torch.float64, device=d["pos_dev"].device)
    hi = torch.tensor([(i0 + ni) / IM, (j0 + nj) / JM, (k0 + nk) / KM], dtype=torch.float64,
                      device=d["pos_dev"].device)
    inside = ((d["pos_dev"] >= lo) & (d["pos_dev"] < hi)).all(1)
    return {"IM": ni, "JM": nj, "KM": nk, "origin": (i0, j0, k0), "coor": c.contiguous().cpu().numpy().reshape(-1, 3),
            "ucat": u.cpu().numpy().reshape(-1, 3), "pos": d["pos_dev"][inside].cpu().numpy()}


def _cpu_rate(o, blk: dict, d: dict, threads: int, runs: int, steps: int) -> dict:
    """Median over `runs` runs of the oracle's per-step loop on a sub-block:
    one untimed warm-up step, then `steps` timed steps, each the workload's
    step (centre->corner, reset, advection with the walking-search locate,
    removal, interpolation, deposition)."""
    import ctypes

    import numpy as np

    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(threads))
    IM, JM, KM, coor = blk["IM"], blk["JM"], blk["KM"], blk["coor"]
    bbox = o.bbox(IM, JM, KM, coor)
    corner = o.center_to_corner(IM, JM, KM, blk["ucat"])
    n = blk["pos"].shape[0]
    cell, w, st, _ = o.locate(IM, JM, KM, coor, blk["pos"])
    s0 = {"pos": blk["pos"].copy(), "vel": o.interp(IM, JM, KM, corner, cell, w), "cell": cell, "weight": w,
          "status": st, "pid": np.arange(n, dtype=np.int64), "P": np.full(n, 101325.0)}
    faces = "faces" in d["deposit"]
    cells = "cells" in d["deposit"] or not faces

    def one_step(s):
        c2 = o.center_to_corner(IM, JM, KM, blk["ucat"])
        o.step(IM, JM, KM, coor, c2, s, d["dt"], rk2=d["rk2"], bbox=bbox)
        if cells:
            cnt = o.count_cells(IM, JM, KM, s["cell"])
            o.scatter_P(IM, JM, KM, s["cell"], s["P"], s["pid"], cnt)
        if faces:
            o.deposit_faces(IM, JM, KM, s["cell"], s["weight"], s["vel"], s["pid"])

    rates, secs = [], []
    for _ in range(runs):
        s = {k: v.copy() for k, v in s0.items()}
        one_step(s)  # warm-up
        ps = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            ps += s["pos"].shape[0]
            one_step(s)
        dt = time.perf_counter() - t0
        rates.append(ps / dt)
        secs.append(dt)
    return {"value": float(np.median(rates)), "runs": [float(r) for r in rates], "seconds": float(sum(secs)),
            "threads": int(threads), "cells": [IM, JM, KM], "particles": int(n)}


def cpu_baseline(d: dict, all_block=(128, 128, 64), one_block=(128, 128, 64), runs: int = 3,
                 steps: int = 3) -> dict:
    """The FP64 oracle built -O3 with OpenMP, timed on the host cores on a
    bounded sample of the same workload: a block of cells from the middle of
    the grid with the workload's particles inside it (same density, field,
    dt, advection and deposition).  Grid work and particle work keep their
    full-size proportions, so the rate (particle-steps/s) is the full
    workload's on these cores.  Median of `runs` runs of `steps` steps after
    a warm-up step, with every available core and with one."""
    from oracle.oracle import Oracle

    host = _host_cpus()
    threads = max(1, min(16, host["nproc"]))
    o = Oracle(fast=True)
    blk_all = _sub_block(d, all_block)
    blk_one = _sub_block(d, one_block) if tuple(one_block) != tuple(all_block) else blk_all
    r_all = _cpu_rate(o, blk_all, d, threads, runs, steps)
    r_one = _cpu_rate(o, blk_one, d, 1, runs, steps)

    def desc(r):
        if list(r["cells"]) == [d["IM"], d["JM"], d["KM"]]:
            return f"the whole {d['IM']}x{d['JM']}x{d['KM']} grid with its {r['particles']} particles"
        return (f"{r['cells'][0]}x{r['cells'][1]}x{r['cells'][2]} cells from the middle of the "
                f"{d['IM']}x{d['JM']}x{d['KM']} grid with its {r['particles']} particles")

    return {"value": r_all["value"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"config {d['config']} ({'RK2' if d['rk2'] else 'Euler'}, "
                      f"deposition {d['deposit']}, centre->corner each step): "
                      f"{desc(r_all)}; median of {runs} runs of {steps} steps after a warm-up step; "
                      f"FP64 oracle -O3 x86-64-v3 OpenMP; omits the reference's PETSc view overheads",
            "one_core": {"value": r_one["value"], "cores": 1, "sample": desc(r_one), "runs": r_one["runs"]},
            "runs": r_all["runs"], "seconds": r_all["seconds"] + r_one["seconds"], "host": host}


def fields_cpu_baseline(d: dict, kplanes: int = 64, reps: int = 3) -> dict:
    """Per-step field preparation (boundaries, Contra2Cart, boundaries) by the
    FP64 oracle built -O3 with OpenMP, on the first `kplanes` cell planes of
    the same grid and Ucont (metrics computed untimed)."""
    import numpy as np

    from oracle.oracle import Oracle

    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = Oracle(fast=True)
    IM, JM, KM = d["IM"], d["JM"], min(kplanes, d["KM"])
    nn = (IM + 1) * (JM + 1) * (KM + 1)
    coor = d["coor"].reshape(-1, 3)[:nn].cpu().numpy()
    ucont = d["ucont"].reshape(-1, 3)[:nn].cpu().numpy().copy()
    csi, eta, zet, _, ori = o.metrics(IM, JM, KM, coor)

    class _Nograd:
        type = [1] * 6
        handler = [1] * 6
        vel = [[0.0, 0.0, 0.0]] * 6
        vmax = [1.0] * 6

    ucat = np.zeros_like(ucont)
    t0 = time.perf_counter()
    for _ in range(reps):
        o.set_eulerian_fields(IM, JM, KM, _Nograd, csi, eta, zet, 1, 0, (0.0, 0.0, 0.0), ori, ucat, ucont)
    dt = (time.perf_counter() - t0) / reps
    ncell_in = (IM - 1) * (JM - 1) * (KM - 1)
    return {"value": ncell_in / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"first {KM} of {d['KM']} cell planes ({IM}x{JM}x{KM}), {reps} preparations, "
                      f"FP64 oracle -O3 x86-64-v3 OpenMP, {dt * 1e3:.1f} ms each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--variants", default="fp32_mlp,fp64")
    ap.add_argument("--cpu-block", default="128,128,64",
                    help="cells of the CPU baseline's sub-block, timed with every core (a block at least "
                         "as large as the grid: the whole workload)")
    ap.add_argument("--cpu-one-block", default=None,
                    help="cells of the sub-block timed with one core (default: --cpu-block)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--call-steps", type=int, default=0,
                    help="time the steps as picgpu_step calls of this many steps each (0: one call)")
    ap.add_argument("--no-relocate", action="store_true",
                    help="skip the config-4 relocation record that follows a config-5 run")
    ap.add_argument("--error-sample", type=int, default=1 << 22,
                    help="particles of the FP32-vs-FP64 one-step error sample (0: skip)")
    ap.add_argument("--fields", action="store_true",
                    help="every step first prepares Ucat from Ucont (boundaries, Contra2Cart, boundaries), "
                         "as the reference's SetEulerianFields does each step")
    ap.add_argument("--io", default=None, metavar="DIR",
                    help="after the timed steps, write one step's output files to DIR and read them back "
                         "as a restart (timed separately, not part of the headline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from fdf_amd import _lib

    _lib.load()
    if args.config == 4:
        main_relocate(args)
        return
    dev = "cuda:0"
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    d = make_inputs(args.config, args.scale, dev)
    d["workload_name"] = f"config{args.config} scale={args.scale:g}"
    if args.io:
        os.makedirs(args.io, exist_ok=True)
        d["io_dir"] = args.io
    if args.fields:
        from fdf_amd import workloads as W

        d["ucont"] = W.contravariant_from_cartesian(d["coor"], d["ucat"], d["IM"], d["JM"], d["KM"]).contiguous()
        d["workload_name"] += " fields"
    torch.cuda.synchronize()
    copy_gbs = calibration_copy()
    log(f"inputs: config {args.config} {d['IM']}x{d['JM']}x{d['KM']} cells, {d['n']} particles, "
        f"{time.perf_counter() - t0:.1f}s")
    steps = args.steps if args.steps is not None else min(d["steps"], 100)
    g = {"nnode": (d["IM"] + 1) * (d["JM"] + 1) * (d["KM"] + 1), "ncell": d["IM"] * d["JM"] * d["KM"]}
    g["nface"] = ((d["IM"] + 1) * d["JM"] * d["KM"] + d["IM"] * (d["JM"] + 1) * d["KM"] +
                  d["IM"] * d["JM"] * (d["KM"] + 1))
    g["nface_avg"] = g["nface"] / 3

    mlp = None
    try:
        from fdf_amd import mlp as M

        mlp = {"train": lambda ctx: M.train_on_swarm(ctx, d, kernel_timing=True)}
    except Exception as e:  # the MLP module is optional in this bench line only
        log("mlp unavailable:", e)

    results = {}
    for v in [s.strip() for s in args.variants.split(",") if s.strip()]:
        r = run_variant(v, d, steps, args.warmup, mlp, args.call_steps)
        r["steps"] = steps
        r["roofline"] = roofline_for(r, d, g)
        rs = 8 if r["precision"] == "fp64" else 4
        n_avg = r["particle_steps"] / max(1, steps)
        r["step_roofline"] = {"algorithmic_bytes_per_step": step_bytes(int(n_avg), g, rs, "faces" in d["deposit"]),
                              "achieved_GBs": step_bytes(int(n_avg), g, rs, "faces" in d["deposit"]) /
                              (r["ms_per_step"] * 1e-3) / 1e9}
        r["step_roofline"]["frac"] = r["step_roofline"]["achieved_GBs"] / HBM_PEAK_GBS
        # the measured copy bandwidth as the second denominator
        r["roofline"]["measured_copy_GBs"] = copy_gbs
        r["roofline"]["frac_of_measured_copy"] = r["roofline"]["achieved"] / copy_gbs
        r["step_roofline"]["frac_of_measured_copy"] = r["step_roofline"]["achieved_GBs"] / copy_gbs
        results[v] = r
        torch.cuda.empty_cache()

    head = results.get("fp32_mlp") or next(iter(results.values()))
    cpu = None
    if not args.no_cpu_baseline:
        try:
            blk = tuple(int(v) for v in args.cpu_block.split(","))
            blk1 = tuple(int(v) for v in (args.cpu_one_block or args.cpu_block).split(","))
            cpu = cpu_baseline(d, blk, blk1)
            log(f"cpu baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} cores")
        except Exception as e:
            log("cpu baseline failed:", e)
        if "ucont" in d:
            try:
                fc = fields_cpu_baseline(d)
                log(f"field preparation cpu baseline: {fc['value']:.4g} cells/s on {fc['cores']} cores")
                for r in results.values():
                    if "fields" in r:
                        r["fields"]["cpu_baseline"] = fc
            except Exception as e:
                log("field preparation cpu baseline failed:", e)
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head["precision"] == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"config{args.config}", "grid_cells": [d["IM"], d["JM"], d["KM"]],
                   "particles": d["n"], "scale": args.scale, "advection": "RK2" if d["rk2"] else "Euler",
                   "deposition": d["deposit"], "dt": d["dt"], "variant": head["variant"],
                   "eulerian_preparation_each_step": bool(args.fields),
                   # particles keep their cell from step to step, so every walk after the
                   # first starts from the previous cell; the MLP guess serves the first
                   # locate (and config 4's relocation of fresh positions)
                   "mlp_guess_per_step": False,
                   "steps_per_call": args.call_steps if args.call_steps > 0 else steps,
                   "mlp_guess_at": "initial locate"},
        "roofline": {k: head["roofline"][k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_kernel": head["roofline"]["kernel"],
        "measured_copy_GBs": copy_gbs,
        "mlp_mfma": head.get("mlp_mfma"),
        "cpu_baseline": cpu,
        "variants": {k: {kk: vv for kk, vv in v.items() if kk not in ("kernels",)} for k, v in results.items()},
        "kernels": {k: v["kernels"] for k, v in results.items()},
    }
    if "fp64" in results and "fp32_mlp" in results:
        line["fp32_mlp_over_fp64"] = results["fp32_mlp"]["value"] / results["fp64"]["value"]
    if args.error_sample > 0:
        try:
            line["fp32_error_vs_fp64"] = precision_errors(d, min(args.error_sample, d["n"]))
            log("fp32 error vs fp64:", line["fp32_error_vs_fp64"])
        except Exception as e:
            log("precision error sample failed:", e)
    if args.config == 5 and not args.no_relocate:
        del d
        torch.cuda.empty_cache()
        try:
            line["config4_relocate"] = relocate_record(args.scale)
        except Exception as e:
            log("config 4 relocation record failed:", e)
    print(json.dumps(line), flush=True)


def relocate_record(scale: float, steps: int = 3) -> dict:
    """The MLP's score on its own workload (config 4: the bent tube at
    128x128x1024, 16M fresh uniform positions located each step, cells
    discarded): FP32 locates/s and walk iterations with the guess and from
    (0,0,0) without it."""
    import torch

    d = make_inputs(4, scale, "cuda:0")
    log(f"config 4 relocation: {d['IM']}x{d['JM']}x{d['KM']} cells, {d['n']} particles")
    with_guess = run_relocate(d, "fp32", True, steps, 1)
    without = run_relocate(d, "fp32", False, 1, 1)
    keep = ("locates_per_s", "ms_per_locate", "walk_iterations", "located", "lost")
    rec = {"workload": "config4", "grid_cells": [d["IM"], d["JM"], d["KM"]], "particles": int(d["n"]),
           "precision": "fp32", "steps_timed": {"guess": steps, "no_guess": 1},
           "guess": {k: with_guess[k] for k in keep}, "no_guess": {k: without[k] for k in keep},
           "mlp_train": with_guess.get("mlp_train", {}).get("train_s")}
    rec["walk_iterations_saved_per_locate"] = (rec["no_guess"]["walk_iterations"]["mean"] -
                                               rec["guess"]["walk_iterations"]["mean"])
    rec["speedup_from_guess"] = rec["guess"]["locates_per_s"] / rec["no_guess"]["locates_per_s"]
    del d
    torch.cuda.empty_cache()
    return rec


def run_relocate(d: dict, precision: str, use_mlp: bool, steps: int, warmup: int) -> dict:
    """Config 4: every step re-locates the whole swarm from positions with no
    cell (the start cell is the MLP guess, or (0,0,0) without it); only the
    locate is timed."""
    import torch

    from fdf_amd import mlp as M
    from fdf_amd.swarm import PicGpu

    IM, JM, KM = d["IM"], d["JM"], d["KM"]
    n = int(d["pos_dev"].shape[0])
    res = {"precision": precision, "mlp": use_mlp}
    with PicGpu(precision=precision, use_mlp_guess=False) as g:
        g.upload_grid_device(IM, JM, KM, d["coor"].data_ptr())
        g.upload_ucat_device(d["ucat"].data_ptr())
        if use_mlp:
            # longer training than the library default for the bent tube
            # (12,000 iterations at batch 16,384, about 1 s, outside the timed
            # locates): the mean walk from the guess is about 3.4 iterations
            res["mlp_train"] = M.train_from_walk(g, d, iters=12000, batch=16384, lr=2e-3)
        g.walk_stats(True)
        secs, iters, hist = 0.0, 0, None
        for s in range(warmup + steps):
            g.init_particles_device(n, d["pos_dev"].data_ptr())  # cells discarded: no temporal coherence
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.LocateAllParticlesInGrid()
            g.synchronize()
            t1 = time.perf_counter()
            if s >= warmup:
                st = g.stats()
                secs += t1 - t0
                iters += st["walk_iterations"]
                hist = st["walk_hist"] if hist is None else hist + st["walk_hist"]
                res["located"], res["lost"] = st["located"], st["lost"]
    res["locates_per_s"] = n * steps / secs
    res["ms_per_locate"] = 1e3 * secs / steps
    res["walk_iterations"] = walk_percentiles(hist)
    # the histogram's last bin holds every walk of 63 or more iterations; the
    # mean is taken from the exact total (walks that cycle count 1000, the cap)
    res["walk_iterations"]["mean"] = iters / (n * steps)
    log(f"relocate {precision} mlp={use_mlp}: {res['locates_per_s']:.4g} locates/s, "
        f"mean walk {res['walk_iterations']['mean']:.2f}, lost {res['lost']}")
    return res


def main_relocate(args) -> None:
    import torch

    torch.cuda.set_device(0)
    d = make_inputs(args.config, args.scale, "cuda:0")
    log(f"inputs: config {args.config} {d['IM']}x{d['JM']}x{d['KM']} cells, {d['n']} particles")
    steps = args.steps if args.steps is not None else d["steps"]
    out = {}
    for prec in ("fp32", "fp64"):
        out[f"{prec}_mlp"] = run_relocate(d, prec, True, steps, args.warmup)
        # without the guess the walks start at (0,0,0): long, so fewer repetitions
        out[f"{prec}_plain"] = run_relocate(d, prec, False, min(steps, 2), 1)
    head = out["fp32_mlp"]
    line = {
        "metric": "particle-locates/sec (fresh positions, walking search from the MLP guess), 1 GPU",
        "value": head["locates_per_s"], "unit": "particle-locates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_locate"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"config{args.config}", "grid_cells": [d["IM"], d["JM"], d["KM"]],
                   "particles": d["n"], "scale": args.scale, "mode": "locate only, cells discarded each step"},
        "walk_iterations_saved_by_guess": {
            p: (out[f"{p}_plain"]["walk_iterations"]["mean"] or 0) - (out[f"{p}_mlp"]["walk_iterations"]["mean"] or 0)
            for p in ("fp32", "fp64")},
        "variants": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
