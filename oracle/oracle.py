"""ctypes wrapper of the CPU oracle (oracle/libpicref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpicref.so")
LIB_FAST = os.path.join(HERE, "libpicref_fast.so")

LOCATED, LOST = 1, 3
MAX_TRAVERSAL = 1000


def interpolation_error_percent(pos, vel, finit: int) -> float:
    """LOG_INTERPOLATION_ERROR (logging.c:533-580): 100 |a - v|_2 / |a|_2 over
    every component of every particle, with the analytical velocity a of
    SetAnalyticalSolution (AnalyticalSolution.c:932-962): sin of each
    position component for finit 1, (0, 0, 1) for finit 0."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    if finit == 1:
        a = np.sin(pos)
    elif finit == 0:
        a = np.zeros_like(pos)
        a[:, 2] = 1.0
    else:
        raise ValueError("finit 0 or 1")
    return 100.0 * float(np.linalg.norm(a - np.asarray(vel, np.float64).reshape(-1, 3)) / np.linalg.norm(a))


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class _Ctx(C.Structure):
    _fields_ = [("IM", C.c_int64), ("JM", C.c_int64), ("KM", C.c_int64), ("coor", C.c_void_p),
                ("corner", C.c_void_p), ("max_traversal", C.c_int), ("bbox", C.c_void_p)]


class _Grid(C.Structure):
    _fields_ = [("IM", C.c_int64), ("JM", C.c_int64), ("KM", C.c_int64), ("coor", C.c_void_p)]


class _BC(C.Structure):
    _fields_ = [("type", C.c_int * 6), ("handler", C.c_int * 6), ("vel", (C.c_double * 3) * 6),
                ("vmax", C.c_double * 6)]


def _bc(cfg) -> _BC:
    """cfg: any object with type[6], handler[6], vel[6][3], vmax[6]."""
    b = _BC()
    for f in range(6):
        b.type[f] = int(cfg.type[f])
        b.handler[f] = int(cfg.handler[f])
        b.vmax[f] = float(cfg.vmax[f])
        for c in range(3):
            b.vel[f][c] = float(cfg.vel[f][c])
    return b


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Thin numpy front end over the C restatement."""

    def __init__(self, fast: bool = False):
        path = LIB_FAST if fast else LIB
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        lib = self.lib
        P, I64, D, I = C.c_void_p, C.c_int64, C.c_double, C.c_int
        sig = {
            "ref_cell_char_size": (D, [P]),
            "ref_face_distances": (I, [P, P, D, P]),
            "ref_point_position": (I, [P]),
            "ref_evaluate_position": (I, [P, P, P, P]),
            "ref_update_weights": (None, [P, P]),
            "ref_locate_all": (I, [C.POINTER(_Grid), I64, P, P, P, P, I, P, P]),
            "ref_center_to_corner": (None, [I64, I64, I64, P, P]),
            "ref_interp_all": (None, [I64, I64, I64, P, I64, P, P, P]),
            "ref_advect_euler": (None, [I64, P, P, D]),
            "ref_reset_statuses": (None, [I64, P]),
            "ref_bbox": (None, [C.POINTER(_Grid), P, P]),
            "ref_remove": (I64, [I64, P, P, P, P, P, P, P, P, P, P, P]),
            "ref_count_cells": (None, [I64, I64, I64, I64, P, P]),
            "ref_scatter_P": (None, [I64, I64, I64, I64, P, P, P, P, P]),
            "ref_deposit_faces": (None, [I64, I64, I64, I64, P, P, P, P, P, P, P]),
            "ref_cell_volume": (D, [P]),
            "ref_face_volfrac": (None, [I64, I64, I64, P, I, P, D, P]),
            "ref_step_euler": (I64, [C.POINTER(_Ctx), I64, P, P, P, P, P, P, P, D, P]),
            "ref_step_rk2": (I64, [C.POINTER(_Ctx), I64, P, P, P, P, P, P, P, D, P]),
            "ref_face_metrics": (None, [I64, I64, I64, P, P, P, P]),
            "ref_jacobian_inverse": (I, [I64, I64, I64, P, P]),
            "ref_grid_orientation": (I, [I64, P, P, P, P, P]),
            "ref_inlet_face": (I, [C.POINTER(_BC)]),
            "ref_initial_ucont": (None, [I64, I64, I64, P, P, P, I, P, I, I, P]),
            "ref_bc_initialize": (I, [I64, I64, I64, C.POINTER(_BC), P, P, P, P, P]),
            "ref_bc_execute": (I, [I64, I64, I64, C.POINTER(_BC), P, P, P, P, P]),
            "ref_contra2cart": (I, [I64, I64, I64, P, P, P, P, P]),
            "ref_set_eulerian_fields": (I, [I64, I64, I64, C.POINTER(_BC), P, P, P, I64, I, P, I, P, P]),
            "ref_srand48_state": (C.c_uint64, [I64]),
            "ref_drand48": (D, [C.POINTER(C.c_uint64)]),
            "ref_seed_particles": (None, [I64, I64, I64, P, I64, I, I, I64, P]),
            "ref_mlp_param_count": (I64, [I]),
            "ref_mlp_forward": (None, [I, P, I64, P, P, I]),
            "ref_mlp_loss_grad": (D, [I, P, I64, P, P, P, I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args

    # ---- primitives ----
    def char_size(self, V):
        V = np.ascontiguousarray(V, np.float64)
        return self.lib.ref_cell_char_size(_p(V))

    def evaluate(self, V, p):
        V = np.ascontiguousarray(V, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        d = np.zeros(6)
        pos = C.c_int()
        r = self.lib.ref_evaluate_position(_p(V), _p(p), _p(d), C.byref(pos))
        return r, d, pos.value

    def update_weights(self, d):
        d = np.array(d, np.float64)
        w = np.zeros(3)
        self.lib.ref_update_weights(_p(d), _p(w))
        return w

    # ---- whole-array stages ----
    def locate(self, IM, JM, KM, coor, pos, cell=None, weight=None, status=None, guess=None,
               max_traversal=MAX_TRAVERSAL):
        coor = np.ascontiguousarray(coor, np.float64)
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = pos.shape[0]
        cell = np.full((n, 3), -1, np.int64) if cell is None else np.array(cell, np.int64).reshape(n, 3)
        weight = np.zeros((n, 3)) if weight is None else np.array(weight, np.float64).reshape(n, 3)
        status = np.full(n, 4, np.int64) if status is None else np.array(status, np.int64)
        guess = None if guess is None else np.ascontiguousarray(guess, np.int64).reshape(n, 3)
        steps = np.zeros(n, np.int32)
        g = _Grid(IM, JM, KM, _p(coor))
        r = self.lib.ref_locate_all(C.byref(g), n, _p(pos), _p(cell), _p(weight), _p(status), max_traversal,
                                    _p(guess), _p(steps))
        if r != 0:
            raise RuntimeError(f"oracle locate error {r}")
        return cell, weight, status, steps

    def center_to_corner(self, IM, JM, KM, ucat):
        ucat = np.ascontiguousarray(ucat, np.float64)
        out = np.zeros_like(ucat)
        self.lib.ref_center_to_corner(IM, JM, KM, _p(ucat), _p(out))
        return out

    def interp(self, IM, JM, KM, corner, cell, weight):
        corner = np.ascontiguousarray(corner, np.float64)
        cell = np.ascontiguousarray(cell, np.int64)
        weight = np.ascontiguousarray(weight, np.float64)
        n = cell.shape[0]
        vel = np.zeros((n, 3))
        self.lib.ref_interp_all(IM, JM, KM, _p(corner), n, _p(cell), _p(weight), _p(vel))
        return vel

    def bbox(self, IM, JM, KM, coor):
        coor = np.ascontiguousarray(coor, np.float64)
        g = _Grid(IM, JM, KM, _p(coor))
        lo, hi = np.zeros(3), np.zeros(3)
        self.lib.ref_bbox(C.byref(g), _p(lo), _p(hi))
        return lo, hi

    def count_cells(self, IM, JM, KM, cell):
        cell = np.ascontiguousarray(cell, np.int64)
        out = np.zeros((IM + 1) * (JM + 1) * (KM + 1))
        self.lib.ref_count_cells(IM, JM, KM, cell.shape[0], _p(cell), _p(out))
        return out

    def scatter_P(self, IM, JM, KM, cell, P, pid, count):
        cell = np.ascontiguousarray(cell, np.int64)
        P = np.ascontiguousarray(P, np.float64)
        order = np.ascontiguousarray(np.argsort(pid, kind="stable"), np.int64)
        count = np.ascontiguousarray(count, np.float64)
        out = np.zeros_like(count)
        self.lib.ref_scatter_P(IM, JM, KM, cell.shape[0], _p(cell), _p(P), _p(order), _p(count), _p(out))
        return out

    def deposit_faces(self, IM, JM, KM, cell, weight, vel, pid):
        cell = np.ascontiguousarray(cell, np.int64)
        weight = np.ascontiguousarray(weight, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        pid = np.ascontiguousarray(pid, np.int64)
        fi = np.zeros(4 * KM * JM * (IM + 1))
        fj = np.zeros(4 * KM * (JM + 1) * IM)
        fk = np.zeros(4 * (KM + 1) * JM * IM)
        self.lib.ref_deposit_faces(IM, JM, KM, cell.shape[0], _p(cell), _p(weight), _p(vel), _p(pid), _p(fi),
                                   _p(fj), _p(fk))
        return fi, fj, fk

    def cell_volume(self, V):
        V = np.ascontiguousarray(V, np.float64)
        return self.lib.ref_cell_volume(_p(V))

    def face_volfrac(self, IM, JM, KM, coor, dir, face, vp):
        """Volume fraction of every face of direction dir from that direction's
        face array (count, mx, my, mz per face)."""
        coor = np.ascontiguousarray(coor, np.float64)
        face = np.ascontiguousarray(face, np.float64)
        out = np.zeros(face.size // 4)
        self.lib.ref_face_volfrac(IM, JM, KM, _p(coor), int(dir), _p(face), float(vp), _p(out))
        return out

    def step(self, IM, JM, KM, coor, corner, state: dict, dt: float, rk2: bool = False,
             max_traversal=MAX_TRAVERSAL, bbox=None):
        """One reference step on a state dict (pos, vel, cell, weight, status, pid, P); compacts in place.
        bbox: the padded bounding box (lo, hi) when known, else it is computed from the nodes."""
        coor = np.ascontiguousarray(coor, np.float64)
        corner = np.ascontiguousarray(corner, np.float64)
        bb = None if bbox is None else np.ascontiguousarray(np.concatenate([bbox[0], bbox[1]]), np.float64)
        s = state
        for k, t in (("pos", np.float64), ("vel", np.float64), ("weight", np.float64), ("P", np.float64),
                     ("cell", np.int64), ("status", np.int64), ("pid", np.int64)):
            s[k] = np.ascontiguousarray(s[k], t)
        n = s["pos"].shape[0]
        ctx = _Ctx(IM, JM, KM, _p(coor), _p(corner), max_traversal, _p(bb))
        removed = C.c_int64()
        fn = self.lib.ref_step_rk2 if rk2 else self.lib.ref_step_euler
        m = fn(C.byref(ctx), n, _p(s["pos"]), _p(s["vel"]), _p(s["cell"]), _p(s["weight"]), _p(s["status"]),
               _p(s["pid"]), _p(s["P"]), dt, C.byref(removed))
        if m < 0:
            raise RuntimeError(f"oracle step error {m}")
        for k in s:
            s[k] = s[k][:m]
        return removed.value

    # ---- Eulerian field preparation (node-slot arrays [(KM+1)(JM+1)(IM+1)](x3)) ----
    def metrics(self, IM, JM, KM, coor):
        """Face metrics, Jacobian inverse and the orientation check (setup.c:279-281).
        Returns (csi, eta, zet, aj, orientation); raises on the reference's errors."""
        coor = np.ascontiguousarray(coor, np.float64)
        nn = (IM + 1) * (JM + 1) * (KM + 1)
        csi, eta, zet = (np.zeros((nn, 3)) for _ in range(3))
        aj = np.zeros(nn)
        self.lib.ref_face_metrics(IM, JM, KM, _p(coor), _p(csi), _p(eta), _p(zet))
        r = self.lib.ref_jacobian_inverse(IM, JM, KM, _p(coor), _p(aj))
        if r != 0:
            raise RuntimeError(f"oracle jacobian error {r}")
        o = C.c_int()
        r = self.lib.ref_grid_orientation(nn, _p(csi), _p(eta), _p(zet), _p(aj), C.byref(o))
        if r != 0:
            raise RuntimeError(f"oracle orientation error {r}")
        return csi, eta, zet, aj, o.value

    def inlet_face(self, cfg):
        b = _bc(cfg)
        return self.lib.ref_inlet_face(C.byref(b))

    def initial_ucont(self, IM, JM, KM, csi, eta, zet, finit, uin, inlet_face, orientation, ucont):
        uin = np.ascontiguousarray(uin, np.float64)
        self.lib.ref_initial_ucont(IM, JM, KM, _p(csi), _p(eta), _p(zet), int(finit), _p(uin), int(inlet_face),
                                   int(orientation), _p(ucont))

    def bc_initialize(self, IM, JM, KM, cfg, csi, eta, zet, ucat, ucont):
        b = _bc(cfg)
        r = self.lib.ref_bc_initialize(IM, JM, KM, C.byref(b), _p(csi), _p(eta), _p(zet), _p(ucat), _p(ucont))
        if r != 0:
            raise RuntimeError(f"oracle bc error {r}")

    def bc_execute(self, IM, JM, KM, cfg, csi, eta, zet, ucat, ucont):
        b = _bc(cfg)
        r = self.lib.ref_bc_execute(IM, JM, KM, C.byref(b), _p(csi), _p(eta), _p(zet), _p(ucat), _p(ucont))
        if r != 0:
            raise RuntimeError(f"oracle bc error {r}")

    def contra2cart(self, IM, JM, KM, ucont, csi, eta, zet, ucat):
        r = self.lib.ref_contra2cart(IM, JM, KM, _p(ucont), _p(csi), _p(eta), _p(zet), _p(ucat))
        if r != 0:
            raise RuntimeError(f"oracle contra2cart error {r}")

    def set_eulerian_fields(self, IM, JM, KM, cfg, csi, eta, zet, step, finit, uin, orientation, ucat, ucont):
        b = _bc(cfg)
        uin = np.ascontiguousarray(uin, np.float64)
        r = self.lib.ref_set_eulerian_fields(IM, JM, KM, C.byref(b), _p(csi), _p(eta), _p(zet), int(step),
                                             int(finit), _p(uin), int(orientation), _p(ucat), _p(ucont))
        if r != 0:
            raise RuntimeError(f"oracle eulerian error {r}")

    # ---- seeding ----
    def drand48_stream(self, seed: int, count: int) -> np.ndarray:
        x = C.c_uint64(self.lib.ref_srand48_state(int(seed)))
        return np.array([self.lib.ref_drand48(C.byref(x)) for _ in range(count)])

    def seed_particles(self, IM, JM, KM, coor, n, volumetric=True, face=4, seed=202403):
        coor = np.ascontiguousarray(coor, np.float64)
        pos = np.zeros((n, 3))
        self.lib.ref_seed_particles(IM, JM, KM, _p(coor), n, 1 if volumetric else 0, int(face), int(seed), _p(pos))
        return pos

    # ---- MLP (spec) ----
    def mlp_forward(self, params, x, round_bf16=True, H=64):
        params = np.ascontiguousarray(params, np.float64)
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        y = np.zeros_like(x)
        self.lib.ref_mlp_forward(H, _p(params), x.shape[0], _p(x), _p(y), int(round_bf16))
        return y

    def mlp_loss_grad(self, params, x, t, round_bf16=True, H=64):
        params = np.ascontiguousarray(params, np.float64)
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.float64).reshape(-1, 3)
        g = np.zeros_like(params)
        loss = self.lib.ref_mlp_loss_grad(H, _p(params), x.shape[0], _p(x), _p(t), _p(g), int(round_bf16))
        return loss, g
