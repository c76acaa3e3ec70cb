"""Host-side mirror of the reference's per-step particle interface, over the C-ABI.

Method names follow the reference entry points they stand in for (one MPI
rank); errors raise PicgpuError (the reference returns a nonzero
PetscErrorCode); per-particle outcomes are statuses.

    ResetAllParticleStatuses          ParticleMotion.c:2149-2176
    UpdateAllParticlePositions        ParticleMotion.c:38-95
    LocateAllParticlesInGrid          ParticleMotion.c:1931-2135 (walk: walkingsearch.c:1094-1242)
    CheckAndRemoveLostParticles       ParticleMotion.c:532-625
    CheckAndRemoveOutOfBoundsParticles ParticleMotion.c:283-383
    InterpolateAllFieldsToSwarm       interpolation.c:1660-1711
    ScatterAllParticleFieldsToEulerFields interpolation.c:2223-2313
    AdvanceSimulation                 the particle stages of AdvanceSimulation_TEST simulation.c:248-292

Eulerian field preparation (the stage before the particle loop):

    compute_metrics                   ComputeFaceMetrics + ComputeCellCenteredJacobianInverse
                                      + CheckAndFixGridOrientation  Metric.c:195-622
    BoundarySystem_Create             Boundaries.c:467-565
    SetInitialInteriorField           AnalyticalSolution.c:986-1176
    BoundarySystem_ExecuteStep        Boundaries.c:566-646
    Contra2Cart                       setup.c:1294-1420
    SetEulerianFields                 simulation.c:38-107
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PicGpu:
    """One GPU context: grid, Eulerian field and swarm resident in HBM."""

    def __init__(self, precision: str = "fp64", use_mlp_guess: bool = False, max_traversal: int = 1000,
                 collect_walk_stats: bool = False, device: int = 0):
        self.lib = L.load()
        opt = L.PicgpuOptions()
        self.lib.picgpu_default_options(C.byref(opt))
        opt.precision = {"fp64": L.FP64, "fp32": L.FP32}[precision]
        opt.use_mlp_guess = int(use_mlp_guess)
        opt.max_traversal = int(max_traversal)
        opt.collect_walk_stats = int(collect_walk_stats)
        opt.device = int(device)
        self.precision = precision
        h = C.c_void_p()
        L.check(self.lib.picgpu_create(C.byref(h), C.byref(opt)))
        self.h = h
        self.IM = self.JM = self.KM = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.picgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- inputs ----
    def upload_grid(self, IM: int, JM: int, KM: int, coor: np.ndarray):
        coor = np.ascontiguousarray(coor, dtype=np.float64)
        assert coor.size == 3 * (IM + 1) * (JM + 1) * (KM + 1)
        L.check(self.lib.picgpu_upload_grid(self.h, IM, JM, KM, _ptr(coor)))
        self.IM, self.JM, self.KM = IM, JM, KM

    def upload_ucat(self, ucat: np.ndarray):
        ucat = np.ascontiguousarray(ucat, dtype=np.float64)
        assert ucat.size == 3 * (self.IM + 1) * (self.JM + 1) * (self.KM + 1)
        L.check(self.lib.picgpu_upload_ucat(self.h, _ptr(ucat)))

    def upload_particles(self, pos, vel=None, cell=None, weight=None, status=None, pid=None, P=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        cell = (np.full((n, 3), -1, np.int64) if cell is None else np.ascontiguousarray(cell, np.int64).reshape(n, 3))
        status = (np.full(n, L.UNINITIALIZED, np.int64) if status is None else np.ascontiguousarray(status, np.int64))
        pid = np.arange(n, dtype=np.int64) if pid is None else np.ascontiguousarray(pid, np.int64)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(n, 3)
        weight = None if weight is None else np.ascontiguousarray(weight, np.float64).reshape(n, 3)
        P = None if P is None else np.ascontiguousarray(P, np.float64)
        L.check(self.lib.picgpu_upload_particles(self.h, n, _ptr(pos), _ptr(vel), _ptr(cell), _ptr(weight),
                                                  _ptr(status), _ptr(pid), _ptr(P)))

    # device-resident inputs (HIP device pointers, e.g. torch.Tensor.data_ptr())
    def upload_grid_device(self, IM: int, JM: int, KM: int, d_coor: int):
        L.check(self.lib.picgpu_upload_grid_device(self.h, IM, JM, KM, C.c_void_p(d_coor)))
        self.IM, self.JM, self.KM = IM, JM, KM

    def upload_ucat_device(self, d_ucat: int):
        L.check(self.lib.picgpu_upload_ucat_device(self.h, C.c_void_p(d_ucat)))

    def init_particles_device(self, n: int, d_pos: int, pid_base: int = 0):
        L.check(self.lib.picgpu_init_particles_device(self.h, n, C.c_void_p(d_pos), pid_base))

    # ---- outputs ----
    @property
    def n(self) -> int:
        v = C.c_int64()
        L.check(self.lib.picgpu_num_particles(self.h, C.byref(v)))
        return v.value

    def download_particles(self) -> dict:
        n = self.n
        out = {
            "pos": np.empty((n, 3), np.float64),
            "vel": np.empty((n, 3), np.float64),
            "cell": np.empty((n, 3), np.int64),
            "weight": np.empty((n, 3), np.float64),
            "status": np.empty(n, np.int64),
            "pid": np.empty(n, np.int64),
        }
        L.check(self.lib.picgpu_download_particles(self.h, _ptr(out["pos"]), _ptr(out["vel"]), _ptr(out["cell"]),
                                                    _ptr(out["weight"]), _ptr(out["status"]), _ptr(out["pid"])))
        return out

    def walk_steps(self) -> np.ndarray:
        s = np.empty(self.n, np.int32)
        L.check(self.lib.picgpu_download_walk_steps(self.h, _ptr(s)))
        return s

    def field(self, which: int) -> np.ndarray:
        cnt = C.c_int64()
        L.check(self.lib.picgpu_grid_field_size(self.h, which, C.byref(cnt)))
        out = np.empty(cnt.value, np.float64)
        L.check(self.lib.picgpu_download_grid_field(self.h, which, _ptr(out)))
        return out

    def particle_count_summary(self) -> tuple[float, float, int]:
        """CalculateParticleCountPerCell's logged diagnostics (ParticleMotion.c:1021-1026):
        (total counted, largest count in a cell, its first flat node-slot index),
        reduced on the device."""
        tot, mx, at = C.c_double(), C.c_double(), C.c_int64()
        L.check(self.lib.picgpu_particle_count_summary(self.h, C.byref(tot), C.byref(mx), C.byref(at)))
        return tot.value, mx.value, at.value

    def LOG_INTERPOLATION_ERROR(self, finit: int) -> float:
        """The reference's interpolation-error diagnostic (logging.c:533-580), in
        percent, reduced on the device: finit 1 compares with sin of the
        positions, finit 0 with (0, 0, 1)."""
        pct = C.c_double()
        L.check(self.lib.picgpu_interpolation_error(self.h, finit, C.byref(pct)))
        return pct.value

    def stats(self) -> dict:
        st = L.PicgpuStats()
        L.check(self.lib.picgpu_get_stats(self.h, C.byref(st)))
        out = {f: getattr(st, f) for f, _ in L.PicgpuStats._fields_}
        out["walk_hist"] = np.array(list(st.walk_hist), np.int64)
        return out

    def walk_stats(self, on: bool = True):
        L.check(self.lib.picgpu_walk_stats_enable(self.h, int(on)))

    # ---- reference entry points ----
    def ResetAllParticleStatuses(self):
        L.check(self.lib.picgpu_reset_statuses(self.h))

    def UpdateAllParticlePositions(self, dt: float):
        L.check(self.lib.picgpu_advect(self.h, dt))

    def LocateAllParticlesInGrid(self):
        L.check(self.lib.picgpu_locate(self.h))

    def CheckAndRemoveLostParticles(self) -> int:
        r = C.c_int64()
        L.check(self.lib.picgpu_remove_lost(self.h, C.byref(r)))
        return r.value

    def CheckAndRemoveOutOfBoundsParticles(self) -> int:
        r = C.c_int64()
        L.check(self.lib.picgpu_remove_out_of_bounds(self.h, C.byref(r)))
        return r.value

    def InterpolateAllFieldsToSwarm(self):
        L.check(self.lib.picgpu_interpolate(self.h))

    def ScatterAllParticleFieldsToEulerFields(self):
        L.check(self.lib.picgpu_scatter(self.h))

    def DepositParticlesToFaces(self, particle_volume: float = 1.0):
        L.check(self.lib.picgpu_deposit_faces(self.h, particle_volume))

    def set_particle_volume(self, particle_volume: float):
        """V_p of the face volume fractions (count * V_p / V_dual)."""
        L.check(self.lib.picgpu_set_particle_volume(self.h, particle_volume))

    def AdvanceSimulation(self, dt: float, nsteps: int, rk2: bool = False, deposit_cells: bool = True,
                          deposit_faces: bool = False, staged: bool = False, fields: bool = False,
                          faces_of_start: bool = False):
        """fields=True: every step first prepares Ucat (boundaries, Contra2Cart,
        boundaries), as SetEulerianFields does for the steps after the first.
        faces_of_start=True: the faces hold the deposition of the state at the
        start of the call's last step (PICGPU_STEP_FACES_OF_START)."""
        flags = (L.STEP_RK2 if rk2 else 0) | (L.STEP_DEPOSIT_CELLS if deposit_cells else 0) | \
                (L.STEP_DEPOSIT_FACES if deposit_faces else 0) | (L.STEP_STAGED if staged else 0) | \
                (L.STEP_FIELDS if fields else 0) | (L.STEP_FACES_OF_START if faces_of_start else 0)
        L.check(self.lib.picgpu_step(self.h, dt, nsteps, flags))

    # ---- Eulerian field preparation ----
    def compute_metrics(self) -> int:
        """Csi/Eta/Zet, Aj and the orientation check; returns GridOrientation."""
        o = C.c_int()
        L.check(self.lib.picgpu_compute_metrics(self.h, C.byref(o)))
        return o.value

    def BoundarySystem_Create(self, config):
        """config: a BoundaryConfig or the path of a bcs.dat-format file."""
        from .boundary import BoundaryConfig

        cfg = config if isinstance(config, BoundaryConfig) else BoundaryConfig.from_file(config)
        b = cfg.to_c()
        L.check(self.lib.picgpu_boundary_create(self.h, C.byref(b)))
        self.boundary = cfg

    def upload_ucont(self, ucont: np.ndarray):
        ucont = np.ascontiguousarray(ucont, dtype=np.float64)
        assert ucont.size == 3 * (self.IM + 1) * (self.JM + 1) * (self.KM + 1)
        L.check(self.lib.picgpu_upload_ucont(self.h, _ptr(ucont)))

    def upload_ucont_device(self, d_ucont: int):
        L.check(self.lib.picgpu_upload_ucont_device(self.h, C.c_void_p(d_ucont)))

    def SetInitialInteriorField(self, finit: int, uin=(0.0, 0.0, 0.0)):
        u = np.ascontiguousarray(uin, np.float64)
        L.check(self.lib.picgpu_set_initial_ucont(self.h, int(finit), _ptr(u)))

    def BoundarySystem_ExecuteStep(self):
        L.check(self.lib.picgpu_boundary_execute(self.h))

    def Contra2Cart(self):
        L.check(self.lib.picgpu_contra2cart(self.h))

    def SetEulerianFields(self, step: int, start_step: int = 0, finit: int = 0, uin=(0.0, 0.0, 0.0)):
        u = np.ascontiguousarray(uin, np.float64)
        L.check(self.lib.picgpu_set_eulerian_fields(self.h, int(step), int(start_step), int(finit), _ptr(u)))

    # ---- seeding (the reference's drand48 stream; seed 202403 + rank there) ----
    def InitializeParticleBasicProperties(self, n: int, volumetric: bool = True, inlet_face: int = 4,
                                          seed: int = 202403, pid_base: int = 0):
        """n particles in the volume (-pinit 1) or on the inlet face (-pinit 0)."""
        L.check(self.lib.picgpu_seed_particles(self.h, int(n), 1 if volumetric else 0, int(inlet_face), int(seed),
                                               int(pid_base)))

    def ReinitializeParticlesOnInletSurface(self, inlet_face: int = 4, seed: int = 202403):
        L.check(self.lib.picgpu_reinit_on_inlet(self.h, int(inlet_face), int(seed)))

    # ---- on-disk formats (PETSc binary Vec files under `directory`) ----
    def WriteOutput(self, step: int, directory: str = "results"):
        """WriteSwarmField(position, velocity) + WriteSimulationFields (simulation.c:304-306)."""
        L.check(self.lib.picgpu_write_output(self.h, directory.encode(), int(step)))

    def WriteVTK(self, step: int, prefix: str = "."):
        """WriteEulerianVTK + WriteParticleVTK (postprocess.c:451-520)."""
        L.check(self.lib.picgpu_write_vtk(self.h, prefix.encode(), int(step)))

    def ReadRestart(self, step: int, directory: str = "results"):
        """ReadSimulationFields + ReadAllSwarmFields + PrepareLoadedSwarmForRelocation."""
        L.check(self.lib.picgpu_read_restart(self.h, directory.encode(), int(step)))

    def synchronize(self):
        L.check(self.lib.picgpu_synchronize(self.h))

    # ---- MLP initial guess ----
    def mlp_param_count(self) -> int:
        return int(self.lib.picgpu_mlp_param_count())

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.mlp_param_count()
        L.check(self.lib.picgpu_mlp_set_params(self.h, _ptr(p)))

    def mlp_get_params(self) -> np.ndarray:
        p = np.empty(self.mlp_param_count(), np.float32)
        L.check(self.lib.picgpu_mlp_get_params(self.h, _ptr(p)))
        return p

    def mlp_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        y = np.empty_like(x)
        L.check(self.lib.picgpu_mlp_forward(self.h, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def mlp_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        g = np.empty(self.mlp_param_count(), np.float32)
        loss = C.c_float()
        L.check(self.lib.picgpu_mlp_loss_grad(self.h, x.shape[0], _ptr(x), _ptr(t), C.byref(loss), _ptr(g)))
        return loss.value, g

    def mlp_train(self, max_pairs: int = 2_097_152, batch: int = 4096, iters: int = 2000, lr: float = 3e-3,
                  seed: int = 3004) -> float:
        loss = C.c_float()
        L.check(self.lib.picgpu_mlp_train(self.h, max_pairs, batch, iters, lr, seed, C.byref(loss)))
        return loss.value

    def mlp_guess(self, pos) -> np.ndarray:
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        cell = np.empty((pos.shape[0], 3), np.int64)
        L.check(self.lib.picgpu_mlp_guess(self.h, pos.shape[0], _ptr(pos), _ptr(cell)))
        return cell

    def mlp_enable(self, on: bool = True):
        L.check(self.lib.picgpu_mlp_enable(self.h, int(on)))

    # ---- timing ----
    def timing(self, on: bool = True):
        L.check(self.lib.picgpu_timing_enable(self.h, int(on)))

    def timing_reset(self):
        L.check(self.lib.picgpu_timing_reset(self.h))

    def timing_select(self, kernel_prefix: str = ""):
        """Time only the kernels whose name starts with kernel_prefix ("" = all)."""
        L.check(self.lib.picgpu_timing_select(self.h, kernel_prefix.encode()))

    def timing_query(self, kernel: str):
        ms = C.c_double()
        cnt = C.c_int64()
        L.check(self.lib.picgpu_timing_query(self.h, kernel.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value
