"""picgpu_locate_device_noalloc (the locate that allocates nothing) against
picgpu_locate, through the C-ABI: same cells, statuses and weights (bit for
bit in both precision modes: the same kernel and arithmetic), same counters,
on two calls in a row, the second with fewer particles.  The device buffers
are the test's own, allocated with the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

from fdf_amd import _lib as L
from fdf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the runtime libpicgpu.so links."""

    H2D, D2H = 1, 2

    def __init__(self):
        L.load()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_locate_matches_locate(precision):
    from fdf_amd.swarm import PicGpu

    IM, JM, KM = 16, 16, 32
    coor = W.perturbed_cube(IM, JM, KM, seed=11)
    rng = np.random.default_rng(41)
    real = np.float64 if precision == "fp64" else np.float32
    hip = _Hip()
    lib = L.load()
    with PicGpu(precision=precision) as g:
        g.upload_grid(IM, JM, KM, coor)
        g.upload_particles(np.full((1, 3), 0.5))
        g.LocateAllParticlesInGrid()  # builds the face-plane records
        for n in (30000, 7000):
            _say("noalloc locate", precision, n)
            pos = rng.random((n, 3))
            pos[: n // 50] = pos[: n // 50] * 1.4 - 0.2  # some outside the domain
            start = np.floor(pos * [IM, JM, KM]).astype(np.int64) + rng.integers(-2, 3, (n, 3))
            start[n // 3:n // 2] = -1
            ok = ((start >= 0) & (start < [IM, JM, KM])).all(1)
            packed = np.where(ok, start[:, 0] + IM * (start[:, 1] + JM * start[:, 2]), -1).astype(np.int32)
            status = np.zeros(n, np.int32)
            status[::7] = 1  # already located: skipped by both
            # the existing entry point on the context's swarm
            g.upload_particles(pos, cell=np.where(ok[:, None], start, -1), status=status.astype(np.int64))
            g.LocateAllParticlesInGrid()
            ref = g.download_particles()
            ref_st = g.stats()
            o = np.argsort(ref["pid"])
            # the same particles in the caller's device buffers
            d_pos, d_cell = hip.put(pos), hip.put(packed)
            d_w, d_st = hip.put(np.full((n, 3), -7.0, real)), hip.put(status)
            st = L.PicgpuStats()
            L.check(lib.picgpu_locate_device_noalloc(g.h, n, d_pos, d_cell, d_w, d_st, C.byref(st)))
            cell = hip.get(d_cell, packed)
            w = hip.get(d_w, np.zeros((n, 3), real))
            stat = hip.get(d_st, status)
            hip.free()
            unpacked = np.where(cell[:, None] >= 0,
                                np.stack([cell % IM, (cell // IM) % JM, cell // (IM * JM)], 1), -1)
            np.testing.assert_array_equal(unpacked, ref["cell"][o])
            np.testing.assert_array_equal(stat, ref["status"][o])
            walked = status == 0
            assert np.array_equal(w[walked & (stat == 1)].astype(np.float64), ref["weight"][o][walked & (stat == 1)])
            assert (w[~walked] == -7.0).all()  # skipped particles keep their weights
            assert st.located == ref_st["located"] and st.lost == ref_st["lost"]
            assert st.walk_iterations == ref_st["walk_iterations"] and st.n_particles == n
            assert st.fp64_fallbacks == ref_st["fp64_fallbacks"]
            assert sum(st.walk_hist) == walked.sum()


def test_noalloc_rejects_what_it_cannot_do():
    from fdf_amd.swarm import PicGpu

    lib = L.load()
    st = L.PicgpuStats()
    with PicGpu(precision="fp64") as g:
        g.upload_grid(4, 4, 4, W.uniform_grid(4, 4, 4))
        # no face-plane records yet
        assert lib.picgpu_locate_device_noalloc(g.h, 0, None, None, None, None, C.byref(st)) == -2
        g.upload_particles(np.full((1, 3), 0.5))
        g.LocateAllParticlesInGrid()
        assert lib.picgpu_locate_device_noalloc(g.h, 1 << 31, None, None, None, None, None) == -1
        assert lib.picgpu_locate_device_noalloc(g.h, 5, None, None, None, None, None) == -1
        assert lib.picgpu_locate_device_noalloc(g.h, 0, None, None, None, None, C.byref(st)) == 0
        assert st.n_particles == 0 and st.located == 0
