"""Synthetic inputs of the five benchmark configurations (grids, velocity, particles).

The reference ships no inputs, so each configuration is a specification
written here.  Conventions follow the reference: IM, JM, KM are cell counts,
nodes are [KM+1][JM+1][IM+1][3] with i fastest (grid.c:386-395), the
cell-centred velocity `ucat` of cell (i,j,k) is stored at node slot (i,j,k)
(README.md:242-260).

All generators are written against an array namespace `xp` (numpy by
default, torch for device-side generation of the large configurations), so
the same code defines the inputs on both paths.

Configurations
  1  unit cube 20x20x280, ucat = (0,0,1), 10,000 particles U[0,1)^3,
     dt = 0.1, 10 steps, forward Euler, cell counts.
  2  bent tube 64x64x512: centreline radius R=5, tube radius a=0.5,
     90-degree sweep, O-grid cross-section (each square ring of (i,j)
     projected radially onto a circle, rings clustered to the wall by ratio
     1.05), Poiseuille
     U0(1 - rho^2/a^2) along the axis + iid N(0, 0.05^2) per cell component,
     4,194,304 particles uniform in the tube volume (rho <= 0.999 a), RK2,
     50 steps, dt = 0.5 * (R * sweep / KM) / U0, face volume fraction.
  3  unit cube 128x128x1024 with interior nodes displaced by
     0.2 h_a sin(2 pi (kappa_a . xi) + phi_a) per axis a (seeded kappa in
     {1,2,3}^3, phi in [0, 2 pi)), Taylor-Green-like field
     (cos X sin Y sin Z, sin X cos Y sin Z, sin X sin Y cos Z), X = 2 pi x
     (AnalyticalSolution.c:41-45), 33,554,432 particles, 20 steps, Euler,
     dt = 0.5 / KM.
  4  bent tube of config 2 at 128x128x1024; 2,097,152 training pairs;
     16,777,216 fresh uniform-in-tube positions per step, locate only.
  5  perturbed cube 256x256x1024, divergence-free random Fourier field
     u = sum_m a_m cos(2 pi k_m . x + phi_m), 64 modes, integer k with
     1 <= |k| <= 8, a_m perpendicular to k_m, |a_m| = 1/4; 134,217,728
     particles; 100 steps RK2, dt = 0.5 * (1/KM) / max|ucat|; face
     deposition of count and momentum.
Seeds: positions 1000+c, velocity perturbation / phases 2000+c, grid 3000+c.
"""
from __future__ import annotations

import math

import numpy as np

TUBE_R, TUBE_A, TUBE_SWEEP, TUBE_RATIO = 5.0, 0.5, math.pi / 2, 1.05
RHO_MAX = 0.999


class _NumpyNS:
    float64 = np.float64
    pi = math.pi

    @staticmethod
    def arange(n):
        return np.arange(n, dtype=np.float64)

    sin = staticmethod(np.sin)
    cos = staticmethod(np.cos)
    sqrt = staticmethod(np.sqrt)
    abs = staticmethod(np.abs)
    where = staticmethod(np.where)
    maximum = staticmethod(np.maximum)
    atan2 = staticmethod(np.arctan2)

    @staticmethod
    def stack(xs, axis=-1):
        return np.stack(xs, axis=axis)

    @staticmethod
    def zeros(shape):
        return np.zeros(shape, np.float64)

    @staticmethod
    def asarray(a):
        return np.asarray(a, dtype=np.float64)


class _TorchNS:
    def __init__(self, device):
        import torch

        self.t = torch
        self.device = device
        self.float64 = torch.float64
        self.pi = math.pi

    def arange(self, n):
        return self.t.arange(n, dtype=self.t.float64, device=self.device)

    def sin(self, x):
        return self.t.sin(x)

    def cos(self, x):
        return self.t.cos(x)

    def sqrt(self, x):
        return self.t.sqrt(x)

    def abs(self, x):
        return self.t.abs(x)

    def where(self, c, a, b):
        return self.t.where(c, a, b)

    def maximum(self, a, b):
        return self.t.maximum(a, b) if hasattr(a, "shape") and hasattr(b, "shape") else self.t.clamp(a, min=b)

    def atan2(self, y, x):
        return self.t.atan2(y, x)

    def stack(self, xs, axis=-1):
        return self.t.stack(xs, dim=axis)

    def zeros(self, shape):
        return self.t.zeros(shape, dtype=self.t.float64, device=self.device)

    def asarray(self, a):
        return self.t.as_tensor(np.asarray(a, dtype=np.float64), device=self.device)


def namespace(backend="numpy", device="cuda"):
    return _NumpyNS() if backend == "numpy" else _TorchNS(device)


# ---------------------------------------------------------------- grids
def uniform_grid(IM, JM, KM, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), xp=None):
    """x_i = xMin + L * (i / IM) (grid.c:265-283 with r = 1, :386-395)."""
    xp = xp or _NumpyNS()
    x = lo[0] + (hi[0] - lo[0]) * (xp.arange(IM + 1) / float(IM))
    y = lo[1] + (hi[1] - lo[1]) * (xp.arange(JM + 1) / float(JM))
    z = lo[2] + (hi[2] - lo[2]) * (xp.arange(KM + 1) / float(KM))
    X = x[None, None, :] + 0.0 * y[None, :, None] + 0.0 * z[:, None, None]
    Y = 0.0 * x[None, None, :] + y[None, :, None] + 0.0 * z[:, None, None]
    Z = 0.0 * x[None, None, :] + 0.0 * y[None, :, None] + z[:, None, None]
    return xp.stack([X, Y, Z], axis=-1).reshape(-1, 3)


def _tube_cross_section(IM, JM, ratio, xp):
    assert IM == JM and IM % 2 == 0, "the O-grid cross-section needs IM == JM, even"
    n = IM // 2
    u = 2.0 * xp.arange(IM + 1) / IM - 1.0
    v = 2.0 * xp.arange(JM + 1) / JM - 1.0
    U = u[None, :] + 0.0 * v[:, None]
    V = 0.0 * u[None, :] + v[:, None]
    t = xp.maximum(xp.abs(U), xp.abs(V))
    rn = ratio ** n
    s = (rn - ratio ** (n * (1.0 - t))) / (rn - 1.0)
    # each square ring t = max(|u|,|v|) is projected radially onto the circle
    # of radius s(t); ring widths shrink by 1/ratio per ring toward the wall
    r = xp.sqrt(U * U + V * V)
    scale = xp.where(r > 0, s / xp.where(r > 0, r, 1.0 + 0.0 * r), 0.0 * r)
    X = U * scale
    Y = V * scale
    return X, Y  # [JM+1][IM+1], unit disk


def tube_grid(IM, JM, KM, R=TUBE_R, a=TUBE_A, sweep=TUBE_SWEEP, ratio=TUBE_RATIO, xp=None):
    """Bent tube: torus section, right-handed (i radial-ish, j along z, k along -theta)."""
    xp = xp or _NumpyNS()
    X, Y = _tube_cross_section(IM, JM, ratio, xp)
    th = -sweep * xp.arange(KM + 1) / KM
    rr = R + a * X
    x = rr[None, :, :] * xp.cos(th)[:, None, None]
    y = rr[None, :, :] * xp.sin(th)[:, None, None]
    z = (a * Y)[None, :, :] + 0.0 * th[:, None, None]
    return xp.stack([x, y, z], axis=-1).reshape(-1, 3)


def perturbed_cube(IM, JM, KM, seed, amp=0.2, xp=None):
    """Unit cube with interior nodes displaced by amp*h_a*sin(2 pi kappa_a.xi + phi_a)."""
    xp = xp or _NumpyNS()
    rng = np.random.Generator(np.random.PCG64(seed))
    kappa = rng.integers(1, 4, size=(3, 3)).astype(np.float64)
    phi = rng.random(3) * 2 * math.pi
    xi = xp.arange(IM + 1) / IM
    eta = xp.arange(JM + 1) / JM
    zeta = xp.arange(KM + 1) / KM
    XI = xi[None, None, :] + 0.0 * eta[None, :, None] + 0.0 * zeta[:, None, None]
    ETA = 0.0 * xi[None, None, :] + eta[None, :, None] + 0.0 * zeta[:, None, None]
    ZETA = 0.0 * xi[None, None, :] + 0.0 * eta[None, :, None] + zeta[:, None, None]
    ii = xp.arange(IM + 1)
    jj = xp.arange(JM + 1)
    kk = xp.arange(KM + 1)
    interior = ((ii > 0) & (ii < IM))[None, None, :] & ((jj > 0) & (jj < JM))[None, :, None] & \
        ((kk > 0) & (kk < KM))[:, None, None]
    h = (1.0 / IM, 1.0 / JM, 1.0 / KM)
    out = []
    for ax, base in enumerate((XI, ETA, ZETA)):
        ph = 2 * math.pi * (kappa[ax, 0] * XI + kappa[ax, 1] * ETA + kappa[ax, 2] * ZETA) + phi[ax]
        disp = amp * h[ax] * xp.sin(ph)
        out.append(base + xp.where(interior, disp, 0.0 * disp))
    return xp.stack(out, axis=-1).reshape(-1, 3)


def cell_centroids(coor, IM, JM, KM):
    """Mean of the 8 vertices of every cell, [KM][JM][IM][3]."""
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    s = (c[:-1, :-1, :-1] + c[:-1, :-1, 1:] + c[:-1, 1:, :-1] + c[:-1, 1:, 1:] +
         c[1:, :-1, :-1] + c[1:, :-1, 1:] + c[1:, 1:, :-1] + c[1:, 1:, 1:])
    return s / 8.0


def to_node_slots(cellvals, IM, JM, KM, xp=None):
    """[KM][JM][IM][3] -> node-slot layout [(KM+1)(JM+1)(IM+1)][3] (zeros in the extra slots)."""
    xp = xp or _NumpyNS()
    out = xp.zeros((KM + 1, JM + 1, IM + 1, 3))
    out[:KM, :JM, :IM] = cellvals
    return out.reshape(-1, 3)


# ---------------------------------------------------------------- velocity
def ucat_uniform(IM, JM, KM, vec=(0.0, 0.0, 1.0), xp=None):
    xp = xp or _NumpyNS()
    cells = xp.zeros((KM, JM, IM, 3))
    for a in range(3):
        cells[..., a] = vec[a]
    return to_node_slots(cells, IM, JM, KM, xp)


def ucat_poiseuille(coor, IM, JM, KM, seed, U0=1.0, noise=0.05, R=TUBE_R, a=TUBE_A, xp=None):
    xp = xp or _NumpyNS()
    cc = cell_centroids(coor, IM, JM, KM)
    x, y, z = cc[..., 0], cc[..., 1], cc[..., 2]
    th = xp.atan2(y, x)
    rho2 = (xp.sqrt(x * x + y * y) - R) ** 2 + z * z
    mag = U0 * xp.maximum(1.0 - rho2 / (a * a), 0.0 * rho2)
    # tangent in the direction of increasing k (theta decreasing)
    cells = xp.stack([mag * xp.sin(th), -mag * xp.cos(th), 0.0 * mag], axis=-1)
    rng = np.random.Generator(np.random.PCG64(seed))
    pert = xp.asarray(rng.normal(0.0, noise * U0, size=(KM, JM, IM, 3)))
    return to_node_slots(cells + pert, IM, JM, KM, xp)


def ucat_taylor_green(coor, IM, JM, KM, xp=None):
    """AnalyticalSolution.c:41-45 with x -> 2 pi x."""
    xp = xp or _NumpyNS()
    cc = cell_centroids(coor, IM, JM, KM)
    X, Y, Z = 2 * math.pi * cc[..., 0], 2 * math.pi * cc[..., 1], 2 * math.pi * cc[..., 2]
    cells = xp.stack([xp.cos(X) * xp.sin(Y) * xp.sin(Z), xp.sin(X) * xp.cos(Y) * xp.sin(Z),
                      xp.sin(X) * xp.sin(Y) * xp.cos(Z)], axis=-1)
    return to_node_slots(cells, IM, JM, KM, xp)


def ucat_sine(coor, IM, JM, KM, xp=None):
    """The finit = 1 field of SetLocalCartesianField_Vector
    (AnalyticalSolution.c:36-40), (sin x, sin y, sin z), at the cell
    centroids: the field LOG_INTERPOLATION_ERROR (logging.c:533-580) checks
    the interpolated particle velocities against."""
    xp = xp or _NumpyNS()
    cc = cell_centroids(coor, IM, JM, KM)
    cells = xp.stack([xp.sin(cc[..., 0]), xp.sin(cc[..., 1]), xp.sin(cc[..., 2])], axis=-1)
    return to_node_slots(cells, IM, JM, KM, xp)


def fourier_modes(seed, nmodes=64, kmax=8):
    rng = np.random.Generator(np.random.PCG64(seed))
    ks, amps, phis = [], [], []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3).astype(np.float64)
        nk = np.linalg.norm(k)
        if nk < 1 or nk > kmax:
            continue
        r = rng.normal(size=3)
        r -= k * (r @ k) / (k @ k)
        nr = np.linalg.norm(r)
        if nr < 1e-6:
            continue
        ks.append(k)
        amps.append(0.25 * r / nr)
        phis.append(rng.random() * 2 * math.pi)
    return np.array(ks), np.array(amps), np.array(phis)


def ucat_random_fourier(coor, IM, JM, KM, seed, nmodes=64, kmax=8, xp=None):
    """Divergence-free: every mode's amplitude is perpendicular to its wavevector."""
    xp = xp or _NumpyNS()
    ks, amps, phis = fourier_modes(seed, nmodes, kmax)
    cc = cell_centroids(coor, IM, JM, KM)
    u = xp.zeros((KM, JM, IM, 3))
    for m in range(nmodes):
        k0, k1, k2 = (float(v) for v in ks[m])
        ph = 2 * math.pi * (k0 * cc[..., 0] + k1 * cc[..., 1] + k2 * cc[..., 2]) + float(phis[m])
        cph = xp.cos(ph)
        for a in range(3):
            u[..., a] += float(amps[m, a]) * cph
    return to_node_slots(u, IM, JM, KM, xp)


# ---------------------------------------------------------------- particles
def particles_uniform_cube(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.random((n, 3))


def particles_uniform_tube(n, seed, R=TUBE_R, a=TUBE_A, sweep=TUBE_SWEEP, rho_max=RHO_MAX):
    """Uniform in the volume of the analytic tube section (rho <= rho_max * a).

    theta uniform, rho = a_max sqrt(u), phi uniform, accepted with
    probability (R + rho cos phi) / (R + a_max) (the volume element).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    amax = rho_max * a
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = int((n - out.shape[0]) * 1.2) + 64
        th = -sweep * rng.random(m)
        rho = amax * np.sqrt(rng.random(m))
        ph = 2 * math.pi * rng.random(m)
        acc = rng.random(m) * (R + amax) <= R + rho * np.cos(ph)
        rr = R + rho * np.cos(ph)
        pts = np.stack([rr * np.cos(th), rr * np.sin(th), rho * np.sin(ph)], axis=-1)[acc]
        out = np.concatenate([out, pts])
    return out[:n].copy()


# ---------------------------------------------------------------- configurations
def config(c: int, scale: float = 1.0, xp=None, with_particles: bool = True) -> dict:
    """Inputs of configuration c.  `scale` < 1 multiplies the number of cells
    and of particles by `scale` (each grid axis by scale**(1/3)), keeping the
    particles-per-cell density of the full configuration (scale = 1)."""
    xp = xp or _NumpyNS()
    axis = scale ** (1.0 / 3.0)

    def sz(v, mult=1, f=None):
        f = axis if f is None else f
        return max(mult, int(round(v * f / mult)) * mult)

    def npart(v):
        return max(1, int(round(v * scale)))

    d: dict = {"config": c}
    if c == 1:
        IM, JM, KM = 20, 20, 280
        coor = uniform_grid(IM, JM, KM, xp=xp)
        ucat = ucat_uniform(IM, JM, KM, (0.0, 0.0, 1.0), xp=xp)
        n = npart(10_000)
        d.update(dt=0.1, steps=10, rk2=False, deposit="cells", particles="cube")
    elif c in (2, 4):
        IM, JM, KM = (64, 64, 512) if c == 2 else (128, 128, 1024)
        IM, JM, KM = sz(IM, 2), sz(JM, 2), sz(KM)
        coor = tube_grid(IM, JM, KM, xp=xp)
        ucat = ucat_poiseuille(coor, IM, JM, KM, seed=2000 + c, xp=xp)
        n = npart(4_194_304) if c == 2 else npart(16_777_216)
        d.update(dt=0.5 * (TUBE_R * TUBE_SWEEP / KM) / 1.0, steps=50 if c == 2 else 10, rk2=(c == 2),
                 deposit="faces_volfrac" if c == 2 else "none", particles="tube")
        if c == 4:
            d["train_pairs"] = npart(2_097_152)
    elif c in (3, 5):
        IM, JM, KM = (128, 128, 1024) if c == 3 else (256, 256, 1024)
        IM, JM, KM = sz(IM), sz(JM), sz(KM)
        coor = perturbed_cube(IM, JM, KM, seed=3000 + c, xp=xp)
        if c == 3:
            ucat = ucat_taylor_green(coor, IM, JM, KM, xp=xp)
            n = npart(33_554_432)
            d.update(dt=0.5 / KM, steps=20, rk2=False, deposit="cells+faces", particles="cube")
        else:
            ucat = ucat_random_fourier(coor, IM, JM, KM, seed=2000 + c, xp=xp)
            n = npart(134_217_728)
            umax = float(abs(ucat).max())
            d.update(dt=0.5 * (1.0 / KM) / umax, steps=100, rk2=True, deposit="faces", particles="cube")
    else:
        raise ValueError(c)
    d.update(IM=IM, JM=JM, KM=KM, coor=coor, ucat=ucat, n=n)
    if with_particles:
        if d["particles"] == "cube":
            d["pos"] = particles_uniform_cube(n, 1000 + c)
        else:
            d["pos"] = particles_uniform_tube(n, 1000 + c)
    return d


def contravariant_from_cartesian(coor, ucat, IM, JM, KM):
    """Ucont [(KM+1)(JM+1)(IM+1)][3] of a node-slot Cartesian field: the flux
    through the i-face stored at node (i,j,k) (nodes i, j-1..j, k-1..k) is
    that face's area vector, half the cross product of its diagonals, dotted
    with ucat at the same slot; likewise j and k.  Face slots whose face is
    undefined (j = 0 or k = 0 for i-faces, etc.) get 0.  Contra2Cart of the
    result returns ucat to within the grid's variation over a cell, which is
    what the per-step field preparation of a synthetic workload needs.
    Works on numpy arrays and torch tensors alike."""
    N = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    U = ucat.reshape(KM + 1, JM + 1, IM + 1, 3)
    out = 0.0 * U  # same type, device and shape

    def cross_dot(d1, d2, u):
        cx = d1[..., 1] * d2[..., 2] - d1[..., 2] * d2[..., 1]
        cy = d1[..., 2] * d2[..., 0] - d1[..., 0] * d2[..., 2]
        cz = d1[..., 0] * d2[..., 1] - d1[..., 1] * d2[..., 0]
        return 0.5 * (cx * u[..., 0] + cy * u[..., 1] + cz * u[..., 2])

    out[1:, 1:, :, 0] = cross_dot(N[1:, 1:, :] - N[:-1, :-1, :], N[1:, :-1, :] - N[:-1, 1:, :], U[1:, 1:, :])
    out[1:, :, 1:, 1] = cross_dot(N[1:, :, 1:] - N[:-1, :, :-1], N[:-1, :, 1:] - N[1:, :, :-1], U[1:, :, 1:])
    out[:, 1:, 1:, 2] = cross_dot(N[:, 1:, 1:] - N[:, :-1, :-1], N[:, 1:, :-1] - N[:, :-1, 1:], U[:, 1:, 1:])
    return out.reshape(-1, 3)
