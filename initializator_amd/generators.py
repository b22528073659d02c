"""Seeded synthetic inputs for the solver: porosity masks, divergence-free
pore velocity fields, and pulsed-field-gradient schedules.

The reference reads these from files that are not shipped with it
(data/Poro.dat, data/Vx.dat, data/Vy.dat, data/Vz.dat); these generators
produce inputs of the same meaning:

* porosity: N(0,1) white noise -> Gaussian smoothing (sigma voxels,
  truncated at 4 sigma, reflecting edges) -> exactly round(phi*N) pores, the
  largest smoothed values (uint8, 1 = pore);
* velocity: three independent smoothed noise fields form a vector potential
  A on the cell edges of the staggered grid; A is zeroed on every edge that
  touches a solid cell or the domain wall, and v = curl A on the faces.  The
  discrete divergence of every cell is then zero to rounding, every
  pore-solid face and every wall face carries exactly zero flux, and the field
  is scaled so that the largest |face velocity| equals vmax;
* gradient: a pulsed-gradient spin-echo pair, +G for the first round(delta/dt)
  steps and -G for as many steps starting at step round(Delta/dt).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np
from scipy import ndimage

# Gyromagnetic constant of the reference (src/NmrRelaxationSolve.cpp:807).
G_REF = 4.78941714e7


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.default_rng([int(seed) & 0xFFFFFFFF, int(stream)])


def _smooth_noise(shape, seed: int, stream: int, sigma: float) -> np.ndarray:
    w = _rng(seed, stream).standard_normal(shape, dtype=np.float32)
    return ndimage.gaussian_filter(w, sigma=sigma, mode="reflect", truncate=4.0)


def porosity(shape, phi: float, seed: int, sigma: float = 3.0) -> np.ndarray:
    """uint8 pore mask of shape (nz, ny, nx) with exactly round(phi*N) pores."""
    if not 0.0 <= phi <= 1.0:
        raise ValueError("phi must lie in [0, 1]")
    s = _smooth_noise(shape, seed, 0, sigma).ravel()
    n = s.size
    npore = int(round(phi * n))
    mask = np.zeros(n, dtype=np.uint8)
    if npore > 0:
        idx = np.argpartition(s, n - npore)[n - npore :]
        mask[idx] = 1
    return mask.reshape(shape)


def _wet_edges(mask: np.ndarray):
    """Edge masks: an edge is wet when all four cells around it are pores
    (cells outside the domain count as solid)."""
    nz, ny, nx = mask.shape
    p = np.zeros((nz + 2, ny + 2, nx + 2), dtype=bool)
    p[1:-1, 1:-1, 1:-1] = mask != 0
    # Ax edge at (i, j-1/2, k-1/2), j = 0..ny, k = 0..nz: cells (i, j-1|j, k-1|k)
    ax = p[:-1, :-1, 1:-1] & p[:-1, 1:, 1:-1] & p[1:, :-1, 1:-1] & p[1:, 1:, 1:-1]
    # Ay edge at (i-1/2, j, k-1/2): cells (i-1|i, j, k-1|k)
    ay = p[:-1, 1:-1, :-1] & p[:-1, 1:-1, 1:] & p[1:, 1:-1, :-1] & p[1:, 1:-1, 1:]
    # Az edge at (i-1/2, j-1/2, k): cells (i-1|i, j-1|j, k)
    az = p[1:-1, :-1, :-1] & p[1:-1, :-1, 1:] & p[1:-1, 1:, :-1] & p[1:-1, 1:, 1:]
    return ax, ay, az  # shapes (nz+1, ny+1, nx), (nz+1, ny, nx+1), (nz, ny+1, nx+1)


def curl_faces(ax, ay, az, d=(1.0, 1.0, 1.0)):
    """Face velocities of v = curl A for edge potentials A (see _wet_edges for
    the staggering); d = (dx, dy, dz)."""
    dx, dy, dz = d
    vx = (az[:, 1:, :] - az[:, :-1, :]) / dy - (ay[1:, :, :] - ay[:-1, :, :]) / dz
    vy = (ax[1:, :, :] - ax[:-1, :, :]) / dz - (az[:, :, 1:] - az[:, :, :-1]) / dx
    vz = (ay[:, :, 1:] - ay[:, :, :-1]) / dx - (ax[:, 1:, :] - ax[:, :-1, :]) / dy
    return vx, vy, vz


def divergence(vx, vy, vz, d=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Cell divergence of compact face velocities."""
    dx, dy, dz = d
    return (
        (vx[:, :, 1:] - vx[:, :, :-1]) / dx
        + (vy[:, 1:, :] - vy[:, :-1, :]) / dy
        + (vz[1:, :, :] - vz[:-1, :, :]) / dz
    )


def velocity(mask: np.ndarray, seed: int, sigma: float = 4.0, vmax: float = 1e-3, d=(1e-6, 1e-6, 1e-6)):
    """Divergence-free face velocities (FP64) confined to the pores."""
    nz, ny, nx = mask.shape
    wx, wy, wz = _wet_edges(mask)
    ax = _smooth_noise((nz + 1, ny + 1, nx), seed, 1, sigma).astype(np.float64) * wx
    ay = _smooth_noise((nz + 1, ny, nx + 1), seed, 2, sigma).astype(np.float64) * wy
    az = _smooth_noise((nz, ny + 1, nx + 1), seed, 3, sigma).astype(np.float64) * wz
    vx, vy, vz = curl_faces(ax, ay, az, d)
    peak = max(np.abs(vx).max(initial=0.0), np.abs(vy).max(initial=0.0), np.abs(vz).max(initial=0.0))
    if peak > 0:
        s = vmax / peak
        vx *= s
        vy *= s
        vz *= s
    return vx, vy, vz


def pgse_schedule(nsteps: int, dt: float, g: float, delta: float, big_delta: float, axis: int = 2) -> np.ndarray:
    """[nsteps, 3] gradient amplitudes (T/m): +g on steps [0, nd), -g on
    steps [nD, nD+nd), nd = round(delta/dt), nD = round(Delta/dt)."""
    sched = np.zeros((nsteps, 3))
    nd = int(round(delta / dt))
    nD = int(round(big_delta / dt))
    sched[: min(nd, nsteps), axis] = g
    if nD < nsteps:
        sched[nD : min(nD + nd, nsteps), axis] = -g
    return sched


def b_value(g: float, delta: float, big_delta: float, gamma: float = G_REF) -> float:
    """Stejskal-Tanner b = gamma^2 g^2 delta^2 (Delta - delta/3) (s/m^2)."""
    return gamma**2 * g**2 * delta**2 * (big_delta - delta / 3.0)


def gradient_for_b(b: float, delta: float, big_delta: float, gamma: float = G_REF) -> float:
    return float(np.sqrt(b / (gamma**2 * delta**2 * (big_delta - delta / 3.0))))


def stable_dt(dx: float, diffusivity: float, frac: float = 0.4) -> float:
    """frac x the explicit FTCS limit dx^2/(6 D) of the 7-point stencil."""
    return frac * dx * dx / (6.0 * diffusivity)


@dataclass
class Case:
    """One synthetic problem: compact inputs plus physics scalars."""

    shape: tuple
    mask: np.ndarray
    vx: np.ndarray | None
    vy: np.ndarray | None
    vz: np.ndarray | None
    dx: float
    dt: float
    diffusivity: float
    t2_bulk: float
    t1_bulk: float
    rho: float
    nsteps: int
    grad_sched: np.ndarray | None
    ncomp: int

    @property
    def nvox(self) -> int:
        return int(np.prod(self.shape))


def case_small(seed: int = 0x4E4D52 + 0, n: int = 64, nsteps: int = 2000, phi: float = 0.30) -> Case:
    """Plain method, small: n^3, 1 um voxels, 30% pores, no velocity, no
    gradient, D = 2.3e-9, T2 = 2 s, rho = 10 um/s, real M = 1 in pores."""
    shape = (n, n, n)
    dx, D = 1e-6, 2.3e-9
    return Case(shape, porosity(shape, phi, seed), None, None, None, dx, stable_dt(dx, D), D, 2.0, 2.5e10,
                1e-5, nsteps, None, 1)


def case_flow(seed: int = 0x4E4D52 + 1000, n: int = 256, phi: float = 0.20, g: float = 0.3, nsteps: int = 5000,
              delta: float = 2e-3, big_delta: float = 20e-3, rho: float = 10e-6, shape=None, progress=None) -> Case:
    """Flow + PFG: porosity phi, divergence-free velocity (max 1 mm/s), PGSE
    gradient g along z, complex (Mx, My).  progress(msg), if given, is
    called after each generation stage (large grids take minutes)."""
    shape = tuple(shape) if shape is not None else (n, n, n)
    dx, D = 1e-6, 2.3e-9
    mask = porosity(shape, phi, seed)
    if progress:
        progress(f"porosity {shape} done")
    vx, vy, vz = velocity(mask, seed, vmax=1e-3, d=(dx, dx, dx))
    if progress:
        progress(f"velocity {shape} done")
    dt = stable_dt(dx, D)
    sched = pgse_schedule(nsteps, dt, g, delta, big_delta)
    return Case(shape, mask, vx, vy, vz, dx, dt, D, 2.0, 2.5e10, rho, nsteps, sched, 2)


def batch_instance(i: int) -> dict:
    """Instance i of the batched workload: cube edge U{32..128}, porosity
    U(0.15, 0.40), rho U(1, 50) um/s, b log-uniform on [1e7, 1e10] s/m^2 with
    delta 2 ms and Delta 20 ms, divergence-free pore velocity (max 1 mm/s),
    1 um voxels; seed 0x4E4D52 + 4000 + i."""
    seed = 0x4E4D52 + 4000 + i
    r = np.random.default_rng(seed)
    n = int(r.integers(32, 129))
    phi = float(r.uniform(0.15, 0.40))
    rho = float(r.uniform(1e-6, 50e-6))
    b = float(10.0 ** r.uniform(7.0, 10.0))
    shape = (n, n, n)
    mask = porosity(shape, phi, seed)
    vel = velocity(mask, seed, vmax=1e-3, d=(1e-6, 1e-6, 1e-6))
    g = gradient_for_b(b, 2e-3, 20e-3)
    return dict(shape=shape, phi=phi, rho1=rho, rho2=rho, grad_amp=g, delta=2e-3, big_delta=20e-3, b=b,
                mask=mask, vel=vel)


def batch_edges(n: int = 256) -> list:
    """Cube edges of the first n batched-workload instances (the first draw
    of batch_instance's generator), without generating the instances."""
    return [int(np.random.default_rng(0x4E4D52 + 4000 + i).integers(32, 129)) for i in range(n)]


def batch_instances(n: int = 256, workers: int = 16) -> list:
    """The first n batched-workload instances, generated in a pool of worker
    processes started fresh (spawn: safe after the GPU is in use)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    workers = max(1, min(workers, os.cpu_count() or 1, n))
    if workers == 1:
        return [batch_instance(i) for i in range(n)]
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(batch_instance, range(n)))
