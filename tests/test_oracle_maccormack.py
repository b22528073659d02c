"""Pins the oracle's MacCormack step (or_maccormack, the GPU path's
ExplicitMakUpgrade, NmrRelaxationSolveKernels.cu:637-967, 1473-1501) and its
boundary conditions (periodic, zero value; cu:969-1053) against closed forms
and invariants.  The race of the reference's in-place solid ghost fill is
resolved as "a solid neighbour reads as the pore's own value"; the uniform
fixed point next to solids of arbitrary value pins that choice.
"""
import numpy as np
import pytest

import oracle
from initializator_amd import generators, layout

MAK, UPW = oracle.MACCORMACK, oracle.UPWIND
PER, DIR = oracle.BC_PERIODIC, oracle.BC_DIRICHLET


def _setup(shape, mask, *, dx=1e-6, dt=1e-5, d_water=2.3e-9, t1=2.5e10, t2=2.5e10, r2=0.0, diffusion=True,
           velocity=False, precession=False, p_eq=0.0, dy=None, scheme=MAK, bc=(0, 0, 0)):
    p = oracle.params(shape, dx, dt, dy=dy, diffusion=diffusion, velocity=velocity, precession=precession,
                      p_eq=p_eq, scheme=scheme, bc=bc)
    act = layout.pad_mask(mask)
    D, T1, T2, _ = oracle.fill_coefficients(p, act, d_water, t1, t2, 0.0, r2)
    return p, act, D, T1, T2


def _run(p, act, D, T1, T2, m, n, vel=None, sched=None):
    vel = vel if vel is not None else [np.zeros(oracle.padded_shape(p)) for _ in range(3)]
    return oracle.run(p, act, D, T1, T2, vel, n, m, grad_sched=sched)


def _field(p, compact):
    return layout.pad_field(np.asarray(compact, dtype=np.float64))


def test_relaxation_heun_factors():
    """Relaxation alone: predictor + averaged corrector is Heun's method, so
    Mx decays by (1 - a + a^2/2)^n, a = dt/T2, and Mz - pEq by the same
    polynomial in dt/T1."""
    shape = (4, 5, 6)
    dt, T1, T2, peq = 1e-4, 0.2, 0.01, 0.3
    p, act, D, t1, t2 = _setup(shape, np.ones(shape, np.uint8), dt=dt, t1=T1, t2=T2, p_eq=peq)
    m = [np.full(oracle.padded_shape(p), v) for v in (1.0, -0.4, 0.9)]
    n = 50
    _run(p, act, D, t1, t2, m, n)
    a, b = dt / T2, dt / T1
    h2, h1 = (1 - a + a * a / 2) ** n, (1 - b + b * b / 2) ** n
    np.testing.assert_allclose(layout.unpad(m[0]), h2, rtol=1e-13)
    np.testing.assert_allclose(layout.unpad(m[1]), -0.4 * h2, rtol=1e-13)
    np.testing.assert_allclose(layout.unpad(m[2]), peq + (0.9 - peq) * h1, rtol=1e-13)


def test_precession_heun_factor():
    """Gradient G along z alone: m = Mx + iMy evolves as
    m (1 - i th - th^2/2)^n per plane, th = dt g G dz (k - 1/2) (padded k)."""
    shape = (5, 2, 3)
    dt, G, dz = 2e-5, 0.4, 1e-6
    p, act, D, t1, t2 = _setup(shape, np.ones(shape, np.uint8), dt=dt, diffusion=False, precession=True)
    m = [np.ones(oracle.padded_shape(p)), np.zeros(oracle.padded_shape(p)), np.zeros(oracle.padded_shape(p))]
    n = 25
    _run(p, act, D, t1, t2, m, n, sched=np.tile([0.0, 0.0, G], (n, 1)))
    k = np.arange(1, shape[0] + 1)
    th = dt * oracle.G_REF * G * dz * (k - 0.5)
    want = (1 - 1j * th - th**2 / 2) ** n
    got = layout.unpad(m[0])[:, 0, 0] + 1j * layout.unpad(m[1])[:, 0, 0]
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_neumann_cosine_eigenmode_heun():
    """Diffusion alone on an all-pore box with zero-derivative ghosts: the
    cosine eigenvector decays by (1 - l + l^2/2)^n with the FTCS eigenvalue
    l = 4 D dt (sin^2(pi/2N)/dx^2 + sin^2(pi/2M)/dy^2)."""
    nz, ny, nx = 2, 6, 9
    dx, dy, Dw, dt = 1e-6, 1.3e-6, 2.3e-9, 4e-5
    p, act, D, t1, t2 = _setup((nz, ny, nx), np.ones((nz, ny, nx), np.uint8), dx=dx, dy=dy, dt=dt)
    i, j = np.arange(nx), np.arange(ny)
    mode = np.cos(np.pi * (j[:, None] + 0.5) / ny) * np.cos(np.pi * (i[None, :] + 0.5) / nx)
    m0 = np.broadcast_to(mode, (nz, ny, nx))
    m = [_field(p, m0), np.zeros(oracle.padded_shape(p)), np.zeros(oracle.padded_shape(p))]
    n = 30
    _run(p, act, D, t1, t2, m, n)
    lam = 4 * Dw * dt * (np.sin(np.pi / (2 * nx)) ** 2 / dx**2 + np.sin(np.pi / (2 * ny)) ** 2 / dy**2)
    np.testing.assert_allclose(layout.unpad(m[0]), m0 * (1 - lam + lam**2 / 2) ** n, rtol=0, atol=3e-13)


@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_zero_value_boundary_sine_eigenmode(scheme):
    """Zero-value ghosts on both x ends (BC_Advanced with Val 0): the
    eigenvector is sin(pi (i+1)/(N+1)), l = 4 D dt sin^2(pi/(2(N+1)))/dx^2;
    forward Euler decays it by (1 - l)^n, MacCormack by (1 - l + l^2/2)^n."""
    nz, ny, nx = 2, 3, 11
    dx, Dw, dt = 1e-6, 2.3e-9, 5e-5
    p, act, D, t1, t2 = _setup((nz, ny, nx), np.ones((nz, ny, nx), np.uint8), dx=dx, dt=dt, scheme=scheme,
                               bc=(DIR, 0, 0))
    mode = np.sin(np.pi * (np.arange(nx) + 1) / (nx + 1))
    m0 = np.broadcast_to(mode, (nz, ny, nx))
    m = [_field(p, m0), np.zeros(oracle.padded_shape(p)), np.zeros(oracle.padded_shape(p))]
    n = 20
    _run(p, act, D, t1, t2, m, n)
    lam = 4 * Dw * dt * np.sin(np.pi / (2 * (nx + 1))) ** 2 / dx**2
    f = (1 - lam) if scheme == UPW else (1 - lam + lam**2 / 2)
    np.testing.assert_allclose(layout.unpad(m[0]), m0 * f**n, rtol=0, atol=2e-13)


def _periodic_line(scheme, nx, sigma, xi, n):
    dx, dt = 1e-6, 1e-5
    u = sigma * dx / dt
    shape = (1, 1, nx)
    p, act, D, t1, t2 = _setup(shape, np.ones(shape, np.uint8), dx=dx, dt=dt, diffusion=False, velocity=True,
                               scheme=scheme, bc=(PER, 0, 0))
    vx = np.zeros(oracle.padded_shape(p))
    vx[1, 1, 1 : nx + 2] = u
    vel = [vx, np.zeros_like(vx), np.zeros_like(vx)]
    j = np.arange(nx)
    m = [_field(p, np.cos(xi * j).reshape(shape)), np.zeros_like(vx), np.zeros_like(vx)]
    _run(p, act, D, t1, t2, m, n, vel=vel)
    return m[0][1, 1, 1:-1], j


@pytest.mark.parametrize("sigma", [0.3, -0.45])
def test_periodic_advection_lax_wendroff_factor(sigma):
    """Constant velocity on a periodic line: MacCormack is Lax-Wendroff, the
    Fourier mode e^{i xi j} is multiplied by 1 - i s sin(xi) - s^2 (1 - cos xi)
    per step, s = u dt/dx (either sign)."""
    nx, n = 16, 7
    xi = 2 * np.pi * 3 / nx
    got, j = _periodic_line(MAK, nx, sigma, xi, n)
    G = 1 - 1j * sigma * np.sin(xi) - sigma**2 * (1 - np.cos(xi))
    np.testing.assert_allclose(got, np.real(G**n * np.exp(1j * xi * j)), rtol=0, atol=1e-13)


@pytest.mark.parametrize("sigma", [0.3, -0.45])
def test_periodic_advection_upwind_factor(sigma):
    """The same line under the upwind step: the mode's factor is
    1 - s (1 - e^{-i xi}) for s > 0 and 1 - s (e^{i xi} - 1) for s < 0."""
    nx, n = 16, 7
    xi = 2 * np.pi * 2 / nx
    got, j = _periodic_line(UPW, nx, sigma, xi, n)
    G = 1 - sigma * (1 - np.exp(-1j * xi)) if sigma > 0 else 1 - sigma * (np.exp(1j * xi) - 1)
    np.testing.assert_allclose(got, np.real(G**n * np.exp(1j * xi * j)), rtol=0, atol=1e-13)


def test_cell_centred_velocity_depends_on_face_sums_only():
    """MacCormack advects with the cell-centred velocity (um + up)/2: adding
    +c, -c, +c, ... to consecutive faces keeps every cell's sum and so the
    whole step."""
    nx, dx, dt = 12, 1e-6, 1e-5
    shape = (1, 1, nx)
    p, act, D, t1, t2 = _setup(shape, np.ones(shape, np.uint8), dx=dx, dt=dt, diffusion=False, velocity=True)
    rng = np.random.default_rng(4)
    faces = rng.uniform(-0.02, 0.02, nx + 1)
    m0 = rng.random(shape)
    out = []
    for f in (faces, faces + 0.013 * (-1.0) ** np.arange(nx + 1)):
        vx = np.zeros(oracle.padded_shape(p))
        vx[1, 1, 1 : nx + 2] = f
        m = [_field(p, m0), np.zeros_like(vx), np.zeros_like(vx)]
        _run(p, act, D, t1, t2, m, 3, vel=[vx, np.zeros_like(vx), np.zeros_like(vx)])
        out.append(layout.unpad(m[0]))
    np.testing.assert_allclose(out[0], out[1], rtol=0, atol=1e-15)
    assert not np.allclose(out[0], m0)


def test_two_pore_channel_hand_values():
    """Pores a, b between two solids, uniform x velocity s = u dt/dx, Ma = 1,
    Mb = 0, no diffusion.  Predictor (backward differences, solid reads as
    own value): Pa = 1, Pb = s.  Corrector (forward differences):
    Ca = (1 + Pa - s (Pb - Pa))/2 = 1 + s/2 - s^2/2, Cb = (0 + Pb)/2 = s/2.
    (Forward-first ordering would give 1 + s/2 and s/2 + s^2/2.)"""
    s, dx, dt = 0.4, 1e-6, 1e-5
    mask = np.array([[[0, 1, 1, 0]]], np.uint8)
    p, act, D, t1, t2 = _setup(mask.shape, mask, dx=dx, dt=dt, diffusion=False, velocity=True)
    vx = np.zeros(oracle.padded_shape(p))
    vx[1, 1, 1:6] = s * dx / dt
    m = [_field(p, np.array([[[0.0, 1.0, 0.0, 0.0]]])), np.zeros_like(vx), np.zeros_like(vx)]
    _run(p, act, D, t1, t2, m, 1, vel=[vx, np.zeros_like(vx), np.zeros_like(vx)])
    got = layout.unpad(m[0])[0, 0]
    np.testing.assert_allclose(got[1:3], [1 + s / 2 - s * s / 2, s / 2], rtol=1e-14)


def test_uniform_fixed_point_next_to_arbitrary_solids():
    """A uniform pore field stays uniform under any velocity (pore-solid
    faces included) and diffusion when a solid neighbour reads as the pore's
    own value, whatever the solids hold; the solids themselves end at 0 (the
    corrector's value there)."""
    shape = (9, 8, 10)
    mask = generators.porosity(shape, 0.5, seed=21, sigma=1.2)
    rng = np.random.default_rng(8)
    p, act, D, t1, t2 = _setup(shape, mask, dt=2e-5, velocity=True)
    vel = [np.ascontiguousarray(v) for v in layout.pad_velocity(
        rng.uniform(-0.05, 0.05, (9, 8, 11)), rng.uniform(-0.05, 0.05, (9, 9, 10)),
        rng.uniform(-0.05, 0.05, (10, 8, 10)), shape)]
    solid_vals = rng.uniform(-50, 50, shape)
    m0 = np.where(mask != 0, 0.7, solid_vals)
    m = [_field(p, m0), np.zeros(oracle.padded_shape(p)), np.zeros(oracle.padded_shape(p))]
    vx, vy, vz = vel
    pc = [np.zeros(oracle.padded_shape(p)) for _ in range(6)]
    nf = oracle.fill_gradient(p, (0.0, 0.0, 0.0))
    for f in m:
        oracle.bc_field(p, f)
    for _ in range(4):
        oracle.maccormack(p, act, D, t1, t2, vel, nf, m, pc)
    out = layout.unpad(m[0])
    np.testing.assert_allclose(out[mask != 0], 0.7, rtol=0, atol=1e-14)
    assert np.all(out[mask == 0] == 0.0)
