"""Pins the oracle's y and z terms and the pulse flags 3 and 4.

The x-axis checks of test_oracle.py / test_oracle_maccormack.py leave the y
and z advection terms of the update (ExplicitSimple cpp:3236-3247; the
MacCormack stages cu:637-967), the y-gradient precession (cpp:3230-3235) and
AddImpulse flags 3 and 4 (cpp:2583-2597) open.  Here:

* closed forms along y and z: 1-D upwind lines of either sign, a zero face
  that blocks transport, periodic Fourier modes at the upwind and
  Lax-Wendroff amplification factors, the MacCormack two-pore channel;
* precession under a y gradient: w = Mz + iMx turns by (1 - i theta)^n
  (upwind) and by the Heun factor (MacCormack);
* axis permutation: a random anisotropic 3-D problem (mask, non-uniform face
  velocities with zero faces, spacings, gradient, tilted main field,
  surface relaxation, diffusion) whose axes are permuted gives the permuted
  result -- for cyclic permutations directly, for odd ones with the
  gradient negated (a reflection flips the cross product M x Bg);
* pulse flags 3 and 4: the rotation angle measured from the rotated vectors
  against g |a| |(r - 1e13 d) . a_hat| and -g |(r - r_c) * a|, and the
  non-unit axis that flag 4 passes to OrtTransform unnormalised.

A plausible slip in the oracle (swapping vy[c] and vy[yp], flipping the sign
of an hy or hz term, taking dx for dy, a wrong pulse centre or distance
sign) fails at least one of these.
"""
import numpy as np
import pytest

import oracle
from initializator_amd import generators, layout

MAK, UPW = oracle.MACCORMACK, oracle.UPWIND
PER = oracle.BC_PERIODIC


def _line(axis, n, scheme, faces, m0, *, d=1e-6, dt=1e-5, bc=0, nsteps=1):
    """One field on a 1-D grid along `axis` (0 x, 1 y, 2 z) with the given
    face velocities (n + 1 of them), advection only."""
    size = [1, 1, 1]
    size[axis] = n
    shape = (size[2], size[1], size[0])
    bcs = [0, 0, 0]
    bcs[axis] = bc
    p = oracle.params(shape, d, dt, diffusion=False, velocity=True, scheme=scheme, bc=bcs)
    act = layout.pad_mask(np.ones(shape, np.uint8))
    D, t1, t2, _ = oracle.fill_coefficients(p, act, 2.3e-9, 2.5e10, 2.5e10, 0.0, 0.0)
    vel = [np.zeros(oracle.padded_shape(p)) for _ in range(3)]
    sl = [1, 1, 1]
    sl[2 - axis] = slice(1, n + 2)
    vel[axis][tuple(sl)] = faces
    m = [layout.pad_field(np.asarray(m0, float).reshape(shape)), np.zeros_like(vel[0]), np.zeros_like(vel[0])]
    oracle.run(p, act, D, t1, t2, vel, nsteps, m)
    out = layout.unpad(m[0])
    return out.reshape(-1)


@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("u", [2e-3, -2e-3])
def test_upwind_line_along_y_and_z(axis, u):
    """Uniform velocity along y or z: m_i -= c (m_i - m_{i-1}) for u > 0 and
    m_i -= c (m_{i+1} - m_i) for u < 0, c = u dt / d, zero-derivative ghosts
    at the ends (cpp:3236-3247 with the roles of x taken by y or z)."""
    n, dt, d = 9, 1e-4, 1e-6
    rng = np.random.default_rng(2 + axis)
    m0 = rng.random(n)
    got = _line(axis, n, UPW, np.full(n + 1, u), m0, d=d, dt=dt)
    c = u * dt / d
    ext = np.concatenate([[m0[0]], m0, [m0[-1]]])
    want = m0 - c * (ext[1:-1] - ext[:-2]) if u > 0 else m0 - c * (ext[2:] - ext[1:-1])
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("axis", [1, 2])
def test_zero_face_blocks_transport_along_y_and_z(axis):
    """A face of zero velocity between cells 3 and 4 of a y or z line: cells
    downstream of it do not see a change upstream of it (the face indicator
    v / (v + 1e-50) of the face itself, not of its neighbour)."""
    n = 8
    faces = np.full(n + 1, 3e-3)
    faces[4] = 0.0
    rng = np.random.default_rng(9 + axis)
    a = rng.random(n)
    b = a.copy()
    b[:4] += rng.random(4)
    ra = _line(axis, n, UPW, faces, a, dt=1e-4)
    rb = _line(axis, n, UPW, faces, b, dt=1e-4)
    np.testing.assert_array_equal(ra[4:], rb[4:])
    assert not np.allclose(ra[:4], rb[:4])
    # flow the other way: cells upstream (now 4..7) drive cells 0..3 only
    # through the blocked face, so perturbing 4..7 leaves 0..3 alone
    c = a.copy()
    c[4:] += rng.random(4)
    ra = _line(axis, n, UPW, -faces, a, dt=1e-4)
    rc = _line(axis, n, UPW, -faces, c, dt=1e-4)
    np.testing.assert_array_equal(ra[:4], rc[:4])


@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("sigma", [0.3, -0.45])
def test_periodic_modes_along_y_and_z(axis, sigma):
    """Constant velocity on a periodic y or z line: the Fourier mode e^{i xi
    j} is multiplied per step by the upwind factor 1 - s(1 - e^{-i xi})
    (s > 0) or 1 - s(e^{i xi} - 1) (s < 0) under ExplicitSimple, and by the
    Lax-Wendroff factor 1 - i s sin xi - s^2 (1 - cos xi) under MacCormack."""
    n, steps, d, dt = 16, 7, 1e-6, 1e-5
    xi = 2 * np.pi * 3 / n
    j = np.arange(n)
    faces = np.full(n + 1, sigma * d / dt)
    for scheme in (UPW, MAK):
        got = _line(axis, n, scheme, faces, np.cos(xi * j), d=d, dt=dt, bc=PER, nsteps=steps)
        if scheme == MAK:
            G = 1 - 1j * sigma * np.sin(xi) - sigma**2 * (1 - np.cos(xi))
        else:
            G = 1 - sigma * (1 - np.exp(-1j * xi)) if sigma > 0 else 1 - sigma * (np.exp(1j * xi) - 1)
        np.testing.assert_allclose(got, np.real(G**steps * np.exp(1j * xi * j)), rtol=0, atol=1e-13)


@pytest.mark.parametrize("axis", [1, 2])
def test_maccormack_two_pore_channel_along_y_and_z(axis):
    """Two pores between solids along y or z, s = u dt/d, Ma = 1, Mb = 0:
    predictor (backward differences, a solid reads as the pore's own value)
    Pa = 1, Pb = s; corrector Ca = 1 + s/2 - s^2/2, Cb = s/2."""
    s, d, dt = 0.4, 1e-6, 1e-5
    size = [1, 1, 1]
    size[axis] = 4
    shape = (size[2], size[1], size[0])
    mask = np.array([0, 1, 1, 0], np.uint8).reshape(shape)
    p = oracle.params(shape, d, dt, diffusion=False, velocity=True, scheme=MAK)
    act = layout.pad_mask(mask)
    D, t1, t2, _ = oracle.fill_coefficients(p, act, 2.3e-9, 2.5e10, 2.5e10, 0.0, 0.0)
    vel = [np.zeros(oracle.padded_shape(p)) for _ in range(3)]
    sl = [1, 1, 1]
    sl[2 - axis] = slice(1, 6)
    vel[axis][tuple(sl)] = s * d / dt
    m = [layout.pad_field(np.array([0.0, 1.0, 0.0, 0.0]).reshape(shape)), np.zeros_like(vel[0]),
         np.zeros_like(vel[0])]
    oracle.run(p, act, D, t1, t2, vel, 1, m)
    got = layout.unpad(m[0]).reshape(-1)
    np.testing.assert_allclose(got[1:3], [1 + s / 2 - s * s / 2, s / 2], rtol=1e-14)


@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_precession_y_gradient(scheme):
    """Gradient G along y alone, no relaxation or transport: M x Bg with Bg =
    (0, G dy (j - 1/2), 0) turns w = Mz + iMx by (1 - i th) per upwind step
    and by the Heun factor 1 - i th - th^2/2 per MacCormack step,
    th = dt g G dy (j - 1/2) with padded j (cpp:1970-1972, 3230-3235)."""
    shape = (2, 6, 3)
    dy, dt, G = 1e-6, 2e-5, 0.25
    p = oracle.params(shape, 1e-6, dt, dy=dy, diffusion=False, precession=True, scheme=scheme)
    act = layout.pad_mask(np.ones(shape, np.uint8))
    D, t1, t2, _ = oracle.fill_coefficients(p, act, 2.3e-9, 2.5e10, 2.5e10, 0.0, 0.0)
    vel = [np.zeros(oracle.padded_shape(p)) for _ in range(3)]
    m = [np.full(oracle.padded_shape(p), 0.4), np.full(oracle.padded_shape(p), 0.3),
         np.full(oracle.padded_shape(p), -0.7)]
    n = 13
    oracle.run(p, act, D, t1, t2, vel, n, m, grad_const=(0.0, G, 0.0))
    for j in range(1, 7):
        th = dt * oracle.G_REF * G * dy * (j - 0.5)
        f = (1 - 1j * th) if scheme == UPW else (1 - 1j * th - th * th / 2)
        want = (-0.7 + 0.4j) * f**n
        got = m[2][1, j, 1] + 1j * m[0][1, j, 1]
        assert got == pytest.approx(want, rel=1e-12)
        assert m[1][1, j, 1] == pytest.approx(0.3, rel=1e-14)  # My is the rotation axis


def _permute_problem(prob, perm, flip):
    """The problem with new axis a taken from old axis perm[a]; flip negates
    the gradient (for odd permutations)."""
    def arr(f):  # compact [z][y][x] array -> permuted
        return np.ascontiguousarray(np.transpose(f, [2 - perm[2 - dim] for dim in range(3)]))

    out = dict(prob)
    out["mask"] = arr(prob["mask"])
    out["vel"] = [arr(prob["vel"][perm[a]]) for a in range(3)]
    out["m"] = [arr(prob["m"][perm[a]]) for a in range(3)]
    out["d"] = [prob["d"][perm[a]] for a in range(3)]
    out["g"] = [(-1.0 if flip else 1.0) * prob["g"][perm[a]] for a in range(3)]
    out["b0"] = [prob["b0"][perm[a]] for a in range(3)]
    out["bc"] = [prob["bc"][perm[a]] for a in range(3)]
    return out


def _solve(prob, scheme, nsteps):
    mask = prob["mask"]
    shape = mask.shape
    dx, dy, dz = prob["d"]
    p = oracle.params(shape, dx, prob["dt"], dy=dy, dz=dz, precession=True, velocity=True, diffusion=True,
                      p_eq=0.3, b0=prob["b0"], scheme=scheme, bc=prob["bc"])
    act = layout.pad_mask(mask)
    D, t1, t2, _ = oracle.fill_coefficients(p, act, 2.3e-9, 0.5, 0.05, 2e-5, 3e-5)
    vel = [np.ascontiguousarray(v) for v in layout.pad_velocity(*prob["vel"], shape)]
    m = [layout.pad_field(f) for f in prob["m"]]
    sig = oracle.run(p, act, D, t1, t2, vel, nsteps, m, grad_const=tuple(prob["g"]))
    return [layout.unpad(f) for f in m], sig


def _anisotropic_problem(seed, bc=(0, 0, 0)):
    rng = np.random.default_rng(seed)
    shape = (5, 7, 6)  # nz, ny, nx
    nz, ny, nx = shape
    mask = generators.porosity(shape, 0.6, seed=seed, sigma=1.0)

    def faces(sh):
        v = rng.uniform(-3e-3, 3e-3, sh)
        v[rng.random(sh) < 0.2] = 0.0  # zero faces switch the upwind indicators
        return v

    vel = [faces((nz, ny, nx + 1)), faces((nz, ny + 1, nx)), faces((nz + 1, ny, nx))]
    m = [np.where(mask != 0, rng.uniform(-1, 1, shape), rng.uniform(-2, 2, shape)) for _ in range(3)]
    return dict(mask=mask, vel=vel, m=m, d=[1.0e-6, 1.3e-6, 0.8e-6], g=[0.11, -0.23, 0.31],
                b0=[0.2, -0.3, 0.93], bc=list(bc), dt=2e-5)


@pytest.mark.parametrize("scheme", [UPW, MAK])
@pytest.mark.parametrize("perm,flip", [((1, 2, 0), False), ((2, 0, 1), False), ((1, 0, 2), True),
                                       ((0, 2, 1), True), ((2, 1, 0), True)])
def test_axis_permutation_gives_the_permuted_result(scheme, perm, flip):
    """Relabelling the axes of a random anisotropic problem relabels its
    result: every term of the update treats the three axes alike.  Swapping
    the roles of a face pair (vy[c] / vy[yp]), a sign or a spacing on one
    axis breaks this."""
    prob = _anisotropic_problem(17, bc=(0, PER, 0) if scheme == MAK else (0, 0, 0))
    n = 4
    base, sig = _solve(prob, scheme, n)
    q = _permute_problem(prob, perm, flip)
    got, sig_q = _solve(q, scheme, n)
    inv = [perm.index(d) for d in range(3)]
    for a in range(3):
        back = np.transpose(got[a], [2 - inv[2 - dim] for dim in range(3)])
        np.testing.assert_allclose(back, base[perm[a]], rtol=0, atol=1e-12 * np.abs(base[perm[a]]).max())
    assert not np.allclose(base[0], prob["m"][0])
    # the signal components are permuted the same way
    np.testing.assert_allclose(sig_q, sig[:, list(perm)], rtol=1e-12, atol=1e-15)


def _angle_about(v0, v1, n):
    """Rotation angle (-pi, pi] taking v0 to v1 about the unit axis n."""
    n = np.asarray(n, float)
    p0 = v0 - n * (v0 @ n)
    p1 = v1 - n * (v1 @ n)
    return np.arctan2(n @ np.cross(p0, p1), p0 @ p1)


def _wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def test_pulse_flag3_far_origin_gradient():
    """Flag 3: every pore turns about a_hat by g |a| |sum_d (idx_d - 1 - 1e13)
    d_d a_hat_d| (padded indices; cpp:2583-2589).  With the axis along z,
    |a| = 3 and dz = 2^-60 the angle is about 1250 rad and moves by g*3*dz
    (1.25e-10 rad) per plane, so the far-origin distance, its sign and the
    index offset are all visible."""
    shape = (6, 3, 4)
    dz = 2.0**-60
    p = oracle.params(shape, 1e-6, 1e-5, dz=dz)
    rng = np.random.default_rng(31)
    mask = (rng.random(shape) < 0.7).astype(np.uint8)
    act = layout.pad_mask(mask)
    m0 = [layout.pad_field(rng.standard_normal(shape)) for _ in range(3)]
    m = [x.copy() for x in m0]
    oracle.add_impulse(p, act, 3, 0.0, (0.0, 0.0, 3.0), m)
    g = oracle.G_REF
    for (k, j, i) in np.argwhere(mask):
        c = (k + 1, j + 1, i + 1)
        v0 = np.array([x[c] for x in m0])
        v1 = np.array([x[c] for x in m])
        want = g * 3.0 * abs((c[0] - 1 - 1e13) * dz)
        got = _angle_about(v0, v1, (0.0, 0.0, 1.0))
        assert abs(_wrap(got - want)) < 2e-11, (c, got, want)
        assert v1[2] == pytest.approx(v0[2], abs=1e-15)


def test_pulse_flag3_general_axis_and_strength():
    """Flag 3 with an oblique, non-unit axis a: the rotation axis is a / |a|
    and the angle scales with |a| (effectiveGradient); the same pulse with
    2a turns by twice the angle (mod 2 pi)."""
    shape = (3, 4, 5)
    d = 2.0**-62
    p = oracle.params(shape, d, 1e-5, dy=2 * d, dz=3 * d)
    act = layout.pad_mask(np.ones(shape, np.uint8))
    rng = np.random.default_rng(5)
    m0 = [layout.pad_field(rng.standard_normal(shape)) for _ in range(3)]
    a = np.array([0.6, -0.48, 0.64]) * 1.25  # |a| = 1.25
    n0 = a / np.linalg.norm(a)
    angles = []
    for scale in (1.0, 2.0):
        m = [x.copy() for x in m0]
        oracle.add_impulse(p, act, 3, 0.0, tuple(scale * a), m)
        row = []
        for c in [(1, 1, 1), (3, 4, 5), (2, 3, 1), (1, 4, 2)]:
            v0 = np.array([x[c] for x in m0])
            v1 = np.array([x[c] for x in m])
            t = sum((c[2 - ax] - 1 - 1e13) * [d, 2 * d, 3 * d][ax] * n0[ax] for ax in range(3))
            want = oracle.G_REF * scale * 1.25 * abs(t)
            got = _angle_about(v0, v1, n0)
            assert abs(_wrap(got - want)) < 1e-9, (scale, c, got, want)
            assert v1 @ n0 == pytest.approx(v0 @ n0, abs=1e-14)
            row.append(got)
        angles.append(row)
    np.testing.assert_allclose(_wrap(np.array(angles[1]) - 2 * np.array(angles[0])), 0.0, atol=2e-9)


def test_pulse_flag4_radial_phase():
    """Flag 4 with a unit axis a: angle = -g sqrt(sum_d ((idx_d - 1/2) d_d a_d
    - n_d d_d a_d / 2)^2) about a (the distance from the domain centre, each
    axis weighted by a_d; cpp:2591-2595)."""
    shape = (4, 5, 6)
    dx, dy, dz = 1e-8, 1.5e-8, 2e-8
    p = oracle.params(shape, dx, 1e-5, dy=dy, dz=dz)
    rng = np.random.default_rng(12)
    mask = (rng.random(shape) < 0.8).astype(np.uint8)
    act = layout.pad_mask(mask)
    m0 = [layout.pad_field(rng.standard_normal(shape)) for _ in range(3)]
    nz, ny, nx = shape
    for a in [(0.0, 0.0, 1.0), (0.6, 0.0, 0.8), (0.36, 0.48, 0.8)]:
        m = [x.copy() for x in m0]
        oracle.add_impulse(p, act, 4, 0.0, a, m)
        for (k, j, i) in np.argwhere(mask):
            c = (k + 1, j + 1, i + 1)
            r = [(c[2] - 0.5) * dx * a[0] - 0.5 * nx * dx * a[0], (c[1] - 0.5) * dy * a[1] - 0.5 * ny * dy * a[1],
                 (c[0] - 0.5) * dz * a[2] - 0.5 * nz * dz * a[2]]
            want = -oracle.G_REF * np.sqrt(sum(x * x for x in r))
            v0 = np.array([x[c] for x in m0])
            v1 = np.array([x[c] for x in m])
            assert abs(_wrap(_angle_about(v0, v1, a) - want)) < 1e-12, (a, c)
    # the centre of an odd axis: the middle plane is not rotated at all
    p3 = oracle.params((5, 1, 1), 1e-6, 1e-5, dz=dz)
    m = [layout.pad_field(np.full((5, 1, 1), v)) for v in (1.0, 0.0, 0.0)]
    oracle.add_impulse(p3, layout.pad_mask(np.ones((5, 1, 1), np.uint8)), 4, 0.0, (0.0, 0.0, 1.0), m)
    assert m[0][3, 1, 1] == 1.0 and m[1][3, 1, 1] == 0.0
    assert m[1][1, 1, 1] == pytest.approx(np.sin(-oracle.G_REF * 2 * dz), abs=1e-14)


def test_pulse_flag4_non_unit_axis_is_not_normalised():
    """Flag 4 hands its axis to OrtTransform as given (cpp:2594): with a =
    (0, 0, 2) the in-plane part turns by 2 s about z and Mz becomes
    Mz (c + 4 (1 - c)), c, s of the angle -2 g dz |k - 1/2 - nz/2|."""
    shape = (4, 1, 1)
    dz = 1e-8
    p = oracle.params(shape, 1e-6, 1e-5, dz=dz)
    act = layout.pad_mask(np.ones(shape, np.uint8))
    m = [layout.pad_field(np.full(shape, v)) for v in (0.5, -0.2, 0.7)]
    oracle.add_impulse(p, act, 4, 0.0, (0.0, 0.0, 2.0), m)
    for k in range(1, 5):
        ang = -oracle.G_REF * 2 * dz * abs(k - 0.5 - 2.0)
        c, s = np.cos(ang), np.sin(ang)
        got = [x[k, 1, 1] for x in m]
        np.testing.assert_allclose(got, [0.5 * c - 2 * s * (-0.2), -0.2 * c + 2 * s * 0.5, 0.7 * (c + 4 * (1 - c))],
                                   rtol=1e-13, atol=1e-15)
