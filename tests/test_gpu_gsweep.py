"""The general z-sweep (MacCormack stages, periodic / zero-value ghosts,
per-voxel coefficients) against the per-voxel kernel (tile variant -2), which
computes every voxel with the same operations: fields bit for bit, signal
rows to FP64 summation order.  Shapes cover partial x, y and z tiles, more
than one x tile (nx > 256), chunks that reach past nx and an nx that is a
multiple of 4 (the last chunk full, its ghost in the next LDS column)."""
import numpy as np
import pytest

from initializator_amd import _lib, generators
from initializator_amd.solver import HipSolver

pytestmark = pytest.mark.gpu

MAK, UPW = _lib.NMR_SCHEME_MACCORMACK, _lib.NMR_SCHEME_UPWIND
NEU, PER, DIR = _lib.NMR_BC_NEUMANN, _lib.NMR_BC_PERIODIC, _lib.NMR_BC_DIRICHLET

SHAPES = [(13, 17, 21), (9, 6, 261), (7, 5, 3), (6, 9, 12)]
BCS = [(NEU, NEU, NEU), (PER, PER, PER), (DIR, DIR, DIR), (PER, DIR, NEU), (DIR, PER, PER)]


def _case(shape, nsteps, seed, phi=0.4):
    return generators.case_flow(seed=seed, shape=shape, phi=phi, g=0.3, nsteps=nsteps, big_delta=4e-3, delta=1e-3)


def _pair(case, m0, scheme, bc, coef=None, weights=None, sched=None, **kw):
    """(signal, fields) of the general sweep and of the per-voxel kernel."""
    out = []
    for variant in (-1, -2):
        s = HipSolver.from_case(case, scheme=scheme, bc=bc, **kw)
        s.set_tile(variant, 0)
        s.upload_magnetization(*m0[: case.ncomp])
        if coef is not None:
            s.upload_coefficients(*coef[:2], r1=coef[2])
        if weights is not None:
            s.upload_phase_weights(*weights)
        sig = s.run(case.nsteps, sched if sched is not None else
                    (None if case.grad_sched is None else case.grad_sched[: case.nsteps]))
        out.append((sig, s.download_magnetization()))
        s.close()
    return out


def _same(name, out, mask=None, mak=False):
    (sa, fa), (sb, fb) = out
    for q, (a, b) in enumerate(zip(fa, fb)):
        if not np.array_equal(a, b):
            d = np.abs(a - b)
            i = np.unravel_index(int(np.argmax(d)), d.shape)
            raise AssertionError(f"{name}: component {q} differs at {i}: {a[i]!r} vs {b[i]!r}")
    np.testing.assert_allclose(sa, sb, rtol=1e-12, atol=1e-300)
    if mak and mask is not None:
        for a in fa:
            assert np.all(a[mask == 0] == 0.0)
    print(f"[gsweep] {name}: fields bit-identical, signal rows within 1e-12", flush=True)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_general_sweep_equals_per_voxel_kernel(scheme, bc, shape):
    case = _case(shape, 12, seed=5 + shape[2])
    rng = np.random.default_rng(shape[0])
    m0 = [np.where(case.mask != 0, 1.0, rng.random(shape)), 0.1 * rng.random(shape), np.zeros(shape)]
    _same(f"scheme {scheme} bc {bc} {shape}", _pair(case, m0, scheme, bc), case.mask, scheme == MAK)


@pytest.mark.parametrize("nc", [1, 3])
@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_general_sweep_components(nc, scheme):
    """One component (no precession), and three with a tilted main field,
    pEq and gradients on every axis."""
    case = _case((11, 14, 19), 10, seed=70 + nc)
    case.ncomp = nc
    rng = np.random.default_rng(nc)
    m0 = [rng.random(case.shape) for _ in range(3)]
    if nc == 1:
        case.grad_sched = None
        sched, kw = None, {}
    else:
        sched = np.tile([0.02, -0.03, 0.3], (case.nsteps, 1))
        kw = dict(b0_dir=(0.2, 0.3, 0.9), p_eq=0.1)
    out = _pair(case, m0, scheme, (PER, DIR, NEU), sched=sched, **kw)
    _same(f"ncomp {nc} scheme {scheme}", out, case.mask, scheme == MAK)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_general_sweep_per_voxel_coefficients(nc, scheme):
    """Per-voxel D (harmonic means), 1/T2 and 1/T1 (given with three
    components; omitted with two, where it is not used)."""
    case = _case((10, 19, 27), 10, seed=90 + nc)
    case.ncomp = nc
    rng = np.random.default_rng(7)
    shape = case.shape
    m0 = [rng.random(shape) for _ in range(3)]
    D = rng.uniform(0.5e-9, 3e-9, shape)
    r2 = rng.uniform(0.2, 40.0, shape)
    r1 = rng.uniform(0.1, 5.0, shape) if nc == 3 else None
    sched = np.tile([0.01, 0.02, 0.3] if nc == 3 else [0.0, 0.0, 0.3], (case.nsteps, 1))
    kw = dict(b0_dir=(0.1, 0.0, 1.0)) if nc == 3 else {}
    out = _pair(case, m0, scheme, (DIR, PER, PER), coef=(D, r2, r1), sched=sched, **kw)
    _same(f"per-voxel coefficients ncomp {nc} scheme {scheme}", out, case.mask, scheme == MAK)


def test_general_sweep_phase_weighted_sums():
    """Two-phase weights: the signal comes from the field sum after each
    step of the general sweep."""
    case = _case((12, 13, 17), 8, seed=12)
    rng = np.random.default_rng(3)
    shape = case.shape
    m0 = [case.mask.astype(np.float64), np.zeros(shape), np.zeros(shape)]
    w1 = rng.random(shape)
    out = _pair(case, m0, UPW, (PER, NEU, DIR), weights=(w1, 1.0 - w1))
    assert out[0][0].shape == (case.nsteps + 1, 4)
    _same("phase-weighted sums", out)


def test_general_sweep_velocity_on_solid_faces():
    """Random face velocities everywhere (solids are read by upwind
    advection, so all-solid chunks must be loaded)."""
    case = _case((9, 14, 23), 10, seed=17)
    rng = np.random.default_rng(9)
    nz, ny, nx = case.shape
    case.vx = rng.uniform(-1e-3, 1e-3, (nz, ny, nx + 1))
    case.vy = rng.uniform(-1e-3, 1e-3, (nz, ny + 1, nx))
    case.vz = rng.uniform(-1e-3, 1e-3, (nz + 1, ny, nx))
    m0 = [rng.random(case.shape), rng.random(case.shape), None]
    _same("velocity on solid faces", _pair(case, m0, UPW, (PER, PER, DIR)))


@pytest.mark.parametrize("bc", [(NEU, NEU, NEU), (PER, DIR, NEU)])
def test_masked_face_advection(bc):
    """NMR_F_ADVECT_MASK (the face masks of NMR_2d): random velocity on
    every face, so pores border flowing solid faces.  The general sweep and
    the per-voxel kernel agree bit for bit, and both follow the oracle's
    masked step; the unmasked step differs."""
    import oracle

    case = _case((12, 15, 21), 60, seed=23)
    rng = np.random.default_rng(12)
    nz, ny, nx = case.shape
    case.vx = rng.uniform(-1e-3, 1e-3, (nz, ny, nx + 1))
    case.vy = rng.uniform(-1e-3, 1e-3, (nz, ny + 1, nx))
    case.vz = rng.uniform(-1e-3, 1e-3, (nz + 1, ny, nx))
    m0 = [np.where(case.mask != 0, 1.0, 3.0), np.zeros(case.shape), np.zeros(case.shape)]
    out = _pair(case, m0, UPW, bc, advect_mask=True)
    _same(f"masked-face advection bc {bc}", out)
    sig_ref, f_ref = oracle.run_case(case, m0=m0, bc=bc, advect_mask=True)
    sig, f = out[0]
    e_sig = float(np.linalg.norm(sig - sig_ref[:, :2]) / np.linalg.norm(sig_ref[:, :2]))
    e_vox = max(float(np.abs(a - b).max()) for a, b in zip(f, f_ref[:2]))
    print(f"[gsweep] masked-face advection vs oracle: signal relL2 = {e_sig:.3e}, max voxel = {e_vox:.3e}",
          flush=True)
    assert e_sig <= 1e-5 and e_vox <= 1e-4
    plain = _pair(case, m0, UPW, bc)[0]
    assert np.abs(plain[0] - sig).max() > 1e-6


@pytest.mark.parametrize("nx", [12, 13, 64])
@pytest.mark.parametrize("nsteps", [1, 20])
def test_maccormack_keeps_uniform_equilibrium(nx, nsteps):
    """Mz = p_eq at every pore, Mx = My = 0, with flow, diffusion and both
    relaxations: every difference, rate and rotation term is zero, so the
    predictor and the corrector return the field unchanged (to the FP32
    value of p_eq) -- in the row layout (1 step) and the brick layout (20
    steps), whatever nx is modulo 4."""
    shape = (7, 9, nx)
    case = _case(shape, nsteps, seed=nx)
    peq = 0.55
    s = HipSolver(shape, 3, case.dx, case.dt, case.diffusivity, 1.0 / case.t2_bulk, r1_bulk=20.0, rho1=1e-5,
                  rho2=2e-5, p_eq=peq, velocity=True, precession=True, scheme=MAK)
    s.upload_mask(case.mask)
    s.upload_velocity(case.vx, case.vy, case.vz)
    pore = case.mask != 0
    s.upload_magnetization(np.zeros(shape), np.zeros(shape), peq * pore)
    s.run(nsteps)
    f = s.download_magnetization()
    s.close()
    assert np.all(f[0] == 0.0) and np.all(f[1] == 0.0)
    np.testing.assert_array_equal(f[2][pore], np.float32(peq))
    assert np.all(f[2][~pore] == 0.0)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_general_sweep_per_voxel_field(nc, scheme):
    """A per-voxel field along the main-field axis (nmr_hip_upload_field, the
    reference's noise field) in place of that axis's gradient: the sweep and
    the per-voxel kernel agree bit for bit, and the field changes the result."""
    shape = (9, 11, 16)
    case = _case(shape, 8, seed=120 + nc)
    case.ncomp = nc
    rng = np.random.default_rng(nc)
    m0 = [rng.random(shape) for _ in range(3)]
    bfield = rng.uniform(-3e-5, 3e-5, shape)
    sched = case.grad_sched[: case.nsteps]
    res = []
    for variant, field in [(-1, True), (-2, True), (-1, False)]:
        s = HipSolver.from_case(case, scheme=scheme, bc=(NEU, NEU, NEU))
        s.set_tile(variant, 0)
        s.upload_magnetization(*m0[:nc])
        if field:
            s.upload_field(bfield)
        sig = s.run(case.nsteps, sched)
        res.append((sig, s.download_magnetization()))
        s.close()
    _same(f"per-voxel field ncomp {nc} scheme {scheme}", res[:2], case.mask, scheme == MAK)
    assert not np.array_equal(res[0][1][0], res[2][1][0])
