"""MacCormack scheme and periodic / zero-value boundaries on the GPU (the
per-voxel kernel) against the FP64 oracle, at the bars of the sweep's parity
tests (signal relative L2 <= 1e-5, max voxel <= 1e-4), and the per-voxel
kernel's upwind step against the z-sweep bit for bit."""
import numpy as np
import pytest

import oracle
from initializator_amd import _lib, generators
from initializator_amd._lib import NMR_LAYOUT_PADDED
from initializator_amd.solver import HipSolver

pytestmark = pytest.mark.gpu

SIG_TOL, VOX_TOL = 1e-5, 1e-4
MAK, UPW = _lib.NMR_SCHEME_MACCORMACK, _lib.NMR_SCHEME_UPWIND
PER, DIR = _lib.NMR_BC_PERIODIC, _lib.NMR_BC_DIRICHLET


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _flow_case(shape, nsteps, phi=0.35, seed=31, g=0.3):
    return generators.case_flow(seed=seed, shape=shape, phi=phi, g=g, nsteps=nsteps, big_delta=4e-3,
                                delta=1e-3)


def _gpu(case, m0, scheme, bc, variant=-1, **kw):
    s = HipSolver.from_case(case, scheme=scheme, bc=bc, **kw)
    s.set_tile(variant, 0)
    s.upload_magnetization(*m0[: case.ncomp])
    sig = s.run(case.nsteps, None if case.grad_sched is None else case.grad_sched[: case.nsteps])
    f = s.download_magnetization()
    padded = s.download_magnetization(NMR_LAYOUT_PADDED)
    s.close()
    return sig, f, padded


def _check(name, case, m0, scheme, bc):
    sig_ref, f_ref = oracle.run_case(case, m0=m0, scheme=scheme, bc=bc)
    sig, f, _ = _gpu(case, m0, scheme, bc)
    nc = case.ncomp
    e_sig = rel_l2(sig, sig_ref[:, :nc])
    e_vox = max(float(np.abs(a - b).max()) for a, b in zip(f, f_ref[:nc]))
    print(f"[schemes] {name}: signal relL2 = {e_sig:.3e}, max voxel = {e_vox:.3e}", flush=True)
    assert e_sig <= SIG_TOL and e_vox <= VOX_TOL, (e_sig, e_vox)
    return f


@pytest.mark.parametrize("bc", [(0, 0, 0), (PER, PER, PER), (DIR, 0, PER)])
def test_maccormack_flow_pgse_against_oracle(bc):
    """Flow + PGSE, complex (Mx, My), nx not a multiple of 4, solids holding
    nonzero values (they end each MacCormack step at 0, as the reference's
    corrector copy leaves them)."""
    case = _flow_case((20, 26, 23), 300)
    rng = np.random.default_rng(2)
    m0 = [np.where(case.mask != 0, 1.0, rng.random(case.shape)), np.zeros(case.shape), np.zeros(case.shape)]
    f = _check(f"MacCormack bc {bc}", case, m0, MAK, bc)
    assert np.all(f[0][case.mask == 0] == 0.0)


@pytest.mark.parametrize("bc", [(PER, PER, PER), (DIR, DIR, 0)])
def test_upwind_nonneumann_boundaries_against_oracle(bc):
    case = _flow_case((18, 21, 25), 300, seed=12)
    m0 = [case.mask.astype(np.float64), np.zeros(case.shape), np.zeros(case.shape)]
    _check(f"upwind bc {bc}", case, m0, UPW, bc)


def test_maccormack_three_components_tilted_field():
    shape = (11, 13, 17)
    case = _flow_case(shape, 120, phi=0.45, seed=5)
    sched = np.zeros((case.nsteps, 3))
    sched[:50] = (0.05, -0.1, 0.2)
    sched[70:100] = (-0.2, 0.1, 0.05)
    case.grad_sched = sched
    case.ncomp = 3
    case.t1_bulk, case.t2_bulk = 0.02, 0.01
    b0, peq = (0.3, -0.5, 0.8), 0.4
    rng = np.random.default_rng(9)
    m0 = [rng.random(shape) * case.mask for _ in range(3)]
    p, act, D, t1, t2, vel, m = oracle.setup_case(case, p_eq=peq, m0=m0, scheme=MAK, bc=(PER, 0, 0))
    for i in range(3):
        p.b0[i] = b0[i]
    sig_ref = oracle.run(p, act, D, t1, t2, vel, case.nsteps, m, grad_sched=sched)
    from initializator_amd import layout

    f_ref = [layout.unpad(x) for x in m]
    s = HipSolver.from_case(case, b0_dir=b0, p_eq=peq, scheme=MAK, bc=(PER, 0, 0))
    s.upload_magnetization(*m0)
    sig = s.run(case.nsteps, sched)
    f = s.download_magnetization()
    s.close()
    e_sig, e_vox = rel_l2(sig, sig_ref), max(float(np.abs(a - b).max()) for a, b in zip(f, f_ref))
    print(f"[schemes] MacCormack ncomp 3 tilted B0: signal relL2 = {e_sig:.3e}, max voxel = {e_vox:.3e}",
          flush=True)
    assert e_sig <= SIG_TOL and e_vox <= VOX_TOL


@pytest.mark.parametrize("nc", [1, 2, 3])
def test_per_voxel_upwind_equals_sweep_bitwise(nc):
    """With zero-derivative boundaries the per-voxel kernel (tile variant -2)
    and the z-sweep compute every voxel with the same operations."""
    case = _flow_case((15, 19, 22), 40, seed=44)
    if nc == 1:
        case.grad_sched = None
    case.ncomp = nc
    rng = np.random.default_rng(nc)
    m0 = [rng.random(case.shape) for _ in range(3)]
    kw = dict(b0_dir=(0.2, 0.3, 0.9), p_eq=0.1) if nc == 3 else {}
    out = []
    for variant in (-1, -2):
        s = HipSolver.from_case(case, **kw)
        s.set_tile(variant, 0)
        s.upload_magnetization(*m0[:nc])
        sched = None if case.grad_sched is None else np.tile([0.01, -0.02, 0.3] if nc == 3 else [0, 0, 0.3],
                                                              (case.nsteps, 1))
        sig = s.run(case.nsteps, sched)
        out.append((sig, s.download_magnetization()))
        s.close()
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=1e-12)


def test_padded_download_ghosts_follow_boundaries():
    case = _flow_case((6, 7, 9), 3, seed=3)
    m0 = [np.random.default_rng(1).random(case.shape) * case.mask, np.zeros(case.shape), np.zeros(case.shape)]
    _, f, pad = _gpu(case, m0, MAK, (PER, DIR, 0))
    mx = pad[0]
    np.testing.assert_array_equal(mx[1:-1, 1:-1, 0], f[0][:, :, -1])    # periodic x
    np.testing.assert_array_equal(mx[1:-1, 1:-1, -1], f[0][:, :, 0])
    assert np.all(mx[1:-1, 0, 1:-1] == 0) and np.all(mx[1:-1, -1, 1:-1] == 0)  # zero y
    np.testing.assert_array_equal(mx[0, 1:-1, 1:-1], f[0][0])            # zero derivative z


def _mixture(case, seed=4):
    """Smooth two-phase concentration, face-wise relaxivities, and the FP64
    per-voxel D, T1, T2 of FillInDiffusinAndTime for it."""
    from initializator_amd import layout

    rng = np.random.default_rng(seed)
    shape = case.shape
    p = oracle.params(shape, case.dx, case.dt, precession=case.grad_sched is not None, velocity=case.vx is not None)
    act = layout.pad_mask(case.mask)
    x = np.clip(generators._smooth_noise(shape, seed, 7, 2.0) * 3.0 + 0.5, 0.0, 1.0).astype(np.float64)
    conc = (layout.pad_field(x), layout.pad_field(1.0 - x))
    faces = [np.ascontiguousarray(rng.uniform(2e-6, 4e-5, oracle.padded_shape(p))) for _ in range(12)]
    fl = oracle.fluids(d=(2.3e-9, 4e-10), t1=(0.8, 0.3), t2=(0.4, 0.05), ncomp=2)
    D, t1, t2, _ = oracle.fill_coefficients_2f(p, act, fl, conc, faces)
    return p, act, D, t1, t2, conc, fl, x


@pytest.mark.parametrize("scheme", [UPW, MAK])
def test_per_voxel_coefficients_two_phase_against_oracle(scheme):
    """Two-phase mixture with face-wise relaxivities: the GPU takes the
    per-voxel D, 1/T1, 1/T2 and the phase weights; fields, the plain pore
    sums and the two-phase weighted sums follow the FP64 restatement."""
    from initializator_amd import layout

    case = _flow_case((14, 17, 19), 120, seed=21)
    p, act, D, t1, t2, conc, fl, x = _mixture(case)
    p.scheme = scheme
    vel = [np.ascontiguousarray(v) for v in layout.pad_velocity(case.vx, case.vy, case.vz, case.shape)]
    m0 = [case.mask.astype(np.float64), np.zeros(case.shape), np.zeros(case.shape)]
    m = [np.ascontiguousarray(layout.pad_field(a)) for a in m0]
    # oracle: step by step to get the two-phase sums of every step
    pc = [np.zeros(oracle.padded_shape(p)) for _ in range(6)]
    a = [q.copy() for q in m]
    ref_w = []
    for f in m:
        oracle.bc_field(p, f)

    def wrow():
        s = [oracle.output_2f(p, act, (0, 0, 0), fl, conc, m[q]) for q in range(2)]
        return [s[0][0], s[1][0], s[0][1], s[1][1]]

    ref_w.append(wrow())
    for n in range(case.nsteps):
        nf = oracle.fill_gradient(p, case.grad_sched[n])
        if scheme == MAK:
            oracle.maccormack(p, act, D, t1, t2, vel, nf, m, pc)
        else:
            oracle.explicit_simple(p, act, D, t1, t2, vel, nf, m, a)
        for f in m:
            oracle.bc_field(p, f)
        ref_w.append(wrow())
    ref_w = np.array(ref_w)
    f_ref = [layout.unpad(q) for q in m[:2]]
    # GPU
    c1, c2 = np.vectorize(lambda v: oracle.phase(fl, v, 1.0 - v))(x)
    gate = (c1 > 0.8) | (c2 > 0.8)
    s = HipSolver.from_case(case, scheme=scheme)
    s.upload_magnetization(*m0[:2])
    pore = case.mask != 0
    s.upload_coefficients(np.where(pore, layout.unpad(D), 0.0), np.where(pore, 1.0 / layout.unpad(t2), 0.0),
                          r1=np.where(pore, 1.0 / layout.unpad(t1), 0.0))
    s.upload_phase_weights(np.where(gate, c1, 0.0), np.where(gate, c2, 0.0))
    sig = s.run(case.nsteps, case.grad_sched[: case.nsteps])
    f = s.download_magnetization()
    s.close()
    e_sig = rel_l2(sig, ref_w)
    e_vox = max(float(np.abs(u - v).max()) for u, v in zip(f, f_ref))
    print(f"[schemes] two-phase per-voxel, scheme {scheme}: weighted signal relL2 = {e_sig:.3e}, "
          f"max voxel = {e_vox:.3e}", flush=True)
    assert e_sig <= SIG_TOL and e_vox <= VOX_TOL


def test_per_voxel_coefficients_matching_the_uniform_ones():
    """Uploading the coefficients the mask implies (uniform D; bulk plus
    rho/d per solid face) reproduces the uniform-coefficient sweep to FP32
    rounding."""
    case = _flow_case((12, 15, 22), 80, seed=8)
    m0 = [case.mask.astype(np.float64), np.zeros(case.shape)]
    s = HipSolver.from_case(case)
    s.upload_magnetization(*m0)
    sig_a = s.run(case.nsteps, case.grad_sched[: case.nsteps])
    fa = s.download_magnetization()
    s.close()
    pad = np.pad(case.mask, 1, mode="edge")
    nsolid = sum((pad[1:-1, 1:-1, 1:-1] != 0) & (np.roll(pad, sh, ax)[1:-1, 1:-1, 1:-1] == 0)
                 for ax in range(3) for sh in (1, -1))
    r2 = 1.0 / case.t2_bulk + case.rho * nsolid / case.dx
    r1 = 1.0 / case.t1_bulk + case.rho * nsolid / case.dx
    s = HipSolver.from_case(case)
    s.upload_magnetization(*m0)
    s.upload_coefficients(np.where(case.mask != 0, case.diffusivity, 0.0), r2, r1=r1)
    sig_b = s.run(case.nsteps, case.grad_sched[: case.nsteps])
    fb = s.download_magnetization()
    s.close()
    assert rel_l2(sig_b, sig_a) < 1e-6
    assert max(float(np.abs(u - v).max()) for u, v in zip(fa, fb)) < 1e-6
