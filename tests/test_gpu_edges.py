"""Edge cases of the GPU solve against the FP64 oracle: degenerate grids
(one voxel, one row, one column, one plane), an all-solid and an all-pore
mask, and a run of zero steps -- through the z-sweep (upwind, zero-derivative
ghosts) and the general sweep (MacCormack, periodic and zero-value ghosts)."""
import numpy as np
import pytest

import oracle
from initializator_amd import _lib, generators
from initializator_amd.solver import HipSolver

pytestmark = pytest.mark.gpu

MAK, UPW = _lib.NMR_SCHEME_MACCORMACK, _lib.NMR_SCHEME_UPWIND
PER, DIR = _lib.NMR_BC_PERIODIC, _lib.NMR_BC_DIRICHLET
SHAPES = [(1, 1, 1), (1, 1, 7), (1, 5, 1), (6, 1, 1), (2, 3, 5), (3, 1, 9)]
MODES = [(UPW, (0, 0, 0)), (MAK, (PER, PER, PER)), (UPW, (DIR, PER, DIR))]


def _case(shape, mask, seed, nsteps=30):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    dx, D = 1e-6, 2.3e-9
    vx = rng.uniform(-1e-3, 1e-3, (nz, ny, nx + 1))
    vy = rng.uniform(-1e-3, 1e-3, (nz, ny + 1, nx))
    vz = rng.uniform(-1e-3, 1e-3, (nz + 1, ny, nx))
    sched = generators.pgse_schedule(nsteps, generators.stable_dt(dx, D), 0.3, 5 * 2.9e-5, 12 * 2.9e-5)
    return generators.Case(shape, mask.astype(np.uint8), vx, vy, vz, dx, generators.stable_dt(dx, D), D, 0.5,
                           2.5e10, 20e-6, nsteps, sched, 2)


def _compare(name, case, m0, scheme, bc, nsteps=None):
    n = case.nsteps if nsteps is None else nsteps
    sig_ref, f_ref = oracle.run_case(case, nsteps=n, m0=m0, scheme=scheme, bc=bc)
    with HipSolver.from_case(case, scheme=scheme, bc=bc) as s:
        s.upload_magnetization(*m0[:2])
        sig = s.run(n, case.grad_sched[:n])
        f = s.download_magnetization()
    scale = max(1e-30, float(np.abs(sig_ref[:, :2]).max()))
    e_sig = float(np.abs(sig - sig_ref[:, :2]).max()) / scale
    e_vox = max(float(np.abs(a - b).max()) for a, b in zip(f, f_ref[:2]))
    print(f"[edges] {name}: signal max err / scale {e_sig:.2e}, max voxel {e_vox:.2e}", flush=True)
    assert e_sig <= 1e-5 and e_vox <= 1e-4, (e_sig, e_vox)
    return sig, f


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", SHAPES)
def test_degenerate_grids(shape, mode):
    rng = np.random.default_rng(sum(shape))
    mask = rng.random(shape) < 0.7
    mask.flat[0] = True
    case = _case(shape, mask, seed=len(shape) + shape[2])
    m0 = [rng.random(shape), rng.random(shape) - 0.5, np.zeros(shape)]
    _compare(f"{shape} scheme {mode[0]} bc {mode[1]}", case, m0, *mode)


@pytest.mark.parametrize("mode", MODES)
def test_all_solid_mask(mode):
    """No pores: the signal is 0 every step; upwind leaves every voxel as it
    was, MacCormack leaves 0 everywhere (its corrector's solid value)."""
    shape = (5, 6, 9)
    case = _case(shape, np.zeros(shape, dtype=bool), seed=3)
    rng = np.random.default_rng(3)
    m0 = [rng.random(shape), rng.random(shape), np.zeros(shape)]
    with HipSolver.from_case(case, scheme=mode[0], bc=mode[1]) as s:
        s.upload_magnetization(*m0[:2])
        sig = s.run(case.nsteps, case.grad_sched)
        f = s.download_magnetization()
    assert np.all(sig == 0.0)
    for q in range(2):
        want = np.zeros(shape) if mode[0] == MAK else np.asarray(m0[q], dtype=np.float32)
        np.testing.assert_array_equal(f[q], want)


@pytest.mark.parametrize("mode", MODES)
def test_all_pore_mask(mode):
    shape = (7, 9, 13)
    case = _case(shape, np.ones(shape, dtype=bool), seed=4)
    rng = np.random.default_rng(4)
    m0 = [rng.random(shape), rng.random(shape) - 0.5, np.zeros(shape)]
    _compare(f"all pores scheme {mode[0]} bc {mode[1]}", case, m0, *mode)


@pytest.mark.parametrize("mode", MODES)
def test_zero_steps(mode):
    """A run of 0 steps returns the signal of the uploaded state and changes
    nothing."""
    shape = (6, 7, 10)
    rng = np.random.default_rng(5)
    mask = rng.random(shape) < 0.5
    case = _case(shape, mask, seed=5)
    m0 = [rng.random(shape), rng.random(shape), np.zeros(shape)]
    with HipSolver.from_case(case, scheme=mode[0], bc=mode[1]) as s:
        s.upload_magnetization(*m0[:2])
        sig = s.run(0)
        f = s.download_magnetization()
    assert sig.shape == (1, 2)
    for q in range(2):
        np.testing.assert_array_equal(f[q], np.asarray(m0[q], dtype=np.float32))
        np.testing.assert_allclose(sig[0, q], np.sum(np.asarray(m0[q], dtype=np.float32) * mask, dtype=np.float64),
                                   rtol=1e-12)
