"""Pins the oracle's masked-face advection (the NMR_2d variant,
NmrRelaxationSolveKernels.cu:90-232: advection across a face only when the
neighbour is active, cu:190-210) by hand values and by its agreement with the
plain upwind step wherever no pore-solid face carries velocity."""
import numpy as np
import pytest

import oracle
from initializator_amd import generators, layout


def _line_step(mask_row, faces, m_row, advect_mask, dt=1e-4, dx=1e-6):
    nx = len(mask_row)
    shape = (1, 1, nx)
    p = oracle.params(shape, dx, dt, velocity=True, diffusion=False, advect_mask=advect_mask)
    act = layout.pad_mask(np.array(mask_row, dtype=np.uint8).reshape(shape))
    D, t1, t2, _ = oracle.fill_coefficients(p, act, 2.3e-9, 2.5e10, 2.5e10, 0.0, 0.0)
    vx = np.zeros(oracle.padded_shape(p))
    vx[1, 1, 1 : nx + 2] = faces
    vel = [vx, np.zeros_like(vx), np.zeros_like(vx)]
    m = [layout.pad_field(np.array(m_row, dtype=np.float64).reshape(shape)), np.zeros_like(vx), np.zeros_like(vx)]
    for f in m:
        oracle.bc_field(p, f)
    a = [x.copy() for x in m]
    oracle.explicit_simple(p, act, D, t1, t2, vel, oracle.fill_gradient(p, (0.0, 0.0, 0.0)), m, a)
    return m[0][1, 1, 1:-1]


@pytest.mark.parametrize("u", [2e-3, -2e-3])
def test_pore_next_to_a_flowing_solid_face(u):
    """Cells 0..5, cell 2 solid, the same velocity on every face.  The plain
    step advects the solid's value into its downstream pore neighbour; the
    masked step leaves that pore unchanged and treats every other pore as
    the plain step does."""
    mask = [1, 1, 0, 1, 1, 1]
    m0 = np.array([0.3, 0.7, 5.0, 0.2, 0.9, 0.4])
    faces = np.full(7, u)
    c = u * 1e-4 / 1e-6
    plain = _line_step(mask, faces, m0, False)
    masked = _line_step(mask, faces, m0, True)
    ext = np.concatenate([[m0[0]], m0, [m0[-1]]])  # zero-derivative ghosts
    want = m0.copy()
    for i in range(6):
        if mask[i] == 0:
            continue
        if u > 0:
            up_i = i - 1
            want[i] = m0[i] - c * (ext[i + 1] - ext[i])
        else:
            up_i = i + 1
            want[i] = m0[i] - c * (ext[i + 2] - ext[i + 1])
        np.testing.assert_allclose(plain[i], want[i], rtol=1e-14)
        if 0 <= up_i < 6 and mask[up_i] == 0:
            # the solid's value does not enter (bulk T2 = 2.5e10 s changes it by 4e-15)
            np.testing.assert_allclose(masked[i], m0[i], rtol=1e-14)
            assert abs(plain[i] - m0[i]) > 1e-3
        else:
            np.testing.assert_allclose(masked[i], want[i], rtol=1e-14)
    assert masked[2] == 5.0 and plain[2] == 5.0           # solids are not updated


def test_masked_equals_plain_without_flow_at_solid_faces():
    """Velocity confined to the pores (no pore-solid face carries flow):
    masked and plain steps agree to FP64 rounding over several steps."""
    case = generators.case_flow(seed=5, shape=(9, 10, 12), phi=0.4, nsteps=6, g=0.2)
    rng = np.random.default_rng(4)
    m0 = [rng.random(case.shape), rng.random(case.shape), np.zeros(case.shape)]
    sa, fa = oracle.run_case(case, m0=m0)
    sb, fb = oracle.run_case(case, m0=m0, advect_mask=True)
    np.testing.assert_allclose(sb, sa, rtol=1e-13)
    for a, b in zip(fa, fb):
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-15)


def test_masked_step_ignores_solid_values():
    """Random velocity on every face: the masked step's pores do not depend
    on the values held by solid cells (the plain step's do)."""
    case = generators.case_flow(seed=6, shape=(8, 9, 11), phi=0.5, nsteps=1, g=0.0)
    rng = np.random.default_rng(8)
    nz, ny, nx = case.shape
    case.vx = rng.uniform(-1e-3, 1e-3, (nz, ny, nx + 1))
    case.vy = rng.uniform(-1e-3, 1e-3, (nz, ny + 1, nx))
    case.vz = rng.uniform(-1e-3, 1e-3, (nz + 1, ny, nx))
    pore = case.mask != 0
    base = rng.random(case.shape)
    other = np.where(pore, base, base + 3.0)
    res = {}
    for am in (False, True):
        for name, f in (("a", base), ("b", other)):
            _, fl = oracle.run_case(case, m0=[f, np.zeros(case.shape), np.zeros(case.shape)], advect_mask=am)
            res[(am, name)] = fl[0][pore]
    np.testing.assert_allclose(res[(True, "a")], res[(True, "b")], rtol=0, atol=1e-30)
    assert np.abs(res[(False, "a")] - res[(False, "b")]).max() > 1e-6
