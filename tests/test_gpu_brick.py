"""The one-step sweep on bricks of 8 x-voxels by 4 rows (nmr_hip_set_layout, the default for
runs of 16 steps or more) against the row layout: the same operations on
the same values, so the fields must agree bit for bit and the per-step
signal to FP64 summation order; shapes cover x, y and z extents that are
not multiples of the brick (4, 4, 2), more than one x tile, every tile
variant with a brick sweep, one to three components and the exact phase."""
import numpy as np
import pytest

from initializator_amd import generators
from initializator_amd.solver import HipSolver

pytestmark = pytest.mark.gpu


def _run(case, m0, brick, variant=-1, nsteps=None, coef=None, field=None, weights=None, **kw):
    n = case.nsteps if nsteps is None else nsteps
    with HipSolver.from_case(case, **kw) as s:
        s.set_layout(brick)
        s.set_tile(variant, 0)
        s.upload_magnetization(*m0[: case.ncomp])
        if coef is not None:
            s.upload_coefficients(coef[0], coef[1], r1=coef[2])
        if field is not None:
            s.upload_field(field)
        if weights is not None:
            s.upload_phase_weights(*weights)
        sched = None if case.grad_sched is None else case.grad_sched[:n]
        sig = s.run(n, sched)
        f = s.download_magnetization()
        sig2 = s.run(n, sched)  # a second run continues from the converted-back state
        f2 = s.download_magnetization()
    return sig, f, sig2, f2


def _same(name, a, b):
    for x, y in zip(a[1] + a[3], b[1] + b[3]):
        assert np.array_equal(x, y), name
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(a[2], b[2], rtol=1e-12, atol=1e-300)
    print(f"[brick] {name}: fields bit-identical, signal within 1e-12", flush=True)


@pytest.mark.parametrize("shape,variant", [((23, 37, 45), -1), ((9, 6, 261), 0), ((17, 14, 90), 2),
                                           ((13, 21, 50), 4), ((11, 33, 30), 6), ((5, 3, 7), -1),
                                           ((9, 19, 261), 1), ((12, 37, 90), 3), ((13, 70, 48), 5),
                                           ((7, 9, 12), -1), ((6, 10, 20), 4), ((5, 13, 36), 2), ((8, 8, 64), 0)])
def test_brick_sweep_equals_row_sweep(shape, variant):
    case = generators.case_flow(seed=sum(shape), shape=shape, phi=0.35, nsteps=40, big_delta=6e-4, delta=2e-4)
    rng = np.random.default_rng(1)
    m0 = [np.where(case.mask != 0, 1.0, rng.random(shape)), rng.random(shape) - 0.5, np.zeros(shape)]
    _same(f"{shape} variant {variant}", _run(case, m0, True, variant), _run(case, m0, False, variant))


@pytest.mark.parametrize("nc", [1, 3])
def test_brick_sweep_components(nc):
    shape = (14, 19, 27)
    case = generators.case_flow(seed=5, shape=shape, phi=0.4, nsteps=24)
    case.ncomp = nc
    rng = np.random.default_rng(nc)
    m0 = [rng.random(shape) for _ in range(3)]
    kw = {}
    if nc == 1:
        case.grad_sched = None
    else:
        case.grad_sched = np.tile([0.02, -0.01, 0.3], (case.nsteps, 1))
        kw = dict(b0_dir=(0.2, 0.1, 0.9), p_eq=0.2)
    _same(f"ncomp {nc}", _run(case, m0, True, **kw), _run(case, m0, False, **kw))


def test_brick_sweep_exact_phase_and_velocity_on_solid_faces():
    shape = (12, 15, 22)
    case = generators.case_flow(seed=8, shape=shape, phi=0.4, nsteps=30)
    rng = np.random.default_rng(4)
    nz, ny, nx = shape
    case.vx = rng.uniform(-1e-3, 1e-3, (nz, ny, nx + 1))
    case.vy = rng.uniform(-1e-3, 1e-3, (nz, ny + 1, nx))
    case.vz = rng.uniform(-1e-3, 1e-3, (nz + 1, ny, nx))
    m0 = [rng.random(shape), rng.random(shape), None]
    _same("exact phase, flowing solid faces", _run(case, m0, True, phase_exact=True),
          _run(case, m0, False, phase_exact=True))


MAK = 1
PER, DIR = 1, 2


@pytest.mark.parametrize("shape", [(17, 14, 90), (9, 6, 261), (5, 7, 3), (6, 9, 64)])
@pytest.mark.parametrize("scheme,bc", [(MAK, (0, 0, 0)), (MAK, (PER, PER, PER)), (MAK, (DIR, PER, 0)),
                                       (0, (PER, DIR, PER))])
def test_brick_general_sweep_equals_row_and_per_voxel(shape, scheme, bc):
    """The general sweep (MacCormack, periodic / zero-value ghosts) on bricks
    against the row layout and against the per-voxel kernel, from a state
    whose solids hold nonzero values (the first MacCormack step reads them
    through periodic ghosts, later steps hold 0 there); two runs in a row."""
    case = generators.case_flow(seed=sum(shape) + 3, shape=shape, phi=0.35, nsteps=24, big_delta=6e-4, delta=2e-4)
    rng = np.random.default_rng(2)
    m0 = [np.where(case.mask != 0, 1.0, rng.random(shape)), rng.random(shape) - 0.5, np.zeros(shape)]
    kw = dict(scheme=scheme, bc=bc)
    brick = _run(case, m0, True, **kw)
    row = _run(case, m0, False, **kw)
    _same(f"general {shape} scheme {scheme} bc {bc}: brick vs row", brick, row)
    _same(f"general {shape} scheme {scheme} bc {bc}: brick vs per-voxel", brick, _run(case, m0, False, -2, **kw))


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("scheme", [MAK, 0])
def test_brick_general_sweep_per_voxel_coefficients_and_field(nc, scheme):
    """Per-voxel D, 1/T1, 1/T2 and a per-voxel field (row-layout arrays) with
    the state on bricks: the same bits as the row layout and the per-voxel
    kernel."""
    shape = (11, 13, 36)
    case = generators.case_flow(seed=77 + nc, shape=shape, phi=0.35, nsteps=20, big_delta=6e-4, delta=2e-4)
    case.ncomp = nc
    rng = np.random.default_rng(nc)
    m0 = [rng.random(shape) for _ in range(3)]
    coef = (rng.uniform(0.5e-9, 3e-9, shape), rng.uniform(0.2, 40.0, shape),
            rng.uniform(0.1, 5.0, shape) if nc == 3 else None)
    field = rng.uniform(-2e-5, 2e-5, shape)
    kw = dict(scheme=scheme, coef=coef, field=field)
    brick = _run(case, m0, True, **kw)
    _same(f"coefficients + field ncomp {nc} scheme {scheme}: brick vs row", brick, _run(case, m0, False, **kw))
    _same(f"coefficients + field ncomp {nc} scheme {scheme}: brick vs per-voxel", brick,
          _run(case, m0, False, -2, **kw))


@pytest.mark.parametrize("nc,scheme,buffer", [(1, 0, (0, 0, 0)), (2, MAK, (2, 1, 3)), (3, 0, (1, 2, 0)),
                                              (2, 0, (0, 0, 0))])
def test_brick_general_sweep_fused_phase_weighted_sums(nc, scheme, buffer):
    """Two-phase weighted sums: on bricks they are fused into the sweep
    (2*ncomp sums per workgroup, weights in rows), in the row layout and the
    per-voxel kernel a field sum follows each step; the fields are the same
    bits and the 2*ncomp signal columns agree to FP64 summation order, also
    with a trimmed signal window."""
    shape = (13, 11, 40)
    case = generators.case_flow(seed=31 + nc, shape=shape, phi=0.4, nsteps=24, big_delta=6e-4, delta=2e-4)
    case.ncomp = nc
    rng = np.random.default_rng(10 + nc)
    m0 = [rng.random(shape) for _ in range(3)]
    w1 = rng.random(shape)
    kw = dict(scheme=scheme, weights=(w1, 1.0 - w1), buffer=buffer,
              coef=(rng.uniform(0.5e-9, 3e-9, shape), rng.uniform(0.2, 40.0, shape), None) if nc == 2 else None)
    if nc == 1:  # one component: no gradient precession
        case.grad_sched = None
    brick = _run(case, m0, True, **kw)
    assert brick[0].shape == (case.nsteps + 1, 2 * nc)
    # the weights split the plain sum: w1 + w2 = 1 at every voxel, up to
    # their rounding to FP32
    plain = _run(case, m0, True, **dict(kw, weights=None))
    np.testing.assert_allclose(brick[0][:, :nc] + brick[0][:, nc:], plain[0], rtol=0,
                               atol=1e-6 * np.abs(plain[0]).max())
    _same(f"weighted sums ncomp {nc} scheme {scheme} buffer {buffer}: brick vs row", brick,
          _run(case, m0, False, **kw))
    _same(f"weighted sums ncomp {nc} scheme {scheme} buffer {buffer}: brick vs per-voxel", brick,
          _run(case, m0, False, -2, **kw))


@pytest.mark.parametrize("scheme", [MAK, 0])
def test_general_sweep_long_z_chunks(scheme):
    """The general sweep at the long z-chunks the automatic choice takes on
    large grids (22, 32, 57 planes, and one chunk longer than the grid) with
    per-voxel coefficients, a field and two-phase sums: the same bits as
    8-plane chunks, in bricks and in rows."""
    shape = (100, 18, 21)
    case = generators.case_flow(seed=5 + scheme, shape=shape, phi=0.35, nsteps=20, big_delta=6e-4, delta=2e-4)
    rng = np.random.default_rng(7)
    m0 = [rng.random(shape) for _ in range(3)]
    w1 = rng.random(shape)
    kw = dict(scheme=scheme, coef=(rng.uniform(0.5e-9, 3e-9, shape), rng.uniform(0.2, 40.0, shape), None),
              field=rng.uniform(-2e-5, 2e-5, shape), weights=(w1, 1.0 - w1))

    def run(brick, lz):
        n = case.nsteps
        with HipSolver.from_case(case, scheme=scheme) as s:
            s.set_layout(brick)
            s.set_tile(-1, lz)
            s.upload_magnetization(*m0[: case.ncomp])
            s.upload_coefficients(kw["coef"][0], kw["coef"][1])
            s.upload_field(kw["field"])
            s.upload_phase_weights(*kw["weights"])
            sig = s.run(n, case.grad_sched[:n])
            return sig, s.download_magnetization()

    for brick in (True, False):
        sig8, f8 = run(brick, 8)
        for lz in (22, 32, 57, 120):
            sig, f = run(brick, lz)
            for a, b in zip(f, f8):
                assert np.array_equal(a, b), (brick, lz)
            np.testing.assert_allclose(sig, sig8, rtol=1e-12, atol=1e-300)
    print(f"[brick] general sweep scheme {scheme}: z-chunks 22/32/57/120 bit-identical to 8", flush=True)


@pytest.mark.parametrize("scheme", [0, MAK])
def test_resident_state_short_runs_pulses_and_signal(scheme):
    """The state stays on the device in the sweep's layout from run to run:
    runs of 3, 7, 1 and 5 steps with RF, gradient and radial pulses between
    them, signal queries and a download in the middle.  On bricks and on rows
    the fields are the same bits after every stage and the per-step signals
    agree to FP64 summation order; row 0 of a run is the signal of the state
    it starts from (the previous run's last row, or the sum after a pulse)."""
    shape = (13, 22, 37)
    case = generators.case_flow(seed=91 + scheme, shape=shape, phi=0.35, nsteps=16, big_delta=6e-4, delta=2e-4)
    rng = np.random.default_rng(3)
    m0 = [np.where(case.mask != 0, 1.0, rng.random(shape)), rng.random(shape) - 0.5]
    g = case.grad_sched

    def stages(brick):
        out = []
        with HipSolver.from_case(case, scheme=scheme) as s:
            s.set_layout(brick)
            s.upload_magnetization(*m0)
            a = s.run(3, g[0:3])
            assert np.array_equal(s.signal(), a[-1])  # kept, not recomputed
            s.pulse(2, np.pi / 3, (0.0, 0.0, 1.0))
            p = s.signal()
            b = s.run(7, g[3:10])
            assert np.array_equal(b[0], p)
            out.append(s.download_magnetization())
            s.pulse(1, 2e-4, (0.0, 0.0, 1.0))
            c = s.run(1, g[10:11])
            s.pulse(4, 0.0, (0.0, 0.0, 1.0))
            d = s.run(5, g[11:16])
            e = s.run(0)
            assert np.array_equal(e[0], d[-1])
            out.append(s.download_magnetization())
            sigs = [a, b, c, d]
        return out, sigs

    fb, sb = stages(True)
    fr, sr = stages(False)
    for stage_b, stage_r in zip(fb, fr):
        for x, y in zip(stage_b, stage_r):
            assert np.array_equal(x, y)
    for x, y in zip(sb, sr):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-300)
    print(f"[brick] resident state, scheme {scheme}: short runs, pulses, signals: bricks == rows", flush=True)
