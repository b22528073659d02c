"""nmr_hip_run_static -- the run entry point that allocates nothing (its
partial sums and signal rows live in the library's fixed device scratch) --
against nmr_hip_run: the same fields bit for bit and the same signal rows,
on two calls in a row, the second on a smaller grid; long runs are split
into sub-runs that fit the scratch.  Named to be collected last."""
import ctypes

import numpy as np
import pytest

from initializator_amd import _lib, generators
from initializator_amd.solver import HipSolver, check

pytestmark = pytest.mark.gpu

MAK = 1


def _dptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _pair(case, m0, scheme):
    out = []
    for _ in range(2):
        s = HipSolver.from_case(case, scheme=scheme)
        s.upload_magnetization(*m0)
        out.append(s)
    return out


def _compare(shape, nsteps, scheme, seed):
    case = generators.case_flow(seed=seed, shape=shape, phi=0.3, nsteps=nsteps, big_delta=6e-4, delta=2e-4)
    rng = np.random.default_rng(seed)
    m0 = [np.where(case.mask != 0, 1.0, rng.random(shape)), rng.random(shape) - 0.5]
    lib = _lib.load()
    a, b = _pair(case, m0, scheme)
    try:
        sched = np.ascontiguousarray(case.grad_sched)
        want = a.run(nsteps, sched)
        before = b.device_bytes()
        got = np.full_like(want, np.nan)
        check(lib.nmr_hip_run_static(b.ctx, nsteps, _dptr(sched), _dptr(got)), "nmr_hip_run_static")
        assert b.device_bytes() == before  # nothing allocated
        fa, fb = a.download_magnetization(), b.download_magnetization()
        for x, y in zip(fa, fb):
            assert np.array_equal(x, y)
        assert np.array_equal(got, want)
        # a second call continues from the state the first left
        want2 = a.run(3, sched[:3])
        got2 = np.empty_like(want2)
        check(lib.nmr_hip_run_static(b.ctx, 3, _dptr(sched[:3]), _dptr(got2)), "nmr_hip_run_static")
        assert np.array_equal(got2, want2)
        for x, y in zip(a.download_magnetization(), b.download_magnetization()):
            assert np.array_equal(x, y)
    finally:
        a.close()
        b.close()
    print(f"[no-alloc] {shape} scheme {scheme} {nsteps} steps: fields and signal identical to nmr_hip_run",
          flush=True)


@pytest.mark.parametrize("scheme", [0, MAK])
def test_static_run_equals_run_large_then_small(scheme):
    # 768 workgroups per step: 600 steps take two sub-runs of the scratch
    _compare((64, 256, 256), 600, scheme, 41 + scheme)
    _compare((12, 14, 17), 9, scheme, 43 + scheme)

