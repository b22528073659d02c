"""Pins the oracle's two-phase coefficients and outputs (or_phase,
or_fill_coefficients_2f, or_output_2f; NmrRelaxationSolve.cpp:1756-1893,
3324-3350) with closed forms and hand-computed cases."""
import numpy as np
import pytest

import oracle
from initializator_amd import layout

DW, DO = 2.3e-9, 2.1e-10


def test_phase_fractions_on_and_off_the_tie_line():
    fl = oracle.fluids(an=(1.0, 0.0), bn=(0.0, 1.0))
    # on the segment between the two points the fraction is the coordinate
    for x in (0.0, 0.13, 0.5, 0.7, 1.0):
        c1, c2 = oracle.phase(fl, x, 1.0 - x)
        assert c1 == pytest.approx(x, abs=1e-15) and c2 == pytest.approx(1 - x, abs=1e-15)
    # equidistant from both points: half and half
    assert oracle.phase(fl, 0.5, 0.5) == pytest.approx((0.5, 0.5))
    assert oracle.phase(fl, 3.0, 3.0) == pytest.approx((0.5, 0.5))
    # NCOMP=1 passes the concentrations through
    assert oracle.phase(oracle.fluids(ncomp=1), 0.25, 0.6) == (0.25, 0.6)


def _box(shape, conc0):
    mask = np.ones(shape, np.uint8)
    p = oracle.params(shape, 1e-6, 1e-5)
    act = layout.pad_mask(mask)
    c0 = layout.pad_field(conc0)
    return p, act, c0, layout.pad_field(1.0 - conc0)


def test_mixing_penalty_special_values():
    """D = (c1 Dw + c2 Do)(1 - 2 c1 c2 / (c1^2 + c2^2))^10: pure phases give
    Dw and Do, an even mixture gives exactly 0."""
    shape = (1, 1, 3)
    conc = np.array([[[1.0, 0.0, 0.5]]])
    p, act, c0, c1 = _box(shape, conc)
    fl = oracle.fluids(d=(DW, DO))
    D, _, _, _ = oracle.fill_coefficients_2f(p, act, fl, (c0, c1))
    d = layout.unpad(D)[0, 0]
    assert d[0] == DW and d[1] == DO and d[2] == 0.0


def test_rates_with_face_relaxivities_per_side_and_fluid():
    """Corner cell of a 3^3 pore block in a 5^3 box with c1 = 0.7: solid
    faces on -x, -y, -z; each face carries its own water and oil
    relaxivity, 1/T2 = c1/T2w + c2/T2o + sum_faces (c1 q2w + c2 q2o)/h."""
    shape = (5, 5, 5)
    mask = np.zeros(shape, np.uint8)
    mask[1:4, 1:4, 1:4] = 1
    p = oracle.params(shape, 1e-6, 1e-5, dy=2e-6, dz=4e-6)
    act = layout.pad_mask(mask)
    conc0 = np.full(shape, 0.7)
    c0, c1 = layout.pad_field(conc0), layout.pad_field(1.0 - conc0)
    faces = [np.zeros(oracle.padded_shape(p)) for _ in range(12)]
    # the corner cell is padded (2, 2, 2); its minus faces are stored at the cell itself
    corner = (2, 2, 2)
    vals = {}
    for f in range(12):
        faces[f][corner] = 1e-6 * (f + 1)  # distinct per (fluid, relaxation, axis)
        vals[f] = 1e-6 * (f + 1)
    fl = oracle.fluids(t1=(0.5, 0.25), t2=(0.2, 0.1), ncomp=2)
    _, T1, T2, _ = oracle.fill_coefficients_2f(p, act, fl, (c0, c1), faces)
    h = (1e-6, 2e-6, 4e-6)
    a, b = 0.7, 0.3
    r1 = a / 0.5 + b / 0.25 + sum((a * vals[ax] + b * vals[6 + ax]) / h[ax] for ax in range(3))
    r2 = a / 0.2 + b / 0.1 + sum((a * vals[3 + ax] + b * vals[9 + ax]) / h[ax] for ax in range(3))
    assert 1.0 / T1[corner] == pytest.approx(r1, rel=1e-13)
    assert 1.0 / T2[corner] == pytest.approx(r2, rel=1e-13)
    # the centre cell touches no solid: bulk rates only
    assert 1.0 / T2[3, 3, 3] == pytest.approx(a / 0.2 + b / 0.1, rel=1e-13)


def test_two_phase_sums_brute_force():
    """Cells on the tie line have c1 = x: the gate keeps x > 0.8 or x < 0.2
    and the sums weigh the field by x and 1 - x."""
    shape = (6, 7, 8)
    rng = np.random.default_rng(2)
    mask = (rng.random(shape) < 0.7).astype(np.uint8)
    x = rng.random(shape)
    f = rng.normal(size=shape)
    p = oracle.params(shape, 1e-6, 1e-5)
    act = layout.pad_mask(mask)
    fl = oracle.fluids(an=(1.0, 0.0), bn=(0.0, 1.0))
    s1, s2 = oracle.output_2f(p, act, (1, 0, 2), fl, (layout.pad_field(x), layout.pad_field(1 - x)),
                              layout.pad_field(f))
    sl = (slice(2, 4), slice(0, 7), slice(1, 7))
    keep = (mask[sl] == 1) & ((x[sl] > 0.8) | (x[sl] < 0.2))
    assert s1 == pytest.approx((f[sl] * x[sl])[keep].sum(), rel=1e-12)
    assert s2 == pytest.approx((f[sl] * (1 - x[sl]))[keep].sum(), rel=1e-12)
