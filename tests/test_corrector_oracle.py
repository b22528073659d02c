"""Pins the FP64 restatement of the corrector net (oracle/corrector.py):
against torch.nn.functional.conv3d + autograd on the CPU (the committed
fixture tests/golden/corrector_fixture.npz and live torch), against central
finite differences, and its BF16 rounding against torch's bfloat16 cast."""
import os

import numpy as np
import pytest

from oracle import corrector as oc

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fix():
    return dict(np.load(os.path.join(HERE, "golden", "corrector_fixture.npz")))


def test_forward_matches_torch_fixture(fix):
    out = oc.forward(fix["params"], fix["features"], emulate=False)
    np.testing.assert_allclose(out, fix["out"], rtol=1e-12, atol=1e-12)


def test_backward_matches_torch_autograd_fixture(fix):
    g = oc.backward(fix["params"], fix["features"], fix["grad_out"], emulate=False)
    np.testing.assert_allclose(g, fix["grad_params"], rtol=1e-10, atol=1e-10)


def test_backward_against_finite_differences(fix):
    rng = np.random.default_rng(3)
    p, f, go = fix["params"].copy(), fix["features"], fix["grad_out"]
    g = oc.backward(p, f, go, emulate=False)
    # a few weights of every layer and the biases
    idx = list(rng.choice(2592, 4, replace=False)) + [2592 + 3] + list(2608 + rng.choice(6912, 4, replace=False)) \
        + [9520 + 7] + list(9536 + rng.choice(864, 4, replace=False)) + [10400, 10401]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = ((oc.forward(pp, f, emulate=False) - oc.forward(pm, f, emulate=False)) * go).sum() / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-7), i


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-20, 100), rng.normal(0, 1e20, 100),
                        [0.0, -0.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, np.inf, -np.inf]]).astype(np.float32)
    want = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(oc.bf16(x), want)


def test_emulated_forward_close_to_exact(fix):
    """BF16 storage changes the output by a few parts in a thousand: the
    scale of error the GPU comparison must allow."""
    a = oc.forward(fix["params"], fix["features"], emulate=False)
    b = oc.forward(fix["params"], fix["features"], emulate=True)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-5 < rel < 2e-2
