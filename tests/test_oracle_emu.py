"""Pins of the emulator's FP64 oracle (oracle/emu_oracle.py): its loss gradient against central finite differences
of its own forward loss (a dropped GELU derivative, a transposed weight or a missing bias term fails it), its GELU
against scipy's normal CDF, and its bf16 rounding against torch's."""
import numpy as np
import pytest
from scipy import special

import emu_oracle as E


def _net(seed=0, dims=(5, 7, 8, 2)):
    rng = np.random.default_rng(seed)
    n_in, n_out, hidden, layers = dims
    sizes = [n_in] + [hidden] * layers + [n_out]
    p = np.concatenate([np.concatenate([rng.standard_normal(sizes[l + 1] * sizes[l]) / np.sqrt(sizes[l]),
                                        0.1 * rng.standard_normal(sizes[l + 1])]) for l in range(layers + 1)])
    scale = (rng.standard_normal(n_in), 1.0 + rng.random(n_in), rng.standard_normal(n_out), 1.0 + rng.random(n_out))
    x = rng.standard_normal((9, n_in))
    t = rng.standard_normal((9, n_out))
    w = rng.random(n_out) + 0.5
    return p, x, t, w, dims, scale


def test_gradient_by_central_differences():
    p, x, t, w, dims, scale = _net()
    loss, g = E.loss_and_grad(p, x, t, dims, scale, w)
    h = 1e-6
    idx = np.random.default_rng(1).choice(p.size, 40, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (E.loss_and_grad(pp, x, t, dims, scale, w)[0] - E.loss_and_grad(pm, x, t, dims, scale, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_against_normal_cdf():
    x = np.linspace(-8, 8, 101)
    # (1 + erf) cancels for large negative x: the form the kernels use too, absolute error ~1e-16
    np.testing.assert_allclose(E.gelu(x), x * special.ndtr(x), rtol=1e-13, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(E.gelu_grad(x), (E.gelu(x + h) - E.gelu(x - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    a = np.concatenate([rng.standard_normal(1000) * 10.0 ** rng.uniform(-20, 20, 1000), [0.0, -0.0, 1.0, 65504.0]])
    ref = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(E.bf16(a), ref)
