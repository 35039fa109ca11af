"""Pins the surrogate-MLP oracle (oracle/mlp_oracle.py) before it checks the
device gradients: its exact-mode gradients against central finite differences
of its own forward pass (a dropped term, a wrong transpose or a wrong GELU
derivative fails), its device-rounding mode against the exact mode, and the
bfloat16 rounding against hand-computed values."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert mo.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    # 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert mo.bf16(np.array([1.0 + 3 * 2.0 ** -9]))[0] == 1.0 + 2.0 ** -7
    assert mo.bf16(np.array([-2.5]))[0] == -2.5
    assert mo.bf16(np.array([np.pi]))[0] == 3.140625


def test_gelu_derivative_matches_finite_difference():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(x))) < 1e-8
    assert abs(mo.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15


def test_exact_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    m = mo.MLPOracle(din=5, hidden=7, layers=3)
    p = rng.normal(0, 0.6, size=m.count())
    x = rng.uniform(-1, 1, size=(9, 5))
    t = rng.uniform(-1, 1, size=9)
    g, loss = m.gradients(p, x, t, device=False)
    y, _ = m.forward(p, x, device=False)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-15
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(len(p)):
        q = p.copy()
        q[i] += h
        lp = np.mean((m.forward(q, x, device=False)[0] - t) ** 2)
        q[i] -= 2 * h
        lm = np.mean((m.forward(q, x, device=False)[0] - t) ** 2)
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))


def test_device_mode_close_to_exact():
    rng = np.random.default_rng(4)
    m = mo.MLPOracle(din=20, hidden=64, layers=3)
    p = rng.normal(0, 0.15, size=m.count())
    x = rng.uniform(-1, 1, size=(64, 20))
    t = rng.uniform(-1, 1, size=64)
    g0, _ = m.gradients(p, x, t, device=False)
    g1, _ = m.gradients(p, x, t, device=True)
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 2e-2
    assert not np.array_equal(g0, g1)
