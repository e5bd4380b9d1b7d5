"""Pins the float64 network reference (tests/mlp_ref.py) that checks the GPU
kernels: its gradient against central finite differences, its BF16 rounding
against hand-computed values, and its layout against the parameter count."""
import numpy as np

from tests import mlp_ref


def test_parameter_layout():
    assert sum(fi * fo + fo for fi, fo in mlp_ref.SIZES) == 5378


def test_bf16_rounding_nearest_even():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0);
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: rounds up to even
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -2.5, 3.0e-3], np.float32)
    got = mlp_ref.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1 + 2 ** -6 and got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, 5378)
    X = rng.normal(0, 1, (40, 16))
    T = rng.normal(0, 1, (40, 2))
    mask = (rng.random(40) > 0.2).astype(np.float64)
    loss, g = mlp_ref.loss_and_grad(p, X, T, mask, bf16_ops=False)
    idx = rng.choice(5378, 60, replace=False)
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp_ref.loss_and_grad(pp, X, T, mask, False)[0] - mlp_ref.loss_and_grad(pm, X, T, mask, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_planted_gradient_mistake_is_caught():
    # a backward pass with the ReLU mask dropped disagrees with finite differences
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, 5378)
    X = rng.normal(0, 1, (20, 16))
    T = rng.normal(0, 1, (20, 2))
    mask = np.ones(20)
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p)
    y, (x, z1, a1, z2, a2) = mlp_ref.forward(p, X, False)
    dy = 2 * (y - T) / 20
    dz2_wrong = dy @ W3.T  # missing * (z2 > 0)
    gb2_wrong = dz2_wrong.sum(0)
    _, g = mlp_ref.loss_and_grad(p, X, T, mask, False)
    o = 16 * 64 + 64 + 64 * 64
    assert not np.allclose(gb2_wrong, g[o:o + 64])
