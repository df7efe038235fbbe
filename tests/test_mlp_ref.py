"""Pins the CPU reference of the warm-start network: bfloat16 rounding on
known bit patterns, and the float64 gradient against central finite
differences (a dropped term, a wrong transpose or a wrong GELU derivative
fails the latter)."""
import numpy as np

from tests import mlp_ref


def _bits(v):
    return np.array([v], dtype=np.uint32).view(np.float32)


def test_bf16_rounding_known_patterns():
    cases = [
        (0x3F800000, 0x3F800000),  # 1.0 exact
        (0x3F808000, 0x3F800000),  # tie, even below: stays
        (0x3F818000, 0x3F820000),  # tie, odd below: rounds up
        (0x3F808001, 0x3F810000),  # above the tie: up
        (0x3F807FFF, 0x3F800000),  # below the tie: down
        (0xBF818000, 0xBF820000),  # sign kept
        (0x7F7FFFFF, 0x7F800000),  # largest float rounds to infinity
    ]
    for src, want in cases:
        got = mlp_ref.bf16(_bits(src)).astype(np.float32).view(np.uint32)[0]
        assert got == want, (hex(src), hex(got), hex(want))


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    dims = (3, 16, 5)
    npar = 3 * 16 + 16 + 16 * 16 + 16 + 16 * 5 + 5
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((7, 3))
    T = rng.standard_normal((7, 5))
    loss, g = mlp_ref.loss_grad(p, dims, X, T)
    h = 1e-6
    fd = np.empty(npar)
    for k in range(npar):
        e = np.zeros(npar)
        e[k] = h
        fd[k] = (mlp_ref.loss_grad(p + e, dims, X, T)[0] - mlp_ref.loss_grad(p - e, dims, X, T)[0]) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-7 * (1 + np.abs(g).max())
    y = mlp_ref.predict(p, dims, X)
    assert abs(loss - ((y - T) ** 2).mean()) <= 1e-14 * (1 + loss)


def test_bf16_products_stay_close_to_float64():
    rng = np.random.default_rng(8)
    dims = (16, 64, 40)
    npar = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 40 + 40
    p = rng.standard_normal(npar) * 0.2
    X = rng.standard_normal((32, 16))
    T = rng.standard_normal((32, 40))
    l0, g0 = mlp_ref.loss_grad(p, dims, X, T)
    l1, g1 = mlp_ref.loss_grad(p, dims, X, T, bf16_products=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 3e-2 * np.linalg.norm(g0)
