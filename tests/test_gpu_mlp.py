"""GPU tests of the warm-start network (include/pdhcg_mlp.h).

The BF16 matrix-core product is compared with a float64 product of the same
bfloat16-rounded operands: the bound is the FP32 accumulation error,
K * 2^-23 * (|A| |B|) elementwise, which any misplaced lane, transposed
operand or dropped k-block exceeds by orders of magnitude.  The network's
loss and gradient are compared with the CPU reference computing with the same
bfloat16 roundings (tests/mlp_ref.py, itself pinned by finite differences).
"""
import numpy as np
import pytest

from pdhcg_amd import warmstart
from tests import mlp_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (48, 80, 32), (64, 64, 64), (128, 192, 1024), (272, 16, 48)])
def test_mfma_gemm_matches_bf16_product(ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K + 100 * ta + 10 * tb)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    Ar, Br = mlp_ref.bf16(A), mlp_ref.bf16(B)
    opA, opB = (Ar.T if ta else Ar), (Br.T if tb else Br)
    ref = opA @ opB
    bound = K * 2.0 ** -23 * (np.abs(opA) @ np.abs(opB)) + 1e-30
    got = warmstart.mfma_gemm(A, B, bool(ta), bool(tb)).astype(np.float64)
    assert got.shape == (M, N)
    assert np.all(np.abs(got - ref) <= bound), np.abs(got - ref).max()


def test_gemm_rejects_unpadded_sizes():
    with pytest.raises(RuntimeError):
        warmstart.mfma_gemm(np.ones((17, 16)), np.ones((16, 16)))


def _net(dims, seed):
    net = warmstart.MLP(*dims, seed=seed)
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(net.num_params) * 0.2).astype(np.float32)
    net.set_params(p)
    return net, p


def test_params_round_trip():
    net, p = _net((5, 32, 7), 1)
    assert net.num_params == 5 * 32 + 32 + 32 * 32 + 32 + 32 * 7 + 7
    assert np.array_equal(net.get_params(), p)
    net.close()


@pytest.mark.parametrize("dims,rows", [((16, 64, 40), 32), ((3, 128, 150), 48), ((20, 1024, 33), 64)])
def test_loss_and_gradient_match_reference(dims, rows):
    net, p = _net(dims, rows)
    rng = np.random.default_rng(rows + 1)
    X = rng.standard_normal((rows, dims[0])).astype(np.float32)
    T = rng.standard_normal((rows, dims[2])).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    rl, rg = mlp_ref.loss_grad(p, dims, X, T, bf16_products=True)
    assert abs(loss - rl) <= 1e-4 * rl
    # per parameter block (W1, b1, W2, b2, W3, b3)
    i, h, o = dims
    sizes = [i * h, h, h * h, h, h * o, o]
    off = 0
    for k, s in enumerate(sizes):
        d, r = g[off:off + s], rg[off:off + s]
        assert np.linalg.norm(d - r) <= 2e-3 * np.linalg.norm(r) + 1e-12, (k, np.linalg.norm(d - r) / np.linalg.norm(r))
        off += s
    # the weights did not change
    assert np.array_equal(net.get_params(), p)
    net.close()


def test_training_fits_a_smooth_map_and_predict_matches_reference():
    dims = (8, 256, 24)
    rng = np.random.default_rng(5)
    Z = rng.standard_normal((1024, 8)).astype(np.float32)
    W = rng.standard_normal((8, 24)) / np.sqrt(8)
    T = (np.tanh(Z @ W) + 0.1 * (Z @ W) ** 2).astype(np.float32)
    net = warmstart.MLP(*dims, seed=3)
    first = net.train(Z, T, epochs=1, batch=128, lr=1e-3, seed=1)
    last = net.train(Z, T, epochs=60, batch=128, lr=1e-3, seed=2)
    assert last < 0.05 * first, (first, last)
    Y = net.predict(Z[:100])  # 100 rows: padded to 112 inside
    ref = mlp_ref.predict(net.get_params(), dims, Z[:100], bf16_products=True)
    assert np.abs(Y - ref).max() <= 1e-3 * (1 + np.abs(ref).max())
    net.close()
