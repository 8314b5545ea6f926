"""The surrogate network's forward and backward passes.

CPU: the oracle's analytic gradient (float64, no rounding) against central
finite differences of its loss, for every parameter block and all three loss
terms (a dropped term, a sign or a transposed operand fails). GPU: the device's
BF16 MFMA forward and backward kernels (through ptg_mlp_probe) against the
oracle with the device's BF16 roundings emulated.
"""
import numpy as np
import pytest

from oracle import mlp


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 8))
    x[:, :5] = rng.uniform(-1, 1, (n, 5))
    t = np.zeros((n, 8))
    t[:, 0] = rng.normal(0.4, 0.3, n)
    t[:, 1] = rng.uniform(0, 1, n)
    t[:, 2] = rng.uniform(0, 1, n)
    t[:, 3] = rng.integers(0, 2, n)
    t[:, 4] = rng.integers(0, 2, n)
    t[:, 5] = rng.uniform(0, 1, n)
    return x, t


def test_layout_matches_the_library():
    from ptarmigan_amd import _abi
    assert mlp.NPARAM == _abi.MLP_NPARAM == 38304


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], dtype=np.float32)
    got = mlp.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1); 1 + 3 * 2^-8 rounds up
    assert got.tolist() == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, float(np.float32(mlp.bf16(np.float32(3.0e38))[0]))]
    assert np.all(np.abs(mlp.bf16(np.linspace(-3, 3, 1001)) - np.linspace(-3, 3, 1001)) <= 2 ** -8 * 3)


@pytest.mark.parametrize("w_t", [1.0, 0.5])
def test_oracle_gradient_matches_finite_differences(w_t):
    p = mlp.init_params(3)
    x, t = _batch(48, 4)
    _, g, _ = mlp.forward_backward(p, x, t, w_t, emulate_bf16=False, dtype=np.float64)
    rng = np.random.default_rng(5)
    blocks = [(mlp.W1, mlp.W1 + 5 * 1), (mlp.B1, mlp.W2), (mlp.W2, mlp.B2), (mlp.B2, mlp.W3), (mlp.W3, mlp.B3),
              (mlp.B3, mlp.W4), (mlp.W4, mlp.W4 + 3 * mlp.H), (mlp.B4, mlp.B4 + 3)]
    h = 1e-6
    for lo, hi in blocks:
        for i in rng.integers(lo, hi, 6):
            q = p.copy()
            q[i] += h
            lp = mlp.mean_loss(q, x, t, w_t)
            q[i] -= 2 * h
            lm = mlp.mean_loss(q, x, t, w_t)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (lo, i, fd, g[i])
    # the padding inputs and the unused output rows get no gradient
    assert np.all(g[mlp.W4 + 3 * mlp.H:mlp.B4] == 0.0)
    assert np.all(g[mlp.B4 + 3:] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 512])
def test_device_mlp_matches_oracle(engine, n):
    p = mlp.init_params(7).astype(np.float32)
    x, t = _batch(n, 8)
    out, g, loss = engine.mlp_probe(p, x, t, 0.7)
    o_ref, g_ref, l_ref = mlp.forward_backward(p, x, t, 0.7, emulate_bf16=True, dtype=np.float32)
    print("mlp probe", n, flush=True)
    assert np.abs(out[:, :3] - o_ref[:, :3]).max() <= 2e-3 * (1.0 + np.abs(o_ref[:, :3]).max())
    assert np.allclose(out[:, 3:], 0.0)
    assert np.allclose(loss, l_ref, rtol=5e-3, atol=1e-5)
    for lo, hi in [(mlp.W1, mlp.B1), (mlp.B1, mlp.W2), (mlp.W2, mlp.B2), (mlp.B2, mlp.W3), (mlp.W3, mlp.B3),
                   (mlp.B3, mlp.W4), (mlp.W4, mlp.B4), (mlp.B4, mlp.NPARAM)]:
        a, b = g[lo:hi], g_ref[lo:hi]
        scale = np.abs(b).max() + 1e-12
        assert np.abs(a - b).max() <= 2e-2 * scale, (lo, np.abs(a - b).max(), scale)
