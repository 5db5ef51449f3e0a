"""Pins the cost model's CPU oracle (oracle/mlp_oracle.py) to what the
mathematics fixes: the analytic gradient against central finite differences
in float64, Adam's first step against its closed form, and the unit features
against hand-computed values.  A planted mistake in a copy of the oracle's
backward pass must fail the same check."""
import numpy as np
import pytest

import mlp_oracle as mo


def fd_grad(p, x, t, shape, eps=1e-6):
    g = np.zeros_like(p)
    for i in range(p.size):
        a = p.copy()
        b = p.copy()
        a[i] += eps
        b[i] -= eps
        g[i] = (mo.loss_grad(a, x, t, shape)[0] - mo.loss_grad(b, x, t, shape)[0]) / (2 * eps)
    return g


def small_case(seed=0, shape=(5, 7, 2), n=13):
    rng = np.random.default_rng(seed)
    n_in, h, n_out = shape
    npar = n_in * h + h + h * h + h + h * n_out + n_out
    p = rng.normal(0, 0.7, npar)
    x = rng.normal(0, 1, (n, n_in))
    t = rng.normal(0, 1, (n, n_out))
    return p, x, t, shape


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_finite_differences(seed):
    p, x, t, shape = small_case(seed)
    loss, g = mo.loss_grad(p, x, t, shape)
    fd = fd_grad(p, x, t, shape)
    assert loss > 0
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_forward_closed_form_single_unit():
    # 1-1-1-1 network: y = w3 * relu(w2 * relu(w1 x + b1) + b2) + b3
    p = np.array([2.0, -1.0, 3.0, 0.5, -4.0, 0.25])
    x = np.array([[1.5], [0.2], [-3.0]])
    y, _ = mo.forward(p, x, (1, 1, 1))
    a1 = np.maximum(2.0 * x[:, 0] - 1.0, 0)
    a2 = np.maximum(3.0 * a1 + 0.5, 0)
    np.testing.assert_allclose(y[:, 0], -4.0 * a2 + 0.25)


def test_planted_backward_mistake_is_caught():
    """A plausible mistake (W2 not transposed in backprop, with a square W2)
    must fail the finite-difference check that pins the oracle."""
    p, x, t, shape = small_case(3, shape=(4, 6, 2))

    def wrong(p, x, t, shape):
        W1, b1, W2, b2, W3, b3 = mo.unpack(p, *shape)
        y, (z1, a1, z2, a2) = mo.forward(p, x, shape)
        dy = 2.0 * (y - t) / y.size
        dz2 = (dy @ W3.T) * (z2 > 0)
        dz1 = (dz2 @ W2) * (z1 > 0)          # planted: missing transpose
        return np.concatenate([a.ravel() for a in (x.T @ dz1, dz1.sum(0), a1.T @ dz2, dz2.sum(0),
                                                   a2.T @ dy, dy.sum(0))])

    fd = fd_grad(p, x, t, shape)
    assert not np.allclose(wrong(p, x, t, shape), fd, rtol=1e-5, atol=1e-7)
    assert np.allclose(mo.loss_grad(p, x, t, shape)[1], fd, rtol=1e-5, atol=1e-7)


def test_adam_first_step_closed_form():
    # step 1: m_hat = g, v_hat = g^2, so the update is lr * g / (|g| + eps)
    rng = np.random.default_rng(5)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    lr = 0.01
    p1, m, v = mo.adam_step(p, g, np.zeros(50), np.zeros(50), 1, lr)
    np.testing.assert_allclose(p1, p - lr * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_unit_features_hand_computed():
    full = np.full(6, 0xFFFFFFFF, np.uint32)
    one = np.array([1, 0, 0, 0, 0, 0], np.uint32)
    mask = np.array([1, 0, 0, 0, 0, 0], np.uint32)        # popcount 1
    orw = np.array([0xFF, 0, 0, 0, 0, 0x3], np.uint32)    # popcount 10
    q = np.stack([full, one])                               # popcounts 192, 1
    f = mo.unit_features(mask, orw, 100, 100 * 12, q)
    assert f[0] == pytest.approx(np.log2(101) / 20)
    assert f[1] == pytest.approx(1 / 64)
    assert f[2] == pytest.approx(np.log2(3) / 12)
    assert f[3] == pytest.approx(96.5 / 192)
    assert f[4] == pytest.approx(95.5 / 192)
    assert f[5] == pytest.approx(1 / 192)
    assert f[6] == pytest.approx(1.0)
    assert list(f[7:11]) == [0.5, 0.0, 0.0, 0.5]
    assert f[11] == pytest.approx(12 / 64)
    assert f[12] == pytest.approx(10 / 192)
    assert f[13] == pytest.approx((10 + 1) / 2 / 192)
    assert f[14] == pytest.approx(10 / 192)
    assert f[15] == pytest.approx(np.log2(1 + 200) / 32)
    # an empty batch keeps only the partition's own features
    e = mo.unit_features(mask, orw, 100, 1200, np.zeros((0, 6), np.uint32))
    assert e[2] == 0 and e[3:11].sum() == 0 and e[13:].sum() == 0 and e[0] > 0


def test_partition_stats_brute_force():
    rng = np.random.default_rng(9)
    w = rng.integers(0, 2**32, (20, 6), dtype=np.uint64).astype(np.uint32)
    off = np.array([0, 5, 5, 20], np.uint32)
    orw, cnt, pop = mo.partition_stats(w, off)
    assert list(cnt) == [5, 0, 15]
    for p, (a, b) in enumerate([(0, 5), (5, 5), (5, 20)]):
        acc = np.zeros(6, np.uint32)
        s = 0
        for i in range(a, b):
            acc |= w[i]
            s += sum(bin(int(v)).count("1") for v in w[i])
        assert np.array_equal(orw[p], acc)
        assert pop[p] == s
