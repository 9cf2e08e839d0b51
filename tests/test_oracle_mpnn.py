"""The CPU statement of the warm-start network (oracle/or_mpnn.py) pinned by
what the mathematics fixes: central finite differences of the loss (FP64, no
rounding) for the backward formulas, the adjoint identity of the message
aggregation, torch's float32 -> bfloat16 conversion for the rounding, and the
invariances the feature scaling is built for."""
import numpy as np
import pytest

import or_mpnn as om
import systems


def _sample(nx, ny, nz, seed):
    nbr, nnbr, ac, anb, b = systems.pressure_like(nx, ny, nz, seed, shift=0.0)
    rng = np.random.default_rng(seed + 100)
    x = rng.standard_normal(len(ac))
    vol = rng.uniform(0.5, 1.5, len(ac))
    nbr0 = nbr.astype(np.int64) - 1
    X0, e, t, sb, sa = om.features(nbr0, b, ac, anb, vol, x)
    return nbr0, X0, e, t


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.0e38, 1e-40], np.float32)])
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(om.bf16(a), ref)


def test_aggregation_adjoint():
    nbr0, X0, e, t = _sample(4, 3, 3, 1)
    rng = np.random.default_rng(0)
    H = rng.standard_normal((X0.shape[0], 64))
    D = rng.standard_normal((X0.shape[0], 64))
    lhs = np.sum(om._agg(nbr0, e, H) * D)
    rhs = np.sum(H * om._agg_t(nbr0, e, D))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_backward_matches_finite_differences():
    samples = [_sample(4, 3, 3, 1), _sample(4, 3, 3, 2)]
    theta = om.init_params(5) + 0.01  # non-zero biases: every relu branch in play
    loss, caches, _ = om.forward(theta, samples, bf16_round=False)
    g = om.backward(theta, samples, caches, bf16_round=False)
    rng = np.random.default_rng(7)
    idx = rng.choice(len(theta), 60, replace=False)
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (om.forward(tp, samples, bf16_round=False)[0] - om.forward(tm, samples, bf16_round=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (k, fd, g[k])


def test_scale_invariance_of_features_and_prediction():
    """Scaling b (and x) or the whole operator leaves the normalised problem unchanged."""
    nbr, nnbr, ac, anb, b = systems.pressure_like(4, 3, 3, 4, shift=0.0)
    nbr0 = nbr.astype(np.int64) - 1
    vol = np.ones(len(ac))
    x = np.linspace(-1, 1, len(ac))
    xp = np.cos(np.arange(len(ac)))
    base = om.features(nbr0, b, ac, anb, vol, x, xp)
    for cb, ca in ((3.0, 1.0), (1.0, 7.0), (0.01, 100.0)):
        f = om.features(nbr0, cb * b, ca * ac, ca * anb, 2.0 * vol, cb / ca * x, cb / ca * (xp + 5.0))
        np.testing.assert_allclose(f[0], base[0], rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(f[1], base[1], rtol=1e-13)
        np.testing.assert_allclose(f[2], base[2], rtol=1e-12, atol=1e-15)


def test_bf16_forward_close_to_fp64():
    samples = [_sample(4, 3, 3, 1)]
    theta = om.init_params(5)
    l64 = om.forward(theta, samples, bf16_round=False)[0]
    l16 = om.forward(theta, samples, bf16_round=True)[0]
    assert abs(l16 - l64) <= 0.05 * l64
