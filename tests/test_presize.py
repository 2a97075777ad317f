"""CPU tests of the learned presizing: the oracle pinned by finite
differences and closed forms, the parameter layout, and the options."""
import numpy as np
import pytest

import wtet_amd
from wtet_amd import _core, presize
from oracle import mlp_ref as R


def _random_params(seed, X, Y):
    p = presize.init_params(X, Y, seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    a, b = R.OFF["b1"]
    p[a:b] = rng.standard_normal(b - a) * 0.1
    a, b = R.OFF["b2"]
    p[a:b] = rng.standard_normal(b - a) * 0.1
    return p


def test_layout_matches_header():
    assert R.NPARAMS == _core.PRESIZE_PARAMS == presize._o
    assert R.OFF == presize.OFF
    assert R.F == _core.PRESIZE_FEATURES and R.O == _core.PRESIZE_OUTPUTS


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((6, R.F))
    Y = rng.standard_normal((6, R.O))
    p = _random_params(4, X, Y)
    _, g = R.loss_grad(p, X, Y)
    lo, hi = R.TRAIN
    idx = np.concatenate([rng.integers(lo, hi, 300), np.arange(R.OFF["b3"][0], R.OFF["b3"][1]),
                          np.arange(R.OFF["W3"][0], R.OFF["W3"][0] + 20)])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = R.loss_grad(q, X, Y)
        q[i] -= 2 * h
        lm, _ = R.loss_grad(q, X, Y)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # outside the trainable range the gradient is zero
    assert not g[:lo].any() and not g[hi:].any()


def test_oracle_features_closed_form_on_a_lattice():
    """16^3 grid i/15: every 3-bit cell of every axis holds 2 of the 16
    values, so each of the 512 bins holds 1/512; mean 1/2; variance
    (16^2 - 1) / (12 * 15^2); no cross covariance."""
    g = wtet_amd.workloads.lattice(16)
    f = R.features(g)
    assert np.allclose(f[:512], 1 / 512, rtol=0, atol=1e-15)
    assert np.allclose(f[512:515], 0.5, atol=1e-7)
    var = (16 ** 2 - 1) / (12 * 15 ** 2)
    assert np.allclose(f[515:518], var, rtol=1e-6)
    assert np.allclose(f[518:521], 0, atol=1e-7)
    assert f[521] == pytest.approx(np.log(4096))


def test_oracle_features_single_cell():
    """All points in one octant cell: one bin holds everything."""
    rng = np.random.default_rng(2)
    p = (rng.random((1000, 3)) * 0.1).astype(np.float32)
    p[0] = 0.0
    p[1] = 0.8  # box corner far away: the rest fall in the lowest cell
    f = R.features(p)
    assert f[0] == pytest.approx(999 / 1000)
    assert f[:512].sum() == pytest.approx(1.0)


def test_prediction_inverts_standardization():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((40, R.F))
    Y = rng.standard_normal((40, R.O)) * 3 + 2
    p = presize.init_params(X, Y, 1).astype(np.float64)
    p[R.OFF["W3"][0]:R.OFF["b3"][1]] = 0.0  # output = b3 = 0 -> target mean
    m = presize.set_margin(p, [0.1, 0.2, 0.3]).astype(np.float64)
    pred = R.predict(m, X[:3])
    assert np.allclose(pred, Y.mean(0) + np.array([0.1, 0.2, 0.3]), atol=1e-5)


def test_presize_option_validation():
    assert wtet_amd.Config(presize="learned")._c().presize_mode == 1
    assert wtet_amd.Config(presize="grow")._c().presize_mode == 2
    assert wtet_amd.Config()._c().presize_mode == 0
    with pytest.raises(ValueError, match="Invalid presize"):
        wtet_amd.Config(presize="huge")
    with pytest.raises(ValueError, match="84896"):
        _core._params(np.zeros(10))
