"""The oracle's learned-inflation network (no reference counterpart): the
analytic gradient is pinned by central finite differences, the output by its
closed form at zero weights."""
import numpy as np
import pytest


def _params(orc, seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(orc.MLP_NPARAM) * scale


def test_mlp_zero_weights_closed_form(orc):
    p = np.zeros(orc.MLP_NPARAM)
    p[-1] = 0.7  # b3
    feat = np.random.default_rng(1).standard_normal((5, 8))
    rho = orc.mlp_forward(p, feat)
    np.testing.assert_allclose(rho, 1.0 + np.log1p(np.exp(0.7)), rtol=1e-15)


def test_mlp_gradient_matches_central_differences(orc):
    p = _params(orc)
    rng = np.random.default_rng(2)
    feat = rng.standard_normal((7, 8))
    target = 1.0 + rng.random(7)
    loss, g = orc.mlp_loss_grad(p, feat, target)
    # loss is the mean squared error of rho
    rho = orc.mlp_forward(p, feat)
    assert loss == pytest.approx(np.mean((rho - target) ** 2), rel=1e-14)
    idx = np.concatenate([rng.choice(512, 6, replace=False),            # W1
                          512 + rng.choice(64, 3, replace=False),       # b1
                          576 + rng.choice(4096, 8, replace=False),     # W2
                          4672 + rng.choice(64, 3, replace=False),      # b2
                          4736 + rng.choice(64, 3, replace=False),      # W3
                          [4800]])                                      # b3
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp_loss_grad(pp, feat, target)[0] - orc.mlp_loss_grad(pm, feat, target)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-10), i
