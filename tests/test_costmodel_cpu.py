"""CPU checks of the cost-model oracle: the float64 gradient against central
finite differences (what the mathematics fixes), and the parameter packing
size the C ABI declares.  No GPU call."""
import numpy as np


def test_packing_size(orc):
    assert orc.MLP_NPARAMS == 4610


def test_oracle_gradient_matches_finite_differences(orc):
    rng = np.random.default_rng(0)
    p = rng.standard_normal(orc.MLP_NPARAMS) * 0.3
    x = rng.standard_normal((17, 4))
    t = rng.standard_normal((17, 2))
    loss, g = orc.mlp_loss_grad(p, x, t)
    h = 1e-6
    idx = rng.choice(orc.MLP_NPARAMS, size=200, replace=False)
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = orc.mlp_loss_grad(p + e, x, t)
        lm, _ = orc.mlp_loss_grad(p - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), (k, fd, g[k])
    # loss is the mean over batch and outputs of the weighted squared residual
    y = orc.mlp_forward(p, x)
    r = y - t
    assert np.isclose(loss, np.mean(np.where(r < 0, 4.0, 1.0) * r ** 2))
