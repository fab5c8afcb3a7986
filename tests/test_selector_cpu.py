"""Pins the FP64 selector oracle by central finite differences (CPU)."""
import numpy as np

import mlp_oracle as M


def test_param_count():
    assert M.NPARAM == 5508


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = M.init(1)
    x = rng.normal(size=(9, 16))
    y = rng.integers(0, 4, size=9)
    sw = rng.uniform(0.5, 2.0, size=9)
    _, g = M.loss_grad(w, x, y, sw)
    idx = rng.choice(M.NPARAM, size=60, replace=False)
    for i in idx:
        e = np.zeros_like(w)
        e[i] = 1e-6
        fd = (M.loss_grad(w + e, x, y, sw)[0] - M.loss_grad(w - e, x, y, sw)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.00390625, 1.01171875, 1.0078125, -2.5, 3.140625], dtype=np.float32)
    r = M.bf16_round(v)
    # BF16 keeps 7 fraction bits: spacing 2^-7 near 1.  1 + 2^-8 is a tie ->
    # even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6); 1 + 2^-7 exact
    assert r.tolist() == [1.0, 1.0, 1.015625, 1.0078125, -2.5, 3.140625]


def test_class_costs_charge_shipping_to_push():
    from grudon_amd.train_selector import class_costs

    c5 = np.array([[5.0, 3.0, 4.0, 2.5, 1.0], [1.0, 2.0, 1.5, 3.0, 9.0]], dtype=np.float32)
    cc = class_costs(c5)
    assert cc.tolist() == [[1.0, 3.0, 1.0, 2.5], [9.0, 2.0, 9.0, 3.0]]
    assert cc.argmin(axis=1).tolist() == [0, 1]
    c4 = c5[:, :4]
    assert class_costs(c4) is c4
