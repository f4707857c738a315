"""The FP64 U-Net reference's backward pass against finite differences (CPU).

Pins the reference that the GPU gradients are checked against: in plain FP64
(no BF16 rounding) analytic gradients of randomly chosen parameters of every
convolution equal central differences of the loss.
"""
import numpy as np

from oracle import unet_ref as U


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    sl, n = U.param_slices()
    p = np.zeros(n)
    for (w, b, ci, co) in sl:
        p[w[0]:w[1]] = rng.normal(0, np.sqrt(2.0 / (9 * ci)), w[1] - w[0])
        p[b[0]:b[1]] = rng.normal(0, 0.05, co)
    inputs = np.abs(rng.normal(0.5, 0.5, (1, 16, 16, 9)))
    inputs[..., 6:9] = rng.normal(0, 1, (1, 16, 16, 3))
    target = np.abs(rng.normal(0.5, 0.4, (1, 16, 16, 3)))

    def loss(params):
        net = U.Net(params, bf16=False)
        s = net.forward(inputs)
        return net.loss(s, target)[0]

    net = U.Net(p, bf16=False)
    s = net.forward(inputs)
    L, d = net.loss(s, target)
    g = net.backward(s, d)
    checked = 0
    for (w, b, ci, co) in sl:
        for idx in list(rng.integers(w[0], w[1], 3)) + [int(rng.integers(b[0], b[1]))]:
            eps = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[idx] += eps
            pm[idx] -= eps
            fd = (loss(pp) - loss(pm)) / (2 * eps)
            assert abs(fd - g[idx]) <= 1e-3 * max(1e-2, abs(fd)) + 1e-6, (idx, fd, g[idx])  # L1 and ReLU kinks
            checked += 1
    assert checked == 4 * len(sl)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e38], np.float32)
    r = U.round_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.0078125 and r[4] == -2.5
