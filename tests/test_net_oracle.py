"""Pins of the network oracle (oracle/net_oracle.py): its gradient against
central finite differences of its loss (float64, no bf16 rounding), the
bf16 rounding helper against exact cases, and the parameter layout."""
import numpy as np

import net_oracle as N


def test_bf16_round_exact_cases():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 0.0, 3.0e38], dtype=np.float32)
    r = N.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0); 1 + 3*2^-8 -> 1 + 2^-6;
    # 3e38 = 1.76325 * 2^127 -> 226/128 * 2^127 (7 fraction bits)
    np.testing.assert_array_equal(r, np.array([1.0, 1.0, 1.0 + 2**-6, -2.5, 0.0, 226 / 128 * 2.0**127],
                                              dtype=np.float32))
    assert N.bf16_bits(np.array([1.0], dtype=np.float32))[0] == 0x3F80


def test_parameter_count():
    assert N.NUM_PARAMS == 32 * 9 * 8 + 32 + 2 * (32 * 9 * 32 + 32) + 9 * 32 + 1 == 21121


def test_gradient_matches_finite_differences():
    H, W = 6, 5
    p = N.init_params(3).astype(np.float64)
    X = N.random_features(H, W, 4).astype(np.float64)
    t = np.random.default_rng(5).normal(0, 1, (H, W))
    loss, g, _ = N.loss_grad(p, X, t, bf16=False)
    rng = np.random.default_rng(6)
    offs = np.cumsum([0] + [n for _, n in N.SIZES])
    # a few parameters of every block
    idx = np.concatenate([rng.choice(np.arange(offs[k], offs[k + 1]), size=min(6, offs[k + 1] - offs[k]),
                                     replace=False) for k in range(len(N.SIZES))])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (N.loss_grad(pp, X, t, bf16=False)[0] - N.loss_grad(pm, X, t, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_forward_is_a_convolution():
    # shifting the input by one pixel (away from the border) shifts the output
    H, W = 12, 12
    p = N.init_params(1).astype(np.float64)
    X = np.zeros((H + 2, W + 2, 8))
    X[6, 6, :6] = [1, -2, 0.5, 3, -1, 2]
    X[1:-1, 1:-1, 6] = 1.0
    X2 = np.zeros_like(X)
    X2[7, 7] = X[6, 6]
    X2[1:-1, 1:-1, 6] = 1.0
    d1, _ = N.forward(p, X, bf16=False)
    d2, _ = N.forward(p, X2, bf16=False)
    # away from the border (receptive field 9x9) the response moves with the input
    np.testing.assert_allclose(d2[5:8, 5:8], d1[4:7, 4:7], rtol=1e-12, atol=1e-12)
