"""Pins of the network oracle (oracle/net_oracle.py): its gradient against
central finite differences of its loss (float64, no bf16 rounding), the
adjoints of its convolution and up-sampling by the dot-product identity, the
block layouts, the bf16 rounding helper against exact cases, and the
parameter layout."""
import numpy as np
import pytest

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
    assert N.NUM_PARAMS == 6 * (32 * 9 * 32 + 32) + 4 * 9 * 32 + 4 == 56644


def test_block_layouts_are_inverse():
    # s2d gathers the features of each 2x2 block of cells (cells beyond an odd
    # edge read 0), feature 7 replaced by the largest |feature 5| over the
    # cell's 3x3 neighbourhood; d2s scatters the four head outputs back
    H, W = 5, 7
    rng = np.random.default_rng(0)
    X = np.zeros((H + 2, W + 2, 8))
    X[1:-1, 1:-1] = rng.normal(size=(H, W, 8))
    XS = N.s2d(X)
    assert XS.shape == (3, 4, 32)
    for y in range(H):
        for x in range(W):
            s = 2 * (y % 2) + (x % 2)
            np.testing.assert_array_equal(XS[y // 2, x // 2, 8 * s : 8 * s + 7], X[1 + y, 1 + x, :7])
            assert XS[y // 2, x // 2, 8 * s + 7] == np.max(np.abs(X[y : y + 3, x : x + 3, 5]))
    assert np.all(XS[2, :, 16:] == 0.0) and np.all(XS[:, 3, 8:16] == 0.0) and np.all(XS[:, 3, 24:] == 0.0)
    D4 = rng.normal(size=(3, 4, 4))
    D = N.d2s(D4, H, W)
    for y in range(H):
        for x in range(W):
            assert D[y, x] == D4[y // 2, x // 2, 2 * (y % 2) + (x % 2)]


def test_residual_operator():
    # A is symmetric; a constant field has A v = v (no flux across any face);
    # with D = e - 1 on every face (features 3, 4 = 1) an isolated unit value
    # becomes 1 + 4 (e - 1) at its cell and -(e - 1) at its four neighbours
    H, W = 5, 6
    X = np.zeros((H + 2, W + 2, 8))
    X[1:-1, 1:-1, 3:5] = 1.0
    De, Dn = N.face_numbers(X)
    rng = np.random.default_rng(2)
    u, v = rng.normal(size=(H, W)), rng.normal(size=(H, W))
    assert abs(np.sum(N.apply_A(u, De, Dn) * v) - np.sum(u * N.apply_A(v, De, Dn))) <= 1e-12
    np.testing.assert_allclose(N.apply_A(np.full((H, W), 2.5), De, Dn), 2.5, rtol=0, atol=1e-12)
    d = np.zeros((H, W))
    d[2, 3] = 1.0
    r = N.apply_A(d, De, Dn)
    D = np.e - 1
    assert abs(r[2, 3] - (1 + 4 * D)) <= 1e-12
    for y, x in ((1, 3), (3, 3), (2, 2), (2, 4)):
        assert abs(r[y, x] + D) <= 1e-12
    assert abs(np.sum(r) - 1.0) <= 1e-12  # conservative
    # the loss is the squared ratio of residuals: a prediction equal to the
    # target gives 0, a zero prediction gives 1
    t = rng.normal(size=(H, W))
    assert N.residual_loss(t, t, X)[0] == 0.0
    assert abs(N.residual_loss(np.zeros((H, W)), t, X)[0] - 1.0) <= 1e-12


def test_adjoints_by_the_dot_product_identity():
    # <conv(X), G> = <X, conv_t(G)> and <up(C), G> = <C, up_t(G)>: the backward
    # pass's transposes are the adjoints of the forward operators
    rng = np.random.default_rng(1)
    H, W = 9, 6
    X = rng.normal(size=(H, W, 32))
    Wt = rng.normal(size=(32, 9, 32))
    G = rng.normal(size=(H, W, 32))
    lhs = np.sum(N._conv(N._pad(X), Wt) * G)
    rhs = np.sum(X * N._conv_t(G, Wt))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    Cc = rng.normal(size=(5, 3, 32))
    lhs = np.sum(N._up(Cc, H, W) * G)
    rhs = np.sum(Cc * N._up_t(G))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # the pool is a quarter of up's adjoint; a constant image pools to itself
    np.testing.assert_allclose(N._pool(G), 0.25 * N._up_t(G), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(N._pool(np.full((4, 6, 2), 3.5)), np.full((2, 3, 2), 3.5))


@pytest.mark.parametrize("kind", [0, 1])
def test_gradient_matches_finite_differences(kind):
    H, W = 7, 10  # odd: every level has a partial block
    p = N.init_params(3).astype(np.float64)
    X = N.random_features(H, W, 4).astype(np.float64)
    X[..., 3:5] = np.abs(X[..., 3:5])
    t = np.random.default_rng(5).normal(0, 1, (H, W))
    loss, g, _ = N.loss_grad(p, X, t, bf16=False, kind=kind)
    rng = np.random.default_rng(6)
    offs = np.cumsum([0] + [n for _, n in N.SIZES])
    # a few parameters of every block
    idx = np.concatenate([rng.choice(np.arange(offs[k], offs[k + 1]), size=min(5, offs[k + 1] - offs[k]),
                                     replace=False) for k in range(len(N.SIZES))])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (N.loss_grad(pp, X, t, bf16=False, kind=kind)[0] - N.loss_grad(pm, X, t, bf16=False, kind=kind)[0]) / (
            2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_forward_moves_with_its_input_by_eight_cells():
    # pooling twice makes the network equivariant to shifts by 8 cells (level
    # 3 is 8 cells per pixel); away from the border the response to a point
    # input moves with it
    H, W = 176, 176
    p = N.init_params(1).astype(np.float64)
    X = np.zeros((H + 2, W + 2, 8))
    X[1:-1, 1:-1, 6] = 1.0
    X[1:-1, 1:-1, 5] = 1.0  # active everywhere
    X1, X2 = X.copy(), X.copy()
    X1[1 + 84, 1 + 84, :5] = [1, -2, 0.5, 3, -1]
    X2[1 + 92, 1 + 92, :5] = [1, -2, 0.5, 3, -1]
    d1, _ = N.forward(p, X1, bf16=False)
    d2, _ = N.forward(p, X2, bf16=False)
    np.testing.assert_allclose(d2[80:104, 80:104], d1[72:96, 72:96], rtol=1e-12, atol=1e-12)
    # and the point input reaches further than one 3x3 layer per level would:
    # cells 24 away along x still respond
    base, _ = N.forward(p, X, bf16=False)
    assert np.max(np.abs(d1[84, 100:112] - base[84, 100:112])) > 0.0
