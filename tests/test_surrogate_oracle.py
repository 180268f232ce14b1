"""The surrogate MLP oracle (oracle/surrogate_oracle.py), pinned on the CPU:
GELU against the exact Gaussian form and its own identities, the analytic
gradient against central finite differences, and BF16 rounding against known
answers."""
import math

import numpy as np

import surrogate_oracle as S


def test_gelu_tanh_form_tracks_exact_gelu():
    # tanh form vs x * Phi(x) (erf), a different closed form: within 1e-3
    xs = np.linspace(-6.0, 6.0, 241)
    exact = np.array([0.5 * x * (1.0 + math.erf(x / math.sqrt(2.0))) for x in xs])
    assert np.max(np.abs(S.gelu(xs) - exact)) < 1e-3
    # gelu(x) - gelu(-x) = x exactly, gelu(0) = 0
    assert np.allclose(S.gelu(xs) - S.gelu(-xs), xs, atol=1e-14)
    assert S.gelu(np.array(0.0)) == 0.0


def test_gelu_grad_matches_central_differences():
    xs = np.linspace(-5.0, 5.0, 101)
    h = 1e-6
    fd = (S.gelu(xs + h) - S.gelu(xs - h)) / (2 * h)
    assert np.max(np.abs(S.gelu_grad(xs) - fd)) < 1e-8


def _norm(rng):
    return dict(x_mean=rng.normal(size=7), x_std=rng.uniform(0.5, 2.0, 7),
                y_mean=rng.normal(size=4), y_std=rng.uniform(0.5, 2.0, 4))


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = S.init_xavier(1) + rng.normal(scale=0.05, size=S.N_PARAMS)
    x = rng.normal(size=(12, 7))
    y = rng.normal(size=(12, 4))
    norm = _norm(rng)
    _, g = S.loss_grad(p, x, y, norm)
    assert g.shape == (S.N_PARAMS,)
    # every bias, and a spread of weights of every layer
    idx = list(range(448, 512)) + list(range(4608, 4672)) + list(range(8768, 8832)) + \
        list(range(9088, 9092)) + list(rng.choice(S.N_PARAMS, 200, replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros(S.N_PARAMS)
        e[i] = h
        fd = (S.loss_grad(p + e, x, y, norm)[0] - S.loss_grad(p - e, x, y, norm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_parameter_layout():
    assert S.N_PARAMS == 9092
    p = np.arange(S.N_PARAMS, dtype=np.float64)
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = S.unpack(p)
    assert W1.shape == (7, 64) and W1[0, 1] == 1 and W1[1, 0] == 64  # W[in][out]
    assert b1[0] == 448 and W2[0, 0] == 512 and b2[0] == 4608 and W3[0, 0] == 4672
    assert b3[0] == 8768 and W4.shape == (64, 4) and W4[0, 0] == 8832 and b4[0] == 9088


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        1.0: 1.0,
        one + 2.0 ** -8: 1.0,                      # halfway, tie to even (down)
        one + 3 * 2.0 ** -8: one + 2 * 2.0 ** -7,  # halfway, tie to even (up)
        one + 2.0 ** -8 + 2.0 ** -12: one + 2.0 ** -7,
        3.14159: 3.140625,
        -2.5: -2.5,
        1e-40: 2.0 ** -133,  # FP32 subnormal 0x000116c2 -> BF16 0x0001
    }
    for a, b in cases.items():
        assert S.bf16(np.array([a]))[0] == b, (a, S.bf16(np.array([a]))[0], b)


def test_bf16_forward_is_close_to_fp64_forward():
    rng = np.random.default_rng(3)
    p = S.init_xavier(2)
    norm = dict(_norm(rng), box_lo=np.zeros(7), box_hi=np.ones(7))
    x = rng.normal(size=(64, 7)) * norm["x_std"] + norm["x_mean"]
    y64 = S.forward(p, S.normalise_x(x, norm))[2] * norm["y_std"] + norm["y_mean"]
    yb = S.forward_bf16(p, x, norm)
    assert np.linalg.norm(yb - y64) / np.linalg.norm(y64) < 3e-2


def test_constrain_clips_condensate_and_closes_with_vapour():
    # known answers, one cell each: (q, raw change) -> applied change
    q = np.array([[10e-3, 1e-3, 0.5e-3],    # condensation and autoconversion
                  [10e-3, 0.0, 0.0],        # no condensate: negative changes clip to 0
                  [10e-3, 1e-3, 0.5e-3],    # cloud asked to lose more than it holds
                  [0.2e-3, 0.0, 0.0]])      # condensation larger than the vapour
    d = np.array([[-0.4e-3, 0.3e-3, 0.1e-3],
                  [0.3e-3, -0.2e-3, -0.1e-3],
                  [0.0, -3e-3, 0.2e-3],
                  [-1e-3, 0.8e-3, 0.0]])
    dth = np.array([1.0, 0.5, -2.0, 3.0])
    qn, th = S.constrain(q, d, dth)
    want = np.array([[9.6e-3, 1.3e-3, 0.6e-3],
                     [10e-3, 0.0, 0.0],
                     [10.8e-3, 0.0, 0.7e-3],
                     [0.0, 0.2e-3, 0.0]])
    assert np.allclose(qn, want, rtol=0, atol=1e-15)
    # dtheta passes unless the vapour is exhausted (lambda = 0.2 / 0.8)
    assert np.allclose(th, [1.0, 0.5, -2.0, 0.75], rtol=0, atol=1e-15)
    # random cells: total water conserved, nothing negative, the applied
    # cloud and rain changes are the raw ones wherever no clip or scaling acts
    rng = np.random.default_rng(11)
    n = 500
    q = np.abs(rng.normal(size=(n, 3))) * 1e-3
    q[rng.random((n, 3)) < 0.3] = 0.0
    d = rng.normal(size=(n, 3)) * 1e-3
    qn, th = S.constrain(q, d, rng.normal(size=n))
    assert np.allclose(qn.sum(axis=1), q.sum(axis=1), rtol=0, atol=1e-15)
    assert np.all(qn >= 0.0)
    free = (d[:, 1] >= -q[:, 1]) & (d[:, 2] >= -q[:, 2]) & (q[:, 0] - d[:, 1] - d[:, 2] >= 0)
    assert free.any() and (~free).any()
    assert np.allclose(qn[free, 1:] - q[free, 1:], d[free, 1:], rtol=0, atol=1e-15)
