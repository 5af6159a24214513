"""The MLP oracle pinned by closed forms and finite differences (CPU)."""
import numpy as np

from oracle import oracle as O

OFF = O.MLP_OFFSETS


def he_params(seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(O.MLP_NPARAM)
    for name, fan in (("W1", 100), ("W2", 128), ("W3", 128)):
        P[OFF[name]:OFF[name] + 16384] = rng.uniform(-1, 1, 16384) * np.sqrt(6 / fan)
    for b in ("b1", "b2", "b3"):
        P[OFF[b]:OFF[b] + 128] = rng.uniform(-0.1, 0.1, 128)
    P[OFF["w4"]:OFF["w4"] + 128] = rng.uniform(-1, 1, 128) * np.sqrt(3 / 128)
    P[OFF["b4"]] = 0.3
    return P


def test_bf16_round_to_nearest_even():
    assert O.bf16(1.0) == 1.0
    assert O.bf16(1.0 + 2.0 ** -8) == 1.0            # tie -> even mantissa
    assert O.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert O.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert O.bf16(0.3) == 0.30078125
    assert np.isinf(O.bf16(np.inf)) and np.isnan(O.bf16(np.nan))
    assert O.bf16(-2.5) == -2.5


def test_features_closed_form():
    f = O.mlp_freqs()
    assert f[0] == 1.0 and f[5] == 4.0 and f[10] == 16.0 and f[15] == 64.0
    assert np.allclose(f, 2.0 ** (0.4 * np.arange(16)), rtol=1e-7)
    phi = O.mlp_features([0, 0, 0], O.MLP_EXACT)
    sines = phi[4:100:2]
    cosines = phi[5:100:2]
    assert np.all(phi[:4] == 0) and np.all(sines == 0) and np.all(cosines == 1)
    assert np.all(phi[100:] == 0)
    phi = O.mlp_features([0.5, -0.25, 0.125], O.MLP_EXACT)
    assert phi[0] == 0.5 and phi[1] == -0.25 and phi[2] == 0.125 and phi[3] == 0
    for k in range(16):
        for a, ua in enumerate((0.5, -0.25, 0.125)):
            assert np.isclose(phi[4 + 6 * k + 2 * a], np.sin(np.pi * float(f[k]) * ua), atol=1e-12)
            assert np.isclose(phi[5 + 6 * k + 2 * a], np.cos(np.pi * float(f[k]) * ua), atol=1e-12)
    # BF16 mode: the same values rounded to bfloat16
    pb = O.mlp_features([0.5, -0.25, 0.125], O.MLP_BF16)
    assert np.all([O.bf16(v) == v for v in pb])
    assert np.max(np.abs(pb - phi)) < 2 ** -8


def test_forward_hand_set_weights():
    """One path u0 -> h1[0] -> h2[0] -> h3[0] -> y pins the out x in layout."""
    P = np.zeros(O.MLP_NPARAM)
    P[OFF["W1"] + 0 * 128 + 0] = 1.0      # W1[0][0]: feature 0 (u0) -> neuron 0
    P[OFF["W2"] + 0] = 1.0
    P[OFF["W3"] + 0] = 1.0
    P[OFF["w4"] + 0] = 1.0
    P[OFF["b4"]] = 0.25
    y = O.mlp_forward(P, [[0.3, 0.9, -0.4], [-0.3, 0.9, 0.4]], O.MLP_BF16)
    assert y[0] == 0.30078125 + 0.25 and y[1] == 0.25
    # a transposed W1 (in x out) would route u1 instead
    P[OFF["W1"]] = 0.0
    P[OFF["W1"] + 1] = 2.0                # W1[0][1]: feature 1 (u1) -> neuron 0
    y = O.mlp_forward(P, [[0.3, 0.5, -0.4]], O.MLP_EXACT)
    assert np.isclose(y[0], 1.25)
    # hidden biases pass through relu
    P = np.zeros(O.MLP_NPARAM)
    P[OFF["b3"] + 7] = 0.5
    P[OFF["w4"] + 7] = -2.0
    assert np.isclose(O.mlp_forward(P, [[0.1, 0.2, 0.3]], O.MLP_EXACT)[0], -1.0)


def test_gradient_finite_differences():
    P = he_params(0)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, (48, 3)).astype(np.float32)
    t = rng.uniform(0, 1, 48).astype(np.float32)
    loss, g = O.mlp_loss_grad(P, u, t, O.MLP_EXACT)
    idx = np.concatenate([rng.choice(O.MLP_NPARAM, 60, replace=False),
                          [OFF["b4"], OFF["w4"] + 3, OFF["b1"] + 5, OFF["b2"] + 9,
                           OFF["b3"] + 11, OFF["W1"] + 7 * 128 + 2, OFF["W3"] + 100 * 128 + 4]])
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (O.mlp_loss_grad(Pp, u, t, O.MLP_EXACT)[0] - O.mlp_loss_grad(Pm, u, t, O.MLP_EXACT)[0])
        fd /= 2 * eps
        assert abs(fd - g[i]) <= 1e-4 * abs(g[i]) + 1e-9, (i, fd, g[i])
    # unused feature columns (phi[100:] = 0) have exactly zero gradient
    for n in (0, 17, 127):
        assert np.all(g[OFF["W1"] + n * 128 + 100:OFF["W1"] + n * 128 + 128] == 0)


def test_bf16_gradient_close_to_exact():
    P = he_params(2)
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, (64, 3)).astype(np.float32)
    t = rng.uniform(0, 1, 64).astype(np.float32)
    _, ge = O.mlp_loss_grad(P, u, t, O.MLP_EXACT)
    _, gb = O.mlp_loss_grad(P, u, t, O.MLP_BF16)
    cos = ge @ gb / (np.linalg.norm(ge) * np.linalg.norm(gb))
    assert cos > 0.999


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=1000).astype(np.float32)
    g = (rng.normal(size=1000) * 10.0 ** rng.uniform(-3, 1, 1000)).astype(np.float32)
    g[:5] = 0.0
    lr = 1e-3
    p1, m, v = O.mlp_adam(p0, np.zeros(1000), np.zeros(1000), g, 1, lr)
    # m_hat = g, v_hat = g^2 after one step: p1 = p0 - lr * g / (|g| + 1e-8)
    expect = p0.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    # tolerance: FP32 rounding of parameters of magnitude <= 4
    assert np.allclose(p1, expect, rtol=0, atol=5e-7)
    assert np.all(p1[:5] == p0[:5])
    # (1 - 0.999) in FP32 is 0.001 * (1 - 1.3e-5)
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, 0.001 * g.astype(np.float64) ** 2, rtol=5e-5)
    # second step with the same gradient: m_hat = g, v_hat = g^2 again
    p2, _, _ = O.mlp_adam(p1, m, v, g, 2, lr)
    assert np.allclose(p2, p1 - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=5e-7)
