"""The float64 MLP oracle pinned by central finite differences of its loss,
the Adam step by its closed form on one coordinate, and bf16 rounding by
hand-checked values."""
import numpy as np

from oracle import mlp as om


def _setup(seed=0, n=16):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.2, om.NPARAM)
    x = rng.normal(0, 1, (n, om.NIN))
    t = rng.normal(0, 1, (n, om.NOUT))
    return p, x, t


def test_gradient_matches_central_differences():
    p, x, t = _setup()
    _, g = om.loss_grad(p, x, t)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(om.NPARAM, 40, replace=False),
                          [0, om.NPARAM - 1, om.NPARAM - 3]])  # incl. W1[0,0], b3
    h = 1e-6
    for i in idx:
        e = np.zeros(om.NPARAM)
        e[i] = h
        fd = (om.loss(p + e, x, t) - om.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_and_its_derivative():
    z = np.linspace(-6, 6, 101)
    # limits: gelu(z) -> z for z >> 0, -> 0 for z << 0, gelu(0) = 0
    assert om.gelu(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(om.gelu(np.array([8.0])), 8.0, rtol=1e-12)
    assert abs(om.gelu(np.array([-8.0]))[0]) < 1e-12
    h = 1e-6
    np.testing.assert_allclose(om.gelu_grad(z), (om.gelu(z + h) - om.gelu(z - h)) / (2 * h),
                               rtol=1e-6, atol=1e-9)


def test_layout_round_trip_and_size():
    p, _, _ = _setup()
    assert om.NPARAM == 128 * 64 + 128 + 128 * 128 + 128 + 2 * 128 + 2
    np.testing.assert_array_equal(om.pack(om.unpack(p)), p)


def test_adam_first_step_moves_by_lr_times_sign():
    # bias-corrected first step: m^ = g, v^ = g^2 -> step = lr * g / (|g| + eps)
    p = np.zeros(3)
    g = np.array([0.5, -2.0, 1e-3])
    p1, m, v = om.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p1, -0.01 * np.sign(g), rtol=1e-4)


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -3.140625, 1e-3])
    got = om.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 -> 1 + 2^-6
    np.testing.assert_array_equal(got[:4], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6])
    assert got[4] == -3.140625
    assert abs(got[5] - 1e-3) / 1e-3 < 2 ** -8
