"""CPU tests of the radiance-cache oracle (oracle/nrc.py): its gradient against central finite
differences of the same loss (y held constant in the denominator), the encoding's closed forms,
Adam's first step, and bf16 rounding against a bit-level reference."""
import numpy as np

from oracle import nrc


def _params(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, nrc.N_PARAMS)
    Ws = nrc.split(p)
    Ws[0][:, 265:] = 0
    Ws[4][4:] = 0
    return np.concatenate([w.reshape(-1) for w in Ws])


def test_encoding_closed_forms():
    x = np.array([[0.0] * 11, [0.5] * 11, [0.25] * 11])
    e = nrc.encode(x)
    assert e.shape == (3, 272) and np.all(e[:, 264] == 1) and np.all(e[:, 265:] == 0)
    assert np.allclose(e[0, 0::2][:264 // 2], 0) and np.allclose(e[0, 1::2][:132], 1)
    # x = 1/2: band 0 -> sin(pi/2) = 1, cos = 0; band 1 -> sin(pi) = 0, cos = -1; higher bands: cos = 1.
    assert np.isclose(e[1, 0], 1) and np.isclose(e[1, 1], 0, atol=1e-12) and np.isclose(e[1, 3], -1)
    assert np.allclose(e[1, 5:24:2], 1)
    assert np.isclose(e[2, 0], np.sqrt(0.5)) and np.isclose(e[2, 2], 1)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1)
    x = rng.random((24, 11))
    t = rng.random((24, 4))
    y0, _ = nrc.forward(p, x)
    g = nrc.gradients(p, x, t)

    def fixed_loss(q):
        y, _ = nrc.forward(q, x)
        return np.mean((y - t) ** 2 / (y0 ** 2 + 0.01))

    offsets = np.cumsum([0] + nrc.SIZES)
    picks = []
    for layer in range(5):
        sub = rng.integers(offsets[layer], offsets[layer + 1], 40)
        picks += [int(i) for i in sub if abs(g[i]) > 1e-7][:6]
    assert len(picks) >= 20
    h = 1e-6
    for i in picks:
        d = np.zeros_like(p)
        d[i] = h
        fd = (fixed_loss(p + d) - fixed_loss(p - d)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-4 * abs(g[i]), (i, fd, g[i])
    # Padded inputs and unused outputs get no gradient.
    Gs = nrc.split(g)
    assert np.all(Gs[0][:, 265:] == 0) and np.all(Gs[4][4:] == 0)


def test_adam_first_step_moves_by_lr():
    p = _params(2)
    g = np.random.default_rng(3).normal(size=nrc.N_PARAMS)
    q, m, v = nrc.adam(p, np.zeros_like(p), np.zeros_like(p), g, 1, lr=1e-3)
    live = np.abs(g) > 1e-3
    assert np.allclose((p - q)[live], 1e-3 * np.sign(g[live]), rtol=1e-4)


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 1e-30, 65504.0], np.float64)
    got = nrc.to_bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0               # tie to even (mantissa ...0)
    assert got[2] == 1.0078125                            # above the tie rounds up
    assert got[3] == 1.0078125
    assert abs(got[4] + 3.140625) < 1e-12
    assert got[5] > 0
