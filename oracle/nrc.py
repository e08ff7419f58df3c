"""Oracle (test infrastructure only): the learned radiance cache's network in float64 numpy -- the
frequency encoding, the 4 x 64 ReLU MLP, the relative-L2 loss, its gradients and one Adam step --
laid out exactly as the library stores it, so the GPU forward/backward kernels can be checked
against it.  The reference renderer has no learned cache; this follows the product's own
definition (metal-vibe-tracer_amd/csrc/nrc.h):

  inputs   11 numbers in [0, 1] per query: position (scene box), direction (d + 1) / 2,
           normal (n + 1) / 2, roughness, wavelength number
  encoding feature 24 i + 2 b + {0: sin, 1: cos} of (2^b pi x_i), b = 0..11, then a constant 1
           (feature 264) and zeros up to 272
  network  y = W5 relu(W4 relu(W3 relu(W2 relu(W1 e)))), W1 64 x 272, W2-W4 64 x 64, W5 16 x 64
           (rows 4..15 unused); parameters stored row-major in that order (30720 numbers)
  loss     mean over queries and the 4 outputs of (y - t)^2 / (y^2 + 0.01), y held constant in
           the denominator (the relative L2 of neural radiance caching)
  Adam     beta1 0.9, beta2 0.99, eps 1e-8, bias-corrected
  EMA      inference weights: ema = 0.95 ema + 0.05 params after each Adam step, ema = params when
           the parameters are set
"""
import numpy as np

N_IN, BANDS, N_FEAT, HIDDEN, N_OUT_PAD, N_OUT = 11, 12, 272, 64, 16, 4
SHAPES = [(HIDDEN, N_FEAT), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (N_OUT_PAD, HIDDEN)]
SIZES = [a * b for a, b in SHAPES]
N_PARAMS = sum(SIZES)


def split(params):
    out, o = [], 0
    for (a, b), n in zip(SHAPES, SIZES):
        out.append(np.asarray(params[o:o + n], np.float64).reshape(a, b))
        o += n
    return out


def encode(x):
    """x: (n, 11) -> (n, 272) features."""
    x = np.asarray(x, np.float64)
    n = len(x)
    e = np.zeros((n, N_FEAT))
    for i in range(N_IN):
        for b in range(BANDS):
            arg = np.pi * (2.0 ** b) * x[:, i]
            e[:, 24 * i + 2 * b] = np.sin(arg)
            e[:, 24 * i + 2 * b + 1] = np.cos(arg)
    e[:, 264] = 1.0
    return e


def forward(params, x, bf16=False):
    """Returns (y (n, 4), activations [e, h1, h2, h3, h4]).  bf16 = True rounds weights and every
    layer input to bfloat16 as the MFMA kernels consume them."""
    Ws = split(params)
    r = to_bf16 if bf16 else (lambda a: a)
    h = r(encode(x))
    acts = [h]
    for W in Ws[:4]:
        h = r(np.maximum(h @ r(W).T, 0.0))
        acts.append(h)
    y = h @ r(Ws[4]).T
    return y[:, :N_OUT], acts


def loss(y, t):
    return float(np.mean((y - t) ** 2 / (y ** 2 + 0.01)))


def gradients(params, x, t, bf16=False):
    """d loss / d params in the storage layout, with y held constant in the loss denominator."""
    Ws = split(params)
    r = to_bf16 if bf16 else (lambda a: a)
    y, acts = forward(params, x, bf16)
    n = len(x)
    dy = np.zeros((n, N_OUT_PAD))
    dy[:, :N_OUT] = 2.0 * (y - t) / (y ** 2 + 0.01) / (n * N_OUT)
    dy = r(dy)
    grads = [None] * 5
    grads[4] = dy.T @ acts[4]
    d = dy
    for layer in range(3, -1, -1):
        d = d @ r(Ws[layer + 1])
        d = r(d * (acts[layer + 1] > 0))
        grads[layer] = d.T @ acts[layer]
    return np.concatenate([g.reshape(-1) for g in grads])


def adam(params, m, v, grads, step, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


EMA_DECAY = 0.95


def ema(prev, params, decay=EMA_DECAY):
    """One update of the inference weights' moving average."""
    return decay * np.asarray(prev, np.float64) + (1 - decay) * np.asarray(params, np.float64)


def to_bf16(a):
    """Round float64 -> float32 -> bfloat16 (nearest, ties to even) -> back."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
