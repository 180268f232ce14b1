"""CPU oracle of the warm-rain surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py as the checker of
the device surrogate (gpu-wm-next_amd/csrc/gwm_surrogate.hip); the product
never calls it.  The surrogate has no counterpart in the reference
(FahrenheitResearch/gpu-wm-next); the network is specified by its layer
sizes 7 -> 64 -> 64 -> 64 -> 4 with GELU, the flat parameter layout
W1 b1 W2 b2 W3 b3 W4 b4 (W[in][out]) and the loss (mean squared error of
the normalised outputs).  Everything here is FP64 numpy.

Pinned by (tests/test_surrogate_oracle.py): central finite differences of
the loss for the gradient, the GELU closed form and its derivative at known
points, BF16 round-to-nearest-even known answers, and for the conservation
layer known answers (clipping, vapour closure, the vapour-exhaustion
scaling) plus exact conservation and non-negativity on random inputs.
"""
import numpy as np

N_IN, N_OUT, HIDDEN = 7, 4, 64
SIZES = [(N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)]
N_PARAMS = sum(a * b + b for a, b in SIZES)  # 9092

GELU_A = np.sqrt(2.0 / np.pi)
GELU_B = 0.044715


def gelu(x):
    """GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(GELU_A * (x + GELU_B * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_A * (x + GELU_B * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_A * (1.0 + 3.0 * GELU_B * x * x)


def unpack(p):
    """Flat parameters -> [(W, b)] with W[in][out]."""
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for a, b in SIZES:
        W = p[o:o + a * b].reshape(a, b)
        o += a * b
        bias = p[o:o + b]
        o += b
        out.append((W, bias))
    assert o == N_PARAMS
    return out


def normalise_x(x, norm, fp32=False):
    """(x - mean) / std; fp32=True rounds like the device, which normalises
    in FP32 before the FP32 layers."""
    if fp32:
        return ((np.asarray(x, np.float32) - np.asarray(norm["x_mean"], np.float32)) /
                np.asarray(norm["x_std"], np.float32)).astype(np.float64)
    return (np.asarray(x, np.float64) - norm["x_mean"]) / norm["x_std"]


def normalise_y(y, norm, fp32=False):
    if fp32:
        return ((np.asarray(y, np.float32) - np.asarray(norm["y_mean"], np.float32)) /
                np.asarray(norm["y_std"], np.float32)).astype(np.float64)
    return (np.asarray(y, np.float64) - norm["y_mean"]) / norm["y_std"]


def forward(p, xn):
    """Returns pre-activations zs, activations acts (acts[0] = xn), output."""
    layers = unpack(p)
    acts, zs = [np.asarray(xn, np.float64)], []
    h = acts[0]
    for li, (W, b) in enumerate(layers):
        z = h @ W + b
        zs.append(z)
        h = gelu(z) if li < 3 else z
        acts.append(h)
    return zs, acts, h


def loss_grad(p, x_raw, y_raw, norm, fp32_normalise=False):
    """Mean squared error of the normalised outputs and its gradient
    (fp32_normalise: features and targets normalised in FP32, as on the
    device; the network itself stays FP64)."""
    xn = normalise_x(x_raw, norm, fp32_normalise)
    yt = normalise_y(y_raw, norm, fp32_normalise)
    zs, acts, y = forward(p, xn)
    B = xn.shape[0]
    d = y - yt
    loss = float(np.mean(d * d))
    dz = 2.0 * d / d.size
    layers = unpack(p)
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        grads[li] = (acts[li].T @ dz, dz.sum(axis=0))
        if li > 0:
            dz = (dz @ W.T) * gelu_grad(zs[li - 1])
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, x_raw, norm):
    """The inference arithmetic of the matrix-core kernel: FP32 normalised
    features, BF16 weights and activations, exact products accumulated (the
    device accumulates in FP32), biases in FP32; raw outputs."""
    xn = (np.asarray(x_raw, np.float32) - np.asarray(norm["x_mean"], np.float32)) / \
        np.asarray(norm["x_std"], np.float32)
    h = bf16(xn)
    for li, (W, b) in enumerate(unpack(p)):
        z = h @ bf16(W) + np.asarray(b, np.float32).astype(np.float64)
        h = bf16(gelu(z)) if li < 3 else z
    return h * norm["y_std"] + norm["y_mean"]


def constrain(q, d, dth):
    """The surrogate step's conservation layer (gwm_surrogate.hip,
    k_warm_rain_sur): the changes of cloud and rain are clipped at minus
    what they hold, vapour takes the opposite of their sum (the network's
    vapour change is not applied), and where that would exhaust the vapour
    the exchange and dtheta are scaled by lambda = qv / -dqv.
    q, d: [n][3] mixing ratios (qv, qc, qr) and raw changes; returns
    (new q, applied dtheta)."""
    q = np.asarray(q, np.float64)
    d = np.asarray(d, np.float64)
    dc = np.maximum(d[:, 1], -q[:, 1])
    dr = np.maximum(d[:, 2], -q[:, 2])
    dv = -(dc + dr)
    with np.errstate(divide="ignore", invalid="ignore"):
        lam = np.where(q[:, 0] + dv < 0.0, np.maximum(q[:, 0] / -dv, 0.0), 1.0)
    ch = np.stack([dv, dc, dr], axis=1) * lam[:, None]
    return np.maximum(q + ch, 0.0), lam * np.asarray(dth, np.float64)


def init_xavier(seed):
    """Not the device initialiser; a host start point for oracle-only tests."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in SIZES:
        lim = np.sqrt(6.0 / (a + b))
        parts += [rng.uniform(-lim, lim, a * b), np.zeros(b)]
    return np.concatenate(parts)
