"""float64 numpy restatement of the learned proposal -- TEST INFRASTRUCTURE ONLY.

Only tests/ load it, as the checker of the library's matrix-core forward and backward
kernels (monki-gpu_amd/csrc/proposal.h, proposal.hip). The network has no reference
counterpart; this file restates its definition:

* parameters: per layer W (row-major, out x in) then b; layers 12 -> 128 -> 128 -> 128 -> 128 -> 40,
  SiLU after the four hidden layers;
* outputs per component k (8): o[5k:5k+3] an unnormalized mean direction, kappa =
  clip(softplus(o[5k+3]), 0.05, 2000), mixture weights softmax(o[5k+4]);
* density q(w) = sum_k pi_k kappa_k / (2 pi (1 - exp(-2 kappa_k))) exp(kappa_k (mu_k . w - 1));
* loss = -sum_i g_i log q(w_i) / sum_i g_i.

``forward(..., bf16=True)`` mimics the device arithmetic: weights, inputs and hidden activations
rounded to bfloat16 (nearest even), float accumulation (here float64). The analytic gradient is
pinned by central finite differences (tests/test_golden.py).
"""

import numpy as np

INPUTS, HIDDEN, COMPONENTS = 12, 128, 8
OUTPUTS = 5 * COMPONENTS
SHAPES = [(HIDDEN, INPUTS), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (OUTPUTS, HIDDEN)]
KAPPA_MIN, KAPPA_MAX = 0.05, 2000.0
PARAMETERS = sum(r * c + r for r, c in SHAPES)


def unpack(params):
    params = np.asarray(params, dtype=np.float64)
    layers, offset = [], 0

    for rows, cols in SHAPES:
        w = params[offset : offset + rows * cols].reshape(rows, cols)
        offset += rows * cols
        b = params[offset : offset + rows]
        offset += rows
        layers.append((w, b))

    assert offset == params.size == PARAMETERS
    return layers


def bf16(x):
    """Round float32 values to bfloat16, nearest even (as the device's bf16_bits)."""

    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(params, features, bf16_compute=False, keep=False):
    """Network outputs [n, 40]; with keep, also the pre-activations and activations."""

    h = np.asarray(features, dtype=np.float64)
    h = bf16(h) if bf16_compute else h
    zs, hs = [], [h]

    for index, (w, b) in enumerate(unpack(params)):
        w = bf16(w) if bf16_compute else w
        z = h @ w.T + b

        if index == len(SHAPES) - 1:
            return (z, zs, hs) if keep else z

        zs.append(z)
        h = silu(z)
        h = bf16(h) if bf16_compute else h
        hs.append(h)


def softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def mixture(outputs):
    """(mu [n,8,3], kappa [n,8], log pi [n,8], norm [n,8]) of network outputs."""

    o = np.asarray(outputs, dtype=np.float64).reshape(-1, COMPONENTS, 5)
    raw = o[..., :3]
    norm = np.linalg.norm(raw, axis=-1)
    valid = norm > 1e-6
    mu = np.where(valid[..., None], raw / np.where(valid, norm, 1)[..., None], [0.0, 0.0, 1.0])
    kappa = np.clip(softplus(o[..., 3]), KAPPA_MIN, KAPPA_MAX)
    logits = o[..., 4]
    log_pi = logits - np.logaddexp.reduce(logits, axis=-1, keepdims=True)
    return mu, kappa, log_pi, norm


def log_vmf(kappa, dot):
    return np.log(kappa) - np.log(2 * np.pi) - np.log1p(-np.exp(-2 * kappa)) + kappa * (dot - 1)


def log_density(outputs, directions):
    mu, kappa, log_pi, _ = mixture(outputs)
    dot = np.einsum("nkc,nc->nk", mu, np.asarray(directions, dtype=np.float64))
    return np.logaddexp.reduce(log_pi + log_vmf(kappa, dot), axis=-1)


def head(outputs, directions, g):
    """Per-sample loss -g log q and its gradient d loss / d outputs [n, 40]."""

    o = np.asarray(outputs, dtype=np.float64).reshape(-1, COMPONENTS, 5)
    w = np.asarray(directions, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    mu, kappa, log_pi, norm = mixture(outputs)
    dot = np.einsum("nkc,nc->nk", mu, w)
    ell = log_pi + log_vmf(kappa, dot)
    log_q = np.logaddexp.reduce(ell, axis=-1)
    r = np.exp(ell - log_q[:, None])
    pi = np.exp(log_pi)
    dell = -g[:, None] * r
    d = np.zeros_like(o)
    d[..., 4] = -g[:, None] * (r - pi)
    soft = softplus(o[..., 3])
    dlog_kappa = 1 / kappa - 2 / np.expm1(2 * kappa) + dot - 1
    free = (soft > KAPPA_MIN) & (soft < KAPPA_MAX)
    d[..., 3] = np.where(free, dell * dlog_kappa / (1 + np.exp(-o[..., 3])), 0.0)
    gvec = dell[..., None] * kappa[..., None] * w[:, None, :]
    along = np.sum(gvec * mu, axis=-1, keepdims=True)
    valid = (norm > 1e-6)[..., None]
    d[..., :3] = np.where(valid, (gvec - along * mu) / np.where(valid, norm[..., None], 1), 0.0)
    return -g * log_q, d.reshape(-1, OUTPUTS)


def loss_and_gradient(params, features, directions, weights):
    """float64 loss -sum g log q / sum g and its gradient in parameter order."""

    weights = np.asarray(weights, dtype=np.float64)
    scale = 1.0 / weights.sum()
    outputs, zs, hs = forward(params, features, keep=True)
    losses, dy = head(outputs, directions, weights * scale)
    grads = []
    delta = dy
    layers = unpack(params)

    for index in range(len(SHAPES) - 1, -1, -1):
        w, _ = layers[index]
        grads.append((delta.T @ hs[index], delta.sum(axis=0)))

        if index:
            z = zs[index - 1]
            s = 1 / (1 + np.exp(-z))
            delta = (delta @ w) * s * (1 + z * (1 - s))

    flat = []

    for dw, db in reversed(grads):
        flat.extend([dw.ravel(), db])

    return losses.sum(), np.concatenate(flat)


def initial_parameters(seed=0):
    rng = np.random.default_rng(seed)
    parts = []

    for index, (rows, cols) in enumerate(SHAPES):
        scale = (0.1 if index == len(SHAPES) - 1 else 1.0) * np.sqrt(2.0 / cols)
        parts.append(rng.normal(0, scale, rows * cols))
        bias = np.zeros(rows)

        if index == len(SHAPES) - 1:
            bias[3::5] = 1.0

        parts.append(bias)

    return np.concatenate(parts)


def sample_vmf(mean, kappa, rng):
    """Directions of vMF(mean, kappa) (Wood 1994), rows of mean unit vectors."""

    mean = np.asarray(mean, dtype=np.float64)
    n = len(mean)
    u = rng.random(n)
    w = 1 + np.log(u + (1 - u) * np.exp(-2 * kappa)) / kappa
    phi = 2 * np.pi * rng.random(n)
    helper = np.where(np.abs(mean[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]])
    e1 = np.cross(mean, helper)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(mean, e1)
    s = np.sqrt(np.maximum(0, 1 - w**2))
    return w[:, None] * mean + s[:, None] * (np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2)
