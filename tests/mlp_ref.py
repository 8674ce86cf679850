"""CPU oracle of the warm-start MLP (test infrastructure): forward, loss,
backward and Adam in float64 numpy, used to check the GPU training and
inference kernels.  Its own gradient is pinned by central finite differences
(tests/test_mlp_cpu.py).

Network (the product's definition, test-regression_amd/csrc/mlp.hip):
  features  f_i = s_i / m, i < 10, m = max(max_i s_i, 1)   (samples already
            pedestal-free); padded with zeros to 16
  layer 1   z1 = f W1 + b1, W1: 10 x 64; h1 = gelu(z1)
  layer 2   z2 = h1 W2 + b2, W2: 64 x 64; h2 = gelu(z2)
  layer 3   z3 = h2 W3 + b3, W3: 64 x 20; logits = z3[:10],
            a = softplus(z3[10:]) (predicted amplitude / m)
  gelu      0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
  loss      mean over the batch of
            mean_i BCE(logits_i, [x_ref_i > 0]) + mean_i (a_i - x_ref_i / m)^2
  params    flat float32 [W1, b1, W2, b2, W3, b3], row-major [in][out]: 6164
  weights   the forward pass uses the parameters rounded to bfloat16
            (round to nearest even) and accumulates in float32
"""
from __future__ import annotations

import numpy as np

SIZES = [(10, 64), (64,), (64, 64), (64,), (64, 20), (20,)]
NPARAM = sum(int(np.prod(s)) for s in SIZES)
C_GELU = np.sqrt(2.0 / np.pi)


def unflatten(p):
    out, k = [], 0
    for s in SIZES:
        n = int(np.prod(s))
        out.append(p[k:k + n].reshape(s))
        k += n
    return out


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * x * x)


def softplus(x):
    return np.logaddexp(0.0, x)


def features(samples):
    """samples [10, B] -> (X [B, 10], m [B])."""
    s = np.asarray(samples, dtype=np.float64)
    m = np.maximum(s.max(axis=0), 1.0)
    return (s / m).T, m


def forward(p, X, quantize=True):
    W1, b1, W2, b2, W3, b3 = unflatten(bf16_round(p) if quantize else p)
    z1 = X @ W1 + b1
    h1 = gelu(z1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    z3 = h2 @ W3 + b3
    return dict(X=X, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3)


def loss_and_grad(p, samples, xref, quantize=True):
    """Mean loss over the batch and its gradient w.r.t. the flat parameters
    (the gradient of the bf16-rounded weights, applied to the masters)."""
    X, m = features(samples)
    B = X.shape[0]
    f = forward(p, X, quantize)
    W1, b1, W2, b2, W3, b3 = unflatten(bf16_round(p) if quantize else p)
    y = (np.asarray(xref).T > 0).astype(np.float64)        # [B, 10]
    t = np.asarray(xref).T / m[:, None]                     # [B, 10]
    lg, ar = f["z3"][:, :10], f["z3"][:, 10:]
    a = softplus(ar)
    bce = np.maximum(lg, 0) - lg * y + np.log1p(np.exp(-np.abs(lg)))
    loss = (bce.mean(axis=1) + ((a - t) ** 2).mean(axis=1)).mean()
    sig = 1.0 / (1.0 + np.exp(-lg))
    dz3 = np.concatenate([(sig - y) / 10.0, 2.0 * (a - t) / 10.0 * (1.0 / (1.0 + np.exp(-ar)))],
                         axis=1) / B
    dW3 = f["h2"].T @ dz3
    db3 = dz3.sum(axis=0)
    dz2 = (dz3 @ W3.T) * gelu_grad(f["z2"])
    dW2 = f["h1"].T @ dz2
    db2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2.T) * gelu_grad(f["z1"])
    dW1 = X.T @ dz1
    db1 = dz1.sum(axis=0)
    g = np.concatenate([a.ravel() for a in (dW1, db1, dW2, db2, dW3, db3)])
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """Xavier-uniform weights, zero biases (the product's mf_mlp_init)."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SIZES:
        if len(s) == 2:
            lim = np.sqrt(6.0 / (s[0] + s[1]))
            parts.append(rng.uniform(-lim, lim, s).ravel())
        else:
            parts.append(np.zeros(s))
    return np.concatenate(parts).astype(np.float32)


def predict(p, samples):
    """(passive mask [B, 10] bool, amplitudes [B, 10]) as the warm start uses them."""
    X, m = features(samples)
    f = forward(p, X)
    lg, ar = f["z3"][:, :10], f["z3"][:, 10:]
    return lg > 0, softplus(ar) * m[:, None]
