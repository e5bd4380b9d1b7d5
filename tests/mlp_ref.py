"""Float64 reference of the learned-radius network (test infrastructure).

Same layout and semantics as the library's kernels: parameters
W1[16][64], b1, W2[64][64], b2, W3[64][2], b3 (row-major [in][out]); every
product takes BF16-rounded operands (round to nearest even) when
``bf16=True``; loss = mean over masked rows of the squared error of both
outputs; gradients pass straight through the roundings.
"""
import numpy as np

SIZES = ((16, 64), (64, 64), (64, 2))


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = [], 0
    for fi, fo in SIZES:
        out.append(p[o:o + fi * fo].reshape(fi, fo))
        o += fi * fo
        out.append(p[o:o + fo])
        o += fo
    return out


def forward(p, X, bf16_ops=True):
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64))
    x = r(X)
    z1 = x @ r(W1) + b1
    a1 = r(np.maximum(z1, 0.0))
    z2 = a1 @ r(W2) + b2
    a2 = r(np.maximum(z2, 0.0))
    y = a2 @ r(W3) + b3
    return y, (x, z1, a1, z2, a2)


def loss_and_grad(p, X, T, mask, bf16_ops=True):
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64))
    y, (x, z1, a1, z2, a2) = forward(p, X, bf16_ops)
    m = (np.asarray(mask) > 0).astype(np.float64)[:, None]
    cnt = max(m.sum(), 1.0)
    e = (y - T) * m
    loss = (e ** 2).sum() / cnt
    dy = 2.0 * e / cnt
    dz2 = (dy @ r(W3).T) * (z2 > 0)
    dz1 = (dz2 @ r(W2).T) * (z1 > 0)
    g = [x.T @ dz1, dz1.sum(0), a1.T @ dz2, dz2.sum(0), a2.T @ dy, dy.sum(0)]
    return loss, np.concatenate([a.ravel() for a in g])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
