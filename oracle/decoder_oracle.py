"""FP64 CPU oracle of the learned probe decoder (test infrastructure only).

The decoder has no counterpart in the reference: it is this project's own
network (a small space-time convolutional network that reads a trajectory's
measurement record and predicts the probe qubit's conditional Bloch vector).
This module restates the GPU kernels' arithmetic in numpy FP64, optionally
rounding to bfloat16 (round to nearest even) at exactly the points where the
GPU holds bfloat16 values, so that GPU gradients can be compared with it:

  input      record int8 [B, H, W] (-1 unmeasured, 0, 1) -> channels
             (record == 0, record == 1, x == W/2 (the probe's site),
             (y + 1) / H (time fraction, FP32 then bf16))
  conv l     3x3, zero padding in time (H), periodic in sites (W, a ring),
             CI -> 64 channels, + bias, ReLU; weights [9][CI][64] (tap =
             3 * (dy + 1) + (dx + 1)), activations stored as bf16
  pool       mean over the H*W positions
  fc1        64 -> 64 + bias, ReLU;  fc2 64 -> 3 + bias  (bf16 weights)
  loss       mean over the batch of h(S(|out|) - S(|y|)) + MU (|y| - out.y/|out|),
             h the smooth-L1 (Huber) function of width HUBER (d^2 / (2 HUBER)
             inside, |d| - HUBER/2 outside; d^2 when HUBER = 0),
             with |v| = sqrt(sum v^2 + 1e-12) and S(r) the qubit entropy
             (nats) at Bloch length r, continued linearly past R_MAX (so a
             length above 1 is pulled back instead of sitting in a flat
             region): the entropy and the
             direction are fitted separately, so a direction that the record
             leaves open does not shrink the predicted length (a squared
             error on the components would, biasing the entropy upwards)

Pinned by finite differences of its own loss and by torch autograd on the
same network (tests/test_decoder_oracle.py).
"""
from __future__ import annotations

import numpy as np

C = 64
MU = 1.0
R_MAX = float(np.float32(1.0 - 1e-6))  # the GPU's FP32 cap
HUBER = 0.05  # default smooth-L1 width of the entropy term (0: squared error)
NORM_EPS = 1e-12
CONV_CI = (4, 64, 64, 64)


def bf16(x, on=True):
    """Round to bfloat16 (nearest even), returned as float64."""
    if not on:
        return np.asarray(x, np.float64)
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def param_layout():
    """[(name, shape)] in the order of the flat parameter vector."""
    out = []
    for l, ci in enumerate(CONV_CI):
        out.append((f"conv{l}.w", (9, ci, C)))
        out.append((f"conv{l}.b", (C,)))
    out += [("fc1.w", (C, C)), ("fc1.b", (C,)), ("fc2.w", (C, 3)), ("fc2.b", (3,))]
    return out


def unflatten(flat):
    p, o = {}, 0
    for name, shape in param_layout():
        n = int(np.prod(shape))
        p[name] = np.asarray(flat[o:o + n], np.float64).reshape(shape)
        o += n
    return p


def flatten(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in param_layout()])


def param_count():
    return sum(int(np.prod(s)) for _, s in param_layout())


def he_init(seed):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_layout():
        if name.endswith(".b"):
            p[name] = np.zeros(shape)
        else:
            fan_in = shape[0] * shape[1] if len(shape) == 3 else shape[0]
            p[name] = rng.normal(0.0, np.sqrt(2.0 / fan_in), size=shape)
    return p


def input_channels(rec, emulate=True):
    rec = np.asarray(rec)
    B, H, W = rec.shape
    probe = np.zeros((B, H, W))
    probe[:, :, W // 2] = 1.0
    tfrac = np.broadcast_to((np.arange(1, H + 1, dtype=np.float32) / np.float32(H))[None, :, None], (B, H, W))
    return np.stack([(rec == 0), (rec == 1), probe, bf16(tfrac, emulate)], axis=-1).astype(np.float64)


def _shift(x, dy, dx):
    """s[b, y, x] = x[b, y + dy, (x + dx) mod W], zero where y + dy is outside."""
    H = x.shape[1]
    s = np.roll(x, -dx, axis=2)
    out = np.zeros_like(s)
    if dy == 0:
        out[:] = s
    elif dy > 0:
        out[:, :H - dy] = s[:, dy:]
    else:
        out[:, -dy:] = s[:, :H + dy]
    return out


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(x, w):
    """out[b,y,x,co] = sum_tap sum_ci w[tap,ci,co] x[b, y+dy, (x+dx)%W, ci]."""
    out = 0.0
    for t, (dy, dx) in enumerate(TAPS):
        out = out + _shift(x, dy, dx) @ w[t]
    return out


def conv_input_grad(dz, w):
    """Gradient of conv wrt its input: sum_tap shift(dz, -dy, -dx) @ w[tap]^T."""
    out = 0.0
    for t, (dy, dx) in enumerate(TAPS):
        out = out + _shift(dz, -dy, -dx) @ w[t].T
    return out


def conv_weight_grad(x, dz):
    B, H, W, ci = x.shape
    g = np.zeros((9, ci, dz.shape[-1]))
    for t, (dy, dx) in enumerate(TAPS):
        g[t] = _shift(x, dy, dx).reshape(-1, ci).T @ dz.reshape(-1, dz.shape[-1])
    return g


def _norm(v):
    return np.sqrt((v * v).sum(-1) + NORM_EPS)


def forward(p, rec, emulate=True):
    q = lambda v: bf16(v, emulate)
    acts = [input_channels(rec, emulate)]
    for l in range(4):
        z = conv(acts[-1], q(p[f"conv{l}.w"])) + p[f"conv{l}.b"]
        acts.append(q(np.maximum(z, 0.0)))
    pooled = acts[-1].mean(axis=(1, 2))
    h = np.maximum(pooled @ q(p["fc1.w"]) + p["fc1.b"], 0.0)
    out = h @ q(p["fc2.w"]) + p["fc2.b"]
    return out, {"acts": acts, "pooled": pooled, "h": h}


def _qubit_entropy(r):
    c = np.minimum(r, R_MAX)
    a, b = 0.5 * (1 + c), 0.5 * (1 - c)
    return -a * np.log(a) - b * np.log(b) - np.arctanh(R_MAX) * np.maximum(r - R_MAX, 0.0)


def _huber(d, delta):
    if delta <= 0:
        return d * d, 2.0 * d
    inside = np.abs(d) < delta
    return (np.where(inside, d * d / (2 * delta), np.abs(d) - delta / 2),
            np.where(inside, d / delta, np.sign(d)))


def loss_fn(out, y, huber=HUBER):
    no, ny = _norm(out), _norm(y)
    h, _ = _huber(_qubit_entropy(no) - _qubit_entropy(ny), huber)
    return float(np.mean(h + MU * (ny - (out * y).sum(-1) / no)))


def gradients(p, rec, y, emulate=True, huber=HUBER):
    """(loss, grads dict) with the GPU's rounding points when emulate=True."""
    q = lambda v: bf16(v, emulate)
    out, cache = forward(p, rec, emulate)
    B = out.shape[0]
    y = np.asarray(y, np.float64)
    loss = loss_fn(out, y, huber)
    no, ny = _norm(out), _norm(y)
    dot = (out * y).sum(-1)
    dsdr = -np.arctanh(np.minimum(no, R_MAX))
    _, dh_dd = _huber(_qubit_entropy(no) - _qubit_entropy(ny), huber)
    gs = dh_dd * dsdr
    dout = (1.0 / B) * ((gs / no)[:, None] * out - MU * (y / no[:, None] - (dot / no ** 3)[:, None] * out))
    g = {}
    h, pooled, acts = cache["h"], cache["pooled"], cache["acts"]
    g["fc2.w"] = h.T @ dout
    g["fc2.b"] = dout.sum(0)
    dh = (dout @ q(p["fc2.w"]).T) * (h > 0)
    g["fc1.w"] = pooled.T @ dh
    g["fc1.b"] = dh.sum(0)
    dpooled = dh @ q(p["fc1.w"]).T
    H, W = acts[-1].shape[1:3]
    da = np.broadcast_to((dpooled / (H * W))[:, None, None, :], acts[-1].shape).copy()
    for l in range(3, -1, -1):
        dz = da * (acts[l + 1] > 0)
        dzq = q(dz)
        g[f"conv{l}.w"] = conv_weight_grad(acts[l], dzq)
        g[f"conv{l}.b"] = dz.sum(axis=(0, 1, 2))
        if l > 0:
            da = conv_input_grad(dzq, q(p[f"conv{l}.w"]))
    return loss, g


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t counts from 1); returns new (p, m, v) dicts."""
    np_, nm, nv = {}, {}, {}
    for k in p:
        nm[k] = b1 * m[k] + (1 - b1) * g[k]
        nv[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = nm[k] / (1 - b1 ** t)
        vh = nv[k] / (1 - b2 ** t)
        np_[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return np_, nm, nv


def entropy_from_bloch(r):
    """Von Neumann entropy (nats) of the qubit with Bloch vector r (|r| clipped to 1)."""
    n = np.minimum(np.sqrt((np.asarray(r) ** 2).sum(-1)), 1.0)
    lp, lm = 0.5 * (1 + n), 0.5 * (1 - n)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = -np.where(lp > 0, lp * np.log(lp), 0.0) - np.where(lm > 0, lm * np.log(lm), 0.0)
    return s
