"""Pins the decoder oracle (oracle/decoder_oracle.py) on the CPU.

- bf16 rounding: round-to-nearest-even ties, and torch's float32 -> bfloat16
  conversion on random values.
- gradients (no rounding): central finite differences of the oracle's own
  loss, and torch autograd of the same network in float64.
- Adam: torch.optim.Adam on the same gradients.
torch is imported inside the tests only (CPU), never in a GPU test process.
"""
import numpy as np
import pytest

from oracle import decoder_oracle as D


def _case(B=3, H=5, W=4, seed=0):
    rng = np.random.default_rng(seed)
    rec = rng.integers(-1, 2, size=(B, H, W)).astype(np.int8)
    y = rng.normal(size=(B, 3)) * 0.5
    p = D.he_init(seed + 1)
    for k in p:  # non-zero biases so their gradients are exercised
        if k.endswith(".b"):
            p[k] = rng.normal(0.0, 0.1, size=p[k].shape)
    return p, rec, y


def test_bf16_rounding_ties_and_torch():
    one = 1.0
    assert D.bf16(one) == 1.0
    assert D.bf16(1 + 2.0 ** -8) == 1.0                    # tie -> even (1.0)
    assert D.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6      # tie -> even (1 + 2^-6)
    assert D.bf16(1 + 2.0 ** -8 + 2.0 ** -20) == 1 + 2.0 ** -7
    assert D.bf16(-3.0) == -3.0
    import torch
    x = np.random.default_rng(3).normal(size=4096).astype(np.float32) * 10.0 ** np.random.default_rng(4).integers(-8, 8, 4096)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(D.bf16(x), want)


@pytest.mark.parametrize("huber", [0.0, 0.05])
def test_gradients_match_finite_differences(huber):
    p, rec, y = _case()
    loss, g = D.gradients(p, rec, y, emulate=False, huber=huber)
    flat = D.flatten(p)
    gflat = D.flatten(g)
    rng = np.random.default_rng(9)
    idx = rng.choice(len(flat), size=40, replace=False)
    # biases of every layer too
    o = 0
    for name, shape in D.param_layout():
        n = int(np.prod(shape))
        if name.endswith(".b"):
            idx = np.append(idx, o)
        o += n
    h = 1e-6
    for i in idx:
        fp = flat.copy()
        fp[i] += h
        fm = flat.copy()
        fm[i] -= h
        lp = D.loss_fn(D.forward(D.unflatten(fp), rec, False)[0], y, huber)
        lm = D.loss_fn(D.forward(D.unflatten(fm), rec, False)[0], y, huber)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, gflat[i])


def _torch_net(p, rec, y, huber):
    import torch
    import torch.nn.functional as F
    t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    x = torch.tensor(D.input_channels(rec, False)).permute(0, 3, 1, 2)  # B, C, H, W
    for l in range(4):
        w = t[f"conv{l}.w"].reshape(3, 3, -1, D.C).permute(3, 2, 0, 1)  # co, ci, ky, kx
        xp = F.pad(x, (1, 1, 0, 0), mode="circular")               # sites: ring
        xp = F.pad(xp, (0, 0, 1, 1))                                 # time: zeros
        x = torch.relu(F.conv2d(xp, w) + t[f"conv{l}.b"][None, :, None, None])
    pooled = x.mean(dim=(2, 3))
    h = torch.relu(pooled @ t["fc1.w"] + t["fc1.b"])
    out = h @ t["fc2.w"] + t["fc2.b"]
    yt = torch.tensor(y)
    n = lambda v: torch.sqrt((v * v).sum(-1) + D.NORM_EPS)
    def S(r):
        c = torch.clamp(r, max=D.R_MAX)
        a, b = 0.5 * (1 + c), 0.5 * (1 - c)
        return -a * torch.log(a) - b * torch.log(b) - float(np.arctanh(D.R_MAX)) * torch.clamp(r - D.R_MAX, min=0.0)
    d = S(n(out)) - S(n(yt))
    h = d * d if huber <= 0 else F.smooth_l1_loss(d, torch.zeros_like(d), reduction="none", beta=huber)
    loss = (h + D.MU * (n(yt) - (out * yt).sum(-1) / n(out))).mean()
    loss.backward()
    return float(loss.detach()), {k: v.grad.numpy() for k, v in t.items()}, out.detach().numpy()


@pytest.mark.parametrize("fc2_scale", [1.0, 0.3])  # |out| above and below R_MAX
@pytest.mark.parametrize("huber", [0.0, 0.05])
def test_gradients_match_torch_autograd(fc2_scale, huber):
    p, rec, y = _case(B=4, H=6, W=5, seed=2)
    p["fc2.w"] = p["fc2.w"] * fc2_scale
    p["fc2.b"] = p["fc2.b"] * fc2_scale
    loss, g = D.gradients(p, rec, y, emulate=False, huber=huber)
    tl, tg, tout = _torch_net(p, rec, y, huber)
    out, _ = D.forward(p, rec, False)
    assert np.allclose(out, tout, atol=1e-12)
    assert abs(loss - tl) < 1e-12
    for k in g:
        assert np.allclose(g[k], tg[k], atol=1e-10, rtol=1e-9), k


def test_adam_matches_torch():
    import torch
    p, rec, y = _case(B=2, H=4, W=4, seed=5)
    m = {k: np.zeros_like(v) for k, v in p.items()}
    v = {k: np.zeros_like(x) for k, x in p.items()}
    tp = {k: torch.tensor(x, dtype=torch.float64, requires_grad=True) for k, x in p.items()}
    opt = torch.optim.Adam(list(tp.values()), lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    cur = p
    for t in range(1, 4):
        _, g = D.gradients(cur, rec, y, emulate=False)
        cur, m, v = D.adam_step(cur, g, m, v, t)
        for k in tp:
            tp[k].grad = torch.tensor(g[k])
        opt.step()
        for k in tp:
            assert np.allclose(cur[k], tp[k].detach().numpy(), atol=1e-12), (t, k)


def test_entropy_from_bloch_closed_forms():
    assert abs(D.entropy_from_bloch(np.zeros(3)) - np.log(2)) < 1e-15
    assert D.entropy_from_bloch(np.array([0.0, 0.0, 1.0])) == 0.0
    assert D.entropy_from_bloch(np.array([0.0, 2.0, 0.0])) == 0.0  # clipped
    r = 0.6
    want = -(0.8 * np.log(0.8) + 0.2 * np.log(0.2))
    assert abs(D.entropy_from_bloch(np.array([r, 0, 0])) - want) < 1e-15
