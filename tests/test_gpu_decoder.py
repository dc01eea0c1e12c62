"""GPU tests of the learned probe decoder (BF16-MFMA forward/backward/Adam).

Bars: loss and predictions within 1e-3 (relative) of the FP64 oracle run with
the GPU's bfloat16 rounding points (oracle/decoder_oracle.py); each gradient
tensor within 2e-2 relative L2 error of the oracle's (FP32 vs FP64
accumulation can move an activation across a bfloat16 rounding boundary);
one Adam step within 1e-6 of the oracle's update on the GPU's own gradients;
and, trained on simulator data, the decoder's probe entropy beats the p-bin
mean entropy (the method without it) on held-out trajectories.
"""
import time

import numpy as np
import pytest

from oracle import decoder_oracle as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dec():
    from mipt_amd import decoder
    return decoder


def _params(seed):
    rng = np.random.default_rng(seed)
    p = D.he_init(seed)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.normal(0.0, 0.05, size=p[k].shape)
    flat = D.flatten(p).astype(np.float32)
    return flat, D.unflatten(flat.astype(np.float64))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_decoder_forward_and_gradients_match_oracle(dec):
    print("decoder gradients vs oracle (B=8, 32x16)", flush=True)
    rng = np.random.default_rng(1)
    B = 8
    rec = rng.integers(-1, 2, size=(B, 32, 16)).astype(np.int8)
    lab = (rng.normal(size=(B, 3)) * 0.5).astype(np.float32)
    flat, P = _params(11)
    d = dec.Decoder(32, 16, 64, seed=3)
    assert d.nparam == D.param_count()
    d.set_params(flat)
    assert np.array_equal(d.get_params(), flat)
    d.set_dataset(rec, lab)
    out = d.predict()
    want, _ = D.forward(P, rec, emulate=True)
    assert _rel(out, want) < 1e-3
    idx = np.array([3, 1, 4, 0, 6, 2, 7, 5])
    for huber in (0.05, 0.0):
        d.set_loss(huber)
        loss, g = d.gradients(idx)
        lo, go = D.gradients(P, rec[idx], lab[idx].astype(np.float64), emulate=True, huber=huber)
        assert abs(loss - lo) < 1e-3 * abs(lo) + 1e-6, (huber, loss, lo)
        gdev = D.unflatten(g.astype(np.float64))
        errs = {k: _rel(gdev[k], go[k]) for k in go}
        print(f"huber {huber}: relative gradient errors:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        for k, e in errs.items():
            assert e < 2e-2, (huber, k, e)
    # a transposed or shifted layout would be O(1) off: the unrounded oracle
    # stays close too, the rounding points only move the last digits
    _, gn = D.gradients(P, rec[idx], lab[idx].astype(np.float64), emulate=False, huber=0.0)
    assert _rel(gdev["conv1.w"], gn["conv1.w"]) < 0.1


def test_decoder_adam_step_matches_oracle(dec):
    print("decoder Adam step vs oracle", flush=True)
    rng = np.random.default_rng(2)
    B = 16
    rec = rng.integers(-1, 2, size=(B, 32, 16)).astype(np.int8)
    lab = (rng.normal(size=(B, 3)) * 0.5).astype(np.float32)
    flat, P = _params(12)
    d = dec.Decoder(32, 16, 32, seed=4)
    d.set_params(flat)
    d.set_dataset(rec, lab)
    idx = np.arange(B)
    _, g = d.gradients(idx)
    d.train(idx, B, lr=1e-3)
    new = d.get_params().astype(np.float64)
    G = D.unflatten(g.astype(np.float64))
    zero = {k: np.zeros_like(v) for k, v in P.items()}
    want, _, _ = D.adam_step(P, G, zero, zero, 1, lr=1e-3)
    assert np.abs(new - D.flatten(want)).max() < 1e-6


def test_decoder_learns_probe_entropy_beyond_p_bin_baseline(dec):
    print("decoder training on simulator data (L=16, 32 scrambling + 32 measured layers)", flush=True)
    n_train, n_test = 65536, 4096
    train = dec.generate_probe_dataset(np.arange(1, n_train + 1), chunk=16384)
    test = dec.generate_probe_dataset(np.arange(1_000_001, 1_000_001 + n_test), chunk=4096)
    print(f"simulated {n_train + n_test} trajectories in {train.seconds + test.seconds:.2f} s", flush=True)
    assert np.all(np.isfinite(train.bloch)) and np.all(np.abs(train.bloch) <= 1.0 + 1e-5)
    d = dec.Decoder(32, 16, 512, seed=5)
    d.set_dataset(np.concatenate([train.records, test.records]), np.concatenate([train.bloch, test.bloch]))
    rng = np.random.default_rng(0)
    for epoch in range(12):
        loss = d.train(rng.permutation(n_train), 512, lr=dec.cosine_lr(1e-3, epoch, 12))
        print(f"epoch {epoch}: mean batch loss {loss:.4f}", flush=True)
    t = time.perf_counter()
    pred = d.predict(n_train, n_test)
    t_inf = time.perf_counter() - t
    s_exact = dec.entropy_from_bloch(test.bloch)
    s_pred = dec.entropy_from_bloch(pred)
    base, bins, edges = dec.p_bin_baseline(train, test)
    err_dec = np.abs(s_pred - s_exact)
    err_base = np.abs(base - s_exact)
    for k in range(len(edges) - 1):
        m = bins == k
        print(f"p in [{edges[k]:.2f},{edges[k + 1]:.2f}): decoder {err_dec[m].mean():.4f}  baseline {err_base[m].mean():.4f}",
              flush=True)
    print(f"mean |S_pred - S_exact|: decoder {err_dec.mean():.4f} nats, p-bin baseline {err_base.mean():.4f} nats; "
          f"{n_test / t_inf:.0f} inferences/s", flush=True)
    assert err_dec.mean() < err_base.mean()
