"""Learned probe decoder: host side (ctypes over libmipt_hip.so).

Data come from the batched engine itself: a chain of L sites with one fixed
Haar gate realisation shared by every trajectory (a donor trajectory's gate
stream), `scramble` unmeasured layers, the probe attached to site L/2
(reset + Bell pair), then `layers` measured layers at a per-trajectory rate
p in [p_min, p_max].  The input is the measurement record of the measured
layers (int8 [layers][L]: -1 unmeasured, 0, 1), the label the probe's exact
conditional Bloch vector.  Training runs on the GPU through the library's
BF16-MFMA forward/backward/Adam kernels.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np

from . import Batch, Spec, _check, _dp, lib

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_i8p = ctypes.POINTER(ctypes.c_int8)


def _bind():
    L = lib()
    if getattr(L, "_decoder_bound", False):
        return L
    vp, i = ctypes.c_void_p, ctypes.c_int
    L.mipt_hip_decoder_create.argtypes = [i, i, i, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.mipt_hip_decoder_destroy.argtypes = [vp]
    L.mipt_hip_decoder_param_count.argtypes = [vp]
    L.mipt_hip_decoder_get_params.argtypes = [vp, _fp]
    L.mipt_hip_decoder_set_params.argtypes = [vp, _fp]
    L.mipt_hip_decoder_set_dataset.argtypes = [vp, _i8p, _fp, i]
    L.mipt_hip_decoder_gradients.argtypes = [vp, _ip, i, _fp, ctypes.POINTER(ctypes.c_double)]
    L.mipt_hip_decoder_train.argtypes = [vp, _ip, i, i, ctypes.c_float, ctypes.POINTER(ctypes.c_double)]
    L.mipt_hip_decoder_predict.argtypes = [vp, i, i, _fp]
    L.mipt_hip_decoder_synchronize.argtypes = [vp]
    L.mipt_hip_decoder_set_loss.argtypes = [vp, ctypes.c_float]
    L._decoder_bound = True
    return L


class Decoder:
    def __init__(self, height: int = 32, width: int = 16, batch_max: int = 1024, seed: int = 1):
        L = _bind()
        h = ctypes.c_void_p()
        _check(L.mipt_hip_decoder_create(height, width, batch_max, seed, ctypes.byref(h)), "decoder_create")
        self._h = h
        self.height, self.width, self.batch_max = height, width, batch_max
        self.count = 0

    def close(self):
        if self._h:
            lib().mipt_hip_decoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparam(self) -> int:
        return lib().mipt_hip_decoder_param_count(self._h)

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparam, np.float32)
        _check(lib().mipt_hip_decoder_get_params(self._h, p.ctypes.data_as(_fp)), "decoder_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam
        _check(lib().mipt_hip_decoder_set_params(self._h, p.ctypes.data_as(_fp)), "decoder_set_params")

    def set_dataset(self, records, labels=None):
        r = np.ascontiguousarray(records, np.int8)
        assert r.shape[1:] == (self.height, self.width)
        lab = None if labels is None else np.ascontiguousarray(labels, np.float32)
        _check(lib().mipt_hip_decoder_set_dataset(self._h, r.ctypes.data_as(_i8p),
                                                  None if lab is None else lab.ctypes.data_as(_fp), len(r)),
               "decoder_set_dataset")
        self.count = len(r)

    def set_loss(self, huber_delta: float):
        """Smooth-L1 width of the entropy term (0: squared error)."""
        _check(lib().mipt_hip_decoder_set_loss(self._h, huber_delta), "decoder_set_loss")

    def gradients(self, idx):
        idx = np.ascontiguousarray(idx, np.int32)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_double()
        _check(lib().mipt_hip_decoder_gradients(self._h, idx.ctypes.data_as(_ip), len(idx), g.ctypes.data_as(_fp),
                                                ctypes.byref(loss)), "decoder_gradients")
        return loss.value, g

    def train(self, idx, batch: int, lr: float = 1e-3, want_loss: bool = True) -> float:
        idx = np.ascontiguousarray(idx, np.int32)
        steps = len(idx) // batch
        loss = ctypes.c_double()
        _check(lib().mipt_hip_decoder_train(self._h, idx.ctypes.data_as(_ip), batch, steps, lr,
                                            ctypes.byref(loss) if want_loss else None), "decoder_train")
        return loss.value

    def predict(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.count - first if count is None else count
        out = np.zeros((count, 3), np.float32)
        _check(lib().mipt_hip_decoder_predict(self._h, first, count, out.ctypes.data_as(_fp)), "decoder_predict")
        return out


def cosine_lr(lr: float, epoch: int, epochs: int) -> float:
    """Learning rate of epoch `epoch` (0-based): a half cosine from lr down
    towards 0 over `epochs` epochs."""
    return lr * 0.5 * (1.0 + math.cos(math.pi * epoch / epochs))


def entropy_from_bloch(r) -> np.ndarray:
    """Qubit von Neumann entropy (nats) from a Bloch vector, |r| clipped to 1."""
    n = np.minimum(np.sqrt((np.asarray(r, np.float64) ** 2).sum(-1)), 1.0)
    lp, lm = 0.5 * (1 + n), 0.5 * (1 - n)
    with np.errstate(divide="ignore", invalid="ignore"):
        return -np.where(lp > 0, lp * np.log(lp), 0.0) - np.where(lm > 0, lm * np.log(lm), 0.0)


def rate_of_seed(seed: int, p_min: float, p_max: float) -> float:
    """A trajectory's measurement rate, uniform in [p_min, p_max] from its seed."""
    x = (seed * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & (2**64 - 1)
    x ^= x >> 31
    x = (x * 0xBF58476D1CE4E5B9) & (2**64 - 1)
    x ^= x >> 29
    return p_min + (p_max - p_min) * ((x >> 11) / float(1 << 53))


@dataclass
class ProbeData:
    records: np.ndarray  # int8 [n, layers, L]
    bloch: np.ndarray    # float32 [n, 3]
    p: np.ndarray        # float64 [n]
    seeds: np.ndarray    # uint64 [n]
    seconds: float       # wall time of the simulation (draws + evolution + readout)


def generate_probe_dataset(seeds, L: int = 16, scramble: int = 32, layers: int = 32, p_min: float = 0.05,
                           p_max: float = 0.30, gate_seed: int = 7, chunk: int = 16384) -> ProbeData:
    """Simulate one trajectory per seed (see the module docstring)."""
    seeds = np.asarray(seeds, np.uint64)
    n = len(seeds)
    rec = np.zeros((n, layers, L), np.int8)
    bloch = np.zeros((n, 3), np.float32)
    ps = np.array([rate_of_seed(int(s), p_min, p_max) for s in seeds])
    total = scramble + layers
    t0 = time.perf_counter()
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        # trajectory 0 is the gate donor (seed gate_seed), not part of the data
        specs = [Spec(L, total, 0.0, gate_seed, True, probe_site=L // 2, p_zero_layers=scramble)]
        specs += [Spec(L, total, float(ps[i]), int(seeds[i]), True, probe_site=L // 2, p_zero_layers=scramble,
                       shared_gates=0) for i in range(c0, c1)]
        b = Batch(specs)
        b.generate()
        b.evolve()
        b.synchronize()
        rec[c0:c1] = b.records(scramble, total)[1:]
        bloch[c0:c1] = b.probe_bloch()[1:].astype(np.float32)
        b.close()
    return ProbeData(rec, bloch, ps, seeds, time.perf_counter() - t0)


def p_bin_baseline(train: ProbeData, test: ProbeData, bins: int = 5, p_min: float = 0.05, p_max: float = 0.30):
    """The method without the decoder: predict a trajectory's probe entropy as
    the mean exact entropy of training trajectories in its p bin."""
    edges = np.linspace(p_min, p_max, bins + 1)
    s_train = entropy_from_bloch(train.bloch)
    bi_train = np.clip(np.digitize(train.p, edges) - 1, 0, bins - 1)
    bi_test = np.clip(np.digitize(test.p, edges) - 1, 0, bins - 1)
    means = np.array([s_train[bi_train == k].mean() if np.any(bi_train == k) else s_train.mean() for k in range(bins)])
    return means[bi_test], bi_test, edges
