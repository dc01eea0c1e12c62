"""mipt_amd: MI355X-native qubit MIPT trajectory engine (host mirror of the C ABI).

The compute path is ``libmipt_hip.so`` (hand-written gfx950 HIP kernels behind
the C ABI declared in ``include/mipt_hip.h``).  This module binds it with
ctypes and mirrors the reference's host-side interfaces for the hot path
(paths relative to the reference repository root):

* ``PlainEngine``     <- ``cusv::Engine`` (cpp/include/mipt/cusv/engine.hpp:100-582)
* ``Batch``           <- the serial loops of ``run_1d_sim_tmi`` / ``run_1d``
                          (cpp/include/mipt/tmi/runner.hpp:269-378,
                          cpp/apps/mipt.cpp:111-190), batched on device
* module functions    <- the device C ABI (cpp/include/mipt/cuda/tmi_cuda_svd.hpp,
                          tmi_cuda_rdm.hpp, free_energy_measure.hpp, ancilla_rdm.hpp)

Errors raise ``MiptHipError`` carrying the status code and the library's
last-error text.  There is no CPU fallback: if the shared library is missing
the import of any entry point fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmipt_hip.so")

_LIB = None


class MiptHipError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what} failed with status {status}: {last_error()}")
        self.status = status


def lib():
    """The loaded libmipt_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mipt_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp, u8p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), \
        ctypes.POINTER(ctypes.c_uint8)
    fp = ctypes.POINTER(ctypes.c_float)
    i = ctypes.c_int
    L.mipt_hip_last_error.restype = ctypes.c_char_p
    L.mipt_hip_version.restype = ctypes.c_char_p
    L.mipt_hip_measure_collapse_last_error.restype = ctypes.c_char_p
    for name in ("mipt_hip_measure_collapse_f32", "mipt_hip_measure_collapse_f64",
                 "mipt_hip_measure_collapse_f32_noalloc"):
        getattr(L, name).argtypes = [vp, i, i, ctypes.c_double, ip, dp]
    for name in ("mipt_hip_entropy_svd_f32", "mipt_hip_entropy_svd_f64", "mipt_hip_renyi2_f32",
                 "mipt_hip_renyi2_f64", "mipt_hip_rdm_subsystem_f32", "mipt_hip_rdm_subsystem_f64"):
        getattr(L, name).argtypes = [vp, i, ip, i, i, dp]
    for name in ("mipt_hip_rho1_complex_f32", "mipt_hip_rho1_complex_f64"):
        getattr(L, name).argtypes = [vp, i, i, dp]
    L.mipt_hip_tmi_f32.argtypes = [vp, i, i, i, i, dp, dp]
    L.mipt_hip_tmi_f64.argtypes = [vp, i, i, i, i, dp, dp]
    L.mipt_hip_set_gram_precision.argtypes = [i]
    L.mipt_hip_set_eigensolver.argtypes = [i]
    L.mipt_hip_entropy_hermitian_batched.argtypes = [dp, i, i, dp, dp]
    L.mipt_hip_set_gram_profiling.argtypes = [i]
    L.mipt_hip_gram_stats.argtypes = [dp, ip, dp]
    L.mipt_hip_engine_create.argtypes = [i, vp, ctypes.POINTER(vp)]
    L.mipt_hip_engine_destroy.argtypes = [vp]
    L.mipt_hip_engine_adopt.argtypes = [vp, vp, i]
    L.mipt_hip_engine_reset.argtypes = [vp, i]
    L.mipt_hip_engine_data.argtypes = [vp]
    L.mipt_hip_engine_data.restype = vp
    L.mipt_hip_engine_apply_matrix.argtypes = [vp, ip, i, dp]
    L.mipt_hip_engine_apply_bonds.argtypes = [vp, ip, ip, dp, i, i]
    L.mipt_hip_engine_measure_layer.argtypes = [vp, ip, i, dp, i, i, ip, dp, dp]
    L.mipt_hip_engine_run_trajectory.argtypes = [vp, i, i, i, i, dp, u8p, dp, i, i, i, ip]
    L.mipt_hip_engine_to_host.argtypes = [vp, vp]
    L.mipt_hip_engine_set_from_host.argtypes = [vp, vp]
    L.mipt_hip_engine_squared_norm.argtypes = [vp, dp]
    L.mipt_hip_batch_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
    L.mipt_hip_batch_destroy.argtypes = [vp]
    L.mipt_hip_batch_setup.argtypes = [vp, vp, i]
    L.mipt_hip_batch_generate.argtypes = [vp]
    L.mipt_hip_batch_set_layers.argtypes = [vp, i, dp, u8p, dp]
    L.mipt_hip_batch_set_state.argtypes = [vp, i, fp]
    L.mipt_hip_batch_get_state.argtypes = [vp, i, fp]
    L.mipt_hip_batch_state_ptr.argtypes = [vp, i]
    L.mipt_hip_batch_state_ptr.restype = vp
    L.mipt_hip_batch_evolve.argtypes = [vp]
    L.mipt_hip_batch_evolve_range.argtypes = [vp, i, i]
    L.mipt_hip_batch_synchronize.argtypes = [vp]
    L.mipt_hip_batch_outcomes.argtypes = [vp, i, ip, u8p, dp, dp]
    L.mipt_hip_batch_observables.argtypes = [vp, i, i, dp, dp, dp, dp]
    L.mipt_hip_batch_stream.argtypes = [vp]
    L.mipt_hip_batch_stream.restype = vp
    L.mipt_hip_batch_reset.argtypes = [vp]
    L.mipt_hip_batch_export_layers.argtypes = [vp, i, dp, u8p, dp]
    L.mipt_hip_batch_set_profiling.argtypes = [vp, i]
    L.mipt_hip_batch_set_streams.argtypes = [vp, i]
    L.mipt_hip_layer_plan.argtypes = [i, i, i, i, i, i, i, i, ip, i, ip]
    L.mipt_hip_compiled_sweeps.argtypes = [ip, i, ip]
    L.mipt_hip_set_layer_fusion.argtypes = [i]
    L.mipt_hip_batch_pass_stats.argtypes = [vp, dp, ip, dp, ctypes.POINTER(ctypes.c_long), dp]
    L.mipt_hip_stream_copy.argtypes = [vp, vp, ctypes.c_uint64, vp]
    L.mipt_hip_batch_tmi.argtypes = [vp, i, i, i, i, dp]
    L.mipt_hip_batch_records.argtypes = [vp, i, i, ctypes.POINTER(ctypes.c_int8)]
    L.mipt_hip_batch_observables_all.argtypes = [vp, i, i, dp, dp]
    L.mipt_hip_batch_autotune.argtypes = [vp, i, ip, ip, dp]
    L.mipt_hip_batch_rho3.argtypes = [vp, i, i, ip, i, dp]
    L.mipt_hip_mipt_main.argtypes = [i, ctypes.POINTER(ctypes.c_char_p)]
    L.mipt_hip_entropy_main.argtypes = [i, ctypes.POINTER(ctypes.c_char_p)]
    L.mipt_hip_batch_translation_entropies.argtypes = [vp, i, i, i, dp, dp]
    L.mipt_hip_batch_set_graph.argtypes = [vp, i]
    L.mipt_hip_batch_probe_bloch.argtypes = [vp, dp]
    L.mipt_hip_batch_probe_rdm.argtypes = [vp, dp]
    L.mipt_hip_probe_information.argtypes = [dp, i, dp]
    L.mipt_hip_run_batch.argtypes = [vp, i, i, i, i, dp, vp, vp]
    L.mipt_hip_sim_tmi_main.argtypes = [i, ctypes.POINTER(ctypes.c_char_p)]
    L.mipt_hip_probed_main.argtypes = [i, ctypes.POINTER(ctypes.c_char_p)]
    L.mipt_hip_free_energy_main.argtypes = [i, ctypes.POINTER(ctypes.c_char_p)]
    L.mipt_hip_free_energy_initial_state.argtypes = [i, i, ctypes.c_uint64, dp]
    _LIB = L
    return L


def last_error() -> str:
    if _LIB is None:
        return ""
    return _LIB.mipt_hip_last_error().decode(errors="replace")


def _check(status: int, what: str) -> None:
    if status != 0:
        raise MiptHipError(status, what)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ------------------------------------------------------------------ drop-ins
def set_gram_profiling(enable: bool) -> None:
    """Time every Gram kernel of the complex64 RDM path (clears earlier timings)."""
    _check(lib().mipt_hip_set_gram_profiling(int(enable)), "set_gram_profiling")


def gram_stats() -> dict:
    """Average Gram kernel ms, count timed, average algorithmic FLOPs."""
    ms, n, fl = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
    _check(lib().mipt_hip_gram_stats(ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl)), "gram_stats")
    return {"avg_ms": ms.value, "timed": n.value, "flops_avg": fl.value}


def set_gram_precision(bits: int) -> None:
    """Gram arithmetic of complex64 RDMs: 64 (FP64 MFMA, default) or 32 (FP32 MFMA)."""
    _check(lib().mipt_hip_set_gram_precision(int(bits)), "set_gram_precision")


def set_eigensolver(which: int) -> None:
    """Eigenvalue path of the entropies: 0 = the library's batched FP64 kernels
    (default), 1 = rocSOLVER zheevd."""
    _check(lib().mipt_hip_set_eigensolver(int(which)), "set_eigensolver")


def entropy_hermitian_batched(rhos, eigenvalues: bool = True):
    """Entropies (bits) and ascending eigenvalues of a stack of Hermitian
    matrices [count, m, m] on the library's batched FP64 eigensolver
    (entropy_hermitian_inplace, linalg.hpp:524-583, for many matrices)."""
    r = np.ascontiguousarray(rhos, np.complex128)
    count, m = r.shape[0], r.shape[1]
    ent = np.zeros(count)
    lam = np.zeros((count, m)) if eigenvalues else None
    _check(lib().mipt_hip_entropy_hermitian_batched(r.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    m, count, _dp(ent), _dp(lam) if eigenvalues else None),
           "entropy_hermitian_batched")
    return (ent, lam) if eigenvalues else ent


def measure_collapse(state_ptr: int, n_qubits: int, qubit: int, uniform: float, fp64: bool = False):
    """mipt_cuda_measure_collapse_* (free_energy_measure.hpp:13-27): (outcome, P(outcome))."""
    o = ctypes.c_int()
    p = ctypes.c_double()
    f = lib().mipt_hip_measure_collapse_f64 if fp64 else lib().mipt_hip_measure_collapse_f32
    _check(f(ctypes.c_void_p(state_ptr), n_qubits, qubit, float(uniform), ctypes.byref(o), ctypes.byref(p)),
           "measure_collapse")
    return o.value, p.value


def entropy(state_ptr: int, n_qubits: int, kept, fp64: bool = False, order: int = 1) -> float:
    """mipt_cuda_entropy_svd_* / renyi2_* (tmi_cuda_svd.hpp:20-58), bits."""
    k = np.ascontiguousarray(kept, np.int32)
    out = ctypes.c_double()
    if order == 2:
        f = lib().mipt_hip_renyi2_f64 if fp64 else lib().mipt_hip_renyi2_f32
    else:
        f = lib().mipt_hip_entropy_svd_f64 if fp64 else lib().mipt_hip_entropy_svd_f32
    _check(f(ctypes.c_void_p(state_ptr), n_qubits, _ip(k), len(k), 0, ctypes.byref(out)), "entropy")
    return out.value


def rdm_subsystem(state_ptr: int, n_qubits: int, kept, fp64: bool = False) -> np.ndarray:
    """mipt_cuda_rdm_subsystem_* (tmi_cuda_rdm.hpp:20-33): complex128 [D, D]."""
    k = np.ascontiguousarray(kept, np.int32)
    d = 1 << len(k)
    out = np.zeros((d, d), np.complex128)
    f = lib().mipt_hip_rdm_subsystem_f64 if fp64 else lib().mipt_hip_rdm_subsystem_f32
    _check(f(ctypes.c_void_p(state_ptr), n_qubits, _ip(k), len(k), 0,
             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "rdm_subsystem")
    return out


def rho1(state_ptr: int, n_qubits: int, qubit: int, fp64: bool = False) -> np.ndarray:
    """mipt_cuda_rho1_complex_* (ancilla_rdm.hpp:19-29): {rho00, rho11, Re rho01, Im rho01}."""
    out = np.zeros(4)
    f = lib().mipt_hip_rho1_complex_f64 if fp64 else lib().mipt_hip_rho1_complex_f32
    _check(f(ctypes.c_void_p(state_ptr), n_qubits, qubit, _dp(out)), "rho1")
    return out


def tmi(state_ptr: int, n_qubits: int, chain_sites: int | None = None, block: int | None = None, cycle: int = 0):
    """I3 in bits with the seven terms (compute.hpp:186-227)."""
    L = n_qubits if chain_sites is None else chain_sites
    b = L // 4 if block is None else block
    out = ctypes.c_double()
    terms = np.zeros(7)
    _check(lib().mipt_hip_tmi_f32(ctypes.c_void_p(state_ptr), n_qubits, L, b, cycle, ctypes.byref(out), _dp(terms)),
           "tmi")
    return out.value, terms


# ------------------------------------------------------------------ plain engine
class PlainEngine:
    """One trajectory, one launch per blocked op, a host round trip per measured
    layer: cusv::Engine (cpp/include/mipt/cusv/engine.hpp:100-582)."""

    def __init__(self, fp64: bool = False, stream: int | None = None):
        self.fp64 = fp64
        h = ctypes.c_void_p()
        _check(lib().mipt_hip_engine_create(int(fp64), ctypes.c_void_p(stream or 0), ctypes.byref(h)),
               "engine_create")
        self._h = h
        self.n = 0

    def close(self):
        if self._h:
            lib().mipt_hip_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, n: int):
        _check(lib().mipt_hip_engine_reset(self._h, n), "engine_reset")
        self.n = n

    def adopt(self, ptr: int, n: int):
        _check(lib().mipt_hip_engine_adopt(self._h, ctypes.c_void_p(ptr), n), "engine_adopt")
        self.n = n

    @property
    def data(self) -> int:
        return lib().mipt_hip_engine_data(self._h)

    def apply_matrix(self, targets, m):
        t = np.ascontiguousarray(targets, np.int32)
        mm = np.ascontiguousarray(m, np.complex128)
        _check(lib().mipt_hip_engine_apply_matrix(self._h, _ip(t), len(t), mm.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "apply_matrix")

    def apply_bonds(self, bonds, gates, block_targets: int = 4):
        """bonds: [(q0, q1)]; gates: complex128 [nb, 4, 4] (bit 0 = q0)."""
        if len(bonds) == 0:
            return
        q0 = np.ascontiguousarray([b[0] for b in bonds], np.int32)
        q1 = np.ascontiguousarray([b[1] for b in bonds], np.int32)
        g = np.ascontiguousarray(gates, np.complex128)
        _check(lib().mipt_hip_engine_apply_bonds(self._h, _ip(q0), _ip(q1), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                 len(bonds), block_targets), "apply_bonds")

    def measure_layer(self, sites, uniforms, mode: int = 0, chunk_bits: int = 10):
        """mode 0: sequential per-site draws; mode 1: joint chunk draws (engine.hpp:293-368).
        Returns (outcomes, p1 (mode 0), surprisal)."""
        s = np.ascontiguousarray(sites, np.int32)
        u = np.ascontiguousarray(uniforms, np.float64)
        o = np.zeros(len(s), np.int32)
        p1 = np.zeros(len(s), np.float64)
        sp = ctypes.c_double()
        _check(lib().mipt_hip_engine_measure_layer(self._h, _ip(s), len(s), _dp(u), mode, chunk_bits, _ip(o), _dp(p1),
                                                   ctypes.byref(sp)), "measure_layer")
        return o, p1, sp.value

    def to_host(self) -> np.ndarray:
        if self.n < 1:
            raise ValueError("engine holds no state (reset, adopt or run a trajectory first)")
        out = np.zeros(1 << self.n, np.complex128 if self.fp64 else np.complex64)
        _check(lib().mipt_hip_engine_to_host(self._h, out.ctypes.data_as(ctypes.c_void_p)), "to_host")
        return out

    def set_from_host(self, psi: np.ndarray):
        a = np.ascontiguousarray(psi, np.complex128 if self.fp64 else np.complex64)
        if self.n < 1 or a.size != (1 << self.n):
            raise ValueError(f"state of {a.size} amplitudes for an engine of {self.n} qubits")
        _check(lib().mipt_hip_engine_set_from_host(self._h, a.ctypes.data_as(ctypes.c_void_p)), "set_from_host")

    def squared_norm(self) -> float:
        out = ctypes.c_double()
        _check(lib().mipt_hip_engine_squared_norm(self._h, ctypes.byref(out)), "squared_norm")
        return out.value


def set_layer_fusion(on) -> None:
    """Process-wide: run a layer's last sweep and the next layer's first in
    one launch for the trajectories that measure at most one site in the
    layer, not bonded by the latter (True / 1, the default), or not (False /
    0); None or -1 goes back to MIPT_HIP_FUSE / the default."""
    if on is None:
        on = -1
    _check(lib().mipt_hip_set_layer_fusion(int(on) if isinstance(on, bool) else on), "set_layer_fusion")


def layer_plan(L: int, nq: int, closed: bool, grid_x: int, layer: int, tile_bits: int = 13, threads: int = 0,
               plan: int = 0) -> dict:
    """The sweeps the batch runs for one layer (host only, no GPU): {"bonds":
    [(q0, q1)], "passes": [{"k", "c", "s", "stages": [{"use_lds", "zpair",
    "bonds": [(bond, q0, q1)] x 2, "reg_u", "lane_u", "wave_u"}]}]}."""
    buf = np.zeros(1 << 14, dtype=np.int32)
    n = ctypes.c_int()
    _check(lib().mipt_hip_layer_plan(L, nq, int(bool(closed)), grid_x, layer, tile_bits, threads, plan, _ip(buf),
                                     len(buf), ctypes.byref(n)), "layer_plan")
    v = buf[: n.value].tolist()
    pos = 0

    def take(k):
        nonlocal pos
        out = v[pos:pos + k]
        pos += k
        return out

    nb = take(1)[0]
    flat = take(2 * nb)
    bonds = [(flat[2 * b], flat[2 * b + 1]) for b in range(nb)]
    passes = []
    for _ in range(take(1)[0]):
        k, c, s, ns = take(4)
        stages = []
        for _ in range(ns):
            use_lds, zpair = take(2)
            sb = [tuple(take(3)) for _ in range(2)]
            stages.append({"use_lds": use_lds, "zpair": zpair, "bonds": sb, "reg_u": take(6), "lane_u": take(6),
                           "wave_u": take(4)})
        passes.append({"k": k, "c": c, "s": s, "stages": stages})
    return {"bonds": bonds, "passes": passes}


def compiled_sweeps() -> list:
    """The sweep kernels compiled for a fixed stage structure: [(zp, pat,
    pat2)] (host only, no GPU)."""
    buf = np.zeros(1024, dtype=np.int32)
    n = ctypes.c_int()
    _check(lib().mipt_hip_compiled_sweeps(_ip(buf), len(buf), ctypes.byref(n)), "compiled_sweeps")
    v = buf[: n.value].tolist()
    return [tuple(v[i:i + 3]) for i in range(0, len(v), 3)]


def stream_copy_calibration(log2_elems: int = 25) -> int:
    """Copies a 2^log2_elems complex64 buffer (256 MiB at 25) once with a plain
    streaming kernel; returns the bytes read (= bytes written)."""
    src = PlainEngine(False)
    dst = PlainEngine(False)
    src.reset(log2_elems)
    dst.reset(log2_elems)
    nbytes = (1 << log2_elems) * 8
    _check(lib().mipt_hip_stream_copy(ctypes.c_void_p(dst.data), ctypes.c_void_p(src.data), nbytes, None),
           "stream_copy")
    src.close()
    dst.close()
    return nbytes


def layer_bonds(n: int, layer: int, closed: bool = True, grid_x: int = 0):
    """Bonds (q0, q1) of a layer in gate-stream order: the brickwork chain
    (cpp/include/mipt/circuits/haar.hpp:226-243), or with grid_x > 0 the four
    colours of an n = grid_x * (n / grid_x) grid (site r*grid_x + c): 0/1
    horizontal from even/odd columns row by row (1 closes each row when
    closed and grid_x is even), 2/3 vertical in each column from rows
    ((c % 2) + colour - 2) % 2 in steps of 2."""
    if grid_x <= 0:
        start = layer & 1
        stop = n if (closed and start == 1 and n > 2) else n - 1
        return [(i, (i + 1) % n) for i in range(start, stop, 2)]
    gx, gy, colour = grid_x, n // grid_x, layer % 4
    out = []
    if colour < 2:
        for r in range(gy):
            out += [(r * gx + c, r * gx + c + 1) for c in range(colour, gx - 1, 2)]
            if colour == 1 and closed and gx % 2 == 0 and gx > 2:
                out.append((r * gx + gx - 1, r * gx))
        return out
    for c in range(gx):
        for r in range(((c % 2) + colour - 2) % 2, gy - 1, 2):
            out.append((r * gx + c, (r + 1) * gx + c))
    return out


def run_plain_trajectory_native(engine: PlainEngine, n: int, gates, flags, uniforms, closed: bool = True,
                                block_targets: int = 4, mode: int = 0, chunk_bits: int = 10, grid_x: int = 0):
    """The try_cusv loop (cpp/include/mipt/circuit.hpp:240-282) in one C-ABI call
    (mipt_hip_engine_run_trajectory).  Same arguments and result as
    run_plain_trajectory."""
    flags = np.ascontiguousarray(flags, np.uint8)
    layers = flags.shape[0]
    u = np.ascontiguousarray(uniforms, np.float64)
    g = np.ascontiguousarray(gates, np.complex128)
    outcomes = np.full((layers, n), -1, np.int32)
    _check(lib().mipt_hip_engine_run_trajectory(engine._h, n, layers, int(closed), grid_x,
                                                g.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(u),
                                                block_targets, mode, chunk_bits, _ip(outcomes)),
           "engine_run_trajectory")
    engine.n = n  # the call reset the engine to n qubits
    return outcomes


def run_plain_trajectory(engine: PlainEngine, n: int, gates, flags, uniforms, closed: bool = True,
                         block_targets: int = 4, mode: int = 0, chunk_bits: int = 10, grid_x: int = 0):
    """The try_cusv loop (cpp/include/mipt/circuit.hpp:240-282) on the plain
    engine, driven layer by layer from Python: per layer, blocked bond gates,
    then the measured sites.
    gates: complex128 [bonds,4,4]; flags/uniforms: [layers, n].  Returns outcomes [layers, n] (-1 unmeasured)."""
    engine.reset(n)
    layers = flags.shape[0]
    outcomes = np.full((layers, n), -1, np.int32)
    g = 0
    for t in range(layers):
        bonds = layer_bonds(n, t, closed, grid_x)
        engine.apply_bonds(bonds, gates[g:g + len(bonds)], block_targets)
        g += len(bonds)
        sites = np.nonzero(flags[t])[0]
        if len(sites):
            u = uniforms[t, sites] if mode == 0 else uniforms[t, :]
            o, _, _ = engine.measure_layer(sites, u, mode, chunk_bits)
            outcomes[t, sites] = o
    return outcomes


# ------------------------------------------------------------------ batched engine
class TrajSpec(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("layers", ctypes.c_int), ("p", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("closed", ctypes.c_int), ("gate_family", ctypes.c_int), ("probe_site", ctypes.c_int),
                ("p_zero_layers", ctypes.c_int), ("shared_gates", ctypes.c_int), ("grid_x", ctypes.c_int),
                ("probe_count", ctypes.c_int), ("probe_sites", ctypes.c_int * 4), ("attach_layer", ctypes.c_int),
                ("attach_stagger", ctypes.c_int), ("fp64_check", ctypes.c_int)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("tile_bits", ctypes.c_int), ("max_hist_bits", ctypes.c_int), ("threads", ctypes.c_int),
                ("bond_arith", ctypes.c_int), ("plan", ctypes.c_int)]


@dataclass
class Spec:
    L: int
    layers: int
    p: float
    seed: int
    closed: bool = True
    probe_site: int = -1
    p_zero_layers: int = 0
    shared_gates: int = -1
    grid_x: int = 0
    probe_sites: tuple = ()  # several references (qubits L, L+1, ...); probe_site stays -1
    attach_layer: int = -1   # with probe_sites: the attach layer (-1: p_zero_layers)
    attach_stagger: int = 0  # reference j attaches at attach_layer + j * attach_stagger
    fp64_check: bool = False  # run_batch states this result's errors against a complex128 replay

    def to_c(self) -> TrajSpec:
        sites = list(self.probe_sites) + [0] * (4 - len(self.probe_sites))
        return TrajSpec(self.L, self.layers, float(self.p), self.seed & (2**64 - 1), int(self.closed), 1,
                        self.probe_site, self.p_zero_layers, self.shared_gates, self.grid_x,
                        len(self.probe_sites), (ctypes.c_int * 4)(*sites[:4]), self.attach_layer, self.attach_stagger,
                        int(self.fp64_check))


class Batch:
    """Device-resident batch of trajectories (mixed chain lengths allowed)."""

    def __init__(self, specs, tile_bits: int = 0, stream: int | None = None, threads: int = 0, bond_arith: int = 0,
                 plan: int = 0):
        cfg = BatchConfig(tile_bits, 0, threads, bond_arith, plan)
        h = ctypes.c_void_p()
        _check(lib().mipt_hip_batch_create(ctypes.byref(cfg), ctypes.c_void_p(stream or 0), ctypes.byref(h)),
               "batch_create")
        self._h = h
        self.specs = list(specs)
        arr = (TrajSpec * len(self.specs))(*[s.to_c() for s in self.specs])
        _check(lib().mipt_hip_batch_setup(self._h, arr, len(self.specs)), "batch_setup")

    def close(self):
        if self._h:
            lib().mipt_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().mipt_hip_batch_stream(self._h) or 0

    def generate(self):
        _check(lib().mipt_hip_batch_generate(self._h), "batch_generate")

    def set_layers(self, traj: int, gates, flags, uniforms):
        g = np.ascontiguousarray(gates, np.complex128)
        f = np.ascontiguousarray(flags, np.uint8)
        u = np.ascontiguousarray(uniforms, np.float64)
        _check(lib().mipt_hip_batch_set_layers(self._h, traj, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(u)),
               "batch_set_layers")

    def nq(self, traj: int) -> int:
        s = self.specs[traj]
        return s.L + (len(s.probe_sites) or (1 if s.probe_site >= 0 else 0))

    def set_state(self, traj: int, psi):
        a = np.ascontiguousarray(psi, np.complex64)
        _check(lib().mipt_hip_batch_set_state(self._h, traj, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "batch_set_state")

    def get_state(self, traj: int) -> np.ndarray:
        out = np.zeros(1 << self.nq(traj), np.complex64)
        _check(lib().mipt_hip_batch_get_state(self._h, traj, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "batch_get_state")
        return out

    def state_ptr(self, traj: int) -> int:
        return lib().mipt_hip_batch_state_ptr(self._h, traj)

    def evolve(self, layer_begin: int | None = None, layer_end: int | None = None):
        if layer_begin is None:
            _check(lib().mipt_hip_batch_evolve(self._h), "batch_evolve")
        else:
            _check(lib().mipt_hip_batch_evolve_range(self._h, layer_begin, layer_end), "batch_evolve_range")

    def synchronize(self):
        _check(lib().mipt_hip_batch_synchronize(self._h), "batch_synchronize")

    def reset(self):
        """Re-arm in-kernel initial states so the same trajectories run again."""
        _check(lib().mipt_hip_batch_reset(self._h), "batch_reset")

    def export_layers(self, traj: int):
        """Device-drawn (gates complex128 [bonds,4,4], flags [layers,L], uniforms [layers,L])."""
        s = self.specs[traj]
        nb = 0
        for t in range(s.layers):
            nb += len(layer_bonds(s.L, t, s.closed, s.grid_x))
        g = np.zeros((nb, 4, 4), np.complex128)
        f = np.zeros((s.layers, s.L), np.uint8)
        u = np.zeros((s.layers, s.L), np.float64)
        _check(lib().mipt_hip_batch_export_layers(self._h, traj, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(u)),
               "batch_export_layers")
        return g, f, u

    def set_streams(self, streams: int):
        """1: every tiled group on the batch stream; 2 (default): two groups in flight."""
        _check(lib().mipt_hip_batch_set_streams(self._h, streams), "batch_set_streams")

    def set_profiling(self, sample_every: int):
        _check(lib().mipt_hip_batch_set_profiling(self._h, sample_every), "batch_set_profiling")

    def pass_stats(self):
        """(avg sweep ms over timed launches, timed count, mean algorithmic bytes, launches, total bytes)."""
        avg = ctypes.c_double()
        n = ctypes.c_int()
        b = ctypes.c_double()
        la = ctypes.c_long()
        bt = ctypes.c_double()
        _check(lib().mipt_hip_batch_pass_stats(self._h, ctypes.byref(avg), ctypes.byref(n), ctypes.byref(b),
                                               ctypes.byref(la), ctypes.byref(bt)), "batch_pass_stats")
        return {"avg_ms": avg.value, "timed": n.value, "bytes_avg": b.value, "launches": la.value,
                "bytes_total": bt.value}

    def outcomes(self, traj: int):
        s = self.specs[traj]
        o = np.zeros((s.layers, s.L), np.int32)
        f = np.zeros((s.layers, s.L), np.uint8)
        p1 = np.zeros((s.layers, s.L), np.float64)
        sp = ctypes.c_double()
        _check(lib().mipt_hip_batch_outcomes(self._h, traj, _ip(o), f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                             _dp(p1), ctypes.byref(sp)), "batch_outcomes")
        return o, f, p1, sp.value

    def observables(self, traj: int, tmi_block: int = 0, want_tmi: bool = True):
        sh = ctypes.c_double()
        tm = ctypes.c_double(float("nan"))
        pe = ctypes.c_double()
        bl = np.zeros(3)
        _check(lib().mipt_hip_batch_observables(self._h, traj, tmi_block, ctypes.byref(sh),
                                                ctypes.byref(tm) if want_tmi else None, ctypes.byref(pe), _dp(bl)),
               "batch_observables")
        return {"s_half": sh.value, "tmi": tm.value, "probe_entropy": pe.value, "bloch": bl}

    def records(self, layer_begin: int, layer_end: int) -> np.ndarray:
        """int8 [count, layers, L]: -1 unmeasured, else the outcome (same L for all)."""
        L = self.specs[0].L
        out = np.zeros((len(self.specs), layer_end - layer_begin, L), np.int8)
        _check(lib().mipt_hip_batch_records(self._h, layer_begin, layer_end,
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))), "batch_records")
        return out

    def observables_all(self, tmi_block: int = 0, tmi_cycles: int = 1):
        """(S(L/2) [count], TMI [count, tmi_cycles]) of every trajectory, bits,
        with the eigensolves of one RDM size batched."""
        n = len(self.specs)
        sh = np.zeros(n)
        tm = np.zeros((n, max(tmi_cycles, 1)))
        _check(lib().mipt_hip_batch_observables_all(self._h, tmi_block, tmi_cycles, _dp(sh),
                                                    _dp(tm) if tmi_cycles > 0 else None), "batch_observables_all")
        return sh, tm[:, :tmi_cycles]

    def rho3(self, subsystems, first: int = 0, count: int | None = None) -> np.ndarray:
        """Trace-normalised three-site RDMs [count, nsub, 8, 8] (complex128);
        subsystem qubit k = reduced-basis bit k."""
        count = len(self.specs) - first if count is None else count
        sub = np.ascontiguousarray(subsystems, np.int32).reshape(-1, 3)
        out = np.zeros((count, len(sub), 8, 8, 2))
        _check(lib().mipt_hip_batch_rho3(self._h, first, count, _ip(sub), len(sub), _dp(out)), "batch_rho3")
        return out[..., 0] + 1j * out[..., 1]

    def translation_entropies(self, min_kept: int, max_kept: int, max_s1_kept: int | None = None):
        """(S1, S2) [count, max_kept-min_kept+1] in nats, averaged over the L cyclic starts."""
        max_s1_kept = max_kept if max_s1_kept is None else max_s1_kept
        n = len(self.specs)
        s1 = np.zeros((n, max_kept - min_kept + 1))
        s2 = np.zeros_like(s1)
        _check(lib().mipt_hip_batch_translation_entropies(self._h, min_kept, max_kept, max_s1_kept, _dp(s1), _dp(s2)),
               "batch_translation_entropies")
        return s1, s2

    def autotune(self, probe_layers: int = 4, with_plan: bool = False):
        """Pick the fastest tile configuration, then layer plan variant, on this
        batch; returns (configuration index, ms per configuration) or, with
        with_plan, (configuration index, plan variant, ms [9])."""
        best, plan = ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(9)
        _check(lib().mipt_hip_batch_autotune(self._h, probe_layers, ctypes.byref(best), ctypes.byref(plan), _dp(ms)),
               "batch_autotune")
        if with_plan:
            return best.value, plan.value, ms
        return best.value, ms[:6]

    def set_graph(self, enable: bool = True):
        _check(lib().mipt_hip_batch_set_graph(self._h, int(enable)), "batch_set_graph")

    def probe_bloch(self) -> np.ndarray:
        """[count, 3] probe Bloch vectors (NaN rows without a probe)."""
        out = np.zeros((len(self.specs), 3))
        _check(lib().mipt_hip_batch_probe_bloch(self._h, _dp(out)), "batch_probe_bloch")
        return out

    def probe_rdm(self) -> np.ndarray:
        """[count, 2^k, 2^k] joint density matrices of the k references
        (reduced-basis bit j = reference j), trace 1."""
        k = len(self.specs[0].probe_sites) or (1 if self.specs[0].probe_site >= 0 else 0)
        D = 1 << k
        out = np.zeros((len(self.specs), D, D, 2))
        _check(lib().mipt_hip_batch_probe_rdm(self._h, _dp(out)), "batch_probe_rdm")
        return out[..., 0] + 1j * out[..., 1]

    def tmi(self, traj: int, cycle: int = 0, block: int = 0, order: int = 1) -> float:
        out = ctypes.c_double()
        _check(lib().mipt_hip_batch_tmi(self._h, traj, block, cycle, order, ctypes.byref(out)), "batch_tmi")
        return out.value


class TrajResult(ctypes.Structure):
    _fields_ = [("s_half", ctypes.c_double), ("probe_entropy", ctypes.c_double), ("bloch", ctypes.c_double * 3),
                ("norm_error", ctypes.c_double), ("measured", ctypes.c_int), ("flagged_draws", ctypes.c_int),
                ("amp_err", ctypes.c_double), ("s_half_err", ctypes.c_double), ("tmi_err", ctypes.c_double)]


PROBE_FIELDS = ("sq", "i2", "i3", "i4", "negativity", "log_negativity", "tmi", "average_mi",
                "min_bipartite_negativity", "joint_purity", "mean_single_purity")


def probe_information(rho) -> dict:
    """Information measures (nats) of a k-reference density matrix
    (mipt_hip_probe_information), keyed by the reference's sample fields."""
    rho = np.asarray(rho, np.complex128)
    k = int(round(np.log2(rho.shape[0])))
    ri = np.ascontiguousarray(np.stack([rho.real, rho.imag], -1))
    out = np.zeros(11)
    _check(lib().mipt_hip_probe_information(_dp(ri), k, _dp(out)), "probe_information")
    return dict(zip(PROBE_FIELDS, out.tolist()))


def run_batch(specs, tmi_cycles: int = 1, tmi_block: int = 0, entropy_order: int = 1):
    """Draw, evolve and read out `specs` in one library call (mipt_hip_run_batch).
    Returns (list of per-trajectory dicts, tmi array [count, tmi_cycles])."""
    specs = list(specs)
    arr = (TrajSpec * len(specs))(*[s.to_c() for s in specs])
    res = (TrajResult * len(specs))()
    tmi = np.zeros((len(specs), max(tmi_cycles, 1)))
    _check(lib().mipt_hip_run_batch(ctypes.cast(arr, ctypes.c_void_p), len(specs), tmi_block, tmi_cycles,
                                    entropy_order, _dp(tmi), ctypes.cast(res, ctypes.c_void_p), None), "run_batch")
    out = [{"s_half": r.s_half, "probe_entropy": r.probe_entropy, "bloch": np.array(r.bloch[:]),
            "norm_error": r.norm_error, "measured": r.measured, "flagged_draws": r.flagged_draws,
            "amp_err": r.amp_err, "s_half_err": r.s_half_err, "tmi_err": r.tmi_err} for r in res]
    return out, tmi[:, :tmi_cycles]


def mipt_main(argv) -> int:
    """The three-site RDM record (same positional arguments as the
    reference's mipt.exe), run in this process; returns its exit code."""
    args = [b"mipt"] + [str(a).encode() for a in argv]
    arr = (ctypes.c_char_p * len(args))(*args)
    return lib().mipt_hip_mipt_main(len(args), arr)


def entropy_main(argv) -> int:
    """entropy.exe's scan (N periods p realizations circ_type output.csv) in this process."""
    args = [b"entropy"] + [str(a).encode() for a in argv]
    arr = (ctypes.c_char_p * len(args))(*args)
    return lib().mipt_hip_entropy_main(len(args), arr)


def free_energy_main(argv) -> int:
    """free_energy.exe ([N] [p] [realizations] [circ_type] [init_state]) in
    this process; returns its exit code."""
    args = [b"free_energy"] + [str(a).encode() for a in argv]
    arr = (ctypes.c_char_p * len(args))(*args)
    return lib().mipt_hip_free_energy_main(len(args), arr)


def free_energy_initial_state(n: int, init_state: int, trajectory_seed: int) -> np.ndarray:
    """The initial state free_energy draws for a trajectory seed (complex128 [2^n])."""
    out = np.zeros((1 << n, 2))
    _check(lib().mipt_hip_free_energy_initial_state(n, init_state, trajectory_seed & (2**64 - 1), _dp(out)),
           "free_energy_initial_state")
    return out[:, 0] + 1j * out[:, 1]


def probed_main(argv) -> int:
    """mipt_probed.exe's modes 0 and 1 ([probes] [N] [realizations]
    [circ_type] [mode] [args]) in this process; returns its exit code."""
    args = [b"mipt_probed"] + [str(a).encode() for a in argv]
    arr = (ctypes.c_char_p * len(args))(*args)
    return lib().mipt_hip_probed_main(len(args), arr)


def sim_tmi_main(argv) -> int:
    """The sim_tmi scan (same positional arguments as the reference's
    sim_tmi.exe), run in this process; returns its exit code."""
    args = [b"sim_tmi"] + [str(a).encode() for a in argv]
    arr = (ctypes.c_char_p * len(args))(*args)
    return lib().mipt_hip_sim_tmi_main(len(args), arr)


__all__ = ["lib", "last_error", "MiptHipError", "measure_collapse", "entropy", "rdm_subsystem", "rho1", "tmi",
           "PlainEngine", "layer_bonds", "run_plain_trajectory", "Spec", "Batch", "run_batch", "sim_tmi_main", "mipt_main", "entropy_main"]
