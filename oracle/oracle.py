"""Python face of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker.  The product path
(``mipt_amd``) never imports it.

Evolution, measurement and the random streams come from ``liboracle.so``
(oracle/mipt_oracle.cpp, complex128).  The dense linear algebra of the
entropy side (RDM of a kept set, Hermitian eigenvalues, log2 entropies, TMI)
is restated here with numpy, following (paths relative to the reference
repository root):

* cpp/include/mipt/tmi/linalg.hpp:885-1026  basis offsets / Schmidt gather
* cpp/include/mipt/tmi/linalg.hpp:125-141   entropy cut-offs (lambda<=1e-15, |S|<1e-12)
* cpp/include/mipt/tmi/compute.hpp:64-113, 186-227  TMI regions and formula
* cpp/include/mipt/free_energy.hpp:274-287  S(L/2): kept = {0 .. n/2-1}
* cpp/include/mipt/ancilla_mi.hpp:24-52     probe entropy (nats)
* cpp/include/mipt/probed/kernels.hpp:55-147, ancilla_info.hpp:46-139,
  ancilla_mi.hpp:222-262, analysis/cut_negativity.hpp:139-200
                                            several references: attach, joint
                                            RDM and information measures
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/libref.so when the reference is present)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_philox_haar.restype = None
        L.orc_philox_haar.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _dp]
        L.orc_philox_uniform.restype = ctypes.c_double
        L.orc_philox_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32]
        L.orc_trajectory_seed.restype = ctypes.c_uint64
        L.orc_trajectory_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_linspace.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_layer_bonds.restype = ctypes.c_int
        L.orc_layer_bonds.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_total_bonds.restype = ctypes.c_int
        L.orc_total_bonds.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("orc_mt_build_layers", "orc_philox_build_layers"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint64, _dp, _u8p, _dp]
        L.orc_apply_dense.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, _dp]
        L.orc_measure_collapse.restype = ctypes.c_int
        L.orc_measure_collapse.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, _dp]
        L.orc_measure_layer_joint.restype = ctypes.c_double
        L.orc_measure_layer_joint.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip]
        L.orc_run_layers.restype = ctypes.c_int
        L.orc_run_layers.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _u8p, _dp,
                                     ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.orc_rdm.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, _dp]
        L.orc_jacobi_eigvalsh.restype = ctypes.c_int
        L.orc_jacobi_eigvalsh.argtypes = [_dp, ctypes.c_int, _dp]
        L.orc_entropy_bits_from_eigs.restype = ctypes.c_double
        L.orc_entropy_bits_from_eigs.argtypes = [_dp, ctypes.c_int]
        L.orc_rho1.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_rho1_entropy_nats.restype = ctypes.c_double
        L.orc_rho1_entropy_nats.argtypes = [_dp]
        _LIB = L
    return _LIB


def ref_available() -> bool:
    """Whether oracle/_ref/libref.so has been built (checked without loading it)."""
    return os.path.exists(os.path.join(_HERE, "_ref", "libref.so"))


def ref_lib():
    """The reference's own host code compiled into oracle/_ref (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        u32p, u64p, cp = _u32p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p
        R.ref_density_header.restype = ctypes.c_int
        R.ref_density_header.argtypes = [cp, u32p, u64p, _dp, ctypes.c_char_p, ctypes.c_int]
        R.ref_density_records.restype = ctypes.c_int
        R.ref_density_records.argtypes = [cp, u32p, u32p, u32p, _dp, _dp, ctypes.c_char_p, ctypes.c_int]
        R.ref_tmi_resume_load.restype = ctypes.c_int
        R.ref_tmi_resume_load.argtypes = [cp, cp, ctypes.c_int, ctypes.c_int, cp, u64p, _ip, ctypes.c_char_p,
                                          ctypes.c_int]
        R.ref_measure_collapse_host.restype = ctypes.c_int
        R.ref_measure_collapse_host.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, _dp, _dp]
        R.ref_entropy_subsystem_rdm.restype = ctypes.c_double
        R.ref_entropy_subsystem_rdm.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int]
        R.ref_entropy_hermitian.restype = ctypes.c_double
        R.ref_entropy_hermitian.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        R.ref_entropy_from_eigenvalue.restype = ctypes.c_double
        R.ref_entropy_from_eigenvalue.argtypes = [ctypes.c_double]
        R.ref_finish_entropy.restype = ctypes.c_double
        R.ref_finish_entropy.argtypes = [ctypes.c_double]
        R.ref_hermitian_eigenvalues.restype = ctypes.c_int
        R.ref_hermitian_eigenvalues.argtypes = [_dp, ctypes.c_int, _dp]
        R.ref_splitmix64.restype = ctypes.c_uint64
        R.ref_splitmix64.argtypes = [ctypes.c_uint64]
        R.ref_trajectory_seed.restype = ctypes.c_uint64
        R.ref_trajectory_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        R.ref_linspace.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        R.ref_tensor_product.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp]
        _REF = R
    return _REF


def ref_density_read(path: str) -> dict:
    """A MIPTRHO2 v2 file read by the reference's own DensityMatrixReader
    (cpp/include/mipt/density.hpp:253-347).  Raises ValueError with the
    reader's message when it rejects the file."""
    R = ref_lib()
    err = ctypes.create_string_buffer(512)
    u = np.zeros(10, np.uint32)
    rc = ctypes.c_uint64(0)
    pr = np.zeros(2)
    if R.ref_density_header(path.encode(), u.ctypes.data_as(_u32p), ctypes.byref(rc), _d(pr), err, 512) != 0:
        raise ValueError(err.value.decode())
    keys = ["spatial_dimension", "total_qubits", "kept_qubits", "matrix_dimension", "periods", "realizations",
            "resolution", "grid_x", "grid_y", "subsystem_count"]
    meta = {k: int(v) for k, v in zip(keys, u)}
    meta["record_count"] = int(rc.value)
    meta["p_min"], meta["p_max"] = float(pr[0]), float(pr[1])
    nrec, nsub, kept, dim = meta["record_count"], meta["subsystem_count"], meta["kept_qubits"], meta["matrix_dimension"]
    subs = np.zeros(max(nsub * kept, 1), np.uint32)
    pidx = np.zeros(max(nrec, 1), np.uint32)
    real = np.zeros(max(nrec, 1), np.uint32)
    ps = np.zeros(max(nrec, 1))
    rho = np.zeros(max(nrec * nsub * 2 * dim * dim, 1))
    if R.ref_density_records(path.encode(), subs.ctypes.data_as(_u32p), pidx.ctypes.data_as(_u32p),
                             real.ctypes.data_as(_u32p), _d(ps), _d(rho), err, 512) != 0:
        raise ValueError(err.value.decode())
    meta["subsystems"] = subs[:nsub * kept].reshape(nsub, kept)
    meta["p_index"], meta["realization"], meta["p"] = pidx[:nrec], real[:nrec], ps[:nrec]
    r = rho[:nrec * nsub * 2 * dim * dim].reshape(nrec, nsub, dim, dim, 2)
    meta["rho"] = r[..., 0] + 1j * r[..., 1]
    return meta


def ref_tmi_resume_load(path: str, expected_name: str, realizations: int, cycles: int, p_texts) -> dict:
    """The reference's sim_tmi checkpoint scan (cpp/include/mipt/tmi/resume.hpp:205-333)
    of an existing CSV.  Raises ValueError("Resume refused: ...") when it refuses."""
    R = ref_lib()
    err = ctypes.create_string_buffer(1024)
    out = np.zeros(5, np.uint64)
    fl = np.zeros(2, np.int32)
    if R.ref_tmi_resume_load(path.encode(), expected_name.encode(), int(realizations), int(cycles),
                             "\n".join(p_texts).encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                             _i(fl), err, 1024) != 0:
        raise ValueError(err.value.decode())
    keys = ["rows", "dropped_rows", "kept_bytes", "start_p_index", "start_realization"]
    d = {k: int(v) for k, v in zip(keys, out)}
    d["name_verified"], d["loaded"] = bool(fl[0]), bool(fl[1])
    return d


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


# ----------------------------------------------------------------- streams

def layer_bonds(n: int, layer: int, closed: bool = True):
    q0 = np.zeros(n, np.int32)
    q1 = np.zeros(n, np.int32)
    nb = lib().orc_layer_bonds(n, layer, int(closed), _i(q0), _i(q1))
    return [(int(q0[b]), int(q1[b])) for b in range(nb)]


def total_bonds(n: int, layers: int, closed: bool = True) -> int:
    return lib().orc_total_bonds(n, layers, int(closed))


def build_layers(n: int, layers: int, p: float, seed: int, closed: bool = True, stream: str = "philox"):
    """Gates (complex128 [bonds,4,4]), flags (uint8 [layers,n]), uniforms ([layers,n]).

    stream "philox": the counter-based stream the batched engine draws on device.
    stream "mt": the reference-compatible mt19937 host stream (circuit.hpp:129-140).
    """
    nb = total_bonds(n, layers, closed)
    gates = np.zeros((nb, 4, 4), np.complex128)
    flags = np.zeros((layers, n), np.uint8)
    unif = np.zeros((layers, n), np.float64)
    f = lib().orc_philox_build_layers if stream == "philox" else lib().orc_mt_build_layers
    f(n, layers, float(p), int(closed), ctypes.c_uint64(seed & (2**64 - 1)),
      gates.ctypes.data_as(_dp), flags.ctypes.data_as(_u8p), _d(unif))
    return gates, flags, unif


def philox4x32_10(key, ctr):
    k = np.asarray(key, np.uint32)
    c = np.asarray(ctr, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(k.ctypes.data_as(_u32p), c.ctypes.data_as(_u32p), out.ctypes.data_as(_u32p))
    return [int(x) for x in out]


# ----------------------------------------------------------------- evolution

def zero_state(n: int) -> np.ndarray:
    psi = np.zeros(1 << n, np.complex128)
    psi[0] = 1.0
    return psi


def apply_dense(psi: np.ndarray, n: int, targets, m: np.ndarray) -> None:
    t = np.asarray(targets, np.int32)
    mm = np.ascontiguousarray(m, np.complex128)
    lib().orc_apply_dense(psi.ctypes.data_as(_dp), n, _i(t), len(t), mm.ctypes.data_as(_dp))


def measure_collapse(psi: np.ndarray, n: int, q: int, u: float):
    o = ctypes.c_int(0)
    pr = ctypes.c_double(0.0)
    st = lib().orc_measure_collapse(psi.ctypes.data_as(_dp), n, q, float(u), ctypes.byref(o), ctypes.byref(pr))
    return st, o.value, pr.value


def run_layers(psi: np.ndarray, n: int, gates, flags, unif, layer_begin: int = 0, layer_end: int | None = None,
               closed: bool = True, mode: int = 0, chunk_bits: int = 16):
    """Evolve psi in place; returns (outcomes[layers,n] (-1 unmeasured), p1[layers,n])."""
    layers = flags.shape[0]
    if layer_end is None:
        layer_end = layers
    outcomes = np.full((layers, n), -1, np.int32)
    p1 = np.full((layers, n), -1.0, np.float64)
    g = np.ascontiguousarray(gates, np.complex128)
    fl = np.ascontiguousarray(flags, np.uint8)
    un = np.ascontiguousarray(unif, np.float64)
    st = lib().orc_run_layers(psi.ctypes.data_as(_dp), n, layer_begin, layer_end, int(closed), g.ctypes.data_as(_dp),
                              fl.ctypes.data_as(_u8p), _d(un), mode, chunk_bits, _i(outcomes), _d(p1))
    if st != 0:
        raise RuntimeError(f"oracle run_layers status {st}")
    return outcomes, p1


def simulate(n: int, layers: int, p: float, seed: int, closed: bool = True, stream: str = "philox", mode: int = 0):
    gates, flags, unif = build_layers(n, layers, p, seed, closed, stream)
    psi = zero_state(n)
    outcomes, p1 = run_layers(psi, n, gates, flags, unif, closed=closed, mode=mode)
    return psi, outcomes, p1, (gates, flags, unif)


# ----------------------------------------------------------------- entropies

LOG2_EPS = 1.0e-15


def schmidt_matrix(psi: np.ndarray, n: int, kept) -> np.ndarray:
    """M[r, c] = psi[env_off[c] | row_off[r]]: bit j of r is kept[j], env ascending."""
    kept = list(kept)
    env = [q for q in range(n) if q not in kept]
    t = psi.reshape([2] * n)  # axis a <-> qubit n-1-a
    axes = [n - 1 - q for q in reversed(kept)] + [n - 1 - q for q in reversed(env)]
    return np.ascontiguousarray(np.transpose(t, axes)).reshape(1 << len(kept), 1 << len(env))


def rdm(psi: np.ndarray, n: int, kept) -> np.ndarray:
    m = schmidt_matrix(psi, n, kept)
    return m @ m.conj().T


def entropy_from_eigs(w) -> float:
    w = np.asarray(w, np.float64)
    w = w[w > LOG2_EPS]
    s = float(-(w * np.log2(w)).sum())
    return 0.0 if abs(s) < 1.0e-12 else s


def renyi2_from_purity(purity: float) -> float:
    if not purity > 0.0 or purity >= 1.0:
        return 0.0
    s = -np.log2(purity)
    return 0.0 if abs(s) < 1.0e-12 else float(s)


def entropy(psi: np.ndarray, n: int, kept, order: int = 1) -> float:
    if len(kept) == 0 or len(kept) == n:
        return 0.0
    rho = rdm(psi, n, kept)
    if order == 2:
        return renyi2_from_purity(float((np.abs(rho) ** 2).sum()))
    rho = 0.5 * (rho + rho.conj().T)
    return entropy_from_eigs(np.linalg.eigvalsh(rho))


def cyclic_range(begin: int, count: int, n: int):
    return [(begin + i) % n for i in range(count)]


def tmi_regions(n: int, block: int, cycle: int = 0):
    """A, B, C, D (compute.hpp:89-113); D is the rest of the ring."""
    A = cyclic_range(cycle, block, n)
    B = cyclic_range(cycle + block, block, n)
    C = cyclic_range(cycle + 2 * block, block, n)
    used = set(A + B + C)
    D = [(cycle + 3 * block + i) % n for i in range(n - 3 * block)]
    assert not used.intersection(D)
    return A, B, C, D


def tmi(psi: np.ndarray, n: int, block: int | None = None, cycle: int = 0, order: int = 1) -> float:
    """I3 = S_A+S_B+S_C+S_D-S_AB-S_AC-S_BC in bits (compute.hpp:186-227)."""
    if block is None:
        block = n // 4
    A, B, C, D = tmi_regions(n, block, cycle)
    s = lambda k: entropy(psi, n, k, order)
    return s(A) + s(B) + s(C) + s(D) - s(A + B) - s(A + C) - s(B + C)


def half_chain_entropy(psi: np.ndarray, n: int, order: int = 1) -> float:
    return entropy(psi, n, list(range(n // 2)), order)


def rho1(psi: np.ndarray, nq: int, qubit: int):
    out = np.zeros(4, np.float64)
    lib().orc_rho1(psi.ctypes.data_as(_dp), nq, qubit, _d(out))
    return out


def rho1_entropy_nats(r) -> float:
    r = np.ascontiguousarray(r, np.float64)
    return lib().orc_rho1_entropy_nats(_d(r))


def bloch_from_rho1(r):
    tr = r[0] + r[1]
    return np.array([2 * r[2] / tr, -2 * r[3] / tr, (r[0] - r[1]) / tr])


def jacobi_eigvalsh(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.complex128)
    n = a.shape[0]
    w = np.zeros(n)
    st = lib().orc_jacobi_eigvalsh(a.ctypes.data_as(_dp), n, _d(w))
    if st != 0:
        raise RuntimeError("oracle Jacobi did not converge")
    return w


# ------------------------------------------------------------- probe protocol
STREAM_ATCH = 0x41544348  # probe-attach reset uniforms


def philox_uniform(seed: int, layer: int, item: int, slot: int, stream: int) -> float:
    return lib().orc_philox_uniform(seed & (2**64 - 1), layer, item, slot, stream)


def attach_probe(psi: np.ndarray, n: int, site: int, u: float) -> int:
    """reset(site); h(site); cx(site, probe) with the probe at qubit n in |0>
    (cpp/include/mipt/probed/kernels.hpp:68-77).  The reset is a sequential
    Born draw on the site (outcome 1 iff u < P1, renormalised), then the site
    is flipped to 0.  In place on the (n+1)-qubit state; returns the outcome."""
    nq = n + 1
    _, o, _ = measure_collapse(psi, nq, site, u)
    S, P = 1 << site, 1 << n
    idx = np.arange(1 << nq)
    base = idx[((idx & S) == 0) & ((idx & P) == 0)]
    a = psi[base | (S if o else 0)].copy() / np.sqrt(2.0)
    psi[base | S] = 0.0
    psi[base | P] = 0.0
    psi[base] = a
    psi[base | S | P] = a
    return o


def probe_trajectory(n: int, layers: int, p: float, seed: int, probe_site: int, p_zero_layers: int,
                     gate_seed: int | None = None, closed: bool = True):
    """The probe protocol the batched engine runs for a spec with a probe:
    unmeasured layers 0..p_zero_layers-1, the probe attached (reset + Bell
    pair) at the start of layer p_zero_layers (or a Bell pair at t=0 when
    p_zero_layers = 0), measured layers after it.  Gates from gate_seed's
    stream when given (a shared realisation), flags and uniforms from seed.
    Returns (psi over n+1 qubits, outcomes [layers,n], p1, attach outcome)."""
    g, f, u = build_layers(n, layers, p, seed, closed)
    if gate_seed is not None:
        g = build_layers(n, layers, 0.0, gate_seed, closed)[0]
    f[:p_zero_layers] = 0
    u[:p_zero_layers] = 0.0
    nq = n + 1
    psi = np.zeros(1 << nq, np.complex128)
    att = -1
    if p_zero_layers == 0:
        psi[0] = psi[(1 << probe_site) | (1 << n)] = 1 / np.sqrt(2.0)
    else:
        psi[0] = 1.0
    outs = np.full((layers, n), -1, np.int32)
    p1 = np.zeros((layers, n))
    gi = 0
    for t in range(layers):
        if p_zero_layers > 0 and t == p_zero_layers:
            att = attach_probe(psi, n, probe_site, philox_uniform(seed, t, probe_site, 0, STREAM_ATCH))
        for (q0, q1) in layer_bonds(n, t, closed):
            apply_dense(psi, nq, [q0, q1], g[gi])
            gi += 1
        for s in range(n):
            if f[t, s]:
                st, o, pr = measure_collapse(psi, nq, s, u[t, s])
                outs[t, s] = o
                p1[t, s] = pr if o == 1 else 1.0 - pr
    return psi, outs, p1, att, (g, f, u)


# --------------------------------------------------------- several references
def attach_reference(psi: np.ndarray, nq: int, site: int, ref: int, u: float) -> int:
    """reset(site); h(site); cx(site, ref) on an nq-qubit state whose
    reference qubit `ref` is |0> (cpp/include/mipt/probed/kernels.hpp:55-77):
    a sequential Born draw on the site (outcome 1 iff u < P1, renormalised),
    the site flipped to 0, then the Bell pair.  In place; returns the outcome."""
    _, o, _ = measure_collapse(psi, nq, site, u)
    S, P = 1 << site, 1 << ref
    idx = np.arange(1 << nq)
    base = idx[((idx & S) == 0) & ((idx & P) == 0)]
    a = psi[base | (S if o else 0)].copy() / np.sqrt(2.0)
    psi[base | S] = 0.0
    psi[base | P] = 0.0
    psi[base] = a
    psi[base | S | P] = a
    return o


def references_trajectory(n: int, layers: int, p: float, seed: int, sites, p_zero_layers: int,
                          closed: bool = True, attach_layer: int | None = None, attach_stagger: int = 0):
    """k = len(sites) references (qubits n .. n+k-1); layers 0..p_zero_layers-1
    unmeasured.  Bell pairs (site_j, n+j) at t = 0 when the attach layer
    (attach_layer, default p_zero_layers) is 0, else the references attached
    in order j = 0, 1, ... at the start of that layer, after the previous
    layer's measurements (cpp/include/mipt/probed/kernels.hpp:55-66, each
    reset's uniform keyed by (layer, site)); with attach_stagger, reference j
    attaches at attach_layer + j * attach_stagger (runner.hpp:230-256).
    Returns (psi over n+k qubits,
    outcomes [layers, n], p1, attach outcomes, (g, f, u))."""
    k = len(sites)
    attach = p_zero_layers if attach_layer is None else attach_layer
    g, f, u = build_layers(n, layers, p, seed, closed)
    f[:p_zero_layers] = 0
    u[:p_zero_layers] = 0.0
    nq = n + k
    psi = np.zeros(1 << nq, np.complex128)
    att = []
    at = [attach + j * attach_stagger for j in range(k)]
    paired = [j for j in range(k) if at[j] == 0]
    for b in range(1 << len(paired)):
        i = 0
        for bit, j in enumerate(paired):
            if (b >> bit) & 1:
                i |= (1 << sites[j]) | (1 << (n + j))
        psi[i] = 2.0 ** (-len(paired) / 2)
    outs = np.full((layers, n), -1, np.int32)
    p1 = np.zeros((layers, n))
    gi = 0
    for t in range(layers):
        for j, site in enumerate(sites):
            if t > 0 and at[j] == t:
                att.append(attach_reference(psi, nq, site, n + j, philox_uniform(seed, t, site, 0, STREAM_ATCH)))
        for (q0, q1) in layer_bonds(n, t, closed):
            apply_dense(psi, nq, [q0, q1], g[gi])
            gi += 1
        for s in range(n):
            if f[t, s]:
                st, o, pr = measure_collapse(psi, nq, s, u[t, s])
                outs[t, s] = o
                p1[t, s] = pr if o == 1 else 1.0 - pr
    return psi, outs, p1, att, (g, f, u)


def reference_rdm(psi: np.ndarray, n: int, k: int) -> np.ndarray:
    """Joint density matrix of qubits n..n+k-1 (reduced-basis bit j = qubit
    n+j), trace 1: the state index is r 2^n + e with r the references' bits."""
    m = np.asarray(psi).reshape(1 << k, 1 << n)
    rho = m @ m.conj().T
    return rho / np.trace(rho).real


def partial_trace_bits(rho: np.ndarray, k: int, keep) -> np.ndarray:
    """Reduced matrix of the bits in `keep` (bit j of the index = party j; the
    result's bit i = keep[i] for keep ascending), trace 1."""
    keep = sorted(keep)
    rows = "abcd"[:k]
    cols = ["efgh"[j] if j in keep else rows[j] for j in range(k)]
    # tensor axis a of a 2^k index is bit k-1-a
    src = "".join(rows[k - 1 - a] for a in range(k)) + "".join(cols[k - 1 - a] for a in range(k))
    dst = "".join(rows[j] for j in reversed(keep)) + "".join(cols[j] for j in reversed(keep))
    out = np.einsum(src + "->" + dst, np.asarray(rho).reshape([2] * (2 * k)))
    d = 1 << len(keep)
    out = out.reshape(d, d)
    return out / np.trace(out).real


def _entropy_nats(rho) -> float:
    w = np.clip(np.linalg.eigvalsh(0.5 * (rho + rho.conj().T)), 0.0, None)
    w = w[w > 0]
    return float(-(w * np.log(w)).sum())


def _purity(rho) -> float:
    return float((np.abs(rho) ** 2).sum() / np.trace(rho).real ** 2)


def cut_negativity(rho: np.ndarray, k: int, j: int) -> float:
    """0.5 (||rho^{T_j}||_1 - 1), partial transpose of party j
    (cpp/include/mipt/analysis/cut_negativity.hpp:139-200, qubits)."""
    t = np.asarray(rho).reshape([2] * (2 * k))
    a = k - 1 - j
    pt = np.swapaxes(t, a, k + a).reshape(1 << k, 1 << k)
    w = np.linalg.eigvalsh(0.5 * (pt + pt.conj().T))
    v = 0.5 * (np.abs(w).sum() - 1.0)
    return 0.0 if -1e-10 < v < 0 else float(v)


def probe_information(rho: np.ndarray) -> dict:
    """The k-reference sample of cpp/include/mipt/probed/kernels.hpp:106-147
    (nats): k=1 sq; k=2 sq, i2 = I(A:B), negativity, log_negativity
    (ancilla_mi.hpp:222-262); k=3 tmi, average_mi, min_bipartite_negativity,
    joint_purity, mean_single_purity (ancilla_info.hpp:46-66); k=4 sq, mean
    I2 over pairs, mean I3 over triples, I4 (ancilla_info.hpp:75-139)."""
    rho = np.asarray(rho, np.complex128)
    k = int(round(np.log2(rho.shape[0])))
    S = {}
    for mask in range(1, 1 << k):
        S[mask] = _entropy_nats(partial_trace_bits(rho, k, [j for j in range(k) if (mask >> j) & 1]))
    out = dict.fromkeys(("sq", "i2", "i3", "i4", "negativity", "log_negativity", "tmi", "average_mi",
                         "min_bipartite_negativity", "joint_purity", "mean_single_purity"), 0.0)
    full = rho / np.trace(rho).real
    if k == 1:
        out["sq"] = S[1]
    elif k == 2:
        out["sq"] = S[3]
        out["i2"] = max(0.0, S[1] + S[2] - S[3])
        out["negativity"] = max(0.0, cut_negativity(full, 2, 0))
        out["log_negativity"] = float(np.log1p(2 * out["negativity"]))
    elif k == 3:
        out["tmi"] = S[1] + S[2] + S[4] - S[3] - S[5] - S[6] + S[7]
        out["average_mi"] = ((S[1] + S[2] - S[3]) + (S[1] + S[4] - S[5]) + (S[2] + S[4] - S[6])) / 3
        out["min_bipartite_negativity"] = max(0.0, min(cut_negativity(full, 3, j) for j in range(3)))
        out["joint_purity"] = _purity(full)
        out["mean_single_purity"] = np.mean([_purity(partial_trace_bits(full, 3, [j])) for j in range(3)])
    else:
        bits = lambda m: [j for j in range(4) if (m >> j) & 1]
        i2 = [S[1 << a] + S[1 << b] - S[m] for m in range(16) if len(bits(m)) == 2 for a, b in [bits(m)]]
        i3 = []
        for m in range(16):
            if len(bits(m)) == 3:
                a, b, c = bits(m)
                i3.append(S[1 << a] + S[1 << b] + S[1 << c] - S[(1 << a) | (1 << b)] - S[(1 << a) | (1 << c)]
                          - S[(1 << b) | (1 << c)] + S[m])
        out["sq"] = S[15]
        out["i2"] = float(np.mean(i2))
        out["i3"] = float(np.mean(i3))
        out["i4"] = sum((1 if len(bits(m)) % 2 else -1) * S[m] for m in range(1, 16))
    return out


# ------------------------------------------------------------------ 2D grid
STREAM_FLAG = 0x464C4147
STREAM_MEAS = 0x4D454153


def grid_layer_bonds(gx: int, gy: int, layer: int, closed: bool = False):
    """Bonds of one colour layer of a gx x gy grid (site r*gx + c), in
    gate-stream order: colour 0/1 = horizontal bonds from even/odd columns,
    row by row (colour 1 closes each row with (r*gx+gx-1, r*gx) when closed
    and gx is even); colour 2/3 = vertical bonds, column by column, from rows
    ((c % 2) + colour - 2) % 2 in steps of 2 (the staggered four-colour
    layering of cpp/include/mipt/circuits/mms.hpp:159-330)."""
    colour = layer % 4
    bonds = []
    if colour in (0, 1):
        for r in range(gy):
            c = colour
            while c + 1 < gx:
                bonds.append((r * gx + c, r * gx + c + 1))
                c += 2
            if colour == 1 and closed and gx % 2 == 0 and gx > 2:
                bonds.append((r * gx + gx - 1, r * gx))
    else:
        for c in range(gx):
            r = ((c % 2) + colour - 2) % 2
            while r + 1 < gy:
                bonds.append((r * gx + c, (r + 1) * gx + c))
                r += 2
    return bonds


def simulate_grid(gx: int, gy: int, layers: int, p: float, seed: int, closed: bool = False):
    """One 2D trajectory from |0...0> on the counter stream: gates
    orc_philox_haar(seed, layer, bond), flags Bernoulli(p) and uniforms per
    (layer, site) as for the chain; each colour layer is followed by the
    sequential measurement of its flagged sites (ascending).
    Returns (psi, outcomes [layers, n], p1 [layers, n], flags, uniforms)."""
    n = gx * gy
    psi = zero_state(n)
    outs = np.full((layers, n), -1, np.int32)
    p1 = np.zeros((layers, n))
    flags = np.zeros((layers, n), np.uint8)
    unif = np.zeros((layers, n))
    g = np.zeros(32)
    for t in range(layers):
        for b, (q0, q1) in enumerate(grid_layer_bonds(gx, gy, t, closed)):
            lib().orc_philox_haar(seed & (2**64 - 1), t, b, g.ctypes.data_as(_dp))
            apply_dense(psi, n, [q0, q1], (g[0::2] + 1j * g[1::2]).reshape(4, 4))
        for s in range(n):
            a = philox_uniform(seed, t, s, 0, STREAM_FLAG)
            f = p >= 1.0 or (0.0 < p and a < p)
            if f:
                flags[t, s] = 1
                unif[t, s] = philox_uniform(seed, t, s, 0, STREAM_MEAS)
                st, o, pr = measure_collapse(psi, n, s, unif[t, s])
                outs[t, s] = o
                p1[t, s] = pr if o == 1 else 1.0 - pr
    return psi, outs, p1, flags, unif
