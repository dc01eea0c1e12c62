"""GPU tests of the library's batched FP64 Hermitian eigensolver (eigen.hip).

The kernels replace the cuSOLVER heevd/gesvd ladder of the reference's device
entropy (cpp/src/cuda/tmi_cuda_svd.cu:663-798) and the host Hermitian entropy
(cpp/include/mipt/tmi/linalg.hpp:524-583) for batches of RDMs.

Pins:
* known spectra: rho = U diag(lambda) U^H with a Haar U (FP64), including
  fully degenerate (maximally mixed), rank-1, rank-k with exact zeros,
  geometric decay down to 1e-14, a rank-16 Kronecker product (the reduction
  runs into pure rounding noise past the rank, as on early-time circuit
  RDMs), diagonal input (every reflector is the identity) and a 2 x 2 with
  a complex off-diagonal; eigenvalues within
  4 m eps of lambda (the backward error of a Householder reduction) and of
  numpy's LAPACK eigvalsh on the same matrix; entropy within 1e-9 bits of
  -sum lambda log2 lambda with the reference's cut-offs;
* the state path: S and TMI of batch observables and single drop-in
  entropies equal the rocSOLVER path to 1e-10 bits and the FP64 oracle to
  1e-5 bits;
* repeat calls are bit-identical (fixed-order sums).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    yield mipt_amd
    mipt_amd.set_eigensolver(0)


def haar(m, rng):
    z = (rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def spectra(m, rng):
    """Named test spectra of trace 1 for an m x m density matrix."""
    out = {"dirichlet": rng.dirichlet(np.ones(m)), "mixed": np.full(m, 1.0 / m)}
    pure = np.zeros(m)
    pure[0] = 1.0
    out["pure"] = pure
    if m >= 4:
        k = m // 4
        rk = np.zeros(m)
        rk[:k] = 1.0 / k
        out["rank_quarter"] = rk
        g = 10.0 ** (-14.0 * np.arange(m) / (m - 1))
        out["decay_1e-14"] = g / g.sum()
    return out


def entropy_bits(lam):
    lam = lam[lam > 1e-15]
    s = float(-(lam * np.log2(lam)).sum())
    return 0.0 if abs(s) < 1e-12 else s


@pytest.mark.parametrize("m", [1, 2, 3, 5, 16, 31, 32, 33, 64, 100, 257, 512, 1024])
def test_known_spectra(M, m):
    rng = np.random.default_rng(100 + m)
    names, mats, lams = [], [], []
    for name, lam in spectra(m, rng).items():
        u = haar(m, rng)
        rho = (u * lam) @ u.conj().T
        rho = 0.5 * (rho + rho.conj().T)
        names.append(name)
        mats.append(rho)
        lams.append(np.sort(lam))
    if m >= 64 and m % 16 == 0:
        # rank 16 with a tensor-product structure, like an early-time circuit
        # RDM: past the rank the reduction only sees rounding noise
        a = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
        sig = a @ a.conj().T
        sig /= np.trace(sig).real
        v = rng.normal(size=m // 16) + 1j * rng.normal(size=m // 16)
        v /= np.linalg.norm(v)
        rho = np.kron(sig, np.outer(v, v.conj()))
        names.append("kron_rank16")
        mats.append(rho)
        lams.append(np.sort(np.concatenate([np.linalg.eigvalsh(sig), np.zeros(m - 16)])))
    # diagonal input: the reduction's reflectors are all the identity
    d = rng.dirichlet(np.ones(m))
    names.append("diagonal")
    mats.append(np.diag(d).astype(np.complex128))
    lams.append(np.sort(d))
    ent, ev = M.entropy_hermitian_batched(np.stack(mats))
    tol = 4 * m * EPS
    for i, name in enumerate(names):
        ref = np.linalg.eigvalsh(mats[i])
        assert np.abs(ev[i] - lams[i]).max() <= tol + 4 * EPS, f"m={m} {name}: eigenvalues vs spectrum"
        assert np.abs(ev[i] - ref).max() <= tol + 4 * EPS, f"m={m} {name}: eigenvalues vs numpy"
        assert abs(ent[i] - entropy_bits(lams[i])) < 1e-9, f"m={m} {name}: entropy {ent[i]} vs {entropy_bits(lams[i])}"


def test_two_by_two_complex_offdiagonal(M):
    # the last reflector of every reduction sees a complex alpha with no x
    rho = np.array([[0.7, 0.2 - 0.3j], [0.2 + 0.3j, 0.3]])
    ent, ev = M.entropy_hermitian_batched(rho[None])
    ref = np.linalg.eigvalsh(rho)
    assert np.abs(ev[0] - ref).max() < 1e-15
    assert abs(ent[0] - entropy_bits(ref)) < 1e-14


def test_repeat_calls_bit_identical(M):
    rng = np.random.default_rng(7)
    mats = []
    for _ in range(6):
        z = rng.normal(size=(256, 300)) + 1j * rng.normal(size=(256, 300))
        r = z @ z.conj().T
        mats.append(r / np.trace(r).real)
    a, la = M.entropy_hermitian_batched(np.stack(mats))
    b, lb = M.entropy_hermitian_batched(np.stack(mats))
    assert np.array_equal(a, b) and np.array_equal(la, lb)


def _batch(M, L, layers, traj, seed0):
    specs = [M.Spec(L, layers, 0.05 + 0.05 * (i % 4), seed0 + i, True) for i in range(traj)]
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    return b


@pytest.mark.parametrize("L", [12, 16, 20])
def test_batch_observables_match_rocsolver_and_oracle(M, L):
    b = _batch(M, L, 8, 4, 4242)
    try:
        M.set_eigensolver(0)
        sh0, tm0 = b.observables_all(L // 4, L // 4)
        M.set_eigensolver(1)
        sh1, tm1 = b.observables_all(L // 4, L // 4)
        M.set_eigensolver(0)
        assert np.abs(sh0 - sh1).max() < 1e-10
        assert np.abs(tm0 - tm1).max() < 1e-10
        for i in range(4):
            psi = b.get_state(i).astype(np.complex128)
            assert abs(sh0[i] - O.half_chain_entropy(psi, L)) < 1e-5
            for c in range(L // 4):
                assert abs(tm0[i, c] - O.tmi(psi, L, L // 4, c)) < 1e-5
    finally:
        M.set_eigensolver(0)
        b.close()


def test_single_drop_in_matches_rocsolver(M):
    b = _batch(M, 18, 6, 1, 777)
    try:
        ptr = b.state_ptr(0)
        for kept in ([0, 1, 2, 3, 4], [3, 7, 11, 15, 2, 9, 0], list(range(9))):
            M.set_eigensolver(0)
            s0 = M.entropy(ptr, 18, kept)
            M.set_eigensolver(1)
            s1 = M.entropy(ptr, 18, kept)
            M.set_eigensolver(0)
            assert abs(s0 - s1) < 1e-10, f"kept {kept}: {s0} vs {s1}"
            psi = b.get_state(0).astype(np.complex128)
            assert abs(s0 - O.entropy(psi, 18, kept)) < 1e-5
    finally:
        M.set_eigensolver(0)
        b.close()


def test_eigensolver_knob_refuses_bad_values(M):
    with pytest.raises(M.MiptHipError):
        M.set_eigensolver(2)
    with pytest.raises(M.MiptHipError):
        M.entropy_hermitian_batched(np.zeros((1, 1025, 1025), np.complex128))


def _device_free_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_batch_destroy_releases_its_readout_workspace(M):
    # the batch owns the entropy workspace of its readouts: 8 trajectories x 3
    # ten-site TMI regions = 24 RDMs of 1024 x 1024 with their Gram planes and
    # eigensolver panels, ~0.6 GB (below the 2 GiB the library trims batch
    # buffers to after a readout); it is held while the batch lives and
    # released with it
    free0 = _device_free_bytes()
    b = _batch(M, 20, 4, 8, 9090)
    try:
        free_states = _device_free_bytes()
        sh, _ = b.observables_all(5, 1)
        assert np.all(np.isfinite(sh))
        held = free_states - _device_free_bytes()
        assert held > (400 << 20), f"the readout holds only {held / 2**20:.0f} MiB"
    finally:
        b.close()
    assert _device_free_bytes() >= free0 - (64 << 20)
