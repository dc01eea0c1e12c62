"""The allocation-free sequential measurement (mipt_hip_measure_collapse_f32_noalloc)
against mipt_hip_measure_collapse_f32 on the same states, through the C ABI:
two calls in a row, the second on a smaller state (the fixed scratch keeps the
first call's contents), outcome and probability equal and the collapsed
states equal; sizes beyond the scratch refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (psi / np.linalg.norm(psi)).astype(np.complex64)


def measure(M, fn, psi, n, q, u):
    eng = M.PlainEngine(False)
    try:
        eng.reset(n)
        eng.set_from_host(psi)
        o, p = ctypes.c_int(-1), ctypes.c_double(-1.0)
        st = getattr(M.lib(), fn)(ctypes.c_void_p(eng.data), n, q, u, ctypes.byref(o), ctypes.byref(p))
        return st, o.value, p.value, eng.to_host()
    finally:
        eng.close()


@pytest.mark.parametrize("first,second", [((20, 7, 0.31), (12, 3, 0.77)), ((24, 23, 0.5), (14, 0, 0.05))])
def test_noalloc_measurement_matches_the_allocating_one(M, first, second):
    for k, (n, q, u) in enumerate((first, second)):
        print(f"no-alloc measurement call {k}: n={n} q={q} u={u}", flush=True)
        psi = random_state(n, 100 * n + q)
        st_a, o_a, p_a, s_a = measure(M, "mipt_hip_measure_collapse_f32", psi, n, q, u)
        st_b, o_b, p_b, s_b = measure(M, "mipt_hip_measure_collapse_f32_noalloc", psi, n, q, u)
        assert st_a == 0 and st_b == 0, M.last_error()
        assert o_a == o_b
        assert p_a == p_b
        assert np.array_equal(s_a, s_b)
        # and the collapse is the measurement it reports
        kept = (np.arange(1 << n) >> q) & 1 == o_b
        assert np.all(s_b[~kept] == 0)
        assert abs(np.sum(np.abs(s_b.astype(np.complex128)) ** 2) - 1.0) < 1e-5  # FP64 sum


def test_noalloc_measurement_refuses_beyond_its_scratch(M):
    psi = random_state(4, 1)
    st, o, p, _ = measure(M, "mipt_hip_measure_collapse_f32_noalloc", psi, 4, 1, 0.5)
    assert st == 0
    eng = M.PlainEngine(False)
    try:
        eng.reset(4)
        o, p = ctypes.c_int(), ctypes.c_double()
        # n = 30 exceeds the fixed scratch: refused before any device work
        st = M.lib().mipt_hip_measure_collapse_f32_noalloc(ctypes.c_void_p(eng.data), 30, 1, 0.5, ctypes.byref(o),
                                                           ctypes.byref(p))
        assert st == -5
        assert "29 qubits" in M.lib().mipt_hip_measure_collapse_last_error().decode()
    finally:
        eng.close()
