"""Full-size parity of the batched engine against the FP64 oracle and closed forms.

* TMI regions on the device: 4-party GHZ states whose I3 is +1 bit exactly
  for the cyclic A, B, C, D of cpp/include/mipt/tmi/compute.hpp:64-113 (and 0
  under an offset, direction or D-width mistake, as tests/test_oracle.py shows
  on the CPU), at every cycle offset, through mipt_hip_tmi_f32 and the
  batched observables; L=26 with 6-site blocks and an 8-site D.
* The reference's published full-pipeline value (cpp/README.md:456-458):
  deep p=0 Haar brickwork at N=16 averages to the Page values of I3,
  -5.847228 bits (von Neumann) and -5.022476 bits (Renyi-2).
* The chains no other test evolves: L=22 and L=24 (three sweeps per layer),
  the metric's workload at full depth (L=20, T=40, 32 trajectories across the
  p sweep, device-drawn streams) and L=26 (512 MiB states).

Bar: outcomes bit-exact except draws within 1e-6 of the Born threshold;
amplitudes within 1e-5 of the complex128 oracle; S and TMI within 1e-5 bits.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from threadpoolctl import threadpool_limits

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FLAG_EPS = 1e-6
POOL = 16


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


def ghz(n, sites, dtype=np.complex64):
    psi = np.zeros(1 << n, dtype)
    psi[0] = psi[sum(1 << q for q in sites)] = 1 / math.sqrt(2)
    return psi


def ghz_sites(n, b, o):
    """Last site of A, first of B, last of C, last of D for offset o."""
    return [(o + b - 1) % n, (o + b) % n, (o + 3 * b - 1) % n, (o + n - 1) % n]


def check_outcomes(got, want, p1_oracle, unif, flags, label):
    m = flags == 1
    near = np.abs(unif - p1_oracle) < FLAG_EPS
    bad = m & (got != want) & ~near
    assert not bad.any(), f"{label}: outcome mismatch at {np.argwhere(bad)[:5]}"
    return bool((m & near).any())


@pytest.mark.parametrize("n,b", [(8, 2), (10, 2), (12, 3), (14, 3), (16, 4)])
def test_tmi_regions_ghz_on_device(M, n, b):
    print(f"GHZ4 TMI regions n={n} b={b}", flush=True)
    # one trajectory per cycle offset, each holding that offset's GHZ state
    specs = [M.Spec(n, 2, 0.0, 900 + o) for o in range(b)]
    batch = M.Batch(specs)
    try:
        for o in range(b):
            batch.set_state(o, ghz(n, ghz_sites(n, b, o)))
        _, tm = batch.observables_all(tmi_block=b, tmi_cycles=b)
        for o in range(b):
            for o2 in range(b):
                want = 1.0 if o2 == o else 0.0
                assert abs(tm[o, o2] - want) < 1e-5, (o, o2, tm[o, o2])
            # the drop-in on the same device state
            got, terms = M.tmi(batch.state_ptr(o), n, n, b, o)
            assert abs(got - 1.0) < 1e-5
            assert abs(batch.tmi(o, o, b, 2) - 1.0) < 1e-5  # Renyi-2: the same +1 for GHZ
    finally:
        batch.close()


@pytest.mark.timeout(300)
def test_tmi_regions_ghz_l26_blocks_of_six(M):
    # n = 26, b = 6: A, B, C of 6 sites and an 8-site D (the reference allows
    # only n % 4 == 0); the GHZ sites include site 25, which a D of width b
    # would leave out of every region
    n, b = 26, 6
    print("GHZ4 TMI regions n=26 b=6", flush=True)
    batch = M.Batch([M.Spec(n, 2, 0.0, 950)])
    try:
        for o in (0, 5):
            batch.set_state(0, ghz(n, ghz_sites(n, b, o)))
            got, _ = M.tmi(batch.state_ptr(0), n, n, b, o)
            assert abs(got - 1.0) < 1e-5, (o, got)
            other, _ = M.tmi(batch.state_ptr(0), n, n, b, 5 - o)
            assert abs(other) < 1e-5, (o, other)
    finally:
        batch.close()


@pytest.mark.timeout(300)
def test_page_value_of_deep_haar_circuits_n16(M):
    # cpp/README.md:456-458: 300 realizations of deep p=0 Haar at N=16 give
    # -5.84743 +- 0.00028 (order 1) and -5.02235 +- 0.00038 (order 2) against
    # the Page closed forms -5.847228 and -5.022476.  Same size and count here,
    # depth 8N; the mean must lie within 4 standard errors of the closed form.
    n, R, layers = 16, 300, 128
    print(f"Page value n={n} x {R} trajectories, {layers} layers", flush=True)
    specs = [M.Spec(n, layers, 0.0, 16000 + i) for i in range(R)]
    batch = M.Batch(specs)
    try:
        batch.generate()
        batch.evolve()
        batch.synchronize()
        _, tm = batch.observables_all(tmi_block=4, tmi_cycles=1)
        i3 = tm[:, 0]
        i3_2 = np.array([batch.tmi(i, 0, 4, 2) for i in range(R)])
    finally:
        batch.close()
    for vals, page in ((i3, -5.847228), (i3_2, -5.022476)):
        se = vals.std(ddof=1) / math.sqrt(R)
        print(f"  mean {vals.mean():.6f} +- {se:.6f} (closed form {page})", flush=True)
        assert se < 0.002
        assert abs(vals.mean() - page) < 4 * se + 2e-5


def _oracle_traj(args):
    n, layers, p, seed, with_obs = args
    psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed)
    obs = None
    if with_obs:
        obs = (O.half_chain_entropy(psi, n), O.tmi(psi, n, n // 4, 0))
    return psi, out, p1, f, u, obs


def _oracle_many(cases):
    """The oracle runs of `cases` on a thread pool.  The oracle library is
    loaded and the BLAS thread limit set here, in the calling thread: a first
    dlopen in one worker while another scans the loaded libraries for
    threadpool_limits (dl_iterate_phdr holds the loader lock and calls back
    into Python for the GIL; dlopen holds the GIL and waits for the loader
    lock) deadlocks the pool."""
    O.lib()
    with threadpool_limits(1), ThreadPoolExecutor(min(POOL, len(cases))) as ex:
        return list(ex.map(_oracle_traj, cases))


def _run_device(M, cases, plan=0):
    specs = [M.Spec(n, layers, p, seed) for (n, layers, p, seed, _) in cases]
    batch = M.Batch(specs, plan=plan)
    batch.generate()
    batch.evolve()
    batch.synchronize()
    return batch


@pytest.mark.timeout(300)
@pytest.mark.parametrize("plan", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [22, 24])
def test_three_sweep_chains_match_oracle(M, n, plan):
    # L = 22 and 24 take three sweeps per layer (windows of 8 high bits over
    # the 5 lowest); device-drawn streams against the oracle's, under every
    # layer plan variant (window placement and bond-to-window assignment)
    cases = [(n, 6, 0.15, 2200 + 10 * n + i, n == 22) for i in range(2)] + [(n, 5, 0.3, 2250 + n, False)]
    print(f"L={n} plan {plan}: {len(cases)} trajectories", flush=True)
    batch = _run_device(M, cases, plan)
    refs = _oracle_many(cases)
    try:
        sh = batch.observables_all(tmi_block=0, tmi_cycles=0)[0] if n == 22 else None
        for i, (psi, out, p1, f, u, obs) in enumerate(refs):
            got, flagged, gp1, _ = batch.outcomes(i)
            near = check_outcomes(got, out, p1, u, f, f"L={n} traj {i}")
            m = f == 1
            assert np.abs(gp1[m] - p1[m]).max(initial=0) < 1e-5
            if not near:
                err = np.abs(batch.get_state(i) - psi).max()
                assert err < 1e-5, (i, err)
            if obs is not None:
                assert abs(sh[i] - obs[0]) < 1e-5
    finally:
        batch.close()


def _oracle_closed(args):
    n, layers, p, seed, closed = args
    psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed, closed)
    return psi, out, p1, f, u


@pytest.mark.timeout(300)
@pytest.mark.parametrize("plan", [0, 1, 2, 3])
def test_layer_fusion_on_and_off_match_oracle(M, plan):
    # With layer fusion a trajectory that measures at most one site in layer
    # t, not bonded in layer t + 1's first sweep, runs that sweep in the same
    # launch as the last sweep of layer t (its histogram taken after it, its
    # collapse applied at layer t + 1's second sweep), and the sweeps of odd
    # layers run in reverse window order.  Both settings against the oracle
    # (device-drawn streams; fusing layers at every p here, all of them at
    # p = 0; at p = 0.6 layers beyond one 10-site histogram chunk, which never
    # fuse, in the same group), and the two device states against each other:
    # L=20 rings, an L=20 chain, an L=18 ring and L=22 (three sweeps).
    ps = [0.0, 0.01, 0.03, 0.06, 0.1, 0.2, 0.6]
    cases = [(20, 16, p, 3300 + i, True) for i, p in enumerate(ps)] + \
            [(20, 12, 0.02, 3310, False), (18, 10, 0.03, 3311, True), (22, 8, 0.03, 3312, True)]
    specs = [M.Spec(n, layers, p, seed, closed) for (n, layers, p, seed, closed) in cases]
    got = {}
    try:
        for on in (0, 1):
            print(f"plan {plan}: layer fusion {on}, {len(cases)} trajectories", flush=True)
            M.set_layer_fusion(on)
            batch = M.Batch(specs, plan=plan)
            try:
                batch.generate()
                batch.evolve()
                batch.synchronize()
                got[on] = [(batch.outcomes(i), batch.get_state(i)) for i in range(len(cases))]
            finally:
                batch.close()
    finally:
        M.set_layer_fusion(None)
    O.lib()
    with threadpool_limits(1), ThreadPoolExecutor(min(POOL, len(cases))) as ex:
        refs = list(ex.map(_oracle_closed, cases))
    for i, (psi, out, p1, f, u) in enumerate(refs):
        tag = f"plan {plan} traj {i} L={cases[i][0]} p={cases[i][2]}"
        near = False
        for on in (0, 1):
            (o, flagged, gp1, _), st = got[on][i]
            near |= check_outcomes(o, out, p1, u, f, f"{tag} fusion {on}")
            m = f == 1
            assert np.abs(gp1[m] - p1[m]).max(initial=0) < 1e-5, tag
        if near:
            continue
        for on in (0, 1):
            assert np.abs(got[on][i][1] - psi).max() < 1e-5, f"{tag} fusion {on}: amplitudes"
        assert np.array_equal(got[0][i][0][0], got[1][i][0][0]), tag
        assert np.abs(got[0][i][1] - got[1][i][1]).max() < 1e-5, tag


def test_split_evolution_ranges_match_one_evolution(M):
    # evolve(0, 3), evolve(3, 7), evolve(7, 12) against one evolve() of the
    # same device draws, layer fusion on: a range's last layer fuses into
    # nothing and its collapse is applied before the range returns, so no
    # deferred collapse crosses a boundary; L=20 rings and an L=22 ring
    # (three sweeps), and the oracle on top
    cases = [(20, 12, p, 3400 + i, True) for i, p in enumerate([0.0, 0.03, 0.1, 0.25])] + [(22, 12, 0.05, 3410, True)]
    specs = [M.Spec(n, layers, p, seed, closed) for (n, layers, p, seed, closed) in cases]
    print(f"split ranges: {len(cases)} trajectories", flush=True)
    runs = []
    for split in (False, True):
        batch = M.Batch(specs)
        try:
            batch.generate()
            if split:
                for a, b in ((0, 3), (3, 7), (7, 12)):
                    batch.evolve(a, b)
            else:
                batch.evolve()
            batch.synchronize()
            runs.append([(batch.outcomes(i), batch.get_state(i)) for i in range(len(cases))])
        finally:
            batch.close()
    O.lib()
    with threadpool_limits(1), ThreadPoolExecutor(min(POOL, len(cases))) as ex:
        refs = list(ex.map(_oracle_closed, cases))
    for i, (psi, out, p1, f, u) in enumerate(refs):
        tag = f"traj {i} L={cases[i][0]} p={cases[i][2]}"
        near = False
        for r in runs:
            near |= check_outcomes(r[i][0][0], out, p1, u, f, tag)
        if near:
            continue
        assert np.array_equal(runs[0][i][0][0], runs[1][i][0][0]), tag
        for r in runs:
            assert np.abs(r[i][1] - psi).max() < 1e-5, tag


@pytest.mark.timeout(600)
def test_metric_workload_full_depth_32_trajectories(M):
    # L=20 ring, T=40, p in {0, 0.02, ..., 0.30}: two trajectories per p value
    # (trajectory j of the 16 x 256 sweep has seed j + 1), device-drawn
    # streams; outcomes, amplitudes, S(10) and TMI (5-site blocks) per result
    ps = np.linspace(0.0, 0.3, 16)
    cases = [(20, 40, float(ps[pi]), pi * 256 + r + 1, True) for pi in range(16) for r in range(2)]
    print("L=20 T=40: 32 trajectories", flush=True)
    batch = _run_device(M, cases)
    try:
        sh, tm = batch.observables_all(tmi_block=5, tmi_cycles=1)
        states = [batch.get_state(i) for i in range(len(cases))]
        outs = [batch.outcomes(i) for i in range(len(cases))]
    finally:
        batch.close()
    refs = _oracle_many(cases)
    worst = {"amp": 0.0, "s": 0.0, "tmi": 0.0}
    for i, (psi, out, p1, f, u, (s_ref, t_ref)) in enumerate(refs):
        got, flagged, gp1, _ = outs[i]
        near = check_outcomes(got, out, p1, u, f, f"traj {i} p={cases[i][2]:.2f}")
        if near:
            continue
        worst["amp"] = max(worst["amp"], float(np.abs(states[i] - psi).max()))
        worst["s"] = max(worst["s"], abs(sh[i] - s_ref))
        worst["tmi"] = max(worst["tmi"], abs(tm[i, 0] - t_ref))
    print(f"  worst errors: {worst}", flush=True)
    assert worst["amp"] < 1e-5 and worst["s"] < 1e-5 and worst["tmi"] < 1e-5


@pytest.mark.timeout(600)
def test_fused_three_sweep_rings_match_oracle(M):
    # three sweeps per layer with most (L=24, p = 0.02) or all (L=26, p = 0)
    # layers fused into the next: the compiled fused sweeps of these rings,
    # the deferred collapse at the middle sweep, against the oracle
    cases = [(24, 6, 0.02, 3501, True), (26, 4, 0.0, 3502, True)]
    print(f"fused three-sweep rings: {cases}", flush=True)
    specs = [M.Spec(n, layers, p, seed, closed) for (n, layers, p, seed, closed) in cases]
    batch = M.Batch(specs)
    try:
        batch.generate()
        batch.evolve()
        batch.synchronize()
        got = [(batch.outcomes(i), batch.get_state(i)) for i in range(len(cases))]
    finally:
        batch.close()
    for i, case in enumerate(cases):
        psi, out, p1, f, u = _oracle_closed(case)
        (o, flagged, gp1, _), st = got[i]
        near = check_outcomes(o, out, p1, u, f, f"L={case[0]}")
        if not near:
            assert np.abs(st - psi).max() < 1e-5, f"L={case[0]}"


@pytest.mark.timeout(600)
def test_l26_evolution_matches_oracle(M):
    # L=26 (2^26 amplitudes, 512 MiB complex64 on the device, 1 GiB complex128
    # in the oracle), four layers at p = 0.16 on the device-drawn stream
    n, layers, p, seed = 26, 4, 0.16, 2601
    print("L=26: 1 trajectory x 4 layers vs the oracle", flush=True)
    batch = _run_device(M, [(n, layers, p, seed, False)])
    try:
        got, flagged, gp1, _ = batch.outcomes(0)
        dev = batch.get_state(0)
    finally:
        batch.close()
    psi, out, p1, f, u, _ = _oracle_traj((n, layers, p, seed, False))
    near = check_outcomes(got, out, p1, u, f, "L=26")
    if not near:
        err = float(np.abs(dev - psi).max())
        print(f"  amplitude error {err:.3e}", flush=True)
        assert err < 1e-5
