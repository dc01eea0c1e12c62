"""free_energy: the measurement-record free energy (cpp/apps/free_energy.cpp).

CPU: the argument contract and the refusal of existing CSVs (both before any
device work); the initial states (free_energy.hpp:127-222): a product state
is normalised with every single-site state pure and its Bloch z uniform on
[-1, 1] (mean 0, mean square 1/3), a Haar state is normalised with S(L/2)
near the Page value.  GPU: every samples row and every time-series row
against the oracle run of the same trajectories (same gates, draws and
initial states): outcomes equal, F to 1e-3 nats, S(L/2) to 1e-4 bits.
"""
import csv
import math

import numpy as np
import pytest

from oracle import oracle as O

M64 = 2**64 - 1


def splitmix64(v):
    v = (v + 0x9E3779B97F4A7C15) & M64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & M64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & M64
    return v ^ (v >> 31)


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


@pytest.mark.parametrize("argv,code", [
    (["--help"], 0),
    ([], 2),
    (["8", "0.1", "4", "1"], 2),
    (["8", "0.1", "4", "2", "0"], 1),    # circ_type 2 not built
    (["7", "0.1", "4", "1", "0"], 1),    # odd ring
    (["28", "0.1", "4", "1", "0"], 1),   # beyond the engine's 26 sites
    (["8", "1.5", "4", "1", "0"], 1),    # p outside [0, 1]
    (["8", "0.1", "0", "1", "0"], 1),    # realizations
    (["8", "0.1", "4", "1", "2"], 1),    # init_state
])
def test_free_energy_argument_contract(M, argv, code, capfd):
    assert M.free_energy_main(argv) == code
    if code == 1:
        assert "free_energy:" in capfd.readouterr().err


def test_free_energy_refuses_existing_csv(M, tmp_path, monkeypatch, capfd):
    monkeypatch.chdir(tmp_path)
    (tmp_path / "free_energy_haar_n_8_p_0.1_init_0_reals_4_samples.csv").write_text("x\n")
    assert M.free_energy_main(["8", "0.1", "4", "1", "0"]) == 1
    assert "resuming is not built" in capfd.readouterr().err


def test_initial_states(M):
    zs = []
    for seed in range(200):
        psi = M.free_energy_initial_state(6, 0, seed)
        assert abs(np.vdot(psi, psi).real - 1) < 1e-12
        for q in range(6):
            r = O.rdm(psi, 6, [q])
            assert abs(np.trace(r @ r).real - 1) < 1e-12  # pure: a product state
        zs.append(np.real(O.rdm(psi, 6, [0])[0, 0] - O.rdm(psi, 6, [0])[1, 1]))
    zs = np.array(zs)
    assert abs(zs.mean()) < 0.15 and abs((zs ** 2).mean() - 1 / 3) < 0.08
    h = M.free_energy_initial_state(12, 1, 5)
    assert abs(np.vdot(h, h).real - 1) < 1e-12
    page = 6 - 1 / (2 * math.log(2))
    assert abs(O.half_chain_entropy(h, 12) - page) < 0.1
    assert np.array_equal(h, M.free_energy_initial_state(12, 1, 5))
    assert not np.allclose(h, M.free_energy_initial_state(12, 1, 6))


def read(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], [dict(zip(rows[0], r)) for r in rows[1:]]


@pytest.mark.gpu
@pytest.mark.parametrize("init_state", [0, 1])
def test_free_energy_csvs_match_oracle(M, tmp_path, monkeypatch, init_state):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_FREE_ENERGY_SEED", "77")
    n, p, R = 6, 0.2, 3
    print(f"free_energy n={n} init={init_state}", flush=True)
    assert M.free_energy_main([n, p, R, 1, init_state]) == 0
    prefix = f"free_energy_haar_n_{n}_p_0.2_init_{init_state}_reals_{R}"
    head_s, samples = read(tmp_path / (prefix + "_samples.csv"))
    head_t, series = read(tmp_path / (prefix + "_timeseries.csv"))
    assert head_s[:6] == ["N", "p", "circ_type", "circuit_name", "init_state", "trajectory"]
    assert head_t[-4:] == ["S_half_mean", "S_half_stddev", "S_half_stderr", "S_half_count"]
    eq, total, max_t = 4 * n, 28 * n, 2 * n
    assert len(samples) == R and len(series) == total + 1
    F = np.zeros((R, total + 1))
    S = np.full((R, max_t + 1), np.nan)
    for i in range(R):
        seed = splitmix64(77 ^ i)
        assert int(samples[i]["trajectory_seed"]) == seed
        psi = M.free_energy_initial_state(n, init_state, seed)
        g, f, u = O.build_layers(n, total, p, seed)
        S[i, 0] = O.half_chain_entropy(psi, n)
        meas = 0
        for t in range(total):
            outs, p1 = O.run_layers(psi, n, g, f, u, layer_begin=t, layer_end=t + 1)
            layer = sum(-math.log(p1[t, s] if outs[t, s] == 1 else 1 - p1[t, s]) for s in range(n) if outs[t, s] >= 0)
            meas += int((outs[t] >= 0).sum())
            F[i, t + 1] = F[i, t] + layer
            if t + 1 <= max_t:
                S[i, t + 1] = O.half_chain_entropy(psi, n)
        row = samples[i]
        assert int(row["measurements_total"]) == meas
        assert abs(float(row["F_total"]) - F[i, total]) < 1e-3
        assert abs(float(row["F_production"]) - (F[i, total] - F[i, eq])) < 1e-3
        x = np.arange(1, total - eq + 1)
        slope = np.polyfit(x, F[i, eq + 1:] - F[i, eq], 1)[0]
        assert abs(float(row["production_slope"]) - slope) < 1e-5
        assert abs(float(row["free_energy_density_tilde"]) - slope / n) < 1e-6
    for t, row in enumerate(series):
        assert int(row["t"]) == t and int(row["completed_realizations"]) == R
        assert row["phase"] == ("initial" if t == 0 else "equilibration" if t <= eq else "production")
        assert abs(float(row["F_mean"]) - F[:, t].mean()) < 1e-3
        if t <= max_t:
            assert int(row["S_half_count"]) == R
            assert abs(float(row["S_half_mean"]) - S[:, t].mean()) < 1e-4
        else:
            assert int(row["S_half_count"]) == 0 and row["S_half_mean"] == "nan"
