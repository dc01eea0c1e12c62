"""Translation-averaged S1/S2 of cyclic blocks (the reference's entropy.exe,
cpp/include/mipt/entropy.hpp:521-640) and its CSV.

CPU: the argument contract of cpp/apps/entropy.cpp:16-80.  GPU: the batched
per-trajectory averages against the FP64 oracle (within 1e-5 nats), and the
driver's CSV: schema, counts, x = (N/pi) sin(pi L_A/N), and Welford means and
standard errors against the oracle's values of the same trajectories.
"""
import csv
import math

import numpy as np
import pytest

from oracle import oracle as O

LN2 = math.log(2.0)


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


@pytest.mark.parametrize("argv,code", [
    (["--help"], 0),
    (["8", "2", "0.1", "4", "1"], 2),        # output missing
    (["2", "2", "0.1", "4", "1", "o.csv"], 1),
    (["8", "-1", "0.1", "4", "1", "o.csv"], 1),
    (["8", "2", "1.5", "4", "1", "o.csv"], 1),
    (["8", "2", "0.1", "0", "1", "o.csv"], 1),
    (["8", "2", "0.1", "4", "0", "o.csv"], 1),
    (["9", "2", "0.1", "4", "1", "o.csv"], 1),
])
def test_entropy_argument_contract(M, argv, code, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    assert M.entropy_main(argv) == code


def oracle_translation(psi, n, lmin, lmax):
    s1, s2 = [], []
    for l in range(lmin, lmax + 1):
        a = b = 0.0
        for i in range(n):
            kept = [(i + k) % n for k in range(l)]
            a += O.entropy(psi, n, kept) * LN2
            b += O.entropy(psi, n, kept, order=2) * LN2
        s1.append(a / n)
        s2.append(b / n)
    return np.array(s1), np.array(s2)


@pytest.mark.gpu
def test_translation_entropies_match_oracle(M):
    cases = [(8, 6, 0.1, 601), (12, 8, 0.2, 602), (16, 4, 0.1, 603)]
    print("batch translation entropies", flush=True)
    b = M.Batch([M.Spec(n, t, p, s) for (n, t, p, s) in cases])
    b.generate()
    b.evolve()
    b.synchronize()
    s1, s2 = b.translation_entropies(2, 4, 3)
    for i, (n, t, p, seed) in enumerate(cases):
        psi, *_ = O.simulate(n, t, p, seed)
        w1, w2 = oracle_translation(psi, n, 2, 4)
        assert np.abs(s2[i] - w2).max() < 1e-5
        assert np.abs(s1[i, :2] - w1[:2]).max() < 1e-5
        assert math.isnan(s1[i, 2])   # above max_s1_kept


@pytest.mark.gpu
def test_entropy_csv_matches_oracle(M, tmp_path, monkeypatch):
    monkeypatch.setenv("MIPT_SEED", "777")
    n, periods, p, R = 8, 3, 0.15, 5
    out = tmp_path / "ent.csv"
    print("entropy scan", flush=True)
    assert M.entropy_main([n, periods, p, R, 1, str(out)]) == 0
    rows = list(csv.DictReader(open(out)))
    assert [int(r["L_A"]) for r in rows] == [2, 3, 4]
    vals1, vals2 = [], []
    for r in range(R):
        psi, *_ = O.simulate(n, 2 * periods, p, O.lib().orc_trajectory_seed(777, r))
        w1, w2 = oracle_translation(psi, n, 2, 4)
        vals1.append(w1)
        vals2.append(w2)
    vals1, vals2 = np.array(vals1), np.array(vals2)
    for k, row in enumerate(rows):
        la = 2 + k
        assert int(row["N"]) == n and int(row["completed_realizations"]) == R
        assert row["circuit_name"] == "Random Haar Unitary"
        assert abs(float(row["x"]) - (n / math.pi) * math.sin(math.pi * la / n)) < 1e-12
        assert abs(float(row["S1_mean"]) - vals1[:, k].mean()) < 1e-5
        assert abs(float(row["S2_mean"]) - vals2[:, k].mean()) < 1e-5
        assert abs(float(row["S2_stderr"]) - vals2[:, k].std(ddof=1) / math.sqrt(R)) < 1e-5
        assert int(row["total_subsystems"]) == n * R
