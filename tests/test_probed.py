"""mipt_probed modes 0, 1, 2 (one, two or four references), 3 (two) and 4 (two or three).

CPU: the argument contract of cpp/include/mipt/probed/config.hpp:261-372
(every case fails before any device work).  GPU: the CSV the driver writes
(cpp/include/mipt/probed/output.hpp:386-476 columns, the generated name of
output.hpp:46-95) against the FP64 oracle run of every trajectory (same
seeds): per-(p, t) means within 1e-4 nats, standard errors within 1e-4; the
readout right after the attach at t0 = 2N is the closed form S = k ln 2 with
no information.
"""
import csv
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


@pytest.mark.parametrize("argv,code", [
    (["--help"], 0),
    ([], 1),
    (["3", "8", "4", "1", "1", "0", "0.3", "2"], 1),       # three references: mode 4 only
    (["1", "8", "4", "2", "1", "0", "0.3", "2"], 1),       # circ_type 2 not built
    (["1", "7", "4", "1", "1", "0", "0.3", "2"], 1),       # odd ring
    (["4", "10", "4", "1", "1", "0", "0.3", "2"], 1),      # four references need N % 4 == 0
    (["2", "8", "4", "1", "0", "0.1", "4", "20"], 1),      # t_min < t0 = 2N
    (["1", "8", "4", "1", "5", "0.1", "4"], 1),            # mode 5 not built
    (["1", "8", "4", "1", "3", "0.1", "0", "2"], 1),       # mode 3 needs two references
    (["2", "8", "4", "1", "3", "0.1", "2", "2"], 1),       # temporal branch needs delta_x = 0
    (["2", "8", "4", "1", "3", "0.1", "3", "0"], 1),       # spatial branch needs delta_x = N/2
    (["3", "8", "4", "1", "4", "0.1", "4", "1.5"], 1),     # B_min outside [0, 1]
    (["3", "8", "4", "1", "4", "0.1", "4", "0.99"], 1),    # no triangle of an 8-ring that balanced
    (["1", "8", "4", "1", "4", "0.1", "4"], 1),            # mode 4 needs two or three references
    (["2", "8", "4", "1", "4", "0.1"], 1),                 # mode 4 needs [p] [tau_max]
    (["1", "8", "4", "1", "2", "0", "4", "2", "0", "1"], 1),        # mode 2 needs six values
    (["1", "8", "4", "1", "2", "0", "2", "5", "0.1", "0.2", "2"], 1),  # duplicate integer times
    (["1", "8", "4", "1", "0", "0.1", "4"], 1),            # mode 0 needs three values
    (["1", "8", "4", "1", "1", "0.3", "0.1", "3"], 1),     # p_min > p_max
    (["1", "8", "4", "1", "1", "0.1", "0.3", "0"], 1),     # p_res <= 0
    (["2", "26", "4", "1", "1", "0.1", "0.3", "2"], 1),    # 28 qubits
])
def test_probed_argument_contract(M, argv, code, capfd):
    assert M.probed_main(argv) == code
    if code == 1:
        assert "error:" in capfd.readouterr().err


def read_csv(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], [dict(zip(rows[0], r)) for r in rows[1:]]


def oracle_samples(n, depth, p, seeds, sites, pz, attach):
    k = len(sites)
    out = []
    for s in seeds:
        psi, *_ = O.references_trajectory(n, depth, p, s, list(sites), pz, attach_layer=attach)
        out.append(O.probe_information(O.reference_rdm(psi, n, k)))
    return out


def mean_se(v):
    v = np.asarray(v, np.float64)
    return v.mean(), (v.std(ddof=1) / math.sqrt(len(v)) if len(v) > 1 else 0.0)


@pytest.mark.gpu
def test_probed_mode1_two_references(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9001")
    n, R = 8, 6
    print("mipt_probed mode 1, two references", flush=True)
    assert M.probed_main([2, n, R, 1, 1, 0.1, 0.3, 2]) == 0
    name = f"probed_2_haar_n_{n}_reals_{R}_mode_1_p_0.1_0.3_res_2_t_32_32.csv"
    head, rows = read_csv(tmp_path / name)
    assert head == ["N", "circ_type", "circuit_name", "realizations", "p", "t", "S_Q_mean", "S_Q_stderr",
                    "I2_mean", "I2_stderr"]
    assert len(rows) == 2
    for pi, row in enumerate(rows):
        p = [0.1, 0.3][pi]
        assert (int(row["N"]), row["circuit_name"], int(row["t"]), float(row["p"])) == (n, "Random Haar Unitary", 32, p)
        seeds = [O.lib().orc_trajectory_seed(9001, pi * R + i) for i in range(R)]
        want = oracle_samples(n, 4 * n, p, seeds, [3, 7], 0, 2 * n)
        for col, key in (("S_Q", "sq"), ("I2", "i2")):
            m, se = mean_se([w[key] for w in want])
            assert abs(float(row[col + "_mean"]) - m) < 1e-4, (col, p)
            assert abs(float(row[col + "_stderr"]) - se) < 1e-4, (col, p)


@pytest.mark.gpu
def test_probed_mode0_one_reference_encoder(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9002")
    n, R, p = 8, 5, 0.2
    print("mipt_probed mode 0, one reference", flush=True)
    assert M.probed_main([1, n, R, 1, 0, p, 0, 3]) == 0
    head, rows = read_csv(tmp_path / f"probed_1_haar_n_{n}_reals_{R}_mode_0_p_0.2_t_0_3.csv")
    assert head == ["N", "circ_type", "circuit_name", "realizations", "p", "t", "S_mean", "S_stderr"]
    seeds = [O.lib().orc_trajectory_seed(9002, i) for i in range(R)]
    for t, row in enumerate(rows):
        assert int(row["t"]) == t
        # reference at site N-1, Bell pair at t=0, 2N unmeasured encoder layers, then t measured layers
        want = oracle_samples(n, 2 * n + t, p, seeds, [n - 1], 2 * n, 0)
        m, se = mean_se([w["sq"] for w in want])
        assert abs(float(row["S_mean"]) - m) < 1e-4 and abs(float(row["S_stderr"]) - se) < 1e-4


@pytest.mark.gpu
def test_probed_mode0_four_references(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9003")
    out = tmp_path / "four.csv"
    monkeypatch.setenv("MIPT_PROBED_INTERNAL_OUTPUT", str(out))
    n, R, p = 8, 4, 0.15
    print("mipt_probed mode 0, four references", flush=True)
    assert M.probed_main([4, n, R, 1, 0, p, 2 * n, 2 * n + 2]) == 0
    head, rows = read_csv(out)
    assert head[5:] == ["t", "t_minus_t0", "scaled_time", "I2_mean", "I2_stderr", "I3_mean", "I3_stderr",
                        "I4_mean", "I4_stderr"]
    seeds = [O.lib().orc_trajectory_seed(9003, i) for i in range(R)]
    sites = [n // 4 - 1, n // 2 - 1, 3 * n // 4 - 1, n - 1]
    for j, row in enumerate(rows):
        t = 2 * n + j
        assert (int(row["t"]), int(row["t_minus_t0"])) == (t, j)
        assert abs(float(row["scaled_time"]) - j / n) < 1e-15
        if j == 0:
            for col in ("I2", "I3", "I4"):
                assert float(row[col + "_mean"]) == 0.0
            continue
        want = oracle_samples(n, t, p, seeds, sites, 0, 2 * n)
        for col, key in (("I2", "i2"), ("I3", "i3"), ("I4", "i4")):
            m, se = mean_se([w[key] for w in want])
            assert abs(float(row[col + "_mean"]) - m) < 1e-4, (t, col)
            assert abs(float(row[col + "_stderr"]) - se) < 1e-4, (t, col)


@pytest.mark.gpu
def test_probed_mode2_grid(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9004")
    n, R = 8, 4
    print("mipt_probed mode 2, two references", flush=True)
    assert M.probed_main([2, n, R, 1, 2, 0, 6, 4, 0.1, 0.2, 2]) == 0
    head, rows = read_csv(tmp_path / f"probed_2_haar_n_{n}_reals_{R}_mode_2_p_0.1_0.2_res_2_t_0_6_res_4.csv")
    assert head[5:] == ["t", "S_Q_mean", "S_Q_stderr"]
    assert [(float(r["p"]), int(r["t"])) for r in rows] == [(p, t) for p in (0.1, 0.2) for t in (0, 2, 4, 6)]
    for k, row in enumerate(rows):
        pi, t = divmod(k, 4)
        t *= 2
        p = [0.1, 0.2][pi]
        if t == 0:  # Bell pairs with their sites: S_Q = 2 ln 2
            assert abs(float(row["S_Q_mean"]) - 2 * math.log(2)) < 1e-12
            continue
        seeds = [O.lib().orc_trajectory_seed(9004, pi * R + i) for i in range(R)]
        want = oracle_samples(n, t, p, seeds, [3, 7], 0, 0)
        m, se = mean_se([w["sq"] for w in want])
        assert abs(float(row["S_Q_mean"]) - m) < 1e-4 and abs(float(row["S_Q_stderr"]) - se) < 1e-4


@pytest.mark.gpu
def test_probed_mode4_pairs(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9005")
    n, R, p, tau_max = 8, 3, 0.15, 3
    print("mipt_probed mode 4, two references", flush=True)
    assert M.probed_main([2, n, R, 1, 4, p, tau_max]) == 0
    head, rows = read_csv(tmp_path / f"probed_2_haar_n_{n}_reals_{R}_mode_4_p_0.15_teq_16_tau_0_3.csv")
    assert head[5:] == ["distance", "chord_length", "t_eq", "tau", "tau_over_chord", "t_absolute", "I_mean",
                        "I_stderr", "negativity_mean", "negativity_stderr", "log_negativity_mean",
                        "log_negativity_stderr", "negativity_positive_count", "negativity_positive_fraction",
                        "origins_randomized"]
    assert len(rows) == (n // 2) * (tau_max + 1)
    M64 = 2**64 - 1

    def splitmix(v):
        v = (v + 0x9E3779B97F4A7C15) & M64
        v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & M64
        v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & M64
        return v ^ (v >> 31)

    for row in rows:
        d, tau = int(row["distance"]), int(row["tau"])
        chord = n / math.pi * math.sin(math.pi * d / n)
        assert abs(float(row["chord_length"]) - chord) < 1e-12 and int(row["t_absolute"]) == 16 + tau
        if tau == 0:
            assert float(row["I_mean"]) == 0.0 and int(row["negativity_positive_count"]) == 0
            continue
        want = []
        for i in range(R):
            seed = O.lib().orc_trajectory_seed(9005, (d - 1) * R + i)
            x0 = splitmix(seed ^ 0x4F524947494E) % n
            want += oracle_samples(n, 16 + tau, p, [seed], [x0, (x0 + d) % n], 0, 16)
        for col, key in (("I", "i2"), ("negativity", "negativity"), ("log_negativity", "log_negativity")):
            m, se = mean_se([w[key] for w in want])
            assert abs(float(row[col + "_mean"]) - m) < 1e-4, (d, tau, col)
            assert abs(float(row[col + "_stderr"]) - se) < 1e-4, (d, tau, col)
        assert int(row["negativity_positive_count"]) == sum(w["negativity"] > 1e-12 for w in want)


@pytest.mark.gpu
@pytest.mark.parametrize("dx,dt", [(0, 2), (3, 0)])
def test_probed_mode3_staggered(M, tmp_path, monkeypatch, dx, dt):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9006")
    n, R, p = 6, 2, 0.15
    print(f"mipt_probed mode 3 dx={dx} dt={dt}", flush=True)
    assert M.probed_main([2, n, R, 1, 3, p, dx, dt]) == 0
    head, rows = read_csv(tmp_path / f"probed_2_haar_n_{n}_reals_{R}_mode_3_p_0.15_dx_{dx}_dt_{dt}_tau_0_{8 * n}.csv")
    assert head[5:] == ["delta_x", "delta_t", "tau1", "tau2", "tau", "tau_over_L", "t_absolute", "I_mean",
                        "I_stderr", "negativity_mean", "negativity_stderr", "log_negativity_mean",
                        "log_negativity_stderr"]
    tau1, tau2 = 4 * n, 4 * n + dt
    assert len(rows) == 8 * n + 1
    seeds = [O.lib().orc_trajectory_seed(9006, i) for i in range(R)]
    for tau, row in enumerate(rows):
        assert (int(row["tau1"]), int(row["tau2"]), int(row["tau"]), int(row["t_absolute"])) == (tau1, tau2, tau,
                                                                                                 tau2 + tau)
        if tau == 0:
            assert float(row["I_mean"]) == 0.0
            continue
        if tau % 8 and tau != 1:
            continue  # the oracle replays every trajectory from t = 0: check a subset of readout times
        want = [O.probe_information(O.reference_rdm(
            O.references_trajectory(n, tau2 + tau, p, s, [0, dx % n], 0, attach_layer=tau1, attach_stagger=dt)[0],
            n, 2)) for s in seeds]
        for col, key in (("I", "i2"), ("negativity", "negativity"), ("log_negativity", "log_negativity")):
            m, se = mean_se([w[key] for w in want])
            assert abs(float(row[col + "_mean"]) - m) < 1e-4, (tau, col)
            assert abs(float(row[col + "_stderr"]) - se) < 1e-4, (tau, col)


@pytest.mark.gpu
def test_probed_mode4_triangles(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "9007")
    n, R, p, tau_max = 6, 2, 0.15, 2
    print("mipt_probed mode 4, three references", flush=True)
    assert M.probed_main([3, n, R, 1, 4, p, tau_max]) == 0
    head, rows = read_csv(tmp_path / f"probed_3_haar_n_{n}_reals_{R}_mode_4_p_0.15_b_0.5_teq_12_tau_0_2.csv")
    assert head[5:9] == ["geometry_id", "distance_1", "distance_2", "distance_3"]
    # triangles of a 6-ring by sorted periodic sides, balance >= 0.5: (1,1,2), (1,2,3), (2,2,2)
    from itertools import combinations
    pd = lambda a, b: min(abs(a - b), n - abs(a - b))
    groups = {}
    for a, b, c in combinations(range(n), 3):
        groups.setdefault(tuple(sorted((pd(a, b), pd(a, c), pd(b, c)))), []).append((a, b, c))
    sig = sorted(groups)
    assert [tuple(int(r["distance_%d" % i]) for i in (1, 2, 3)) for r in rows[::tau_max + 1]] == sig
    M64 = 2**64 - 1

    def splitmix(v):
        v = (v + 0x9E3779B97F4A7C15) & M64
        v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & M64
        v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & M64
        return v ^ (v >> 31)

    for row in rows:
        g, tau = int(row["geometry_id"]), int(row["tau"])
        emb = groups[sig[g]]
        assert int(row["embedding_count"]) == len(emb) and row["GMN_mean"] == "nan"
        if tau == 0:
            assert float(row["TMI_mean"]) == 0.0 and float(row["joint_purity_mean"]) == 0.125
            continue
        want = []
        for i in range(R):
            seed = O.lib().orc_trajectory_seed(9007, g * R + i)
            sites = list(emb[splitmix(seed ^ 0x4F524947494E) % len(emb)])
            want += oracle_samples(n, 12 + tau, p, [seed], sites, 0, 12)
        for col, key in (("TMI", "tmi"), ("average_MI", "average_mi"),
                         ("min_bipartite_negativity", "min_bipartite_negativity"),
                         ("joint_purity", "joint_purity"), ("mean_single_purity", "mean_single_purity")):
            m, se = mean_se([w[key] for w in want])
            assert abs(float(row[col + "_mean"]) - m) < 1e-4, (g, tau, col)
            assert abs(float(row[col + "_stderr"]) - se) < 1e-4, (g, tau, col)
