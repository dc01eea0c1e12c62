"""The batched run entry point and the sim_tmi scan driver.

CPU tests: the argument contract of sim_tmi (cpp/apps/sim_tmi.cpp:11-90,
cpp/include/mipt/tmi/runner.hpp:52-76), which fails before any GPU call.
GPU tests: mipt_hip_run_batch and the scan's CSV against the FP64 oracle on
the same device-drawn streams (S, TMI within 1e-5 bits; probe within 1e-5).
"""
import csv
import math
import os

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
    (["8"] * 10, 2),                                   # too many arguments
    (["10"], 1),                                       # n % 4 != 0
    (["8", "0"], 1),                                   # periods <= 0
    (["8", "2", "0"], 1),                              # realizations <= 0
    (["8", "2", "2", "0"], 1),                         # resolution <= 0
    (["8", "2", "2", "2", "0", "1.5"], 1),             # p outside [0, 1]
    (["8", "2", "2", "2", "0", "1", "0"], 1),          # MMS is not built here
    (["8", "2", "2", "2", "0", "1", "1", "x.csv", "7"], 1),  # all_cycles not 0/1
    (["28"], 1),                                       # beyond 2^26 amplitudes
])
def test_sim_tmi_argument_contract(M, argv, code, capfd):
    assert M.sim_tmi_main(argv) == code
    err = capfd.readouterr().err
    if code == 1:
        assert "error:" in err
    if code == 2:
        assert "Too many arguments" in err


def test_sim_tmi_entropy_order_env_is_checked(M, monkeypatch, capfd):
    monkeypatch.setenv("SIM_TMI_ENTROPY_ORDER", "3")
    assert M.sim_tmi_main(["8", "2", "1", "1"]) == 1
    assert "SIM_TMI_ENTROPY_ORDER" in capfd.readouterr().err


@pytest.mark.parametrize("seed", ["abc", "12x", "-5", "", "99999999999999999999999"])
def test_malformed_seed_is_refused(M, monkeypatch, capfd, seed):
    # MIPT_SEED must parse completely as an unsigned 64-bit value; a typo is
    # refused before any work instead of becoming seed 0 or a prefix
    monkeypatch.setenv("MIPT_SEED", seed)
    assert M.sim_tmi_main(["8", "1", "1", "1"]) == 1
    assert "MIPT_SEED" in capfd.readouterr().err
    assert M.mipt_main(["1", "8", "1", "1", "1"]) == 1
    assert "MIPT_SEED" in capfd.readouterr().err


def _write(path, text):
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(text)
    return str(path)


def test_sim_tmi_resume_refuses_a_foreign_file(M, tmp_path, capfd):
    # the resume contract of cpp/include/mipt/tmi/resume.hpp checks the
    # generated name, the header and the p grid before any simulation
    args = ["8", "2", "3", "2", "0", "1", "1", "0"]
    # a generated name for other realizations
    f = _write(tmp_path / "tmi_haar_n_8_real_5_0_1_res_2.csv", "p,tmi\n")
    assert M.sim_tmi_main(args + [f]) == 1
    assert "was produced with realizations=5" in capfd.readouterr().err
    f = _write(tmp_path / "a.csv", "x,y\n")
    assert M.sim_tmi_main(args + [f]) == 1
    assert "expected header" in capfd.readouterr().err
    f = _write(tmp_path / "b.csv", "p,tmi\n0.5,0.1\n")
    assert M.sim_tmi_main(args + [f]) == 1
    assert "switches away from p=0" in capfd.readouterr().err
    f = _write(tmp_path / "c.csv", "p,tmi\n0,0.1\n0,0.2\n0,0.3\n0,0.4\n")  # 4 rows for a 3-row group
    assert M.sim_tmi_main(args + [f]) == 1
    assert "more rows" in capfd.readouterr().err


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference tree absent (oracle/_ref not built)")


@needs_ref
def test_sim_tmi_resume_refusals_agree_with_reference_reader(M, tmp_path, capfd):
    # every file the driver refuses, the reference's own checkpoint scan
    # (cpp/include/mipt/tmi/resume.hpp:205-333) refuses with the same reason
    args = ["8", "2", "3", "2", "0", "1", "1", "0"]
    expected = "tmi_haar_n_8_real_3_0_1_res_2.csv"
    cases = [("tmi_haar_n_8_real_5_0_1_res_2.csv", "p,tmi\n", "was produced with realizations=5"),
             ("a.csv", "x,y\n", "expected header"),
             ("b.csv", "p,tmi\n0.5,0.1\n", "switches away from p=0"),
             ("c.csv", "p,tmi\n0,0.1\n0,0.2\n0,0.3\n0,0.4\n", "more than the 3")]
    for name, text, why in cases:
        f = _write(tmp_path / name, text)
        assert M.sim_tmi_main(args + [f]) == 1
        ours = capfd.readouterr().err
        with pytest.raises(ValueError) as e:
            O.ref_tmi_resume_load(f, expected, 3, 1, ["0", "1"])
        assert why in str(e.value), str(e.value)
        assert why.split(" ")[0] in ours, ours


@needs_ref
def test_sim_tmi_resume_checkpoints_agree_with_reference_reader(M, tmp_path):
    # complete and whole-circuit files the driver accepts; the reference's scan
    # finds the same checkpoint the driver acts on (nothing left to do for a
    # complete file; whole circuits kept, a trailing partial one dropped)
    text = "p,tmi\n" + "".join(f"{p},{0.1 * i}\n" for p in ("0", "1") for i in range(3))
    f = _write(tmp_path / "done.csv", text)
    ck = O.ref_tmi_resume_load(f, "tmi_haar_n_8_real_3_0_1_res_2.csv", 3, 1, ["0", "1"])
    assert ck["loaded"] and ck["rows"] == 6 and ck["start_p_index"] == 2 and ck["kept_bytes"] == len(text)
    # two cycles per circuit: 5 rows = 2 whole circuits + 1 row of a third
    lines = ["p,tmi"] + [f"0,{0.1 * i}" for i in range(5)]
    f = _write(tmp_path / "cut.csv", "\n".join(lines) + "\n0,0.9")
    ck = O.ref_tmi_resume_load(f, "tmi_haar_all_n_8_real_3_0_1_res_2.csv", 3, 2, ["0", "1"])
    assert (ck["rows"], ck["dropped_rows"], ck["start_p_index"], ck["start_realization"]) == (4, 1, 0, 2)
    assert ck["kept_bytes"] == len("\n".join(lines[:5]) + "\n")


def test_sim_tmi_resume_complete_file_is_left_alone(M, tmp_path, monkeypatch, capfd):
    # 2 p values x 3 realizations x 1 cycle, already complete: nothing to do
    text = "p,tmi\n" + "".join(f"{p},{0.1 * i}\n" for p in ("0", "1") for i in range(3))
    f = _write(tmp_path / "done.csv", text)
    assert M.sim_tmi_main(["8", "2", "3", "2", "0", "1", "1", "0", f]) == 0
    assert "nothing to do" in capfd.readouterr().err
    assert (tmp_path / "done.csv").read_text() == text
    monkeypatch.setenv("SIM_TMI_RESUME", "maybe")
    assert M.sim_tmi_main(["8", "2", "3", "2", "0", "1", "1", "0", f]) == 1


def _oracle_run(n, layers, p, seed):
    psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed)
    return psi, out, p1, f, u


@pytest.mark.gpu
def test_run_batch_matches_oracle(M):
    specs = [M.Spec(8, 8, 0.2, 501), M.Spec(12, 10, 0.1, 502), M.Spec(16, 6, 0.25, 503), M.Spec(8, 8, 0.0, 504)]
    print("run_batch L=8,12,16", flush=True)
    res, tmi = M.run_batch(specs, tmi_cycles=2, entropy_order=1)
    for i, s in enumerate(specs):
        psi, out, p1, f, u = _oracle_run(s.L, s.layers, s.p, s.seed)
        assert res[i]["flagged_draws"] == 0 or np.any(np.abs(u - p1)[f == 1] < 1e-6)
        assert res[i]["measured"] == int(f.sum())
        assert abs(res[i]["s_half"] - O.half_chain_entropy(psi, s.L)) < 1e-5
        for c in range(2):
            assert abs(tmi[i, c] - O.tmi(psi, s.L, s.L // 4, c)) < 1e-5
        assert res[i]["norm_error"] < 1e-5
        assert math.isnan(res[i]["probe_entropy"])
    # Renyi-2 TMI
    _, tmi2 = M.run_batch(specs[:1], tmi_cycles=1, entropy_order=2)
    psi, *_ = _oracle_run(8, 8, 0.2, 501)
    assert abs(tmi2[0, 0] - O.tmi(psi, 8, 2, 0, order=2)) < 1e-5


@pytest.mark.gpu
def test_run_batch_states_errors_per_result(M):
    # fp64_check: each listed result carries its error against a complex128
    # replay of the same FP64 gates, flags and uniforms; the replay is the
    # oracle's computation, so the stated errors equal the errors against the
    # oracle (to FP64 rounding) and stay within the 1e-5 bar; unlisted results
    # and trajectories the plain engine does not replay carry NaN
    specs = [M.Spec(12, 10, 0.15, 611, fp64_check=True), M.Spec(16, 8, 0.2, 612, fp64_check=True),
             M.Spec(20, 6, 0.1, 613, fp64_check=True), M.Spec(12, 8, 0.3, 614, False, grid_x=4, fp64_check=True),
             M.Spec(16, 6, 0.2, 615), M.Spec(12, 6, 0.2, 616, probe_site=6, fp64_check=True)]
    print("run_batch with per-result FP64 errors", flush=True)
    res, tmi = M.run_batch(specs, tmi_cycles=2, entropy_order=1)
    for i, s in enumerate(specs[:4]):
        if s.grid_x:
            psi, out, p1, f, u = O.simulate_grid(s.grid_x, s.L // s.grid_x, s.layers, s.p, s.seed, s.closed)
        else:
            psi, out, p1, f, u = _oracle_run(s.L, s.layers, s.p, s.seed)
        r = res[i]
        if r["flagged_draws"]:
            continue
        assert r["amp_err"] < 1e-5 and r["s_half_err"] < 1e-5 and r["tmi_err"] < 1e-5, r
        s_or = O.half_chain_entropy(psi, s.L)
        t_or = max(abs(tmi[i, c] - O.tmi(psi, s.L, s.L // 4, c)) for c in range(2))
        assert abs(r["s_half_err"] - abs(r["s_half"] - s_or)) < 1e-9
        assert abs(r["tmi_err"] - t_or) < 1e-9
    for r in res[4:]:
        assert math.isnan(r["amp_err"]) and math.isnan(r["s_half_err"]) and math.isnan(r["tmi_err"])


@pytest.mark.gpu
def test_sim_tmi_csv_matches_oracle(M, tmp_path, monkeypatch):
    n, periods, R, res, pmin, pmax = 8, 3, 2, 2, 0.1, 0.3
    master = 12345
    monkeypatch.setenv("MIPT_SEED", str(master))
    path = tmp_path / "scan.csv"
    print("sim_tmi n=8 scan", flush=True)
    assert M.sim_tmi_main([n, periods, R, res, pmin, pmax, 1, 1, str(path)]) == 0
    rows = list(csv.reader(open(path)))
    assert rows[0] == ["p", "tmi"]
    rows = rows[1:]
    cycles = n // 4
    assert len(rows) == res * R * cycles
    ps = [pmin, pmax]
    k = 0
    for pi, p in enumerate(ps):
        for r in range(R):
            seed = O.lib().orc_trajectory_seed(master, pi * R + r)
            psi, *_ = _oracle_run(n, 2 * periods, p, seed)
            for c in range(cycles):
                assert float(rows[k][0]) == p
                assert abs(float(rows[k][1]) - O.tmi(psi, n, n // 4, c)) < 1e-5
                k += 1


@pytest.mark.gpu
def test_sim_tmi_default_output_path(M, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "7")
    print("sim_tmi default path", flush=True)
    assert M.sim_tmi_main(["8", "1", "1", "2", "0", "0.25"]) == 0
    want = tmp_path / "csv/tmi/haar/tmi_haar_n_8_real_1_0_0.25_res_2.csv"
    assert want.exists(), os.listdir(tmp_path)
    lines = want.read_text().splitlines()
    assert lines[0] == "p,tmi" and len(lines) == 3
    assert lines[1].split(",")[0] == "0" and lines[2].split(",")[0] == "0.25"


@pytest.mark.gpu
def test_sim_tmi_resume_after_interruption_reproduces_the_scan(M, tmp_path, monkeypatch):
    # a complete scan, then the same file cut inside a circuit (and mid-line):
    # the re-issued command keeps the whole circuits, trims the rest and
    # recomputes exactly the same rows (fixed MIPT_SEED)
    monkeypatch.setenv("MIPT_SEED", "99")
    args = ["8", "2", "3", "2", "0.1", "0.2", "1", "1"]
    full = tmp_path / "full.csv"
    print("sim_tmi resume", flush=True)
    assert M.sim_tmi_main(args + [str(full)]) == 0
    text = full.read_text()
    lines = text.splitlines(keepends=True)
    assert len(lines) == 1 + 2 * 3 * 2
    cut = tmp_path / "cut.csv"
    cut.write_text("".join(lines[:6]) + lines[6][:5])  # header + 5 rows + a partial line
    if O.ref_available():
        # the reference's checkpoint scan of the driver's own file: whole
        # circuits (2 cycles each) are kept, the fifth row is dropped
        p_texts = list(dict.fromkeys(l.split(",")[0] for l in lines[1:]))
        ck = O.ref_tmi_resume_load(str(cut), "tmi_haar_all_n_8_real_3_0.1_0.2_res_2.csv", 3, 2, p_texts)
        assert (ck["rows"], ck["dropped_rows"], ck["start_p_index"], ck["start_realization"]) == (4, 1, 0, 2)
        assert ck["kept_bytes"] == len("".join(lines[:5]))
        full_ck = O.ref_tmi_resume_load(str(full), "tmi_haar_all_n_8_real_3_0.1_0.2_res_2.csv", 3, 2, p_texts)
        assert full_ck["rows"] == 12 and full_ck["start_p_index"] == 2
    assert M.sim_tmi_main(args + [str(cut)]) == 0
    assert cut.read_text() == text
    monkeypatch.setenv("SIM_TMI_RESUME", "0")
    assert M.sim_tmi_main(args + [str(cut)]) == 0
    assert cut.read_text() == text
